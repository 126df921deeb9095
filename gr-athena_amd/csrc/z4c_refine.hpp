// Vertex-centred restriction and prolongation between refinement levels
// (src/mesh/mesh_refinement_vc.cpp, 3-D, NGHOST = 4, NCGHOST = 4).
//
// Coarse buffer of a block (MeshRefinement's coarse representation,
// src/mesh/meshblock.cpp:611-640, 851-895): nx_d/2 + 2*NCGHOST + 1 vertices
// per axis, coarse interior civs = NCGHOST .. cive = civs + nx_d/2, coarse
// spacing 2 dx, coarse vertex civs at the fine vertex il.  Fine vertex f and
// coarse vertex c coincide when f = 2 (c - RF_H) (RF_H = H_NCGHOST = 2); the
// odd fine vertices between them are prolongated.
//
// Prolongation is ProlongateVertexCenteredValues (:283-772) with even NGHOST
// and DBG_SYMMETRIZE_P_OP / DBG_SYMMETRIZE_CHEAP on (defs.hpp.in:17,20): per
// fine vertex one of eight cases by the parity of its three indices
// (even: injected coordinate, odd: prolongated coordinate), each a tensor
// product of the 6-point Lagrange midpoint weights
// InterpolateLagrangeUniform<3>::coeff (interp_univariate_stencils.tpp:99-101)
// whose mirrored coarse values are added first (FloatingPoint::sum_corners,
// utils/floating_point.hpp:44-55) in the argument order of each case.  The
// library is built without FMA contraction, so every sum and product rounds
// as in that loop nest.
#pragma once

#define RF_NCG 4                     // NCGHOST
#define RF_H ((2 * RF_NCG - 4) / 2)  // H_NCGHOST (mesh_refinement_vc.cpp:305)
#define RF_HSZ (RF_H + 1)            // H_SZ: half-width of the Lagrange stencil

// 6-point Lagrange weights at the midpoint, outermost first
__host__ __device__ inline double rf_lag(int q) {
  return q == 0 ? 3.0 / 256.0 : (q == 1 ? -25.0 / 256.0 : 75.0 / 128.0);
}

// RestrictVertexCenteredIndicialHelper (:24-45): fine index of coarse c
__host__ __device__ inline int rf_fine_of(int c, int cvs, int cve, int vs, int ve) {
  if (c < cvs) return vs - 2 * (cvs - c);
  if (c > cve) return ve + 2 * (c - cve);
  return 2 * (c - cvs) + vs;
}

// One prolongation entry: the coarse index box [s, e] of
// ProlongateVertexCenteredValues on one block, with its derived prolongated
// ranges [sp, ep) (si_prl / ei_prl of :300-340).  Per axis, fine vertex
// 2 (c - RF_H) is injected from coarse c in [s, e] and fine vertex
// 2 (c - RF_H) + 1 is prolongated around coarse c in [sp, ep); a vertex is
// written when all three of its indices are.
struct RfEntry {
  int blk;
  int s[3], e[3];
  int sp[3], ep[3];
};

// Coarse range [c0, c0 + n) of case bit `prl` (0 injected, 1 prolongated)
// on axis a of an entry; n <= 0 when the entry has no such vertex.
__host__ __device__ inline void rf_case_range(const RfEntry& E, int a, int prl, int& c0, int& n) {
  if (prl) {
    c0 = E.sp[a];
    n = E.ep[a] - E.sp[a];
  } else {
    c0 = E.s[a];
    n = E.e[a] - E.s[a] + 1;
  }
}

// Value of one fine vertex from the coarse field C (strides 1, s2, s3) at
// coarse base (i, j, k), case bits m = prl_i | prl_j << 1 | prl_k << 2.
__host__ __device__ inline double rf_prolong_point(const double* __restrict__ C, long s2,
                                                   long s3, int i, int j, int k, int m) {
  auto CO = [&](int kk, int jj, int ii) { return C[ii + s2 * jj + s3 * kk]; };
  double acc = 0.0;
  switch (m) {
    case 7:  // (prl, prl, prl)
      for (int dk = 0; dk < RF_HSZ; ++dk) {
        const int ku = k + dk + 1, kl = k - dk;
        const double lck = rf_lag(RF_HSZ - dk - 1);
        for (int dj = 0; dj < RF_HSZ; ++dj) {
          const int ju = j + dj + 1, jl = j - dj;
          const double lckj = lck * rf_lag(RF_HSZ - dj - 1);
          for (int di = 0; di < RF_HSZ; ++di) {
            const int iu = i + di + 1, il = i - di;
            const double w = lckj * rf_lag(RF_HSZ - di - 1);
            const double s = ((CO(ku, ju, iu) + CO(kl, jl, il)) + (CO(ku, ju, il) + CO(kl, jl, iu))) +
                             ((CO(kl, ju, il) + CO(ku, jl, iu)) + (CO(kl, ju, iu) + CO(ku, jl, il)));
            acc += w * s;
          }
        }
      }
      return acc;
    case 3:  // (inj, prl, prl): k injected
      for (int dj = 0; dj < RF_HSZ; ++dj) {
        const int ju = j + dj + 1, jl = j - dj;
        const double lcj = rf_lag(RF_HSZ - dj - 1);
        for (int di = 0; di < RF_HSZ; ++di) {
          const int iu = i + di + 1, il = i - di;
          const double w = lcj * rf_lag(RF_HSZ - di - 1);
          acc += w * ((CO(k, ju, iu) + CO(k, jl, il)) + (CO(k, jl, iu) + CO(k, ju, il)));
        }
      }
      return acc;
    case 5:  // (prl, inj, prl): j injected
      for (int dk = 0; dk < RF_HSZ; ++dk) {
        const int ku = k + dk + 1, kl = k - dk;
        const double lck = rf_lag(RF_HSZ - dk - 1);
        for (int di = 0; di < RF_HSZ; ++di) {
          const int iu = i + di + 1, il = i - di;
          const double w = lck * rf_lag(RF_HSZ - di - 1);
          acc += w * ((CO(kl, j, iu) + CO(ku, j, il)) + (CO(ku, j, iu) + CO(kl, j, il)));
        }
      }
      return acc;
    case 6:  // (prl, prl, inj): i injected
      for (int dk = 0; dk < RF_HSZ; ++dk) {
        const int ku = k + dk + 1, kl = k - dk;
        const double lck = rf_lag(RF_HSZ - dk - 1);
        for (int dj = 0; dj < RF_HSZ; ++dj) {
          const int ju = j + dj + 1, jl = j - dj;
          const double w = lck * rf_lag(RF_HSZ - dj - 1);
          acc += w * ((CO(ku, ju, i) + CO(kl, jl, i)) + (CO(kl, ju, i) + CO(ku, jl, i)));
        }
      }
      return acc;
    case 1:  // (inj, inj, prl)
      for (int di = 0; di < RF_HSZ; ++di)
        acc += rf_lag(RF_HSZ - di - 1) * (CO(k, j, i - di) + CO(k, j, i + di + 1));
      return acc;
    case 2:  // (inj, prl, inj)
      for (int dj = 0; dj < RF_HSZ; ++dj)
        acc += rf_lag(RF_HSZ - dj - 1) * (CO(k, j - dj, i) + CO(k, j + dj + 1, i));
      return acc;
    case 4:  // (prl, inj, inj)
      for (int dk = 0; dk < RF_HSZ; ++dk)
        acc += rf_lag(RF_HSZ - dk - 1) * (CO(k - dk, j, i) + CO(k + dk + 1, j, i));
      return acc;
    default:  // (inj, inj, inj)
      return CO(k, j, i);
  }
}
