// z4c_stencil.hpp -- stencil pass of the Z4c right-hand side.
//
// One workgroup (256 threads) takes one field of one x-y tile of a block's
// interior (at most 256 points, one per thread) and marches along z:
//
//  * each thread keeps the 9-point z column of its point in a register ring
//    and loads one new value per plane;
//  * the field on the tile +-4 in x and y goes to LDS per plane, for the x
//    and y stencils;
//  * for fields with second derivatives, the derived planes Dx_z(u) (from
//    7-point register rings over the tile +-3, at most two columns per
//    thread) and Dx_y(u) (from the LDS plane) go to LDS too, for the mixed
//    derivatives;
//  * LDS is double-buffered (two barriers per plane) and the next plane's
//    global loads are issued before the current plane's outputs are formed.
//
// The mixed second derivatives are the reference's nested Dx(a, Dx(b, u)),
// a < b (src/z4c/calculate_z4c_rhs.cpp:132-175, src/z4c/adm_z4c.cpp:977-1037),
// with the same operation order, so the values equal those of first
// preparing 3-D first-derivative arrays and differentiating them again.
//
// Per interior point the pass writes the derivative slots the algebra needs
// (layout in z4c_device.hpp): first derivatives, second derivatives, the
// advective sum sum_a Lx(a, beta^a, u) and the KO sum.
#pragma once

#include "z4c_device.hpp"

namespace z4c {

constexpr int ST_THREADS = 256;  // workgroup size; tiles hold <= 256 points
constexpr int ST_CUR = 3;        // LDS-plane elements per thread: (ni+8)(nj+8) <= 768
constexpr int ST_EXT = 2;        // Dx_z columns per thread: (ni+6)(nj+6) <= 512

struct StencilArgs {
  const DevBlock* blocks;
  const PlaneTile* tiles;
  int ntiles;
  const double* u;  // state, 22 fields
  double* dv;       // derivative slots, NDV fields per interior point
  double diss;
};

// Whether a tile of ni x nj interior points fits the thread mapping.
__host__ __device__ constexpr bool stencil_tile_fits(int ni, int nj) {
  return ni * nj <= ST_THREADS && (ni + 8) * (nj + 8) <= ST_CUR * ST_THREADS &&
         (ni + 6) * (nj + 6) <= ST_EXT * ST_THREADS;
}

// LDS doubles of one buffer (plane + two derived planes) of a tile.
__host__ __device__ constexpr int stencil_lds_buffer(int ni, int nj) {
  return (ni + 8) * (nj + 8) + 2 * (ni + 6) * (nj + 6);
}

// LDS doubles needed by a tile (double-buffered).
__host__ __device__ constexpr int stencil_lds_doubles(int ni, int nj) {
  return 2 * stencil_lds_buffer(ni, nj);
}

// The tile and block descriptors are taken by value: held in registers, they
// are not reloaded (and waited for) after every store of the z loop.
__device__ __forceinline__ void stencil_tile(const StencilArgs& a, const PlaneTile T,
                                             int f, double* sm) {
  const DevBlock B = a.blocks[T.block];
  const int N = B.npts;
  const int sz = B.nn1 * B.nn2;  // z stride
  const double* __restrict__ uf = a.u + B.off * NVAR + (long long)f * N;
  const double* __restrict__ ubeta = a.u + B.off * NVAR + (long long)V_BETAX * N;
  double* __restrict__ dv = a.dv + B.ioff * NDV;
  const int db = dbase(f), sl = s_L(f), sd = s_D(f);
  const int NI = B.nint;
  const int PX = T.ni + 8, PP = PX * (T.nj + 8);
  const int EX = T.ni + 6, EE = EX * (T.nj + 6);
  const int SB = stencil_lds_buffer(T.ni, T.nj);
  const bool full = dkind(f) == DK_FULL;
  const bool grad = dkind(f) != DK_ADV;
  const int tid = threadIdx.x;
  const int gi0 = B.il + T.i0 - 4, gj0 = B.jl + T.j0 - 4;  // padded coords of plane (0,0)

  // this thread's output point
  const bool active = tid < T.ni * T.nj;
  const int tj = active ? tid / T.ni : 0, ti = active ? tid - tj * T.ni : 0;
  const int o = (ti + 4) + PX * (tj + 4);  // in the LDS plane
  const int w = (ti + 3) + EX * (tj + 3);  // in the derived planes
  const int pxy = (gi0 + ti + 4) + B.nn1 * (gj0 + tj + 4);

  // this thread's LDS-plane elements and Dx_z columns (xy offsets)
  int cofs[ST_CUR], eofs[ST_EXT];
#pragma unroll
  for (int r = 0; r < ST_CUR; ++r) {
    const int e = tid + r * ST_THREADS;
    const int jj = e / PX, ii = e - jj * PX;
    cofs[r] = e < PP ? (gi0 + ii) + B.nn1 * (gj0 + jj) : -1;
  }
#pragma unroll
  for (int r = 0; r < ST_EXT; ++r) {
    const int e = tid + r * ST_THREADS;
    const int ey = e / EX, ex = e - ey * EX;
    eofs[r] = (full && e < EE) ? (gi0 + ex + 1) + B.nn1 * (gj0 + ey + 1) : -1;
  }

  // registers: next plane, z rings (zr: z-4..z+4 at the point; er: z-3..z+3)
  double pc[ST_CUR], zr[9], er[ST_EXT][7], pz = 0.0, pe[ST_EXT], pb[3] = {0.0, 0.0, 0.0};
  {
    const int k0 = B.kl;
#pragma unroll
    for (int r = 0; r < ST_CUR; ++r) pc[r] = cofs[r] >= 0 ? uf[cofs[r] + sz * k0] : 0.0;
#pragma unroll
    for (int m = 0; m < 9; ++m) zr[m] = 0.0;
    if (active) {
#pragma unroll
      for (int m = 0; m < 8; ++m) zr[m] = uf[pxy + sz * (k0 + m - 4)];
      pz = uf[pxy + sz * (k0 + 4)];
#pragma unroll
      for (int c = 0; c < 3; ++c) pb[c] = ubeta[c * N + pxy + sz * k0];
    }
#pragma unroll
    for (int r = 0; r < ST_EXT; ++r) {
#pragma unroll
      for (int m = 0; m < 7; ++m) er[r][m] = 0.0;
      pe[r] = 0.0;
      if (eofs[r] >= 0) {
#pragma unroll
        for (int m = 0; m < 6; ++m) er[r][m] = uf[eofs[r] + sz * (k0 + m - 3)];
        pe[r] = uf[eofs[r] + sz * (k0 + 3)];
      }
    }
  }

  for (int kz = B.kl; kz <= B.ku; ++kz) {
    double* cur = sm + ((kz - B.kl) & 1) * SB;  // field on the tile +-4 (x, y)
    double* dyp = cur + PP;                     // Dx_y(u) on the tile +-3
    double* dzp = dyp + EE;                     // Dx_z(u) on the tile +-3
    zr[8] = pz;
    const double bx = pb[0], by = pb[1], bz = pb[2];  // shift at the point
#pragma unroll
    for (int r = 0; r < ST_CUR; ++r)
      if (tid + r * ST_THREADS < PP) cur[tid + r * ST_THREADS] = pc[r];
    if (full) {
#pragma unroll
      for (int r = 0; r < ST_EXT; ++r) {
        er[r][6] = pe[r];
        if (eofs[r] >= 0) {
          double v[9];
          v[0] = v[8] = 0.0;
#pragma unroll
          for (int m = 0; m < 7; ++m) v[m + 1] = er[r][m];
          dzp[tid + r * ST_THREADS] = st_dx(v, B.cidx1[2]);
        }
      }
    }
    __syncthreads();
    if (full) {
      // Dx_y(u) on the rows of the tile, x in [-3, ni+3)
      for (int e = tid; e < EX * T.nj; e += ST_THREADS) {
        const int ey = e / EX + 3, ex = e - (ey - 3) * EX;
        const int oc = (ex + 1) + PX * (ey + 1);
        double yv[9];
#pragma unroll
        for (int m = 1; m < 8; ++m) yv[m] = cur[oc + (m - 4) * PX];
        yv[0] = yv[8] = 0.0;
        dyp[ex + EX * ey] = st_dx(yv, B.cidx1[1]);
      }
      __syncthreads();
    }
    // issue the next plane's loads
    if (kz < B.ku) {
      const int k1 = kz + 1;
#pragma unroll
      for (int r = 0; r < ST_CUR; ++r) pc[r] = cofs[r] >= 0 ? uf[cofs[r] + sz * k1] : 0.0;
      if (active) {
        pz = uf[pxy + sz * (k1 + 4)];
#pragma unroll
        for (int c = 0; c < 3; ++c) pb[c] = ubeta[c * N + pxy + sz * k1];
      }
#pragma unroll
      for (int r = 0; r < ST_EXT; ++r)
        if (eofs[r] >= 0) pe[r] = uf[eofs[r] + sz * (k1 + 3)];
    }
    if (active) {
      const int qi = (T.i0 + ti) + B.ni * ((T.j0 + tj) + B.nj * (kz - B.kl));
      double xv[9], yv[9];
#pragma unroll
      for (int m = 0; m < 9; ++m) {
        xv[m] = cur[o + (m - 4)];
        yv[m] = cur[o + (m - 4) * PX];
      }
      if (grad) {
        dv[(long long)(db + 0) * NI + qi] = st_dx(xv, B.cidx1[0]);
        dv[(long long)(db + 1) * NI + qi] = st_dx(yv, B.cidx1[1]);
        dv[(long long)(db + 2) * NI + qi] = st_dx(zr, B.cidx1[2]);
      }
      if (full) {
        double v[9];
        dv[(long long)(db + 3 + 0) * NI + qi] = st_dxx(xv, B.cidx2[0]);
        dv[(long long)(db + 3 + 3) * NI + qi] = st_dxx(yv, B.cidx2[1]);
        dv[(long long)(db + 3 + 5) * NI + qi] = st_dxx(zr, B.cidx2[2]);
        v[0] = v[8] = 0.0;
#pragma unroll
        for (int m = 1; m < 8; ++m) v[m] = dyp[w + (m - 4)];
        dv[(long long)(db + 3 + 1) * NI + qi] = st_dx(v, B.cidx1[0]);  // d_x(d_y u)
#pragma unroll
        for (int m = 1; m < 8; ++m) v[m] = dzp[w + (m - 4)];
        dv[(long long)(db + 3 + 2) * NI + qi] = st_dx(v, B.cidx1[0]);  // d_x(d_z u)
#pragma unroll
        for (int m = 1; m < 8; ++m) v[m] = dzp[w + (m - 4) * EX];
        dv[(long long)(db + 3 + 4) * NI + qi] = st_dx(v, B.cidx1[1]);  // d_y(d_z u)
      }
      // advective sum with the shift at the point (calculate_z4c_rhs.cpp:184-226)
      double L = 0.0;
      L += st_lx(xv, B.lidx[0], bx);
      L += st_lx(yv, B.lidx[1], by);
      L += st_lx(zr, B.lidx[2], bz);
      dv[(long long)sl * NI + qi] = L;
      double D = st_diss(xv, B.cidxd[0], a.diss);
      D += st_diss(yv, B.cidxd[1], a.diss);
      D += st_diss(zr, B.cidxd[2], a.diss);
      dv[(long long)sd * NI + qi] = D;
    }
#pragma unroll
    for (int m = 0; m < 8; ++m) zr[m] = zr[m + 1];
#pragma unroll
    for (int r = 0; r < ST_EXT; ++r)
#pragma unroll
      for (int m = 0; m < 6; ++m) er[r][m] = er[r][m + 1];
  }
}

}  // namespace z4c
