// z4c_stencil.hpp -- stencil pass of the Z4c right-hand side.
//
// One workgroup (256 threads) takes one field of one x-y tile of a block's
// interior (at most 256 points, one per thread) and marches along z:
//
//  * each thread keeps the 9-point z column of its point in a register
//    ring and loads one new value per plane;
//  * the field on the tile +-4 in x and y goes to LDS per plane, for the x
//    and y stencils;
//  * for fields with second derivatives, the derived planes Dx_z(u) (from
//    7-point register rings over the tile +-3, at most two columns per
//    thread) and Dx_y(u) (from the LDS plane) go to LDS too, for the mixed
//    derivatives;
//  * LDS is double-buffered (two barriers per plane) and the next plane's
//    global loads are issued before the current plane's outputs are formed.
//
// The loop body has no data-dependent branches: every thread issues the
// same loads and stores every plane (threads past the tile read clamped,
// in-range addresses; their stores carry a byte offset past the slot
// buffer's range, which the buffer unit drops), so the memory counter waits
// only for the next plane's loads, not for the stores just issued.
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

// Workgroup configurations of the stencil pass: THREADS threads, a tile of
// at most THREADS interior points, CUR LDS-plane elements and EXT Dx_z
// columns per thread.  The host picks one per block size (by measurement,
// z4c_autotune, or by the fewest thread slots).
template <int TH, int CUR_, int EXT_>
struct StencilCfg {
  static constexpr int THREADS = TH;
  static constexpr int CUR = CUR_;
  static constexpr int EXT = EXT_;
  // LDS doubles of one buffer: plane, Dx_y plane, Dx_z plane
  static constexpr int BUF = CUR * THREADS + 2 * EXT * THREADS;
  // whether a tile of ni x nj interior points fits the thread mapping
  __host__ __device__ static constexpr bool fits(int ni, int nj) {
    return ni * nj <= THREADS && (ni + 8) * (nj + 8) <= CUR * THREADS &&
           (ni + 6) * (nj + 6) <= EXT * THREADS;
  }
};
using StCfg256 = StencilCfg<256, 3, 2>;
using StCfg320 = StencilCfg<320, 3, 2>;
using StCfg192 = StencilCfg<192, 4, 3>;
using StCfg384 = StencilCfg<384, 2, 2>;
// one tile per 25 x 25 plane (24^3-cell blocks), two per 33 x 33 plane
using StCfg640 = StencilCfg<640, 2, 2>;
using StCfg576 = StencilCfg<576, 2, 2>;
constexpr int NSTCFG = 6;

struct StencilArgs {
  const DevBlock* blocks;
  const PlaneTile* tiles;
  int ntiles;
  const double* u;  // state, 22 fields
  double* dv;       // derivative slots (the linearisation buffer in LIN mode)
  double diss;
  double adv_alpha, adv_beta;  // lapse_advect, shift_advect
  int lin_stride;   // LIN mode: doubles per interior vertex of the buffer
};

// Selected lop-sided derivative: st_lx without the shift factor, the
// one-sided stencil chosen by the sign of vx.
__device__ __forceinline__ double st_lsel(const double* v, double lidx, double vx) {
  double dl = 0.0;
  dl += 1.0 * v[0];
  dl += -8.0 * v[1];
  dl += 30.0 * v[2];
  dl += -80.0 * v[3];
  dl += 35.0 * v[4];
  dl += 24.0 * v[5];
  dl += -2.0 * v[6];
  double dr = 0.0;
  dr += -1.0 * v[8];
  dr += 8.0 * v[7];
  dr += -30.0 * v[6];
  dr += 80.0 * v[5];
  dr += -35.0 * v[4];
  dr += -24.0 * v[3];
  dr += 2.0 * v[2];
  return ((vx < 0) ? dl : dr) * lidx;
}

// LIN: the linearisation the adjoint reads instead of the derivative
// slots (first derivatives, all six second derivatives, the selected
// lop-sided derivative per axis; slot stride nint), same stencils.
template <class Cfg, int KIND, bool LIN = false>
__device__ __forceinline__ void stencil_tile(const StencilArgs& a, const PlaneTile T,
                                             int f, double* sm) {
  constexpr int ST_THREADS = Cfg::THREADS;
  constexpr int ST_CUR = Cfg::CUR;
  constexpr int ST_EXT = Cfg::EXT;
  constexpr int ST_BUF = Cfg::BUF;
  constexpr bool metric = KIND == DK_METRIC && !LIN;  // P_ab instead of d_a d_b
  constexpr bool full = KIND == DK_FULL || KIND == DK_METRIC;  // second derivatives
  constexpr bool grad = KIND != DK_ADV;
  const DevBlock B = a.blocks[T.block];
  const int N = B.npts;
  const int sz = B.nn1 * B.nn2;  // z stride
  // buffer addressing: per-thread byte offsets in VGPRs, plane and slot
  // offsets in SGPRs
  const Buf U = make_buf(a.u + B.off * NVAR + (long long)f * N, (long long)N * 8);
  const Buf BETA = make_buf(a.u + B.off * NVAR + (long long)V_BETAX * N, 3LL * N * 8);
  const Buf GM = make_buf(a.u + B.off * NVAR + (long long)V_GXX * N, 6LL * N * 8);
  const Buf DVB = LIN ? make_buf(a.dv + B.ioff * a.lin_stride, (long long)NLIN * B.nint * 8)
                      : make_buf(a.dv + B.doff, (long long)NDV * B.nis * 8);
  const int NS8 = (LIN ? B.nint : B.nis) * 8;
  const int db = dbase(f), sld = s_LD(f);
  const double lc = (f == V_ALPHA) ? a.adv_alpha : ((f >= V_BETAX) ? a.adv_beta : 1.0);
  const int PX = T.ni + 8, PP = PX * (T.nj + 8);
  const int EX = T.ni + 6, EE = EX * (T.nj + 6);
  const int EY = EX * T.nj;  // Dx_y points: tile rows, x in [-3, ni+3)
  const int tid = threadIdx.x;
  const int gi0 = B.il + T.i0 - 4, gj0 = B.jl + T.j0 - 4;  // padded coords of plane (0,0)

  // this thread's output point; threads past the tile mirror point 0 and
  // their stores are dropped
  const int npt = T.ni * T.nj;
  const bool active = tid < npt;
  const int tj = active ? tid / T.ni : 0, ti = active ? tid - tj * T.ni : 0;
  const int o = (ti + 4) + PX * (tj + 4);  // in the LDS plane
  const int w = (ti + 3) + EX * (tj + 3);  // in the derived planes
  const unsigned pxy8 = 8u * ((gi0 + ti + 4) + B.nn1 * (gj0 + tj + 4));
  unsigned q8 = active ? 8u * ((T.i0 + ti) + B.ni * (T.j0 + tj)) : DV_DROP;
  const unsigned qstep8 = active ? 8u * B.ni * B.nj : 0u;

  // this thread's LDS-plane elements and Dx_z columns (byte offsets, clamped)
  unsigned cofs[ST_CUR], eofs[ST_EXT];
#pragma unroll
  for (int r = 0; r < ST_CUR; ++r) {
    const int e = min(tid + r * ST_THREADS, PP - 1);
    const int jj = e / PX, ii = e - jj * PX;
    cofs[r] = 8u * ((gi0 + ii) + B.nn1 * (gj0 + jj));
  }
#pragma unroll
  for (int r = 0; r < ST_EXT; ++r) {
    const int e = min(tid + r * ST_THREADS, EE - 1);
    const int ey = e / EX, ex = e - ey * EX;
    eofs[r] = 8u * ((gi0 + ex + 1) + B.nn1 * (gj0 + ey + 1));
  }
  // Dx_y points of this thread (clamped; duplicates write equal values)
  int yofs[ST_EXT], ydst[ST_EXT];
#pragma unroll
  for (int r = 0; r < ST_EXT; ++r) {
    const int e = min(tid + r * ST_THREADS, EY - 1);
    const int ey = e / EX + 3, ex = e - (ey - 3) * EX;
    yofs[r] = (ex + 1) + PX * (ey + 1);
    ydst[r] = ex + EX * ey;
  }

  // registers: next plane, z rings (zr: z-4..z+4 at the point; er: z-3..z+3)
  double pc[ST_CUR], zr[9], er[ST_EXT][7], pz, pe[ST_EXT], pb[3];
  double bx, by, bz;  // shift at the point, current plane
  double pg[6], guu[6];  // metric fields: g~_ab at the point (next plane), g~^ab (current)

  // Places a loaded plane (pc, pz, pe, pb) into LDS buffer `buf` and the
  // rings.  The loop is rotated so that loads are issued at the top of an
  // iteration and placed at its bottom: no load is pending at the loop
  // header, and waiting for the loads at the bottom does not wait for the
  // stores issued in between.
  auto place = [&](int buf) {
    double* cur = sm + buf * ST_BUF;
    double* dzp = cur + ST_CUR * ST_THREADS + ST_EXT * ST_THREADS;
    zr[8] = pz;
    bx = pb[0];
    by = pb[1];
    bz = pb[2];
    if (metric) {
      const double detg = det3(pg);
      inv3(1.0 / detg, pg, guu);
    }
#pragma unroll
    for (int r = 0; r < ST_CUR; ++r) cur[tid + r * ST_THREADS] = pc[r];
    if (full) {
#pragma unroll
      for (int r = 0; r < ST_EXT; ++r) {
        er[r][6] = pe[r];
        double v[9];
        v[0] = v[8] = 0.0;
#pragma unroll
        for (int m = 0; m < 7; ++m) v[m + 1] = er[r][m];
        dzp[tid + r * ST_THREADS] = st_dx(v, B.cidx1[2]);
      }
    }
  };
  auto load_plane = [&](int k1) {
#pragma unroll
    for (int r = 0; r < ST_CUR; ++r) pc[r] = bload(U, cofs[r], 8 * sz * k1);
    pz = bload(U, pxy8, 8 * sz * (k1 + 4));
#pragma unroll
    for (int c = 0; c < 3; ++c) pb[c] = bload(BETA, pxy8, 8 * (c * N + sz * k1));
    if (full) {
#pragma unroll
      for (int r = 0; r < ST_EXT; ++r) pe[r] = bload(U, eofs[r], 8 * sz * (k1 + 3));
    }
    if (metric) {
#pragma unroll
      for (int c = 0; c < 6; ++c) pg[c] = bload(GM, pxy8, 8 * (c * N + sz * k1));
    }
  };

  {
    const int k0 = B.kl;
#pragma unroll
    for (int m = 0; m < 8; ++m) zr[m] = bload(U, pxy8, 8 * sz * (k0 + m - 4));
#pragma unroll
    for (int r = 0; r < ST_EXT; ++r) {
#pragma unroll
      for (int m = 0; m < 7; ++m) er[r][m] = 0.0;
      pe[r] = 0.0;
      if (full) {
#pragma unroll
        for (int m = 0; m < 6; ++m) er[r][m] = bload(U, eofs[r], 8 * sz * (k0 + m - 3));
      }
    }
    load_plane(k0);
    place(0);
  }

  for (int kz = B.kl; kz <= B.ku; ++kz) {
    const int buf = (kz - B.kl) & 1;
    double* cur = sm + buf * ST_BUF;          // field on the tile +-4 (x, y)
    double* dyp = cur + ST_CUR * ST_THREADS;  // Dx_y(u) on the tile +-3
    double* dzp = dyp + ST_EXT * ST_THREADS;  // Dx_z(u) on the tile +-3
    // the next plane's loads (the last plane re-reads itself)
    load_plane(min(kz + 1, B.ku));
    __syncthreads();
    if (full) {
#pragma unroll
      for (int r = 0; r < ST_EXT; ++r) {
        double yv[9];
        yv[0] = yv[8] = 0.0;
#pragma unroll
        for (int m = 1; m < 8; ++m) yv[m] = cur[yofs[r] + (m - 4) * PX];
        dyp[ydst[r]] = st_dx(yv, B.cidx1[1]);
      }
      __syncthreads();
    }
    // slot stores of this plane (dropped for lanes past the tile, whose
    // offset lies beyond the buffer's range)
    auto put = [&](double v, int slot) { bstore_dv(DVB, v, q8, slot * NS8); };
    // one direction at a time (x, y, then z from the ring), so that a
    // single 9-point array is live; L and D accumulate in x, y, z order
    double L = 0.0, D;
    double dd[6];  // second derivatives, SYM2 order
    {
      double xv[9];
#pragma unroll
      for (int m = 0; m < 9; ++m) xv[m] = cur[o + (m - 4)];
      if (grad) put(st_dx(xv, B.cidx1[0]), LIN ? l_d1(f, 0) : db + 0);
      if (full) dd[0] = st_dxx(xv, B.cidx2[0]);
      if (LIN) {
        put(st_lsel(xv, B.lidx[0], bx), l_ls(f, 0));
      } else {
        L += st_lx(xv, B.lidx[0], bx);
        D = st_diss(xv, B.cidxd[0], a.diss);
      }
    }
    {
      double yv[9];
#pragma unroll
      for (int m = 0; m < 9; ++m) yv[m] = cur[o + (m - 4) * PX];
      if (grad) put(st_dx(yv, B.cidx1[1]), LIN ? l_d1(f, 1) : db + 1);
      if (full) dd[3] = st_dxx(yv, B.cidx2[1]);
      if (LIN) {
        put(st_lsel(yv, B.lidx[1], by), l_ls(f, 1));
      } else {
        L += st_lx(yv, B.lidx[1], by);
        D += st_diss(yv, B.cidxd[1], a.diss);
      }
    }
    if (grad) put(st_dx(zr, B.cidx1[2]), LIN ? l_d1(f, 2) : db + 2);
    if (full) dd[5] = st_dxx(zr, B.cidx2[2]);
    if (LIN) {
      put(st_lsel(zr, B.lidx[2], bz), l_ls(f, 2));
    } else {
      // advective sum with the shift at the point (calculate_z4c_rhs.cpp:184-226)
      L += st_lx(zr, B.lidx[2], bz);
      D += st_diss(zr, B.cidxd[2], a.diss);
      put(lc * L + D, sld);
    }
    if (full) {
      double v[9];
      v[0] = v[8] = 0.0;
#pragma unroll
      for (int m = 1; m < 8; ++m) v[m] = dyp[w + (m - 4)];
      dd[1] = st_dx(v, B.cidx1[0]);  // d_x(d_y u)
#pragma unroll
      for (int m = 1; m < 8; ++m) v[m] = dzp[w + (m - 4)];
      dd[2] = st_dx(v, B.cidx1[0]);  // d_x(d_z u)
#pragma unroll
      for (int m = 1; m < 8; ++m) v[m] = dzp[w + (m - 4) * EX];
      dd[4] = st_dx(v, B.cidx1[1]);  // d_y(d_z u)
      if (metric) {
        // sum_{c<=d} w_cd g~^cd d_c d_d g~_ab in the reference's c, d order
        double P = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int d = c; d < 3; ++d)
            P += ((c == d) ? 0.5 : 1.0) * guu[sym(c, d)] * dd[sym(c, d)];
        put(P, db + 3);
      } else {
#pragma unroll
        for (int ab = 0; ab < 6; ++ab) put(dd[ab], LIN ? l_d2(f, ab) : db + 3 + ab);
      }
    }
    q8 += qstep8;
#pragma unroll
    for (int m = 0; m < 8; ++m) zr[m] = zr[m + 1];
#pragma unroll
    for (int r = 0; r < ST_EXT; ++r)
#pragma unroll
      for (int m = 0; m < 6; ++m) er[r][m] = er[r][m + 1];
    // the other LDS buffer is free: every thread passed this plane's first
    // barrier after finishing the previous plane, which read it
    place(buf ^ 1);
  }
}

}  // namespace z4c
