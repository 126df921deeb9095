// z4c_device.hpp -- gfx950 device building blocks of the Z4c hot path:
// block table, finite-difference stencils and 3x3 symmetric algebra.
//
// Arithmetic is FP64 throughout.  Every stencil and algebraic expression is
// evaluated in the operation order of the reference so results agree with
// the CPU path to rounding (the build disables FMA contraction):
//   stencils   src/utils/finite_differencing.tpp (DBG_SYMMETRIZE_FD branches,
//              on by src/defs.hpp.in:18), NGHOST = 4
//   algebra    src/utils/linear_algebra.hpp:47-118,271-290,
//              src/utils/floating_point.hpp:110-131
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace z4c {

constexpr int NG = 4;
constexpr int NVAR = 22;
constexpr int ND1 = 33;  // d(alpha, chi, beta^c, g~_cd)/dx^b, index b*11 + f

// Z4c variable indices (src/z4c/z4c.hpp:54-88)
enum : int {
  V_CHI = 0, V_GXX = 1, V_KHAT = 7, V_AXX = 8, V_GAMX = 14, V_THETA = 17,
  V_ALPHA = 18, V_BETAX = 19
};
// ADM / constraint / matter indices (src/z4c/z4c.hpp:99-190)
enum : int { A_ALPHA = 0, A_BETAX = 1, A_GXX = 4, A_KXX = 10, A_PSI4 = 16, NADM = 17 };
enum : int { C_C = 0, C_H = 1, C_M = 2, C_Z = 3, C_MX = 4, NCON = 7 };
enum : int { M_RHO = 0, M_SX = 1, M_SXX = 4, NMAT = 10 };
// prepared first-derivative fields, per direction b: f = 0 alpha, 1 chi,
// 2..4 beta^x..z, 5..10 g~_xx..g~_zz
enum : int { F_ALPHA = 0, F_CHI = 1, F_BETA = 2, F_G = 5 };
// prepared-derivative field index of direction b, field f
__host__ __device__ constexpr int d1f(int b, int f) { return b * 11 + f; }

// SYM2 component order xx,xy,xz,yy,yz,zz (src/athena_tensor.hpp:57-179)
__host__ __device__ constexpr int sym(int a, int b) {
  return (a == b) ? (a == 0 ? 0 : (a == 1 ? 3 : 5))
                  : ((a + b == 1) ? 1 : ((a + b == 2) ? 2 : 4));
}

// Device-side description of one MeshBlock.
// Derivative slots of a block: NDV rows of nis doubles, nint + DV_PAD rounded
// up to 16 (each row starts on a 128-byte line); the entries past the
// interior are never read or written.
constexpr int DV_PAD = 16;
// Byte offset of a store that the buffer unit drops: past every slot
// buffer's range (< 2^31 bytes, checked at registration), and with any slot
// offset (< 2^31) added it stays below 2^32.
constexpr unsigned DV_DROP = 0x80000000u;

struct DevBlock {
  long long off;     // point offset in state buffers (block data at off*nvar)
  long long ioff;    // interior-point offset in interior-indexed buffers
  long long doff;    // offset of the block's derivative slots (NDV * nis each)
  int nis;           // derivative-slot row length (nint + DV_PAD, multiple of 16)
  int nn1, nn2, nn3;
  int il, iu, jl, ju, kl, ku;
  int npts;          // nn1*nn2*nn3
  int ni, nj, nk;    // interior extents
  int nint;          // ni*nj*nk
  int faces;         // Sommerfeld face bits
  int pfaces;        // faces with the physical ghost fill (Sommerfeld | outflow)
  double dx[3];
  double x0[3];
  double cidx1[3];   // 1/(dx*60)        Dx
  double cidx2[3];   // (1/dx)^2/180     Dxx
  double cidx1lo[3]; // 1/(dx*2)         Ds
  double cidxd[3];   // -2^-8/(dx*1)     Diss (times diss)
  double lidx[3];    // 1/(dx*60)        Lx
};

// ---------------------------------------------------------------------------
// Derivative slots per interior point, written by the stencil pass and read
// by the algebra pass (field stride = interior points of the block).
//   FULL fields (chi, alpha, beta^i): d_a (3), d_a d_b (6, SYM2 order;
//        off-diagonal = nested Dx(a, Dx(b, .)) with a < b), LD
//   METRIC fields (g~_ab): d_a (3), P_ab, LD with
//        P_ab = sum_{c<=d} w_cd g~^cd d_c d_d g~_ab  (w = 1/2 on the
//        diagonal, 1 off it; SYM2 order), the second-derivative part of the
//        Ricci tensor (calculate_z4c_rhs.cpp:270-286), formed where the
//        second derivatives are
//   GRAD fields (Khat, Gam~^i, Theta):       d_a (3), LD
//   ADV  fields (A~_ij):                      LD
// LD = c_n L + D with L = sum_a Lx(a, beta^a, .) in a order, D = (Diss_x +
// Diss_y) + Diss_z and c_n the advection coefficient of the field's
// right-hand side (lapse_advect for alpha, shift_advect for beta^i, 1
// otherwise): every right-hand side takes its advective and KO terms as
// plain additive terms (calculate_z4c_rhs.cpp:694-902, 908-915).
// ---------------------------------------------------------------------------
enum : int { DK_FULL = 0, DK_GRAD = 1, DK_ADV = 2, DK_METRIC = 3 };
__host__ __device__ constexpr int dkind(int n) {
  return (n >= V_GXX && n < V_GXX + 6)
             ? DK_METRIC
             : ((n == V_CHI || n == V_ALPHA || n >= V_BETAX)
                    ? DK_FULL
                    : ((n == V_KHAT || n == V_THETA || (n >= V_GAMX && n < V_GAMX + 3))
                           ? DK_GRAD
                           : DK_ADV));
}
__host__ __device__ constexpr int dnslots(int kind) {
  return kind == DK_FULL ? 10 : (kind == DK_METRIC ? 5 : (kind == DK_GRAD ? 4 : 1));
}
__host__ __device__ constexpr int dbase(int n) {
  int s = 0;
  for (int m = 0; m < n; ++m) s += dnslots(dkind(m));
  return s;
}
constexpr int NDV = 106;
static_assert(dbase(NVAR) == NDV, "derivative slot layout");
__host__ __device__ constexpr int s_d(int n, int a) { return dbase(n) + a; }
__host__ __device__ constexpr int s_dd(int n, int ab) { return dbase(n) + 3 + ab; }
__host__ __device__ constexpr int s_lap(int n) { return dbase(n) + 3; }
__host__ __device__ constexpr int s_LD(int n) { return dbase(n) + dnslots(dkind(n)) - 1; }

// Linearisation of the state at an interior vertex (the finite-difference
// values the right-hand side reads), NLIN doubles per vertex:
//   first derivatives d_a u_n of the 16 fields with one (lin1 index), 48;
//   second derivatives d_a d_b u_n of the 11 fields with them (lin2), 66;
//   selected lop-sided derivatives Lsel_a(u_n) of all 22 fields, 66.
// Derivative cotangents per vertex (NADJ doubles): the 48 first- and 66
// second-derivative cotangents, in the same order.
__host__ __device__ constexpr int lin1(int n) {  // chi, g~, Khat, Gam~, Theta, alpha, beta
  return n <= V_KHAT ? n : (n >= V_GAMX ? n - 6 : -1);
}
__host__ __device__ constexpr int lin2(int n) {  // chi, g~, alpha, beta
  return n < V_KHAT ? n : (n >= V_ALPHA ? n - 11 : -1);
}
constexpr int NLIN1 = 16 * 3;
constexpr int NLIN2 = 11 * 6;
constexpr int NLIN = NLIN1 + NLIN2 + NVAR * 3;
constexpr int NADJ = NLIN1 + NLIN2;
__host__ __device__ constexpr int l_d1(int n, int a) { return lin1(n) * 3 + a; }
__host__ __device__ constexpr int l_d2(int n, int ab) { return NLIN1 + lin2(n) * 6 + ab; }
__host__ __device__ constexpr int l_ls(int n, int a) { return NLIN1 + NLIN2 + n * 3 + a; }

// Tile of the interior x-y plane of a block for the z-marching stencil pass
// (interior-relative origin i0, j0 and extents ni, nj).
struct PlaneTile {
  int block;
  int i0, j0, ni, nj;
};

// A tile = up to TILE consecutive points of one box of one block.
struct Tile {
  int block;
  int box;  // index into the launch's box table
  int start;
};

constexpr int TILE = 256;

// Box kinds used by the launch tables.
// BOX_GHOST holds six boxes per block that together cover exactly the
// ghost vertices (the padded block minus the interior).
enum BoxKind : int { BOX_INTERIOR = 0, BOX_PREP = 1, BOX_WHOLE = 2, BOX_GHOST = 3, NBOX = 4 };

struct Box {
  int lo[3];   // i, j, k
  int n[3];    // extents
};

// ---------------------------------------------------------------------------
// Buffer addressing.  A workgroup works on one block, so the block's field
// buffer is a wave-uniform buffer resource (SGPRs).  Each thread keeps one
// 32-bit byte offset for its point (v0) and one per axis shifted four points
// back (vd[d] = (p - 4*stride_d)*8), so every stencil access is a
// buffer_load with a VGPR offset and a uniform SGPR offset
// (field*N + (m+4)*stride_d)*8 >= 0 -- no 64-bit per-lane addresses.
// Stencil points stay inside the block for interior points (index >= 4).
// ---------------------------------------------------------------------------
constexpr int RSRC_WORD3 = 0x00020000;  // gfx9 family (incl. gfx950) raw buffer

struct Buf {
  __amdgpu_buffer_rsrc_t r;
};

__device__ __forceinline__ Buf make_buf(const double* base, long long nbytes) {
  const int nr = (nbytes > 0x7fffffffLL) ? 0x7fffffff : (int)nbytes;
  return Buf{__builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(base), (short)0, nr,
                                               RSRC_WORD3)};
}

__device__ __forceinline__ double bload(const Buf& b, unsigned vbyte, int sbyte) {
  return __builtin_bit_cast(double,
                            __builtin_amdgcn_raw_buffer_load_b64(b.r, vbyte, sbyte, 0));
}

__device__ __forceinline__ void bstore(const Buf& b, double v, unsigned vbyte, int sbyte) {
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), b.r, vbyte, sbyte, 0);
}

// Cache policy of the derivative-slot stores (stencil pass) and loads
// (assembly).  The slots are written once and read once per stage, far more
// than the L2 holds: the stores are non-temporal (NT bit, 2), which measured
// 5% faster on the stencil pass; NT loads measured slower, so loads keep the
// default policy.
#ifndef Z4C_DV_STORE_CPOL
#define Z4C_DV_STORE_CPOL 2
#endif
#ifndef Z4C_DV_LOAD_CPOL
#define Z4C_DV_LOAD_CPOL 0
#endif

__device__ __forceinline__ double bload_dv(const Buf& b, unsigned vbyte, int sbyte) {
  return __builtin_bit_cast(
      double, __builtin_amdgcn_raw_buffer_load_b64(b.r, vbyte, sbyte, Z4C_DV_LOAD_CPOL));
}

__device__ __forceinline__ void bstore_dv(const Buf& b, double v, unsigned vbyte, int sbyte) {
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), b.r, vbyte, sbyte,
                                        Z4C_DV_STORE_CPOL);
}

// Per-thread addressing of one point.
struct Pt {
  unsigned v0;     // p*8
  unsigned vd[3];  // (p - 4*st[d])*8
  int sd[3];       // st[d]*8
  int fb;          // field stride N*8
};

__device__ __forceinline__ Pt make_pt(int p, int nn1, int nn2, int N) {
  Pt P;
  const int st[3] = {1, nn1, nn1 * nn2};
  P.v0 = (unsigned)p * 8u;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    P.vd[d] = (unsigned)(p - 4 * st[d]) * 8u;
    P.sd[d] = st[d] * 8;
  }
  P.fb = N * 8;
  return P;
}

// value of field n at the point
__device__ __forceinline__ double ldc(const Buf& b, const Pt& P, int n) {
  return bload(b, P.v0, n * P.fb);
}
// value of field n at offset m (-4..4) along axis d
__device__ __forceinline__ double ldm(const Buf& b, const Pt& P, int n, int d, int m) {
  return bload(b, P.vd[d], n * P.fb + (m + 4) * P.sd[d]);
}

// ---------------------------------------------------------------------------
// Finite differences of field n along axis d at the point.
// ---------------------------------------------------------------------------
// FDStencilCenteredDegreeOdd<1,3> {-1, 9, -45}/60
__device__ __forceinline__ double fd_dx(const Buf& b, const Pt& P, int n, int d,
                                        double cidx) {
  double out = 0.0;
  out += -1.0 * (ldm(b, P, n, d, -3) - ldm(b, P, n, d, 3));
  out += 9.0 * (ldm(b, P, n, d, -2) - ldm(b, P, n, d, 2));
  out += -45.0 * (ldm(b, P, n, d, -1) - ldm(b, P, n, d, 1));
  return out * cidx;
}

// FDStencilCenteredDegreeEven<2,3> {2, -27, 270, -490}/180
__device__ __forceinline__ double fd_dxx(const Buf& b, const Pt& P, int n, int d,
                                         double cidx2) {
  double out = 0.0;
  out += 2.0 * (ldm(b, P, n, d, -3) + ldm(b, P, n, d, 3));
  out += -27.0 * (ldm(b, P, n, d, -2) + ldm(b, P, n, d, 2));
  out += 270.0 * (ldm(b, P, n, d, -1) + ldm(b, P, n, d, 1));
  out += -490.0 * ldm(b, P, n, d, 0);
  return out * cidx2;
}

// FDStencilCenteredDegreeOdd<1,1> {-1}/2
__device__ __forceinline__ double fd_ds(const Buf& b, const Pt& P, int n, int d,
                                        double cidxlo) {
  double out = 0.0;
  out += -1.0 * (ldm(b, P, n, d, -1) - ldm(b, P, n, d, 1));
  return out * cidxlo;
}

// FDStencilCenteredDegreeEven<8,4> {1,-8,28,-56,70}, scaled -2^-8/dx * diss
__device__ __forceinline__ double fd_diss(const Buf& b, const Pt& P, int n, int d,
                                          double cidxd, double diss) {
  double out = 0.0;
  out += 1.0 * (ldm(b, P, n, d, -4) + ldm(b, P, n, d, 4));
  out += -8.0 * (ldm(b, P, n, d, -3) + ldm(b, P, n, d, 3));
  out += 28.0 * (ldm(b, P, n, d, -2) + ldm(b, P, n, d, 2));
  out += -56.0 * (ldm(b, P, n, d, -1) + ldm(b, P, n, d, 1));
  out += 70.0 * ldm(b, P, n, d, 0);
  return out * cidxd * diss;
}

// Lop-sided advective derivative: FDStencilBiasedLeft<1,4,2> on -4..+2 if
// vx < 0, FDStencilBiasedRight<1,4,2> on -2..+4 otherwise; both sums are
// formed (as the reference does) and one is selected.
__device__ __forceinline__ double fd_lx(const Buf& b, const Pt& P, int n, int d,
                                        double lidx, double vx) {
  const double um4 = ldm(b, P, n, d, -4), um3 = ldm(b, P, n, d, -3),
               um2 = ldm(b, P, n, d, -2), um1 = ldm(b, P, n, d, -1),
               u0 = ldm(b, P, n, d, 0), up1 = ldm(b, P, n, d, 1),
               up2 = ldm(b, P, n, d, 2), up3 = ldm(b, P, n, d, 3),
               up4 = ldm(b, P, n, d, 4);
  double dl = 0.0;
  dl += 1.0 * um4;
  dl += -8.0 * um3;
  dl += 30.0 * um2;
  dl += -80.0 * um1;
  dl += 35.0 * u0;
  dl += 24.0 * up1;
  dl += -2.0 * up2;
  double dr = 0.0;
  dr += -1.0 * up4;
  dr += 8.0 * up3;
  dr += -30.0 * up2;
  dr += 80.0 * up1;
  dr += -35.0 * u0;
  dr += -24.0 * um1;
  dr += 2.0 * um2;
  return ((vx < 0) ? (vx * dl) : (vx * dr)) * lidx;
}

// ---------------------------------------------------------------------------
// The same stencils on 9 gathered values v[0..8] (offsets -4..+4), used by
// the LDS stencil pass.  Identical operations in identical order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double st_dx(const double* v, double cidx) {
  double out = 0.0;
  out += -1.0 * (v[1] - v[7]);
  out += 9.0 * (v[2] - v[6]);
  out += -45.0 * (v[3] - v[5]);
  return out * cidx;
}
__device__ __forceinline__ double st_dxx(const double* v, double cidx2) {
  double out = 0.0;
  out += 2.0 * (v[1] + v[7]);
  out += -27.0 * (v[2] + v[6]);
  out += 270.0 * (v[3] + v[5]);
  out += -490.0 * v[4];
  return out * cidx2;
}
__device__ __forceinline__ double st_diss(const double* v, double cidxd, double diss) {
  double out = 0.0;
  out += 1.0 * (v[0] + v[8]);
  out += -8.0 * (v[1] + v[7]);
  out += 28.0 * (v[2] + v[6]);
  out += -56.0 * (v[3] + v[5]);
  out += 70.0 * v[4];
  return out * cidxd * diss;
}
__device__ __forceinline__ double st_lx(const double* v, double lidx, double vx) {
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
  return ((vx < 0) ? (vx * dl) : (vx * dr)) * lidx;
}

// ---------------------------------------------------------------------------
// 3x3 symmetric algebra (SYM2 order)
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ double det3(const double* g) {
  const double gxx = g[0], gxy = g[1], gxz = g[2], gyy = g[3], gyz = g[4],
               gzz = g[5];
  return (-(gxz * gxz) * gyy + 2 * gxy * gxz * gyz - gxx * (gyz * gyz) -
          (gxy * gxy) * gzz + gxx * gyy * gzz);
}

__host__ __device__ __forceinline__ void inv3(double oodetg, const double* g, double* u) {
  const double gxx = g[0], gxy = g[1], gxz = g[2], gyy = g[3], gyz = g[4],
               gzz = g[5];
  u[0] = (-(gyz * gyz) + gyy * gzz) * oodetg;
  u[1] = (gxz * gyz - gxy * gzz) * oodetg;
  u[3] = (-(gxz * gxz) + gxx * gzz) * oodetg;
  u[2] = (-gxz * gyy + gxy * gyz) * oodetg;
  u[4] = (gxy * gxz - gxx * gyz) * oodetg;
  u[5] = (-(gxy * gxy) + gxx * gyy) * oodetg;
}

__host__ __device__ __forceinline__ double trace_rank2(double detginv, const double* g,
                                              const double* A) {
  const double gxx = g[0], gxy = g[1], gxz = g[2], gyy = g[3], gyz = g[4],
               gzz = g[5];
  const double Axx = A[0], Axy = A[1], Axz = A[2], Ayy = A[3], Ayz = A[4],
               Azz = A[5];
  return detginv * (-2. * Ayz * gxx * gyz + Axx * gyy * gzz +
                    gxx * (Azz * gyy + Ayy * gzz) +
                    2. * (gxz * (Ayz * gxy - Axz * gyy + Axy * gyz) +
                          gxy * (Axz * gyz - Axy * gzz)) -
                    Azz * (gxy * gxy) - Ayy * (gxz * gxz) - Axx * (gyz * gyz));
}

__host__ __device__ __forceinline__ double dmax(double a, double b) {
  return (a < b) ? b : a;  // std::max semantics
}

__host__ __device__ __forceinline__ double regularize_nth_root(double value,
                                                      double value_reg,
                                                      double eps_floor,
                                                      double npow) {
  const double eps = value - value_reg;
  return (eps < eps_floor) ? (1.0 + eps_floor / npow) : pow(value, 1.0 / npow);
}

__host__ __device__ __forceinline__ double regularize_near_zero(double v, double e) {
  return (v >= 0) ? ((v < e) ? e : v) : ((v > -e) ? -e : v);
}

}  // namespace z4c
