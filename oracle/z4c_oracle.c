/*
 * z4c_oracle.c -- CPU restatement of the GR-Athena++ Z4c hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product; see the header).
 *
 * Every function restates the per-point arithmetic of the reference in the
 * reference's operation order, one grid point at a time instead of one
 * (k,j) pencil at a time.  The pencil structure of the reference does not
 * change any value: each pencil entry i is computed from point-local inputs
 * by the same sequence of IEEE operations.  Build with -ffp-contract=off so
 * that no a*b+c is fused (the reference is built that way, configure.py:1145).
 */
#include "z4c_oracle.h"

#include <float.h>
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Marks the finite-difference functions: the FLOP-counting build
   (flops/fdouble.hpp) attributes their operations to the stencil share;
   empty in the plain build. */
#ifndef ORACLE_FD_SCOPE
#define ORACLE_FD_SCOPE
#endif

/* SYM2 component order xx,xy,xz,yy,yz,zz (src/athena_tensor.hpp:57-179). */
static const int S[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};

/* ------------------------------------------------------------------------ */
/* Finite differences, NGHOST = 4, DBG_SYMMETRIZE_FD (defs.hpp.in:18).       */
/* ------------------------------------------------------------------------ */
typedef struct {
  long s[3];
  double cidx1[3];    /* 1/(dx*60)          Dx     (finite_differencing.tpp:34-36) */
  double cidx2[3];    /* (1/dx)^2/180       Dxx    (:38-40)                        */
  double cidx1_lo[3]; /* 1/(dx*2)           Ds     (:42-44)                        */
  double cidxd[3];    /* -2^-8/(dx*1)       Diss   (:46-48, diss_scaling :24)      */
  double lidx[3];     /* 1/(dx*60)          Lx     (:51-58)                        */
} fdop;

static void fd_init(const oracle_geom* g, fdop* f) {
  /* diss_scaling = pow(2,-2*NGHOST)*(NGHOST%2==0 ? -1 : 1)  (.tpp:24) */
  const double diss_scaling = pow(2, -2 * 4) * (4 % 2 == 0 ? -1 : 1);
  f->s[0] = 1;
  f->s[1] = g->nn1;
  f->s[2] = (long)g->nn1 * g->nn2;
  for (int d = 0; d < 3; ++d) {
    const double dx = g->dx[d];
    const double t = 1.0 / dx;
    f->cidx1[d] = 1.0 / (dx * 60.0);
    f->cidx2[d] = (t * t) / 180.0;
    f->cidx1_lo[d] = 1.0 / (dx * 2.0);
    f->cidxd[d] = diss_scaling / (dx * 1.0);
    f->lidx[d] = 1.0 / (dx * 60.0);
  }
}

/* FDStencilCenteredDegreeOdd<1,3>: {-1, 9, -45} / 60  (stencils.tpp:264-270);
   Dx symmetrised branch finite_differencing.tpp:99-110. */
static inline double Dx(const fdop* f, const double* u, int d) {
  ORACLE_FD_SCOPE;
  const long s = f->s[d];
  const double* pu = u - 3 * s;
  double out = 0.0;
  out += -1.0 * (pu[0] - pu[6 * s]);
  out += 9.0 * (pu[1 * s] - pu[5 * s]);
  out += -45.0 * (pu[2 * s] - pu[4 * s]);
  return out * f->cidx1[d];
}

/* FDStencilCenteredDegreeEven<2,3>: {2,-27,270,-490} / 180 (stencils.tpp:330-336);
   Dxx symmetrised branch finite_differencing.tpp:225-234. */
static inline double Dxx(const fdop* f, const double* u, int d) {
  ORACLE_FD_SCOPE;
  const long s = f->s[d];
  const double* pu = u - 3 * s;
  double out = 0.0;
  out += 2.0 * (pu[0] + pu[6 * s]);
  out += -27.0 * (pu[1 * s] + pu[5 * s]);
  out += 270.0 * (pu[2 * s] + pu[4 * s]);
  out += -490.0 * pu[3 * s];
  return out * f->cidx2[d];
}

/* FDStencilCenteredDegreeOdd<1,1>: {-1} / 2; Ds symmetrised branch .tpp:134-143. */
static inline double Ds(const fdop* f, const double* u, int d) {
  ORACLE_FD_SCOPE;
  const long s = f->s[d];
  const double* pu = u - s;
  double out = 0.0;
  out += -1.0 * (pu[0] - pu[2 * s]);
  return out * f->cidx1_lo[d];
}

/* FDStencilCenteredDegreeEven<8,4>: {1,-8,28,-56,70} (stencils.tpp:395-401);
   Diss symmetrised branch finite_differencing.tpp:339-350. */
static inline double Diss(const fdop* f, const double* u, int d, double diss) {
  ORACLE_FD_SCOPE;
  const long s = f->s[d];
  const double* pu = u - 4 * s;
  double out = 0.0;
  out += 1.0 * (pu[0] + pu[8 * s]);
  out += -8.0 * (pu[1 * s] + pu[7 * s]);
  out += 28.0 * (pu[2 * s] + pu[6 * s]);
  out += -56.0 * (pu[3 * s] + pu[5 * s]);
  out += 70.0 * pu[4 * s];
  return out * f->cidxd[d] * diss;
}

/* FDStencilBiasedLeft<1,4,2> {1,-8,30,-80,35,24,-2}/60 on offsets -4..+2 and
   FDStencilBiasedRight<1,4,2> {2,-24,-35,80,-30,8,-1}/60 on offsets -2..+4
   (stencils.tpp:476-489); Lx symmetrised branch finite_differencing.tpp:156-174. */
static const double LX_L[7] = {1, -8, 30, -80, 35, 24, -2};
static const double LX_R[7] = {2, -24, -35, 80, -30, 8, -1};

static inline double Lx_left(const fdop* f, const double* u, int d) {
  ORACLE_FD_SCOPE;
  const long s = f->s[d];
  double dl = 0.0;
  for (int n = 0; n < 7; ++n) dl += LX_L[n] * u[(n - 4) * s];
  return dl;
}
static inline double Lx_right(const fdop* f, const double* u, int d) {
  ORACLE_FD_SCOPE;
  const long s = f->s[d];
  double dr = 0.0;
  for (int n = 6; n >= 0; --n) dr += LX_R[n] * u[(n - 2) * s];
  return dr;
}
static inline double Lx(const fdop* f, const double* u, int d, double vx) {
  ORACLE_FD_SCOPE;
  const double dl = Lx_left(f, u, d);
  const double dr = Lx_right(f, u, d);
  return ((vx < 0) ? (vx * dl) : (vx * dr)) * f->lidx[d];
}

double oracle_fd_point(int kind, const oracle_geom* g, const double* u, long p,
                       int dir) {
  fdop f;
  fd_init(g, &f);
  switch (kind) {
    case 0: return Dx(&f, u + p, dir);
    case 1: return Dxx(&f, u + p, dir);
    case 2: return Ds(&f, u + p, dir);
    case 3: return Lx_left(&f, u + p, dir) * f.lidx[dir];
    case 4: return Lx_right(&f, u + p, dir) * f.lidx[dir];
    case 5: return Diss(&f, u + p, dir, 1.0);
    default: return NAN;
  }
}

double oracle_lx_point(const oracle_geom* g, const double* u, long p, int dir,
                       double vx) {
  fdop f;
  fd_init(g, &f);
  return Lx(&f, u + p, dir, vx);
}

/* ------------------------------------------------------------------------ */
/* Linear algebra (src/utils/linear_algebra.hpp:47-118, 271-290) and        */
/* regularisers (src/utils/floating_point.hpp:110-131).                      */
/* ------------------------------------------------------------------------ */
static inline double det3(double gxx, double gxy, double gxz, double gyy,
                          double gyz, double gzz) {
  return (-(gxz * gxz) * gyy + 2 * gxy * gxz * gyz - gxx * (gyz * gyz) -
          (gxy * gxy) * gzz + gxx * gyy * gzz);
}
static inline double det3v(const double* g6) {
  return det3(g6[0], g6[1], g6[2], g6[3], g6[4], g6[5]);
}
static inline void inv3(double oodetg, const double* g6, double* u6) {
  const double gxx = g6[0], gxy = g6[1], gxz = g6[2], gyy = g6[3],
               gyz = g6[4], gzz = g6[5];
  u6[0] = (-(gyz * gyz) + gyy * gzz) * oodetg;
  u6[1] = (gxz * gyz - gxy * gzz) * oodetg;
  u6[3] = (-(gxz * gxz) + gxx * gzz) * oodetg;
  u6[2] = (-gxz * gyy + gxy * gyz) * oodetg;
  u6[4] = (gxy * gxz - gxx * gyz) * oodetg;
  u6[5] = (-(gxy * gxy) + gxx * gyy) * oodetg;
}
static inline double trace_rank2(double detginv, const double* g6,
                                 const double* A6) {
  const double gxx = g6[0], gxy = g6[1], gxz = g6[2], gyy = g6[3],
               gyz = g6[4], gzz = g6[5];
  const double Axx = A6[0], Axy = A6[1], Axz = A6[2], Ayy = A6[3],
               Ayz = A6[4], Azz = A6[5];
  return detginv * (-2. * Ayz * gxx * gyz + Axx * gyy * gzz +
                    gxx * (Azz * gyy + Ayy * gzz) +
                    2. * (gxz * (Ayz * gxy - Axz * gyy + Axy * gyz) +
                          gxy * (Axz * gyz - Axy * gzz)) -
                    Azz * (gxy * gxy) - Ayy * (gxz * gxz) - Axx * (gyz * gyz));
}
static inline double regularize_nth_root(double value, double value_reg,
                                         double eps_floor, double npow) {
  const double eps = value - value_reg;
  return ((eps < eps_floor) ? (1.0 + eps_floor / npow)
                            : pow(value, 1.0 / npow));
}
static inline double regularize_near_zero(double value, double value_eps) {
  return (value >= 0) ? ((value < value_eps) ? value_eps : value)
                      : ((value > -value_eps) ? -value_eps : value);
}
/* std::max(a, b) == (a < b) ? b : a */
static inline double stdmax(double a, double b) { return (a < b) ? b : a; }

/* ------------------------------------------------------------------------ */
/* PrepareZ4cDerivatives (src/z4c/adm_z4c.cpp:972-1120)                       */
/* ------------------------------------------------------------------------ */
void oracle_prep_derivs(const oracle_params* o, const oracle_geom* g,
                        const double* u, double* d1, double* ddchi,
                        double* aux) {
  fdop f;
  fd_init(g, &f);
  const long N = (long)g->nn1 * g->nn2 * g->nn3;
  const int ng_ext = 4 - 1;
  for (int b = 0; b < 3; ++b) {
    const int klo = (b == 2) ? g->kl : g->kl - ng_ext;
    const int khi = (b == 2) ? g->ku : g->ku + ng_ext;
    const int jlo = (b == 1) ? g->jl : g->jl - ng_ext;
    const int jhi = (b == 1) ? g->ju : g->ju + ng_ext;
    const int ilo = (b == 0) ? g->il : g->il - ng_ext;
    const int ihi = (b == 0) ? g->iu : g->iu + ng_ext;
    for (int k = klo; k <= khi; ++k)
      for (int j = jlo; j <= jhi; ++j)
        for (int i = ilo; i <= ihi; ++i) {
          const long p = i + (long)g->nn1 * (j + (long)g->nn2 * k);
          d1[(b * 11 + 0) * N + p] = Dx(&f, u + OR_ALPHA * N + p, b);
          d1[(b * 11 + 1) * N + p] = Dx(&f, u + OR_CHI * N + p, b);
          for (int c = 0; c < 3; ++c)
            d1[(b * 11 + 2 + c) * N + p] =
                Dx(&f, u + (OR_BETAX + c) * N + p, b);
          for (int c = 0; c < 6; ++c)
            d1[(b * 11 + 5 + c) * N + p] = Dx(&f, u + (OR_GXX + c) * N + p, b);
        }
  }
  for (int k = g->kl; k <= g->ku; ++k)
    for (int j = g->jl; j <= g->ju; ++j)
      for (int i = g->il; i <= g->iu; ++i) {
        const long p = i + (long)g->nn1 * (j + (long)g->nn2 * k);
        for (int a = 0; a < 3; ++a) {
          ddchi[S[a][a] * N + p] = Dxx(&f, u + OR_CHI * N + p, a);
          for (int b = a + 1; b < 3; ++b)
            ddchi[S[a][b] * N + p] = Dx(&f, d1 + (b * 11 + 1) * N + p, a);
        }
      }
  if (!aux) return;
  const double pw = o->chi_psi_power;
  for (int k = g->kl; k <= g->ku; ++k)
    for (int j = g->jl; j <= g->ju; ++j)
      for (int i = g->il; i <= g->iu; ++i) {
        const long p = i + (long)g->nn1 * (j + (long)g->nn2 * k);
        const double chi_g = stdmax(u[OR_CHI * N + p], o->chi_div_floor);
        const double oochi = 1.0 / chi_g;
        double chi_pow = 0.0;
        const double fac = 4.0 / pw;
        if (pw != -4.) chi_pow = pow(chi_g, fac);
        for (int c = 0; c < 3; ++c) {
          const double dchi_c = d1[(c * 11 + 1) * N + p];
          for (int ab = 0; ab < 6; ++ab) {
            const double dg = d1[(c * 11 + 5 + ab) * N + p];
            const double gab = u[(OR_GXX + ab) * N + p];
            double v;
            if (pw == -4.)
              v = oochi * (dg - oochi * gab * dchi_c);
            else
              v = chi_pow * (dg + fac * oochi * gab * dchi_c);
            aux[(12 + c * 6 + ab) * N + p] = v;
          }
        }
        for (int c = 0; c < 3; ++c)
          for (int a = 0; a < 3; ++a)
            aux[(3 + c * 3 + a) * N + p] = d1[(c * 11 + 2 + a) * N + p];
        for (int c = 0; c < 3; ++c)
          aux[(0 + c) * N + p] = d1[(c * 11 + 0) * N + p];
      }
}

/* ------------------------------------------------------------------------ */
/* Z4cRHS (src/z4c/calculate_z4c_rhs.cpp:29-916), vacuum build, default    */
/* constant shift damping (no Z4C_ETA_CONF / Z4C_ETA_TRACK_TP).              */
/* ------------------------------------------------------------------------ */
static void rhs_point(const oracle_params* o, const fdop* f, long N, long p,
                      const double* u, const double* mat, const double* d1,
                      const double* ddchi3, double* rhs) {
  const double pw = o->chi_psi_power;
  const double oochi_psi_power = 1.0 / pw;
  const double chi_pow_DA_u = -6.0 / pw;
#define U(n) (u + (long)(n) * N + p)
#define D1(b, fld) (d1 + ((long)(b) * 11 + (fld)) * N + p)

  const double alpha = *U(OR_ALPHA), chi = *U(OR_CHI), Khat = *U(OR_KHAT),
               Theta = *U(OR_THETA);
  double beta[3], Gam[3], gdd[6], Add[6];
  for (int a = 0; a < 3; ++a) {
    beta[a] = *U(OR_BETAX + a);
    Gam[a] = *U(OR_GAMX + a);
  }
  for (int ab = 0; ab < 6; ++ab) {
    gdd[ab] = *U(OR_GXX + ab);
    Add[ab] = *U(OR_AXX + ab);
  }
  double rho = 0.0, Sd[3] = {0.0, 0.0, 0.0}, Sdd[6] = {0, 0, 0, 0, 0, 0};
  if (mat) {
    rho = mat[OR_MAT_RHO * N + p];
    for (int a = 0; a < 3; ++a) Sd[a] = mat[(OR_MAT_SX + a) * N + p];
    for (int ab = 0; ab < 6; ++ab) Sdd[ab] = mat[(OR_MAT_SXX + ab) * N + p];
  }

  /* 1st derivatives (:84-114); dA_ddd of the reference is never read. */
  double dalpha[3], dchi[3], dKhat[3], dTheta[3], dbeta[3][3], dGam[3][3],
      dg[3][6];
  for (int a = 0; a < 3; ++a) {
    dalpha[a] = *D1(a, 0);
    dchi[a] = *D1(a, 1);
    dKhat[a] = Dx(f, U(OR_KHAT), a);
    dTheta[a] = Dx(f, U(OR_THETA), a);
  }
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      dbeta[b][a] = *D1(b, 2 + a);
      dGam[b][a] = Dx(f, U(OR_GAMX + a), b);
    }
  for (int c = 0; c < 3; ++c)
    for (int ab = 0; ab < 6; ++ab) dg[c][ab] = *D1(c, 5 + ab);

  /* 2nd derivatives (:121-178): diagonal Dxx, mixed as nested Dx. */
  double ddalpha[6], ddchi[6], ddbeta[6][3], ddg[6][6];
  for (int a = 0; a < 3; ++a) {
    ddalpha[S[a][a]] = Dxx(f, U(OR_ALPHA), a);
    ddchi[S[a][a]] = ddchi3[S[a][a] * N + p];
    for (int b = a + 1; b < 3; ++b) {
      ddalpha[S[a][b]] = Dx(f, D1(b, 0), a);
      ddchi[S[a][b]] = ddchi3[S[a][b] * N + p];
    }
  }
  for (int c = 0; c < 3; ++c)
    for (int a = 0; a < 3; ++a)
      for (int b = a; b < 3; ++b)
        ddbeta[S[a][b]][c] = (a == b) ? Dxx(f, U(OR_BETAX + c), a)
                                      : Dx(f, D1(b, 2 + c), a);
  for (int cd = 0; cd < 6; ++cd)
    for (int a = 0; a < 3; ++a)
      for (int b = a; b < 3; ++b)
        ddg[S[a][b]][cd] = (a == b) ? Dxx(f, U(OR_GXX + cd), a)
                                    : Dx(f, D1(b, 5 + cd), a);

  /* Advective derivatives (:184-226). */
  double Lalpha = 0.0, Lchi = 0.0, LKhat = 0.0, LTheta = 0.0;
  for (int a = 0; a < 3; ++a) {
    Lalpha += Lx(f, U(OR_ALPHA), a, beta[a]);
    Lchi += Lx(f, U(OR_CHI), a, beta[a]);
    LKhat += Lx(f, U(OR_KHAT), a, beta[a]);
    LTheta += Lx(f, U(OR_THETA), a, beta[a]);
  }
  double Lbeta[3] = {0, 0, 0}, LGam[3] = {0, 0, 0};
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      Lbeta[b] += Lx(f, U(OR_BETAX + b), a, beta[a]);
      LGam[b] += Lx(f, U(OR_GAMX + b), a, beta[a]);
    }
  double Lg[6] = {0, 0, 0, 0, 0, 0}, LA[6] = {0, 0, 0, 0, 0, 0};
  for (int ab = 0; ab < 6; ++ab)
    for (int c = 0; c < 3; ++c) {
      Lg[ab] += Lx(f, U(OR_GXX + ab), c, beta[c]);
      LA[ab] += Lx(f, U(OR_AXX + ab), c, beta[c]);
    }

  const double K = Khat + 2. * Theta;
  const double chi_guarded = stdmax(chi, o->chi_div_floor);
  const double oochi = 1.0 / chi_guarded;
  const double oopsi4 = (pw == -4.) ? chi_guarded : pow(chi_guarded, -4. / pw);

  /* Inverse conformal metric, vacuum branch (:292-310). */
  double guu[6];
  const double detg = det3v(gdd);
  inv3(1.0 / detg, gdd, guu);

  /* Christoffels (:316-352). */
  double Gddd[3][6], Gudd[3][6], Gu[3];
  for (int c = 0; c < 3; ++c)
    for (int a = 0; a < 3; ++a)
      for (int b = a; b < 3; ++b)
        Gddd[c][S[a][b]] =
            0.5 * (dg[a][S[b][c]] + dg[b][S[a][c]] - dg[c][S[a][b]]);
  for (int c = 0; c < 3; ++c)
    for (int a = 0; a < 3; ++a)
      for (int b = a; b < 3; ++b) {
        double t = 0.0;
        for (int d = 0; d < 3; ++d) t += guu[S[c][d]] * Gddd[d][S[a][b]];
        Gudd[c][S[a][b]] = t;
      }
  for (int a = 0; a < 3; ++a) {
    double t = 0.0;
    for (int b = 0; b < 3; ++b)
      for (int c = b; c < 3; ++c)
        t += ((b == c) ? 1.0 : 2.0) * guu[S[b][c]] * Gudd[a][S[b][c]];
    Gu[a] = t;
  }

  /* Conformal Ricci (:357-396). */
  double Rdd[6];
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b) {
      double r = 0.0;
      for (int c = 0; c < 3; ++c)
        r += 0.5 * (gdd[S[c][a]] * dGam[b][c] + gdd[S[c][b]] * dGam[a][c] +
                    Gu[c] * (Gddd[a][S[b][c]] + Gddd[b][S[a][c]]));
      for (int c = 0; c < 3; ++c)
        for (int d = c; d < 3; ++d)
          r -= ((c == d) ? 0.5 : 1.0) * guu[S[c][d]] * ddg[S[c][d]][S[a][b]];
      for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d)
          for (int e = 0; e < 3; ++e)
            r += guu[S[c][d]] * (Gudd[e][S[c][a]] * Gddd[b][S[e][d]] +
                                 Gudd[e][S[c][b]] * Gddd[a][S[e][d]] +
                                 Gudd[e][S[a][d]] * Gddd[e][S[c][b]]);
      Rdd[S[a][b]] = r;
    }

  /* Conformal factor derivatives (:402-425). */
  double dphi[3], Ddphi[6];
  for (int a = 0; a < 3; ++a) dphi[a] = dchi[a] * oochi * oochi_psi_power;
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b) {
      double t = ddchi[S[a][b]] * oochi * oochi_psi_power -
                 pw * dphi[a] * dphi[b];
      for (int c = 0; c < 3; ++c) t -= Gudd[c][S[a][b]] * dphi[c];
      Ddphi[S[a][b]] = t;
    }
  /* Rphi (:434-458) */
  double trphi = 0.0;
  for (int c = 0; c < 3; ++c)
    for (int d = c; d < 3; ++d)
      trphi += ((c == d) ? 1.0 : 2.0) * guu[S[c][d]] *
               (Ddphi[S[c][d]] + 2. * dphi[c] * dphi[d]);
  double Rphi[6];
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b)
      Rphi[S[a][b]] = 4. * dphi[a] * dphi[b] - 2. * Ddphi[S[a][b]] -
                      2. * gdd[S[a][b]] * trphi;

  /* Matter trace (:466-476). */
  double Str = 0.0;
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b)
      Str += ((a == b) ? 1.0 : 2.0) * oopsi4 * guu[S[a][b]] * Sdd[S[a][b]];

  /* D_a D_b alpha (:484-532). */
  double Dda = 0.0;
  for (int c = 0; c < 3; ++c)
    for (int d = 0; d < 3; ++d) Dda += guu[S[c][d]] * dphi[c] * dalpha[d];
  double Ddalpha_dd[6];
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b) {
      double t = ddalpha[S[a][b]] -
                 2. * (dphi[a] * dalpha[b] + dphi[b] * dalpha[a]) +
                 2. * gdd[S[a][b]] * Dda;
      for (int c = 0; c < 3; ++c) t -= Gudd[c][S[a][b]] * dalpha[c];
      Ddalpha_dd[S[a][b]] = t;
    }
  double Ddalpha = 0.0;
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b)
      Ddalpha += ((a == b) ? 1.0 : 2.0) * oopsi4 * guu[S[a][b]] *
                 Ddalpha_dd[S[a][b]];

  /* A contractions (:536-601). */
  double Aud[3][3];
  for (int c = 0; c < 3; ++c)
    for (int b = 0; b < 3; ++b) {
      double t = 0.0;
      for (int d = 0; d < 3; ++d) t += guu[S[c][d]] * Add[S[d][b]];
      Aud[c][b] = t;
    }
  double AA[6];
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b) {
      double t = 0.0;
      for (int c = 0; c < 3; ++c) t += Add[S[a][c]] * Aud[c][b];
      AA[S[a][b]] = t;
    }
  double trAA = 0.0;
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b)
      trAA += ((a == b) ? 1.0 : 2.0) * guu[S[a][b]] * AA[S[a][b]];
  double Auu[6];
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b) {
      double t = 0.0;
      for (int c = 0; c < 3; ++c) t += guu[S[b][c]] * Aud[a][c];
      Auu[S[a][b]] = t;
    }
  double DA[3];
  for (int a = 0; a < 3; ++a) {
    double t = 0.0;
    for (int b = 0; b < 3; ++b) {
      t -= chi_pow_DA_u * Auu[S[a][b]] * dchi[b] * oochi;
      t -= (1. / 3.) * guu[S[a][b]] * (2. * dKhat[b] + dTheta[b]);
    }
    for (int b = 0; b < 3; ++b)
      for (int c = b; c < 3; ++c)
        t += ((b == c) ? 1.0 : 2.0) * Gudd[a][S[b][c]] * Auu[S[b][c]];
    DA[a] = t;
  }

  /* Ricci scalar, H~ (:606-624). */
  double R = 0.0;
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b)
      R += ((a == b) ? 1.0 : 2.0) * oopsi4 * guu[S[a][b]] *
           (Rdd[S[a][b]] + Rphi[S[a][b]]);
  const double Ht = R + (2. / 3.) * (K * K) - trAA;

  /* Lie derivative completion (:630-687). */
  double dbeta_div = 0.0;
  for (int a = 0; a < 3; ++a) dbeta_div += dbeta[a][a];
  double ddbeta_d[3];
  for (int a = 0; a < 3; ++a) {
    double t = 0.0;
    for (int b = 0; b < 3; ++b) t += (1. / 3.) * ddbeta[S[a][b]][b];
    ddbeta_d[a] = t;
  }
  Lchi += (1. / 6.) * pw * chi_guarded * dbeta_div;
  for (int a = 0; a < 3; ++a) {
    LGam[a] += (2. / 3.) * Gu[a] * dbeta_div;
    for (int b = 0; b < 3; ++b) {
      LGam[a] += guu[S[a][b]] * ddbeta_d[b] - Gu[b] * dbeta[b][a];
      for (int c = b; c < 3; ++c)
        LGam[a] += ((b == c) ? 1.0 : 2.0) * guu[S[b][c]] * ddbeta[S[b][c]][a];
    }
  }
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b) {
      const int ab = S[a][b];
      Lg[ab] -= (2. / 3.) * gdd[ab] * dbeta_div;
      LA[ab] -= (2. / 3.) * Add[ab] * dbeta_div;
      for (int c = 0; c < 3; ++c) {
        Lg[ab] += dbeta[a][c] * gdd[S[b][c]];
        LA[ab] += dbeta[b][c] * Add[S[a][c]];
        Lg[ab] += dbeta[b][c] * gdd[S[a][c]];
        LA[ab] += dbeta[a][c] * Add[S[b][c]];
      }
    }

  /* Assemble (:694-811). */
  const double k1 = o->damp_kappa1, k2 = o->damp_kappa2;
  double r_Khat = -Ddalpha + alpha * (trAA + (1. / 3.) * (K * K)) + LKhat +
                  k1 * (1 - k2) * alpha * Theta;
  r_Khat += 4 * M_PI * alpha * (Str + rho);
  const double r_chi =
      Lchi - (1. / 6.) * pw * chi_guarded * alpha * K;
  double r_Theta = LTheta + alpha * (0.5 * Ht - (2. + k2) * k1 * Theta);
  r_Theta -= 8. * M_PI * alpha * rho;
  double r_Gam[3];
  for (int a = 0; a < 3; ++a) {
    double t = 2. * alpha * DA[a] + LGam[a];
    t -= 2. * alpha * k1 * (Gam[a] - Gu[a]);
    for (int b = 0; b < 3; ++b) {
      t -= 2. * Auu[S[a][b]] * dalpha[b];
      t -= 16. * M_PI * alpha * guu[S[a][b]] * Sd[b];
    }
    r_Gam[a] = t;
  }
  double r_g[6], r_A[6];
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b) {
      const int ab = S[a][b];
      r_g[ab] = -2. * alpha * Add[ab] + Lg[ab];
      double t = oopsi4 * (-Ddalpha_dd[ab] + alpha * (Rdd[ab] + Rphi[ab]));
      t -= (1. / 3.) * gdd[ab] * (-Ddalpha + alpha * R);
      t += alpha * (K * Add[ab] - 2. * AA[ab]);
      t += LA[ab];
      t -= 8. * M_PI * alpha * (oopsi4 * Sdd[ab] - (1. / 3.) * Str * gdd[ab]);
      r_A[ab] = t;
    }
  const double fl = o->lapse_oplog * o->lapse_harmonicf + o->lapse_harmonic * alpha;
  const double r_alpha = (o->lapse_advect * Lalpha -
                          fl * alpha * (Khat + 2 * o->lapse_K * Theta));
  double r_beta[3];
  for (int a = 0; a < 3; ++a)
    r_beta[a] = o->shift_Gamma * Gam[a] + o->shift_advect * Lbeta[a];
  if (fabs(o->shift_alpha2Gamma) > 0.0)
    for (int a = 0; a < 3; ++a)
      r_beta[a] += o->shift_alpha2Gamma * (alpha * alpha) * Gam[a];
  if (fabs(o->shift_H) > 0.0)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        r_beta[a] += o->shift_H * alpha * chi_guarded *
                     (0.5 * alpha * dchi[b] - dalpha[b]) * guu[S[a][b]];
  for (int a = 0; a < 3; ++a) r_beta[a] -= o->shift_eta * beta[a];

  rhs[OR_CHI * N + p] = r_chi;
  rhs[OR_KHAT * N + p] = r_Khat;
  rhs[OR_THETA * N + p] = r_Theta;
  rhs[OR_ALPHA * N + p] = r_alpha;
  for (int a = 0; a < 3; ++a) {
    rhs[(OR_GAMX + a) * N + p] = r_Gam[a];
    rhs[(OR_BETAX + a) * N + p] = r_beta[a];
  }
  for (int ab = 0; ab < 6; ++ab) {
    rhs[(OR_GXX + ab) * N + p] = r_g[ab];
    rhs[(OR_AXX + ab) * N + p] = r_A[ab];
  }
#undef U
#undef D1
}

void oracle_rhs(const oracle_params* o, const oracle_geom* g, const double* u,
                const double* mat, const double* d1, const double* ddchi,
                double* rhs) {
  const long N = (long)g->nn1 * g->nn2 * g->nn3;
  if (o->cowling) {
    memset(rhs, 0, sizeof(double) * N * OR_NZ4C);
    return;
  }
  fdop f;
  fd_init(g, &f);
  for (int k = g->kl; k <= g->ku; ++k)
    for (int j = g->jl; j <= g->ju; ++j)
      for (int i = g->il; i <= g->iu; ++i) {
        const long p = i + (long)g->nn1 * (j + (long)g->nn2 * k);
        rhs_point(o, &f, N, p, u, mat, d1, ddchi, rhs);
      }
  /* KO dissipation (:908-915). */
  for (int n = 0; n < OR_NZ4C; ++n)
    for (int k = g->kl; k <= g->ku; ++k)
      for (int j = g->jl; j <= g->ju; ++j)
        for (int i = g->il; i <= g->iu; ++i) {
          const long p = i + (long)g->nn1 * (j + (long)g->nn2 * k);
          for (int a = 0; a < 3; ++a)
            rhs[n * N + p] += Diss(&f, u + n * N + p, a, o->diss);
        }
}

/* ------------------------------------------------------------------------ */
/* Z4cBoundaryRHS / Z4cSommerfeld_ (src/z4c/boundary_conditions.cpp:16-174)   */
/* ------------------------------------------------------------------------ */
static void sommerfeld_range(const oracle_geom* g, const fdop* f,
                             const double* u, double* rhs, int is, int ie,
                             int js, int je, int ks, int ke) {
  const long N = (long)g->nn1 * g->nn2 * g->nn3;
  const double sq2 = sqrt(2.);
  for (int k = ks; k <= ke; ++k)
    for (int j = js; j <= je; ++j)
      for (int i = is; i <= ie; ++i) {
        const long p = i + (long)g->nn1 * (j + (long)g->nn2 * k);
        double dKhat[3], dTheta[3], dGam[3][3], dA[3][6];
        for (int a = 0; a < 3; ++a) {
          dKhat[a] = Ds(f, u + OR_KHAT * N + p, a);
          dTheta[a] = Ds(f, u + OR_THETA * N + p, a);
        }
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            dGam[b][a] = Ds(f, u + (OR_GAMX + a) * N + p, b);
        for (int ab = 0; ab < 6; ++ab)
          for (int c = 0; c < 3; ++c)
            dA[c][ab] = Ds(f, u + (OR_AXX + ab) * N + p, c);
        const double x1 = g->x0[0] + i * g->dx[0];
        const double x2 = g->x0[1] + j * g->dx[1];
        const double x3 = g->x0[2] + k * g->dx[2];
        const double r = sqrt(x1 * x1 + x2 * x2 + x3 * x3);
        const double su[3] = {x1 / r, x2 / r, x3 / r};
        double th = -u[OR_THETA * N + p] / r;
        double kh = -sq2 * u[OR_KHAT * N + p] / r;
        for (int a = 0; a < 3; ++a) {
          th -= su[a] * dTheta[a];
          kh -= sq2 * su[a] * dKhat[a];
        }
        rhs[OR_THETA * N + p] = th;
        rhs[OR_KHAT * N + p] = kh;
        for (int a = 0; a < 3; ++a) {
          double t = -u[(OR_GAMX + a) * N + p] / r;
          for (int b = 0; b < 3; ++b) t -= su[b] * dGam[b][a];
          rhs[(OR_GAMX + a) * N + p] = t;
        }
        for (int ab = 0; ab < 6; ++ab) {
          double t = -u[(OR_AXX + ab) * N + p] / r;
          for (int c = 0; c < 3; ++c) t -= su[c] * dA[c][ab];
          rhs[(OR_AXX + ab) * N + p] = t;
        }
      }
}

void oracle_boundary_rhs(const oracle_params* o, const oracle_geom* g,
                         const double* u, const int* faces, double* rhs) {
  (void)o;
  fdop f;
  fd_init(g, &f);
  if (faces[0]) sommerfeld_range(g, &f, u, rhs, g->il, g->il, g->jl, g->ju, g->kl, g->ku);
  if (faces[1]) sommerfeld_range(g, &f, u, rhs, g->iu, g->iu, g->jl, g->ju, g->kl, g->ku);
  if (faces[2]) sommerfeld_range(g, &f, u, rhs, g->il, g->iu, g->jl, g->jl, g->kl, g->ku);
  if (faces[3]) sommerfeld_range(g, &f, u, rhs, g->il, g->iu, g->ju, g->ju, g->kl, g->ku);
  if (faces[4]) sommerfeld_range(g, &f, u, rhs, g->il, g->iu, g->jl, g->ju, g->kl, g->kl);
  if (faces[5]) sommerfeld_range(g, &f, u, rhs, g->il, g->iu, g->jl, g->ju, g->ku, g->ku);
}

/* ------------------------------------------------------------------------ */
/* AlgConstr (src/z4c/z4c.cpp:677-799 whole block; :801-991 ranged)           */
/* ------------------------------------------------------------------------ */
static void algconstr_point(const oracle_params* o, long N, long p, double* u) {
  double gdd[6], Add[6];
  for (int ab = 0; ab < 6; ++ab) {
    gdd[ab] = u[(OR_GXX + ab) * N + p];
    Add[ab] = u[(OR_AXX + ab) * N + p];
  }
  double detg = det3v(gdd);
  if (o->tf_force_regularization) {
    detg = detg > 0. ? detg : 1.;
  } else if (!(detg > o->eps_floor)) {
    goto regular;
  }
  {
    const double oopsi4 = regularize_nth_root(detg, 1.0, o->eps_floor, -3.0);
    for (int ab = 0; ab < 6; ++ab) gdd[ab] *= oopsi4;
    const double A = trace_rank2(1.0, gdd, Add);
    for (int ab = 0; ab < 6; ++ab) Add[ab] -= (1.0 / 3.0) * A * gdd[ab];
    for (int ab = 0; ab < 6; ++ab) {
      u[(OR_GXX + ab) * N + p] = gdd[ab];
      u[(OR_AXX + ab) * N + p] = Add[ab];
    }
  }
regular:
  if (o->force_regularization) {
    u[OR_CHI * N + p] = stdmax(u[OR_CHI * N + p], o->eps_floor);
    u[OR_ALPHA * N + p] = regularize_near_zero(u[OR_ALPHA * N + p], o->eps_floor);
  }
}

void oracle_algconstr_range(const oracle_params* o, const oracle_geom* g,
                            double* u, int il, int iu, int jl, int ju, int kl,
                            int ku, int skip_physical) {
  const long N = (long)g->nn1 * g->nn2 * g->nn3;
  for (int k = kl; k <= ku; ++k)
    for (int j = jl; j <= ju; ++j) {
      const int sp_kj = skip_physical && g->jl <= j && j <= g->ju &&
                        g->kl <= k && k <= g->ku;
      for (int i = il; i <= iu; ++i) {
        if (sp_kj && g->il <= i && i <= g->iu) continue;
        algconstr_point(o, N, i + (long)g->nn1 * (j + (long)g->nn2 * k), u);
      }
    }
}

void oracle_algconstr(const oracle_params* o, const oracle_geom* g, double* u) {
  oracle_algconstr_range(o, g, u, 0, g->nn1 - 1, 0, g->nn2 - 1, 0, g->nn3 - 1, 0);
}

/* ------------------------------------------------------------------------ */
/* PrepareAuxExtended (src/z4c/adm_z4c.cpp:1139-1231), vacuum build.          */
/* ------------------------------------------------------------------------ */
static void aux_ext_point(const oracle_params* o, long N, long p,
                          const double* u, const double* adm, double* aux_ext) {
  const double chi = fabs(u[OR_CHI * N + p]);
  const double chi_g = stdmax(chi, o->chi_div_floor);
  double sdg;
  if (o->chi_psi_power == -4.)
    sdg = 1.0 / (chi_g * sqrt(chi_g));
  else
    sdg = pow(chi_g, (12.0 / o->chi_psi_power) / 2.0);
  aux_ext[0 * N + p] = sdg;
  double g6[6], u6[6];
  for (int ab = 0; ab < 6; ++ab) g6[ab] = adm[(OR_ADM_GXX + ab) * N + p];
  inv3(1.0 / (sdg * sdg), g6, u6);
  for (int ab = 0; ab < 6; ++ab) aux_ext[(1 + ab) * N + p] = u6[ab];
}

/* ------------------------------------------------------------------------ */
/* ADMToZ4c (src/z4c/adm_z4c.cpp:46-192)                                     */
/* ------------------------------------------------------------------------ */
void oracle_adm_to_z4c(const oracle_params* o, const oracle_geom* g,
                       const double* adm, double* u, double* aux_ext) {
  const long N = (long)g->nn1 * g->nn2 * g->nn3;
  const double pw = o->chi_psi_power;
  for (long p = 0; p < N; ++p) {
    u[OR_ALPHA * N + p] = adm[OR_ADM_ALPHA * N + p];
    for (int a = 0; a < 3; ++a)
      u[(OR_BETAX + a) * N + p] = adm[(OR_ADM_BETAX + a) * N + p];
    double G[6], Kt[6], gt[6];
    for (int ab = 0; ab < 6; ++ab) G[ab] = adm[(OR_ADM_GXX + ab) * N + p];
    double detg = det3v(G), oopsi4;
    if (pw == -4.) {
      oopsi4 = cbrt(1.0 / detg);
      u[OR_CHI * N + p] = oopsi4;
    } else {
      oopsi4 = pow(detg, -1. / 3.);
      u[OR_CHI * N + p] = pow(detg, 1. / 12. * pw);
    }
    for (int ab = 0; ab < 6; ++ab) {
      gt[ab] = oopsi4 * G[ab];
      Kt[ab] = oopsi4 * adm[(OR_ADM_KXX + ab) * N + p];
      u[(OR_GXX + ab) * N + p] = gt[ab];
    }
    detg = det3v(gt);
    const double Khat = trace_rank2(1.0 / detg, gt, Kt);
    u[OR_KHAT * N + p] = Khat;
    for (int ab = 0; ab < 6; ++ab)
      u[(OR_AXX + ab) * N + p] = Kt[ab] - (1. / 3.) * Khat * gt[ab];
  }
  /* Gamma^a = - sum_b Dx_b g~^{ba} on the interior only (:145-176). */
  double* guu = (double*)malloc(sizeof(double) * 6 * N);
  for (long p = 0; p < N; ++p) {
    double gt[6], u6[6];
    for (int ab = 0; ab < 6; ++ab) gt[ab] = u[(OR_GXX + ab) * N + p];
    inv3(1.0 / det3v(gt), gt, u6);
    for (int ab = 0; ab < 6; ++ab) guu[ab * N + p] = u6[ab];
  }
  fdop f;
  fd_init(g, &f);
  for (int a = 0; a < 3; ++a)
    for (long p = 0; p < N; ++p) u[(OR_GAMX + a) * N + p] = 0.0;
  for (int k = g->kl; k <= g->ku; ++k)
    for (int j = g->jl; j <= g->ju; ++j)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          for (int i = g->il; i <= g->iu; ++i) {
            const long p = i + (long)g->nn1 * (j + (long)g->nn2 * k);
            u[(OR_GAMX + a) * N + p] -= Dx(&f, guu + S[b][a] * N + p, b);
          }
  free(guu);
  for (long p = 0; p < N; ++p) u[OR_THETA * N + p] = 0.0;
  oracle_algconstr(o, g, u);
  if (aux_ext)
    for (long p = 0; p < N; ++p) aux_ext_point(o, N, p, u, adm, aux_ext);
}

/* ------------------------------------------------------------------------ */
/* Z4cToADM (src/z4c/adm_z4c.cpp:200-260)                                    */
/* ------------------------------------------------------------------------ */
void oracle_z4c_to_adm(const oracle_params* o, const oracle_geom* g,
                       const double* u, double* adm, double* aux_ext) {
  const long N = (long)g->nn1 * g->nn2 * g->nn3;
  for (long p = 0; p < N; ++p) {
    adm[OR_ADM_ALPHA * N + p] = u[OR_ALPHA * N + p];
    for (int a = 0; a < 3; ++a)
      adm[(OR_ADM_BETAX + a) * N + p] = u[(OR_BETAX + a) * N + p];
    const double chi = u[OR_CHI * N + p];
    const double psi4 = (o->chi_psi_power == -4.)
                            ? 1.0 / chi
                            : pow(chi, 4. / o->chi_psi_power);
    adm[OR_ADM_PSI4 * N + p] = psi4;
    double G[6];
    for (int ab = 0; ab < 6; ++ab) {
      G[ab] = psi4 * u[(OR_GXX + ab) * N + p];
      adm[(OR_ADM_GXX + ab) * N + p] = G[ab];
    }
    const double th = u[OR_THETA * N + p], kh = u[OR_KHAT * N + p];
    for (int ab = 0; ab < 6; ++ab)
      adm[(OR_ADM_KXX + ab) * N + p] =
          psi4 * u[(OR_AXX + ab) * N + p] + (1. / 3.) * (kh + 2. * th) * G[ab];
  }
  if (aux_ext)
    for (long p = 0; p < N; ++p) aux_ext_point(o, N, p, u, adm, aux_ext);
}

/* ------------------------------------------------------------------------ */
/* ADMConstraints (src/z4c/adm_z4c.cpp:382-930)                              */
/* ------------------------------------------------------------------------ */
void oracle_adm_constraints(const oracle_params* o, const oracle_geom* g,
                            const double* u, const double* adm,
                            const double* mat, const double* d1,
                            const double* ddchi3, double* con) {
  const long N = (long)g->nn1 * g->nn2 * g->nn3;
  memset(con, 0, sizeof(double) * OR_NCON * N);
  fdop f;
  fd_init(g, &f);
  const double pw = o->chi_psi_power;
  const double r2_max = o->r_max_con * o->r_max_con;
  for (int k = g->kl; k <= g->ku; ++k)
    for (int j = g->jl; j <= g->ju; ++j)
      for (int i = g->il; i <= g->iu; ++i) {
        const long p = i + (long)g->nn1 * (j + (long)g->nn2 * k);
#define D1(b, fld) (d1 + ((long)(b) * 11 + (fld)) * N + p)
        double ddchi[6], gt[6], G[6], Kdd[6];
        for (int ab = 0; ab < 6; ++ab) {
          ddchi[ab] = ddchi3[ab * N + p];
          gt[ab] = u[(OR_GXX + ab) * N + p];
          G[ab] = adm[(OR_ADM_GXX + ab) * N + p];
          Kdd[ab] = adm[(OR_ADM_KXX + ab) * N + p];
        }
        double rho = 0.0, Sd[3] = {0, 0, 0};
        if (mat) {
          rho = mat[OR_MAT_RHO * N + p];
          for (int a = 0; a < 3; ++a) Sd[a] = mat[(OR_MAT_SX + a) * N + p];
        }
        const double chi_g = stdmax(u[OR_CHI * N + p], o->chi_div_floor);
        const double oochi = 1.0 / chi_g;
        double dg[3][6], dK[3][6], ddg[6][6];
        if (pw == -4.) {
          for (int c = 0; c < 3; ++c)
            for (int ab = 0; ab < 6; ++ab) {
              dg[c][ab] = oochi * (*D1(c, 5 + ab) - oochi * gt[ab] * *D1(c, 1));
              dK[c][ab] = Dx(&f, adm + (OR_ADM_KXX + ab) * N + p, c);
            }
          for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b)
              for (int cd = 0; cd < 6; ++cd) {
                const double dchi_a = *D1(a, 1);
                const double gtil = gt[cd];
                if (a == b) {
                  const double ddg_conf = Dxx(&f, u + (OR_GXX + cd) * N + p, a);
                  ddg[S[a][a]][cd] =
                      oochi * (ddg_conf -
                               oochi * (2.0 * dchi_a * *D1(a, 5 + cd) +
                                        ddchi[S[a][a]] * gtil) +
                               2.0 * oochi * oochi * dchi_a * dchi_a * gtil);
                } else {
                  const double dchi_b = *D1(b, 1);
                  const double ddg_conf = Dx(&f, D1(b, 5 + cd), a);
                  ddg[S[a][b]][cd] =
                      oochi * (ddg_conf -
                               oochi * (dchi_a * *D1(b, 5 + cd) +
                                        dchi_b * *D1(a, 5 + cd) +
                                        ddchi[S[a][b]] * gtil) +
                               2.0 * oochi * oochi * dchi_a * dchi_b * gtil);
                }
              }
        } else {
          const double fac = 4.0 / pw, fac2 = fac - 1.0;
          const double psi4 = pow(chi_g, fac);
          for (int c = 0; c < 3; ++c)
            for (int ab = 0; ab < 6; ++ab) {
              dg[c][ab] = psi4 * (*D1(c, 5 + ab) + fac * oochi * gt[ab] * *D1(c, 1));
              dK[c][ab] = Dx(&f, adm + (OR_ADM_KXX + ab) * N + p, c);
            }
          for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b)
              for (int cd = 0; cd < 6; ++cd) {
                const double dchi_a = *D1(a, 1);
                const double gtil = gt[cd];
                if (a == b) {
                  const double ddg_conf = Dxx(&f, u + (OR_GXX + cd) * N + p, a);
                  ddg[S[a][a]][cd] =
                      psi4 * (ddg_conf +
                              fac * oochi *
                                  (2.0 * dchi_a * *D1(a, 5 + cd) +
                                   ddchi[S[a][a]] * gtil) +
                              fac * fac2 * oochi * oochi * dchi_a * dchi_a * gtil);
                } else {
                  const double dchi_b = *D1(b, 1);
                  const double ddg_conf = Dx(&f, D1(b, 5 + cd), a);
                  ddg[S[a][b]][cd] =
                      psi4 * (ddg_conf +
                              fac * oochi *
                                  (dchi_a * *D1(b, 5 + cd) +
                                   dchi_b * *D1(a, 5 + cd) +
                                   ddchi[S[a][b]] * gtil) +
                              fac * fac2 * oochi * oochi * dchi_a * dchi_b * gtil);
                }
              }
        }
        /* physical inverse metric, Christoffels */
        double guu[6];
        inv3(1. / det3v(G), G, guu);
        double Gddd[3][6], Gudd[3][6];
        for (int c = 0; c < 3; ++c)
          for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b)
              Gddd[c][S[a][b]] =
                  0.5 * (dg[a][S[b][c]] + dg[b][S[a][c]] - dg[c][S[a][b]]);
        for (int c = 0; c < 3; ++c)
          for (int ab = 0; ab < 6; ++ab) {
            double t = 0.0;
            for (int d = 0; d < 3; ++d) t += guu[S[c][d]] * Gddd[d][ab];
            Gudd[c][ab] = t;
          }
        /* Ricci tensor and scalar (:654-693) */
        double R = 0.0;
        for (int a = 0; a < 3; ++a)
          for (int b = a; b < 3; ++b) {
            double r = 0.0;
            for (int c = 0; c < 3; ++c)
              for (int d = 0; d < 3; ++d)
                for (int e = 0; e < 3; ++e)
                  r += guu[S[d][e]] * (Gudd[c][S[a][e]] * Gddd[c][S[b][d]] -
                                       Gudd[c][S[a][b]] * Gddd[c][S[e][d]]);
            for (int c = 0; c < 3; ++c)
              for (int d = 0; d < 3; ++d)
                r += 0.5 * guu[S[c][d]] *
                     (-ddg[S[b][a]][S[c][d]] + ddg[S[d][a]][S[b][c]] +
                      ddg[S[d][b]][S[a][c]] - ddg[S[d][c]][S[a][b]]);
            R += ((a == b) ? 1.0 : 2.0) * guu[S[a][b]] * r;
          }
        /* Extrinsic curvature traces (:697-721) */
        double Kud[3][3], K = 0.0;
        for (int a = 0; a < 3; ++a) {
          for (int b = 0; b < 3; ++b) {
            double t = 0.0;
            for (int c = 0; c < 3; ++c) t += guu[S[a][c]] * Kdd[S[c][b]];
            Kud[a][b] = t;
          }
          K += Kud[a][a];
        }
        double KK = 0.0;
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) KK += Kud[a][b] * Kud[b][a];
        double DKddd[3][6], DKudd[3][6];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            for (int c = b; c < 3; ++c) {
              double t = dK[a][S[b][c]];
              for (int d = 0; d < 3; ++d) {
                t -= Gudd[d][S[a][b]] * Kdd[S[d][c]];
                t -= Gudd[d][S[a][c]] * Kdd[S[b][d]];
              }
              DKddd[a][S[b][c]] = t;
            }
        for (int a = 0; a < 3; ++a)
          for (int bc = 0; bc < 6; ++bc) {
            double t = 0.0;
            for (int d = 0; d < 3; ++d) t += guu[S[a][d]] * DKddd[d][bc];
            DKudd[a][bc] = t;
          }
        const double H = R + (K * K) - KK - 16 * M_PI * rho;
        double Mu[3] = {0, 0, 0};
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) {
            Mu[a] -= 8 * M_PI * guu[S[a][b]] * Sd[b];
            for (int c = 0; c < 3; ++c) {
              Mu[a] += guu[S[a][b]] * DKudd[c][S[b][c]];
              Mu[a] -= guu[S[b][c]] * DKudd[a][S[b][c]];
            }
          }
        double Md[3];
        for (int a = 0; a < 3; ++a) {
          double t = 0;
          for (int b = 0; b < 3; ++b) t += G[S[a][b]] * Mu[b];
          Md[a] = t;
        }
        double M = 0.;
        for (int a = 0; a < 3; ++a)
          for (int b = a; b < 3; ++b)
            M += ((a == b) ? 1.0 : 2.0) * G[S[a][b]] * Mu[a] * Mu[b];
        /* conformal Gamma_u from stored conformal derivatives (:813-876) */
        double cguu[6];
        inv3(1. / det3v(gt), gt, cguu);
        double cGddd[3][6], cGudd[3][6], cGu[3];
        for (int c = 0; c < 3; ++c)
          for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b)
              cGddd[c][S[a][b]] = 0.5 * (*D1(a, 5 + S[b][c]) + *D1(b, 5 + S[a][c]) -
                                         *D1(c, 5 + S[a][b]));
        for (int c = 0; c < 3; ++c)
          for (int ab = 0; ab < 6; ++ab) {
            double t = 0.0;
            for (int d = 0; d < 3; ++d) t += cguu[S[c][d]] * cGddd[d][ab];
            cGudd[c][ab] = t;
          }
        for (int a = 0; a < 3; ++a) {
          double t = 0.0;
          for (int b = 0; b < 3; ++b)
            for (int c = b; c < 3; ++c)
              t += ((b == c) ? 1.0 : 2.0) * cguu[S[b][c]] * cGudd[a][S[b][c]];
          cGu[a] = t;
        }
        double Z = 0.;
        for (int a = 0; a < 3; ++a)
          for (int b = a; b < 3; ++b)
            Z += ((a == b) ? 0.25 : 0.5) * G[S[a][b]] *
                 (u[(OR_GAMX + a) * N + p] - cGu[a]) *
                 (u[(OR_GAMX + b) * N + p] - cGu[b]);
        const double x1 = g->x0[0] + i * g->dx[0];
        const double x2 = g->x0[1] + j * g->dx[1];
        const double x3 = g->x0[2] + k * g->dx[2];
        const double R2 = x1 * x1 + x2 * x2 + x3 * x3;
        const double th = u[OR_THETA * N + p];
        if (R2 > r2_max) {
          /* all zero (already cleared) */
        } else {
          con[OR_CON_H * N + p] = H;
          con[OR_CON_M * N + p] = M;
          con[OR_CON_Z * N + p] = Z;
          for (int a = 0; a < 3; ++a) con[(OR_CON_MX + a) * N + p] = Md[a];
          con[OR_CON_C * N + p] = (H * H) + M + (th * th) + 4.0 * Z;
        }
#undef D1
      }
}

/* ------------------------------------------------------------------------ */
/* Time integrators                                                          */
/* ------------------------------------------------------------------------ */
void oracle_bt_sum(const oracle_geom* g, int nvar, double* u_out,
                   const double* u0, int nk, const double* const* k,
                   const double* w, double dt) {
  const long N = (long)g->nn1 * g->nn2 * g->nn3 * nvar;
  for (long p = 0; p < N; ++p) u_out[p] = u0[p];
  for (int s = 0; s < nk; ++s) {
    if (w[s] != 0) {
      const double a_dt = w[s] * dt;
      for (long p = 0; p < N; ++p) u_out[p] += a_dt * k[s][p];
    }
  }
}

void oracle_weighted_ave(const oracle_geom* g, int nvar, double* u_out,
                         const double* u_in1, const double* u_in2,
                         const double* w) {
  const long N = (long)g->nn1 * g->nn2 * g->nn3;
  for (int n = 0; n < nvar; ++n)
    for (int k = g->kl; k <= g->ku; ++k)
      for (int j = g->jl; j <= g->ju; ++j)
        for (int i = g->il; i <= g->iu; ++i) {
          const long p = n * N + i + (long)g->nn1 * (j + (long)g->nn2 * k);
          if (w[0] == 1.0) {
            if (w[2] != 0.0)
              u_out[p] += w[1] * u_in1[p] + w[2] * u_in2[p];
            else if (w[1] != 0.0)
              u_out[p] += w[1] * u_in1[p];
          } else if (w[0] == 0.0) {
            if (w[2] != 0.0)
              u_out[p] = w[1] * u_in1[p] + w[2] * u_in2[p];
            else if (w[1] == 1.0)
              u_out[p] = u_in1[p];
            else
              u_out[p] = w[1] * u_in1[p];
          } else {
            if (w[2] != 0.0)
              u_out[p] = w[0] * u_out[p] + w[1] * u_in1[p] + w[2] * u_in2[p];
            else if (w[1] != 0.0)
              u_out[p] = w[0] * u_out[p] + w[1] * u_in1[p];
            else
              u_out[p] *= w[0];
          }
        }
}

void oracle_add_rhs(const oracle_geom* g, int nvar, const double* rhs,
                    double wght, double* u_out) {
  const long N = (long)g->nn1 * g->nn2 * g->nn3;
  for (int n = 0; n < nvar; ++n)
    for (int k = g->kl; k <= g->ku; ++k)
      for (int j = g->jl; j <= g->ju; ++j)
        for (int i = g->il; i <= g->iu; ++i) {
          const long p = n * N + i + (long)g->nn1 * (j + (long)g->nn2 * k);
          u_out[p] += wght * rhs[p];
        }
}

double oracle_new_dt(const oracle_geom* g, double cfl) {
  double min_dt = (double)FLT_MAX;
  for (int d = 0; d < 3; ++d) min_dt = (min_dt < g->dx[d]) ? min_dt : g->dx[d];
  return min_dt * cfl;
}

/* ------------------------------------------------------------------------ */
/* Periodic vertex-centred self-exchange of one block: ghosts take the       */
/* periodic image, shared boundary vertices take the mean of their images    */
/* (additive unpack then division by node multiplicity, src/comm/            */
/* comm_channel.cpp:1277-1330, src/comm/node_multiplicity.hpp:1-60,          */
/* ranges src/comm/index_utilities.hpp:116-139).                             */
/* ------------------------------------------------------------------------ */
static int images(int i, int lo, int hi, int* src) {
  const int n = hi - lo;
  if (i < lo) { src[0] = i + n; return 1; }
  if (i > hi) { src[0] = i - n; return 1; }
  if (i == lo) { src[0] = lo; src[1] = hi; return 2; }
  if (i == hi) { src[0] = hi; src[1] = lo; return 2; }
  src[0] = i;
  return 1;
}

void oracle_periodic_fill(const oracle_geom* g, int nvar, double* u) {
  const long N = (long)g->nn1 * g->nn2 * g->nn3;
  double* snap = (double*)malloc(sizeof(double) * N);
  for (int n = 0; n < nvar; ++n) {
    double* un = u + n * N;
    memcpy(snap, un, sizeof(double) * N);
    for (int k = 0; k < g->nn3; ++k)
      for (int j = 0; j < g->nn2; ++j)
        for (int i = 0; i < g->nn1; ++i) {
          int sk[2], sj[2], si[2];
          const int nk = images(k, g->kl, g->ku, sk);
          const int nj = images(j, g->jl, g->ju, sj);
          const int ni = images(i, g->il, g->iu, si);
          double acc = 0.0;
          for (int a = 0; a < nk; ++a)
            for (int b = 0; b < nj; ++b)
              for (int c = 0; c < ni; ++c)
                acc += snap[si[c] + (long)g->nn1 * (sj[b] + (long)g->nn2 * sk[a])];
          un[i + (long)g->nn1 * (j + (long)g->nn2 * k)] = acc / (double)(nk * nj * ni);
        }
  }
  free(snap);
}

/* ------------------------------------------------------------------------ */
/* History sums of the Z4c constraints (src/outputs/history.cpp:178-205).    */
/* The reference loops over cell indices is..ie of the vertex arrays and     */
/* tests the cell-centre radius x1v = x1f + dx/2.                            */
/* ------------------------------------------------------------------------ */
void oracle_history(const oracle_params* o, const oracle_geom* g,
                    const double* u, const double* con, double* out) {
  const long N = (long)g->nn1 * g->nn2 * g->nn3;
  const double r2max = o->r_max_con * o->r_max_con;
  const double vol = g->dx[0] * g->dx[1] * g->dx[2];
  for (int k = g->kl; k < g->ku; ++k)
    for (int j = g->jl; j < g->ju; ++j)
      for (int i = g->il; i < g->iu; ++i) {
        const double x1 = g->x0[0] + (i + 0.5) * g->dx[0];
        const double x2 = g->x0[1] + (j + 0.5) * g->dx[1];
        const double x3 = g->x0[2] + (k + 0.5) * g->dx[2];
        if (x1 * x1 + x2 * x2 + x3 * x3 > r2max) continue;
        const long p = i + (long)g->nn1 * (j + (long)g->nn2 * k);
        const double H = fabs(con[OR_CON_H * N + p]);
        const double Mx = fabs(con[(OR_CON_MX + 0) * N + p]);
        const double My = fabs(con[(OR_CON_MX + 1) * N + p]);
        const double Mz = fabs(con[(OR_CON_MX + 2) * N + p]);
        const double th = fabs(u[OR_THETA * N + p]);
        out[0] += vol * (H * H);
        out[1] += vol * fabs(con[OR_CON_M * N + p]);
        out[2] += vol * (Mx * Mx);
        out[3] += vol * (My * My);
        out[4] += vol * (Mz * Mz);
        out[5] += vol * fabs(con[OR_CON_Z * N + p]);
        out[6] += vol * (th * th);
        out[7] += vol * fabs(con[OR_CON_C * N + p]);
      }
}

/* ------------------------------------------------------------------------ */
/* Same-level vertex-centred exchange between blocks (additive unpack of     */
/* every owner's interior value, then multiplication by 1/multiplicity:     */
/* src/comm/comm_channel.cpp:1277-1330, src/comm/node_multiplicity.cpp:     */
/* 349-420, ranges src/comm/index_utilities.hpp:116-139).                   */
/* Brute force: for every vertex of every block, the owners are the         */
/* interior vertices of all blocks at the same lattice point (periodic axes */
/* wrap); the vertex takes the mean of the owners' values, summed in block  */
/* order and then k, j, i order.  Vertices without an owner are untouched.  */
/* ------------------------------------------------------------------------ */
static long lattice(double x, double lo, double h) { return lround((x - lo) / h); }

static int owners_axis(long W, long Lo, long nx, long Nd, int per, long il, long* out) {
  long d = W - Lo;
  if (per) d = ((d % Nd) + Nd) % Nd;
  int n = 0;
  if (d >= 0 && d <= nx) out[n++] = il + d;
  if (per && d == 0 && nx == Nd) out[n++] = il + Nd; /* block spans the period */
  return n;
}

void oracle_exchange(int nb, const oracle_geom* g, double* const* u, int nvar,
                     const double* lo, const double* hi, const int* periodic) {
  const double h[3] = {g[0].dx[0], g[0].dx[1], g[0].dx[2]};
  long Nd[3];
  for (int d = 0; d < 3; ++d) Nd[d] = lround((hi[d] - lo[d]) / h[d]);
  double** snap = (double**)malloc(sizeof(double*) * nb);
  for (int b = 0; b < nb; ++b) {
    const long N = (long)g[b].nn1 * g[b].nn2 * g[b].nn3;
    snap[b] = (double*)malloc(sizeof(double) * N * nvar);
    memcpy(snap[b], u[b], sizeof(double) * N * nvar);
  }
  for (int b = 0; b < nb; ++b) {
    const oracle_geom* G = &g[b];
    const long N = (long)G->nn1 * G->nn2 * G->nn3;
    const int nn[3] = {G->nn1, G->nn2, G->nn3};
    for (int k = 0; k < nn[2]; ++k)
      for (int j = 0; j < nn[1]; ++j)
        for (int i = 0; i < nn[0]; ++i) {
          const int ijk[3] = {i, j, k};
          long W[3];
          for (int d = 0; d < 3; ++d) {
            W[d] = lattice(G->x0[d] + ijk[d] * h[d], lo[d], h[d]);
            if (periodic[d]) W[d] = ((W[d] % Nd[d]) + Nd[d]) % Nd[d];
          }
          /* owners: all blocks, in block order, then k, j, i */
          double acc[64];
          for (int n = 0; n < nvar && n < 64; ++n) acc[n] = 0.0;
          int cnt = 0;
          for (int c = 0; c < nb; ++c) {
            const oracle_geom* C = &g[c];
            const long il[3] = {C->il, C->jl, C->kl};
            const long nx[3] = {C->iu - C->il, C->ju - C->jl, C->ku - C->kl};
            long o[3][2];
            int no[3];
            for (int d = 0; d < 3; ++d) {
              const long Lo = lattice(C->x0[d] + il[d] * h[d], lo[d], h[d]);
              no[d] = owners_axis(W[d], Lo, nx[d], Nd[d], periodic[d], il[d], o[d]);
            }
            const long NC = (long)C->nn1 * C->nn2 * C->nn3;
            for (int a = 0; a < no[2]; ++a)
              for (int bb = 0; bb < no[1]; ++bb)
                for (int cc = 0; cc < no[0]; ++cc) {
                  const long p = o[0][cc] + (long)C->nn1 * (o[1][bb] + (long)C->nn2 * o[2][a]);
                  for (int n = 0; n < nvar && n < 64; ++n) acc[n] += snap[c][n * NC + p];
                  ++cnt;
                }
          }
          if (cnt == 0) continue;
          const long p = i + (long)nn[0] * (j + (long)nn[1] * k);
          const double inv = 1.0 / (double)cnt;
          for (int n = 0; n < nvar && n < 64; ++n)
            u[b][n * N + p] = (cnt > 1) ? acc[n] * inv : acc[n];
        }
  }
  for (int b = 0; b < nb; ++b) free(snap[b]);
  free(snap);
}

/* ------------------------------------------------------------------------ */
/* Physical ghost fill of a vertex-centred block (comm::ApplyPhysicalBCs,   */
/* src/comm/physical_bcs.cpp:506-676, with ExtrapolateOutflowBC :187-290,   */
/* which gr_sommerfeld and extrapolate_outflow faces of VC data use,        */
/* :439-453).  NEXTRAPOLATE = nghost + 1 = 5 points (configure.py:891-895), */
/* weights InterpolateLagrangeUniformBiasR<5> = {5,-10,10,-5,1}             */
/* (src/utils/interp_univariate_stencils.tpp:26-28).  The reference indexes */
/* with the cell-centred ranges is..ie = ng..ng+nx-1 (meshblock.cpp:874-    */
/* 877) on the vertex arrays, so an outer face fills ie+1..ie+ng (the       */
/* boundary vertex through the second-outermost ghost) and the transverse   */
/* ranges stop at je / ke.  Faces in order x1, x2, x3; x2 faces span the x1 */
/* ghosts, x3 faces the x1 and x2 ghosts.  faces: 6 flags ix1..ox3.         */
/* ------------------------------------------------------------------------ */
static const double EXTRAP5[5] = {5., -10., 10., -5., 1.};

void oracle_physical_bc(const oracle_geom* g, int nvar, double* u, const int* faces) {
  const int ng = g->ng;
  const int is = g->il, ie = g->iu - 1, js = g->jl, je = g->ju - 1, ks = g->kl, ke = g->ku - 1;
  const long s1 = 1, s2 = g->nn1, s3 = (long)g->nn1 * g->nn2;
  const long N = s3 * g->nn3;
  const int bis = is - ng, bie = ie + ng;
  int bjs = faces[2] ? js : js - ng, bje = faces[3] ? je : je + ng;
  const int bks = faces[4] ? ks : ks - ng, bke = faces[5] ? ke : ke + ng;
  for (int n = 0; n < nvar; ++n) {
    double* v = u + n * N;
#define AT(i, j, k) v[(long)(i) * s1 + (long)(j) * s2 + (long)(k) * s3]
    if (faces[0])
      for (int k = bks; k <= bke; ++k)
        for (int j = bjs; j <= bje; ++j)
          for (int i = is - 1; i >= is - ng; --i) {
            double x = 0.0;
            for (int p = 0; p < 5; ++p) x += EXTRAP5[p] * AT(i + p + 1, j, k);
            AT(i, j, k) = x;
          }
    if (faces[1])
      for (int k = bks; k <= bke; ++k)
        for (int j = bjs; j <= bje; ++j)
          for (int i = ie + 1; i <= ie + ng; ++i) {
            double x = 0.0;
            for (int p = 0; p < 5; ++p) x += EXTRAP5[p] * AT(i - p - 1, j, k);
            AT(i, j, k) = x;
          }
    if (faces[2])
      for (int k = bks; k <= bke; ++k)
        for (int j = js - 1; j >= js - ng; --j)
          for (int i = bis; i <= bie; ++i) {
            double x = 0.0;
            for (int p = 0; p < 5; ++p) x += EXTRAP5[p] * AT(i, j + p + 1, k);
            AT(i, j, k) = x;
          }
    if (faces[3])
      for (int k = bks; k <= bke; ++k)
        for (int j = je + 1; j <= je + ng; ++j)
          for (int i = bis; i <= bie; ++i) {
            double x = 0.0;
            for (int p = 0; p < 5; ++p) x += EXTRAP5[p] * AT(i, j - p - 1, k);
            AT(i, j, k) = x;
          }
    {
      const int tjs = js - ng, tje = je + ng;  /* x3 faces span the x2 ghosts */
      if (faces[4])
        for (int k = ks - 1; k >= ks - ng; --k)
          for (int j = tjs; j <= tje; ++j)
            for (int i = bis; i <= bie; ++i) {
              double x = 0.0;
              for (int p = 0; p < 5; ++p) x += EXTRAP5[p] * AT(i, j, k + p + 1);
              AT(i, j, k) = x;
            }
      if (faces[5])
        for (int k = ke + 1; k <= ke + ng; ++k)
          for (int j = tjs; j <= tje; ++j)
            for (int i = bis; i <= bie; ++i) {
              double x = 0.0;
              for (int p = 0; p < 5; ++p) x += EXTRAP5[p] * AT(i, j, k - p - 1);
              AT(i, j, k) = x;
            }
    }
#undef AT
  }
  (void)bjs;
}


/* ------------------------------------------------------------------------ */
/* Z4c::Z4cWeyl (src/z4c/calculate_weyl_scalars.cpp:44-628): Re/Im psi4 at  */
/* the interior vertices from the ADM variables, the Z4c state and the      */
/* prepared conformal derivatives (PrepareZ4cDerivatives, d1 and ddchi of   */
/* oracle_prep_derivs).  out: 2 fields (rpsi4, ipsi4), NaN outside the      */
/* interior (u_weyl.Fill(NAN), :64-65).                                     */
/* ------------------------------------------------------------------------ */
void oracle_weyl(const oracle_params* o, const oracle_geom* g, const double* u,
                 const double* adm, const double* d1, const double* ddchi, double* out) {
  fdop f;
  fd_init(g, &f);
  const long N = (long)g->nn1 * g->nn2 * g->nn3;
  for (long e = 0; e < 2 * N; ++e) out[e] = NAN;
  const double pw = o->chi_psi_power;
  const double FR4 = 0.25;
  for (int k = g->kl; k <= g->ku; ++k)
    for (int j = g->jl; j <= g->ju; ++j)
      for (int i = g->il; i <= g->iu; ++i) {
        const long p = i + (long)g->nn1 * (j + (long)g->nn2 * k);
        const double chi_g = stdmax(u[OR_CHI * N + p], o->chi_div_floor);
        double dchi[3], ddc[6], gt[6], dg3[3][6];
        for (int a = 0; a < 3; ++a) dchi[a] = d1[(a * 11 + 1) * N + p];
        for (int ab = 0; ab < 6; ++ab) {
          ddc[ab] = ddchi[ab * N + p];
          gt[ab] = u[(OR_GXX + ab) * N + p];
        }
        for (int c = 0; c < 3; ++c)
          for (int ab = 0; ab < 6; ++ab) dg3[c][ab] = d1[(c * 11 + 5 + ab) * N + p];
        /* physical metric derivatives by the chain rule (:136-262) */
        double dg[3][6], dK[3][6], ddg[6][6]; /* ddg[sym(a,b)][sym(c,d)] */
        for (int c = 0; c < 3; ++c)
          for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b) {
              const int ab = S[a][b];
              if (pw == -4.) {
                const double oochi = 1.0 / chi_g;
                dg[c][ab] = oochi * (dg3[c][ab] - oochi * gt[ab] * dchi[c]);
              } else {
                const double fac = 4.0 / pw;
                const double psi4 = pow(chi_g, fac);
                const double oochi = 1.0 / chi_g;
                dg[c][ab] = psi4 * (dg3[c][ab] + fac * oochi * gt[ab] * dchi[c]);
              }
              dK[c][ab] = Dx(&f, adm + (OR_ADM_KXX + ab) * N + p, c);
            }
        for (int c = 0; c < 3; ++c)
          for (int d = c; d < 3; ++d)
            for (int a = 0; a < 3; ++a)
              for (int b = a; b < 3; ++b) {
                const int cd = S[c][d];
                const double gtil_cd = gt[cd];
                double v;
                if (pw == -4.) {
                  const double oochi = 1.0 / chi_g;
                  if (a == b) {
                    const double ddg_conf = Dxx(&f, u + (OR_GXX + cd) * N + p, a);
                    v = oochi * (ddg_conf -
                                 oochi * (2.0 * dchi[a] * dg3[a][cd] + ddc[S[a][a]] * gtil_cd) +
                                 2.0 * oochi * oochi * dchi[a] * dchi[a] * gtil_cd);
                  } else {
                    const double ddg_conf = Dx(&f, d1 + (b * 11 + 5 + cd) * N + p, a);
                    v = oochi * (ddg_conf -
                                 oochi * (dchi[a] * dg3[b][cd] + dchi[b] * dg3[a][cd] +
                                          ddc[S[a][b]] * gtil_cd) +
                                 2.0 * oochi * oochi * dchi[a] * dchi[b] * gtil_cd);
                  }
                } else {
                  const double fac = 4.0 / pw, fac2 = fac - 1.0;
                  const double psi4 = pow(chi_g, fac);
                  const double oochi = 1.0 / chi_g;
                  if (a == b) {
                    const double ddg_conf = Dxx(&f, u + (OR_GXX + cd) * N + p, a);
                    v = psi4 * (ddg_conf +
                                fac * oochi * (2.0 * dchi[a] * dg3[a][cd] + ddc[S[a][a]] * gtil_cd) +
                                fac * fac2 * oochi * oochi * dchi[a] * dchi[a] * gtil_cd);
                  } else {
                    const double ddg_conf = Dx(&f, d1 + (b * 11 + 5 + cd) * N + p, a);
                    v = psi4 * (ddg_conf +
                                fac * oochi *
                                    (dchi[a] * dg3[b][cd] + dchi[b] * dg3[a][cd] +
                                     ddc[S[a][b]] * gtil_cd) +
                                fac * fac2 * oochi * oochi * dchi[a] * dchi[b] * gtil_cd);
                  }
                }
                ddg[S[a][b]][cd] = v;
              }
        /* inverse metric, Christoffels (:264-306) */
        double gd[6], Kd[6], gu[6];
        for (int ab = 0; ab < 6; ++ab) {
          gd[ab] = adm[(OR_ADM_GXX + ab) * N + p];
          Kd[ab] = adm[(OR_ADM_KXX + ab) * N + p];
        }
        inv3(1. / det3v(gd), gd, gu);
        double Gd[3][6], Gu[3][6];
        for (int c = 0; c < 3; ++c)
          for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b)
              Gd[c][S[a][b]] = 0.5 * (dg[a][S[b][c]] + dg[b][S[a][c]] - dg[c][S[a][b]]);
        for (int c = 0; c < 3; ++c)
          for (int ab = 0; ab < 6; ++ab) Gu[c][ab] = 0.0;
        for (int c = 0; c < 3; ++c)
          for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b)
              for (int d = 0; d < 3; ++d) Gu[c][S[a][b]] += gu[S[c][d]] * Gd[d][S[a][b]];
        /* Ricci tensor and scalar (:308-347) */
        double Rd[6] = {0, 0, 0, 0, 0, 0}, R = 0.0;
        for (int a = 0; a < 3; ++a)
          for (int b = a; b < 3; ++b) {
            for (int c = 0; c < 3; ++c)
              for (int d = 0; d < 3; ++d) {
                for (int e = 0; e < 3; ++e) {
                  Rd[S[a][b]] += gu[S[c][d]] * Gu[e][S[a][c]] * Gd[e][S[b][d]];
                  Rd[S[a][b]] -= gu[S[c][d]] * Gu[e][S[a][b]] * Gd[e][S[c][d]];
                }
                Rd[S[a][b]] += 0.5 * gu[S[c][d]] *
                               (-ddg[S[c][d]][S[a][b]] - ddg[S[a][b]][S[c][d]] +
                                ddg[S[a][c]][S[b][d]] + ddg[S[b][c]][S[a][d]]);
              }
            R += ((a == b) ? 1.0 : 2.0) * gu[S[a][b]] * Rd[S[a][b]];
          }
        /* K traces and covariant derivative (:349-390) */
        double Kud[3][3], K = 0.0;
        for (int a = 0; a < 3; ++a) {
          for (int b = 0; b < 3; ++b) {
            Kud[a][b] = 0.0;
            for (int c = 0; c < 3; ++c) Kud[a][b] += gu[S[a][c]] * Kd[S[c][b]];
          }
          K += Kud[a][a];
        }
        double DK[3][6]; /* DK[a][sym(b,c)]: the value of the last write, b >= c */
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            for (int c = 0; c < 3; ++c) {
              double t = dK[a][S[b][c]];
              for (int d = 0; d < 3; ++d) {
                t -= Gu[d][S[a][b]] * Kd[S[d][c]];
                t -= Gu[d][S[a][c]] * Kd[S[b][d]];
              }
              DK[a][S[b][c]] = t;
            }
        /* tetrad (:392-512) */
        const double x = g->x0[0] + i * g->dx[0], y = g->x0[1] + j * g->dx[1],
                     z = g->x0[2] + k * g->dx[2];
        double xx = x;
        if (x * x + y * y < 1e-10) xx = xx + 1e-8;
        double uv[3] = {xx, y, z}, vv[3] = {xx * z, y * z, -(xx * xx) - y * y},
               wv[3] = {y * -1.0, xx, 0.0};
        double dp = 0.0;
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) dp += gd[S[a][b]] * wv[a] * wv[b];
        for (int a = 0; a < 3; ++a) wv[a] = wv[a] / sqrt(dp);
        dp = 0.0;
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) dp += gd[S[a][b]] * wv[a] * uv[b];
        for (int a = 0; a < 3; ++a) uv[a] -= dp * wv[a];
        dp = 0.0;
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) dp += gd[S[a][b]] * uv[a] * uv[b];
        for (int a = 0; a < 3; ++a) uv[a] = uv[a] / sqrt(dp);
        dp = 0.0;
        double dp2 = 0.0;
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) dp += gd[S[a][b]] * wv[a] * vv[b];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) dp2 += gd[S[a][b]] * uv[a] * vv[b];
        for (int a = 0; a < 3; ++a) vv[a] -= dp * wv[a] + dp2 * uv[a];
        dp = 0.0;
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) dp += gd[S[a][b]] * vv[a] * vv[b];
        for (int a = 0; a < 3; ++a) vv[a] = vv[a] / sqrt(dp);
        /* psi4 (:514-628) */
        double rp = 0.0, ip = 0.0;
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) {
            double R4 = Rd[S[a][b]] + K * Kd[S[a][b]];
            for (int c = 0; c < 3; ++c)
              for (int d = 0; d < 3; ++d) R4 -= gu[S[c][d]] * Kd[S[a][c]] * Kd[S[d][b]];
            const double Tr = vv[a] * vv[b] - wv[a] * wv[b];
            const double Ti = -vv[a] * wv[b] - wv[a] * vv[b];
            rp += -FR4 * R4 * Tr;
            ip += -FR4 * R4 * Ti;
          }
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            for (int c = 0; c < 3; ++c) {
              const double R4 = -(DK[b][S[a][c]] - DK[c][S[a][b]]);
              const double Tr = vv[a] * vv[b] - wv[a] * wv[b];
              const double Ti = -vv[a] * wv[b] - wv[a] * vv[b];
              rp += 0.5 * R4 * uv[c] * Tr;
              ip += 0.5 * R4 * uv[c] * Ti;
            }
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            for (int c = 0; c < 3; ++c)
              for (int d = 0; d < 3; ++d) {
                const double R4 = gd[S[d][c]] * Rd[S[a][b]] + gd[S[a][b]] * Rd[S[d][c]] -
                                  gd[S[d][b]] * Rd[S[a][c]] - gd[S[a][c]] * Rd[S[d][b]] -
                                  0.5 * R * (gd[S[d][c]] * gd[S[a][b]] - gd[S[d][b]] * gd[S[a][c]]) +
                                  Kd[S[d][c]] * Kd[S[a][b]] - Kd[S[d][b]] * Kd[S[a][c]];
                const double Tr = vv[a] * vv[b] - wv[a] * wv[b];
                const double Ti = -vv[a] * wv[b] - wv[a] * vv[b];
                rp += -FR4 * R4 * uv[d] * uv[c] * Tr;
                ip += -FR4 * R4 * uv[d] * uv[c] * Ti;
              }
        out[p] = rp;
        out[N + p] = ip;
      }
}


/* ------------------------------------------------------------------------ */
/* VC restriction and prolongation (src/mesh/mesh_refinement_vc.cpp), 3-D,  */
/* NGHOST = 4, NCGHOST = 4 (configure.py:876-881 default for nghost 4).     */
/* Coarse buffer per axis: nx/2 + 2*NCGHOST + 1 vertices, coarse interior   */
/* civs = NCGHOST .. cive = civs + nx/2, midpoint cimp = civs + (nx/2)/2    */
/* (meshblock.cpp:611-640, 851-895).  Fine indices ivs = NGHOST, ive.       */
/* ------------------------------------------------------------------------ */
#define OR_NCG 4
#define OR_HNCG ((2 * OR_NCG - 4) / 2) /* H_NCGHOST, :305 */
#define OR_HSZ (OR_HNCG + 1)           /* H_SZ, :307 */
/* InterpolateLagrangeUniform<3>::coeff (interp_univariate_stencils.tpp:99-101) */
static const double LAG3[3] = {3. / 256., -25. / 256., 75. / 128.};

/* RestrictVertexCenteredIndicialHelper (:24-45) */
static int restrict_fine_index(int c, int cvs, int cve, int vs, int ve) {
  if (c < cvs) return vs - 2 * (cvs - c);
  if (c > cve) return ve + 2 * (c - cve);
  return 2 * (c - cvs) + vs;
}

/* RestrictVertexCenteredValues (:56-100) over the coarse interior, as
   RestrictInterior calls it (comm/refinement_ops.cpp:178-205). */
void oracle_restrict_vc(const oracle_geom* g, const oracle_geom* cg, int nvar,
                        const double* fine, double* coarse) {
  const long N = (long)g->nn1 * g->nn2 * g->nn3, CN = (long)cg->nn1 * cg->nn2 * cg->nn3;
  for (int n = 0; n < nvar; ++n)
    for (int ck = cg->kl; ck <= cg->ku; ++ck) {
      const int k = restrict_fine_index(ck, cg->kl, cg->ku, g->kl, g->ku);
      for (int cj = cg->jl; cj <= cg->ju; ++cj) {
        const int j = restrict_fine_index(cj, cg->jl, cg->ju, g->jl, g->ju);
        for (int ci = cg->il; ci <= cg->iu; ++ci) {
          const int i = restrict_fine_index(ci, cg->il, cg->iu, g->il, g->iu);
          coarse[n * CN + ci + (long)cg->nn1 * (cj + (long)cg->nn2 * ck)] =
              fine[n * N + i + (long)g->nn1 * (j + (long)g->nn2 * k)];
        }
      }
    }
}

/* ProlongateVertexCenteredValues (:283-772), 3-D branch, even NGHOST
   (eo_offset = 0, injection ranges = si..ei), DBG_SYMMETRIZE_P_OP and
   DBG_SYMMETRIZE_CHEAP on (defs.hpp.in:17,20): the corner sums are
   FloatingPoint::sum_corners (utils/floating_point.hpp:44-55) with the
   argument orders of each case.  cimp/cjmp/ckmp: coarse midpoints. */
static double sc8(double a, double b, double c, double d, double e, double f, double g8, double h) {
  return (((a + b) + (c + d)) + ((e + f) + (g8 + h)));
}
static double sc4(double a, double b, double c, double d) { return ((a + b) + (c + d)); }

void oracle_prolongate_vc(const oracle_geom* g, const oracle_geom* cg, int nvar,
                          const double* coarse, double* fine, int si, int ei, int sj,
                          int ej, int sk, int ek) {
  const long N = (long)g->nn1 * g->nn2 * g->nn3, CN = (long)cg->nn1 * cg->nn2 * cg->nn3;
  const int cimp = cg->il + (cg->iu - cg->il) / 2, cjmp = cg->jl + (cg->ju - cg->jl) / 2,
            ckmp = cg->kl + (cg->ku - cg->kl) / 2;
#define CO(n, k, j, i) coarse[(n) * CN + (i) + (long)cg->nn1 * ((j) + (long)cg->nn2 * (k))]
#define FI(n, k, j, i) fine[(n) * N + (i) + (long)g->nn1 * ((j) + (long)g->nn2 * (k))]
  const int si_prl = (si > cimp) ? si - 1 : si, ei_prl = (ei < cimp) ? ei + 1 : ei;
  const int sj_prl = (sj > cjmp) ? sj - 1 : sj, ej_prl = (ej < cjmp) ? ej + 1 : ej;
  const int sk_prl = (sk > ckmp) ? sk - 1 : sk, ek_prl = (ek < ckmp) ? ek + 1 : ek;
  const int H = OR_HNCG;
  for (int n = 0; n < nvar; ++n) {
    /* (prl, prl, prl) */
    for (int k = sk_prl; k < ek_prl; ++k)
      for (int j = sj_prl; j < ej_prl; ++j)
        for (int i = si_prl; i < ei_prl; ++i) {
          const int fk = 2 * (k - H) + 1, fj = 2 * (j - H) + 1, fi = 2 * (i - H) + 1;
          double acc = 0.;
          for (int dk = 0; dk < OR_HSZ; ++dk) {
            const int ku = k + dk + 1, kl = k - dk;
            const double lck = LAG3[OR_HSZ - dk - 1];
            for (int dj = 0; dj < OR_HSZ; ++dj) {
              const int ju = j + dj + 1, jl = j - dj;
              const double lckj = lck * LAG3[OR_HSZ - dj - 1];
              for (int di = 0; di < OR_HSZ; ++di) {
                const int iu = i + di + 1, il = i - di;
                const double w = lckj * LAG3[OR_HSZ - di - 1];
                acc += w * sc8(CO(n, ku, ju, iu), CO(n, kl, jl, il), CO(n, ku, ju, il),
                               CO(n, kl, jl, iu), CO(n, kl, ju, il), CO(n, ku, jl, iu),
                               CO(n, kl, ju, iu), CO(n, ku, jl, il));
              }
            }
          }
          FI(n, fk, fj, fi) = acc;
        }
    /* (inj, prl, prl) */
    for (int k = sk; k <= ek; ++k)
      for (int j = sj_prl; j < ej_prl; ++j)
        for (int i = si_prl; i < ei_prl; ++i) {
          const int fk = 2 * (k - H), fj = 2 * (j - H) + 1, fi = 2 * (i - H) + 1;
          double acc = 0.;
          for (int dj = 0; dj < OR_HSZ; ++dj) {
            const int ju = j + dj + 1, jl = j - dj;
            const double lcj = LAG3[OR_HSZ - dj - 1];
            for (int di = 0; di < OR_HSZ; ++di) {
              const int iu = i + di + 1, il = i - di;
              const double w = lcj * LAG3[OR_HSZ - di - 1];
              acc += w * sc4(CO(n, k, ju, iu), CO(n, k, jl, il), CO(n, k, jl, iu), CO(n, k, ju, il));
            }
          }
          FI(n, fk, fj, fi) = acc;
        }
    /* (prl, inj, prl) */
    for (int k = sk_prl; k < ek_prl; ++k)
      for (int j = sj; j <= ej; ++j)
        for (int i = si_prl; i < ei_prl; ++i) {
          const int fk = 2 * (k - H) + 1, fj = 2 * (j - H), fi = 2 * (i - H) + 1;
          double acc = 0.;
          for (int dk = 0; dk < OR_HSZ; ++dk) {
            const int ku = k + dk + 1, kl = k - dk;
            const double lck = LAG3[OR_HSZ - dk - 1];
            for (int di = 0; di < OR_HSZ; ++di) {
              const int iu = i + di + 1, il = i - di;
              const double w = lck * LAG3[OR_HSZ - di - 1];
              acc += w * sc4(CO(n, kl, j, iu), CO(n, ku, j, il), CO(n, ku, j, iu), CO(n, kl, j, il));
            }
          }
          FI(n, fk, fj, fi) = acc;
        }
    /* (prl, prl, inj) */
    for (int k = sk_prl; k < ek_prl; ++k)
      for (int j = sj_prl; j < ej_prl; ++j)
        for (int i = si; i <= ei; ++i) {
          const int fk = 2 * (k - H) + 1, fj = 2 * (j - H) + 1, fi = 2 * (i - H);
          double acc = 0.;
          for (int dk = 0; dk < OR_HSZ; ++dk) {
            const int ku = k + dk + 1, kl = k - dk;
            const double lck = LAG3[OR_HSZ - dk - 1];
            for (int dj = 0; dj < OR_HSZ; ++dj) {
              const int ju = j + dj + 1, jl = j - dj;
              const double w = lck * LAG3[OR_HSZ - dj - 1];
              acc += w * sc4(CO(n, ku, ju, i), CO(n, kl, jl, i), CO(n, kl, ju, i), CO(n, ku, jl, i));
            }
          }
          FI(n, fk, fj, fi) = acc;
        }
    /* (inj, inj, prl) */
    for (int k = sk; k <= ek; ++k)
      for (int j = sj; j <= ej; ++j)
        for (int i = si_prl; i < ei_prl; ++i) {
          const int fk = 2 * (k - H), fj = 2 * (j - H), fi = 2 * (i - H) + 1;
          double acc = 0.;
          for (int di = 0; di < OR_HSZ; ++di)
            acc += LAG3[OR_HSZ - di - 1] * (CO(n, k, j, i - di) + CO(n, k, j, i + di + 1));
          FI(n, fk, fj, fi) = acc;
        }
    /* (inj, prl, inj) */
    for (int k = sk; k <= ek; ++k)
      for (int j = sj_prl; j < ej_prl; ++j)
        for (int i = si; i <= ei; ++i) {
          const int fk = 2 * (k - H), fj = 2 * (j - H) + 1, fi = 2 * (i - H);
          double acc = 0.;
          for (int dj = 0; dj < OR_HSZ; ++dj)
            acc += LAG3[OR_HSZ - dj - 1] * (CO(n, k, j - dj, i) + CO(n, k, j + dj + 1, i));
          FI(n, fk, fj, fi) = acc;
        }
    /* (prl, inj, inj) */
    for (int k = sk_prl; k < ek_prl; ++k)
      for (int j = sj; j <= ej; ++j)
        for (int i = si; i <= ei; ++i) {
          const int fk = 2 * (k - H) + 1, fj = 2 * (j - H), fi = 2 * (i - H);
          double acc = 0.;
          for (int dk = 0; dk < OR_HSZ; ++dk)
            acc += LAG3[OR_HSZ - dk - 1] * (CO(n, k - dk, j, i) + CO(n, k + dk + 1, j, i));
          FI(n, fk, fj, fi) = acc;
        }
    /* (inj, inj, inj) */
    for (int k = sk; k <= ek; ++k)
      for (int j = sj; j <= ej; ++j)
        for (int i = si; i <= ei; ++i)
          FI(n, 2 * (k - H), 2 * (j - H), 2 * (i - H)) = CO(n, k, j, i);
  }
#undef CO
#undef FI
}

/* ProlongationRange (comm/refinement_ops.hpp:48-72) for VC: pcng = ng/2. */
void oracle_prolongation_range(int parity, int ox, int cvs, int cve, int* s, int* e) {
  const int pcng = 4 / 2;
  if (ox > 0) {
    *s = cve + 1;
    *e = cve + pcng;
  } else if (ox < 0) {
    *s = cvs - pcng;
    *e = cvs - 1;
  } else {
    *s = cvs;
    *e = cve;
    if ((parity & 1) == 0)
      *e += pcng;
    else
      *s -= pcng;
  }
}


/* ------------------------------------------------------------------------ */
/* Restart records (src/outputs/restart.cpp:141-147, 206-214).              */
/* ------------------------------------------------------------------------ */
typedef struct {
  double x1min, x2min, x3min;
  double x1max, x2max, x3max;
  double x1rat, x2rat, x3rat;
  int nx1, nx2, nx3;
} or_region_size; /* RegionSize, src/athena.hpp:88-96 */

typedef struct {
  long long lx1, lx2, lx3;
  int level;
} or_logical_location; /* LogicalLocation, src/athena.hpp:58-66 */

unsigned long oracle_rst_header(unsigned char* out, int nbtotal, int root_level,
                                const double* size9, const int* nx, double time, double dt,
                                int ncycle, unsigned long long datasize) {
  or_region_size rs;
  memset(&rs, 0, sizeof rs);
  rs.x1min = size9[0]; rs.x2min = size9[1]; rs.x3min = size9[2];
  rs.x1max = size9[3]; rs.x2max = size9[4]; rs.x3max = size9[5];
  rs.x1rat = size9[6]; rs.x2rat = size9[7]; rs.x3rat = size9[8];
  rs.nx1 = nx[0]; rs.nx2 = nx[1]; rs.nx3 = nx[2];
  unsigned long o = 0;
  memcpy(out + o, &nbtotal, sizeof(int)); o += sizeof(int);
  memcpy(out + o, &root_level, sizeof(int)); o += sizeof(int);
  memcpy(out + o, &rs, sizeof rs); o += sizeof rs;
  memcpy(out + o, &time, sizeof(double)); o += sizeof(double);
  memcpy(out + o, &dt, sizeof(double)); o += sizeof(double);
  memcpy(out + o, &ncycle, sizeof(int)); o += sizeof(int);
  memcpy(out + o, &datasize, sizeof datasize); o += sizeof datasize;
  return o;
}

unsigned long oracle_rst_idrec(unsigned char* out, const long long* lx, int level,
                               double cost) {
  or_logical_location ll;
  memset(&ll, 0, sizeof ll);
  ll.lx1 = lx[0]; ll.lx2 = lx[1]; ll.lx3 = lx[2];
  ll.level = level;
  memcpy(out, &ll, sizeof ll);
  memcpy(out + sizeof ll, &cost, sizeof(double));
  return sizeof ll + sizeof(double);
}
