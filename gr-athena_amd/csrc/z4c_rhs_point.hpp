// z4c_rhs_point.hpp -- the Z4c right-hand side at one interior vertex, in
// two passes with bounded register footprints:
//   curvature pass : R~_ab + R^phi_ab (conformal + conformal-factor Ricci)
//                    and D_a D_b alpha -> 12 values per point, kept in HBM;
//   assembly pass  : A~ contractions, Lie/advective terms, KO dissipation
//                    and the 22 right-hand sides.
// One thread evaluates every stencil its point needs directly from the
// state and the prepared first derivatives (buffer loads, L1/L2 resident
// for the tile).  The arithmetic follows Z4c::Z4cRHS
// (src/z4c/calculate_z4c_rhs.cpp:29-916) term by term and in the same
// order; the reference's pencil loops only batch this per-point sequence.
// Vacuum build and constant shift damping (no Z4C_ETA_CONF /
// Z4C_ETA_TRACK_TP); matter terms are kept for a non-null matter buffer.
#pragma once

#include "z4c_device.hpp"

namespace z4c {

struct RhsParams {
  double chi_psi_power, chi_div_floor, diss, damp_kappa1, damp_kappa2;
  double lapse_oplog, lapse_harmonicf, lapse_harmonic, lapse_advect, lapse_K;
  double shift_Gamma, shift_alpha2Gamma, shift_H, shift_advect, shift_eta;
};

// d1 field index of direction b, prepared field f
__device__ __forceinline__ constexpr int d1f(int b, int f) { return b * 11 + f; }

// Conformal metric inverse and Christoffels from the state and the prepared
// first derivatives (calculate_z4c_rhs.cpp:292-352).
struct Metric {
  double gdd[6], guu[6], Gudd[3][6], Gu[3];
};

__device__ __forceinline__ void metric_at(const Buf& U, const Buf& D, const Pt& P,
                                          Metric& M, double (*Gddd)[6]) {
#pragma unroll
  for (int ab = 0; ab < 6; ++ab) M.gdd[ab] = ldc(U, P, V_GXX + ab);
  {
    const double detg = det3(M.gdd);
    inv3(1.0 / detg, M.gdd, M.guu);
  }
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = a; b < 3; ++b)
        Gddd[c][sym(a, b)] = 0.5 * (ldc(D, P, d1f(a, F_G + sym(b, c))) +
                                    ldc(D, P, d1f(b, F_G + sym(a, c))) -
                                    ldc(D, P, d1f(c, F_G + sym(a, b))));
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int ab = 0; ab < 6; ++ab) {
      double t = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) t += M.guu[sym(c, d)] * Gddd[d][ab];
      M.Gudd[c][ab] = t;
    }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double t = 0.0;
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int c = b; c < 3; ++c)
        t += ((b == c) ? 1.0 : 2.0) * M.guu[sym(b, c)] * M.Gudd[a][sym(b, c)];
    M.Gu[a] = t;
  }
}

// ---------------------------------------------------------------------------
// Curvature pass: cv[0..5] = R~_ab + R^phi_ab, cv[6..11] = D_a D_b alpha.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void rhs_curv_point(const RhsParams& o, const DevBlock& B,
                                               const Buf& U, const Buf& D,
                                               const Pt& P, double* __restrict__ cv,
                                               int p, int N) {
  const double pw = o.chi_psi_power;
  const double oochi_psi_power = 1.0 / pw;
  Metric M;
  double Gddd[3][6];
  metric_at(U, D, P, M, Gddd);

  // conformal Ricci R~_ab (:357-396)
  double dGam[3][3];  // [direction][component]
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) dGam[b][a] = fd_dx(U, P, V_GAMX + a, b, B.cidx1[b]);
  double Rdd[6];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b) {
      double r = 0.0;
#pragma unroll
      for (int c = 0; c < 3; ++c)
        r += 0.5 * (M.gdd[sym(c, a)] * dGam[b][c] + M.gdd[sym(c, b)] * dGam[a][c] +
                    M.Gu[c] * (Gddd[a][sym(b, c)] + Gddd[b][sym(a, c)]));
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = c; d < 3; ++d) {
          // d_c d_d g~_ab: Dxx on the diagonal, nested Dx(c, d_d g~) else
          const double ddg = (c == d)
                                 ? fd_dxx(U, P, V_GXX + sym(a, b), c, B.cidx2[c])
                                 : fd_dx(D, P, d1f(d, F_G + sym(a, b)), c, B.cidx1[c]);
          r -= ((c == d) ? 0.5 : 1.0) * M.guu[sym(c, d)] * ddg;
        }
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
          for (int e = 0; e < 3; ++e)
            r += M.guu[sym(c, d)] * (M.Gudd[e][sym(c, a)] * Gddd[b][sym(e, d)] +
                                     M.Gudd[e][sym(c, b)] * Gddd[a][sym(e, d)] +
                                     M.Gudd[e][sym(a, d)] * Gddd[e][sym(c, b)]);
      Rdd[sym(a, b)] = r;
    }

  // conformal factor terms R^phi_ab (:402-458)
  const double chi = ldc(U, P, V_CHI);
  const double chi_guarded = dmax(chi, o.chi_div_floor);
  const double oochi = 1.0 / chi_guarded;
  double dphi[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) dphi[a] = ldc(D, P, d1f(a, F_CHI)) * oochi * oochi_psi_power;
  {
    double Ddphi[6];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = a; b < 3; ++b) {
        const double ddchi = (a == b) ? fd_dxx(U, P, V_CHI, a, B.cidx2[a])
                                      : fd_dx(D, P, d1f(b, F_CHI), a, B.cidx1[a]);
        double t = ddchi * oochi * oochi_psi_power - pw * dphi[a] * dphi[b];
#pragma unroll
        for (int c = 0; c < 3; ++c) t -= M.Gudd[c][sym(a, b)] * dphi[c];
        Ddphi[sym(a, b)] = t;
      }
    double trphi = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int d = c; d < 3; ++d)
        trphi += ((c == d) ? 1.0 : 2.0) * M.guu[sym(c, d)] *
                 (Ddphi[sym(c, d)] + 2. * dphi[c] * dphi[d]);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = a; b < 3; ++b) {
        const double Rphi = 4. * dphi[a] * dphi[b] - 2. * Ddphi[sym(a, b)] -
                            2. * M.gdd[sym(a, b)] * trphi;
        cv[sym(a, b) * N + p] = Rdd[sym(a, b)] + Rphi;
      }
  }

  // D_a D_b alpha (:484-521)
  double dalpha[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) dalpha[a] = ldc(D, P, d1f(a, F_ALPHA));
  double Dda = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int d = 0; d < 3; ++d) Dda += M.guu[sym(c, d)] * dphi[c] * dalpha[d];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b) {
      const double dda = (a == b) ? fd_dxx(U, P, V_ALPHA, a, B.cidx2[a])
                                  : fd_dx(D, P, d1f(b, F_ALPHA), a, B.cidx1[a]);
      double t = dda - 2. * (dphi[a] * dalpha[b] + dphi[b] * dalpha[a]) +
                 2. * M.gdd[sym(a, b)] * Dda;
#pragma unroll
      for (int c = 0; c < 3; ++c) t -= M.Gudd[c][sym(a, b)] * dalpha[c];
      cv[(6 + sym(a, b)) * N + p] = t;
    }
}

// ---------------------------------------------------------------------------
// Assembly pass: the 22 right-hand sides (incl. KO dissipation) at point p.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void rhs_assemble_point(const RhsParams& o, const DevBlock& B,
                                                   const Buf& U, const Buf& D,
                                                   const Buf& CV, const Buf MT, bool has_mat,
                                                   const Pt& P, double* __restrict__ rhs,
                                                   int p, int N) {
  const double pw = o.chi_psi_power;
  const double chi_pow_DA_u = -6.0 / pw;
  Metric M;
  {
    double Gddd[3][6];
    metric_at(U, D, P, M, Gddd);
  }
  const double alpha = ldc(U, P, V_ALPHA), chi = ldc(U, P, V_CHI),
               Khat = ldc(U, P, V_KHAT), Theta = ldc(U, P, V_THETA);
  const double chi_guarded = dmax(chi, o.chi_div_floor);
  const double oochi = 1.0 / chi_guarded;
  const double oopsi4 = (pw == -4.) ? chi_guarded : pow(chi_guarded, -4. / pw);
  double dchi[3], dalpha[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    dchi[a] = ldc(D, P, d1f(a, F_CHI));
    dalpha[a] = ldc(D, P, d1f(a, F_ALPHA));
  }

  // matter (zero in vacuum) and its trace (:466-476)
  double rho = 0.0, Sd[3] = {0.0, 0.0, 0.0}, Sdd[6] = {0, 0, 0, 0, 0, 0};
  if (has_mat) {
    rho = ldc(MT, P, M_RHO);
#pragma unroll
    for (int a = 0; a < 3; ++a) Sd[a] = ldc(MT, P, M_SX + a);
#pragma unroll
    for (int ab = 0; ab < 6; ++ab) Sdd[ab] = ldc(MT, P, M_SXX + ab);
  }
  double Str = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b)
      Str += ((a == b) ? 1.0 : 2.0) * oopsi4 * M.guu[sym(a, b)] * Sdd[sym(a, b)];

  // traces of the curvature pass output (:522-532, :606-613)
  double Ddalpha = 0.0, R = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b)
      Ddalpha += ((a == b) ? 1.0 : 2.0) * oopsi4 * M.guu[sym(a, b)] *
                 ldc(CV, P, 6 + sym(a, b));
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b)
      R += ((a == b) ? 1.0 : 2.0) * oopsi4 * M.guu[sym(a, b)] * ldc(CV, P, sym(a, b));

  // A~ contractions (:536-601)
  double Add[6];
#pragma unroll
  for (int ab = 0; ab < 6; ++ab) Add[ab] = ldc(U, P, V_AXX + ab);
  double Aud[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      double t = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) t += M.guu[sym(c, d)] * Add[sym(d, b)];
      Aud[c][b] = t;
    }
  double AA[6];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b) {
      double t = 0.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) t += Add[sym(a, c)] * Aud[c][b];
      AA[sym(a, b)] = t;
    }
  double trAA = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b)
      trAA += ((a == b) ? 1.0 : 2.0) * M.guu[sym(a, b)] * AA[sym(a, b)];
  double Auu[6];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b) {
      double t = 0.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) t += M.guu[sym(b, c)] * Aud[a][c];
      Auu[sym(a, b)] = t;
    }
  const double K = Khat + 2. * Theta;
  const double Ht = R + (2. / 3.) * (K * K) - trAA;

  // shift derivatives
  double beta[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) beta[a] = ldc(U, P, V_BETAX + a);
  double dbeta[3][3];  // [direction][component]
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int a = 0; a < 3; ++a) dbeta[b][a] = ldc(D, P, d1f(b, F_BETA + a));
  double dbeta_div = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) dbeta_div += dbeta[a][a];

  const double k1 = o.damp_kappa1, k2 = o.damp_kappa2;
  const double diss = o.diss;
  // advective derivative sum_a Lx(a, beta^a, v), in a order (:184-226)
  auto advect = [&](int v) {
    double L = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) L += fd_lx(U, P, v, a, B.lidx[a], beta[a]);
    return L;
  };
  // KO dissipation added in direction order after the RHS (:908-915)
  auto dissip = [&](int v, double r) {
#pragma unroll
    for (int a = 0; a < 3; ++a) r += fd_diss(U, P, v, a, B.cidxd[a], diss);
    return r;
  };

  // chi, Khat, Theta (:694-711)
  {
    double Lchi = advect(V_CHI);
    Lchi += (1. / 6.) * pw * chi_guarded * dbeta_div;
    const double r_chi = Lchi - (1. / 6.) * pw * chi_guarded * alpha * K;
    rhs[V_CHI * N + p] = dissip(V_CHI, r_chi);
  }
  {
    const double LKhat = advect(V_KHAT);
    double r = -Ddalpha + alpha * (trAA + (1. / 3.) * (K * K)) + LKhat +
               k1 * (1 - k2) * alpha * Theta;
    r += 4 * M_PI * alpha * (Str + rho);
    rhs[V_KHAT * N + p] = dissip(V_KHAT, r);
  }
  {
    const double LTheta = advect(V_THETA);
    double r = LTheta + alpha * (0.5 * Ht - (2. + k2) * k1 * Theta);
    r -= 8. * M_PI * alpha * rho;
    rhs[V_THETA * N + p] = dissip(V_THETA, r);
  }

  // Gamma~ (:568-589 DA, :640-668 Lie completion, :713-729 assembly)
  {
    double DA[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double t = 0.0;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const double dKhat = fd_dx(U, P, V_KHAT, b, B.cidx1[b]);
        const double dTheta = fd_dx(U, P, V_THETA, b, B.cidx1[b]);
        t -= chi_pow_DA_u * Auu[sym(a, b)] * dchi[b] * oochi;
        t -= (1. / 3.) * M.guu[sym(a, b)] * (2. * dKhat + dTheta);
      }
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int c = b; c < 3; ++c)
          t += ((b == c) ? 1.0 : 2.0) * M.Gudd[a][sym(b, c)] * Auu[sym(b, c)];
      DA[a] = t;
    }
    // ddbeta_d^a = (1/3) sum_b d_a d_b beta^b; mixed second derivatives are
    // the reference's nested Dx(lower index, d_(higher index) beta)
    double ddbeta_d[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double t = 0.0;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const int lo = (a < b) ? a : b, hi = (a < b) ? b : a;
        const double v = (a == b) ? fd_dxx(U, P, V_BETAX + b, a, B.cidx2[a])
                                  : fd_dx(D, P, d1f(hi, F_BETA + b), lo, B.cidx1[lo]);
        t += (1. / 3.) * v;
      }
      ddbeta_d[a] = t;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double LGam = advect(V_GAMX + a);
      LGam += (2. / 3.) * M.Gu[a] * dbeta_div;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        LGam += M.guu[sym(a, b)] * ddbeta_d[b] - M.Gu[b] * dbeta[b][a];
#pragma unroll
        for (int c = b; c < 3; ++c) {
          const double ddb = (b == c) ? fd_dxx(U, P, V_BETAX + a, b, B.cidx2[b])
                                      : fd_dx(D, P, d1f(c, F_BETA + a), b, B.cidx1[b]);
          LGam += ((b == c) ? 1.0 : 2.0) * M.guu[sym(b, c)] * ddb;
        }
      }
      const double Gam_a = ldc(U, P, V_GAMX + a);
      double t = 2. * alpha * DA[a] + LGam;
      t -= 2. * alpha * k1 * (Gam_a - M.Gu[a]);
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        t -= 2. * Auu[sym(a, b)] * dalpha[b];
        t -= 16. * M_PI * alpha * M.guu[sym(a, b)] * Sd[b];
      }
      rhs[(V_GAMX + a) * N + p] = dissip(V_GAMX + a, t);
    }
  }

  // g~ and A~ (:671-687 Lie completion, :732-754 assembly)
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b) {
      const int ab = sym(a, b);
      double Lg = advect(V_GXX + ab);
      double LA = advect(V_AXX + ab);
      Lg -= (2. / 3.) * M.gdd[ab] * dbeta_div;
      LA -= (2. / 3.) * Add[ab] * dbeta_div;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Lg += dbeta[a][c] * M.gdd[sym(b, c)];
        LA += dbeta[b][c] * Add[sym(a, c)];
        Lg += dbeta[b][c] * M.gdd[sym(a, c)];
        LA += dbeta[a][c] * Add[sym(b, c)];
      }
      const double r_g = -2. * alpha * Add[ab] + Lg;
      double t = oopsi4 * (-ldc(CV, P, 6 + ab) + alpha * ldc(CV, P, ab));
      t -= (1. / 3.) * M.gdd[ab] * (-Ddalpha + alpha * R);
      t += alpha * (K * Add[ab] - 2. * AA[ab]);
      t += LA;
      t -= 8. * M_PI * alpha * (oopsi4 * Sdd[ab] - (1. / 3.) * Str * M.gdd[ab]);
      rhs[(V_GXX + ab) * N + p] = dissip(V_GXX + ab, r_g);
      rhs[(V_AXX + ab) * N + p] = dissip(V_AXX + ab, t);
    }

  // lapse (:757-766) and shift (:768-902)
  {
    const double Lalpha = advect(V_ALPHA);
    const double f = o.lapse_oplog * o.lapse_harmonicf + o.lapse_harmonic * alpha;
    const double r = (o.lapse_advect * Lalpha - f * alpha * (Khat + 2 * o.lapse_K * Theta));
    rhs[V_ALPHA * N + p] = dissip(V_ALPHA, r);
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double Lbeta = advect(V_BETAX + a);
    const double Gam_a = ldc(U, P, V_GAMX + a);
    double r = o.shift_Gamma * Gam_a + o.shift_advect * Lbeta;
    if (fabs(o.shift_alpha2Gamma) > 0.0) r += o.shift_alpha2Gamma * (alpha * alpha) * Gam_a;
    if (fabs(o.shift_H) > 0.0) {
#pragma unroll
      for (int b = 0; b < 3; ++b)
        r += o.shift_H * alpha * chi_guarded * (0.5 * alpha * dchi[b] - dalpha[b]) *
             M.guu[sym(a, b)];
    }
    r -= o.shift_eta * beta[a];
    rhs[(V_BETAX + a) * N + p] = dissip(V_BETAX + a, r);
  }
}

}  // namespace z4c
