// z4c_adjoint.hpp -- pointwise adjoints (vector-Jacobian products) of the
// vacuum Z4c right-hand side and of AlgConstr.
//
// The right-hand side at an interior vertex p is F(u(p), X(p)): the state at
// the point and the finite-difference values X(p) of the state around it
// (src/z4c/calculate_z4c_rhs.cpp:29-916).  The VJP with a cotangent w splits
// into
//   * the pointwise part, (dF/du(p))^T w and (dF/dX(p))^T w, evaluated here
//     by reverse-mode differentiation written out by hand: the forward
//     quantities of the right-hand side (inverse metric, Christoffels, the
//     Ricci tensor, D_aD_b alpha, the A~ contractions) are recomputed and
//     their adjoints accumulated in reverse order;
//   * the stencils, which are linear: the cotangents of the derivative
//     values are carried back to the vertices by the transposed stencils
//     (a separate kernel).
// The advective terms sum_a beta^a Lx_a(u_f) and the KO terms enter F with
// constant coefficients, so their cotangents are w itself (times lapse_advect
// / shift_advect for alpha / beta); the beta^a factor contributes
// sum_f Lbar_f * Lsel_a(u_f) to the shift's point cotangent, with Lsel_a the
// lop-sided derivative selected by the sign of beta^a.
#pragma once

#include "z4c_device.hpp"
#include "z4c_rhs_point.hpp"

namespace z4c {

// Parameter gradients per cotangent: d<w, rhs>/d(shift_eta, damp_kappa1, damp_kappa2).
constexpr int NPBAR = 3;

// coefficient of the advective sum of field n in its right-hand side
__host__ __device__ __forceinline__ double adv_coef(const RhsParams& o, int n) {
  return n == V_ALPHA ? o.lapse_advect : (n >= V_BETAX ? o.shift_advect : 1.0);
}

// multiplicity of a SYM2 component in a full symmetric contraction
__host__ __device__ constexpr double mult(int a, int b) { return a == b ? 1.0 : 2.0; }

// Forward intermediates of the right-hand side at one vertex (independent
// of the cotangent, so formed once per state and stored, NFWD per vertex).
__host__ __device__ constexpr int f_gu(int ab) { return ab; }
__host__ __device__ constexpr int f_Gddd(int c, int ab) { return 6 + c * 6 + ab; }
__host__ __device__ constexpr int f_Gudd(int c, int ab) { return 24 + c * 6 + ab; }
__host__ __device__ constexpr int f_Gu(int a) { return 42 + a; }
__host__ __device__ constexpr int f_dphi(int a) { return 45 + a; }
__host__ __device__ constexpr int f_Ddphi(int ab) { return 48 + ab; }
__host__ __device__ constexpr int f_RR(int ab) { return 54 + ab; }
__host__ __device__ constexpr int f_DDa(int ab) { return 60 + ab; }
__host__ __device__ constexpr int f_Aud(int c, int b) { return 66 + c * 3 + b; }
__host__ __device__ constexpr int f_AA(int ab) { return 75 + ab; }
__host__ __device__ constexpr int f_Auu(int ab) { return 81 + ab; }
__host__ __device__ constexpr int f_DA(int a) { return 87 + a; }
__host__ __device__ constexpr int f_ddbd(int a) { return 90 + a; }
enum : int {
  F_CHIG = 93, F_OOCHI = 94, F_OOPSI4 = 95, F_TRPHI = 96, F_DDA = 97, F_DDALPHA = 98,
  F_R = 99, F_TRAA = 100, F_DIV = 101, NFWD = 102
};
// per-vertex linearisation buffer: NLIN derivative values, then NFWD
// forward intermediates
constexpr int NLINF = NLIN + NFWD;

template <class XIn, class FOut>
__host__ __device__ __forceinline__ void rhs_fwd_point(const RhsParams& o, const XIn& X, FOut&& F) {
  const double pw = o.chi_psi_power;
  // ------------------------------------------------------ forward values
  double g[6], gu[6];
#pragma unroll
  for (int ab = 0; ab < 6; ++ab) g[ab] = X.u(V_GXX + ab);
  {
    const double detg = det3(g);
    inv3(1.0 / detg, g, gu);
  }
  double Gddd[3][6], Gudd[3][6], Gu[3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = a; b < 3; ++b)
        Gddd[c][sym(a, b)] = 0.5 * (X.d1(V_GXX + sym(b, c), a) + X.d1(V_GXX + sym(a, c), b) -
                                    X.d1(V_GXX + sym(a, b), c));
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int ab = 0; ab < 6; ++ab) {
      double t = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) t += gu[sym(c, d)] * Gddd[d][ab];
      Gudd[c][ab] = t;
    }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double t = 0.0;
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int c = b; c < 3; ++c) t += mult(b, c) * gu[sym(b, c)] * Gudd[a][sym(b, c)];
    Gu[a] = t;
  }
  const double chi = X.u(V_CHI);
  const bool chi_floored = chi < o.chi_div_floor;
  const double chi_g = chi_floored ? o.chi_div_floor : chi;
  const double oochi = 1.0 / chi_g;
  const double oopsi4 = (pw == -4.) ? chi_g : pow(chi_g, -4. / pw);
  double dchi[3], dphi[3], dalpha[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    dchi[a] = X.d1(V_CHI, a);
    dphi[a] = dchi[a] * oochi * (1.0 / pw);
    dalpha[a] = X.d1(V_ALPHA, a);
  }
  double Ddphi[6];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b) {
      double t = X.d2(V_CHI, sym(a, b)) * oochi * (1.0 / pw) - pw * dphi[a] * dphi[b];
#pragma unroll
      for (int c = 0; c < 3; ++c) t -= Gudd[c][sym(a, b)] * dphi[c];
      Ddphi[sym(a, b)] = t;
    }
  double trphi = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int d = c; d < 3; ++d)
      trphi += mult(c, d) * gu[sym(c, d)] * (Ddphi[sym(c, d)] + 2. * dphi[c] * dphi[d]);
  double RR[6];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b) {
      double r = 0.0;
#pragma unroll
      for (int c = 0; c < 3; ++c)
        r += 0.5 * (g[sym(c, a)] * X.d1(V_GAMX + c, b) + g[sym(c, b)] * X.d1(V_GAMX + c, a) +
                    Gu[c] * (Gddd[a][sym(b, c)] + Gddd[b][sym(a, c)]));
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = c; d < 3; ++d)
          r -= 0.5 * mult(c, d) * gu[sym(c, d)] * X.d2(V_GXX + sym(a, b), sym(c, d));
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
          for (int e = 0; e < 3; ++e)
            r += gu[sym(c, d)] * (Gudd[e][sym(c, a)] * Gddd[b][sym(e, d)] +
                                  Gudd[e][sym(c, b)] * Gddd[a][sym(e, d)] +
                                  Gudd[e][sym(a, d)] * Gddd[e][sym(c, b)]);
      RR[sym(a, b)] = r + (4. * dphi[a] * dphi[b] - 2. * Ddphi[sym(a, b)] - 2. * g[sym(a, b)] * trphi);
    }
  double Dda = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int d = 0; d < 3; ++d) Dda += gu[sym(c, d)] * dphi[c] * dalpha[d];
  double DDa[6];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b) {
      double t = X.d2(V_ALPHA, sym(a, b)) - 2. * (dphi[a] * dalpha[b] + dphi[b] * dalpha[a]) +
                 2. * g[sym(a, b)] * Dda;
#pragma unroll
      for (int c = 0; c < 3; ++c) t -= Gudd[c][sym(a, b)] * dalpha[c];
      DDa[sym(a, b)] = t;
    }
  double Ddalpha = 0.0, R = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b) {
      Ddalpha += mult(a, b) * oopsi4 * gu[sym(a, b)] * DDa[sym(a, b)];
      R += mult(a, b) * oopsi4 * gu[sym(a, b)] * RR[sym(a, b)];
    }
  double A[6];
#pragma unroll
  for (int ab = 0; ab < 6; ++ab) A[ab] = X.u(V_AXX + ab);
  double Aud[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      double t = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) t += gu[sym(c, d)] * A[sym(d, b)];
      Aud[c][b] = t;
    }
  double AA[6], Auu[6];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b) {
      double t = 0.0, s = 0.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        t += A[sym(a, c)] * Aud[c][b];
        s += gu[sym(b, c)] * Aud[a][c];
      }
      AA[sym(a, b)] = t;
      Auu[sym(a, b)] = s;
    }
  double trAA = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b) trAA += mult(a, b) * gu[sym(a, b)] * AA[sym(a, b)];
  const double cpow = -6.0 / pw;
  double DA[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double t = 0.0;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      t -= cpow * Auu[sym(a, b)] * dchi[b] * oochi;
      t -= (1. / 3.) * gu[sym(a, b)] * (2. * X.d1(V_KHAT, b) + X.d1(V_THETA, b));
    }
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int c = b; c < 3; ++c) t += mult(b, c) * Gudd[a][sym(b, c)] * Auu[sym(b, c)];
    DA[a] = t;
  }
  const double alpha = X.u(V_ALPHA), Khat = X.u(V_KHAT), Theta = X.u(V_THETA);
  const double K = Khat + 2. * Theta;
  double dbeta[3][3];  // [direction][component]
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int a = 0; a < 3; ++a) dbeta[b][a] = X.d1(V_BETAX + a, b);
  const double div = dbeta[0][0] + dbeta[1][1] + dbeta[2][2];
  double Gam[3], beta[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    Gam[a] = X.u(V_GAMX + a);
    beta[a] = X.u(V_BETAX + a);
  }

  double ddbd[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double t = 0.0;
#pragma unroll
    for (int b = 0; b < 3; ++b) t += (1. / 3.) * X.d2(V_BETAX + b, sym(a, b));
    ddbd[a] = t;
  }
#pragma unroll
  for (int ab = 0; ab < 6; ++ab) {
    F(f_gu(ab), gu[ab]);
    F(f_Ddphi(ab), Ddphi[ab]);
    F(f_RR(ab), RR[ab]);
    F(f_DDa(ab), DDa[ab]);
    F(f_AA(ab), AA[ab]);
    F(f_Auu(ab), Auu[ab]);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      F(f_Gddd(c, ab), Gddd[c][ab]);
      F(f_Gudd(c, ab), Gudd[c][ab]);
    }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    F(f_Gu(a), Gu[a]);
    F(f_dphi(a), dphi[a]);
    F(f_DA(a), DA[a]);
    F(f_ddbd(a), ddbd[a]);
#pragma unroll
    for (int b = 0; b < 3; ++b) F(f_Aud(a, b), Aud[a][b]);
  }
  F(F_CHIG, chi_g);
  F(F_OOCHI, oochi);
  F(F_OOPSI4, oopsi4);
  F(F_TRPHI, trphi);
  F(F_DDA, Dda);
  F(F_DDALPHA, Ddalpha);
  F(F_R, R);
  F(F_TRAA, trAA);
  F(F_DIV, div);
}

// Accumulators of rhs_vjp_point live between its two phases (split before
// the Ricci adjoint); phase a also adds the Gudd adjoints bG[0..17].
struct AdjCarry {
  double b_RR[6], b_g[6], b_gu[6], b_dphi[3], b_Gu[3], b_dchi[3];
  double b_oochi, b_chig, b_oopsi4;
};
constexpr int NCARRY = 30;  // doubles of AdjCarry
// element e of the carry in a fixed order (e a compile-time constant once
// the callers' loops are unrolled)
__host__ __device__ __forceinline__ double& carry_at(AdjCarry& cy, int e) {
  if (e < 6) return cy.b_RR[e];
  if (e < 12) return cy.b_g[e - 6];
  if (e < 18) return cy.b_gu[e - 12];
  if (e < 21) return cy.b_dphi[e - 18];
  if (e < 24) return cy.b_Gu[e - 21];
  if (e < 27) return cy.b_dchi[e - 24];
  if (e == 27) return cy.b_oochi;
  if (e == 28) return cy.b_chig;
  return cy.b_oopsi4;
}

// Phase a of rhs_vjp_point: the right-hand side terms down to D_aD_b alpha
// (every w-dependent term, the parameter partials, the Gudd part of bG).
template <class XIn, class FIn, class UOut, class SOut>
__host__ __device__ __forceinline__ void rhs_vjp_point_a(const RhsParams& o, const XIn& X,
                                                         const FIn& Fw, const double* w,
                                                         UOut&& ub, SOut&& S, double* pb,
                                                         double* bG, int bgs, double wH,
                                                         AdjCarry& cy) {
  const double pw = o.chi_psi_power;
  const double k1 = o.damp_kappa1, k2 = o.damp_kappa2;
  auto g = [&](int ab) { return X.u(V_GXX + ab); };
  auto A = [&](int ab) { return X.u(V_AXX + ab); };
  auto Gam = [&](int a) { return X.u(V_GAMX + a); };
  auto beta = [&](int a) { return X.u(V_BETAX + a); };
  auto dchi = [&](int a) { return X.d1(V_CHI, a); };
  auto dalpha = [&](int a) { return X.d1(V_ALPHA, a); };
  auto dbeta = [&](int b, int a) { return X.d1(V_BETAX + a, b); };
  auto gu = [&](int ab) { return Fw(f_gu(ab)); };
  auto Gddd = [&](int c, int ab) { return Fw(f_Gddd(c, ab)); };
  auto Gudd = [&](int c, int ab) { return Fw(f_Gudd(c, ab)); };
  auto Gu = [&](int a) { return Fw(f_Gu(a)); };
  auto dphi = [&](int a) { return Fw(f_dphi(a)); };
  auto Ddphi = [&](int ab) { return Fw(f_Ddphi(ab)); };
  auto RR = [&](int ab) { return Fw(f_RR(ab)); };
  auto DDa = [&](int ab) { return Fw(f_DDa(ab)); };
  auto Aud = [&](int c, int b) { return Fw(f_Aud(c, b)); };
  auto AA = [&](int ab) { return Fw(f_AA(ab)); };
  auto Auu = [&](int ab) { return Fw(f_Auu(ab)); };
  auto DA = [&](int a) { return Fw(f_DA(a)); };
  auto ddbd = [&](int a) { return Fw(f_ddbd(a)); };
  const bool chi_floored = X.u(V_CHI) < o.chi_div_floor;
  const double chi_g = Fw(F_CHIG), oochi = Fw(F_OOCHI), oopsi4 = Fw(F_OOPSI4);
  const double trphi = Fw(F_TRPHI), Dda = Fw(F_DDA), Ddalpha = Fw(F_DDALPHA), R = Fw(F_R);
  const double trAA = Fw(F_TRAA), div = Fw(F_DIV);
  const double alpha = X.u(V_ALPHA), Khat = X.u(V_KHAT), Theta = X.u(V_THETA);
  const double K = Khat + 2. * Theta;
  const double cpow = -6.0 / pw;

  // ------------------------------------------------------------ adjoints
  double b_alpha = 0, b_K = 0, b_Khat = 0, b_Theta = 0, b_chig = 0, b_div = 0, b_R = 0,
         b_trAA = 0, b_Ddalpha = 0, b_oopsi4 = 0, b_oochi = 0;
  double b_g[6] = {}, b_gu[6] = {}, b_A[6] = {}, b_RR[6] = {}, b_DDa[6] = {}, b_AA[6] = {},
         b_Auu[6] = {};
  double b_DA[3] = {}, b_Gu[3] = {}, b_dphi[3] = {}, b_dchi[3] = {}, b_dalpha[3] = {},
         b_Gam[3] = {}, b_beta[3];
  // the shift's factor in the advective terms: sum_n Lbar_n Lsel_a(u_n)
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double t = 0.0;
#pragma unroll
    for (int n = 0; n < NVAR; ++n) t += adv_coef(o, n) * w[n] * X.ls(n, a);
    b_beta[a] = t;
  }
  double b_dbeta[3][3] = {};

  // shift (calculate_z4c_rhs.cpp:774-901)
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double wa = w[V_BETAX + a];
    b_Gam[a] += wa * o.shift_Gamma;
    if (fabs(o.shift_alpha2Gamma) > 0.0) {
      b_Gam[a] += wa * o.shift_alpha2Gamma * (alpha * alpha);
      b_alpha += wa * o.shift_alpha2Gamma * 2. * alpha * Gam(a);
    }
    if (fabs(o.shift_H) > 0.0) {
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const double T = 0.5 * alpha * dchi(b) - dalpha(b);
        const double c = wa * o.shift_H;
        b_alpha += c * (chi_g * T * gu(sym(a, b)) + alpha * chi_g * 0.5 * dchi(b) * gu(sym(a, b)));
        b_chig += c * alpha * T * gu(sym(a, b));
        b_dchi[b] += c * alpha * chi_g * 0.5 * alpha * gu(sym(a, b));
        b_dalpha[b] -= c * alpha * chi_g * gu(sym(a, b));
        b_gu[sym(a, b)] += c * alpha * chi_g * T;
      }
    }
    b_beta[a] -= o.shift_eta * wa;
    pb[0] -= wa * beta(a);
  }
  // lapse (:757-772)
  {
    const double wl = w[V_ALPHA];
    const double f = o.lapse_oplog * o.lapse_harmonicf + o.lapse_harmonic * alpha;
    const double Q = Khat + 2 * o.lapse_K * Theta;
    b_alpha -= wl * (f * Q + o.lapse_harmonic * alpha * Q);
    b_Khat -= wl * f * alpha;
    b_Theta -= wl * f * alpha * 2 * o.lapse_K;
  }
  // g~ and A~ (:671-754)
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b) {
      const int ab = sym(a, b);
      const double wg = w[V_GXX + ab], wA = w[V_AXX + ab];
      b_alpha += -2. * A(ab) * wg +
                 wA * (oopsi4 * RR(ab) - (1. / 3.) * g(ab) * R + K * A(ab) - 2. * AA(ab));
      b_A[ab] += -2. * alpha * wg + wA * (alpha * K - (2. / 3.) * div);
      b_g[ab] += -(2. / 3.) * div * wg - (1. / 3.) * (-Ddalpha + alpha * R) * wA;
      b_div += -(2. / 3.) * g(ab) * wg - (2. / 3.) * A(ab) * wA;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        b_dbeta[a][c] += wg * g(sym(b, c)) + wA * A(sym(b, c));
        b_dbeta[b][c] += wg * g(sym(a, c)) + wA * A(sym(a, c));
        b_g[sym(b, c)] += wg * dbeta(a, c);
        b_g[sym(a, c)] += wg * dbeta(b, c);
        b_A[sym(a, c)] += wA * dbeta(b, c);
        b_A[sym(b, c)] += wA * dbeta(a, c);
      }
      b_oopsi4 += wA * (-DDa(ab) + alpha * RR(ab));
      b_DDa[ab] -= oopsi4 * wA;
      b_RR[ab] += oopsi4 * alpha * wA;
      b_Ddalpha += (1. / 3.) * g(ab) * wA;
      b_R -= (1. / 3.) * g(ab) * alpha * wA;
      b_K += alpha * A(ab) * wA;
      b_AA[ab] -= 2. * alpha * wA;
    }
  // Gamma~ (:640-729); second-derivative cotangents of beta
  {
    double b_ddbd[3] = {0, 0, 0};
    double b_ddb[3][6] = {};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double wa = w[V_GAMX + a];
      b_alpha += wa * (2. * DA(a) - 2. * k1 * (Gam(a) - Gu(a)));
      b_DA[a] += 2. * alpha * wa;
      b_Gu[a] += wa * ((2. / 3.) * div + 2. * alpha * k1);
      b_div += wa * (2. / 3.) * Gu(a);
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        b_gu[sym(a, b)] += wa * ddbd(b);
        b_ddbd[b] += wa * gu(sym(a, b));
        b_Gu[b] -= wa * dbeta(b, a);
        b_dbeta[b][a] -= wa * Gu(b);
#pragma unroll
        for (int c = b; c < 3; ++c) {
          b_gu[sym(b, c)] += wa * mult(b, c) * X.d2(V_BETAX + a, sym(b, c));
          b_ddb[a][sym(b, c)] += wa * mult(b, c) * gu(sym(b, c));
        }
      }
      b_Gam[a] -= 2. * alpha * k1 * wa;
      pb[1] -= 2. * alpha * (Gam(a) - Gu(a)) * wa;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        b_Auu[sym(a, b)] -= 2. * dalpha(b) * wa;
        b_dalpha[b] -= 2. * Auu(sym(a, b)) * wa;
      }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) b_ddb[b][sym(a, b)] += (1. / 3.) * b_ddbd[a];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int bc = 0; bc < 6; ++bc) S.d2(V_BETAX + a, bc, b_ddb[a][bc]);
  }
  // Theta, Khat, chi (:694-711)
  {
    const double Ht = R + (2. / 3.) * (K * K) - trAA;
    const double wt = w[V_THETA];
    b_alpha += wt * (0.5 * Ht - (2. + k2) * k1 * Theta);
    const double b_Ht = 0.5 * alpha * wt + wH;
    b_Theta -= alpha * (2. + k2) * k1 * wt;
    pb[1] -= alpha * (2. + k2) * Theta * wt;
    pb[2] -= alpha * k1 * Theta * wt;
    const double wk = w[V_KHAT];
    b_Ddalpha -= wk;
    b_alpha += wk * (trAA + (1. / 3.) * (K * K) + k1 * (1 - k2) * Theta);
    b_trAA += alpha * wk;
    b_K += (2. / 3.) * alpha * K * wk;
    b_Theta += k1 * (1 - k2) * alpha * wk;
    pb[1] += (1 - k2) * alpha * Theta * wk;
    pb[2] -= k1 * alpha * Theta * wk;
    const double wc = w[V_CHI];
    b_chig += wc * ((1. / 6.) * pw * div - (1. / 6.) * pw * alpha * K);
    b_div += (1. / 6.) * pw * chi_g * wc;
    b_alpha -= (1. / 6.) * pw * chi_g * K * wc;
    b_K -= (1. / 6.) * pw * chi_g * alpha * wc;
    // Ht = R + 2/3 K^2 - trAA
    b_R += b_Ht;
    b_K += (4. / 3.) * K * b_Ht;
    b_trAA -= b_Ht;
  }
  b_Khat += b_K;
  b_Theta += 2. * b_K;
  ub(V_KHAT, b_Khat);
  ub(V_THETA, b_Theta);
  ub(V_ALPHA, b_alpha);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    ub(V_GAMX + a, b_Gam[a]);
    ub(V_BETAX + a, b_beta[a]);
  }
  // R = sum oopsi4 g^ab RR_ab
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b) {
      const int ab = sym(a, b);
      b_oopsi4 += b_R * mult(a, b) * gu(ab) * RR(ab);
      b_gu[ab] += b_R * oopsi4 * mult(a, b) * RR(ab);
      b_RR[ab] += b_R * oopsi4 * mult(a, b) * gu(ab);
    }
#pragma unroll
  for (int a = 0; a < 3; ++a) b_dbeta[a][a] += b_div;
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int a = 0; a < 3; ++a) S.d1(V_BETAX + a, b, b_dbeta[b][a]);

  // adjoints of the Christoffels, bG[(6c + ab) * bgs] (Gudd) and
  // bG[(18 + 6c + ab) * bgs] (Gddd), and bG[(36 + cd) * bgs] for the
  // inverse-metric adjoint of the quadratic Ricci terms: a per-thread column
  // of LDS on the device (42 doubles)
#pragma unroll
  for (int e = 0; e < 42; ++e) bG[e * bgs] = 0.0;
  // DA (:580-601)
  {
    double b_dK[3] = {0, 0, 0}, b_dTh[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        b_Auu[sym(a, b)] -= cpow * dchi(b) * oochi * b_DA[a];
        b_dchi[b] -= cpow * Auu(sym(a, b)) * oochi * b_DA[a];
        b_oochi -= cpow * Auu(sym(a, b)) * dchi(b) * b_DA[a];
        b_gu[sym(a, b)] -= (1. / 3.) * (2. * X.d1(V_KHAT, b) + X.d1(V_THETA, b)) * b_DA[a];
        b_dK[b] -= (2. / 3.) * gu(sym(a, b)) * b_DA[a];
        b_dTh[b] -= (1. / 3.) * gu(sym(a, b)) * b_DA[a];
      }
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int c = b; c < 3; ++c) {
          bG[(6 * (a) + (sym(b, c))) * bgs] += mult(b, c) * Auu(sym(b, c)) * b_DA[a];
          b_Auu[sym(b, c)] += mult(b, c) * Gudd(a, sym(b, c)) * b_DA[a];
        }
    }
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      S.d1(V_KHAT, b, b_dK[b]);
      S.d1(V_THETA, b, b_dTh[b]);
    }
  }
  // trAA, Auu, AA, Aud (:536-575)
  {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = a; b < 3; ++b) {
        b_gu[sym(a, b)] += mult(a, b) * AA(sym(a, b)) * b_trAA;
        b_AA[sym(a, b)] += mult(a, b) * gu(sym(a, b)) * b_trAA;
      }
    double b_Aud[3][3] = {};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = a; b < 3; ++b)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          b_gu[sym(b, c)] += Aud(a, c) * b_Auu[sym(a, b)];
          b_Aud[a][c] += gu(sym(b, c)) * b_Auu[sym(a, b)];
          b_A[sym(a, c)] += Aud(c, b) * b_AA[sym(a, b)];
          b_Aud[c][b] += A(sym(a, c)) * b_AA[sym(a, b)];
        }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          b_gu[sym(c, d)] += A(sym(d, b)) * b_Aud[c][b];
          b_A[sym(d, b)] += gu(sym(c, d)) * b_Aud[c][b];
        }
  }
#pragma unroll
  for (int ab = 0; ab < 6; ++ab) ub(V_AXX + ab, b_A[ab]);
  // D_aD_b alpha and its trace (:484-532)
  double b_Dda = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b) {
      const int ab = sym(a, b);
      b_oopsi4 += b_Ddalpha * mult(a, b) * gu(ab) * DDa(ab);
      b_gu[ab] += b_Ddalpha * oopsi4 * mult(a, b) * DDa(ab);
      const double v = b_DDa[ab] + b_Ddalpha * oopsi4 * mult(a, b) * gu(ab);
      S.d2(V_ALPHA, ab, v);
      b_dphi[a] -= 2. * dalpha(b) * v;
      b_dalpha[b] -= 2. * dphi(a) * v;
      b_dphi[b] -= 2. * dalpha(a) * v;
      b_dalpha[a] -= 2. * dphi(b) * v;
      b_g[ab] += 2. * Dda * v;
      b_Dda += 2. * g(ab) * v;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        bG[(6 * (c) + (ab)) * bgs] -= dalpha(c) * v;
        b_dalpha[c] -= Gudd(c, ab) * v;
      }
    }
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      b_gu[sym(c, d)] += dphi(c) * dalpha(d) * b_Dda;
      b_dphi[c] += gu(sym(c, d)) * dalpha(d) * b_Dda;
      b_dalpha[d] += gu(sym(c, d)) * dphi(c) * b_Dda;
    }
#pragma unroll
  for (int a = 0; a < 3; ++a) S.d1(V_ALPHA, a, b_dalpha[a]);

  // the accumulators live across the split
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    cy.b_RR[q] = b_RR[q];
    cy.b_g[q] = b_g[q];
    cy.b_gu[q] = b_gu[q];
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    cy.b_dphi[q] = b_dphi[q];
    cy.b_Gu[q] = b_Gu[q];
    cy.b_dchi[q] = b_dchi[q];
  }
  cy.b_oochi = b_oochi;
  cy.b_chig = b_chig;
  cy.b_oopsi4 = b_oopsi4;
}

// Phase b of rhs_vjp_point: the Ricci tensor, the conformal factor, the
// Christoffels and the inverse metric; bG holds phase a's Gudd adjoints and
// zeros above.
template <class XIn, class FIn, class UOut, class SOut>
__host__ __device__ __forceinline__ void rhs_vjp_point_b(const RhsParams& o, const XIn& X,
                                                         const FIn& Fw, UOut&& ub, SOut&& S,
                                                         double* bG, int bgs,
                                                         const AdjCarry& cy) {
  const double pw = o.chi_psi_power;
  const double k1 = o.damp_kappa1, k2 = o.damp_kappa2;
  auto g = [&](int ab) { return X.u(V_GXX + ab); };
  auto A = [&](int ab) { return X.u(V_AXX + ab); };
  auto Gam = [&](int a) { return X.u(V_GAMX + a); };
  auto beta = [&](int a) { return X.u(V_BETAX + a); };
  auto dchi = [&](int a) { return X.d1(V_CHI, a); };
  auto dalpha = [&](int a) { return X.d1(V_ALPHA, a); };
  auto dbeta = [&](int b, int a) { return X.d1(V_BETAX + a, b); };
  auto gu = [&](int ab) { return Fw(f_gu(ab)); };
  auto Gddd = [&](int c, int ab) { return Fw(f_Gddd(c, ab)); };
  auto Gudd = [&](int c, int ab) { return Fw(f_Gudd(c, ab)); };
  auto Gu = [&](int a) { return Fw(f_Gu(a)); };
  auto dphi = [&](int a) { return Fw(f_dphi(a)); };
  auto Ddphi = [&](int ab) { return Fw(f_Ddphi(ab)); };
  auto RR = [&](int ab) { return Fw(f_RR(ab)); };
  auto DDa = [&](int ab) { return Fw(f_DDa(ab)); };
  auto Aud = [&](int c, int b) { return Fw(f_Aud(c, b)); };
  auto AA = [&](int ab) { return Fw(f_AA(ab)); };
  auto Auu = [&](int ab) { return Fw(f_Auu(ab)); };
  auto DA = [&](int a) { return Fw(f_DA(a)); };
  auto ddbd = [&](int a) { return Fw(f_ddbd(a)); };
  const bool chi_floored = X.u(V_CHI) < o.chi_div_floor;
  const double chi_g = Fw(F_CHIG), oochi = Fw(F_OOCHI), oopsi4 = Fw(F_OOPSI4);
  const double trphi = Fw(F_TRPHI), Dda = Fw(F_DDA), Ddalpha = Fw(F_DDALPHA), R = Fw(F_R);
  const double trAA = Fw(F_TRAA), div = Fw(F_DIV);
  const double alpha = X.u(V_ALPHA), Khat = X.u(V_KHAT), Theta = X.u(V_THETA);
  const double K = Khat + 2. * Theta;
  const double cpow = -6.0 / pw;

  double b_RR[6], b_g[6], b_gu[6], b_dphi[3], b_Gu[3], b_dchi[3];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    b_RR[q] = cy.b_RR[q];
    b_g[q] = cy.b_g[q];
    b_gu[q] = cy.b_gu[q];
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    b_dphi[q] = cy.b_dphi[q];
    b_Gu[q] = cy.b_Gu[q];
    b_dchi[q] = cy.b_dchi[q];
  }
  double b_oochi = cy.b_oochi, b_chig = cy.b_chig, b_oopsi4 = cy.b_oopsi4;
  // Ricci (:357-458)
  double b_Ddphi[6] = {};
  double b_trphi = 0.0;
  {
    double b_dGam[3][3] = {};  // [component][direction]
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = a; b < 3; ++b) {
              const int ab = sym(a, b);
        const double r = b_RR[ab];
        b_dphi[a] += 4. * dphi(b) * r;
        b_dphi[b] += 4. * dphi(a) * r;
        b_Ddphi[ab] -= 2. * r;
        b_g[ab] -= 2. * trphi * r;
        b_trphi -= 2. * g(ab) * r;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          b_g[sym(c, a)] += 0.5 * X.d1(V_GAMX + c, b) * r;
          b_dGam[c][b] += 0.5 * g(sym(c, a)) * r;
          b_g[sym(c, b)] += 0.5 * X.d1(V_GAMX + c, a) * r;
          b_dGam[c][a] += 0.5 * g(sym(c, b)) * r;
          b_Gu[c] += 0.5 * (Gddd(a, sym(b, c)) + Gddd(b, sym(a, c))) * r;
          bG[(18 + 6 * (a) + (sym(b, c))) * bgs] += 0.5 * Gu(c) * r;
          bG[(18 + 6 * (b) + (sym(a, c))) * bgs] += 0.5 * Gu(c) * r;
        }
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int d = c; d < 3; ++d) {
            b_gu[sym(c, d)] -= 0.5 * mult(c, d) * X.d2(V_GXX + ab, sym(c, d)) * r;
            S.d2(V_GXX + ab, sym(c, d), -0.5 * mult(c, d) * gu(sym(c, d)) * r);
          }
      }
    // The quadratic Christoffel terms, sum_{a<=b} R_ab Q_ab with
    // Q_ab = g^cd (G^e_ca G_b,ed + G^e_cb G_a,ed + G^e_ad G_e,cb).  Q is
    // symmetric in (a, b), so the sum equals the sum over ordered (a, b) of
    // w_ab g^cd (2 G^e_ca G_b,ed + G^e_ad G_e,cb), w_aa = R_aa and
    // w_ab = R_ab / 2, whose derivatives contract through
    // Acon[c][b][e] = sum_d g^cd G_b,ed and Ncon[a][d][e] = sum_c g^dc G^e_ca.
    // Cotangents of a symmetric pair add both orders (the packing of bG).
    {
      double w[3][3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) w[a][b] = (a == b) ? b_RR[sym(a, a)] : 0.5 * b_RR[sym(a, b)];
      {
        // G^e_ca (term 1) and G^e_ad (term 3), through Acon
        double Acon[3][3][3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int e = 0; e < 3; ++e) {
              double t = 0.0;
#pragma unroll
              for (int d = 0; d < 3; ++d) t += gu(sym(c, d)) * Gddd(b, sym(e, d));
              Acon[c][b][e] = t;  // sum_d g^cd G_b,ed
            }
#pragma unroll
        for (int e = 0; e < 3; ++e)
#pragma unroll
          for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              double t1 = 0.0, t3 = 0.0;
#pragma unroll
              for (int b = 0; b < 3; ++b) {
                t1 += w[a][b] * Acon[c][b][e];
                t3 += w[c][b] * Acon[a][e][b];  // term 3's G^e_ad with (a, d) = (c, a)
              }
              bG[(6 * e + sym(c, a)) * bgs] += 2.0 * t1 + t3;
            }
      }
      {
        // G_b,ed (term 1) and G_e,cb (term 3), through Ncon
        double Ncon[3][3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int e = 0; e < 3; ++e) {
              double t = 0.0;
#pragma unroll
              for (int c = 0; c < 3; ++c) t += gu(sym(d, c)) * Gudd(e, sym(c, a));
              Ncon[a][d][e] = t;  // sum_c g^dc G^e_ca
            }
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
          for (int e = 0; e < 3; ++e)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              double t1 = 0.0, t3 = 0.0;
#pragma unroll
              for (int a = 0; a < 3; ++a) {
                t1 += w[a][b] * Ncon[a][d][e];
                t3 += w[a][d] * Ncon[a][e][b];  // term 3's G_e,cb with (e, c, b) = (b, e, d)
              }
              bG[(18 + 6 * b + sym(e, d)) * bgs] += 2.0 * t1 + t3;
            }
      }
      {
        // g^cd, term 1: 2 sum_{a,e} G^e_ca P[a][e][d], P[a][e][d] = sum_b w_ab G_b,ed
        double P[3][3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int e = 0; e < 3; ++e)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              double t = 0.0;
#pragma unroll
              for (int b = 0; b < 3; ++b) t += w[a][b] * Gddd(b, sym(e, d));
              P[a][e][d] = t;
            }
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            double t = 0.0;
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
              for (int e = 0; e < 3; ++e) t += Gudd(e, sym(c, a)) * P[a][e][d];
            bG[(36 + sym(c, d)) * bgs] += 2.0 * t;
          }
      }
      {
        // g^cd, term 3: sum_{a,e} G^e_ad Q2[e][c][a], Q2[e][c][a] = sum_b w_ab G_e,cb
        double Q2[3][3][3];
#pragma unroll
        for (int e = 0; e < 3; ++e)
#pragma unroll
          for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              double t = 0.0;
#pragma unroll
              for (int b = 0; b < 3; ++b) t += w[a][b] * Gddd(e, sym(c, b));
              Q2[e][c][a] = t;
            }
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            double t = 0.0;
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
              for (int e = 0; e < 3; ++e) t += Gudd(e, sym(a, d)) * Q2[e][c][a];
            bG[(36 + sym(c, d)) * bgs] += t;
          }
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int b = 0; b < 3; ++b) S.d1(V_GAMX + c, b, b_dGam[c][b]);
  }
#pragma unroll
  for (int cd = 0; cd < 6; ++cd) b_gu[cd] += bG[(36 + cd) * bgs];
  // trphi, Ddphi, dphi
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int d = c; d < 3; ++d) {
      const int cd = sym(c, d);
      b_gu[cd] += mult(c, d) * (Ddphi(cd) + 2. * dphi(c) * dphi(d)) * b_trphi;
      b_Ddphi[cd] += mult(c, d) * gu(cd) * b_trphi;
      b_dphi[c] += 2. * mult(c, d) * gu(cd) * dphi(d) * b_trphi;
      b_dphi[d] += 2. * mult(c, d) * gu(cd) * dphi(c) * b_trphi;
    }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b) {
      const int ab = sym(a, b);
      const double v = b_Ddphi[ab];
      S.d2(V_CHI, ab, oochi * (1.0 / pw) * v);
      b_oochi += X.d2(V_CHI, ab) * (1.0 / pw) * v;
      b_dphi[a] -= pw * dphi(b) * v;
      b_dphi[b] -= pw * dphi(a) * v;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        bG[(6 * (c) + (ab)) * bgs] -= dphi(c) * v;
        b_dphi[c] -= Gudd(c, ab) * v;
      }
    }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    b_dchi[a] += oochi * (1.0 / pw) * b_dphi[a];
    b_oochi += dchi(a) * (1.0 / pw) * b_dphi[a];
    S.d1(V_CHI, a, b_dchi[a]);
  }
  b_chig += -oochi * oochi * b_oochi + ((pw == -4.) ? b_oopsi4 : (-4. / pw) * oopsi4 * oochi * b_oopsi4);
  // Christoffels: Gu, Gudd, Gddd (:300-340)
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int c = b; c < 3; ++c) {
        b_gu[sym(b, c)] += mult(b, c) * Gudd(a, sym(b, c)) * b_Gu[a];
        bG[(6 * (a) + (sym(b, c))) * bgs] += mult(b, c) * gu(sym(b, c)) * b_Gu[a];
      }
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int ab = 0; ab < 6; ++ab)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        b_gu[sym(c, d)] += Gddd(d, ab) * bG[(6 * (c) + (ab)) * bgs];
        bG[(18 + 6 * (d) + (ab)) * bgs] += gu(sym(c, d)) * bG[(6 * (c) + (ab)) * bgs];
      }
  {
    double b_dg[3][6] = {};  // [direction][component]
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = a; b < 3; ++b) {
          const double v = 0.5 * bG[(18 + 6 * (c) + (sym(a, b))) * bgs];
          b_dg[a][sym(b, c)] += v;
          b_dg[b][sym(a, c)] += v;
          b_dg[c][sym(a, b)] -= v;
        }
#pragma unroll
    for (int ab = 0; ab < 6; ++ab)
#pragma unroll
      for (int c = 0; c < 3; ++c) S.d1(V_GXX + ab, c, b_dg[c][ab]);
  }
  // inverse metric: b_g -= G^-1 Sbar G^-1 (symmetric packing)
  {
    double T[3][3];
    double Sb[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) Sb[i][j] = b_gu[sym(i, j)] * (i == j ? 1.0 : 0.5);
    double M[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) t += gu(sym(i, k)) * Sb[k][j];
        M[i][j] = t;
      }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) t += M[i][k] * gu(sym(k, j));
        T[i][j] = t;
      }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = i; j < 3; ++j) b_g[sym(i, j)] -= mult(i, j) * T[i][j];
  }

  // point cotangents of chi and g~
  ub(V_CHI, chi_floored ? 0.0 : b_chig);
#pragma unroll
  for (int ab = 0; ab < 6; ++ab) ub(V_GXX + ab, b_g[ab]);
}

// Reverse mode of rhs_curvature_point + rhs_assembly_point (vacuum), from
// the stored forward intermediates.
//   X.u(n)       state at the point
//   X.d1(n, a)   d_a u_n, X.d2(n, ab) d_a d_b u_n, X.ls(n, a) Lsel_a(u_n)
//   Fw(s)        forward intermediate s (rhs_fwd_point)
//   w[n]         cotangent of rhs_n at the point
//   wH           cotangent of the Hamiltonian constraint Ht = R + 2/3 K^2 -
//                A~_ab A~^ab formed inside the right-hand side (:606-624)
// Outputs: ub(n, v) stores the pointwise cotangent of u_n(p), S.d1(n, a, v)
// and S.d2(n, ab, v) the derivative cotangents (each exactly once, as soon
// as final); pb[0..2] += parameter partials.  Forward values are read where
// they are used rather than held in registers.
template <class XIn, class FIn, class UOut, class SOut>
__host__ __device__ __forceinline__ void rhs_vjp_point(const RhsParams& o, const XIn& X,
                                                       const FIn& Fw, const double* w,
                                                       UOut&& ub, SOut&& S, double* pb,
                                                       double* bG, int bgs, double wH) {
  AdjCarry cy;
  rhs_vjp_point_a(o, X, Fw, w, ub, S, pb, bG, bgs, wH, cy);
  rhs_vjp_point_b(o, X, Fw, ub, S, bG, bgs, cy);
}

// Hamiltonian constraint of the Z4c right-hand side at a vertex
// (calculate_z4c_rhs.cpp:606-624), from the stored forward intermediates.
__host__ __device__ __forceinline__ double ham_from_fwd(double R, double trAA, double Khat,
                                                        double Theta) {
  const double K = Khat + 2. * Theta;
  return R + (2. / 3.) * (K * K) - trAA;
}
// the reference's cell loop of the history sums: vertices il..iu-1 (per axis)
__host__ __device__ __forceinline__ bool in_cell_loop(const DevBlock& B, int i, int j, int k) {
  return i < B.iu && j < B.ju && k < B.ku;
}

// Reverse mode of algconstr_vals (z4c.cpp:677-799) at one point: on entry
// bg, bA, bchi, balpha hold the cotangents of the outputs, on exit those of
// the inputs g, A, chi, alpha (the other variables pass through).
__host__ __device__ __forceinline__ void algconstr_vjp(const double* g, const double* A, double chi,
                                              double alpha, double* bg, double* bA,
                                              double& bchi, double& balpha, double eps_floor,
                                              int tf_force, int force_reg) {
  double detg = det3(g);
  bool project = true, det_const = false;
  if (tf_force) {
    if (!(detg > 0.)) {
      detg = 1.;
      det_const = true;
    }
  } else {
    project = detg > eps_floor;
  }
  if (project) {
    const bool root_const = (detg - 1.0) < eps_floor;
    const double s = root_const ? 1.0 + eps_floor / -3.0 : pow(detg, 1.0 / -3.0);
    double g1[6];
#pragma unroll
    for (int ab = 0; ab < 6; ++ab) g1[ab] = g[ab] * s;
    const double tr = trace_rank2(1.0, g1, A);
    double btr = 0.0;
#pragma unroll
    for (int ab = 0; ab < 6; ++ab) btr -= (1.0 / 3.0) * bA[ab] * g1[ab];
    double bg1[6];
#pragma unroll
    for (int ab = 0; ab < 6; ++ab) bg1[ab] = bg[ab] - (1.0 / 3.0) * tr * bA[ab];
    const double gxx = g1[0], gxy = g1[1], gxz = g1[2], gyy = g1[3], gyz = g1[4], gzz = g1[5];
    const double Axx = A[0], Axy = A[1], Axz = A[2], Ayy = A[3], Ayz = A[4], Azz = A[5];
    // d tr / d g1 and d tr / d A of TraceRank2 (linear_algebra.hpp)
    bg1[0] += btr * (-2. * Ayz * gyz + Azz * gyy + Ayy * gzz);
    bg1[1] += btr * (2. * (gxz * Ayz + Axz * gyz - Axy * gzz) - 2. * Azz * gxy);
    bg1[2] += btr * (2. * (Ayz * gxy - Axz * gyy + Axy * gyz) - 2. * Ayy * gxz);
    bg1[3] += btr * (Axx * gzz + gxx * Azz - 2. * gxz * Axz);
    bg1[4] += btr * (-2. * Ayz * gxx + 2. * gxz * Axy + 2. * gxy * Axz - 2. * Axx * gyz);
    bg1[5] += btr * (Axx * gyy + gxx * Ayy - 2. * gxy * Axy);
    bA[0] += btr * (gyy * gzz - gyz * gyz);
    bA[1] += btr * (2. * gxz * gyz - 2. * gxy * gzz);
    bA[2] += btr * (2. * gxy * gyz - 2. * gxz * gyy);
    bA[3] += btr * (gxx * gzz - gxz * gxz);
    bA[4] += btr * (2. * gxy * gxz - 2. * gxx * gyz);
    bA[5] += btr * (gxx * gyy - gxy * gxy);
    double bs = 0.0;
#pragma unroll
    for (int ab = 0; ab < 6; ++ab) {
      bs += bg1[ab] * g[ab];
      bg[ab] = s * bg1[ab];
    }
    if (!root_const && !det_const) {
      const double bdet = bs * (-1.0 / 3.0) * s / detg;
      const double hxx = g[0], hxy = g[1], hxz = g[2], hyy = g[3], hyz = g[4], hzz = g[5];
      bg[0] += bdet * (hyy * hzz - hyz * hyz);
      bg[1] += bdet * (2. * hxz * hyz - 2. * hxy * hzz);
      bg[2] += bdet * (2. * hxy * hyz - 2. * hxz * hyy);
      bg[3] += bdet * (hxx * hzz - hxz * hxz);
      bg[4] += bdet * (2. * hxy * hxz - 2. * hxx * hyz);
      bg[5] += bdet * (hxx * hyy - hxy * hxy);
    }
  }
  if (force_reg) {
    if (chi < eps_floor) bchi = 0.0;
    const bool moved = (alpha >= 0) ? (alpha < eps_floor) : (alpha > -eps_floor);
    if (moved) balpha = 0.0;
  }
}

// ---------------------------------------------------------------------------
// Per-vertex bodies of the adjoint kernels, on raw pointers so that the same
// code runs in the kernels and in host loops.  Buffers of one block:
//   u   state, u[n*npts + p]
//   lin linearisation, lin[slot*nint + q] (q = interior index)
//   w   cotangent (padded layout like u; interior values used)
//   sb  derivative cotangents, sb[slot*nint + q]
//   ubar output (padded layout like u)
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ int iidx(const DevBlock& B, int i, int j, int k) {
  return ((k - B.kl) * B.nj + (j - B.jl)) * B.ni + (i - B.il);
}
__host__ __device__ __forceinline__ bool is_interior(const DevBlock& B, int i, int j, int k) {
  return i >= B.il && i <= B.iu && j >= B.jl && j <= B.ju && k >= B.kl && k <= B.ku;
}

// centred first/second derivative weights (x60, x180) on -3..3, lop-sided
// weights (x60) on -4..2 / -2..4, KO weights on -4..4
__host__ __device__ __forceinline__ double w_d1(int o) {
  return o == -3 ? -1. : o == -2 ? 9. : o == -1 ? -45. : o == 1 ? 45. : o == 2 ? -9. : o == 3 ? 1. : 0.;
}
__host__ __device__ __forceinline__ double w_d2(int o) {
  return (o == -3 || o == 3) ? 2. : (o == -2 || o == 2) ? -27. : (o == -1 || o == 1) ? 270. : -490.;
}
__host__ __device__ __forceinline__ double w_left(int o) {  // o in -4..2
  return o == -4 ? 1. : o == -3 ? -8. : o == -2 ? 30. : o == -1 ? -80. : o == 0 ? 35. : o == 1 ? 24. : o == 2 ? -2. : 0.;
}
__host__ __device__ __forceinline__ double w_right(int o) {  // o in -2..4
  return o == 4 ? -1. : o == 3 ? 8. : o == 2 ? -30. : o == 1 ? 80. : o == 0 ? -35. : o == -1 ? -24. : o == -2 ? 2. : 0.;
}
__host__ __device__ __forceinline__ double w_ko(int o) {
  const int a = o < 0 ? -o : o;
  return a == 4 ? 1. : a == 3 ? -8. : a == 2 ? 28. : a == 1 ? -56. : 70.;
}

struct LinPoint {
  const double* u;
  const double* lin;
  int N, p, nint, q;
  __host__ __device__ double u_(int n) const { return u[(long long)n * N + p]; }
};
struct LinIn {
  LinPoint P;
  __host__ __device__ double u(int n) const { return P.u_(n); }
  __host__ __device__ double d1(int n, int a) const { return P.lin[l_d1(n, a) * P.nint + P.q]; }
  __host__ __device__ double d2(int n, int ab) const { return P.lin[l_d2(n, ab) * P.nint + P.q]; }
  __host__ __device__ double ls(int n, int a) const { return P.lin[l_ls(n, a) * P.nint + P.q]; }
};
struct SlotOut {
  double* sb;
  int nint, q;
  __host__ __device__ void d1(int n, int a, double v) const { sb[l_d1(n, a) * nint + q] = v; }
  __host__ __device__ void d2(int n, int ab, double v) const { sb[l_d2(n, ab) * nint + q] = v; }
};

// Forward intermediates of the right-hand side at interior vertex (i, j, k),
// from u and the derivatives the linearisation pass of the stencil z-march
// stored (a separate pass: the stencils run as a streaming z-march, the
// register-heavy algebra at its own occupancy).
__host__ __device__ inline void adj_lin_fwd_at(const RhsParams& o, const DevBlock& B,
                                               const double* u, double* lin, int i, int j, int k) {
  const int N = B.npts, nint = B.nint;
  const int p = i + B.nn1 * (j + B.nn2 * k);
  const int q = iidx(B, i, j, k);
  LinIn X{LinPoint{u, lin, N, p, nint, q}};
  double* fw = lin + (long long)NLIN * nint + q;
  rhs_fwd_point(o, X, [&](int s, double v) { fw[(long long)s * nint] = v; });
}


// Pointwise adjoint at interior vertex (i, j, k): writes the point part of
// ubar and the derivative cotangents, adds the parameter partials to pb.
// bG: 36 accumulators with stride bgs (see rhs_vjp_point).
__host__ __device__ inline void adj_point_at(const RhsParams& o, const DevBlock& B,
                                             const double* u, const double* lin,
                                             const double* w, double* ubar, double* sb,
                                             int i, int j, int k, double* pb, double* bG,
                                             int bgs, double wH = 0.0) {
  const int N = B.npts;
  const int p = i + B.nn1 * (j + B.nn2 * k);
  const int q = iidx(B, i, j, k);
  double wp[NVAR];
  for (int n = 0; n < NVAR; ++n) wp[n] = w ? w[(long long)n * N + p] : 0.0;
  auto ub = [&](int n, double v) { ubar[(long long)n * N + p] = v; };
  LinIn X{LinPoint{u, lin, N, p, B.nint, q}};
  const double* fw = lin + (long long)NLIN * B.nint + q;
  const int nint = B.nint;
  auto Fw = [&](int s) { return fw[(long long)s * nint]; };
  rhs_vjp_point(o, X, Fw, wp, ub, SlotOut{sb, B.nint, q}, pb, bG, bgs, wH);
}

// Transposed stencils: the stencil parts of ubar_n (all 22 fields) at vertex
// (i, j, k) of the padded block, gathered from the derivative cotangents
// (first and second derivatives), the advective cotangent adv_coef*w (times
// beta^a and the lop-sided weights selected at the source vertex) and the KO
// cotangent w at the interior vertices whose stencils reach (i, j, k).
// Axis a of the gather: source coordinate c_a - o must lie in [0, n_a) with
// the other two coordinates interior; qs/ps are the interior / padded
// strides of the axis.  The bounds, the shift values and the weights of a
// source are formed once for all fields.  Accessors: S(slot, q) derivative
// cotangent at interior index q, (*W)(n, p) the cotangent of rhs_n at padded
// index p (W == nullptr: zero, no advective / KO part), Bx(a, p) beta^a.
// out(n, v) receives the sums.  Fields N0 <= n < N1 (a subset keeps fewer
// accumulators live).
struct AdjAxis {
  int c, n, qs, ps;
  bool others;  // the other two coordinates are interior
};

template <int N0 = 0, int N1 = NVAR, class SIn, class WIn, class BIn, class Out>
__host__ __device__ inline void adj_stencil_at(const RhsParams& o, const DevBlock& B,
                                               const SIn& S, const WIn* W, const BIn& Bx,
                                               int i, int j, int k, Out&& out) {
  const int ci = i - B.il, cj = j - B.jl, ck = k - B.kl;
  const bool ix = ci >= 0 && ci < B.ni, iy = cj >= 0 && cj < B.nj, iz = ck >= 0 && ck < B.nk;
  const int q0 = (ck * B.nj + cj) * B.ni + ci;  // used only at interior sources
  const int p0 = i + B.nn1 * (j + B.nn2 * k);
  const AdjAxis ax0{ci, B.ni, 1, 1, iy && iz}, ax1{cj, B.nj, B.ni, B.nn1, ix && iz},
      ax2{ck, B.nk, B.ni * B.nj, B.nn1 * B.nn2, ix && iy};
  // selected by value (no indexed local array, which would live in scratch)
  auto AX = [&](int a) -> const AdjAxis& { return a == 0 ? ax0 : (a == 1 ? ax1 : ax2); };
  double acc[N1 - N0];
#pragma unroll
  for (int n = 0; n < N1 - N0; ++n) acc[n] = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (!AX(a).others) continue;
#pragma unroll
    for (int oa = -3; oa <= 3; ++oa) {
      const int c = AX(a).c - oa;
      if (c < 0 || c >= AX(a).n) continue;
      const int sq = q0 - oa * AX(a).qs;
      const double c1 = w_d1(oa) * B.cidx1[a], c2 = w_d2(oa) * B.cidx2[a];
#pragma unroll
      for (int n = N0; n < N1; ++n) {
        if (lin1(n) >= 0 && oa != 0) acc[n - N0] += c1 * S(l_d1(n, a), sq);
        if (lin2(n) >= 0) acc[n - N0] += c2 * S(l_d2(n, sym(a, a)), sq);
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a + 1; b < 3; ++b) {
      const int e = 3 - a - b;  // the third axis
      if (!((e == 0) ? ix : (e == 1) ? iy : iz)) continue;
      const double cab = B.cidx1[a] * B.cidx1[b];
#pragma unroll
      for (int oa = -3; oa <= 3; ++oa) {
        const int ca = AX(a).c - oa;
        if (oa == 0 || ca < 0 || ca >= AX(a).n) continue;
#pragma unroll
        for (int ob = -3; ob <= 3; ++ob) {
          const int cb = AX(b).c - ob;
          if (ob == 0 || cb < 0 || cb >= AX(b).n) continue;
          const int sq = q0 - oa * AX(a).qs - ob * AX(b).qs;
          const double cw = (w_d1(oa) * w_d1(ob)) * cab;
#pragma unroll
          for (int n = N0; n < N1; ++n)
            if (lin2(n) >= 0) acc[n - N0] += cw * S(l_d2(n, sym(a, b)), sq);
        }
      }
    }
  if (W) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (!AX(a).others) continue;
#pragma unroll
      for (int oa = -4; oa <= 4; ++oa) {
        const int c = AX(a).c - oa;
        if (c < 0 || c >= AX(a).n) continue;
        const int ps = p0 - oa * AX(a).ps;
        const double vx = Bx(a, ps);
        const double cl = (vx < 0 ? w_left(oa) : w_right(oa)) * vx * B.lidx[a];
        const double ck2 = w_ko(oa) * (B.cidxd[a] * o.diss);
#pragma unroll
        for (int n = N0; n < N1; ++n) {
          const double wn = (*W)(n, ps);
          acc[n - N0] += (adv_coef(o, n) * cl + ck2) * wn;
        }
      }
    }
  }
#pragma unroll
  for (int n = N0; n < N1; ++n) out(n, acc[n - N0]);
}

}  // namespace z4c
