// z4c_rhs_point.hpp -- algebra pass of the Z4c right-hand side.
//
// Pointwise: reads the state at the point and the derivative slots written
// by the stencil pass (coalesced, one buffer offset per thread) and
// evaluates Z4c::Z4cRHS (src/z4c/calculate_z4c_rhs.cpp:29-916) term by term
// in the reference's operation order.  Reorderings (rounding level): the
// three KO terms, which the reference adds to the finished RHS one direction
// at a time (:908-915), arrive pre-summed as (D_x + D_y) + D_z and added to
// the advective sum in the stencil pass (slot LD = c L + D, entering where
// the reference adds L); the Ricci second-derivative term arrives as P_ab.
// Vacuum build and constant shift damping (no Z4C_ETA_CONF /
// Z4C_ETA_TRACK_TP); matter sources enter through `mat` when present.
#pragma once

#include "z4c_device.hpp"

namespace z4c {

struct RhsParams {
  double chi_psi_power, chi_div_floor, diss, damp_kappa1, damp_kappa2;
  double lapse_oplog, lapse_harmonicf, lapse_harmonic, lapse_advect, lapse_K;
  double shift_Gamma, shift_alpha2Gamma, shift_H, shift_advect, shift_eta;
};

// Accessors of one point: state (padded index) and derivative slots
// (interior index).
struct PointIO {
  Buf U;        // state, field stride N
  Buf DV;       // derivative slots, field stride NI
  Buf MT;       // matter (valid if has_mat)
  unsigned vu;  // p*8
  unsigned vd;  // qi*8
  int fu;       // N*8
  int fd;       // NI*8
  bool has_mat;
  __device__ __forceinline__ double u(int n) const { return bload(U, vu, n * fu); }
  __device__ __forceinline__ double d(int slot) const { return bload_dv(DV, vd, slot * fd); }
  __device__ __forceinline__ double m(int n) const { return bload(MT, vu, n * fu); }
};

// Curvature part: cv(0..5) = R~_ab + R^phi_ab, cv(6..11) = D_a D_b alpha.
template <class Out>
__device__ __forceinline__ void rhs_curvature_point(const RhsParams& o, const PointIO& X,
                                                    Out&& cv) {
  const double pw = o.chi_psi_power;
  const double oochi_psi_power = 1.0 / pw;

  // ---------------------------------------------------------------- metric
  double gdd[6], guu[6];
#pragma unroll
  for (int ab = 0; ab < 6; ++ab) gdd[ab] = X.u(V_GXX + ab);
  {
    const double detg = det3(gdd);
    inv3(1.0 / detg, gdd, guu);
  }
  double Gddd[3][6], Gudd[3][6], Gu[3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = a; b < 3; ++b)
        Gddd[c][sym(a, b)] = 0.5 * (X.d(s_d(V_GXX + sym(b, c), a)) +
                                    X.d(s_d(V_GXX + sym(a, c), b)) -
                                    X.d(s_d(V_GXX + sym(a, b), c)));
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int ab = 0; ab < 6; ++ab) {
      double t = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) t += guu[sym(c, d)] * Gddd[d][ab];
      Gudd[c][ab] = t;
    }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double t = 0.0;
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int c = b; c < 3; ++c)
        t += ((b == c) ? 1.0 : 2.0) * guu[sym(b, c)] * Gudd[a][sym(b, c)];
    Gu[a] = t;
  }

  // ------------------------------------------------ conformal Ricci (:357-396)
  double RR[6];  // R~_ab, then + R^phi_ab
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b) {
      double r = 0.0;
#pragma unroll
      for (int c = 0; c < 3; ++c)
        r += 0.5 * (gdd[sym(c, a)] * X.d(s_d(V_GAMX + c, b)) +
                    gdd[sym(c, b)] * X.d(s_d(V_GAMX + c, a)) +
                    Gu[c] * (Gddd[a][sym(b, c)] + Gddd[b][sym(a, c)]));
      // sum_{c<=d} w_cd g~^cd d_c d_d g~_ab, formed by the stencil pass
      r -= X.d(s_lap(V_GXX + sym(a, b)));
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
          for (int e = 0; e < 3; ++e)
            r += guu[sym(c, d)] * (Gudd[e][sym(c, a)] * Gddd[b][sym(e, d)] +
                                   Gudd[e][sym(c, b)] * Gddd[a][sym(e, d)] +
                                   Gudd[e][sym(a, d)] * Gddd[e][sym(c, b)]);
      RR[sym(a, b)] = r;
    }

  // --------------------------------------- conformal factor terms (:402-458)
  const double chi = X.u(V_CHI);
  const double chi_guarded = dmax(chi, o.chi_div_floor);
  const double oochi = 1.0 / chi_guarded;
  const double oopsi4 = (pw == -4.) ? chi_guarded : pow(chi_guarded, -4. / pw);
  double dchi[3], dphi[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    dchi[a] = X.d(s_d(V_CHI, a));
    dphi[a] = dchi[a] * oochi * oochi_psi_power;
  }
  {
    double Ddphi[6];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = a; b < 3; ++b) {
        double t = X.d(s_dd(V_CHI, sym(a, b))) * oochi * oochi_psi_power -
                   pw * dphi[a] * dphi[b];
#pragma unroll
        for (int c = 0; c < 3; ++c) t -= Gudd[c][sym(a, b)] * dphi[c];
        Ddphi[sym(a, b)] = t;
      }
    double trphi = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int d = c; d < 3; ++d)
        trphi += ((c == d) ? 1.0 : 2.0) * guu[sym(c, d)] *
                 (Ddphi[sym(c, d)] + 2. * dphi[c] * dphi[d]);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = a; b < 3; ++b)
        RR[sym(a, b)] = RR[sym(a, b)] + (4. * dphi[a] * dphi[b] - 2. * Ddphi[sym(a, b)] -
                                         2. * gdd[sym(a, b)] * trphi);
  }

  // ------------------------------------------------ D_a D_b alpha (:484-532)
  double dalpha[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) dalpha[a] = X.d(s_d(V_ALPHA, a));
  double Ddalpha_dd[6];
  {
    double Dda = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int d = 0; d < 3; ++d) Dda += guu[sym(c, d)] * dphi[c] * dalpha[d];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = a; b < 3; ++b) {
        double t = X.d(s_dd(V_ALPHA, sym(a, b))) -
                   2. * (dphi[a] * dalpha[b] + dphi[b] * dalpha[a]) +
                   2. * gdd[sym(a, b)] * Dda;
#pragma unroll
        for (int c = 0; c < 3; ++c) t -= Gudd[c][sym(a, b)] * dalpha[c];
        Ddalpha_dd[sym(a, b)] = t;
      }
  }

#pragma unroll
  for (int ab = 0; ab < 6; ++ab) {
    cv(ab, RR[ab]);
    cv(6 + ab, Ddalpha_dd[ab]);
  }
}

// Assembly part: the 22 right-hand sides at the point from the state, the
// derivative slots and the curvature part cv(0..11); out(n, value) stores.
template <class In, class Out>
__device__ __forceinline__ void rhs_assembly_point(const RhsParams& o, const PointIO& X,
                                                   In&& cv, Out&& out) {
  const double pw = o.chi_psi_power;
  const double chi_pow_DA_u = -6.0 / pw;

  // ---------------------------------------------------------------- metric
  double gdd[6], guu[6];
#pragma unroll
  for (int ab = 0; ab < 6; ++ab) gdd[ab] = X.u(V_GXX + ab);
  {
    const double detg = det3(gdd);
    inv3(1.0 / detg, gdd, guu);
  }
  double Gudd[3][6], Gu[3];
  {
  double Gddd[3][6];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = a; b < 3; ++b)
        Gddd[c][sym(a, b)] = 0.5 * (X.d(s_d(V_GXX + sym(b, c), a)) +
                                    X.d(s_d(V_GXX + sym(a, c), b)) -
                                    X.d(s_d(V_GXX + sym(a, b), c)));
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int ab = 0; ab < 6; ++ab) {
      double t = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) t += guu[sym(c, d)] * Gddd[d][ab];
      Gudd[c][ab] = t;
    }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double t = 0.0;
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int c = b; c < 3; ++c)
        t += ((b == c) ? 1.0 : 2.0) * guu[sym(b, c)] * Gudd[a][sym(b, c)];
    Gu[a] = t;
  }

  const double chi = X.u(V_CHI);
  const double chi_guarded = dmax(chi, o.chi_div_floor);
  const double oochi = 1.0 / chi_guarded;
  const double oopsi4 = (pw == -4.) ? chi_guarded : pow(chi_guarded, -4. / pw);
  double dchi[3], dalpha[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    dchi[a] = X.d(s_d(V_CHI, a));
    dalpha[a] = X.d(s_d(V_ALPHA, a));
  }
  double RR[6], Ddalpha_dd[6];
#pragma unroll
  for (int ab = 0; ab < 6; ++ab) {
    RR[ab] = cv(ab);
    Ddalpha_dd[ab] = cv(6 + ab);
  }
  double Ddalpha = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b)
      Ddalpha += ((a == b) ? 1.0 : 2.0) * oopsi4 * guu[sym(a, b)] * Ddalpha_dd[sym(a, b)];

  // ------------------------------------------------------- matter (:466-476)
  double rho = 0.0, Sd[3] = {0.0, 0.0, 0.0}, Sdd[6] = {0, 0, 0, 0, 0, 0};
  if (X.has_mat) {
    rho = X.m(M_RHO);
#pragma unroll
    for (int a = 0; a < 3; ++a) Sd[a] = X.m(M_SX + a);
#pragma unroll
    for (int ab = 0; ab < 6; ++ab) Sdd[ab] = X.m(M_SXX + ab);
  }
  double Str = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b)
      Str += ((a == b) ? 1.0 : 2.0) * oopsi4 * guu[sym(a, b)] * Sdd[sym(a, b)];

  // ------------------------------------------------ A~ contractions (:536-601)
  double Add[6];
#pragma unroll
  for (int ab = 0; ab < 6; ++ab) Add[ab] = X.u(V_AXX + ab);
  double Aud[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      double t = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) t += guu[sym(c, d)] * Add[sym(d, b)];
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
      trAA += ((a == b) ? 1.0 : 2.0) * guu[sym(a, b)] * AA[sym(a, b)];
  double Auu[6];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b) {
      double t = 0.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) t += guu[sym(b, c)] * Aud[a][c];
      Auu[sym(a, b)] = t;
    }
  double DA[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double t = 0.0;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      t -= chi_pow_DA_u * Auu[sym(a, b)] * dchi[b] * oochi;
      t -= (1. / 3.) * guu[sym(a, b)] * (2. * X.d(s_d(V_KHAT, b)) + X.d(s_d(V_THETA, b)));
    }
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int c = b; c < 3; ++c)
        t += ((b == c) ? 1.0 : 2.0) * Gudd[a][sym(b, c)] * Auu[sym(b, c)];
    DA[a] = t;
  }

  // ------------------------------------------------- R, H~, K (:606-624)
  double R = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b)
      R += ((a == b) ? 1.0 : 2.0) * oopsi4 * guu[sym(a, b)] * RR[sym(a, b)];
  const double alpha = X.u(V_ALPHA), Khat = X.u(V_KHAT), Theta = X.u(V_THETA);
  const double K = Khat + 2. * Theta;
  const double Ht = R + (2. / 3.) * (K * K) - trAA;

  // ---------------------------------------- shift derivatives (:630-640)
  double dbeta[3][3];  // [direction][component]
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int a = 0; a < 3; ++a) dbeta[b][a] = X.d(s_d(V_BETAX + a, b));
  double dbeta_div = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) dbeta_div += dbeta[a][a];
  const double k1 = o.damp_kappa1, k2 = o.damp_kappa2;

  // ----------------------------------------- chi, Khat, Theta (:694-711)
  {
    double Lchi = X.d(s_LD(V_CHI));
    Lchi += (1. / 6.) * pw * chi_guarded * dbeta_div;
    const double r = Lchi - (1. / 6.) * pw * chi_guarded * alpha * K;
    out(V_CHI, r);
  }
  {
    double r = -Ddalpha + alpha * (trAA + (1. / 3.) * (K * K)) + X.d(s_LD(V_KHAT)) +
               k1 * (1 - k2) * alpha * Theta;
    r += 4 * M_PI * alpha * (Str + rho);
    out(V_KHAT, r);
  }
  {
    double r = X.d(s_LD(V_THETA)) + alpha * (0.5 * Ht - (2. + k2) * k1 * Theta);
    r -= 8. * M_PI * alpha * rho;
    out(V_THETA, r);
  }

  // -------------------------------------------------- Gamma~ (:640-729)
  {
    double ddbeta_d[3];  // (1/3) sum_b d_a d_b beta^b
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double t = 0.0;
#pragma unroll
      for (int b = 0; b < 3; ++b) t += (1. / 3.) * X.d(s_dd(V_BETAX + b, sym(a, b)));
      ddbeta_d[a] = t;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double LGam = X.d(s_LD(V_GAMX + a));
      LGam += (2. / 3.) * Gu[a] * dbeta_div;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        LGam += guu[sym(a, b)] * ddbeta_d[b] - Gu[b] * dbeta[b][a];
#pragma unroll
        for (int c = b; c < 3; ++c)
          LGam += ((b == c) ? 1.0 : 2.0) * guu[sym(b, c)] * X.d(s_dd(V_BETAX + a, sym(b, c)));
      }
      double t = 2. * alpha * DA[a] + LGam;
      t -= 2. * alpha * k1 * (X.u(V_GAMX + a) - Gu[a]);
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        t -= 2. * Auu[sym(a, b)] * dalpha[b];
        t -= 16. * M_PI * alpha * guu[sym(a, b)] * Sd[b];
      }
      out(V_GAMX + a, t);
    }
  }

  // ------------------------------------------- g~ and A~ (:671-754)
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b) {
      const int ab = sym(a, b);
      double Lg = X.d(s_LD(V_GXX + ab));
      double LA = X.d(s_LD(V_AXX + ab));
      Lg -= (2. / 3.) * gdd[ab] * dbeta_div;
      LA -= (2. / 3.) * Add[ab] * dbeta_div;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Lg += dbeta[a][c] * gdd[sym(b, c)];
        LA += dbeta[b][c] * Add[sym(a, c)];
        Lg += dbeta[b][c] * gdd[sym(a, c)];
        LA += dbeta[a][c] * Add[sym(b, c)];
      }
      const double r_g = -2. * alpha * Add[ab] + Lg;
      double t = oopsi4 * (-Ddalpha_dd[ab] + alpha * RR[ab]);
      t -= (1. / 3.) * gdd[ab] * (-Ddalpha + alpha * R);
      t += alpha * (K * Add[ab] - 2. * AA[ab]);
      t += LA;
      t -= 8. * M_PI * alpha * (oopsi4 * Sdd[ab] - (1. / 3.) * Str * gdd[ab]);
      out(V_GXX + ab, r_g);
      out(V_AXX + ab, t);
    }

  // ------------------------------------------- lapse and shift (:757-902)
  {
    const double f = o.lapse_oplog * o.lapse_harmonicf + o.lapse_harmonic * alpha;
    const double r = (X.d(s_LD(V_ALPHA)) -
                      f * alpha * (Khat + 2 * o.lapse_K * Theta));
    out(V_ALPHA, r);
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double Gam_a = X.u(V_GAMX + a);
    double r = o.shift_Gamma * Gam_a + X.d(s_LD(V_BETAX + a));
    if (fabs(o.shift_alpha2Gamma) > 0.0) r += o.shift_alpha2Gamma * (alpha * alpha) * Gam_a;
    if (fabs(o.shift_H) > 0.0) {
#pragma unroll
      for (int b = 0; b < 3; ++b)
        r += o.shift_H * alpha * chi_guarded * (0.5 * alpha * dchi[b] - dalpha[b]) *
             guu[sym(a, b)];
    }
    r -= o.shift_eta * X.u(V_BETAX + a);
    out(V_BETAX + a, r);
  }
}

}  // namespace z4c
