// FP64 reaction kinetics shared by the chemistry kernels: integer powers,
// Gibbs energies, production rates with the analytic species Jacobian
// (reactions.jl:311-460) and the pivoted dense solve (reactions.jl:463-503).
#pragma once

#include "f3d_internal.h"

namespace f3d {

// exp and log of the chemistry, written with +, -, *, /, explicit fused
// multiply-adds (correctly rounded on any IEEE-754 machine) and exact binary
// scaling only, so that a CPU and this GPU (chem.hip is compiled without
// implicit contraction of a*b+c) produce the same bits; the oracle restates
// them.
// exp: Cody and Waite's argument reduction (W. J. Cody and W. Waite,
// "Software Manual for the Elementary Functions", 1980): x = k ln2 + r with
// |r| <= ln2/2 and ln2 split in a 32-bit head (k * head exact for |k| < 2^11)
// and a tail; exp(r) by its Taylor series to r^13 (remainder < 5e-18
// relative) in Horner form with fma; 2^k by ldexp.  log: x = m 2^e with m in
// [sqrt(1/2), sqrt(2)), log m = 2 atanh(s), s = (m - 1)/(m + 1) (|s| <= 0.172),
// by its series to s^21 (remainder < 1e-18).
__device__ inline double f3d_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return __builtin_inf();
  if (x < -745.1332191019412) return 0.0;
  const double kd = floor(x * 1.4426950408889634 + 0.5);
  const double r = (x - kd * 6.93147180369123816490e-01) - kd * 1.90821492927058770002e-10;
  double p = 1.0 / 6227020800.0;  // 1/13!
  p = fma(p, r, 1.0 / 479001600.0);
  p = fma(p, r, 1.0 / 39916800.0);
  p = fma(p, r, 1.0 / 3628800.0);
  p = fma(p, r, 1.0 / 362880.0);
  p = fma(p, r, 1.0 / 40320.0);
  p = fma(p, r, 1.0 / 5040.0);
  p = fma(p, r, 1.0 / 720.0);
  p = fma(p, r, 1.0 / 120.0);
  p = fma(p, r, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)kd);
}

__device__ inline double f3d_log(double x) {
  if (!(x > 0.0)) return x == 0.0 ? -__builtin_inf() : __builtin_nan("");
  if (x == __builtin_inf()) return x;
  int e = 0;
  double m = frexp(x, &e);  // x = m 2^e, m in [0.5, 1)
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e -= 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  double p = 1.0 / 21.0;
  p = fma(p, z, 1.0 / 19.0);
  p = fma(p, z, 1.0 / 17.0);
  p = fma(p, z, 1.0 / 15.0);
  p = fma(p, z, 1.0 / 13.0);
  p = fma(p, z, 1.0 / 11.0);
  p = fma(p, z, 1.0 / 9.0);
  p = fma(p, z, 1.0 / 7.0);
  p = fma(p, z, 1.0 / 5.0);
  p = fma(p, z, 1.0 / 3.0);
  p = fma(p, z, 1.0);
  const double ed = (double)e;
  return ed * 6.93147180369123816490e-01 + (ed * 1.90821492927058770002e-10 + (2.0 * s) * p);
}

__device__ inline double f3d_log10(double x) { return f3d_log(x) * 0.43429448190325182765; }
__device__ inline double f3d_exp10(double y) { return f3d_exp(y * 2.30258509299404568402); }

__device__ inline double ipow(double x, int n) {
  // Julia Float64^Int for the stoichiometric/order exponents that occur (|n| <= 3)
  if (n == 0) return 1.0;
  if (n < 0) {
    double r = x;
    for (int m = 1; m < -n; ++m) r = r * x;
    return 1.0 / r;
  }
  double r = x;
  for (int m = 1; m < n; ++m) r = r * x;
  return r;
}

// gibbs, thermo.jl:163-191 (g/RT)
template <int NS>
__device__ inline void gibbs(const DevParams* __restrict__ P, double T, double lgT, double invT,
                             double* gi) {
  const double mlogT = 1 - lgT;
  const double T1 = 0.5 * T;
  const double T2 = 1.0 / 6.0 * T * T;
  const double T3 = T1 * T2;
  const double T4 = 1.8 * T2 * T2;
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    const double* c = (T < P->sep[n]) ? P->lo[n] : P->hi[n];
    gi[n] = c[0] * mlogT - c[1] * T1 - c[2] * T2 - c[3] * T3 - c[4] * T4 + c[5] * invT - c[6];
  }
}

// Reaction parameters of the rate evaluation.  KinRT reads them from the
// uniform parameter block (any mechanism: scalar loads and wave-uniform
// branches on the stoichiometry); KinCT<T> takes them from the compile-time
// tables T of a shipped mechanism (f3d_mech_tables.h), so the reaction loop
// unrolls and every stoichiometric test and zero term is resolved by the
// compiler.  Both perform the same arithmetic on the same values.
struct KinRT {
  const DevParams* __restrict__ P;
  __device__ explicit KinRT(const DevParams* p) : P(p) {}
  static constexpr bool kConst = false;
  __device__ int nr() const { return P->Nr; }
  __device__ int rtype(int n) const { return P->rtype[n]; }
  __device__ int sgm(int n) const { return P->sgm[n]; }
  __device__ int vf(int n, int m) const { return P->vf[n][m]; }
  __device__ int vr(int n, int m) const { return P->vr[n][m]; }
  __device__ double Arr(int n, int k) const { return P->Arr[n][k]; }
  __device__ double ef(int n, int m) const { return P->ef[n][m]; }
  __device__ double loP(int n, int k) const { return P->loP[n][k]; }
  __device__ double Troe(int n, int k) const { return P->Troe[n][k]; }
};
template <class T>
struct KinCT {
  __device__ explicit constexpr KinCT(const DevParams*) {}
  static constexpr bool kConst = true;
  __device__ constexpr int nr() const { return T::NR; }
  __device__ constexpr int rtype(int n) const { return T::rtype[n]; }
  __device__ constexpr int sgm(int n) const { return T::sgm[n]; }
  __device__ constexpr int vf(int n, int m) const { return T::vf[n][m]; }
  __device__ constexpr int vr(int n, int m) const { return T::vr[n][m]; }
  __device__ constexpr double Arr(int n, int k) const { return T::Arr[n][k]; }
  __device__ constexpr double ef(int n, int m) const { return T::ef[n][m]; }
  __device__ constexpr double loP(int n, int k) const { return T::loP[n][k]; }
  __device__ constexpr double Troe(int n, int k) const { return T::Troe[n][k]; }
};

// for (n = 0; n < K.nr(); ++n) body(n): unrolled for compile-time tables
template <class K, class F>
__device__ __forceinline__ void for_reactions(const K& Kp, F&& body) {
  if constexpr (K::kConst) {
#pragma unroll
    for (int n = 0; n < Kp.nr(); ++n) {
      body(n);
      // one reaction at a time: interleaving the unrolled reactions would
      // hold several reactions' temporaries next to the Jacobian
      __builtin_amdgcn_sched_barrier(0);
    }
  } else {
    const int Nr = Kp.nr();
    for (int n = 0; n < Nr; ++n) body(n);
  }
}

// vproductionRate(_Jac), reactions.jl:311-460.  A[l][m] accumulates
// Arate[l, m] = d wdot_l / d c_m; the reference neglects the third-body and
// T dependence, FULL (Newton mode) adds the [M] derivative of three-body and
// falloff rates (Lindemann part).
// A is any matrix with at least NS x NS entries (the Newton solve passes its
// (NS+1) x (NS+1) system so the rate Jacobian is built in place).
// TDER (Newton mode, NA = NS + 1): the last column A[l][NS] receives
// d wdot_l / dT at fixed concentrations, analytically: d ln kf/dT =
// (b + (E/R)/T)/T, d ln Kc/dT = sgm/T - sum_m nu_m (h_m/RT)/T (Gibbs-Helmholtz
// on the NASA-7 fits), and for a Troe falloff the derivatives of Pr/(1 + Pr)
// and of F through Fcent(T) and Pr(T).
// RCP (Newton mode): the reverse rate constant as kf (RsT)^-sgm exp(-dg)
// with 1/RsT formed once, instead of kf / Kc per reaction.
template <int NS, bool JAC, bool FULL = false, int NA = NS, bool TDER = false, bool RCP = false,
          class K = KinRT>
__device__ inline void production_rate(const DevParams* __restrict__ P, const double* sc, double T,
                                       double (&A)[NA][NA], double* wdot) {
  const K Kp(P);
  const double lgT = f3d_log(T);
  const double invT = 1.0 / T;
  double gi[NS];
  gibbs<NS>(P, T, lgT, invT, gi);
  double hRT[TDER ? NS : 1];  // h/RT of each species (NASA-7), for d ln Kc / dT
  if (TDER) {
    const double T1 = 0.5 * T;
    const double T2 = 1.0 / 3.0 * T * T;
    const double T3 = 0.25 * T * T * T;
    const double T4 = 0.2 * T * T * T * T;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const double* c = (T < P->sep[n]) ? P->lo[n] : P->hi[n];
      hRT[n] = c[0] + c[1] * T1 + c[2] * T2 + c[3] * T3 + c[4] * T4 + c[5] * invT;
    }
#pragma unroll
    for (int l = 0; l < NS; ++l) A[l][NA - 1] = 0.0;
  }
  const double RsT = P->Ru / P->atm * 1e6 * T;
  const double iRsT = RCP ? 1.0 / RsT : 0.0;
#pragma unroll
  for (int m = 0; m < NS; ++m) {
    wdot[m] = 0.0;
    if (JAC) {
#pragma unroll
      for (int l = 0; l < NS; ++l) A[l][m] = 0.0;
    }
  }
  for_reactions(Kp, [&](const int n) {
    const double kf = Kp.Arr(n, 0) * f3d_exp(Kp.Arr(n, 1) * lgT - Kp.Arr(n, 2) * invT);
    double dg = 0;
#pragma unroll
    for (int m = 0; m < NS; ++m) dg += gi[m] * (double)(Kp.vf(n, m) - Kp.vr(n, m));
    double kr;  // reverse rate constant kf / Kc, Kc = (RsT)^sgm exp(dg)
    if (RCP) {
      const int sg = Kp.sgm(n);
      kr = kf * ((sg > 0 ? ipow(iRsT, sg) : ipow(RsT, -sg)) * f3d_exp(-dg));
    } else {
      kr = kf / (ipow(RsT, Kp.sgm(n)) * f3d_exp(dg));
    }
    double q_f = kf;
    double q_r = kr;
#pragma unroll
    for (int m = 0; m < NS; ++m) {
      const int vf = Kp.vf(n, m), vr = Kp.vr(n, m);
      if (vf != 0) q_f *= ipow(sc[m], vf);
      if (vr != 0) q_r *= ipow(sc[m], vr);
    }
    const int rt = Kp.rtype(n);
    double dlnq_dM = 0.0;
    double fac = 1.0;  // third-body / falloff factor multiplying both directions
    double dlnfac = 0.0;  // d ln fac / dT (TDER)
    if (rt == 2) {
      double mix = 0.0;
#pragma unroll
      for (int m = 0; m < NS; ++m) mix += Kp.ef(n, m) * sc[m];
      q_f *= mix;
      q_r *= mix;
      dlnq_dM = 1.0 / mix;
      fac = mix;
    } else if (rt == 3) {
      double mix = 0.0;
#pragma unroll
      for (int m = 0; m < NS; ++m) mix += Kp.ef(n, m) * sc[m];
      const double redP = mix / kf * Kp.loP(n, 0) * f3d_exp(Kp.loP(n, 1) * lgT - Kp.loP(n, 2) * invT);
      const double logPred = f3d_log10(redP);
      const double A_ = Kp.Troe(n, 0);
      const double e3 = f3d_exp(-T / Kp.Troe(n, 1));
      const double e1 = f3d_exp(-T / Kp.Troe(n, 2));
      const double e2 = f3d_exp(-Kp.Troe(n, 3) * invT);
      const double Fcent = (1 - A_) * e3 + A_ * e1 + e2;
      const double logFcent = f3d_log10(Fcent);
      const double troe_c = -0.4 - 0.67 * logFcent;
      const double troe_n = 0.75 - 1.27 * logFcent;
      const double troe = (troe_c + logPred) / (troe_n - 0.14 * (troe_c + logPred));
      const double F_troe = f3d_exp10(logFcent / (1.0 + troe * troe));
      const double k_1 = (redP / (1.0 + redP)) * F_troe;
      q_f *= k_1;
      q_r *= k_1;
      dlnq_dM = 1.0 / ((1.0 + redP) * mix);
      fac = k_1;
      if (TDER) {
        const double ln10 = 2.30258509299404568402;
        // d ln Pr / dT: k0 and kinf Arrhenius
        const double dlnPr = (Kp.loP(n, 1) - Kp.Arr(n, 1)) * invT +
                             (Kp.loP(n, 2) - Kp.Arr(n, 2)) * invT * invT;
        const double dL = dlnPr / ln10;
        const double dFc = -(1 - A_) / Kp.Troe(n, 1) * e3 - A_ / Kp.Troe(n, 2) * e1 +
                           Kp.Troe(n, 3) * invT * invT * e2;
        const double dLc = dFc / (Fcent * ln10);
        const double u = troe_c + logPred;
        const double v = troe_n - 0.14 * u;
        const double du = -0.67 * dLc + dL;
        const double dv = -1.27 * dLc - 0.14 * du;
        const double dtroe = (du * v - u * dv) / (v * v);
        const double w = 1.0 + troe * troe;
        const double dlog10F = dLc / w - logFcent * (2.0 * troe * dtroe) / (w * w);
        dlnfac = dlnPr / (1.0 + redP) + ln10 * dlog10F;
      }
    }
    if (JAC && FULL && rt != 1) {
      const double dqM = (q_f - q_r) * dlnq_dM;
#pragma unroll
      for (int m = 0; m < NS; ++m) {
        const double e = Kp.ef(n, m);
        if (e == 0.0) continue;
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          const int dl = Kp.vr(n, l) - Kp.vf(n, l);
          if (dl != 0) A[l][m] += dqM * e * (double)dl;
        }
      }
    }
    if (JAC && FULL) {
      // exact mass-action derivatives: dq/dc_m = k nu_m c_m^(nu_m-1) prod_{j!=m} c_j^nu_j
#pragma unroll
      for (int m = 0; m < NS; ++m) {
        const int vfm = Kp.vf(n, m), vrm = Kp.vr(n, m);
        if (vfm == 0 && vrm == 0) continue;
        double dqf = 0.0, dqr = 0.0;
        if (vfm != 0) {
          double pr = kf * (double)vfm * ipow(sc[m], vfm - 1);
#pragma unroll
          for (int j = 0; j < NS; ++j)
            if (j != m && Kp.vf(n, j) != 0) pr *= ipow(sc[j], Kp.vf(n, j));
          dqf = pr;
        }
        if (vrm != 0) {
          double pr = kr * (double)vrm * ipow(sc[m], vrm - 1);
#pragma unroll
          for (int j = 0; j < NS; ++j)
            if (j != m && Kp.vr(n, j) != 0) pr *= ipow(sc[j], Kp.vr(n, j));
          dqr = pr;
        }
        const double dd = (dqf - dqr) * fac;
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          const int dl = Kp.vr(n, l) - Kp.vf(n, l);
          if (dl != 0) A[l][m] += dd * (double)dl;
        }
      }
    }
    const double dq = q_f - q_r;
    if (TDER) {
      double dgs = 0.0;  // sum_m (nu_f - nu_r)_m d(g_m/RT)/dT = -sum_m nu_m h_m/RT / T
#pragma unroll
      for (int m = 0; m < NS; ++m) dgs += hRT[m] * (double)(Kp.vf(n, m) - Kp.vr(n, m));
      const double dlnkf = Kp.Arr(n, 1) * invT + Kp.Arr(n, 2) * invT * invT;
      const double dlnKc = (double)Kp.sgm(n) * invT - dgs * invT;
      const double dqT = q_f * (dlnkf + dlnfac) - q_r * (dlnkf - dlnKc + dlnfac);
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        const int dl = Kp.vr(n, l) - Kp.vf(n, l);
        if (dl != 0) A[l][NA - 1] += dqT * (double)dl;
      }
    }
#pragma unroll
    for (int m = 0; m < NS; ++m) {
      const int vf = Kp.vf(n, m), vr = Kp.vr(n, m);
      if (vf != vr) wdot[m] += dq * (double)(vr - vf);
      if (JAC && !FULL && (vf != 0 || vr != 0)) {
        const double invsc = 1 / (sc[m] + kEps);
        const double Awf = (double)vf * q_f * invsc;
        const double Awr = (double)vr * q_r * invsc;
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          const int dl = Kp.vr(n, l) - Kp.vf(n, l);
          if (dl != 0) A[l][m] += (Awf - Awr) * (double)dl;
        }
      }
    }
  });
}

// Production rates (as production_rate without the Jacobian) together with
// the 3x3 block M[i][j] = d wdot_{ci} / d c_{cj} over three chosen species
// (ci < 0: unused row/column, left zero): exact mass-action derivatives
// (finite at zero concentration) plus the [M] derivative of three-body and
// falloff rates with the falloff factor held fixed (the Lindemann part).
// One pass over the reactions: the transcendentals are shared with wdot.
template <int NS, class K = KinRT>
__device__ inline void production_rate_block3(const DevParams* __restrict__ P, const double* sc,
                                              double T, const int (&ci)[3], double* wdot,
                                              double (&M)[3][3]) {
  const K Kp(P);
  const double lgT = f3d_log(T);
  const double invT = 1.0 / T;
  double gi[NS], cc[3];
  gibbs<NS>(P, T, lgT, invT, gi);
  const double RsT = P->Ru / P->atm * 1e6 * T;
  const double iRsT = 1.0 / RsT;
#pragma unroll
  for (int m = 0; m < NS; ++m) wdot[m] = 0.0;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    double v = 0.0;
#pragma unroll
    for (int m = 0; m < NS; ++m) v += (m == ci[j]) ? sc[m] : 0.0;  // register select, no scratch
    cc[j] = v;
#pragma unroll
    for (int i = 0; i < 3; ++i) M[i][j] = 0.0;
  }
  for_reactions(Kp, [&](const int n) {
    const double kf = Kp.Arr(n, 0) * f3d_exp(Kp.Arr(n, 1) * lgT - Kp.Arr(n, 2) * invT);
    double dg = 0;
#pragma unroll
    for (int m = 0; m < NS; ++m) dg += gi[m] * (double)(Kp.vf(n, m) - Kp.vr(n, m));
    const int sg = Kp.sgm(n);  // kr = kf / Kc, Kc = (RsT)^sgm exp(dg), with 1/RsT formed once
    const double kr = kf * ((sg > 0 ? ipow(iRsT, sg) : ipow(RsT, -sg)) * f3d_exp(-dg));
    // q_f, q_r: the full mass-action products (same order as production_rate,
    // so wdot is bit-identical to it); o_f, o_r: the products over the
    // species that are not chain carriers, shared by the three derivatives
    double q_f = kf, q_r = kr, o_f = kf, o_r = kr;
#pragma unroll
    for (int m = 0; m < NS; ++m) {
      const int vf = Kp.vf(n, m), vr = Kp.vr(n, m);
      const bool carrier = (m == ci[0]) || (m == ci[1]) || (m == ci[2]);
      if (vf != 0) {
        const double t = ipow(sc[m], vf);
        q_f *= t;
        if (!carrier) o_f *= t;
      }
      if (vr != 0) {
        const double t = ipow(sc[m], vr);
        q_r *= t;
        if (!carrier) o_r *= t;
      }
    }
    int vfc[3], vrc[3];  // carrier stoichiometry of this reaction
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      vfc[j] = ci[j] >= 0 ? Kp.vf(n, ci[j]) : 0;
      vrc[j] = ci[j] >= 0 ? Kp.vr(n, ci[j]) : 0;
    }
    const int rt = Kp.rtype(n);
    double dlnq_dM = 0.0;
    double fac = 1.0;
    if (rt == 2) {
      double mix = 0.0;
#pragma unroll
      for (int m = 0; m < NS; ++m) mix += Kp.ef(n, m) * sc[m];
      q_f *= mix;
      q_r *= mix;
      dlnq_dM = 1.0 / mix;
      fac = mix;
    } else if (rt == 3) {
      double mix = 0.0;
#pragma unroll
      for (int m = 0; m < NS; ++m) mix += Kp.ef(n, m) * sc[m];
      const double redP = mix / kf * Kp.loP(n, 0) * f3d_exp(Kp.loP(n, 1) * lgT - Kp.loP(n, 2) * invT);
      const double logPred = f3d_log10(redP);
      const double A_ = Kp.Troe(n, 0);
      const double logFcent = f3d_log10((1 - A_) * f3d_exp(-T / Kp.Troe(n, 1)) +
                                        A_ * f3d_exp(-T / Kp.Troe(n, 2)) +
                                        f3d_exp(-Kp.Troe(n, 3) * invT));
      const double troe_c = -0.4 - 0.67 * logFcent;
      const double troe_n = 0.75 - 1.27 * logFcent;
      const double troe = (troe_c + logPred) / (troe_n - 0.14 * (troe_c + logPred));
      const double F_troe = f3d_exp10(logFcent / (1.0 + troe * troe));
      const double k_1 = (redP / (1.0 + redP)) * F_troe;
      q_f *= k_1;
      q_r *= k_1;
      dlnq_dM = 1.0 / ((1.0 + redP) * mix);
      fac = k_1;
    }
    const double dq = q_f - q_r;
#pragma unroll
    for (int m = 0; m < NS; ++m) {
      const int vf = Kp.vf(n, m), vr = Kp.vr(n, m);
      if (vf != vr) wdot[m] += dq * (double)(vr - vf);
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int cj = ci[j];
      if (cj < 0) continue;
      const int vfj = vfc[j], vrj = vrc[j];
      double d = (rt != 1) ? dq * dlnq_dM * Kp.ef(n, cj) : 0.0;
      if (vfj != 0 || vrj != 0) {
        // d(k c_j^v_j prod_{m != j} c_m^v_m)/dc_j: the non-carrier product
        // times the other two carriers' powers
        double pf = (vfj != 0) ? o_f * (double)vfj * ipow(cc[j], vfj - 1) : 0.0;
        double pr = (vrj != 0) ? o_r * (double)vrj * ipow(cc[j], vrj - 1) : 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          if (k == j) continue;
          if (vfc[k] != 0) pf *= ipow(cc[k], vfc[k]);
          if (vrc[k] != 0) pr *= ipow(cc[k], vrc[k]);
        }
        d += (pf - pr) * fac;
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const int cr = ci[i];
        if (cr < 0) continue;
        const int dl = Kp.vr(n, cr) - Kp.vf(n, cr);
        if (dl != 0) M[i][j] += d * (double)dl;
      }
    }
  });
}

// solve, reactions.jl:463-503 (partial pivoting; row exchanges done with
// selects so the matrix stays in registers, skipped for a column where no
// lane of the wave exchanges rows).  Only the columns >= k that the back
// substitution reads are eliminated; they are bit-identical to the
// reference's full-row update.  RCP (Newton mode): each row's factor is
// -a[i][k] times 1/a[k][k], formed once per column, instead of divided.
template <int N, bool RCP = false>
__device__ inline void gauss_solve(double (&a)[N][N], double (&b)[N], double (&x)[N]) {
#pragma unroll
  for (int k = 0; k < N - 1; ++k) {
    double fmax_ = fabs(a[k][k]);
    int kk = k;
#pragma unroll
    for (int i = k + 1; i < N; ++i) {
      const double absa = fabs(a[i][k]);
      if (absa > fmax_) {
        fmax_ = absa;
        kk = i;
      }
    }
    if (__any(kk != k)) {
#pragma unroll
      for (int i = k + 1; i < N; ++i) {
        const bool sw = (kk == i);
#pragma unroll
        for (int j = k; j < N; ++j) {
          const double t = a[k][j];
          a[k][j] = sw ? a[i][j] : t;
          a[i][j] = sw ? t : a[i][j];
        }
        const double t = b[k];
        b[k] = sw ? b[i] : t;
        b[i] = sw ? t : b[i];
      }
    }
    const double inv = RCP ? 1.0 / a[k][k] : 0.0;
#pragma unroll
    for (int i = k + 1; i < N; ++i) {
      const double f = RCP ? -a[i][k] * inv : -a[i][k] / a[k][k];
#pragma unroll
      for (int j = k + 1; j < N; ++j) a[i][j] += f * a[k][j];
      b[i] += f * b[k];
    }
  }
  x[N - 1] = b[N - 1] / a[N - 1][N - 1];
#pragma unroll
  for (int i = N - 2; i >= 0; --i) {
    double xi = b[i];
#pragma unroll
    for (int j = i + 1; j < N; ++j) xi -= a[i][j] * x[j];
    x[i] = xi / a[i][i];
  }
}

}  // namespace f3d
