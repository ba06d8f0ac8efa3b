// FP64 CPU oracle: restatement of the reference Julia kernels.
//
// TEST INFRASTRUCTURE ONLY (loaded by tests/, the smoke check and the
// benchmark's CPU-baseline leg).  Every function cites the reference lines it
// follows.  The reference's 1-based indices i (padded) map to 0-based p = i-1;
// a face "i-1/2" of the x sweep (schemes.jl:86) with Julia index i in
// [NG+1, Nx+NG+1] is face f = i-NG-1 in [0, Nx], whose right neighbour cell is
// the padded point ic = NG + f (0-based) and whose left neighbour is ic - 1.
//
// Parity status of each group (how it is pinned) is listed in the tests; see
// tests/test_oracle_*.py.

#include "f3d_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kEps = 2.220446049250313e-16;  // eps(Float64)
constexpr int kNewtonMaxSub = 64;  // most backward-Euler substeps the Newton mode tries

struct Grid {
  int Nx, Ny, Nz, NG, Ns, Nr;
  long Nxt, Nyt, Nzt, NP;
  explicit Grid(const or_dims* d)
      : Nx(d->Nx), Ny(d->Ny), Nz(d->Nz), NG(d->NG), Ns(d->Ns), Nr(d->Nr) {
    Nxt = Nx + 2 * NG;
    Nyt = Ny + 2 * NG;
    Nzt = Nz + 2 * NG;
    NP = Nxt * Nyt * Nzt;
  }
  long P(long i, long j, long k) const { return i + Nxt * (j + Nyt * k); }
};

// face-array geometry per sweep direction: extents (nfx, nfy, nfz)
struct FaceGeom {
  long nx, ny, nz, n;
  FaceGeom(const Grid& g, int dir) {
    nx = g.Nx + (dir == 0);
    ny = g.Ny + (dir == 1);
    nz = g.Nz + (dir == 2);
    n = nx * ny * nz;
  }
  long F(long a, long b, long c) const { return a + nx * (b + ny * c); }
};

inline bool solid(const int64_t* tag, long p) { return tag && tag[p] == 1; }
inline bool not_fluid(const int64_t* tag, long p) { return tag && tag[p] != 0; }

// integer power with the semantics of Julia's Float64^Int for |n| <= 2
inline double ipow(double x, long n) {
  if (n == 0) return 1.0;
  if (n < 0) return 1.0 / ipow(x, -n);
  double r = x;
  for (long m = 1; m < n; ++m) r = r * x;
  return r;
}

// ---------------------------------------------------------------------------
// The chemistry path (thermo, rates, Jacobian, pivoted solve, backward-Euler
// Newton) as templates over the scalar type: instantiated with double for
// the oracle itself and with CountedReal for the operation count of one
// Newton solve (or_newton_opcount), so the count is taken from this very
// operation sequence.
// ---------------------------------------------------------------------------
struct OpTally {
  double add = 0, mul = 0, div = 0, trans = 0, sqrt = 0;
};
OpTally g_tally;

// double that tallies its arithmetic: add/sub, mul, div, and transcendental
// calls (exp, log, log10, pow) in g_tally; comparisons and fabs are free
struct CountedReal {
  double v = 0;
  CountedReal() = default;
  CountedReal(double x) : v(x) {}
  explicit operator double() const { return v; }
  CountedReal& operator+=(CountedReal b) { g_tally.add += 1; v += b.v; return *this; }
  CountedReal& operator-=(CountedReal b) { g_tally.add += 1; v -= b.v; return *this; }
  CountedReal& operator*=(CountedReal b) { g_tally.mul += 1; v *= b.v; return *this; }
  CountedReal& operator/=(CountedReal b) { g_tally.div += 1; v /= b.v; return *this; }
};
using CR = CountedReal;
inline CR operator+(CR a, CR b) { g_tally.add += 1; return CR(a.v + b.v); }
inline CR operator-(CR a, CR b) { g_tally.add += 1; return CR(a.v - b.v); }
inline CR operator*(CR a, CR b) { g_tally.mul += 1; return CR(a.v * b.v); }
inline CR operator/(CR a, CR b) { g_tally.div += 1; return CR(a.v / b.v); }
inline CR operator-(CR a) { return CR(-a.v); }
inline bool operator<(CR a, CR b) { return a.v < b.v; }
inline bool operator>(CR a, CR b) { return a.v > b.v; }
inline bool operator<=(CR a, CR b) { return a.v <= b.v; }
inline bool operator>=(CR a, CR b) { return a.v >= b.v; }
inline bool operator==(CR a, CR b) { return a.v == b.v; }
inline bool operator!=(CR a, CR b) { return a.v != b.v; }
inline CR exp(CR a) { g_tally.trans += 1; return CR(std::exp(a.v)); }
inline CR log(CR a) { g_tally.trans += 1; return CR(std::log(a.v)); }
inline CR log10(CR a) { g_tally.trans += 1; return CR(std::log10(a.v)); }
inline CR pow(CR a, CR b) { g_tally.trans += 1; return CR(std::pow(a.v, b.v)); }
inline CR sqrt(CR a) { g_tally.sqrt += 1; return CR(std::sqrt(a.v)); }
inline CR fabs(CR a) { return CR(std::fabs(a.v)); }
inline double dval(double x) { return x; }
inline double dval(CR x) { return x.v; }

template <class D>
D ipow_t(D x, long n) {
  if (n == 0) return D(1.0);
  if (n < 0) return D(1.0) / ipow_t(x, -n);
  D r = x;
  for (long m = 1; m < n; ++m) r = r * x;
  return r;
}

// cp_specs, thermo.jl:82-104
template <class D>
void cp_specs_t(const Grid& g, const or_thermo* th, D T, D* cp) {
  D T2 = T * T;
  D T3 = T2 * T;
  D T4 = T2 * T2;
  for (int n = 0; n < g.Ns; ++n) {
    const double* c = (T < th->coeffs_sep[n]) ? th->coeffs_lo + 7 * n : th->coeffs_hi + 7 * n;
    cp[n] = c[0] + c[1] * T + c[2] * T2 + c[3] * T3 + c[4] * T4;
  }
}

// ei_specs, thermo.jl:136-160 (e/R per species, K)
template <class D>
void ei_specs_t(const Grid& g, const or_thermo* th, D T, D* ei) {
  D T2 = 0.5 * T * T;
  D T3 = 2.0 / 3.0 * T2 * T;
  D T4 = T2 * T2;
  D T5 = 0.8 * T4 * T;
  for (int n = 0; n < g.Ns; ++n) {
    const double* lo = th->coeffs_lo + 7 * n;
    const double* hc = th->coeffs_hi + 7 * n;
    if (T < th->coeffs_sep[n]) {
      ei[n] = (lo[0] - 1) * T + lo[1] * T2 + lo[2] * T3 + lo[3] * T4 + lo[4] * T5;
    } else {
      ei[n] = (hc[0] - 1) * T + hc[1] * T2 + hc[2] * T3 + hc[3] * T4 + hc[4] * T5 +
              (hc[5] - lo[5]);
    }
  }
}

// gibbs, thermo.jl:163-191 (g/RT)
template <class D>
void gibbs_t(const Grid& g, const or_thermo* th, D T, D lgT, D invT, D* gi) {
  D mlogT = 1 - lgT;
  D T1 = 0.5 * T;
  D T2 = 1.0 / 6.0 * T * T;
  D T3 = T1 * T2;
  D T4 = 1.8 * T2 * T2;
  for (int n = 0; n < g.Ns; ++n) {
    const double* c = (T < th->coeffs_sep[n]) ? th->coeffs_lo + 7 * n : th->coeffs_hi + 7 * n;
    gi[n] = c[0] * mlogT - c[1] * T1 - c[2] * T2 - c[3] * T3 - c[4] * T4 + c[5] * invT - c[6];
  }
}

// CV, thermo.jl:194-203 (J/m^3/K).  RCP: each term times 1/W_n (the
// quotient rounded once) instead of divided by W_n, as the library's Newton
// chemistry and warm-started GetT evaluate it (one rounding differs from the
// reference's order).
template <class D, bool RCP = false>
D cv_t(const Grid& g, const or_thermo* th, D T, const D* rhoi) {
  D cp[64];
  cp_specs_t(g, th, T, cp);
  D result = 0;
  for (int n = 0; n < g.Ns; ++n)
    result += RCP ? (cp[n] - 1) * rhoi[n] * (1.0 / th->mw[n]) : (cp[n] - 1) * rhoi[n] / th->mw[n];
  return result * th->Ru;
}

// InternalEnergy, thermo.jl:219-228 (J/m^3); RCP as cv_t
template <class D, bool RCP = false>
D internal_energy_t(const Grid& g, const or_thermo* th, D T, const D* rhoi) {
  D ei[64];
  ei_specs_t(g, th, T, ei);
  D result = 0;
  for (int n = 0; n < g.Ns; ++n)
    result += RCP ? rhoi[n] * ei[n] * (1.0 / th->mw[n]) : rhoi[n] * ei[n] / th->mw[n];
  return result * th->Ru;
}

// exp and log of the chemistry: the library's f3d_exp / f3d_log
// (flame3d.jl_amd/csrc/f3d_chem.h) restated with the same constants and
// operation order, so that the oracle's chemistry and the GPU's agree bit
// for bit (neither contracts a*b+c implicitly; the polynomials use explicit,
// correctly rounded fma).  exp: Cody and Waite's reduction
// x = k ln2 + r (ln2 split so k * head is exact), Taylor series of exp(r) to
// r^13, ldexp; log: x = m 2^e, m in [sqrt(1/2), sqrt(2)), 2 atanh((m-1)/(m+1))
// by its series to s^21.  Accuracy against libm: tests/.
double dexp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return HUGE_VAL;
  if (x < -745.1332191019412) return 0.0;
  const double kd = std::floor(x * 1.4426950408889634 + 0.5);
  const double r = (x - kd * 6.93147180369123816490e-01) - kd * 1.90821492927058770002e-10;
  double p = 1.0 / 6227020800.0;
  p = std::fma(p, r, 1.0 / 479001600.0);
  p = std::fma(p, r, 1.0 / 39916800.0);
  p = std::fma(p, r, 1.0 / 3628800.0);
  p = std::fma(p, r, 1.0 / 362880.0);
  p = std::fma(p, r, 1.0 / 40320.0);
  p = std::fma(p, r, 1.0 / 5040.0);
  p = std::fma(p, r, 1.0 / 720.0);
  p = std::fma(p, r, 1.0 / 120.0);
  p = std::fma(p, r, 1.0 / 24.0);
  p = std::fma(p, r, 1.0 / 6.0);
  p = std::fma(p, r, 0.5);
  p = std::fma(p, r, 1.0);
  p = std::fma(p, r, 1.0);
  return std::ldexp(p, (int)kd);
}

double dlog(double x) {
  if (!(x > 0.0)) return x == 0.0 ? -HUGE_VAL : std::nan("");
  if (x == HUGE_VAL) return x;
  int e = 0;
  double m = std::frexp(x, &e);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e -= 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  double p = 1.0 / 21.0;
  p = std::fma(p, z, 1.0 / 19.0);
  p = std::fma(p, z, 1.0 / 17.0);
  p = std::fma(p, z, 1.0 / 15.0);
  p = std::fma(p, z, 1.0 / 13.0);
  p = std::fma(p, z, 1.0 / 11.0);
  p = std::fma(p, z, 1.0 / 9.0);
  p = std::fma(p, z, 1.0 / 7.0);
  p = std::fma(p, z, 1.0 / 5.0);
  p = std::fma(p, z, 1.0 / 3.0);
  p = std::fma(p, z, 1.0);
  const double ed = (double)e;
  return ed * 6.93147180369123816490e-01 + (ed * 1.90821492927058770002e-10 + (2.0 * s) * p);
}

inline double xexp(double x) { return dexp(x); }
inline double xlog(double x) { return dlog(x); }
inline double xlog10(double x) { return dlog(x) * 0.43429448190325182765; }
inline double xexp10(double y) { return dexp(y * 2.30258509299404568402); }
// the counting type tallies each as one transcendental
inline CR xexp(CR a) { return exp(a); }
inline CR xlog(CR a) { return log(a); }
inline CR xlog10(CR a) { return log10(a); }
inline CR xexp10(CR a) { return pow(CR(10.0), a); }

// vproductionRate / vproductionRate_Jac, reactions.jl:311-460 (see or_production_rate)
// dwdT (may be null): d wdot / dT at fixed concentrations, analytically
// (d ln kf/dT = (b + (E/R)/T)/T; d ln Kc/dT = sgm/T - sum_m nu_m (h_m/RT)/T by
// Gibbs-Helmholtz on the NASA-7 fits; a Troe falloff adds the derivatives of
// Pr/(1 + Pr) and of F through Fcent(T) and Pr(T)), in the library's order.
// rcp (the library's Newton mode): the reverse rate constant as
// kf (RsT)^-sgm exp(-dg) with 1/RsT formed once, instead of kf / Kc.
template <class D>
void prate_t(const Grid& g, const or_thermo* th, const or_kinetics* kin, const D* sc, D T,
             double jac_eps, int32_t full_jac, D* wdot, D* Arate, D* dwdT = nullptr,
             bool rcp = false) {
  const int Ns = g.Ns, Nr = g.Nr;
  std::vector<D> gi(Ns), kf(Nr), kr(Nr), kfac_den(Nr);
  for (int m = 0; m < Ns; ++m) wdot[m] = 0.0;
  if (Arate)
    for (int m = 0; m < Ns * Ns; ++m) Arate[m] = 0.0;
  D lgT = xlog(T);
  D invT = 1.0 / T;
  for (int n = 0; n < Nr; ++n)
    kf[n] = kin->Arr[3 * n] * xexp(kin->Arr[3 * n + 1] * lgT - kin->Arr[3 * n + 2] * invT);
  gibbs_t<D>(g, th, T, lgT, invT, gi.data());
  std::vector<D> hRT(dwdT ? Ns : 0);
  if (dwdT) {
    const D T1 = 0.5 * T;
    const D T2 = 1.0 / 3.0 * T * T;
    const D T3 = 0.25 * T * T * T;
    const D T4 = 0.2 * T * T * T * T;
    for (int m = 0; m < Ns; ++m) {
      const double* c = (dval(T) < th->coeffs_sep[m]) ? th->coeffs_lo + 7 * m : th->coeffs_hi + 7 * m;
      hRT[m] = c[0] + c[1] * T1 + c[2] * T2 + c[3] * T3 + c[4] * T4 + c[5] * invT;
    }
    for (int l = 0; l < Ns; ++l) dwdT[l] = 0.0;
  }
  D RsT = th->Ru / kin->atm * 1e6 * T;
  D iRsT = rcp ? D(1.0) / RsT : D(0.0);
  for (int n = 0; n < Nr; ++n) {
    D dg = 0;
    for (int m = 0; m < Ns; ++m)
      dg += gi[m] * (double)(kin->vf[m + Ns * n] - kin->vr[m + Ns * n]);
    const long sg = kin->sgm[n];
    if (rcp)
      kr[n] = kf[n] * ((sg > 0 ? ipow_t(iRsT, sg) : ipow_t(RsT, -sg)) * xexp(-dg));
    else
      kr[n] = kf[n] / (ipow_t(RsT, sg) * xexp(dg));  // kf / Kc
  }
  for (int n = 0; n < Nr; ++n) {
    D q_f = kf[n];
    D q_r = kr[n];
    for (int m = 0; m < Ns; ++m) {
      long vf = kin->vf[m + Ns * n], vr = kin->vr[m + Ns * n];
      if (vf != 0) q_f *= ipow_t(sc[m], vf);
      if (vr != 0) q_r *= ipow_t(sc[m], vr);
    }
    D mixture = 0.0;
    D dlnq_dM = 0.0;  // d ln(q) / d[M], used by the full Jacobian only
    D dlnfac = 0.0;   // d ln(third-body / falloff factor) / dT
    kfac_den[n] = 0.0;
    if (kin->reaction_type[n] == 2) {
      for (int m = 0; m < Ns; ++m) mixture += kin->ef[m + Ns * n] * sc[m];
      q_f *= mixture;
      q_r *= mixture;
      dlnq_dM = 1.0 / mixture;
      kfac_den[n] = mixture;
    } else if (kin->reaction_type[n] == 3) {
      for (int m = 0; m < Ns; ++m) mixture += kin->ef[m + Ns * n] * sc[m];
      const double* lp = kin->loP + 3 * n;
      const double* tr = kin->Troe + 4 * n;
      D redP = mixture / kf[n] * lp[0] * xexp(lp[1] * lgT - lp[2] * invT);
      D logPred = xlog10(redP);
      double A = tr[0];
      D e3 = xexp(-T / tr[1]);
      D e1 = xexp(-T / tr[2]);
      D e2 = xexp(-tr[3] * invT);
      D Fcent = (1 - A) * e3 + A * e1 + e2;
      D logFcent = xlog10(Fcent);
      D troe_c = -0.4 - 0.67 * logFcent;
      D troe_n = 0.75 - 1.27 * logFcent;
      D troe = (troe_c + logPred) / (troe_n - 0.14 * (troe_c + logPred));
      D F_troe = xexp10(logFcent / (1.0 + troe * troe));
      D k_1 = (redP / (1.0 + redP)) * F_troe;
      q_f *= k_1;
      q_r *= k_1;
      dlnq_dM = 1.0 / ((1.0 + redP) * mixture);
      kfac_den[n] = k_1;
      if (dwdT) {
        const double ln10 = 2.30258509299404568402;
        D dlnPr = (lp[1] - kin->Arr[3 * n + 1]) * invT + (lp[2] - kin->Arr[3 * n + 2]) * invT * invT;
        D dL = dlnPr / ln10;
        D dFc = -(1 - A) / tr[1] * e3 - A / tr[2] * e1 + tr[3] * invT * invT * e2;
        D dLc = dFc / (Fcent * ln10);
        D u = troe_c + logPred;
        D v = troe_n - 0.14 * u;
        D du = -0.67 * dLc + dL;
        D dv = -1.27 * dLc - 0.14 * du;
        D dtroe = (du * v - u * dv) / (v * v);
        D w = 1.0 + troe * troe;
        D dlog10F = dLc / w - logFcent * (2.0 * troe * dtroe) / (w * w);
        dlnfac = dlnPr / (1.0 + redP) + ln10 * dlog10F;
      }
    }
    if (Arate && full_jac && dlnq_dM != 0.0) {
      // third-body / falloff dependence through [M] = sum ef_m c_m
      // (Lindemann part of the falloff; F_troe's dependence on Pr is neglected)
      const D dq = (q_f - q_r) * dlnq_dM;
      for (int m = 0; m < Ns; ++m) {
        const double e = kin->ef[m + Ns * n];
        if (e == 0.0) continue;
        for (int l = 0; l < Ns; ++l)
          Arate[l + Ns * m] += dq * e * (double)(kin->vr[l + Ns * n] - kin->vf[l + Ns * n]);
      }
    }
    if (Arate && full_jac) {
      // exact concentration derivatives of the mass-action products:
      // dq/dc_m = q0 * nu_m c_m^(nu_m-1) prod_{j != m} c_j^nu_j, q0 = q / prod
      const D kfac = q_f, krac = q_r;  // after the [M] / falloff factors
      for (int m = 0; m < Ns; ++m) {
        long vfm = kin->vf[m + Ns * n], vrm = kin->vr[m + Ns * n];
        if (vfm == 0 && vrm == 0) continue;
        D dqf = 0.0, dqr = 0.0;
        if (vfm != 0) {
          D pr = kf[n] * (double)vfm * ipow_t(sc[m], vfm - 1);
          for (int j = 0; j < Ns; ++j)
            if (j != m && kin->vf[j + Ns * n] != 0) pr *= ipow_t(sc[j], kin->vf[j + Ns * n]);
          dqf = pr;
        }
        if (vrm != 0) {
          D pr = kr[n] * (double)vrm * ipow_t(sc[m], vrm - 1);
          for (int j = 0; j < Ns; ++j)
            if (j != m && kin->vr[j + Ns * n] != 0) pr *= ipow_t(sc[j], kin->vr[j + Ns * n]);
          dqr = pr;
        }
        // the third-body / falloff factor multiplies both directions
        const D fac = (kfac_den[n] != 0.0) ? kfac_den[n] : D(1.0);
        const D dd = (dqf - dqr) * fac;
        for (int l = 0; l < Ns; ++l)
          Arate[l + Ns * m] += dd * (double)(kin->vr[l + Ns * n] - kin->vf[l + Ns * n]);
        (void)kfac;
        (void)krac;
      }
    }
    if (dwdT) {
      D dgs = 0.0;
      for (int m = 0; m < Ns; ++m)
        dgs += hRT[m] * (double)(kin->vf[m + Ns * n] - kin->vr[m + Ns * n]);
      D dlnkf = kin->Arr[3 * n + 1] * invT + kin->Arr[3 * n + 2] * invT * invT;
      D dlnKc = (double)kin->sgm[n] * invT - dgs * invT;
      D dqT = q_f * (dlnkf + dlnfac) - q_r * (dlnkf - dlnKc + dlnfac);
      for (int l = 0; l < Ns; ++l) {
        long dl = kin->vr[l + Ns * n] - kin->vf[l + Ns * n];
        if (dl != 0) dwdT[l] += dqT * (double)dl;
      }
    }
    for (int m = 0; m < Ns; ++m) {
      wdot[m] += (q_f - q_r) * (double)(kin->vr[m + Ns * n] - kin->vf[m + Ns * n]);
      if (Arate && !full_jac) {
        D invsc = 1 / (sc[m] + jac_eps);
        D Awf = (double)kin->vf[m + Ns * n] * q_f * invsc;
        D Awr = (double)kin->vr[m + Ns * n] * q_r * invsc;
        for (int l = 0; l < Ns; ++l)
          Arate[l + Ns * m] += (Awf - Awr) * (double)(kin->vr[l + Ns * n] - kin->vf[l + Ns * n]);
      }
    }
  }
}

// solve, reactions.jl:463-503 (see or_gauss_solve).  rcp (the library's
// Newton mode): each row's factor is -A(i,k) times 1/A(k,k), formed once per
// column, instead of divided by A(k,k).
template <class D>
void gauss_t(int32_t n, D* a, D* b, D* x, bool rcp = false) {
  using std::fabs;
  auto A = [&](int i, int j) -> D& { return a[i + n * j]; };
  for (int k = 0; k < n - 1; ++k) {
    D fmax = fabs(A(k, k));
    int kk = k;
    for (int i = k + 1; i < n; ++i) {
      D absa = fabs(A(i, k));
      if (absa > fmax) {
        fmax = absa;
        kk = i;
      }
    }
    for (int j = k; j < n; ++j) std::swap(A(k, j), A(kk, j));
    std::swap(b[k], b[kk]);
    const D inv = rcp ? D(1.0) / A(k, k) : D(0.0);
    for (int i = k + 1; i < n; ++i) {
      D f = rcp ? -A(i, k) * inv : -A(i, k) / A(k, k);
      for (int j = 0; j < n; ++j) A(i, j) += f * A(k, j);
      b[i] += f * b[k];
    }
  }
  x[n - 1] = b[n - 1] / A(n - 1, n - 1);
  for (int i = n - 2; i >= 0; --i) {
    x[i] = b[i];
    for (int j = i + 1; j < n; ++j) x[i] -= A(i, j) * x[j];
    x[i] /= A(i, i);
  }
}

// backward-Euler Newton solve of or_chem_newton (see be_solve)
template <class D>
bool be_solve_t(const Grid& g, const or_thermo* th, const or_kinetics* kin, D* r, D& T,
                double h, double E0, const double* hf, double rho_tot, double T_lo,
                double T_hi, int maxiter, double rtol, double ttol, int& it) {
  using std::fabs;
  const int Ns = g.Ns;
  const int N1 = Ns + 1;
  D rs[64], s[64], wdot[64], wdot2[64], eis[64], rhs[64], dx[64];
  std::vector<D> Arate(Ns * Ns), A(N1 * N1);
  for (int n = 0; n < Ns; ++n) rs[n] = r[n];
  D dT_prev = 0.0;
  for (int k = 0; k < maxiter; ++k) {
    // the constant quotients 1e-6/W_n, W_n/W_l, Ru/W_n and 1/W_n multiply
    // (each rounded once, as the library precomputes them)
    for (int n = 0; n < Ns; ++n) s[n] = r[n] * (1e-6 / th->mw[n]);
    prate_t<D>(g, th, kin, s, T, kEps, 1, wdot, Arate.data(), wdot2, true);  // wdot2: d wdot / dT
    for (int n = 0; n < Ns; ++n) {
      D src = wdot[n] * th->mw[n] * 1e6 * h;
      rhs[n] = -(r[n] - rs[n] - src);
      for (int l = 0; l < Ns; ++l)
        A[n + N1 * l] = (n == l ? 1.0 : 0.0) - Arate[n + Ns * l] * (th->mw[n] / th->mw[l]) * h;
      A[n + N1 * Ns] = -wdot2[n] * th->mw[n] * 1e6 * h;
    }
    D E = internal_energy_t<D, true>(g, th, T, r);
    for (int n = 0; n < Ns; ++n) E += hf[n] * r[n];
    rhs[Ns] = -(E - E0);
    ei_specs_t<D>(g, th, T, eis);
    for (int l = 0; l < Ns; ++l) A[Ns + N1 * l] = eis[l] * (th->Ru / th->mw[l]) + hf[l];
    A[Ns + N1 * Ns] = cv_t<D, true>(g, th, T, r);
    gauss_t<D>(N1, A.data(), rhs, dx, true);
    D rho = 0;
    for (int n = 0; n < Ns; ++n) rho += r[n];
    double mx = 0;
    bool bad = false;
    for (int n = 0; n < Ns; ++n) {
      r[n] += dx[n];
      mx = std::max(mx, dval(fabs(dx[n])));
      bad = bad || !(r[n] > -rho_tot);
    }
    T += dx[Ns];
    ++it;
    // NASA-7 polynomials jump at T_mid: when E0 falls in such a jump the T
    // update flips sign without contracting; with the species converged the
    // step is accepted (the final T comes from GetT of the energy)
    const bool flip = k > 0 && dval(dx[Ns]) * dval(dT_prev) < 0 &&
                      std::fabs(dval(dx[Ns])) >= 0.5 * std::fabs(dval(dT_prev));
    dT_prev = dx[Ns];
    if (bad || !(T > T_lo && T < T_hi)) return false;
    if (mx <= rtol * dval(rho) && (std::fabs(dval(dx[Ns])) <= ttol || flip)) return true;
  }
  return false;
}

}  // namespace

extern "C" {

int32_t or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// ---------------------------------------------------------------------------
// Thermo (thermo.jl)
// ---------------------------------------------------------------------------

// cp_specs, thermo.jl:82-104
static void cp_specs(const Grid& g, const or_thermo* th, double T, double* cp) {
  cp_specs_t<double>(g, th, T, cp);
}

// h_specs, thermo.jl:107-133 (J/kg, formation enthalpy excluded)
void or_h_specs(const or_dims* d, const or_thermo* th, double T, double* hi) {
  Grid g(d);
  double T2 = 0.5 * T * T;
  double T3 = 2.0 / 3.0 * T2 * T;
  double T4 = T2 * T2;
  double T5 = 0.8 * T4 * T;
  for (int n = 0; n < g.Ns; ++n) {
    const double* lo = th->coeffs_lo + 7 * n;
    const double* hc = th->coeffs_hi + 7 * n;
    if (T < th->coeffs_sep[n]) {
      hi[n] = lo[0] * T + lo[1] * T2 + lo[2] * T3 + lo[3] * T4 + lo[4] * T5;
    } else {
      hi[n] = hc[0] * T + hc[1] * T2 + hc[2] * T3 + hc[3] * T4 + hc[4] * T5 + (hc[5] - lo[5]);
    }
    hi[n] *= th->Ru / th->mw[n];
  }
}

// ei_specs, thermo.jl:136-160 (e/R per species, K)
static void ei_specs(const Grid& g, const or_thermo* th, double T, double* ei) {
  ei_specs_t<double>(g, th, T, ei);
}

// gibbs, thermo.jl:163-191 (g/RT)
static void gibbs_impl(const Grid& g, const or_thermo* th, double T, double lgT, double invT,
                       double* gi) {
  gibbs_t<double>(g, th, T, lgT, invT, gi);
}

double or_exp(double x) { return dexp(x); }
double or_log(double x) { return dlog(x); }

void or_gibbs(const or_dims* d, const or_thermo* th, double T, double* g) {
  Grid gr(d);
  gibbs_impl(gr, th, T, dlog(T), 1.0 / T, g);
}

// CV, thermo.jl:194-203 (J/m^3/K)
double or_cv(const or_dims* d, const or_thermo* th, double T, const double* rhoi) {
  return cv_t<double>(Grid(d), th, T, rhoi);
}

// InternalEnergy, thermo.jl:219-228 (J/m^3)
double or_internal_energy(const or_dims* d, const or_thermo* th, double T, const double* rhoi) {
  return internal_energy_t<double>(Grid(d), th, T, rhoi);
}

// Pmixture, thermo.jl:20-26
static double pmixture(const Grid& g, const or_thermo* th, double T, const double* rhoi) {
  double YOW = 0;
  for (int n = 0; n < g.Ns; ++n) YOW += rhoi[n] / th->mw[n];
  return th->Ru * T * YOW;
}

// GetT, thermo.jl:231-282; the reference fixes tol = 1e-3, maxiter = 30,
// tmin/tmax = thermo.min_temp/max_temp.  iters = Newton updates taken
// (0 for the extrapolation branches).
double or_get_T(const or_dims* d, const or_thermo* th, double ein, const double* rhoi,
                double tol, int32_t maxiter, double tmin, double tmax, int32_t* iters) {
  Grid g(d);
  if (iters) *iters = 0;
  double emin = or_internal_energy(d, th, tmin, rhoi);
  double emax = or_internal_energy(d, th, tmax, rhoi);
  if (ein < emin) {
    double cv = or_cv(d, th, tmin, rhoi);
    return tmin - (emin - ein) / cv;
  }
  if (ein > emax) {
    double cv = or_cv(d, th, tmax, rhoi);
    return tmax - (emax - ein) / cv;
  }
  double As = 0;
  for (int n = 0; n < g.Ns; ++n) As += (th->coeffs_lo[7 * n] - 1) * th->Ru / th->mw[n] * rhoi[n];
  double t1 = ein / As;
  if (t1 < tmin || t1 > tmax) t1 = tmin + (tmax - tmin) / (emax - emin) * (ein - emin);
  int it = 0;
  for (int m = 0; m < maxiter; ++m) {
    double e1 = or_internal_energy(d, th, t1, rhoi);
    double cv = or_cv(d, th, t1, rhoi);
    double dt = (ein - e1) / cv;
    if (dt > 100.0) {
      dt = 100.0;
    } else if (dt < -100.0) {
      dt = -100.0;
    } else if (std::fabs(dt) < tol) {
      break;
    } else if (t1 + dt == t1) {
      break;
    }
    t1 += dt;
    ++it;
  }
  if (iters) *iters = it;
  return t1;
}

// GetT for an energy inside the jump of the NASA-7 fits at their breakpoint
// (no root): or_get_T's iteration, which there alternates between the roots
// of the two fits' continuations until its updates run out; once x_m repeats
// x_{m-2} to 1e-9 K the point x_maxiter is returned by parity (x_m for an
// even number of remaining updates, x_{m-1} for an odd one).  The library's
// get_T_gap, whose updates take the energy and c_v with 1/W_n multiplied
// (cv_t<RCP>); pinned by agreement with or_get_T inside the jump (tests/).
double or_get_T_gap(const or_dims* d, const or_thermo* th, double ein, const double* rhoi,
                    double tol, int32_t maxiter, double tmin, double tmax) {
  Grid g(d);
  double emin = or_internal_energy(d, th, tmin, rhoi);
  double emax = or_internal_energy(d, th, tmax, rhoi);
  if (ein < emin) return tmin - (emin - ein) / or_cv(d, th, tmin, rhoi);
  if (ein > emax) return tmax - (emax - ein) / or_cv(d, th, tmax, rhoi);
  double As = 0;
  for (int n = 0; n < g.Ns; ++n) As += (th->coeffs_lo[7 * n] - 1) * th->Ru / th->mw[n] * rhoi[n];
  double t1 = ein / As;
  if (t1 < tmin || t1 > tmax) t1 = tmin + (tmax - tmin) / (emax - emin) * (ein - emin);
  double x1 = 0.0, x2 = 0.0;
  for (int m = 0; m < maxiter; ++m) {
    double e1 = internal_energy_t<double, true>(g, th, t1, rhoi);
    double cv = cv_t<double, true>(g, th, t1, rhoi);
    double dt = (ein - e1) / cv;
    if (dt > 100.0) {
      dt = 100.0;
    } else if (dt < -100.0) {
      dt = -100.0;
    } else if (std::fabs(dt) < tol) {
      break;
    } else if (t1 + dt == t1) {
      break;
    }
    if (m >= 2 && std::fabs(t1 - x2) <= 1e-9) return ((maxiter - m) % 2 == 0) ? t1 : x1;
    x2 = x1;
    x1 = t1;
    t1 += dt;
  }
  return t1;
}

// GetT warm-started from T_guess (the library's get_T_from, used on the FP32
// stage's c2Prim and by the surrogate step): the Newton update of GetT from
// T_guess (energy and c_v with 1/W_n multiplied, as cv_t<RCP>); a guess or result outside [tmin, tmax] takes or_get_T, an energy
// inside the NASA-7 breakpoint jump (an update that flips sign without
// shrinking) or a result within 0.5 K of a breakpoint or_get_T_gap.  Parity unpinned against the reference (it has no
// warm start); pinned here by agreement with or_get_T to the 1e-3 K
// tolerance, and inside the jump to 1e-6 K (tests/).
double or_get_T_from(const or_dims* d, const or_thermo* th, double ein, const double* rhoi,
                     double T_guess, double tol, int32_t maxiter, double tmin, double tmax) {
  if (!(T_guess >= tmin && T_guess <= tmax))
    return or_get_T(d, th, ein, rhoi, tol, maxiter, tmin, tmax, nullptr);
  double t1 = T_guess;
  bool done = false;
  double dt_prev = 0.0;
  Grid g(d);
  for (int m = 0; m < maxiter; ++m) {
    const double e1 = internal_energy_t<double, true>(g, th, t1, rhoi);
    const double cv = cv_t<double, true>(g, th, t1, rhoi);
    double dt = (ein - e1) / cv;
    if (dt > 100.0) {
      dt = 100.0;
    } else if (dt < -100.0) {
      dt = -100.0;
    } else if (std::fabs(dt) < tol) {
      done = true;
      break;
    } else if (t1 + dt == t1) {
      done = true;
      break;
    } else if (dt * dt_prev < 0.0 && std::fabs(dt) >= 0.5 * std::fabs(dt_prev)) {
      // the update flips sign without shrinking: the energy sits in the
      // jump of the NASA-7 fits at their breakpoint (no exact root)
      return or_get_T_gap(d, th, ein, rhoi, tol, maxiter, tmin, tmax);
    }
    dt_prev = dt;
    t1 += dt;
  }
  if (!done || !(t1 >= tmin && t1 <= tmax)) return or_get_T(d, th, ein, rhoi, tol, maxiter, tmin, tmax, nullptr);
  // within 0.5 K of a breakpoint: the reference's own end point
  for (int n = 0; n < d->Ns; ++n)
    if (std::fabs(t1 - th->coeffs_sep[n]) <= 0.5)
      return or_get_T_gap(d, th, ein, rhoi, tol, maxiter, tmin, tmax);
  return t1;
}

static double getT_ref(const Grid& g, const or_dims* d, const or_thermo* th, double ein,
                       const double* rhoi) {
  (void)g;
  return or_get_T(d, th, ein, rhoi, 1e-3, 30, th->min_temp, th->max_temp, nullptr);
}

// dot5 + mixtureProperties, thermo.jl:284-369
static inline double dot5(double lgT, double lgT2, double lgT3, double lgT4, const double* p) {
  return p[0] + lgT * p[1] + lgT2 * p[2] + lgT3 * p[3] + lgT4 * p[4];
}

// 1: mixture-averaged D_k with the reference's literal numerator 1 - X_k
// (thermo.jl:359-368); 0 (default): sum_{l != k} X_l, the form the library uses
static int g_diff_reference_form = 0;
void or_set_diffusivity_reference_form(int32_t on) { g_diff_reference_form = on ? 1 : 0; }

void or_mixture_properties(const or_dims* d, const or_thermo* th, double T, double P,
                           const double* X, double* Diff, double* lambda_out, double* mu_out) {
  Grid g(d);
  const int Ns = g.Ns;
  std::vector<double> mui(Ns), D(Ns * Ns);
  double sqT = std::sqrt(T);
  double sqsqT = std::sqrt(sqT);
  double lgT = std::log(T);
  double lgT2 = lgT * lgT;
  double lgT3 = lgT * lgT2;
  double lgT4 = lgT2 * lgT2;
  for (int n = 0; n < Ns; ++n) mui[n] = sqT * dot5(lgT, lgT2, lgT3, lgT4, th->conduct_poly + 5 * n);
  double sum1 = 0, sum2 = 0;
  for (int k = 0; k < Ns; ++k) {
    sum1 += X[k] * mui[k];
    sum2 += X[k] / mui[k];
  }
  double lambda = 0.5 * (sum1 + 1 / sum2);
  for (int n = 0; n < Ns; ++n) {
    double sqmui = sqsqT * dot5(lgT, lgT2, lgT3, lgT4, th->visc_poly + 5 * n);
    mui[n] = sqmui * sqmui;
  }
  // Wilke, D[(n-1)*Ns + l] in 1-based flat indexing
  for (int n = 0; n < Ns; ++n) {
    double tmp1 = 1 / th->mw[n];
    double tmp2 = mui[n];
    for (int l = 0; l <= n; ++l) {
      double wratioln = th->mw[l] * tmp1;
      double vrationl = tmp2 / mui[l];
      double factor1 = 1 + std::sqrt(vrationl * std::sqrt(wratioln));
      double tmp = factor1 * factor1 / std::sqrt(8 + 8 / wratioln);
      D[n * Ns + l] = tmp;
      D[l * Ns + n] = tmp / (vrationl * wratioln);
    }
    D[n * Ns + n] = 1.0;
  }
  double mu = 0;
  for (int n = 0; n < Ns; ++n) {
    double tmp = 0.0;
    for (int l = 0; l < Ns; ++l) tmp += X[l] * D[n * Ns + l];
    mu += X[n] * mui[n] / tmp;
  }
  for (int n = 0; n < Ns; ++n) {
    for (int nn = n; nn < Ns; ++nn) {
      double tmp = T * sqT * dot5(lgT, lgT2, lgT3, lgT4, th->binarydiff_poly + 5 * (nn + Ns * n));
      D[nn * Ns + n] = tmp;
      D[n * Ns + nn] = tmp;
    }
  }
  // thermo.jl:359-368 with 1 - X_k evaluated as sum_{l!=k} X_l (equal in
  // exact arithmetic since sum X = 1; the difference cancels catastrophically
  // in nearly pure cells, where it is rounding noise).  g_diff_reference_form
  // selects the reference's literal 1 - X_k, for measuring that difference.
  for (int n = 0; n < Ns; ++n) {
    double s = 0.0, sx = 0.0;
    for (int nn = 0; nn < Ns; ++nn) {
      if (nn == n) continue;
      s += X[nn] / D[n * Ns + nn];
      sx += X[nn];
    }
    s *= P;
    if (g_diff_reference_form) sx = 1 - X[n];
    Diff[n] = sx / (s + kEps);
  }
  *lambda_out = lambda;
  *mu_out = mu;
}

// ---------------------------------------------------------------------------
// Kinetics (reactions.jl)
// ---------------------------------------------------------------------------

// vproductionRate / vproductionRate_Jac, reactions.jl:311-460.  Arate may be
// NULL (explicit path).  jac_eps is eps(Float64) in the reference
// (reactions.jl:451) and 1e-20 in NN/H2/chemical_test.cpp:233.
void or_production_rate(const or_dims* d, const or_thermo* th, const or_kinetics* kin,
                        const double* sc, double T, double jac_eps, int32_t full_jac,
                        double* wdot, double* Arate) {
  prate_t<double>(Grid(d), th, kin, sc, T, jac_eps, full_jac, wdot, Arate);
}

// the rates with their analytic temperature derivative at fixed
// concentrations (the backward-Euler Newton's last Jacobian column)
void or_production_rate_dT(const or_dims* d, const or_thermo* th, const or_kinetics* kin,
                           const double* sc, double T, double* wdot, double* dwdT) {
  prate_t<double>(Grid(d), th, kin, sc, T, kEps, 1, wdot, static_cast<double*>(nullptr), dwdT);
}

// solve, reactions.jl:463-503: Gaussian elimination with partial pivoting,
// a[i + n*j] column-major.
void or_gauss_solve(int32_t n, double* a, double* b, double* x) {
  gauss_t<double>(n, a, b, x);
}

// ---------------------------------------------------------------------------
// Splitting (split.jl)
// ---------------------------------------------------------------------------

// fluxSplit, split.jl:17-89 (all padded points)
void or_flux_split(const or_dims* d, const double* Q, double* Fp, double* Fm, const double* Ax,
                   const double* Ay, const double* Az, const int64_t* tag, const or_scheme* sc) {
  Grid g(d);
  const long NP = g.NP;
  const double SWeps2 = sc->SWeps2;
#pragma omp parallel for schedule(static)
  for (long p = 0; p < NP; ++p) {
    if (solid(tag, p)) continue;
    double rho = Q[p], u = Q[p + NP], v = Q[p + 2 * NP], w = Q[p + 3 * NP];
    double pr = Q[p + 4 * NP], ei = Q[p + 6 * NP];
    double A1 = Ax[p], A2 = Ay[p], A3 = Az[p];
    double gam = pr / ei + 1;
    double c = std::sqrt(gam * pr / rho);
    double ss = std::sqrt(A1 * A1 + A2 * A2 + A3 * A3);
    double E1 = A1 * u + A2 * v + A3 * w;
    double E2 = E1 - c * ss;
    double E3 = E1 + c * ss;
    ss = 1 / ss;
    A1 *= ss;
    A2 *= ss;
    A3 *= ss;
    double E1P = (E1 + std::sqrt(E1 * E1 + SWeps2)) * 0.5;
    double E2P = (E2 + std::sqrt(E2 * E2 + SWeps2)) * 0.5;
    double E3P = (E3 + std::sqrt(E3 * E3 + SWeps2)) * 0.5;
    double E1M = E1 - E1P, E2M = E2 - E2P, E3M = E3 - E3P;
    double uc1 = u - c * A1, uc2 = u + c * A1;
    double vc1 = v - c * A2, vc2 = v + c * A2;
    double wc1 = w - c * A3, wc2 = w + c * A3;
    double vvc1 = (uc1 * uc1 + vc1 * vc1 + wc1 * wc1) * 0.50;
    double vvc2 = (uc2 * uc2 + vc2 * vc2 + wc2 * wc2) * 0.50;
    double vv = (gam - 1.0) * (u * u + v * v + w * w);
    double W2 = (3 - gam) / (2 * (gam - 1)) * (c * c);
    double tmp1 = rho / (2 * gam);
    double tmp2 = 2 * (gam - 1);
    Fp[p] = tmp1 * (tmp2 * E1P + E2P + E3P);
    Fp[p + NP] = tmp1 * (tmp2 * E1P * u + E2P * uc1 + E3P * uc2);
    Fp[p + 2 * NP] = tmp1 * (tmp2 * E1P * v + E2P * vc1 + E3P * vc2);
    Fp[p + 3 * NP] = tmp1 * (tmp2 * E1P * w + E2P * wc1 + E3P * wc2);
    Fp[p + 4 * NP] = tmp1 * (E1P * vv + E2P * vvc1 + E3P * vvc2 + W2 * (E2P + E3P));
    Fm[p] = tmp1 * (tmp2 * E1M + E2M + E3M);
    Fm[p + NP] = tmp1 * (tmp2 * E1M * u + E2M * uc1 + E3M * uc2);
    Fm[p + 2 * NP] = tmp1 * (tmp2 * E1M * v + E2M * vc1 + E3M * vc2);
    Fm[p + 3 * NP] = tmp1 * (tmp2 * E1M * w + E2M * wc1 + E3M * wc2);
    Fm[p + 4 * NP] = tmp1 * (E1M * vv + E2M * vvc1 + E3M * vvc2 + W2 * (E2M + E3M));
  }
}

// split, split.jl:101-130
void or_species_split(const or_dims* d, const double* rhoi, const double* Q, double* Fp,
                      double* Fm, const double* Ax, const double* Ay, const double* Az,
                      const int64_t* tag) {
  Grid g(d);
  const long NP = g.NP;
#pragma omp parallel for schedule(static)
  for (long p = 0; p < NP; ++p) {
    if (solid(tag, p)) continue;
    double u = Q[p + NP], v = Q[p + 2 * NP], w = Q[p + 3 * NP];
    double un = Ax[p] * u + Ay[p] * v + Az[p] * w;
    double Ep = 0.5 * (un + std::fabs(un));
    double Em = 0.5 * (un - std::fabs(un));
    for (int n = 0; n < g.Ns; ++n) {
      Fp[p + n * NP] = Ep * rhoi[p + n * NP];
      Fm[p + n * NP] = Em * rhoi[p + n * NP];
    }
  }
}

// ---------------------------------------------------------------------------
// Reconstruction (schemes.jl)
// ---------------------------------------------------------------------------

// shockSensor, schemes.jl:1-24; range 0-based [1, Ntot-2] in each direction
void or_shock_sensor(const or_dims* d, double* phi, const double* Q) {
  Grid g(d);
  const double* P = Q + 4 * g.NP;
  const long sy = g.Nxt, sz = g.Nxt * g.Nyt;
#pragma omp parallel for schedule(static)
  for (long k = 1; k < g.Nzt - 1; ++k)
    for (long j = 1; j < g.Nyt - 1; ++j)
      for (long i = 1; i < g.Nxt - 1; ++i) {
        long p = g.P(i, j, k);
        double Px1 = P[p - 1], Px2 = P[p], Px3 = P[p + 1];
        double Py1 = P[p - sy], Py2 = P[p], Py3 = P[p + sy];
        double Pz1 = P[p - sz], Pz2 = P[p], Pz3 = P[p + sz];
        double px = std::fabs(-Px1 + 2 * Px2 - Px3) / (Px1 + 2 * Px2 + Px3);
        double py = std::fabs(-Py1 + 2 * Py2 - Py3) / (Py1 + 2 * Py2 + Py3);
        double pz = std::fabs(-Pz1 + 2 * Pz2 - Pz3) / (Pz1 + 2 * Pz2 + Pz3);
        phi[p] = px + py + pz;
      }
}

// Ducros sensor (no counterpart in the reference, which switches on the
// pressure sensor above): theta = (div u)^2 / ((div u)^2 + |curl u|^2 + 1e-30)
// from 2nd-order central differences of u, v, w along xi, eta, zeta and the
// metrics (dxi_a/dx_j = M[3a + j] * J, the J^-1-weighted cofactors of
// parse_mesh.jl:210-221), on the same points as shockSensor.
void or_ducros_sensor(const or_dims* d, double* phi, const double* Q, const double* const* M,
                      const double* J) {
  Grid g(d);
  const long sy = g.Nxt, sz = g.Nxt * g.Nyt;
  const long st[3] = {1, sy, sz};
#pragma omp parallel for schedule(static)
  for (long k = 1; k < g.Nzt - 1; ++k)
    for (long j = 1; j < g.Nyt - 1; ++j)
      for (long i = 1; i < g.Nxt - 1; ++i) {
        const long p = g.P(i, j, k);
        double G[3][3];  // G[c][x] = d u_c / d x_x
        for (int c = 0; c < 3; ++c) {
          const double* u = Q + (1 + c) * g.NP;
          double dxi[3];
          for (int a = 0; a < 3; ++a) dxi[a] = 0.5 * (u[p + st[a]] - u[p - st[a]]);
          for (int x = 0; x < 3; ++x)
            G[c][x] = J[p] * (M[x][p] * dxi[0] + M[3 + x][p] * dxi[1] + M[6 + x][p] * dxi[2]);
        }
        const double dv = G[0][0] + G[1][1] + G[2][2];
        const double wx = G[2][1] - G[1][2], wy = G[0][2] - G[2][0], wz = G[1][0] - G[0][1];
        phi[p] = dv * dv / (dv * dv + wx * wx + wy * wy + wz * wz + 1e-30);
      }
}

static inline double minmod(double a, double b) {
  // schemes.jl:26-28
  return (a * b > 0) ? ((std::fabs(a) > std::fabs(b)) ? b : a) : 0.0;
}

// WENO7-Z reconstruction (no counterpart in the reference, whose middle
// branch is WENO5-Z): the face value from the seven values v[0..6] =
// v_{j-3..j+3} of the upwind cell j (Balsara & Shu 2000 candidates and
// smoothness indicators, linear weights (1, 12, 18, 4)/35; Z weights
// alpha_k = d_k (1 + (tau7 / (eps + beta_k))^2) with tau7 = |b0 + 3 b1 - 3 b2 - b3|,
// Castro, Costa & Don 2011, the reference's WENO5-Z weight form).
double or_weno7z(const double* v, double eps) {
  const double a = v[0], b = v[1], c = v[2], d = v[3], e = v[4], f = v[5], g = v[6];
  // the indicators as quadratic forms in the first differences of each
  // candidate's four values (constants drop out exactly; no cancellation of
  // the values' magnitude), the Balsara-Shu forms times 1/240
  const double D0 = b - a, D1 = c - b, D2 = d - c, D3 = e - d, D4 = f - e, D5 = g - f;
  const double r = 1.0 / 240.0;
  const double b0 = r * (D0 * (547 * D0 - 2788 * D1 + 1854 * D2) + D1 * (3708 * D1 - 5188 * D2) +
                         2107 * D2 * D2);
  const double b1 = r * (D1 * (267 * D1 - 1108 * D2 + 494 * D3) + D2 * (1468 * D2 - 1428 * D3) +
                         547 * D3 * D3);
  const double b2 = r * (D2 * (547 * D2 - 1428 * D3 + 494 * D4) + D3 * (1468 * D3 - 1108 * D4) +
                         267 * D4 * D4);
  const double b3 = r * (D3 * (2107 * D3 - 5188 * D4 + 1854 * D5) + D4 * (3708 * D4 - 2788 * D5) +
                         547 * D5 * D5);
  const double tau = std::fabs(b0 + 3 * b1 - 3 * b2 - b3);
  double t;
  t = tau / (eps + b0);
  const double w0 = (1 + t * t) * 1;
  t = tau / (eps + b1);
  const double w1 = (1 + t * t) * 12;
  t = tau / (eps + b2);
  const double w2 = (1 + t * t) * 18;
  t = tau / (eps + b3);
  const double w3 = (1 + t * t) * 4;
  const double q0 = -3 * a + 13 * b - 23 * c + 25 * d;
  const double q1 = b - 5 * c + 13 * d + 3 * e;
  const double q2 = -c + 7 * d + 7 * e - f;
  const double q3 = 3 * d + 13 * e - 5 * f + g;
  return (w0 * q0 + w1 * q1 + w2 * q2 + w3 * q3) / (12 * (w0 + w1 + w2 + w3));
}

// WENO_x / WENO_y / WENO_z, schemes.jl:86-437.  V(o) = value at offset o
// along the sweep from the face's right cell ic.
static inline double weno_face(const double* fp, const double* fm, long ic, long s, double phif,
                               const or_scheme* sc) {
  auto VP = [&](int o) { return fp[ic + o * s]; };
  auto VM = [&](int o) { return fm[ic + o * s]; };
  const double tmp1 = 13.0 / 12.0, tmp2 = 1.0 / 6.0;
  const double* c = sc->UP7;
  if (phif < sc->phi1) {
    double fpv = c[0] * VP(-4) + c[1] * VP(-3) + c[2] * VP(-2) + c[3] * VP(-1) + c[4] * VP(0) +
                 c[5] * VP(1) + c[6] * VP(2);
    double fmv = c[6] * VM(-3) + c[5] * VM(-2) + c[4] * VM(-1) + c[3] * VM(0) + c[2] * VM(1) +
                 c[1] * VM(2) + c[0] * VM(3);
    return fpv + fmv;
  } else if (phif < sc->phi2 && sc->recon_kind == 1) {
    // upwind cell ic-1 for F+ (v_{j+m} = VP(m-1)), mirrored for F- (v_{j+m} = VM(-m))
    double vp[7], vm[7];
    for (int m = 0; m < 7; ++m) {
      vp[m] = VP(m - 4);
      vm[m] = VM(3 - m);
    }
    return or_weno7z(vp, sc->WENOeps) + or_weno7z(vm, sc->WENOeps);
  } else if (phif < sc->phi2) {
    const double e = sc->WENOeps;
    double V1 = VP(-3), V2 = VP(-2), V3 = VP(-1), V4 = VP(0), V5 = VP(1);
    double t, s11, s22, s33, tau, invsum;
    t = V1 - 2 * V2 + V3;
    double t2 = V1 - 4 * V2 + 3 * V3;
    s11 = tmp1 * (t * t) + 0.25 * (t2 * t2);
    t = V2 - 2 * V3 + V4;
    t2 = V2 - V4;
    s22 = tmp1 * (t * t) + 0.25 * (t2 * t2);
    t = V3 - 2 * V4 + V5;
    t2 = 3 * V3 - 4 * V4 + V5;
    s33 = tmp1 * (t * t) + 0.25 * (t2 * t2);
    tau = std::fabs(s11 - s33);
    t = tau / (e + s11);
    s11 = (1 + t * t) * 1;
    t = tau / (e + s22);
    s22 = (1 + t * t) * 6;
    t = tau / (e + s33);
    s33 = (1 + t * t) * 3;
    invsum = 1 / (s11 + s22 + s33);
    double v1 = 2 * V1 - 7 * V2 + 11 * V3;
    double v2 = -V2 + 5 * V3 + 2 * V4;
    double v3 = 2 * V3 + 5 * V4 - V5;
    double fpv = tmp2 * invsum * (s11 * v1 + s22 * v2 + s33 * v3);

    V1 = VM(-2);
    V2 = VM(-1);
    V3 = VM(0);
    V4 = VM(1);
    V5 = VM(2);
    t = V5 - 2 * V4 + V3;
    t2 = V5 - 4 * V4 + 3 * V3;
    s11 = tmp1 * (t * t) + 0.25 * (t2 * t2);
    t = V2 - 2 * V3 + V4;
    t2 = V4 - V2;
    s22 = tmp1 * (t * t) + 0.25 * (t2 * t2);
    t = V3 - 2 * V2 + V1;
    t2 = 3 * V3 - 4 * V2 + V1;
    s33 = tmp1 * (t * t) + 0.25 * (t2 * t2);
    tau = std::fabs(s11 - s33);
    t = tau / (e + s11);
    s11 = (1 + t * t) * 1;
    t = tau / (e + s22);
    s22 = (1 + t * t) * 6;
    t = tau / (e + s33);
    s33 = (1 + t * t) * 3;
    invsum = 1 / (s11 + s22 + s33);
    v1 = 11 * V3 - 7 * V4 + 2 * V5;
    v2 = -V4 + 5 * V3 + 2 * V2;
    v3 = 2 * V3 + 5 * V2 - V1;
    double fmv = tmp2 * invsum * (s11 * v1 + s22 * v2 + s33 * v3);
    return fpv + fmv;
  } else {
    double fpv = VP(-1) + 0.5 * minmod(VP(0) - VP(-1), VP(-1) - VP(-2));
    double fmv = VM(0) - 0.5 * minmod(VM(1) - VM(0), VM(0) - VM(-1));
    return fpv + fmv;
  }
}

void or_weno(const or_dims* d, int32_t dir, double* F, const double* phi, const double* Fp,
             const double* Fm, int32_t NV, const int64_t* tag, const or_scheme* sc) {
  Grid g(d);
  FaceGeom fg(g, dir);
  const long s = (dir == 0) ? 1 : (dir == 1) ? g.Nxt : g.Nxt * g.Nyt;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < fg.nz; ++c)
    for (long b = 0; b < fg.ny; ++b)
      for (long a = 0; a < fg.nx; ++a) {
        long ic = g.P(a + g.NG, b + g.NG, c + g.NG);  // right cell of face
        if (solid(tag, ic)) continue;
        double phif = std::max(phi[ic - s], phi[ic]);
        long f = fg.F(a, b, c);
        for (int n = 0; n < NV; ++n)
          F[f + n * fg.n] = weno_face(Fp + n * g.NP, Fm + n * g.NP, ic, s, phif, sc);
      }
}

// ---------------------------------------------------------------------------
// Viscous fluxes (viscous.jl)
// ---------------------------------------------------------------------------

// metrics order: dxi/dx, dxi/dy, dxi/dz, deta/dx, deta/dy, deta/dz,
// dzeta/dx, dzeta/dy, dzeta/dz.  For sweep direction dir the face lies
// between ic - s[dir] and ic.  Derivative along the sweep: 4th-order face
// stencil (viscous.jl:31-34); across: averaged 4th-order central
// (viscous.jl:36-52).
static inline double d_normal(const double* q, long ic, long s) {
  const double c12 = 1.0 / 12.0;
  return 1.25 * (q[ic] - q[ic - s]) - c12 * (q[ic + s] - q[ic - 2 * s]);
}
static inline double d_cross4(const double* q, long ic, long s, long t) {
  // face between ic-s and ic; derivative along t
  const double c23 = 2.0 / 3.0, c12 = 1.0 / 12.0;
  return 0.5 * (c23 * (q[ic + t] + q[ic - s + t] - q[ic - t] - q[ic - s - t]) -
                c12 * (q[ic + 2 * t] + q[ic - s + 2 * t] - q[ic - 2 * t] - q[ic - s - 2 * t]));
}
static inline double d_cross2(const double* q, long ic, long s, long t) {
  // viscous.jl:476-479 (2nd order, 0.25 factor)
  return 0.25 * (q[ic + t] + q[ic - s + t] - q[ic - t] - q[ic - s - t]);
}

void or_viscous_flux(const or_dims* d, int32_t dir, double* Fv, const double* Q,
                     const double* const* M, const double* J, const double* lambda,
                     const double* mu, const double* Fh, const int64_t* tag) {
  Grid g(d);
  FaceGeom fg(g, dir);
  const long NP = g.NP;
  const long st[3] = {1, g.Nxt, g.Nxt * g.Nyt};
  const long s = st[dir];
  const double c23 = 2.0 / 3.0;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < fg.nz; ++c)
    for (long b = 0; b < fg.ny; ++b)
      for (long a = 0; a < fg.nx; ++a) {
        long ic = g.P(a + g.NG, b + g.NG, c + g.NG);
        if (solid(tag, ic)) continue;
        long im = ic - s;
        double m[9];
        for (int q = 0; q < 9; ++q) m[q] = (M[q][im] + M[q][ic]) * 0.5;
        double Jac = (J[im] + J[ic]) * 0.5;
        double mui = (mu[im] + mu[ic]) * 0.5;
        double lami = (lambda[im] + lambda[ic]) * 0.5;
        // derivatives of u, v, w, T (Q vars 2, 3, 4, 6) along xi, eta, zeta
        const int qv[4] = {1, 2, 3, 5};
        double dq[4][3];
        for (int q = 0; q < 4; ++q) {
          const double* f = Q + qv[q] * NP;
          for (int t = 0; t < 3; ++t)
            dq[q][t] = (t == dir) ? d_normal(f, ic, s) : d_cross4(f, ic, s, st[t]);
        }
        double u = (Q[im + NP] + Q[ic + NP]) * 0.5;
        double v = (Q[im + 2 * NP] + Q[ic + 2 * NP]) * 0.5;
        double w = (Q[im + 3 * NP] + Q[ic + 3 * NP]) * 0.5;
        // grad = (d/dxi * dxi/dx_j + d/deta * deta/dx_j + d/dzeta * dzeta/dx_j) * Jac
        double gr[4][3];
        for (int q = 0; q < 4; ++q)
          for (int j = 0; j < 3; ++j)
            gr[q][j] = (dq[q][0] * m[j] + dq[q][1] * m[3 + j] + dq[q][2] * m[6 + j]) * Jac;
        double dudx = gr[0][0], dudy = gr[0][1], dudz = gr[0][2];
        double dvdx = gr[1][0], dvdy = gr[1][1], dvdz = gr[1][2];
        double dwdx = gr[2][0], dwdy = gr[2][1], dwdz = gr[2][2];
        double dTdx = gr[3][0], dTdy = gr[3][1], dTdz = gr[3][2];
        double divu = dudx + dvdy + dwdz;
        double t11 = mui * (2 * dudx - c23 * divu);
        double t12 = mui * (dudy + dvdx);
        double t13 = mui * (dudz + dwdx);
        double t22 = mui * (2 * dvdy - c23 * divu);
        double t23 = mui * (dwdy + dvdz);
        double t33 = mui * (2 * dwdz - c23 * divu);
        long f = fg.F(a, b, c);
        double E1 = u * t11 + v * t12 + w * t13 + lami * dTdx + Fh[f];
        double E2 = u * t12 + v * t22 + w * t23 + lami * dTdy + Fh[f + fg.n];
        double E3 = u * t13 + v * t23 + w * t33 + lami * dTdz + Fh[f + 2 * fg.n];
        const double* n3 = m + 3 * dir;  // face-normal metrics
        Fv[f] = n3[0] * t11 + n3[1] * t12 + n3[2] * t13;
        Fv[f + fg.n] = n3[0] * t12 + n3[1] * t22 + n3[2] * t23;
        Fv[f + 2 * fg.n] = n3[0] * t13 + n3[1] * t23 + n3[2] * t33;
        Fv[f + 3 * fg.n] = n3[0] * E1 + n3[1] * E2 + n3[2] * E3;
      }
}

// specViscousFlux_x/y/z, viscous.jl:430-665
void or_spec_viscous_flux(const or_dims* d, int32_t dir, double* Fd, const double* Q,
                          const double* Yi, const double* const* M, const double* J,
                          const double* D, double* Fh, const or_thermo* th,
                          const int64_t* tag) {
  Grid g(d);
  FaceGeom fg(g, dir);
  const long NP = g.NP;
  const int Ns = g.Ns;
  const long st[3] = {1, g.Nxt, g.Nxt * g.Nyt};
  const long s = st[dir];
#pragma omp parallel for schedule(static)
  for (long c = 0; c < fg.nz; ++c)
    for (long b = 0; b < fg.ny; ++b)
      for (long a = 0; a < fg.nx; ++a) {
        long ic = g.P(a + g.NG, b + g.NG, c + g.NG);
        if (solid(tag, ic)) continue;
        long im = ic - s;
        double m[9];
        for (int q = 0; q < 9; ++q) m[q] = (M[q][im] + M[q][ic]) * 0.5;
        double Jac = (J[im] + J[ic]) * 0.5;
        double rho = (Q[im] + Q[ic]) * 0.5;
        double T = (Q[im + 5 * NP] + Q[ic + 5 * NP]) * 0.5;
        double Vk1[64], Vk2[64], Vk3[64], hi[64];
        or_h_specs(d, th, T, hi);
        long f = fg.F(a, b, c);
        double fh1 = 0, fh2 = 0, fh3 = 0;
        double sum1 = 0, sum2 = 0, sum3 = 0;
        for (int n = 0; n < Ns; ++n) {
          const double* y = Yi + n * NP;
          double rhoDi = (D[im + n * NP] + D[ic + n * NP]) * rho;
          double dY[3];
          for (int t = 0; t < 3; ++t) dY[t] = (t == dir) ? d_normal(y, ic, s) : d_cross2(y, ic, s, st[t]);
          double Vx = (dY[0] * m[0] + dY[1] * m[3] + dY[2] * m[6]) * Jac * rhoDi;
          double Vy = (dY[0] * m[1] + dY[1] * m[4] + dY[2] * m[7]) * Jac * rhoDi;
          double Vz = (dY[0] * m[2] + dY[1] * m[5] + dY[2] * m[8]) * Jac * rhoDi;
          Vk1[n] = Vx;
          Vk2[n] = Vy;
          Vk3[n] = Vz;
          sum1 += Vx;
          sum2 += Vy;
          sum3 += Vz;
        }
        const double* n3 = m + 3 * dir;
        for (int n = 0; n < Ns; ++n) {
          double Yn = 0.5 * (Yi[im + n * NP] + Yi[ic + n * NP]);
          double hn = hi[n];
          double V1 = Vk1[n] - sum1 * Yn;
          double V2 = Vk2[n] - sum2 * Yn;
          double V3 = Vk3[n] - sum3 * Yn;
          Fd[f + n * fg.n] = V1 * n3[0] + V2 * n3[1] + V3 * n3[2];
          fh1 += V1 * hn;
          fh2 += V2 * hn;
          fh3 += V3 * hn;
        }
        Fh[f] = fh1;
        Fh[f + fg.n] = fh2;
        Fh[f + 2 * fg.n] = fh3;
      }
}

// ---------------------------------------------------------------------------
// Divergence (div.jl)
// ---------------------------------------------------------------------------

// div, div.jl:16-54
void or_div(const or_dims* d, double* U, const double* Fx, const double* Fy, const double* Fz,
            const double* Fvx, const double* Fvy, const double* Fvz, double dt, const double* J,
            const int64_t* tag) {
  Grid g(d);
  FaceGeom gx(g, 0), gy(g, 1), gz(g, 2);
  const long NP = g.NP;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < g.Nz; ++c)
    for (long b = 0; b < g.Ny; ++b)
      for (long a = 0; a < g.Nx; ++a) {
        long p = g.P(a + g.NG, b + g.NG, c + g.NG);
        if (not_fluid(tag, p)) continue;
        double Jact = J[p] * dt;
        long fx = gx.F(a, b, c), fy = gy.F(a, b, c), fz = gz.F(a, b, c);
        long ox = 1, oy = gy.nx, oz = gz.nx * gz.ny;
        double dV[4];
        for (int q = 0; q < 4; ++q) {
          double v1 = Fvx[fx + ox + q * gx.n] - Fvx[fx + q * gx.n];
          double v2 = Fvy[fy + oy + q * gy.n] - Fvy[fy + q * gy.n];
          double v3 = Fvz[fz + oz + q * gz.n] - Fvz[fz + q * gz.n];
          dV[q] = v1 + v2 + v3;
        }
        for (int n = 0; n < 5; ++n) {
          U[p + n * NP] += (Fx[fx + n * gx.n] - Fx[fx + ox + n * gx.n] + Fy[fy + n * gy.n] -
                            Fy[fy + oy + n * gy.n] + Fz[fz + n * gz.n] - Fz[fz + oz + n * gz.n]) *
                           Jact;
        }
        for (int q = 0; q < 4; ++q) U[p + (q + 1) * NP] += dV[q] * Jact;
      }
}

// divSpecs, div.jl:61-84
void or_div_specs(const or_dims* d, double* rhoi, const double* Fx, const double* Fy,
                  const double* Fz, const double* Fdx, const double* Fdy, const double* Fdz,
                  double dt, const double* J, const int64_t* tag) {
  Grid g(d);
  FaceGeom gx(g, 0), gy(g, 1), gz(g, 2);
  const long NP = g.NP;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < g.Nz; ++c)
    for (long b = 0; b < g.Ny; ++b)
      for (long a = 0; a < g.Nx; ++a) {
        long p = g.P(a + g.NG, b + g.NG, c + g.NG);
        if (not_fluid(tag, p)) continue;
        double Jact = J[p] * dt;
        long fx = gx.F(a, b, c), fy = gy.F(a, b, c), fz = gz.F(a, b, c);
        long ox = 1, oy = gy.nx, oz = gz.nx * gz.ny;
        for (int n = 0; n < g.Ns; ++n) {
          long nx = n * gx.n, ny = n * gy.n, nz = n * gz.n;
          rhoi[p + n * NP] += (Fx[fx + nx] - Fx[fx + ox + nx] + Fy[fy + ny] - Fy[fy + oy + ny] +
                               Fz[fz + nz] - Fz[fz + oz + nz] + Fdx[fx + ox + nx] - Fdx[fx + nx] +
                               Fdy[fy + oy + ny] - Fdy[fy + ny] + Fdz[fz + oz + nz] - Fdz[fz + nz]) *
                              Jact;
        }
      }
}

// ---------------------------------------------------------------------------
// mixture, c2Prim, prim2c, linComb, getY (thermo.jl:371-403, utils.jl)
// ---------------------------------------------------------------------------

void or_mixture(const or_dims* d, const double* Q, const double* rhoi, double* Yi,
                double* lambda, double* mu, double* D, const or_thermo* th, const int64_t* tag) {
  Grid g(d);
  const long NP = g.NP;
  const int Ns = g.Ns;
#pragma omp parallel for schedule(static)
  for (long p = 0; p < NP; ++p) {
    if (solid(tag, p)) continue;
    double T = Q[p + 5 * NP];
    double P = Q[p + 4 * NP];
    double rinv = 1 / std::max(Q[p], kEps);
    double Y1[64], X1[64], Df[64];
    for (int n = 0; n < Ns; ++n) {
      Y1[n] = std::max(rhoi[p + n * NP] * rinv, 0.0);
      Yi[p + n * NP] = Y1[n];
    }
    // Y2X, thermo.jl:38-52
    double YOW = 0;
    for (int n = 0; n < Ns; ++n) YOW += Y1[n] / th->mw[n];
    double YOWINV = 1 / YOW;
    for (int n = 0; n < Ns; ++n) X1[n] = Y1[n] / th->mw[n] * YOWINV;
    double lam, m;
    or_mixture_properties(d, th, T, P, X1, Df, &lam, &m);
    for (int n = 0; n < Ns; ++n) D[p + n * NP] = Df[n];
    lambda[p] = lam;
    mu[p] = m;
  }
}

// c2Prim, utils.jl:2-48 (interior points)
void or_c2prim(const or_dims* d, double* U, double* Q, double* rhoi, const or_thermo* th,
               const int64_t* tag) {
  Grid g(d);
  const long NP = g.NP;
  const int Ns = g.Ns;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < g.Nz; ++c)
    for (long b = 0; b < g.Ny; ++b)
      for (long a = 0; a < g.Nx; ++a) {
        long p = g.P(a + g.NG, b + g.NG, c + g.NG);
        if (not_fluid(tag, p)) continue;
        double rho = std::max(U[p], kEps);
        double rinv = 1 / rho;
        double r[64];
        double sr = 0.0;
        for (int n = 0; n < Ns; ++n) {
          r[n] = std::max(rhoi[p + n * NP], 0.0);
          sr += r[n];
        }
        for (int n = 0; n < Ns; ++n) r[n] *= rho / sr;
        double u = U[p + NP] * rinv;
        double v = U[p + 2 * NP] * rinv;
        double w = U[p + 3 * NP] * rinv;
        double ei = std::max(U[p + 4 * NP] - 0.5 * rho * (u * u + v * v + w * w), kEps);
        double T = std::max(getT_ref(g, d, th, ei, r), kEps);
        double pr = std::max(pmixture(g, th, T, r), kEps);
        for (int n = 0; n < Ns; ++n) rhoi[p + n * NP] = r[n];
        Q[p] = rho;
        Q[p + NP] = u;
        Q[p + 2 * NP] = v;
        Q[p + 3 * NP] = w;
        Q[p + 4 * NP] = pr;
        Q[p + 5 * NP] = T;
        Q[p + 6 * NP] = ei;
      }
}

// prim2c, utils.jl:71-92 (all points)
void or_prim2c(const or_dims* d, double* U, const double* Q) {
  Grid g(d);
  const long NP = g.NP;
#pragma omp parallel for schedule(static)
  for (long p = 0; p < NP; ++p) {
    double rho = Q[p], u = Q[p + NP], v = Q[p + 2 * NP], w = Q[p + 3 * NP], ei = Q[p + 6 * NP];
    U[p] = rho;
    U[p + NP] = u * rho;
    U[p + 2 * NP] = v * rho;
    U[p + 3 * NP] = w * rho;
    U[p + 4 * NP] = ei + 0.5 * rho * (u * u + v * v + w * w);
  }
}

// linComb, utils.jl:95-112 (interior points)
void or_lincomb(const or_dims* d, double* U, const double* Un, int32_t NV, double a, double b,
                const int64_t* tag) {
  Grid g(d);
  const long NP = g.NP;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < g.Nz; ++c)
    for (long bb = 0; bb < g.Ny; ++bb)
      for (long aa = 0; aa < g.Nx; ++aa) {
        long p = g.P(aa + g.NG, bb + g.NG, c + g.NG);
        if (not_fluid(tag, p)) continue;
        for (int n = 0; n < NV; ++n) U[p + n * NP] = U[p + n * NP] * a + Un[p + n * NP] * b;
      }
}

// getY, utils.jl:51-68 (all points)
void or_get_y(const or_dims* d, double* Yi, const double* rhoi, const double* Q,
              const int64_t* tag) {
  Grid g(d);
  const long NP = g.NP;
#pragma omp parallel for schedule(static)
  for (long p = 0; p < NP; ++p) {
    if (not_fluid(tag, p)) continue;
    double rinv = 1 / std::max(Q[p], kEps);
    for (int n = 0; n < g.Ns; ++n) Yi[p + n * NP] = std::max(rhoi[p + n * NP] * rinv, 0.0);
  }
}

// ---------------------------------------------------------------------------
// Boundary fills.  The reference has case-specific fills (boundary.jl:1-191):
// inflow at x-min / z-min (first interior node included), zero-gradient
// copies from the second interior node (j <= NG+1 <- NG+2, j >= N+NG <-
// N+NG-1 in 1-based indexing), x then y then z so edges and corners are
// filled.  Periodic wrap is new (the Taylor-Green case).
// ---------------------------------------------------------------------------
void or_fill(const or_dims* d, const int32_t* kinds, double* Q, double* U, double* rhoi,
             const double* inflow_Q, const double* inflow_rhoi) {
  Grid g(d);
  const long NP = g.NP;
  const int Ns = g.Ns;
  const long N[3] = {g.Nx, g.Ny, g.Nz};
  const long Nt[3] = {g.Nxt, g.Nyt, g.Nzt};
  const long NG = g.NG;
  // inflow conserved state
  double inU[5] = {0, 0, 0, 0, 0};
  if (inflow_Q) {
    double rho = inflow_Q[0], u = inflow_Q[1], v = inflow_Q[2], w = inflow_Q[3], ei = inflow_Q[6];
    inU[0] = rho;
    inU[1] = u * rho;
    inU[2] = v * rho;
    inU[3] = w * rho;
    inU[4] = ei + 0.5 * rho * (u * u + v * v + w * w);
  }
  for (int dir = 0; dir < 3; ++dir) {
    for (int side = 0; side < 2; ++side) {
      int kind = kinds[2 * dir + side];
      if (kind == 0) continue;
#pragma omp parallel for schedule(static)
      for (long k = 0; k < g.Nzt; ++k)
        for (long j = 0; j < g.Nyt; ++j)
          for (long i = 0; i < g.Nxt; ++i) {
            long idx[3] = {i, j, k};
            long x = idx[dir];
            long src[3] = {i, j, k};
            bool apply = false, inflow = false;
            if (kind == 1) {  // periodic: ghosts only
              if (side == 0 && x < NG) {
                src[dir] = x + N[dir];
                apply = true;
              } else if (side == 1 && x >= NG + N[dir]) {
                src[dir] = x - N[dir];
                apply = true;
              }
            } else if (kind == 2) {  // copy incl. first interior node
              if (side == 0 && x <= NG) {
                src[dir] = NG + 1;
                apply = true;
              } else if (side == 1 && x >= NG + N[dir] - 1) {
                src[dir] = NG + N[dir] - 2;
                apply = true;
              }
            } else if (kind == 3) {
              if ((side == 0 && x <= NG) || (side == 1 && x >= NG + N[dir] - 1)) {
                apply = true;
                inflow = true;
              }
            }
            (void)Nt;
            if (!apply) continue;
            long p = g.P(i, j, k);
            if (inflow) {
              for (int n = 0; n < 7; ++n) Q[p + n * NP] = inflow_Q[n];
              for (int n = 0; n < 5; ++n) U[p + n * NP] = inU[n];
              for (int n = 0; n < Ns; ++n) rhoi[p + n * NP] = inflow_rhoi[n];
            } else {
              long q = g.P(src[0], src[1], src[2]);
              for (int n = 0; n < 7; ++n) Q[p + n * NP] = Q[q + n * NP];
              for (int n = 0; n < 5; ++n) U[p + n * NP] = U[q + n * NP];
              for (int n = 0; n < Ns; ++n) rhoi[p + n * NP] = rhoi[q + n * NP];
            }
          }
    }
  }
}

// ---------------------------------------------------------------------------
// Chemistry (reactions.jl:176-309)
// ---------------------------------------------------------------------------

static void renorm_and_T(const Grid& g, const or_dims* d, const or_thermo* th, double* Q,
                         double* r, long p, double dei) {
  const long NP = g.NP;
  double rho = std::max(Q[p], kEps);
  double sr = 0.0;
  for (int n = 0; n < g.Ns; ++n) {
    r[n] = std::max(r[n], 0.0);
    sr += r[n];
  }
  for (int n = 0; n < g.Ns; ++n) r[n] *= rho / sr;
  double ein = Q[p + 6 * NP];
  ein += dei;
  double T = std::max(getT_ref(g, d, th, ein, r), kEps);
  double pr = std::max(pmixture(g, th, T, r), kEps);
  Q[p + 4 * NP] = pr;
  Q[p + 5 * NP] = T;
  Q[p + 6 * NP] = ein;
}

// eval_gpu_stiff, reactions.jl:235-309.  bug_compat & 1 reproduces the
// transposed coeffs_lo[n, 6] read of reactions.jl:282 (valid for Ns >= 6);
// otherwise coeffs_lo[6, n] as in eval_gpu (reactions.jl:205).
void or_chem_stiff(const or_dims* d, double* U, double* Q, double* rhoi, double dt,
                   const or_thermo* th, const or_kinetics* kin, const or_scheme* sc,
                   const int64_t* tag, int32_t bug_compat) {
  Grid g(d);
  const long NP = g.NP;
  const int Ns = g.Ns;
#pragma omp parallel for schedule(dynamic, 64)
  for (long c = 0; c < g.Nz; ++c)
    for (long b = 0; b < g.Ny; ++b)
      for (long a = 0; a < g.Nx; ++a) {
        long p = g.P(a + g.NG, b + g.NG, c + g.NG);
        if (not_fluid(tag, p)) continue;
        double T = Q[p + 5 * NP];
        if (!(T > sc->T_criteria)) continue;
        double s[64], drho[64], drhon[64], wdot[64], r[64];
        std::vector<double> Arate(Ns * Ns), A1(Ns * Ns);
        for (int n = 0; n < Ns; ++n) {
          r[n] = rhoi[p + n * NP];
          s[n] = r[n] / th->mw[n] * 1e-6;
        }
        or_production_rate(d, th, kin, s, T, kEps, 0, wdot, Arate.data());
        for (int l = 0; l < Ns; ++l)
          for (int n = 0; n < Ns; ++n)
            A1[n + Ns * l] = (n == l ? 1.0 : 0.0) - Arate[n + Ns * l] * th->mw[n] / th->mw[l] * dt;
        for (int n = 0; n < Ns; ++n) drho[n] = wdot[n] * th->mw[n] * 1e6 * dt;
        or_gauss_solve(Ns, A1.data(), drho, drhon);
        double dei = 0;
        for (int n = 0; n < Ns; ++n) {
          double h6 = (bug_compat & 1) ? th->coeffs_lo[35 + n] : th->coeffs_lo[5 + 7 * n];
          dei += -h6 * drhon[n] * th->Ru / th->mw[n];
          r[n] += drhon[n];
        }
        U[p + 4 * NP] += dei;
        renorm_and_T(g, d, th, Q, r, p, dei);
        for (int n = 0; n < Ns; ++n) rhoi[p + n * NP] = r[n];
      }
}

// eval_gpu, reactions.jl:176-232
void or_chem_explicit(const or_dims* d, double* U, double* Q, double* rhoi, double dt,
                      const or_thermo* th, const or_kinetics* kin, const or_scheme* sc,
                      const int64_t* tag) {
  Grid g(d);
  const long NP = g.NP;
  const int Ns = g.Ns;
#pragma omp parallel for schedule(dynamic, 64)
  for (long c = 0; c < g.Nz; ++c)
    for (long b = 0; b < g.Ny; ++b)
      for (long a = 0; a < g.Nx; ++a) {
        long p = g.P(a + g.NG, b + g.NG, c + g.NG);
        if (not_fluid(tag, p)) continue;
        double T = Q[p + 5 * NP];
        if (!(T > sc->T_criteria)) continue;
        double s[64], wdot[64], r[64];
        for (int n = 0; n < Ns; ++n) {
          r[n] = rhoi[p + n * NP];
          s[n] = r[n] / th->mw[n] * 1e-6;
        }
        or_production_rate(d, th, kin, s, T, kEps, 0, wdot, nullptr);
        double dei = 0;
        for (int n = 0; n < Ns; ++n) {
          double drho = wdot[n] * th->mw[n] * 1e6 * dt;
          dei += -th->coeffs_lo[5 + 7 * n] * drho * th->Ru / th->mw[n];
          r[n] += drho;
        }
        U[p + 4 * NP] += dei;
        renorm_and_T(g, d, th, Q, r, p, dei);
        for (int n = 0; n < Ns; ++n) rhoi[p + n * NP] = r[n];
      }
}

// Backward-Euler Newton on x = (rho_1..rho_Ns, T) at constant total internal
// energy E = InternalEnergy(T, rho) + sum_k a6_lo,k Ru/W_k rho_k (the
// reference's sensible energy plus the formation term eval_gpu_stiff moves
// into U[5]).  Residual  R_k = rho_k - rho0_k - dt W_k 1e6 wdot_k(c, T),
// R_T = E(x) - E0.  Jacobian: the species block of eval_gpu_stiff
// (I - W_n/W_l J_nl dt, third-body and T derivatives neglected as in
// reactions.jl:448-456), dR_k/dT by a forward difference of wdot in T,
// dR_T/drho_l = (e_l(T) + a6_lo,l) Ru/W_l, dR_T/dT = CV(T, rho).  The first
// iterate from (rho0, T0) reproduces the point-implicit update.  Converged
// when max_k |d rho_k| <= rtol * rho and |dT| <= ttol; iters = updates taken.
// One backward-Euler solve over h of the reactor (rho_k, T) starting at
// (r, T) = (rs, Ts), with the total energy fixed at E0: Newton on
//   F_k = rho_k - rs_k - h W_k wdot_k(rho, T) = 0,  F_E = E(rho, T) - E0 = 0
// with the exact mass-action Jacobian (+ third-body terms) and a forward
// difference in T.  Returns false when the iterate leaves the admissible
// set (non-finite, T outside [T_lo, T_hi], rho_k < -rho_tot) or maxiter
// iterations pass without convergence; *it counts the iterations taken.
static bool be_solve(const Grid& g, const or_dims* d, const or_thermo* th,
                     const or_kinetics* kin, double* r, double& T, double h, double E0,
                     const double* hf, double rho_tot, double T_lo, double T_hi, int maxiter,
                     double rtol, double ttol, int& it) {
  (void)d;
  return be_solve_t<double>(g, th, kin, r, T, h, E0, hf, rho_tot, T_lo, T_hi, maxiter, rtol, ttol,
                            it);
}

void or_chem_newton(const or_dims* d, double* U, double* Q, double* rhoi, double dt,
                    const or_thermo* th, const or_kinetics* kin, const or_scheme* sc,
                    const int64_t* tag, int32_t maxiter, double rtol, double ttol,
                    int32_t* iters) {
  Grid g(d);
  const long NP = g.NP;
  const int Ns = g.Ns;
  const double T_lo = 0.5 * th->min_temp, T_hi = 1.5 * th->max_temp;
#pragma omp parallel for schedule(dynamic, 64)
  for (long c = 0; c < g.Nz; ++c)
    for (long b = 0; b < g.Ny; ++b)
      for (long a = 0; a < g.Nx; ++a) {
        long p = g.P(a + g.NG, b + g.NG, c + g.NG);
        long cell = a + g.Nx * (b + g.Ny * c);
        if (iters) iters[cell] = 0;
        if (not_fluid(tag, p)) continue;
        double T0 = Q[p + 5 * NP];
        if (!(T0 > sc->T_criteria)) continue;
        double r0[64], r[64], hf[64];
        double rho_tot = 0;
        for (int n = 0; n < Ns; ++n) {
          r0[n] = rhoi[p + n * NP];
          hf[n] = th->coeffs_lo[5 + 7 * n] * th->Ru / th->mw[n];
          rho_tot += r0[n];
        }
        double E0 = internal_energy_t<double, true>(g, th, T0, r0);
        for (int n = 0; n < Ns; ++n) E0 += hf[n] * r0[n];
        // one step of dt; on failure restart from the initial state with
        // twice as many backward-Euler substeps (at most kNewtonMaxSub)
        int it = 0;
        bool ok = false;
        double T = T0;
        for (int nsub = 1; nsub <= kNewtonMaxSub && !ok; nsub *= 2) {
          for (int n = 0; n < Ns; ++n) r[n] = r0[n];
          T = T0;
          ok = true;
          const double h = dt / nsub;
          for (int k = 0; k < nsub && ok; ++k)
            ok = be_solve(g, d, th, kin, r, T, h, E0, hf, rho_tot, T_lo, T_hi, maxiter, rtol,
                          ttol, it);
        }
        if (!ok) {  // no admissible solution: the reactor is left unreacted
          for (int n = 0; n < Ns; ++n) r[n] = r0[n];
          it = -it;
        }
        if (iters) iters[cell] = it;
        // energy bookkeeping as eval_gpu_stiff: U[5] and e move by the
        // formation-energy change; T from GetT of the new state
        double dei = 0;
        for (int n = 0; n < Ns; ++n) dei += -hf[n] * (r[n] - r0[n]);
        U[p + 4 * NP] += dei;
        renorm_and_T(g, d, th, Q, r, p, dei);
        for (int n = 0; n < Ns; ++n) rhoi[p + n * NP] = r[n];
      }
}

// ---------------------------------------------------------------------------
// One RK stage, solver.jl:298-326 with mixture/shockSensor/specAdvance/
// flowAdvance (solver.jl:22-56).
// ---------------------------------------------------------------------------
void or_rk3_stage(const or_dims* d, int32_t KRK, double* U, double* Un, double* rhoi,
                  double* rhon, double* Q, double* const* ws, const double* const* M,
                  const double* J, double dt, const int64_t* tag, const or_thermo* th,
                  const or_scheme* sc, const int32_t* fill_kinds, const double* inflow_Q,
                  const double* inflow_rhoi) {
  Grid g(d);
  const long NP = g.NP;
  double *Yi = ws[0], *lam = ws[1], *mu = ws[2], *D = ws[3], *phi = ws[4];
  double *Fp = ws[5], *Fm = ws[6], *Fpi = ws[7], *Fmi = ws[8];
  double *Fx = ws[9], *Fy = ws[10], *Fz = ws[11], *Fvx = ws[12], *Fvy = ws[13], *Fvz = ws[14];
  double *Fxi = ws[15], *Fyi = ws[16], *Fzi = ws[17], *Fdx = ws[18], *Fdy = ws[19], *Fdz = ws[20];
  double *Fhx = ws[21], *Fhy = ws[22], *Fhz = ws[23];
  if (KRK == 1) {
    std::memcpy(Un, U, sizeof(double) * NP * 5);
    std::memcpy(rhon, rhoi, sizeof(double) * NP * g.Ns);
  }
  or_mixture(d, Q, rhoi, Yi, lam, mu, D, th, tag);
  if (sc->sensor_kind == 1)
    or_ducros_sensor(d, phi, Q, M, J);
  else
    or_shock_sensor(d, phi, Q);
  // specAdvance
  or_species_split(d, rhoi, Q, Fpi, Fmi, M[0], M[1], M[2], tag);
  or_weno(d, 0, Fxi, phi, Fpi, Fmi, g.Ns, tag, sc);
  or_species_split(d, rhoi, Q, Fpi, Fmi, M[3], M[4], M[5], tag);
  or_weno(d, 1, Fyi, phi, Fpi, Fmi, g.Ns, tag, sc);
  or_species_split(d, rhoi, Q, Fpi, Fmi, M[6], M[7], M[8], tag);
  or_weno(d, 2, Fzi, phi, Fpi, Fmi, g.Ns, tag, sc);
  or_spec_viscous_flux(d, 0, Fdx, Q, Yi, M, J, D, Fhx, th, tag);
  or_spec_viscous_flux(d, 1, Fdy, Q, Yi, M, J, D, Fhy, th, tag);
  or_spec_viscous_flux(d, 2, Fdz, Q, Yi, M, J, D, Fhz, th, tag);
  or_div_specs(d, rhoi, Fxi, Fyi, Fzi, Fdx, Fdy, Fdz, dt, J, tag);
  // flowAdvance
  or_flux_split(d, Q, Fp, Fm, M[0], M[1], M[2], tag, sc);
  or_weno(d, 0, Fx, phi, Fp, Fm, 5, tag, sc);
  or_flux_split(d, Q, Fp, Fm, M[3], M[4], M[5], tag, sc);
  or_weno(d, 1, Fy, phi, Fp, Fm, 5, tag, sc);
  or_flux_split(d, Q, Fp, Fm, M[6], M[7], M[8], tag, sc);
  or_weno(d, 2, Fz, phi, Fp, Fm, 5, tag, sc);
  or_viscous_flux(d, 0, Fvx, Q, M, J, lam, mu, Fhx, tag);
  or_viscous_flux(d, 1, Fvy, Q, M, J, lam, mu, Fhy, tag);
  or_viscous_flux(d, 2, Fvz, Q, M, J, lam, mu, Fhz, tag);
  or_div(d, U, Fx, Fy, Fz, Fvx, Fvy, Fvz, dt, J, tag);
  if (KRK == 2) {
    or_lincomb(d, U, Un, 5, 0.25, 0.75, tag);
    or_lincomb(d, rhoi, rhon, g.Ns, 0.25, 0.75, tag);
  } else if (KRK == 3) {
    or_lincomb(d, U, Un, 5, 2.0 / 3.0, 1.0 / 3.0, tag);
    or_lincomb(d, rhoi, rhon, g.Ns, 2.0 / 3.0, 1.0 / 3.0, tag);
  }
  or_c2prim(d, U, Q, rhoi, th, tag);
  or_fill(d, fill_kinds, Q, U, rhoi, inflow_Q, inflow_rhoi);
}

// ---------------------------------------------------------------------------
// Operation counts of the chemistry path, from the same templates run on
// CountedReal: out = [add/sub, mul, div, transcendental (exp, log, log10,
// pow), sqrt, iterations].
// ---------------------------------------------------------------------------
static void tally_out(double* out, double iters) {
  out[0] = g_tally.add;
  out[1] = g_tally.mul;
  out[2] = g_tally.div;
  out[3] = g_tally.trans;
  out[4] = g_tally.sqrt;
  out[5] = iters;
}

// One backward-Euler solve over h of one reactor (rho_k, T), as be_solve
// inside or_chem_newton (a single attempt, no substep restart).
void or_newton_opcount(const or_dims* d, const or_thermo* th, const or_kinetics* kin,
                       const double* rhoi, double T0, double h, int32_t maxiter, double rtol,
                       double ttol, double* out) {
  Grid g(d);
  const int Ns = g.Ns;
  double hf[64], rho_tot = 0;
  CR r[64];
  for (int n = 0; n < Ns; ++n) {
    hf[n] = th->coeffs_lo[5 + 7 * n] * th->Ru / th->mw[n];
    rho_tot += rhoi[n];
    r[n] = rhoi[n];
  }
  double E0 = internal_energy_t<double, true>(g, th, T0, rhoi);
  for (int n = 0; n < Ns; ++n) E0 += hf[n] * rhoi[n];
  CR T = T0;
  int it = 0;
  g_tally = OpTally();
  be_solve_t<CR>(g, th, kin, r, T, h, E0, hf, rho_tot, 0.5 * th->min_temp, 1.5 * th->max_temp,
                 maxiter, rtol, ttol, it);
  tally_out(out, it);
}

// One production-rate evaluation (reactions.jl:311-460) as the library's
// Newton mode and surrogate gates run it (reverse rates by the reciprocal
// form): full_jac = 0 rates only, 1 rates and the exact mass-action Jacobian.
void or_rates_opcount(const or_dims* d, const or_thermo* th, const or_kinetics* kin,
                      const double* rhoi, double T, int32_t full_jac, double* out) {
  Grid g(d);
  const int Ns = g.Ns;
  CR s[64], wdot[64];
  std::vector<CR> A(Ns * Ns);
  for (int n = 0; n < Ns; ++n) s[n] = rhoi[n] * (1e-6 / th->mw[n]);
  g_tally = OpTally();
  prate_t<CR>(g, th, kin, s, CR(T), kEps, 1, wdot, full_jac ? A.data() : nullptr, nullptr, true);
  tally_out(out, 1);
}

// One warm-started e -> T Newton update (GetT's update, thermo.jl:231-282, with
// the reciprocal molar masses): InternalEnergy and CV at T, (ein - e)/cv.
void or_getT_update_opcount(const or_dims* d, const or_thermo* th, const double* rhoi, double T,
                            double ein, double* out) {
  Grid g(d);
  CR r[64];
  for (int n = 0; n < g.Ns; ++n) r[n] = rhoi[n];
  g_tally = OpTally();
  CR e1 = internal_energy_t<CR, true>(g, th, CR(T), r);
  CR cv = cv_t<CR, true>(g, th, CR(T), r);
  CR dT = (CR(ein) - e1) / cv;
  (void)dT;
  tally_out(out, 1);
}

// One pivoted solve of an n x n system (reactions.jl:463-503).
void or_gauss_opcount(int32_t n, const double* a, const double* b, double* out) {
  std::vector<CR> A(a, a + (size_t)n * n), B(b, b + n), X(n);
  g_tally = OpTally();
  gauss_t<CR>(n, A.data(), B.data(), X.data());
  tally_out(out, 1);
}

// NN/H2/chemical_test.cpp:950-1007: T from GetT(E), rates and Jacobian,
// I - W_i/W_j J dt, pivoted solve, rho += drho (clip at 0), E -= sum h_f drho.
void or_reactor0d_step(const or_dims* d, const or_thermo* th, const or_kinetics* kin,
                       double* rhoi, double* E, double dt, double jac_eps, double gt_tol,
                       int32_t gt_maxiter, double gt_tmin, double gt_tmax, double* T_out,
                       int32_t* gt_iters) {
  Grid g(d);
  const int Ns = g.Ns;
  double T = or_get_T(d, th, *E, rhoi, gt_tol, gt_maxiter, gt_tmin, gt_tmax, gt_iters);
  double s[64], wdot[64], b[64], x[64];
  std::vector<double> Ar(Ns * Ns), A1(Ns * Ns);
  for (int n = 0; n < Ns; ++n) s[n] = rhoi[n] / th->mw[n] * 1e-6;
  or_production_rate(d, th, kin, s, T, jac_eps, 0, wdot, Ar.data());
  for (int i = 0; i < Ns; ++i)
    for (int j = 0; j < Ns; ++j) {
      A1[i + Ns * j] = (i == j ? 1.0 : 0.0);
      A1[i + Ns * j] -= Ar[i + Ns * j] * th->mw[i] / th->mw[j] * dt;
    }
  for (int n = 0; n < Ns; ++n) b[n] = wdot[n] * th->mw[n] * 1e6 * dt;
  or_gauss_solve(Ns, A1.data(), b, x);
  for (int n = 0; n < Ns; ++n) {
    rhoi[n] += x[n];
    if (rhoi[n] < 0) rhoi[n] = 0;
  }
  for (int n = 0; n < Ns; ++n) *E -= th->coeffs_lo[5 + 7 * n] * th->Ru / th->mw[n] * x[n];
  *T_out = T;
}

}  // extern "C"
