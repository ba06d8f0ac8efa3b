// FP64 stiff chemistry (reactions.jl:176-503): production rates with the
// analytic species Jacobian, point-implicit update with a pivoted dense solve
// in registers, and a backward-Euler Newton mode on (rho_k, T).
//
// One thread per interior cell.  The species count is a template parameter so
// the Ns x Ns (or (Ns+1)^2) matrix, the rates and the Gibbs energies live in
// registers with compile-time indices; the reaction loop runs over the
// mechanism held in the uniform parameter block, so its stoichiometry tests
// are wave-uniform branches that skip the zero terms the reference adds.
#include "f3d_internal.h"

namespace f3d {

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

// vproductionRate(_Jac), reactions.jl:311-460.  A[l][m] accumulates
// Arate[l, m] = d wdot_l / d c_m; the reference neglects the third-body and
// T dependence, FULL (Newton mode) adds the [M] derivative of three-body and
// falloff rates (Lindemann part).
template <int NS, bool JAC, bool FULL = false>
__device__ inline void production_rate(const DevParams* __restrict__ P, const double* sc, double T,
                                       double (&A)[NS][NS], double* wdot) {
  const double lgT = log(T);
  const double invT = 1.0 / T;
  double gi[NS];
  gibbs<NS>(P, T, lgT, invT, gi);
  const double RsT = P->Ru / P->atm * 1e6 * T;
#pragma unroll
  for (int m = 0; m < NS; ++m) {
    wdot[m] = 0.0;
    if (JAC) {
#pragma unroll
      for (int l = 0; l < NS; ++l) A[l][m] = 0.0;
    }
  }
  const int Nr = P->Nr;
  for (int n = 0; n < Nr; ++n) {
    const double kf = P->Arr[n][0] * exp(P->Arr[n][1] * lgT - P->Arr[n][2] * invT);
    double dg = 0;
#pragma unroll
    for (int m = 0; m < NS; ++m) dg += gi[m] * (double)(P->vf[n][m] - P->vr[n][m]);
    const double Kc = ipow(RsT, P->sgm[n]) * exp(dg);
    double q_f = kf;
    double q_r = q_f / Kc;
#pragma unroll
    for (int m = 0; m < NS; ++m) {
      const int vf = P->vf[n][m], vr = P->vr[n][m];
      if (vf != 0) q_f *= ipow(sc[m], vf);
      if (vr != 0) q_r *= ipow(sc[m], vr);
    }
    const int rt = P->rtype[n];
    double dlnq_dM = 0.0;
    double fac = 1.0;  // third-body / falloff factor multiplying both directions
    if (rt == 2) {
      double mix = 0.0;
#pragma unroll
      for (int m = 0; m < NS; ++m) mix += P->ef[n][m] * sc[m];
      q_f *= mix;
      q_r *= mix;
      dlnq_dM = 1.0 / mix;
      fac = mix;
    } else if (rt == 3) {
      double mix = 0.0;
#pragma unroll
      for (int m = 0; m < NS; ++m) mix += P->ef[n][m] * sc[m];
      const double redP = mix / kf * P->loP[n][0] * exp(P->loP[n][1] * lgT - P->loP[n][2] * invT);
      const double logPred = log10(redP);
      const double A_ = P->Troe[n][0];
      const double logFcent = log10((1 - A_) * exp(-T / P->Troe[n][1]) +
                                    A_ * exp(-T / P->Troe[n][2]) + exp(-P->Troe[n][3] * invT));
      const double troe_c = -0.4 - 0.67 * logFcent;
      const double troe_n = 0.75 - 1.27 * logFcent;
      const double troe = (troe_c + logPred) / (troe_n - 0.14 * (troe_c + logPred));
      const double F_troe = pow(10.0, logFcent / (1.0 + troe * troe));
      const double k_1 = (redP / (1.0 + redP)) * F_troe;
      q_f *= k_1;
      q_r *= k_1;
      dlnq_dM = 1.0 / ((1.0 + redP) * mix);
      fac = k_1;
    }
    if (JAC && FULL && rt != 1) {
      const double dqM = (q_f - q_r) * dlnq_dM;
#pragma unroll
      for (int m = 0; m < NS; ++m) {
        const double e = P->ef[n][m];
        if (e == 0.0) continue;
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          const int dl = P->vr[n][l] - P->vf[n][l];
          if (dl != 0) A[l][m] += dqM * e * (double)dl;
        }
      }
    }
    if (JAC && FULL) {
      // exact mass-action derivatives: dq/dc_m = k nu_m c_m^(nu_m-1) prod_{j!=m} c_j^nu_j
#pragma unroll
      for (int m = 0; m < NS; ++m) {
        const int vfm = P->vf[n][m], vrm = P->vr[n][m];
        if (vfm == 0 && vrm == 0) continue;
        double dqf = 0.0, dqr = 0.0;
        if (vfm != 0) {
          double pr = kf * (double)vfm * ipow(sc[m], vfm - 1);
#pragma unroll
          for (int j = 0; j < NS; ++j)
            if (j != m && P->vf[n][j] != 0) pr *= ipow(sc[j], P->vf[n][j]);
          dqf = pr;
        }
        if (vrm != 0) {
          double pr = kf / Kc * (double)vrm * ipow(sc[m], vrm - 1);
#pragma unroll
          for (int j = 0; j < NS; ++j)
            if (j != m && P->vr[n][j] != 0) pr *= ipow(sc[j], P->vr[n][j]);
          dqr = pr;
        }
        const double dd = (dqf - dqr) * fac;
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          const int dl = P->vr[n][l] - P->vf[n][l];
          if (dl != 0) A[l][m] += dd * (double)dl;
        }
      }
    }
    const double dq = q_f - q_r;
#pragma unroll
    for (int m = 0; m < NS; ++m) {
      const int vf = P->vf[n][m], vr = P->vr[n][m];
      if (vf != vr) wdot[m] += dq * (double)(vr - vf);
      if (JAC && !FULL && (vf != 0 || vr != 0)) {
        const double invsc = 1 / (sc[m] + kEps);
        const double Awf = (double)vf * q_f * invsc;
        const double Awr = (double)vr * q_r * invsc;
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          const int dl = P->vr[n][l] - P->vf[n][l];
          if (dl != 0) A[l][m] += (Awf - Awr) * (double)dl;
        }
      }
    }
  }
}

// solve, reactions.jl:463-503 (partial pivoting; row exchanges done with
// selects so the matrix stays in registers).  Only the columns >= k that
// the back substitution reads are eliminated; they are bit-identical to the
// reference's full-row update.
template <int N>
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
#pragma unroll
    for (int i = k + 1; i < N; ++i) {
      const double f = -a[i][k] / a[k][k];
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

// common epilogue of eval_gpu / eval_gpu_stiff (reactions.jl:209-229, :286-306)
template <int NS>
__device__ inline void chem_epilogue(const DevParams* __restrict__ P, double* __restrict__ U,
                                     double* __restrict__ Q, double* __restrict__ rhoi, long p,
                                     long NP, double* r, double dei) {
  U[p + 4 * NP] += dei;
  const double rho = fmax(Q[p], kEps);
  double sr = 0.0;
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    r[n] = fmax(r[n], 0.0);
    sr += r[n];
  }
#pragma unroll
  for (int n = 0; n < NS; ++n) r[n] *= rho / sr;
  double ein = Q[p + 6 * NP];
  ein += dei;
  const double T = fmax(get_T<NS>(P, ein, r, nullptr), kEps);
  const double pr = fmax(pmixture<NS>(P, T, r), kEps);
  Q[p + 4 * NP] = pr;
  Q[p + 5 * NP] = T;
  Q[p + 6 * NP] = ein;
#pragma unroll
  for (int n = 0; n < NS; ++n) rhoi[p + n * NP] = r[n];
}

// mode 0: eval_gpu_stiff, 1: eval_gpu, 2: backward-Euler Newton
template <int NS, int MODE>
__global__ void __launch_bounds__(kBlock)
    k_chem(double* __restrict__ U, double* __restrict__ Q, double* __restrict__ rhoi, double dt,
           int maxiter, double rtol, double ttol, int32_t* __restrict__ iters,
           const int64_t* __restrict__ tag, const DevParams* __restrict__ P, Geom g, int bug) {
  const long id = gid();
  if (id >= (long)g.Nx * g.Ny * g.Nz) return;
  const int a = (int)(id % g.Nx);
  const int t = (int)(id / g.Nx);
  const int b = t % g.Ny, c = t / g.Ny;
  const long p = (a + NG) + (long)g.Nxt * ((b + NG) + (long)g.Nyt * (c + NG));
  if (MODE == 2 && iters) iters[id] = 0;
  if (tag && tag[p] != 0) return;
  const long NP = g.NP;
  const double T0 = Q[p + 5 * NP];
  if (!(T0 > P->T_criteria)) return;
  double r[NS], sc[NS], wdot[NS];
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    r[n] = rhoi[p + n * NP];
    sc[n] = r[n] / P->mw[n] * 1e-6;
  }
  if (MODE == 1) {
    double A[NS][NS];
    production_rate<NS, false>(P, sc, T0, A, wdot);
    double dei = 0;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const double drho = wdot[n] * P->mw[n] * 1e6 * dt;
      dei += -P->lo[n][5] * drho * P->Ru / P->mw[n];
      r[n] += drho;
    }
    chem_epilogue<NS>(P, U, Q, rhoi, p, NP, r, dei);
  } else if (MODE == 0) {
    double A[NS][NS], bb[NS], x[NS];
    production_rate<NS, true>(P, sc, T0, A, wdot);
#pragma unroll
    for (int l = 0; l < NS; ++l)
#pragma unroll
      for (int n = 0; n < NS; ++n)
        A[n][l] = (n == l ? 1.0 : 0.0) - A[n][l] * P->mw[n] / P->mw[l] * dt;
#pragma unroll
    for (int n = 0; n < NS; ++n) bb[n] = wdot[n] * P->mw[n] * 1e6 * dt;
    gauss_solve<NS>(A, bb, x);
    double dei = 0;
    const double* lo_flat = &P->lo[0][0];
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const double h6 = bug ? lo_flat[35 + n] : P->lo[n][5];
      dei += -h6 * x[n] * P->Ru / P->mw[n];
      r[n] += x[n];
    }
    chem_epilogue<NS>(P, U, Q, rhoi, p, NP, r, dei);
  } else {
    constexpr int N1 = NS + 1;
    double r0[NS], hf[NS];
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      r0[n] = r[n];
      hf[n] = P->lo[n][5] * P->Ru / P->mw[n];
    }
    double E0 = internal_energy<NS>(P, T0, r0);
#pragma unroll
    for (int n = 0; n < NS; ++n) E0 += hf[n] * r0[n];
    double T = T0;
    int it = 0;
    while (it < maxiter) {
#pragma unroll
      for (int n = 0; n < NS; ++n) sc[n] = r[n] / P->mw[n] * 1e-6;
      double Ar[NS][NS], wdot2[NS];
      production_rate<NS, true, true>(P, sc, T, Ar, wdot);
      const double hT = 1e-6 * T;
      production_rate<NS, false>(P, sc, T + hT, Ar, wdot2);  // Ar untouched (JAC=false)
      double A[N1][N1], rhs[N1], dx[N1];
#pragma unroll
      for (int n = 0; n < NS; ++n) {
        const double src = wdot[n] * P->mw[n] * 1e6 * dt;
        rhs[n] = -(r[n] - r0[n] - src);
#pragma unroll
        for (int l = 0; l < NS; ++l)
          A[n][l] = (n == l ? 1.0 : 0.0) - Ar[n][l] * P->mw[n] / P->mw[l] * dt;
        A[n][NS] = -((wdot2[n] - wdot[n]) / hT) * P->mw[n] * 1e6 * dt;
      }
      double E = internal_energy<NS>(P, T, r);
#pragma unroll
      for (int n = 0; n < NS; ++n) E += hf[n] * r[n];
      rhs[NS] = -(E - E0);
      {
        // ei_specs (thermo.jl:136-160) for dE/drho_l
        const double T2 = 0.5 * T * T;
        const double T3 = 2.0 / 3.0 * T2 * T;
        const double T4 = T2 * T2;
        const double T5 = 0.8 * T4 * T;
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          const double* lo = P->lo[l];
          const double* hc = P->hi[l];
          double e;
          if (T < P->sep[l])
            e = (lo[0] - 1) * T + lo[1] * T2 + lo[2] * T3 + lo[3] * T4 + lo[4] * T5;
          else
            e = (hc[0] - 1) * T + hc[1] * T2 + hc[2] * T3 + hc[3] * T4 + hc[4] * T5 + (hc[5] - lo[5]);
          A[NS][l] = e * P->Ru / P->mw[l] + hf[l];
        }
      }
      A[NS][NS] = cv_mix<NS>(P, T, r);
      gauss_solve<N1>(A, rhs, dx);
      double rho = 0;
#pragma unroll
      for (int n = 0; n < NS; ++n) rho += r[n];
      double mx = 0;
#pragma unroll
      for (int n = 0; n < NS; ++n) {
        r[n] += dx[n];
        mx = fmax(mx, fabs(dx[n]));
      }
      T += dx[NS];
      ++it;
      if (mx <= rtol * rho && fabs(dx[NS]) <= ttol) break;
    }
    if (iters) iters[id] = it;
    double dei = 0;
#pragma unroll
    for (int n = 0; n < NS; ++n) dei += -hf[n] * (r[n] - r0[n]);
    chem_epilogue<NS>(P, U, Q, rhoi, p, NP, r, dei);
  }
}

template <int NS>
static void chem_ns(f3d_ctx* c, int mode, double* U, double* Q, double* rhoi, double dt,
                    int maxiter, double rtol, double ttol, int32_t* iters, const int64_t* tag,
                    hipStream_t s) {
  const long n = (long)c->g.Nx * c->g.Ny * c->g.Nz;
  const int bug = c->bug_compat & 1;
  if (mode == 0)
    hipLaunchKernelGGL((k_chem<NS, 0>), dim3(nblocks(n)), dim3(kBlock), 0, s, U, Q, rhoi, dt,
                       maxiter, rtol, ttol, iters, tag, c->dparams, c->g, bug);
  else if (mode == 1)
    hipLaunchKernelGGL((k_chem<NS, 1>), dim3(nblocks(n)), dim3(kBlock), 0, s, U, Q, rhoi, dt,
                       maxiter, rtol, ttol, iters, tag, c->dparams, c->g, bug);
  else
    hipLaunchKernelGGL((k_chem<NS, 2>), dim3(nblocks(n)), dim3(kBlock), 0, s, U, Q, rhoi, dt,
                       maxiter, rtol, ttol, iters, tag, c->dparams, c->g, bug);
}

int launch_chem(f3d_ctx* c, int mode, double* U, double* Q, double* rhoi, double dt,
                int maxiter, double rtol, double ttol, int32_t* iters, const int64_t* tag,
                hipStream_t s) {
  if (mode == 0 && (c->bug_compat & 1) && c->g.Ns < 6) {
    set_error("bug_compat coeffs_lo[n,6] read needs Ns >= 6 (it reads past the array for Ns < 6)");
    return F3D_ERR_ARG;
  }
  switch (c->g.Ns) {
    case 5: chem_ns<5>(c, mode, U, Q, rhoi, dt, maxiter, rtol, ttol, iters, tag, s); break;
    case 9: chem_ns<9>(c, mode, U, Q, rhoi, dt, maxiter, rtol, ttol, iters, tag, s); break;
    default: set_error("chemistry: Ns must be 5 or 9"); return F3D_ERR_UNSUPPORTED;
  }
  return hip_check(hipGetLastError(), "chem");
}

}  // namespace f3d
