// FP64 stiff chemistry (reactions.jl:176-503): production rates with the
// analytic species Jacobian, point-implicit update with a pivoted dense solve
// in registers, and a backward-Euler Newton mode on (rho_k, T).
//
// One thread per interior cell.  The species count is a template parameter so
// the Ns x Ns (or (Ns+1)^2) matrix, the rates and the Gibbs energies live in
// registers with compile-time indices; the reaction loop runs over the
// mechanism held in the uniform parameter block, so its stoichiometry tests
// are wave-uniform branches that skip the zero terms the reference adds.
#include "f3d_chem.h"
#include "f3d_internal.h"
#include "f3d_mech_tables.h"

namespace f3d {

// Most backward-Euler substeps the Newton mode tries before giving up.
constexpr int kNewtonMaxSub = 64;

// One backward-Euler solve over h of the reactor (rho_k, T) from (r, T),
// total energy fixed at E0: Newton on
//   F_k = rho_k - rs_k - h W_k wdot_k(rho, T) = 0,  F_E = E(rho, T) - E0 = 0
// with the exact mass-action Jacobian (+ third-body terms) and the analytic
// temperature derivative of the rates.  Constant quotients (1e-6/W_n,
// W_n/W_l, Ru/W_n, 1/W_n, the reverse rates' 1/RsT, the pivots' reciprocals)
// are multiplied where the reference's point-implicit step divides.  Returns false when the iterate leaves the admissible set
// (non-finite, T outside [T_lo, T_hi], rho_k < -rho_tot) or maxiter
// iterations pass without convergence; `it` counts the iterations taken.
template <int NS, class K>
__device__ inline bool be_solve(const DevParams* __restrict__ P, double (&r)[NS], double& T,
                                double h, double E0, const double (&hf)[NS], double rho_tot,
                                double T_lo, double T_hi, int maxiter, double rtol, double ttol,
                                int& it) {
  constexpr int N1 = NS + 1;
  double rs[NS], sc[NS], wdot[NS];
#pragma unroll
  for (int n = 0; n < NS; ++n) rs[n] = r[n];
  double dT_prev = 0.0;
  for (int k = 0; k < maxiter; ++k) {
#pragma unroll
    for (int n = 0; n < NS; ++n) sc[n] = r[n] * P->cmw[n];
    // the rate Jacobian is built in the top-left block of the system matrix
    // and turned into I - W_n/W_l h dwdot/dc in place (one matrix in registers)
    double A[N1][N1], rhs[N1], dx[N1];
    // rate Jacobian in A[0..NS-1][0..NS-1], d wdot / dT in A[0..NS-1][NS]
    production_rate<NS, true, true, N1, true, true, K>(P, sc, T, A, wdot);
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const double src = wdot[n] * P->mw[n] * 1e6 * h;
      rhs[n] = -(r[n] - rs[n] - src);
#pragma unroll
      for (int l = 0; l < NS; ++l) A[n][l] = (n == l ? 1.0 : 0.0) - A[n][l] * P->wnl[n][l] * h;
      A[n][NS] = -A[n][NS] * P->mw[n] * 1e6 * h;
    }
    double E = internal_energy<NS, true>(P, T, r);
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
        A[NS][l] = e * P->Rmw[l] + hf[l];
      }
    }
    A[NS][NS] = cv_mix<NS, true>(P, T, r);
    gauss_solve<N1, true>(A, rhs, dx);
    double rho = 0;
#pragma unroll
    for (int n = 0; n < NS; ++n) rho += r[n];
    double mx = 0;
    bool bad = false;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      r[n] += dx[n];
      mx = fmax(mx, fabs(dx[n]));
      bad = bad || !(r[n] > -rho_tot);
    }
    T += dx[NS];
    ++it;
    // NASA-7 polynomials jump at T_mid: when E0 falls in such a jump the T
    // update flips sign without contracting; with the species converged the
    // step is accepted (the final T comes from GetT of the energy)
    const bool flip = k > 0 && dx[NS] * dT_prev < 0 && fabs(dx[NS]) >= 0.5 * fabs(dT_prev);
    dT_prev = dx[NS];
    if (bad || !(T > T_lo && T < T_hi)) return false;
    if (mx <= rtol * rho && (fabs(dx[NS]) <= ttol || flip)) return true;
  }
  return false;
}

// mode 0: eval_gpu_stiff, 1: eval_gpu, 2: backward-Euler Newton.  K: the
// reaction parameters, from the parameter block or compiled in (KinCT).
template <int NS, int MODE, class K = KinRT>
__global__ void __launch_bounds__(kBlock)
    k_chem(double* __restrict__ U, double* __restrict__ Q, double* __restrict__ rhoi, double dt,
           int maxiter, double rtol, double ttol, int32_t* __restrict__ iters,
           const int64_t* __restrict__ tag, const DevParams* __restrict__ P, Geom g, int bug,
           const int32_t* __restrict__ list, long nwork) {
  // one thread per interior cell, or per entry of an ordered cell list (the
  // surrogate's rejected cells)
  const long t_ = gid();
  if (t_ >= nwork) return;
  const long id = list ? (long)list[t_] : t_;
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
    production_rate<NS, false, false, NS, false, false, K>(P, sc, T0, A, wdot);
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
    production_rate<NS, true, false, NS, false, false, K>(P, sc, T0, A, wdot);
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
    double r0[NS], hf[NS];
    double rho_tot = 0.0;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      r0[n] = r[n];
      hf[n] = P->hfm[n];
      rho_tot += r0[n];
    }
    double E0 = internal_energy<NS, true>(P, T0, r0);
#pragma unroll
    for (int n = 0; n < NS; ++n) E0 += hf[n] * r0[n];
    const double T_lo = 0.5 * P->min_temp, T_hi = 1.5 * P->max_temp;
    // one step of dt; on failure restart from the initial state with twice
    // as many backward-Euler substeps (at most kNewtonMaxSub)
    int it = 0;
    bool ok = false;
    double T = T0;
    for (int nsub = 1; nsub <= kNewtonMaxSub && !ok; nsub *= 2) {
#pragma unroll
      for (int n = 0; n < NS; ++n) r[n] = r0[n];
      T = T0;
      ok = true;
      const double h = dt / nsub;
      for (int k = 0; k < nsub && ok; ++k)
        ok = be_solve<NS, K>(P, r, T, h, E0, hf, rho_tot, T_lo, T_hi, maxiter, rtol, ttol, it);
    }
    if (!ok) {  // no admissible solution: the reactor is left unreacted
#pragma unroll
      for (int n = 0; n < NS; ++n) r[n] = r0[n];
      it = -it;
    }
    if (iters) iters[id] = it;
    double dei = 0;
#pragma unroll
    for (int n = 0; n < NS; ++n) dei += -hf[n] * (r[n] - r0[n]);
    chem_epilogue<NS>(P, U, Q, rhoi, p, NP, r, dei);
  }
}

template <int NS, class K>
static void chem_k(f3d_ctx* c, int mode, double* U, double* Q, double* rhoi, double dt,
                   int maxiter, double rtol, double ttol, int32_t* iters, const int64_t* tag,
                   hipStream_t s, const int32_t* list, long nlist) {
  const long n = list ? nlist : (long)c->g.Nx * c->g.Ny * c->g.Nz;
  if (n <= 0) return;
  const int bug = c->bug_compat & 1;
  if (mode == 0)
    hipLaunchKernelGGL((k_chem<NS, 0, K>), dim3(nblocks(n)), dim3(kBlock), 0, s, U, Q, rhoi, dt,
                       maxiter, rtol, ttol, iters, tag, c->dparams, c->g, bug, list, n);
  else if (mode == 1)
    hipLaunchKernelGGL((k_chem<NS, 1, K>), dim3(nblocks(n)), dim3(kBlock), 0, s, U, Q, rhoi, dt,
                       maxiter, rtol, ttol, iters, tag, c->dparams, c->g, bug, list, n);
  else
    hipLaunchKernelGGL((k_chem<NS, 2, K>), dim3(nblocks(n)), dim3(kBlock), 0, s, U, Q, rhoi, dt,
                       maxiter, rtol, ttol, iters, tag, c->dparams, c->g, bug, list, n);
}

// the compiled-in tables of the context's mechanism when it is one of the
// shipped ones (f3d_create compares every entry), else the parameter block
template <int NS>
static void chem_ns(f3d_ctx* c, int mode, double* U, double* Q, double* rhoi, double dt,
                    int maxiter, double rtol, double ttol, int32_t* iters, const int64_t* tag,
                    hipStream_t s, const int32_t* list, long nlist) {
  if constexpr (NS == MechH2LiDryer::NS) {
    if (c->kin_table == F3D_KIN_H2_LIDRYER) {
      chem_k<NS, KinCT<MechH2LiDryer>>(c, mode, U, Q, rhoi, dt, maxiter, rtol, ttol, iters, tag,
                                       s, list, nlist);
      return;
    }
  }
  if constexpr (NS == MechAir5::NS) {
    if (c->kin_table == F3D_KIN_AIR5) {
      chem_k<NS, KinCT<MechAir5>>(c, mode, U, Q, rhoi, dt, maxiter, rtol, ttol, iters, tag, s,
                                  list, nlist);
      return;
    }
  }
  chem_k<NS, KinRT>(c, mode, U, Q, rhoi, dt, maxiter, rtol, ttol, iters, tag, s, list, nlist);
}

int launch_chem(f3d_ctx* c, int mode, double* U, double* Q, double* rhoi, double dt,
                int maxiter, double rtol, double ttol, int32_t* iters, const int64_t* tag,
                hipStream_t s, const int32_t* list, long nlist) {
  if (mode == 0 && (c->bug_compat & 1) && c->g.Ns < 6) {
    set_error("bug_compat coeffs_lo[n,6] read needs Ns >= 6 (it reads past the array for Ns < 6)");
    return F3D_ERR_ARG;
  }
  switch (c->g.Ns) {
    case 5: chem_ns<5>(c, mode, U, Q, rhoi, dt, maxiter, rtol, ttol, iters, tag, s, list, nlist); break;
    case 9: chem_ns<9>(c, mode, U, Q, rhoi, dt, maxiter, rtol, ttol, iters, tag, s, list, nlist); break;
    default: set_error("chemistry: Ns must be 5 or 9"); return F3D_ERR_UNSUPPORTED;
  }
  return hip_check(hipGetLastError(), "chem");
}

}  // namespace f3d
