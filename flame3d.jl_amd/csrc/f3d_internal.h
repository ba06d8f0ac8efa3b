// Internal declarations shared by the libflame3d translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/flame3d.h"

namespace f3d {

constexpr int MAXNS = 16;
constexpr int MAXNR = 48;
constexpr int NG = 4;
constexpr double kEps = 2.220446049250313e-16;  // eps(Float64)

// Parameter block resident in device memory; kernels read it through a
// uniform pointer, so the loads are scalar.  Arrays are transposed to
// [species][coefficient] so one species' coefficients are contiguous.
struct DevParams {
  int Ns, Nr;
  double Ru, min_temp, max_temp, atm;
  double mw[MAXNS], sep[MAXNS];
  double lo[MAXNS][7], hi[MAXNS][7];
  double visc[MAXNS][5], cond[MAXNS][5];
  double bdiff[MAXNS][MAXNS][5];  // [n][nn][c] = binarydiff_poly[c, nn, n]
  int rtype[MAXNR], sgm[MAXNR];
  int vf[MAXNR][MAXNS], vr[MAXNR][MAXNS];
  double Arr[MAXNR][3], ef[MAXNR][MAXNS], loP[MAXNR][3], Troe[MAXNR][4];
  double SWeps2, WENOeps, phi1, phi2, UP7[7], T_criteria;
  // Quotients of constants, rounded once on the host as the device would
  // round them: 1/W_n, Ru/W_n, 1e-6/W_n, W_n/W_l, and the formation energy
  // per mass lo[n][5] Ru / W_n in the reference's operation order.  The
  // backward-Euler Newton, the warm-started e -> T and the surrogate's
  // epilogue multiply by them where the reference divides (a correctly
  // rounded FP64 division is a ten-instruction sequence on the GPU).
  double rmw[MAXNS], Rmw[MAXNS], cmw[MAXNS], hfm[MAXNS], wnl[MAXNS][MAXNS];
  int recon_kind;  // 0 UP7 / WENO5-Z / NND (the reference), 1 WENO7-Z in the middle branch
  // float copies for the FP32 flux path
  float mw_f[MAXNS], sep_f[MAXNS], lo_f[MAXNS][7], hi_f[MAXNS][7];
  float visc_f[MAXNS][5], cond_f[MAXNS][5], bdiff_f[MAXNS][MAXNS][5];
  float Ru_f, WENOeps_f, phi1_f, phi2_f, SWeps2_f, UP7_f[7];
  float Rmw_f[MAXNS];  // Ru / W_n rounded to float
  // Wilke's species-pair constants for the FP32 mixture rule (W_l / W_n)^(1/4),
  // 1 / sqrt(8 (1 + W_n / W_l)) and W_n / W_l
  float wr4_f[MAXNS][MAXNS], c8_f[MAXNS][MAXNS], iwr_f[MAXNS][MAXNS];
};

// Padded-block geometry (0-based).  Point p = i + Nxt*(j + Nyt*k); variable
// n of a padded array lives at p + n*NP.
struct Geom {
  int Nx, Ny, Nz, Ns, Nr;
  int Nxt, Nyt, Nzt;
  int NP;
};

}  // namespace f3d

struct f3d_ctx {
  f3d::Geom g;
  int precision;
  int bug_compat;
  int sensor_kind = 0;  // F3D_SENSOR_PRESSURE / F3D_SENSOR_DUCROS
  int stage_kind = 0;   // F3D_STAGE_FUSED / F3D_STAGE_GRANULAR (F3D_FP32 contexts)
  int kin_table = 0;    // F3D_KIN_*: the chemistry kernels' reaction-parameter source
  long spec_waves = 0;  // species-diffusion march waves resident on the device (0: not queried)
  f3d::DevParams host_params;
  f3d::DevParams* dparams = nullptr;
  // boundary fills
  int bc[6] = {0, 0, 0, 0, 0, 0};
  double inflow_Q[7] = {0, 0, 0, 0, 0, 0, 0};
  double inflow_rhoi[f3d::MAXNS] = {0};
  // metrics registered by f3d_set_metrics
  const double* metrics[9] = {nullptr};
  const double* J = nullptr;
  // F3D_FP32: float copies of the 9 metrics and J (same values as the
  // float conversions the kernels would make; half the bytes)
  float* mf[10] = {nullptr};
  void* mf_base = nullptr;
  int32_t* nan_flag = nullptr;  // f3d_nan_check's device flag
  // context-owned workspace for f3d_rk3_stage
  void* ws_base = nullptr;
  size_t ws_bytes = 0;
  void* ws[32] = {nullptr};
  // chemistry surrogate (f3d_nn_*)
  bool nn_ready = false;
  float* nn_master = nullptr;  // FP32 parameters, then Adam moments m1, m2
  float* nn_m1 = nullptr;
  float* nn_m2 = nullptr;
  void* nn_Wb = nullptr;   // BF16 weights [o][k]
  void* nn_Wtb = nullptr;  // BF16 weights transposed [k][o]
  float* nn_partial = nullptr;
  long nn_partial_tiles = 0;
  void* nn_cfg = nullptr;  // device NNDev
  double nn_dt = 0;
  long nn_step = 0;
  int8_t* nn_status = nullptr;  // per-cell gate codes [Nx*Ny*Nz] when the caller passes none
  int32_t* nn_fbcount = nullptr;  // rejected cells per 64-cell tile, then their offsets
  int32_t* nn_fblist = nullptr;   // rejected cell ids in cell order [Nx*Ny*Nz]
  int32_t* nn_fbtotal = nullptr;  // number of rejected cells
  float* nn_out = nullptr;        // the network's raw outputs [Ns][Nx*Ny*Nz] for the epilogue
  double* nn_cand = nullptr;      // the epilogue's candidate states [Ns + 3][Nx*Ny*Nz]
  // f3d_nn_timing: events around the last f3d_nn_chem's network kernel, its
  // gate epilogue kernel, its scan / compaction and its Newton fallback, on
  // the call's stream
  bool nn_timing = false;
  bool nn_timed = false;
  hipEvent_t nn_ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
};

namespace f3d {

void set_error(const std::string& msg);
void nn_free(f3d_ctx* c);
int hip_check(hipError_t e, const char* what);

// ---- launch helpers ---------------------------------------------------------
constexpr int kBlock = 256;

inline unsigned nblocks(long n, int block = kBlock) { return (unsigned)((n + block - 1) / block); }

// Bijective remap of hardware block b of nb so that consecutive logical
// blocks run on the same XCD (workgroups are dealt round-robin over the 8
// XCDs, each with its own L2): neighbouring tiles then share an L2.
__device__ inline unsigned xcd_remap(unsigned b, unsigned nb) {
  const unsigned q = nb >> 3, r = nb & 7u, x = b & 7u, k = b >> 3;
  const unsigned base = (x < r) ? x * (q + 1) : r * (q + 1) + (x - r) * q;
  return base + k;
}

__device__ inline unsigned xcd_block() { return xcd_remap(blockIdx.x, gridDim.x); }

__device__ inline long gid() { return (long)xcd_block() * blockDim.x + threadIdx.x; }

// ---- typed parameter access -----------------------------------------------
template <typename R>
struct PT;
template <>
struct PT<double> {
  __device__ static const double* lo(const DevParams* P, int n) { return P->lo[n]; }
  __device__ static const double* hi(const DevParams* P, int n) { return P->hi[n]; }
  __device__ static double sep(const DevParams* P, int n) { return P->sep[n]; }
  __device__ static double mw(const DevParams* P, int n) { return P->mw[n]; }
  __device__ static const double* visc(const DevParams* P, int n) { return P->visc[n]; }
  __device__ static const double* cond(const DevParams* P, int n) { return P->cond[n]; }
  __device__ static const double* bdiff(const DevParams* P, int n, int nn) { return P->bdiff[n][nn]; }
  __device__ static double Ru(const DevParams* P) { return P->Ru; }
  __device__ static double Rmw(const DevParams* P, int n) { return P->Rmw[n]; }
  __device__ static double weps(const DevParams* P) { return P->WENOeps; }
  __device__ static double phi1(const DevParams* P) { return P->phi1; }
  __device__ static double phi2(const DevParams* P) { return P->phi2; }
  __device__ static double sweps2(const DevParams* P) { return P->SWeps2; }
  __device__ static const double* up7(const DevParams* P) { return P->UP7; }
  __device__ static double eps() { return kEps; }
};
template <>
struct PT<float> {
  __device__ static const float* lo(const DevParams* P, int n) { return P->lo_f[n]; }
  __device__ static const float* hi(const DevParams* P, int n) { return P->hi_f[n]; }
  __device__ static float sep(const DevParams* P, int n) { return P->sep_f[n]; }
  __device__ static float mw(const DevParams* P, int n) { return P->mw_f[n]; }
  __device__ static const float* visc(const DevParams* P, int n) { return P->visc_f[n]; }
  __device__ static const float* cond(const DevParams* P, int n) { return P->cond_f[n]; }
  __device__ static const float* bdiff(const DevParams* P, int n, int nn) { return P->bdiff_f[n][nn]; }
  __device__ static float Ru(const DevParams* P) { return P->Ru_f; }
  __device__ static float Rmw(const DevParams* P, int n) { return P->Rmw_f[n]; }
  __device__ static float weps(const DevParams* P) { return P->WENOeps_f; }
  __device__ static float phi1(const DevParams* P) { return P->phi1_f; }
  __device__ static float phi2(const DevParams* P) { return P->phi2_f; }
  __device__ static float sweps2(const DevParams* P) { return P->SWeps2_f; }
  __device__ static const float* up7(const DevParams* P) { return P->UP7_f; }
  // eps(Float64) is used as a floor in the reference; kept as the same value
  __device__ static float eps() { return (float)kEps; }
};

// ---- thermo device functions (thermo.jl) -----------------------------------

// h_specs, thermo.jl:107-133 (J/kg)
template <int NS, typename R>
__device__ inline void h_specs(const DevParams* P, R T, R* h) {
  const R T2 = R(0.5) * T * T;
  const R T3 = R(2.0 / 3.0) * T2 * T;
  const R T4 = T2 * T2;
  const R T5 = R(0.8) * T4 * T;
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    const R* lo = PT<R>::lo(P, n);
    const R* hc = PT<R>::hi(P, n);
    R v;
    if (T < PT<R>::sep(P, n)) {
      v = lo[0] * T + lo[1] * T2 + lo[2] * T3 + lo[3] * T4 + lo[4] * T5;
    } else {
      v = hc[0] * T + hc[1] * T2 + hc[2] * T3 + hc[3] * T4 + hc[4] * T5 + (hc[5] - lo[5]);
    }
    h[n] = v * PT<R>::Rmw(P, n);  // Ru / W_n (the same quotient, rounded once on the host)
  }
}

// ei_specs, thermo.jl:136-160 and InternalEnergy, thermo.jl:219-228 (FP64).
// RCP: rho_k e_k times 1/W_k instead of divided by W_k (one rounding
// differs; the Newton chemistry and the warm-started e -> T use it).
template <int NS, bool RCP = false>
__device__ inline double internal_energy(const DevParams* P, double T, const double* rhoi) {
  const double T2 = 0.5 * T * T;
  const double T3 = 2.0 / 3.0 * T2 * T;
  const double T4 = T2 * T2;
  const double T5 = 0.8 * T4 * T;
  double result = 0;
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    const double* lo = P->lo[n];
    const double* hc = P->hi[n];
    double e;
    if (T < P->sep[n]) {
      e = (lo[0] - 1) * T + lo[1] * T2 + lo[2] * T3 + lo[3] * T4 + lo[4] * T5;
    } else {
      e = (hc[0] - 1) * T + hc[1] * T2 + hc[2] * T3 + hc[3] * T4 + hc[4] * T5 + (hc[5] - lo[5]);
    }
    result += RCP ? rhoi[n] * e * P->rmw[n] : rhoi[n] * e / P->mw[n];
  }
  return result * P->Ru;
}

// CV, thermo.jl:194-203 (FP64); RCP as internal_energy
template <int NS, bool RCP = false>
__device__ inline double cv_mix(const DevParams* P, double T, const double* rhoi) {
  const double T2 = T * T;
  const double T3 = T2 * T;
  const double T4 = T2 * T2;
  double result = 0;
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    const double* c = (T < P->sep[n]) ? P->lo[n] : P->hi[n];
    const double cp = c[0] + c[1] * T + c[2] * T2 + c[3] * T3 + c[4] * T4;
    result += RCP ? (cp - 1) * rhoi[n] * P->rmw[n] : (cp - 1) * rhoi[n] / P->mw[n];
  }
  return result * P->Ru;
}

// Pmixture, thermo.jl:20-26; RCP as internal_energy
template <int NS, bool RCP = false>
__device__ inline double pmixture(const DevParams* P, double T, const double* rhoi) {
  double YOW = 0;
#pragma unroll
  for (int n = 0; n < NS; ++n) YOW += RCP ? rhoi[n] * P->rmw[n] : rhoi[n] / P->mw[n];
  return P->Ru * T * YOW;
}

// GetT, thermo.jl:231-282 (FP64 Newton e -> T; tol 1e-3 K, <= 30 updates,
// +-100 K step clamp, linear extrapolation outside [min_temp, max_temp]).
template <int NS>
__device__ inline double get_T(const DevParams* P, double ein, const double* rhoi, int* iters) {
  const double tol = 1e-3;
  const double tmin = P->min_temp, tmax = P->max_temp;
  const double emin = internal_energy<NS>(P, tmin, rhoi);
  const double emax = internal_energy<NS>(P, tmax, rhoi);
  if (iters) *iters = 0;
  if (ein < emin) return tmin - (emin - ein) / cv_mix<NS>(P, tmin, rhoi);
  if (ein > emax) return tmax - (emax - ein) / cv_mix<NS>(P, tmax, rhoi);
  double As = 0;
#pragma unroll
  for (int n = 0; n < NS; ++n) As += (P->lo[n][0] - 1) * P->Ru / P->mw[n] * rhoi[n];
  double t1 = ein / As;
  if (t1 < tmin || t1 > tmax) t1 = tmin + (tmax - tmin) / (emax - emin) * (ein - emin);
  int it = 0;
  double x1 = 0.0, x2 = 0.0;  // t1 one and two updates ago
  for (int m = 0; m < 30; ++m) {
    const double e1 = internal_energy<NS>(P, t1, rhoi);
    const double cv = cv_mix<NS>(P, t1, rhoi);
    double dt = (ein - e1) / cv;
    if (dt > 100.0) {
      dt = 100.0;
    } else if (dt < -100.0) {
      dt = -100.0;
    } else if (fabs(dt) < tol) {
      break;
    } else if (t1 + dt == t1) {
      break;
    }
    // Each update is a function of t1 alone, so once t1 equals the iterate
    // two updates back exactly (an energy inside the jump of the NASA-7
    // fits at a breakpoint, where the updates alternate across it) the rest
    // cycles with period two and the point the 30 updates end on is known:
    // the same result, bit for bit, without the remaining updates.
    if (m >= 2 && t1 == x2) {
      if (iters) *iters = 30;
      return ((30 - m) % 2 == 0) ? t1 : x1;
    }
    x2 = x1;
    x1 = t1;
    t1 += dt;
    ++it;
  }
  if (iters) *iters = it;
  return t1;
}

// GetT for an energy inside the small jump of the NASA-7 fits at their
// breakpoint, where e(T) has no root.  get_T's update then alternates
// between the roots of the two fits' continuations across the breakpoint (a
// 2-cycle, reached within a few updates) until its 30 updates run out, and
// returns the cycle point the parity of the remaining updates selects.  This
// is get_T's iteration (same start, update, clamp and stopping tests) that,
// once x_m repeats x_{m-2} to 1e-9 K, returns the point x_30 would be
// (x_m for an even number of remaining updates, x_{m-1} for an odd one)
// instead of doing the remaining updates.  Off the cycle it is get_T.  Its
// updates evaluate the energy and c_v with the reciprocal molar masses (one
// rounding each differs from get_T's; the cycle point it returns agrees with
// get_T's to that rounding).
template <int NS>
__device__ inline double get_T_gap(const DevParams* P, double ein, const double* rhoi) {
  const double tol = 1e-3;
  const double tmin = P->min_temp, tmax = P->max_temp;
  const double emin = internal_energy<NS>(P, tmin, rhoi);
  const double emax = internal_energy<NS>(P, tmax, rhoi);
  if (ein < emin) return tmin - (emin - ein) / cv_mix<NS>(P, tmin, rhoi);
  if (ein > emax) return tmax - (emax - ein) / cv_mix<NS>(P, tmax, rhoi);
  double As = 0;
#pragma unroll
  for (int n = 0; n < NS; ++n) As += (P->lo[n][0] - 1) * P->Ru / P->mw[n] * rhoi[n];
  double t1 = ein / As;
  if (t1 < tmin || t1 > tmax) t1 = tmin + (tmax - tmin) / (emax - emin) * (ein - emin);
  double x1 = 0.0, x2 = 0.0;  // t1 one and two updates ago
  for (int m = 0; m < 30; ++m) {
    const double e1 = internal_energy<NS, true>(P, t1, rhoi);
    const double cv = cv_mix<NS, true>(P, t1, rhoi);
    double dt = (ein - e1) / cv;
    if (dt > 100.0) {
      dt = 100.0;
    } else if (dt < -100.0) {
      dt = -100.0;
    } else if (fabs(dt) < tol) {
      break;
    } else if (t1 + dt == t1) {
      break;
    }
    if (m >= 2 && fabs(t1 - x2) <= 1e-9) return ((30 - m) % 2 == 0) ? t1 : x1;
    x2 = x1;
    x1 = t1;
    t1 += dt;
  }
  return t1;
}

// GetT warm-started from a nearby earlier temperature (the cell's T before
// a stage or a chemistry step): the same Newton update, 1e-3 K tolerance,
// +-100 K clamp and 30-update limit as get_T, from T_guess instead of the
// reference's e / A_s start, and without the two bracketing energies.  A
// guess outside [min_temp, max_temp] (or not finite), no convergence within
// the 30 updates, or a result outside the range
// (the reference's linear extrapolation region), takes get_T itself.  Both
// stop within the tolerance of the same root, so they differ by < 1e-3 K.
// An update that flips sign without shrinking marks an energy inside the
// small jump of the NASA-7 fits at their breakpoint (e.g. a stream at
// 1000 K), which has no exact root: there, and for any result within 0.5 K
// of a breakpoint, the result is get_T's own end point (get_T_gap).  The
// warm updates evaluate the energy and c_v with the reciprocal molar masses.
template <int NS>
__device__ inline double get_T_from(const DevParams* P, double ein, const double* rhoi,
                                    double T_guess) {
  const double tol = 1e-3;
  const double tmin = P->min_temp, tmax = P->max_temp;
  if (!(T_guess >= tmin && T_guess <= tmax)) return get_T<NS>(P, ein, rhoi, nullptr);
  double t1 = T_guess;
  bool done = false;
  double dt_prev = 0.0;
  for (int m = 0; m < 30; ++m) {
    const double e1 = internal_energy<NS, true>(P, t1, rhoi);
    const double cv = cv_mix<NS, true>(P, t1, rhoi);
    double dt = (ein - e1) / cv;
    if (dt > 100.0) {
      dt = 100.0;
    } else if (dt < -100.0) {
      dt = -100.0;
    } else if (fabs(dt) < tol) {
      done = true;
      break;
    } else if (t1 + dt == t1) {
      done = true;
      break;
    } else if (dt * dt_prev < 0.0 && fabs(dt) >= 0.5 * fabs(dt_prev)) {
      // the update flips sign without shrinking: the energy sits in the
      // jump of the NASA-7 fits at their breakpoint (no exact root)
      return get_T_gap<NS>(P, ein, rhoi);
    }
    dt_prev = dt;
    t1 += dt;
  }
  if (!done || !(t1 >= tmin && t1 <= tmax)) return get_T<NS>(P, ein, rhoi, nullptr);
  // within 0.5 K of a breakpoint (the jumps of the fits are <= 0.2 K wide):
  // where the jump is narrower than the tolerance, or the fits overlap (two
  // roots), the start decides where the update stops; take get_T's end point
#pragma unroll
  for (int n = 0; n < NS; ++n)
    if (fabs(t1 - P->sep[n]) <= 0.5) return get_T_gap<NS>(P, ein, rhoi);
  return t1;
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

// ---- internal launchers (defined in the .hip files) ------------------------
// Metric arguments: `M` points at 9 metric arrays (and `J` at the Jacobian)
// of type double, or of type float when `mf` is set (the FP32 stage's copies).
int launch_flux_split(f3d_ctx* c, const double* Q, void* Fp, void* Fm, const void* Ax,
                      const void* Ay, const void* Az, const int64_t* tag, hipStream_t s,
                      bool mf = false);
int launch_species_split(f3d_ctx* c, const double* rhoi, const double* Q, void* Fp, void* Fm,
                         const void* Ax, const void* Ay, const void* Az,
                         const int64_t* tag, hipStream_t s, bool mf = false);
int launch_shock_sensor(f3d_ctx* c, void* phi, const double* Q, hipStream_t s);
int launch_weno(f3d_ctx* c, int dir, void* F, const void* phi, const void* Fp, const void* Fm,
                int NV, const int64_t* tag, hipStream_t s);
int launch_viscous(f3d_ctx* c, int dir, void* Fv, const double* Q, const void* const* M,
                   const void* J, const void* lam, const void* mu, const void* Fh,
                   const int64_t* tag, hipStream_t s, bool mf = false);
int launch_spec_viscous(f3d_ctx* c, int dir, void* Fd, const double* Q, const void* Yi,
                        const void* const* M, const void* J, const void* D, void* Fh,
                        const int64_t* tag, hipStream_t s, bool mf = false);
int launch_to_float(f3d_ctx* c, float* dst, const double* src, long n, hipStream_t s);
int launch_div(f3d_ctx* c, double* U, const void* Fx, const void* Fy, const void* Fz,
               const void* Fvx, const void* Fvy, const void* Fvz, double dt, const double* J,
               const int64_t* tag, hipStream_t s);
int launch_div_specs(f3d_ctx* c, double* rhoi, const void* Fx, const void* Fy, const void* Fz,
                     const void* Fdx, const void* Fdy, const void* Fdz, double dt,
                     const double* J, const int64_t* tag, hipStream_t s);
int launch_mixture(f3d_ctx* c, const double* Q, const double* rhoi, void* Yi, void* lam,
                   void* mu, void* D, const int64_t* tag, hipStream_t s);
int launch_c2prim(f3d_ctx* c, double* U, double* Q, double* rhoi, const int64_t* tag,
                  hipStream_t s);
int launch_stage_end(f3d_ctx* c, int KRK, double* U, const double* Un, double* rhoi,
                     const double* rhon, double* Q, void* const* F, double dt, const double* J,
                     const int64_t* tag, hipStream_t s);
int launch_prim2c(f3d_ctx* c, double* U, const double* Q, hipStream_t s);
// fused FP32 sweeps (sweep.hip): every direction's convective and diffusive
// face-flux differences into rhs[5 + Ns][Nx*Ny*Nz], then the stage end
int launch_sweep_conv(f3d_ctx* c, int dir, float* rhs, const double* Q, const double* rhoi,
                      const void* phi, int acc, hipStream_t s);
int launch_sweep_diff(f3d_ctx* c, int dir, float* rhs, const double* Q, const void* Yi,
                      const void* lam, const void* mu, const void* D, hipStream_t s);
int launch_sweeps(f3d_ctx* c, float* rhs, const double* Q, const double* rhoi, const void* phi,
                  const void* Yi, const void* lam, const void* mu, const void* D,
                  hipStream_t s);
int launch_sweep_stage_end(f3d_ctx* c, int KRK, double* U, const double* Un, double* rhoi,
                           const double* rhon, double* Q, const float* rhs, double dt,
                           const double* J, hipStream_t s);
int launch_lincomb(f3d_ctx* c, double* U, const double* Un, int NV, double a, double b,
                   const int64_t* tag, hipStream_t s);
int launch_get_y(f3d_ctx* c, double* Yi, const double* rhoi, const double* Q,
                 const int64_t* tag, hipStream_t s);
int launch_fill(f3d_ctx* c, double* Q, double* U, double* rhoi, hipStream_t s);
int launch_chem(f3d_ctx* c, int mode, double* U, double* Q, double* rhoi, double dt,
                int maxiter, double rtol, double ttol, int32_t* iters, const int64_t* tag,
                hipStream_t s, const int32_t* list = nullptr, long nlist = 0);

}  // namespace f3d
