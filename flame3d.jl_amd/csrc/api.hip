// C-ABI of libflame3d.so: context management, argument checks and the
// RK3-stage orchestration of solver.jl:22-56 / :298-326.
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "f3d_internal.h"
#include "f3d_mech_tables.h"

namespace f3d {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return F3D_OK;
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return F3D_ERR_HIP;
}

static int fill_params(DevParams& P, const f3d_dims* d, const f3d_thermo* th,
                       const f3d_kinetics* kin, const f3d_scheme* sc) {
  std::memset(&P, 0, sizeof(P));
  const int Ns = d->Ns, Nr = d->Nr;
  P.Ns = Ns;
  P.Nr = Nr;
  P.Ru = th->Ru;
  P.min_temp = th->min_temp;
  P.max_temp = th->max_temp;
  for (int n = 0; n < Ns; ++n) {
    P.mw[n] = th->mw[n];
    P.sep[n] = th->coeffs_sep[n];
    for (int c = 0; c < 7; ++c) {
      P.lo[n][c] = th->coeffs_lo[c + 7 * n];
      P.hi[n][c] = th->coeffs_hi[c + 7 * n];
    }
    for (int c = 0; c < 5; ++c) {
      P.visc[n][c] = th->visc_poly ? th->visc_poly[c + 5 * n] : 0.0;
      P.cond[n][c] = th->conduct_poly ? th->conduct_poly[c + 5 * n] : 0.0;
    }
    for (int nn = 0; nn < Ns; ++nn)
      for (int c = 0; c < 5; ++c)
        P.bdiff[n][nn][c] = th->binarydiff_poly ? th->binarydiff_poly[c + 5 * (nn + Ns * n)] : 0.0;
  }
  if (Nr > 0) {
    if (!kin || !kin->reaction_type || !kin->sgm || !kin->vf || !kin->vr || !kin->Arr) {
      set_error("f3d_create: Nr > 0 needs the kinetics arrays");
      return F3D_ERR_ARG;
    }
    P.atm = kin->atm;
    for (int r = 0; r < Nr; ++r) {
      P.rtype[r] = (int)kin->reaction_type[r];
      if (P.rtype[r] < 1 || P.rtype[r] > 3) {
        set_error("f3d_create: reaction_type must be 1, 2 or 3");
        return F3D_ERR_ARG;
      }
      P.sgm[r] = (int)kin->sgm[r];
      for (int c = 0; c < 3; ++c) {
        P.Arr[r][c] = kin->Arr[c + 3 * r];
        P.loP[r][c] = kin->loP ? kin->loP[c + 3 * r] : 0.0;
      }
      for (int c = 0; c < 4; ++c) P.Troe[r][c] = kin->Troe ? kin->Troe[c + 4 * r] : 0.0;
      for (int n = 0; n < Ns; ++n) {
        P.vf[r][n] = (int)kin->vf[n + Ns * r];
        P.vr[r][n] = (int)kin->vr[n + Ns * r];
        P.ef[r][n] = kin->ef ? kin->ef[n + Ns * r] : 0.0;
      }
    }
  } else {
    P.atm = kin ? kin->atm : 101325.0;
  }
  P.SWeps2 = sc->SWeps2;
  P.WENOeps = sc->WENOeps;
  P.phi1 = sc->hybrid_phi1;
  P.phi2 = sc->hybrid_phi2;
  P.recon_kind = sc->recon_kind;
  for (int c = 0; c < 7; ++c) P.UP7[c] = sc->UP7[c];
  P.T_criteria = sc->T_criteria;
  // float copies
  for (int n = 0; n < MAXNS; ++n) {
    P.mw_f[n] = (float)P.mw[n];
    P.sep_f[n] = (float)P.sep[n];
    for (int c = 0; c < 7; ++c) {
      P.lo_f[n][c] = (float)P.lo[n][c];
      P.hi_f[n][c] = (float)P.hi[n][c];
    }
    for (int c = 0; c < 5; ++c) {
      P.visc_f[n][c] = (float)P.visc[n][c];
      P.cond_f[n][c] = (float)P.cond[n][c];
    }
    for (int nn = 0; nn < MAXNS; ++nn)
      for (int c = 0; c < 5; ++c) P.bdiff_f[n][nn][c] = (float)P.bdiff[n][nn][c];
  }
  for (int n = 0; n < d->Ns; ++n)
    for (int l = 0; l < d->Ns; ++l) {
      const double wr = P.mw[l] / P.mw[n];
      P.wr4_f[n][l] = (float)std::sqrt(std::sqrt(wr));
      P.c8_f[n][l] = (float)(1.0 / std::sqrt(8.0 + 8.0 / wr));
      P.iwr_f[n][l] = (float)(1.0 / wr);
    }
  for (int n = 0; n < d->Ns; ++n) {
    P.rmw[n] = 1.0 / P.mw[n];
    P.Rmw[n] = P.Ru / P.mw[n];
    P.Rmw_f[n] = (float)P.Rmw[n];
    P.cmw[n] = 1e-6 / P.mw[n];
    P.hfm[n] = P.lo[n][5] * P.Ru / P.mw[n];
    for (int l = 0; l < d->Ns; ++l) P.wnl[n][l] = P.mw[n] / P.mw[l];
  }
  P.Ru_f = (float)P.Ru;
  P.WENOeps_f = (float)P.WENOeps;
  P.phi1_f = (float)P.phi1;
  P.phi2_f = (float)P.phi2;
  P.SWeps2_f = (float)P.SWeps2;
  for (int c = 0; c < 7; ++c) P.UP7_f[c] = (float)P.UP7[c];
  return F3D_OK;
}

// The compiled-in reaction tables (f3d_mech_tables.h) that equal the
// context's kinetics in every entry, if any.
template <class T>
static bool same_kinetics(const DevParams& P) {
  if (P.Ns != T::NS || P.Nr != T::NR) return false;
  for (int r = 0; r < T::NR; ++r) {
    if (P.rtype[r] != T::rtype[r] || P.sgm[r] != T::sgm[r]) return false;
    for (int k = 0; k < 3; ++k)
      if (P.Arr[r][k] != T::Arr[r][k] || P.loP[r][k] != T::loP[r][k]) return false;
    for (int k = 0; k < 4; ++k)
      if (P.Troe[r][k] != T::Troe[r][k]) return false;
    for (int n = 0; n < T::NS; ++n)
      if (P.vf[r][n] != T::vf[r][n] || P.vr[r][n] != T::vr[r][n] || P.ef[r][n] != T::ef[r][n])
        return false;
  }
  return true;
}

static int kinetics_table(const DevParams& P) {
  if (same_kinetics<MechH2LiDryer>(P)) return F3D_KIN_H2_LIDRYER;
  if (same_kinetics<MechAir5>(P)) return F3D_KIN_AIR5;
  return F3D_KIN_RUNTIME;
}

static size_t real_size(const f3d_ctx* c) { return c->precision == F3D_FP64 ? 8 : 4; }

// Workspace layout of f3d_rk3_stage (names as solver.jl:141-167).
enum WsId {
  W_Yi, W_lam, W_mu, W_D, W_phi, W_Fp, W_Fm, W_Fpi, W_Fmi,
  W_Fx, W_Fy, W_Fz, W_Fvx, W_Fvy, W_Fvz, W_Fxi, W_Fyi, W_Fzi,
  W_Fdx, W_Fdy, W_Fdz, W_Fhx, W_Fhy, W_Fhz, W_RHS, W_COUNT
};
static const char* kWsNames[W_COUNT] = {
    "Yi", "lambda", "mu", "D", "phi", "Fp", "Fm", "Fp_i", "Fm_i", "Fx", "Fy", "Fz",
    "Fvx", "Fvy", "Fvz", "Fx_i", "Fy_i", "Fz_i", "Fdx", "Fdy", "Fdz", "Fhx", "Fhy", "Fhz",
    "rhs"};

static int ensure_workspace(f3d_ctx* c) {
  if (c->ws_base) return F3D_OK;
  const Geom& g = c->g;
  const size_t NP = g.NP, Ns = g.Ns;
  const size_t nf[3] = {(size_t)(g.Nx + 1) * g.Ny * g.Nz, (size_t)g.Nx * (g.Ny + 1) * g.Nz,
                        (size_t)g.Nx * g.Ny * (g.Nz + 1)};
  size_t elems[W_COUNT];
  elems[W_Yi] = NP * Ns;
  elems[W_lam] = NP;
  elems[W_mu] = NP;
  elems[W_D] = NP * Ns;
  elems[W_phi] = NP;
  elems[W_Fp] = elems[W_Fm] = NP * 5;
  elems[W_Fpi] = elems[W_Fmi] = NP * Ns;
  for (int d = 0; d < 3; ++d) {
    elems[W_Fx + d] = nf[d] * 5;
    elems[W_Fvx + d] = nf[d] * 4;
    elems[W_Fxi + d] = nf[d] * Ns;
    elems[W_Fdx + d] = nf[d] * Ns;
    elems[W_Fhx + d] = nf[d] * 3;
  }
  // the fused FP32 stage's right-hand side [5 + Ns][Nx*Ny*Nz]
  elems[W_RHS] = c->precision == F3D_FP32 ? (size_t)g.Nx * g.Ny * g.Nz * (5 + Ns) : 0;
  const size_t rs = real_size(c);
  size_t off[W_COUNT], total = 0;
  for (int w = 0; w < W_COUNT; ++w) {
    off[w] = total;
    total += (elems[w] * rs + 255) / 256 * 256;
  }
  void* base = nullptr;
  if (hipMalloc(&base, total) != hipSuccess) {
    set_error("f3d_rk3_stage: workspace allocation failed");
    return F3D_ERR_ALLOC;
  }
  // phi is only written on [1, N+2NG-2]; the reference allocates it zeroed
  (void)hipMemset(base, 0, total);
  c->ws_base = base;
  c->ws_bytes = total;
  for (int w = 0; w < W_COUNT; ++w) c->ws[w] = (char*)base + off[w];
  return F3D_OK;
}

}  // namespace f3d

using namespace f3d;

#define CHECK_CTX(c)                      \
  do {                                    \
    if (!(c)) {                           \
      set_error("null f3d_ctx");          \
      return F3D_ERR_ARG;                 \
    }                                     \
  } while (0)
#define CHECK_PTR(p, name)                           \
  do {                                               \
    if (!(p)) {                                      \
      set_error(std::string("null argument ") + name); \
      return F3D_ERR_ARG;                            \
    }                                                \
  } while (0)

extern "C" {

const char* f3d_last_error(void) { return g_err.c_str(); }

int f3d_create(const f3d_dims* d, const f3d_thermo* th, const f3d_kinetics* kin,
               const f3d_scheme* sc, f3d_ctx** out) {
  CHECK_PTR(d, "dims");
  CHECK_PTR(th, "thermo");
  CHECK_PTR(sc, "scheme");
  CHECK_PTR(out, "out");
  *out = nullptr;
  if (d->NG != NG) {
    set_error("f3d_create: NG must be 4");
    return F3D_ERR_UNSUPPORTED;
  }
  if (d->Nx < 4 || d->Ny < 4 || d->Nz < 4) {
    set_error("f3d_create: Nx, Ny, Nz must be >= 4");
    return F3D_ERR_ARG;
  }
  if (d->Ns < 1 || d->Ns > MAXNS || d->Nr < 0 || d->Nr > MAXNR) {
    set_error("f3d_create: Ns must be in [1, 16] and Nr in [0, 48]");
    return F3D_ERR_UNSUPPORTED;
  }
  const long NP = (long)(d->Nx + 2 * NG) * (d->Ny + 2 * NG) * (d->Nz + 2 * NG);
  if (NP * (long)(d->Ns > 7 ? d->Ns : 7) >= (1L << 31)) {
    set_error("f3d_create: block too large for 32-bit point indexing");
    return F3D_ERR_UNSUPPORTED;
  }
  if (sc->precision != F3D_FP64 && sc->precision != F3D_FP32) {
    set_error("f3d_create: precision must be F3D_FP64 or F3D_FP32");
    return F3D_ERR_ARG;
  }
  if (sc->recon_kind != F3D_RECON_REFERENCE && sc->recon_kind != F3D_RECON_WENO7) {
    set_error("f3d_create: recon_kind must be F3D_RECON_REFERENCE or F3D_RECON_WENO7");
    return F3D_ERR_ARG;
  }
  if (sc->stage_kind != F3D_STAGE_FUSED && sc->stage_kind != F3D_STAGE_GRANULAR) {
    set_error("f3d_create: stage_kind must be F3D_STAGE_FUSED or F3D_STAGE_GRANULAR");
    return F3D_ERR_ARG;
  }
  if (sc->sensor_kind != F3D_SENSOR_PRESSURE && sc->sensor_kind != F3D_SENSOR_DUCROS) {
    set_error("f3d_create: sensor_kind must be F3D_SENSOR_PRESSURE or F3D_SENSOR_DUCROS");
    return F3D_ERR_ARG;
  }
  f3d_ctx* c = new (std::nothrow) f3d_ctx();
  if (!c) {
    set_error("f3d_create: out of host memory");
    return F3D_ERR_ALLOC;
  }
  c->g.Nx = d->Nx;
  c->g.Ny = d->Ny;
  c->g.Nz = d->Nz;
  c->g.Ns = d->Ns;
  c->g.Nr = d->Nr;
  c->g.Nxt = d->Nx + 2 * NG;
  c->g.Nyt = d->Ny + 2 * NG;
  c->g.Nzt = d->Nz + 2 * NG;
  c->g.NP = (int)NP;
  c->precision = sc->precision;
  c->bug_compat = sc->bug_compat;
  c->sensor_kind = sc->sensor_kind;
  c->stage_kind = sc->stage_kind;
  int rc = fill_params(c->host_params, d, th, kin, sc);
  if (rc) {
    delete c;
    return rc;
  }
  c->kin_table = kinetics_table(c->host_params);
  if (hipMalloc(&c->dparams, sizeof(DevParams)) != hipSuccess) {
    delete c;
    set_error("f3d_create: device allocation of parameters failed");
    return F3D_ERR_ALLOC;
  }
  rc = hip_check(hipMemcpy(c->dparams, &c->host_params, sizeof(DevParams), hipMemcpyHostToDevice),
                 "f3d_create: parameter upload");
  if (rc) {
    (void)hipFree(c->dparams);
    delete c;
    return rc;
  }
  *out = c;
  return F3D_OK;
}

int f3d_destroy(f3d_ctx* c) {
  if (!c) return F3D_OK;
  nn_free(c);
  if (c->ws_base) (void)hipFree(c->ws_base);
  if (c->mf_base) (void)hipFree(c->mf_base);
  if (c->nan_flag) (void)hipFree(c->nan_flag);
  if (c->dparams) (void)hipFree(c->dparams);
  delete c;
  return F3D_OK;
}

// On the FP32 path, metric arrays that are the ones registered by
// f3d_set_metrics are read from the context's float copies (same values as
// the kernels' own float conversion; half the bytes).  Returns the copy of
// registered metric `m`, or nullptr.
static const float* registered_float(const f3d_ctx* c, const double* m) {
  if (c->precision != F3D_FP32 || !c->mf_base || !m) return nullptr;
  for (int q = 0; q < 9; ++q)
    if (c->metrics[q] == m) return c->mf[q];
  return (c->J == m) ? c->mf[9] : nullptr;
}

int f3d_real_size(const f3d_ctx* c) { return c ? (int)real_size(c) : F3D_ERR_ARG; }

int f3d_kinetics_source(const f3d_ctx* c) { return c ? c->kin_table : F3D_ERR_ARG; }

int f3d_flux_split(f3d_ctx* c, const double* Q, void* Fp, void* Fm, const double* Ax,
                   const double* Ay, const double* Az, const int64_t* tag, hipStream_t s) {
  CHECK_CTX(c);
  CHECK_PTR(Q, "Q");
  CHECK_PTR(Fp, "Fp");
  CHECK_PTR(Fm, "Fm");
  CHECK_PTR(Ax, "Ax");
  CHECK_PTR(Ay, "Ay");
  CHECK_PTR(Az, "Az");
  const float *fx = registered_float(c, Ax), *fy = registered_float(c, Ay),
              *fz = registered_float(c, Az);
  if (fx && fy && fz) return launch_flux_split(c, Q, Fp, Fm, fx, fy, fz, tag, s, true);
  return launch_flux_split(c, Q, Fp, Fm, Ax, Ay, Az, tag, s);
}

int f3d_species_split(f3d_ctx* c, const double* rhoi, const double* Q, void* Fp, void* Fm,
                      const double* Ax, const double* Ay, const double* Az, const int64_t* tag,
                      hipStream_t s) {
  CHECK_CTX(c);
  CHECK_PTR(rhoi, "rhoi");
  CHECK_PTR(Q, "Q");
  CHECK_PTR(Fp, "Fp");
  CHECK_PTR(Fm, "Fm");
  CHECK_PTR(Ax, "Ax");
  CHECK_PTR(Ay, "Ay");
  CHECK_PTR(Az, "Az");
  const float *fx = registered_float(c, Ax), *fy = registered_float(c, Ay),
              *fz = registered_float(c, Az);
  if (fx && fy && fz) return launch_species_split(c, rhoi, Q, Fp, Fm, fx, fy, fz, tag, s, true);
  return launch_species_split(c, rhoi, Q, Fp, Fm, Ax, Ay, Az, tag, s);
}

int f3d_shock_sensor(f3d_ctx* c, void* phi, const double* Q, hipStream_t s) {
  CHECK_CTX(c);
  CHECK_PTR(phi, "phi");
  CHECK_PTR(Q, "Q");
  return launch_shock_sensor(c, phi, Q, s);
}

int f3d_weno(f3d_ctx* c, int32_t dir, void* F, const void* phi, const void* Fp, const void* Fm,
             int32_t NV, const int64_t* tag, hipStream_t s) {
  CHECK_CTX(c);
  CHECK_PTR(F, "F");
  CHECK_PTR(phi, "phi");
  CHECK_PTR(Fp, "Fp");
  CHECK_PTR(Fm, "Fm");
  if (dir < 0 || dir > 2 || NV < 1 || NV > MAXNS) {
    set_error("f3d_weno: dir must be 0..2 and NV in [1, 16]");
    return F3D_ERR_ARG;
  }
  return launch_weno(c, dir, F, phi, Fp, Fm, NV, tag, s);
}

int f3d_viscous_flux(f3d_ctx* c, int32_t dir, void* Fv, const double* Q,
                     const double* const* metrics, const double* J, const void* lambda,
                     const void* mu, const void* Fh, const int64_t* tag, hipStream_t s) {
  CHECK_CTX(c);
  CHECK_PTR(Fv, "Fv");
  CHECK_PTR(Q, "Q");
  CHECK_PTR(metrics, "metrics");
  for (int q = 0; q < 9; ++q) CHECK_PTR(metrics[q], "metrics[q]");
  CHECK_PTR(J, "J");
  CHECK_PTR(lambda, "lambda");
  CHECK_PTR(mu, "mu");
  CHECK_PTR(Fh, "Fh");
  if (dir < 0 || dir > 2) {
    set_error("f3d_viscous_flux: dir must be 0..2");
    return F3D_ERR_ARG;
  }
  const void* Mv[9];
  const void* Mf[9];
  bool all = true;
  for (int q = 0; q < 9; ++q) {
    Mv[q] = metrics[q];
    Mf[q] = registered_float(c, metrics[q]);
    all = all && Mf[q];
  }
  const float* Jf = registered_float(c, J);
  if (all && Jf) return launch_viscous(c, dir, Fv, Q, Mf, Jf, lambda, mu, Fh, tag, s, true);
  return launch_viscous(c, dir, Fv, Q, Mv, J, lambda, mu, Fh, tag, s);
}

int f3d_spec_viscous_flux(f3d_ctx* c, int32_t dir, void* Fd, const double* Q, const void* Yi,
                          const double* const* metrics, const double* J, const void* D, void* Fh,
                          const int64_t* tag, hipStream_t s) {
  CHECK_CTX(c);
  CHECK_PTR(Fd, "Fd");
  CHECK_PTR(Q, "Q");
  CHECK_PTR(Yi, "Yi");
  CHECK_PTR(metrics, "metrics");
  for (int q = 0; q < 9; ++q) CHECK_PTR(metrics[q], "metrics[q]");
  CHECK_PTR(J, "J");
  CHECK_PTR(D, "D");
  CHECK_PTR(Fh, "Fh");
  if (dir < 0 || dir > 2) {
    set_error("f3d_spec_viscous_flux: dir must be 0..2");
    return F3D_ERR_ARG;
  }
  const void* Mv[9];
  const void* Mf[9];
  bool all = true;
  for (int q = 0; q < 9; ++q) {
    Mv[q] = metrics[q];
    Mf[q] = registered_float(c, metrics[q]);
    all = all && Mf[q];
  }
  const float* Jf = registered_float(c, J);
  if (all && Jf) return launch_spec_viscous(c, dir, Fd, Q, Yi, Mf, Jf, D, Fh, tag, s, true);
  return launch_spec_viscous(c, dir, Fd, Q, Yi, Mv, J, D, Fh, tag, s);
}

int f3d_div(f3d_ctx* c, double* U, const void* Fx, const void* Fy, const void* Fz,
            const void* Fvx, const void* Fvy, const void* Fvz, double dt, const double* J,
            const int64_t* tag, hipStream_t s) {
  CHECK_CTX(c);
  CHECK_PTR(U, "U");
  CHECK_PTR(Fx, "Fx");
  CHECK_PTR(Fy, "Fy");
  CHECK_PTR(Fz, "Fz");
  CHECK_PTR(Fvx, "Fvx");
  CHECK_PTR(Fvy, "Fvy");
  CHECK_PTR(Fvz, "Fvz");
  CHECK_PTR(J, "J");
  return launch_div(c, U, Fx, Fy, Fz, Fvx, Fvy, Fvz, dt, J, tag, s);
}

int f3d_div_specs(f3d_ctx* c, double* rhoi, const void* Fx, const void* Fy, const void* Fz,
                  const void* Fdx, const void* Fdy, const void* Fdz, double dt, const double* J,
                  const int64_t* tag, hipStream_t s) {
  CHECK_CTX(c);
  CHECK_PTR(rhoi, "rhoi");
  CHECK_PTR(Fx, "Fx");
  CHECK_PTR(Fy, "Fy");
  CHECK_PTR(Fz, "Fz");
  CHECK_PTR(Fdx, "Fdx");
  CHECK_PTR(Fdy, "Fdy");
  CHECK_PTR(Fdz, "Fdz");
  CHECK_PTR(J, "J");
  return launch_div_specs(c, rhoi, Fx, Fy, Fz, Fdx, Fdy, Fdz, dt, J, tag, s);
}

int f3d_mixture(f3d_ctx* c, const double* Q, const double* rhoi, void* Yi, void* lambda, void* mu,
                void* D, const int64_t* tag, hipStream_t s) {
  CHECK_CTX(c);
  CHECK_PTR(Q, "Q");
  CHECK_PTR(rhoi, "rhoi");
  CHECK_PTR(Yi, "Yi");
  CHECK_PTR(lambda, "lambda");
  CHECK_PTR(mu, "mu");
  CHECK_PTR(D, "D");
  return launch_mixture(c, Q, rhoi, Yi, lambda, mu, D, tag, s);
}

int f3d_c2prim(f3d_ctx* c, double* U, double* Q, double* rhoi, const int64_t* tag,
               hipStream_t s) {
  CHECK_CTX(c);
  CHECK_PTR(U, "U");
  CHECK_PTR(Q, "Q");
  CHECK_PTR(rhoi, "rhoi");
  return launch_c2prim(c, U, Q, rhoi, tag, s);
}

int f3d_prim2c(f3d_ctx* c, double* U, const double* Q, hipStream_t s) {
  CHECK_CTX(c);
  CHECK_PTR(U, "U");
  CHECK_PTR(Q, "Q");
  return launch_prim2c(c, U, Q, s);
}

int f3d_lincomb(f3d_ctx* c, double* U, const double* Un, int32_t NV, double a, double b,
                const int64_t* tag, hipStream_t s) {
  CHECK_CTX(c);
  CHECK_PTR(U, "U");
  CHECK_PTR(Un, "Un");
  if (NV < 1 || NV > MAXNS) {
    set_error("f3d_lincomb: NV must be in [1, 16]");
    return F3D_ERR_ARG;
  }
  return launch_lincomb(c, U, Un, NV, a, b, tag, s);
}

int f3d_get_y(f3d_ctx* c, double* Yi, const double* rhoi, const double* Q, const int64_t* tag,
              hipStream_t s) {
  CHECK_CTX(c);
  CHECK_PTR(Yi, "Yi");
  CHECK_PTR(rhoi, "rhoi");
  CHECK_PTR(Q, "Q");
  return launch_get_y(c, Yi, rhoi, Q, tag, s);
}

int f3d_set_boundary(f3d_ctx* c, const int32_t* kinds, const double* inflow_Q,
                     const double* inflow_rhoi) {
  CHECK_CTX(c);
  CHECK_PTR(kinds, "kinds");
  for (int f = 0; f < 6; ++f) {
    if (kinds[f] < F3D_BC_KEEP || kinds[f] > F3D_BC_INFLOW) {
      set_error("f3d_set_boundary: unknown fill kind");
      return F3D_ERR_ARG;
    }
    if (kinds[f] == F3D_BC_INFLOW && (!inflow_Q || !inflow_rhoi)) {
      set_error("f3d_set_boundary: F3D_BC_INFLOW needs inflow_Q and inflow_rhoi");
      return F3D_ERR_ARG;
    }
    const int N = (f < 2) ? c->g.Nx : (f < 4) ? c->g.Ny : c->g.Nz;
    if (kinds[f] == F3D_BC_PERIODIC && N < NG) {
      set_error("f3d_set_boundary: periodic direction needs N >= NG");
      return F3D_ERR_ARG;
    }
  }
  for (int f = 0; f < 6; ++f) c->bc[f] = kinds[f];
  if (inflow_Q)
    for (int q = 0; q < 7; ++q) c->inflow_Q[q] = inflow_Q[q];
  if (inflow_rhoi)
    for (int q = 0; q < c->g.Ns; ++q) c->inflow_rhoi[q] = inflow_rhoi[q];
  return F3D_OK;
}

int f3d_fill(f3d_ctx* c, double* Q, double* U, double* rhoi, hipStream_t s) {
  CHECK_CTX(c);
  CHECK_PTR(Q, "Q");
  CHECK_PTR(U, "U");
  CHECK_PTR(rhoi, "rhoi");
  return launch_fill(c, Q, U, rhoi, s);
}

int f3d_chem_stiff(f3d_ctx* c, double* U, double* Q, double* rhoi, double dt, const int64_t* tag,
                   hipStream_t s) {
  CHECK_CTX(c);
  CHECK_PTR(U, "U");
  CHECK_PTR(Q, "Q");
  CHECK_PTR(rhoi, "rhoi");
  if (c->g.Nr < 1) {
    set_error("f3d_chem_stiff: context has no reactions");
    return F3D_ERR_STATE;
  }
  return launch_chem(c, 0, U, Q, rhoi, dt, 0, 0, 0, nullptr, tag, s);
}

int f3d_chem_explicit(f3d_ctx* c, double* U, double* Q, double* rhoi, double dt,
                      const int64_t* tag, hipStream_t s) {
  CHECK_CTX(c);
  CHECK_PTR(U, "U");
  CHECK_PTR(Q, "Q");
  CHECK_PTR(rhoi, "rhoi");
  if (c->g.Nr < 1) {
    set_error("f3d_chem_explicit: context has no reactions");
    return F3D_ERR_STATE;
  }
  return launch_chem(c, 1, U, Q, rhoi, dt, 0, 0, 0, nullptr, tag, s);
}

int f3d_chem_newton(f3d_ctx* c, double* U, double* Q, double* rhoi, double dt, int32_t maxiter,
                    double rtol, double ttol, int32_t* iters, const int64_t* tag, hipStream_t s) {
  CHECK_CTX(c);
  CHECK_PTR(U, "U");
  CHECK_PTR(Q, "Q");
  CHECK_PTR(rhoi, "rhoi");
  if (c->g.Nr < 1) {
    set_error("f3d_chem_newton: context has no reactions");
    return F3D_ERR_STATE;
  }
  if (maxiter < 1 || maxiter > 1000 || !(rtol > 0) || !(ttol > 0)) {
    set_error("f3d_chem_newton: need 1 <= maxiter <= 1000, rtol > 0, ttol > 0");
    return F3D_ERR_ARG;
  }
  return launch_chem(c, 2, U, Q, rhoi, dt, maxiter, rtol, ttol, iters, tag, s);
}

#define RET(x)              \
  do {                      \
    int _rc = (x);          \
    if (_rc) return _rc;    \
  } while (0)

int f3d_set_metrics(f3d_ctx* c, const double* const* metrics, const double* J) {
  CHECK_CTX(c);
  CHECK_PTR(metrics, "metrics");
  for (int q = 0; q < 9; ++q) CHECK_PTR(metrics[q], "metrics[q]");
  CHECK_PTR(J, "J");
  for (int q = 0; q < 9; ++q) c->metrics[q] = metrics[q];
  c->J = J;
  if (c->precision == F3D_FP32) {
    const long NP = c->g.NP;
    if (!c->mf_base) {
      if (hipMalloc(&c->mf_base, sizeof(float) * NP * 10) != hipSuccess) {
        c->mf_base = nullptr;
        set_error("f3d_set_metrics: allocation of the float metric copies failed");
        return F3D_ERR_ALLOC;
      }
      for (int q = 0; q < 10; ++q) c->mf[q] = (float*)c->mf_base + (size_t)q * NP;
    }
    for (int q = 0; q < 9; ++q) RET(launch_to_float(c, c->mf[q], metrics[q], NP, nullptr));
    RET(launch_to_float(c, c->mf[9], J, NP, nullptr));
    RET(hip_check(hipDeviceSynchronize(), "f3d_set_metrics"));
  }
  return F3D_OK;
}

static int sweep_ready(f3d_ctx* c, const char* what) {
  if (c->precision != F3D_FP32 || !c->mf[0]) {
    set_error(std::string(what) + ": needs an F3D_FP32 context after f3d_set_metrics");
    return F3D_ERR_STATE;
  }
  return F3D_OK;
}

int f3d_sweep_conv(f3d_ctx* c, int32_t dir, float* rhs, const double* Q, const double* rhoi,
                   const float* phi, int32_t accumulate, hipStream_t s) {
  CHECK_CTX(c);
  CHECK_PTR(rhs, "rhs");
  CHECK_PTR(Q, "Q");
  CHECK_PTR(rhoi, "rhoi");
  CHECK_PTR(phi, "phi");
  if (dir < 0 || dir > 2) {
    set_error("f3d_sweep_conv: dir must be 0, 1 or 2");
    return F3D_ERR_ARG;
  }
  RET(sweep_ready(c, "f3d_sweep_conv"));
  return launch_sweep_conv(c, dir, rhs, Q, rhoi, phi, accumulate != 0, s);
}

int f3d_sweep_diff(f3d_ctx* c, int32_t dir, float* rhs, const double* Q, const float* Yi,
                   const float* lambda, const float* mu, const float* D, hipStream_t s) {
  CHECK_CTX(c);
  CHECK_PTR(rhs, "rhs");
  CHECK_PTR(Q, "Q");
  CHECK_PTR(Yi, "Yi");
  CHECK_PTR(lambda, "lambda");
  CHECK_PTR(mu, "mu");
  CHECK_PTR(D, "D");
  if (dir < 0 || dir > 2) {
    set_error("f3d_sweep_diff: dir must be 0, 1 or 2");
    return F3D_ERR_ARG;
  }
  RET(sweep_ready(c, "f3d_sweep_diff"));
  return launch_sweep_diff(c, dir, rhs, Q, Yi, lambda, mu, D, s);
}

int f3d_sweep_stage_end(f3d_ctx* c, int32_t KRK, double* U, const double* Un, double* rhoi,
                        const double* rhon, double* Q, const float* rhs, double dt,
                        hipStream_t s) {
  CHECK_CTX(c);
  CHECK_PTR(U, "U");
  CHECK_PTR(rhoi, "rhoi");
  CHECK_PTR(Q, "Q");
  CHECK_PTR(rhs, "rhs");
  if (KRK < 1 || KRK > 3 || (KRK > 1 && (!Un || !rhon))) {
    set_error("f3d_sweep_stage_end: KRK must be 1, 2 or 3 (2 and 3 need Un, rhon)");
    return F3D_ERR_ARG;
  }
  RET(sweep_ready(c, "f3d_sweep_stage_end"));
  return launch_sweep_stage_end(c, KRK, U, Un, rhoi, rhon, Q, rhs, dt, c->J, s);
}

void* f3d_workspace(f3d_ctx* c, const char* name) {
  if (!c || !name || !c->ws_base) return nullptr;
  for (int w = 0; w < W_COUNT; ++w)
    if (std::strcmp(name, kWsNames[w]) == 0) return c->ws[w];
  return nullptr;
}

int f3d_rk3_stage(f3d_ctx* c, int32_t KRK, double* U, double* Un, double* rhoi, double* rhon,
                  double* Q, double dt, const int64_t* tag, hipStream_t s) {
  CHECK_CTX(c);
  CHECK_PTR(U, "U");
  CHECK_PTR(Un, "Un");
  CHECK_PTR(rhoi, "rhoi");
  CHECK_PTR(rhon, "rhon");
  CHECK_PTR(Q, "Q");
  if (KRK < 1 || KRK > 3) {
    set_error("f3d_rk3_stage: KRK must be 1, 2 or 3");
    return F3D_ERR_ARG;
  }
  if (!c->J) {
    set_error("f3d_rk3_stage: call f3d_set_metrics first");
    return F3D_ERR_STATE;
  }
  RET(ensure_workspace(c));
  const Geom& g = c->g;
  void** w = c->ws;
  // metrics: the caller's FP64 arrays, or the float copies on the FP32 path
  const bool mf = (c->precision == F3D_FP32);
  const void* M[9];
  for (int q = 0; q < 9; ++q) M[q] = mf ? (const void*)c->mf[q] : (const void*)c->metrics[q];
  const void* Jm = mf ? (const void*)c->mf[9] : (const void*)c->J;
  const double* J = c->J;
  if (KRK == 1) {
    RET(hip_check(hipMemcpyAsync(Un, U, sizeof(double) * (size_t)g.NP * 5,
                                 hipMemcpyDeviceToDevice, s), "copy Un"));
    RET(hip_check(hipMemcpyAsync(rhon, rhoi, sizeof(double) * (size_t)g.NP * g.Ns,
                                 hipMemcpyDeviceToDevice, s), "copy rhon"));
  }
  RET(launch_mixture(c, Q, rhoi, w[W_Yi], w[W_lam], w[W_mu], w[W_D], tag, s));
  RET(launch_shock_sensor(c, w[W_phi], Q, s));
  if (mf && !tag && c->stage_kind == F3D_STAGE_FUSED) {
    // fused FP32 stage: per-direction sweeps into one right-hand side, then
    // divSpecs + div + linComb + c2Prim in one pass
    float* rhs = (float*)w[W_RHS];
    RET(launch_sweeps(c, rhs, Q, rhoi, w[W_phi], w[W_Yi], w[W_lam], w[W_mu], w[W_D], s));
    RET(launch_sweep_stage_end(c, KRK, U, Un, rhoi, rhon, Q, rhs, dt, J, s));
    return launch_fill(c, Q, U, rhoi, s);
  }
  // specAdvance, solver.jl:40-56
  for (int d = 0; d < 3; ++d) {
    RET(launch_species_split(c, rhoi, Q, w[W_Fpi], w[W_Fmi], M[3 * d], M[3 * d + 1],
                             M[3 * d + 2], tag, s, mf));
    RET(launch_weno(c, d, w[W_Fxi + d], w[W_phi], w[W_Fpi], w[W_Fmi], g.Ns, tag, s));
  }
  for (int d = 0; d < 3; ++d)
    RET(launch_spec_viscous(c, d, w[W_Fdx + d], Q, w[W_Yi], M, Jm, w[W_D], w[W_Fhx + d], tag, s,
                            mf));
  // (divSpecs is applied with div in launch_stage_end: flowAdvance does not read rho_k)
  // flowAdvance, solver.jl:22-38
  for (int d = 0; d < 3; ++d) {
    RET(launch_flux_split(c, Q, w[W_Fp], w[W_Fm], M[3 * d], M[3 * d + 1], M[3 * d + 2], tag, s,
                          mf));
    RET(launch_weno(c, d, w[W_Fx + d], w[W_phi], w[W_Fp], w[W_Fm], 5, tag, s));
  }
  for (int d = 0; d < 3; ++d)
    RET(launch_viscous(c, d, w[W_Fvx + d], Q, M, Jm, w[W_lam], w[W_mu], w[W_Fhx + d], tag, s,
                       mf));
  // divSpecs + div + linComb (stages 2, 3) + c2Prim in one pass (k_stage_end)
  void* const F[12] = {w[W_Fx],  w[W_Fy],  w[W_Fz],  w[W_Fvx], w[W_Fvy], w[W_Fvz],
                       w[W_Fxi], w[W_Fyi], w[W_Fzi], w[W_Fdx], w[W_Fdy], w[W_Fdz]};
  RET(launch_stage_end(c, KRK, U, Un, rhoi, rhon, Q, F, dt, J, tag, s));
  RET(launch_fill(c, Q, U, rhoi, s));
  return F3D_OK;
}

// One time step, solver.jl:240-378: chemistry over dt/2 (sub_step substeps
// for the stiff / explicit forms, run.jl:23) and the fills, the three RK3
// stages, chemistry over dt/2 and the fills.
int f3d_time_step(f3d_ctx* c, double* U, double* Un, double* rhoi, double* rhon, double* Q,
                  double dt, int32_t chem, int32_t sub_step, int32_t maxiter, double rtol,
                  double ttol, int64_t* n_fallback, const int64_t* tag, hipStream_t s) {
  CHECK_CTX(c);
  if (chem < F3D_CHEM_NONE || chem > F3D_CHEM_SURROGATE || sub_step < 1) {
    set_error("f3d_time_step: chem must be one of F3D_CHEM_* and sub_step >= 1");
    return F3D_ERR_ARG;
  }
  auto half = [&]() -> int {
    const double h = 0.5 * dt;
    switch (chem) {
      case F3D_CHEM_NONE: return F3D_OK;
      case F3D_CHEM_STIFF:
        for (int k = 0; k < sub_step; ++k)
          RET(f3d_chem_stiff(c, U, Q, rhoi, h / sub_step, tag, s));
        break;
      case F3D_CHEM_EXPLICIT:
        for (int k = 0; k < sub_step; ++k)
          RET(f3d_chem_explicit(c, U, Q, rhoi, h / sub_step, tag, s));
        break;
      case F3D_CHEM_NEWTON:
        RET(f3d_chem_newton(c, U, Q, rhoi, h, maxiter, rtol, ttol, nullptr, tag, s));
        break;
      default: {
        int64_t nfb = 0;
        RET(f3d_nn_chem(c, U, Q, rhoi, h, maxiter, rtol, ttol, nullptr, &nfb, tag, s));
        if (n_fallback) *n_fallback += nfb;
      }
    }
    return f3d_fill(c, Q, U, rhoi, s);
  };
  RET(half());
  for (int k = 1; k <= 3; ++k) RET(f3d_rk3_stage(c, k, U, Un, rhoi, rhon, Q, dt, tag, s));
  return half();
}

__global__ void __launch_bounds__(kBlock)
    k_nan_check(const double* __restrict__ U, long n, int32_t* __restrict__ flag) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && U[i] != U[i]) flag[0] = 1;  // racing stores of the same value
}

int f3d_nan_check(f3d_ctx* c, const double* U, int32_t* flag, hipStream_t s) {
  CHECK_CTX(c);
  CHECK_PTR(U, "U");
  CHECK_PTR(flag, "flag");
  if (!c->nan_flag && hipMalloc((void**)&c->nan_flag, sizeof(int32_t)) != hipSuccess) {
    c->nan_flag = nullptr;
    set_error("f3d_nan_check: allocation failed");
    return F3D_ERR_ALLOC;
  }
  int32_t* d = c->nan_flag;
  RET(hip_check(hipMemsetAsync(d, 0, sizeof(int32_t), s), "f3d_nan_check"));
  const long n = (long)c->g.NP * 5;
  hipLaunchKernelGGL(k_nan_check, dim3(nblocks(n)), dim3(kBlock), 0, s, U, n, d);
  int rc = hip_check(hipGetLastError(), "f3d_nan_check");
  int32_t h = 0;
  if (!rc) rc = hip_check(hipMemcpyAsync(&h, d, sizeof(int32_t), hipMemcpyDeviceToHost, s),
                          "f3d_nan_check");
  if (!rc) rc = hip_check(hipStreamSynchronize(s), "f3d_nan_check");
  *flag = h;
  return rc;
}

}  // extern "C"
