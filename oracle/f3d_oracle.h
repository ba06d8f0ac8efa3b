/* FP64 CPU oracle for the Flame3D.jl per-stage right-hand side.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates the reference's Julia GPU
 * kernels as plain sequential C++ (OpenMP over the slowest loop) so that the
 * HIP kernels can be checked against it.  Only tests/, the smoke check and the
 * benchmark's CPU-baseline leg load it; nothing in the product links it.
 *
 * Arrays use the reference layout: Julia column-major [i, j, k, var] with NG
 * ghost layers, i fastest, variable slowest (solver.jl:81-167).  Face arrays
 * are [(Nx+1), Ny, Nz, NV] etc. without ghosts.  All arithmetic is FP64 with
 * the reference's operation order and no contraction (-ffp-contract=off).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t Nx, Ny, Nz, NG, Ns, Nr;
} or_dims;

/* thermoProperty (run.jl:71-82) */
typedef struct {
  double Ru, min_temp, max_temp;
  const double *mw, *coeffs_sep, *coeffs_lo, *coeffs_hi;
  const double *visc_poly, *conduct_poly, *binarydiff_poly;
} or_thermo;

/* reactionProperty (run.jl:84-94) */
typedef struct {
  double atm;
  const int64_t *reaction_type, *sgm, *vf, *vr;
  const double *Arr, *ef, *loP, *Troe;
} or_kinetics;

/* scheme constants (run.jl:51-57, :12) */
typedef struct {
  double SWeps2, WENOeps, phi1, phi2;
  double UP7[7];
  double T_criteria;
  int32_t sensor_kind;  /* 0 the reference's pressure sensor, 1 Ducros */
  int32_t recon_kind;   /* 0 the reference's UP7 / WENO5-Z / NND, 1 WENO7-Z in the middle branch */
} or_scheme;

/* ---- pointwise thermo / kinetics (thermo.jl, reactions.jl) ---- */
double or_internal_energy(const or_dims* d, const or_thermo* th, double T, const double* rhoi);
double or_cv(const or_dims* d, const or_thermo* th, double T, const double* rhoi);
double or_get_T(const or_dims* d, const or_thermo* th, double e, const double* rhoi,
                double tol, int32_t maxiter, double tmin, double tmax, int32_t* iters);
double or_get_T_gap(const or_dims* d, const or_thermo* th, double e, const double* rhoi,
                    double tol, int32_t maxiter, double tmin, double tmax);
double or_get_T_from(const or_dims* d, const or_thermo* th, double e, const double* rhoi,
                     double T_guess, double tol, int32_t maxiter, double tmin, double tmax);
void or_h_specs(const or_dims* d, const or_thermo* th, double T, double* h);
/* exp / log of the chemistry rates (the library's f3d_exp / f3d_log restated) */
double or_exp(double x);
double or_log(double x);
void or_gibbs(const or_dims* d, const or_thermo* th, double T, double* g);
void or_production_rate(const or_dims* d, const or_thermo* th, const or_kinetics* kin,
                        const double* sc, double T, double jac_eps, int32_t full_jac,
                        double* wdot, double* Arate /* [Ns,Ns] or NULL */);
void or_production_rate_dT(const or_dims* d, const or_thermo* th, const or_kinetics* kin,
                           const double* sc, double T, double* wdot, double* dwdT);
void or_gauss_solve(int32_t n, double* a /* [n,n] col-major, destroyed */,
                    double* b /* destroyed */, double* x);
void or_set_diffusivity_reference_form(int32_t on);
/* operation counts of the chemistry path: out[6] = add/sub, mul, div,
   transcendental, sqrt, iterations */
void or_newton_opcount(const or_dims* d, const or_thermo* th, const or_kinetics* kin,
                       const double* rhoi, double T0, double h, int32_t maxiter, double rtol,
                       double ttol, double* out);
void or_rates_opcount(const or_dims* d, const or_thermo* th, const or_kinetics* kin,
                      const double* rhoi, double T, int32_t full_jac, double* out);
void or_getT_update_opcount(const or_dims* d, const or_thermo* th, const double* rhoi, double T,
                            double ein, double* out);
void or_gauss_opcount(int32_t n, const double* a, const double* b, double* out);
void or_mixture_properties(const or_dims* d, const or_thermo* th, double T, double P,
                           const double* X, double* Diff, double* lambda, double* mu);

/* ---- kernels (one per @cuda line of solver.jl) ---- */
void or_flux_split(const or_dims* d, const double* Q, double* Fp, double* Fm,
                   const double* Ax, const double* Ay, const double* Az,
                   const int64_t* tag, const or_scheme* sc);
void or_species_split(const or_dims* d, const double* rhoi, const double* Q, double* Fp,
                      double* Fm, const double* Ax, const double* Ay, const double* Az,
                      const int64_t* tag);
void or_shock_sensor(const or_dims* d, double* phi, const double* Q);
double or_weno7z(const double* v7, double eps);
void or_ducros_sensor(const or_dims* d, double* phi, const double* Q, const double* const* M,
                      const double* J);
void or_weno(const or_dims* d, int32_t dir, double* F, const double* phi, const double* Fp,
             const double* Fm, int32_t NV, const int64_t* tag, const or_scheme* sc);
void or_viscous_flux(const or_dims* d, int32_t dir, double* Fv, const double* Q,
                     const double* const* metrics, const double* J, const double* lambda,
                     const double* mu, const double* Fh, const int64_t* tag);
void or_spec_viscous_flux(const or_dims* d, int32_t dir, double* Fd, const double* Q,
                          const double* Yi, const double* const* metrics, const double* J,
                          const double* D, double* Fh, const or_thermo* th,
                          const int64_t* tag);
void or_div(const or_dims* d, double* U, const double* Fx, const double* Fy,
            const double* Fz, const double* Fvx, const double* Fvy, const double* Fvz,
            double dt, const double* J, const int64_t* tag);
void or_div_specs(const or_dims* d, double* rhoi, const double* Fx, const double* Fy,
                  const double* Fz, const double* Fdx, const double* Fdy, const double* Fdz,
                  double dt, const double* J, const int64_t* tag);
void or_mixture(const or_dims* d, const double* Q, const double* rhoi, double* Yi,
                double* lambda, double* mu, double* D, const or_thermo* th,
                const int64_t* tag);
void or_c2prim(const or_dims* d, double* U, double* Q, double* rhoi, const or_thermo* th,
               const int64_t* tag);
void or_prim2c(const or_dims* d, double* U, const double* Q);
void or_lincomb(const or_dims* d, double* U, const double* Un, int32_t NV, double a,
                double b, const int64_t* tag);
void or_get_y(const or_dims* d, double* Yi, const double* rhoi, const double* Q,
              const int64_t* tag);
/* boundary fills; kinds[6] = {x-, x+, y-, y+, z-, z+}:
 * 0 keep, 1 periodic, 2 copy (zero gradient, first interior node included),
 * 3 inflow (fixed state inflow_Q[7], inflow_rhoi[Ns], first node included) */
void or_fill(const or_dims* d, const int32_t* kinds, double* Q, double* U, double* rhoi,
             const double* inflow_Q, const double* inflow_rhoi);
void or_chem_stiff(const or_dims* d, double* U, double* Q, double* rhoi, double dt,
                   const or_thermo* th, const or_kinetics* kin, const or_scheme* sc,
                   const int64_t* tag, int32_t bug_compat);
void or_chem_explicit(const or_dims* d, double* U, double* Q, double* rhoi, double dt,
                      const or_thermo* th, const or_kinetics* kin, const or_scheme* sc,
                      const int64_t* tag);
/* backward-Euler Newton on (rho_1..rho_Ns, T) with a dense (Ns+1) LU; a
 * cell whose iterate leaves the admissible set restarts with 2, 4, .. 64
 * substeps; iters = total iterations (negated if every attempt failed, the
 * cell then left unreacted) */
void or_chem_newton(const or_dims* d, double* U, double* Q, double* rhoi, double dt,
                    const or_thermo* th, const or_kinetics* kin, const or_scheme* sc,
                    const int64_t* tag, int32_t maxiter, double rtol, double ttol,
                    int32_t* iters);

/* one Shu-Osher RK3 stage (solver.jl:298-326): mixture, sensor, specAdvance,
 * flowAdvance, linComb, c2Prim, fills.  ws holds the intermediate arrays:
 * Yi, lambda, mu, D, phi, Fp, Fm, Fp_i, Fm_i, Fx, Fy, Fz, Fvx, Fvy, Fvz, Fx_i, Fy_i,
 * Fz_i, Fdx, Fdy, Fdz, Fhx, Fhy, Fhz (24 pointers, reference shapes). */
void or_rk3_stage(const or_dims* d, int32_t KRK, double* U, double* Un, double* rhoi,
                  double* rhon, double* Q, double* const* ws, const double* const* metrics,
                  const double* J, double dt, const int64_t* tag, const or_thermo* th,
                  const or_scheme* sc, const int32_t* fill_kinds, const double* inflow_Q,
                  const double* inflow_rhoi);

/* 0-D constant-volume point-implicit reactor step in the form of
 * NN/H2/chemical_test.cpp:950-1007 (for cross-checking the primitives) */
void or_reactor0d_step(const or_dims* d, const or_thermo* th, const or_kinetics* kin,
                       double* rhoi, double* E, double dt, double jac_eps, double gt_tol,
                       int32_t gt_maxiter, double gt_tmin, double gt_tmax, double* T_out,
                       int32_t* gt_iters);

int32_t or_num_threads(void);

#ifdef __cplusplus
}
#endif
