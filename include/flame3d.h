/* flame3d.h -- C-ABI of libflame3d.so, the MI355X (gfx950) per-RK-stage
 * right-hand side of Flame3D.jl.
 *
 * Every entry point replaces one `@cuda` launch (or one host function) of the
 * reference's Julia driver, with the same array arguments in the same layout:
 * Julia column-major [i, j, k, var] with NG = 4 ghost layers (i fastest,
 * variable slowest), Float64 state arrays owned by the caller on the device.
 * Face arrays are [(Nx+1), Ny, Nz, NV], [Nx, (Ny+1), Nz, NV], [Nx, Ny, (Nz+1), NV].
 * Intermediate arrays (split fluxes, face fluxes, sensor, transport) use the
 * context's flux precision: double for F3D_FP64, float for F3D_FP32 - they
 * are passed as void*.  The state U, Q, rhoi, metrics and J are always double.
 *
 * All functions return 0 on success and a negative F3D_ERR_* code on failure
 * (f3d_last_error() gives the message); nothing throws across the ABI.
 * Launches are asynchronous on the given HIP stream (0 = default stream),
 * in stream order, exactly as the reference's launches are on its stream.
 * `tag` may be NULL (no immersed boundary: every point is fluid); when given
 * it is the reference's Int64 IBM tag array and kernels skip solid points
 * as the reference does.
 *
 * A Julia caller binds these with ccall.
 */
#ifndef FLAME3D_H
#define FLAME3D_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define F3D_OK 0
#define F3D_ERR_ARG (-1)
#define F3D_ERR_HIP (-2)
#define F3D_ERR_ALLOC (-3)
#define F3D_ERR_UNSUPPORTED (-4)
#define F3D_ERR_STATE (-5)

#define F3D_FP64 0
#define F3D_FP32 1

/* fill kinds per face (f3d_set_boundary) */
#define F3D_BC_KEEP 0     /* ghosts untouched (fixed Dirichlet state) */
#define F3D_BC_PERIODIC 1 /* ghost <- opposite interior */
#define F3D_BC_COPY 2     /* zero gradient, boundary node included (boundary.jl:59-84) */
#define F3D_BC_INFLOW 3   /* fixed uniform state, boundary node included (boundary.jl:1-43) */

typedef struct f3d_ctx f3d_ctx;

/* Block geometry: Nx (= Nxp, this GPU's share), Ny, Nz interior points,
 * NG ghost layers (must be 4, parse_mesh.jl:5), Ns species, Nr reactions
 * (run.jl:13-14, :59-63). */
typedef struct {
  int32_t Nx, Ny, Nz, NG, Ns, Nr;
} f3d_dims;

/* thermoProperty (run.jl:71-82) as initThermo fills it (thermo.jl:405-436):
 * host pointers, column-major. */
typedef struct {
  double Ru, min_temp, max_temp;
  const double* mw;              /* [Ns] kg/mol */
  const double* coeffs_sep;      /* [Ns] NASA-7 mid temperature */
  const double* coeffs_lo;       /* [7, Ns] */
  const double* coeffs_hi;       /* [7, Ns] */
  const double* visc_poly;       /* [5, Ns] */
  const double* conduct_poly;    /* [5, Ns] */
  const double* binarydiff_poly; /* [5, Ns, Ns] */
} f3d_thermo;

/* reactionProperty (run.jl:84-94) as initReact fills it (reactions.jl:506-576). */
typedef struct {
  double atm;
  const int64_t* reaction_type; /* [Nr] 1 Arrhenius, 2 three-body, 3 Troe */
  const int64_t* sgm;           /* [Nr] */
  const int64_t* vf;            /* [Ns, Nr] */
  const int64_t* vr;            /* [Ns, Nr] */
  const double* Arr;            /* [3, Nr] A, b, Ea/R */
  const double* ef;             /* [Ns, Nr] */
  const double* loP;            /* [3, Nr] */
  const double* Troe;           /* [4, Nr] */
} f3d_kinetics;

/* Scheme constants (run.jl:12, :51-57) plus the choices this library adds. */
typedef struct {
  double SWeps2, WENOeps, hybrid_phi1, hybrid_phi2;
  double UP7[7];
  double T_criteria;
  int32_t precision;   /* F3D_FP64 or F3D_FP32 (flux/viscous/divergence arithmetic) */
  int32_t bug_compat;  /* bit 0: eval_gpu_stiff's transposed coeffs_lo[n,6] read
                          (reactions.jl:282), valid for Ns >= 6 */
  int32_t sensor_kind; /* F3D_SENSOR_PRESSURE: the reference's shockSensor (schemes.jl:1-24);
                          F3D_SENSOR_DUCROS: theta = (div u)^2 / ((div u)^2 + |curl u|^2 + 1e-30)
                          from central differences and the metrics (needs f3d_set_metrics);
                          theta in [0, 1] then drives the same switch: hybrid_phi1 = 0.1,
                          hybrid_phi2 > 1 gives UP7 below 0.1 and WENO5-Z above */
  int32_t recon_kind;  /* F3D_RECON_REFERENCE: UP7 / WENO5-Z / NND (schemes.jl:86-437);
                          F3D_RECON_WENO7: WENO7-Z (Balsara-Shu indicators, Z weights with
                          tau7 = |b0 + 3 b1 - 3 b2 - b3|) in the middle branch */
  int32_t stage_kind;  /* F3D_FP32 contexts: F3D_STAGE_FUSED (default) runs f3d_rk3_stage as
                          per-direction fused sweeps that accumulate every face-flux
                          difference into one float right-hand side (split and face fluxes
                          stay on chip); F3D_STAGE_GRANULAR replays the reference's launch
                          sequence (solver.jl:22-56) with the kernel-granular entry points.
                          F3D_FP64 contexts and calls with a tag always take the latter. */
  int32_t reserved[3];
} f3d_scheme;
#define F3D_SENSOR_PRESSURE 0
#define F3D_SENSOR_DUCROS 1
#define F3D_RECON_REFERENCE 0
#define F3D_RECON_WENO7 1
#define F3D_STAGE_FUSED 0
#define F3D_STAGE_GRANULAR 1

/* ---- context ------------------------------------------------------------ */
/* Copies the parameter bundles to the device.  Replaces the Adapt-ed structs
 * of run.jl:96-97 and the compile-time constants of run.jl:6-69. */
int f3d_create(const f3d_dims* dims, const f3d_thermo* thermo, const f3d_kinetics* kin,
               const f3d_scheme* scheme, f3d_ctx** out);
int f3d_destroy(f3d_ctx* ctx);
const char* f3d_last_error(void);
/* precision of the context's intermediate arrays (sizeof element: 8 or 4) */
int f3d_real_size(const f3d_ctx* ctx);
/* Where the context's chemistry kernels take the reaction parameters from
 * (no counterpart in the reference, whose kernels read reactionProperty):
 * F3D_KIN_RUNTIME, the parameter block copied from f3d_kinetics (any
 * mechanism), or the tables compiled in for a mechanism shipped with the
 * library, used when every kinetics parameter of the context equals the
 * table's.  Results are the same either way; the tables let the compiler
 * resolve the stoichiometry. */
#define F3D_KIN_RUNTIME 0
#define F3D_KIN_H2_LIDRYER 1
#define F3D_KIN_AIR5 2
int f3d_kinetics_source(const f3d_ctx* ctx);

/* ---- kernel-granular entry points (one per @cuda line of solver.jl) ----- */
/* fluxSplit, split.jl:17-89: Steger-Warming F+/F- at every padded point.
 * Q [..,7], Fp/Fm [..,5]; Ax/Ay/Az the three metrics of the sweep. */
int f3d_flux_split(f3d_ctx* ctx, const double* Q, void* Fp, void* Fm, const double* Ax,
                   const double* Ay, const double* Az, const int64_t* tag, hipStream_t s);
/* split, split.jl:101-130: species F+/F- = 1/2(Un +- |Un|) rho_k. */
int f3d_species_split(f3d_ctx* ctx, const double* rhoi, const double* Q, void* Fp, void* Fm,
                      const double* Ax, const double* Ay, const double* Az, const int64_t* tag,
                      hipStream_t s);
/* shockSensor, schemes.jl:1-24: phi [padded] from Q[.., 5]. */
int f3d_shock_sensor(f3d_ctx* ctx, void* phi, const double* Q, hipStream_t s);
/* WENO_x / WENO_y / WENO_z, schemes.jl:86-437 (dir = 0, 1, 2): hybrid
 * UP7 / WENO5-Z / NND face flux; NV = 5 (flow) or Ns (species). */
int f3d_weno(f3d_ctx* ctx, int32_t dir, void* F, const void* phi, const void* Fp, const void* Fm,
             int32_t NV, const int64_t* tag, hipStream_t s);
/* viscousFlux_x/y/z, viscous.jl:1-428.  metrics[9] = dxi/dx, dxi/dy, dxi/dz,
 * deta/dx, deta/dy, deta/dz, dzeta/dx, dzeta/dy, dzeta/dz (device pointers);
 * lambda, mu [padded]; Fh [faces, 3] from f3d_spec_viscous_flux. */
int f3d_viscous_flux(f3d_ctx* ctx, int32_t dir, void* Fv, const double* Q,
                     const double* const* metrics, const double* J, const void* lambda,
                     const void* mu, const void* Fh, const int64_t* tag, hipStream_t s);
/* specViscousFlux_x/y/z, viscous.jl:430-665: Fd [faces, Ns], Fh [faces, 3]. */
int f3d_spec_viscous_flux(f3d_ctx* ctx, int32_t dir, void* Fd, const double* Q, const void* Yi,
                          const double* const* metrics, const double* J, const void* D, void* Fh,
                          const int64_t* tag, hipStream_t s);
/* div, div.jl:16-54: U += J dt (sum of face-flux differences). */
int f3d_div(f3d_ctx* ctx, double* U, const void* Fx, const void* Fy, const void* Fz,
            const void* Fvx, const void* Fvy, const void* Fvz, double dt, const double* J,
            const int64_t* tag, hipStream_t s);
/* divSpecs, div.jl:61-84. */
int f3d_div_specs(f3d_ctx* ctx, double* rhoi, const void* Fx, const void* Fy, const void* Fz,
                  const void* Fdx, const void* Fdy, const void* Fdz, double dt, const double* J,
                  const int64_t* tag, hipStream_t s);
/* mixture, thermo.jl:371-403: Yi, lambda, mu, D at every padded point. */
int f3d_mixture(f3d_ctx* ctx, const double* Q, const double* rhoi, void* Yi, void* lambda,
                void* mu, void* D, const int64_t* tag, hipStream_t s);
/* c2Prim, utils.jl:2-48 (FP64, e->T Newton). */
int f3d_c2prim(f3d_ctx* ctx, double* U, double* Q, double* rhoi, const int64_t* tag,
               hipStream_t s);
/* prim2c, utils.jl:71-92. */
int f3d_prim2c(f3d_ctx* ctx, double* U, const double* Q, hipStream_t s);
/* linComb, utils.jl:95-112: U = a U + b Un on NV variables. */
int f3d_lincomb(f3d_ctx* ctx, double* U, const double* Un, int32_t NV, double a, double b,
                const int64_t* tag, hipStream_t s);
/* getY, utils.jl:51-68 (Yi in double). */
int f3d_get_y(f3d_ctx* ctx, double* Yi, const double* rhoi, const double* Q, const int64_t* tag,
              hipStream_t s);
/* ---- boundary fills (replace fillGhost/fillSpec, boundary.jl:180-191, and
 * the MPI halo exchange mpi.jl:2-37 on a single GPU) ---------------------- */
/* kinds[6] = {x-, x+, y-, y+, z-, z+} from F3D_BC_*; inflow_Q[7] and
 * inflow_rhoi[Ns] (host) give the F3D_BC_INFLOW state. */
int f3d_set_boundary(f3d_ctx* ctx, const int32_t* kinds, const double* inflow_Q,
                     const double* inflow_rhoi);
int f3d_fill(f3d_ctx* ctx, double* Q, double* U, double* rhoi, hipStream_t s);

/* ---- chemistry (FP64) --------------------------------------------------- */
/* eval_gpu_stiff, reactions.jl:235-309: one point-implicit step of dt. */
int f3d_chem_stiff(f3d_ctx* ctx, double* U, double* Q, double* rhoi, double dt,
                   const int64_t* tag, hipStream_t s);
/* eval_gpu, reactions.jl:176-232: one explicit Euler step of dt. */
int f3d_chem_explicit(f3d_ctx* ctx, double* U, double* Q, double* rhoi, double dt,
                      const int64_t* tag, hipStream_t s);
/* Backward-Euler Newton on (rho_1..rho_Ns, T) with a dense (Ns+1)x(Ns+1)
 * pivoted LU per cell; maxiter bounds the iterations of one solve.  A cell
 * whose iterate leaves the admissible set (non-finite, T outside
 * [min_temp/2, 1.5 max_temp], rho_k < -rho) or does not converge restarts
 * from its initial state with 2, 4, .., 64 backward-Euler substeps of dt.
 * iters [Nx*Ny*Nz] (device, may be NULL) receives the total iteration count
 * of each interior cell, negated when every attempt failed (that cell is
 * then left unreacted).  No reference counterpart (extends eval_gpu_stiff,
 * whose first iterate it reproduces; run.jl:23 sub-cycles chemistry too). */
int f3d_chem_newton(f3d_ctx* ctx, double* U, double* Q, double* rhoi, double dt,
                    int32_t maxiter, double rtol, double ttol, int32_t* iters,
                    const int64_t* tag, hipStream_t s);

/* ---- learned chemistry surrogate (BF16 MFMA MLP 11-64-64-64-9, GELU) -----
 * Reference pattern: pre_input / evalModel / post_predict, reactions.jl:2-72
 * (Box-Cox inputs, z-scored inputs and labels, GEMM + GELU layers).  Inputs
 * are (T, log P, BoxCox(Y_1..Y_9)), outputs dY_1..dY_9 over the step dt; the
 * network is trained here, on the GPU, by f3d_nn_train_step.  Contexts with
 * Ns = 9 (H2/O2: 11 inputs, 9 outputs) or Ns = 5 (air: 7 inputs, 5 outputs);
 * the configuration arrays are sized for the larger and use their first
 * Ns + 2 / Ns entries (mahal_inv keeps its row stride of 11). */
#define F3D_NN_NPARAM 12512 /* packed FP32 parameters, layout in csrc/f3d_nn.h */
typedef struct {
  float in_mean[11], in_std[11], out_mean[9], out_std[9];
  float mahal_mean[11];
  float mahal_inv[121];  /* inverse covariance of the normalized inputs, row-major */
  float d_max;           /* cells with Mahalanobis distance > d_max fall back */
  float elem_tol;        /* max_e |net change of element e| / gross elemental turnover
                            sum_e sum_k a_ek |dY_k| / W_k > elem_tol falls back */
  float neg_tol;         /* Y_k + dY_k < -neg_tol falls back */
  float res_tol;         /* backward-Euler residual max_k |R_k|/rho > res_tol falls back
                            (R_k = rho_k' - rho_k - dt W_k wdot_k(rho', T')); <= 0: off */
  float res_T_tol;       /* ... or sum_k |h_f,k R_k| / (rho c_v) > res_T_tol [K], the
                            temperature error those residuals imply; <= 0: off */
  float dT_max;          /* predicted chemical temperature change sum_k |h_f,k dY_k| / c_v
                            above dT_max [K] falls back (fast chemistry); <= 0: off */
  float chi_max;         /* thermal explosiveness at the predicted state,
                            chi = dt (d HRR / dT) / (rho c_v), HRR = -sum_k h_f,k W_k wdot_k,
                            above chi_max falls back (where the backward-Euler step can
                            jump between roots); <= 0: off */
  float chain_max;       /* chain explosiveness at the predicted state: cells where the
                            3x3 block of dt d wdot / d c over the chain carriers
                            (chain_species) has an eigenvalue with real part >=
                            chain_max fall back (near dt lambda = 1 the backward-Euler
                            step of a radical-free hot mixture jumps to the ignited
                            root while a near-root with a tiny residual exists);
                            <= 0: off */
  int32_t chain_species; /* bitmask of the (at most 3) chain-carrier species, bit k =
                            species k (H2/O2: H, O, OH) */
  float boxcox_lambda;
  int32_t label_mode;    /* outputs: 0 dY_k; 1 d BoxCox(Y_k) (post_predict form,
                            Y' = (lambda (BoxCox(Y) + out) + 1)^(1/lambda));
                            2 asinh(dY_k / label_scale_k);
                            3 asinh(dY_k / s_k) with the per-cell scale
                            s_k = min(label_scale_k, max(label_floor, label_rel Y_k)):
                            relative to Y_k for trace species, so a species at or
                            near zero stays there (the element projection is then
                            weighted per cell by (s_k + |dY_k|)^2) */
  float label_scale[9];  /* label_mode 2: dY_k = label_scale_k sinh(out_k); mode 3: the cap
                            of s_k */
  float label_rel;       /* label_mode 3 (see above) */
  float label_floor;     /* label_mode 3 (see above) */
  float y_floor;         /* inputs use BoxCox(max(Y_k, y_floor)); 0: BoxCox(Y_k) */
  int32_t n_elem;
  float elem[4][9];      /* atoms of element e in species k */
  double dt;             /* chemistry step the network predicts */
} f3d_nn_config;
/* Allocates the network (parameters zero) and stores the configuration;
 * the context must have Ns = 9 or Ns = 5 (inputs Ns + 2, outputs Ns; the
 * arrays above hold 11 and 9 entries, of which the first Ns + 2 and Ns are
 * used).  May be called again to change the config.
 * With n_elem > 0 every prediction is projected onto the element-conserving
 * subspace after the element-balance gate (weighted by out_std^2, or per cell
 * in label_mode 3). */
int f3d_nn_init(f3d_ctx* ctx, const f3d_nn_config* cfg);
/* Host FP32 parameter vector of F3D_NN_NPARAM floats; set resets Adam. */
int f3d_nn_set_params(f3d_ctx* ctx, const float* params);
int f3d_nn_get_params(f3d_ctx* ctx, float* params);
/* One Adam step on the MSE of normalized outputs for a device batch
 * X[B][Ns+2] (normalized inputs), Y[B][Ns] (normalized targets), B % 64 == 0;
 * loss (device float, may be NULL) receives sum (y - t)^2 / (B * 9). */
int f3d_nn_train_step(f3d_ctx* ctx, const float* X, const float* Y, int32_t B, float lr,
                      float* loss, hipStream_t s);
/* f3d_nn_train_step without any device allocation (no counterpart in the
 * reference: for callers that cannot allocate, such as code under HIP graph
 * capture or a training loop that forbids hipMalloc).  The per-tile partial
 * gradients live in a static device array of the library (one per GPU, 6.4 MB,
 * held until the process ends) instead of the context's buffer that
 * f3d_nn_train_step allocates on first use or growth: B <= 8192 (B % 64 == 0),
 * else F3D_ERR_ARG.  Every call on a GPU uses that GPU's one array, so calls
 * that overlap on one GPU (two streams, two contexts) race; calls in stream
 * order do not.  Same kernels and result as f3d_nn_train_step. */
int f3d_nn_train_step_static(f3d_ctx* ctx, const float* X, const float* Y, int32_t B, float lr,
                             float* loss, hipStream_t s);
/* Gradient of that loss w.r.t. the packed parameters (device, F3D_NN_NPARAM
 * floats, followed by the loss) without updating the weights. */
int f3d_nn_grad(f3d_ctx* ctx, const float* X, const float* Y, int32_t B, float* grad,
                hipStream_t s);
/* Forward on normalized inputs: Y[B][Ns] = net(X[B][Ns+2]), B % 64 == 0. */
int f3d_nn_forward(f3d_ctx* ctx, const float* X, float* Y, int32_t B, hipStream_t s);
/* Surrogate chemistry step of dt (= cfg.dt) over the interior: cells with
 * T > T_criteria are predicted by the network; cells failing the gates are
 * integrated by the FP64 backward-Euler Newton (f3d_chem_newton settings).
 * status [Nx*Ny*Nz] (device, may be NULL): 0 untouched, 1 surrogate; Newton
 * fallback after the gate that rejected the cell: 2 Mahalanobis distance,
 * 3 element balance, 4 positivity, 5 predicted temperature change,
 * 6 backward-Euler residual, 7 thermal explosiveness, 8 chain explosiveness.
 * *n_fallback (host) receives the number of fallback cells; this call
 * synchronizes the stream to size the fallback launch. */
int f3d_nn_chem(f3d_ctx* ctx, double* U, double* Q, double* rhoi, double dt, int32_t maxiter,
                double rtol, double ttol, int8_t* status, int64_t* n_fallback, const int64_t* tag,
                hipStream_t s);

/* Timing of f3d_nn_chem's parts with HIP events recorded on the call's stream
 * (no reference counterpart: a measurement hook).  enable = 1 records events
 * in later calls, 0 stops; ms (host, may be NULL, 4 doubles) receives the
 * last timed call's [network kernel, gate epilogue kernel, tile scan +
 * compaction, read-back + Newton fallback] in milliseconds (synchronizes on
 * that call's end). */
int f3d_nn_timing(f3d_ctx* ctx, int32_t enable, double* ms);

/* ---- stage level -------------------------------------------------------- */
/* Registers the block's metrics (9 device pointers, order as above) and J;
 * the context keeps the pointers and, for F3D_FP32, float copies of the
 * arrays, which the FP32 stage and the kernel-granular split / viscous calls
 * given these same pointers read instead (the values the kernels would
 * convert to anyway).  Call again after changing the arrays' contents. */
int f3d_set_metrics(f3d_ctx* ctx, const double* const* metrics, const double* J);

/* ---- mesh (parse_mesh.jl) ----------------------------------------------- */
/* Ghost node coordinates of padded x, y, z (interior nodes given), in place:
 * point reflection through the boundary node along x and y for interior k,
 * the x-y corner columns from the two adjacent ghost rows, then the z ghost
 * planes over the whole x-y extent (generateXYZ, parse_mesh.jl:52-113).
 * Needs Nx, Ny, Nz >= 5. */
int f3d_mesh_ghosts(f3d_ctx* ctx, double* x, double* y, double* z, hipStream_t s);
/* Cofactor metrics (9 device arrays, order as f3d_set_metrics) and J at
 * every padded point from padded node coordinates (computeMetrics,
 * parse_mesh.jl:132-221): sixth-order central differences (CD6) of x, y, z
 * along each index, second-order one-sided (CD2_L / CD2_R) in the three
 * outermost layers; J = 1/det of the coordinate Jacobian; metric (a, b) =
 * d(index a)/d(coordinate b) / J. */
int f3d_mesh_metrics(f3d_ctx* ctx, const double* x, const double* y, const double* z,
                     double* const* metrics, double* J, hipStream_t s);
/* One Shu-Osher RK3 stage KRK = 1, 2, 3 exactly as solver.jl:298-326:
 * (KRK == 1: Un <- U, rhon <- rhoi), mixture, shockSensor, specAdvance,
 * flowAdvance, linComb (KRK 2, 3), c2Prim, fills.  Intermediates live in
 * context-owned device memory (allocated on first use). */
int f3d_rk3_stage(f3d_ctx* ctx, int32_t KRK, double* U, double* Un, double* rhoi, double* rhon,
                  double* Q, double dt, const int64_t* tag, hipStream_t s);
/* Pieces of the fused float stage (F3D_FP32 contexts with F3D_STAGE_FUSED,
 * after f3d_set_metrics), for callers that drive or time them one by one;
 * f3d_rk3_stage calls them in this order after f3d_mixture and
 * f3d_shock_sensor.  Together they replace the launches of flowAdvance and
 * specAdvance (solver.jl:22-56) and the divergence, linComb and c2Prim of
 * the stage body (solver.jl:304-317).  rhs [5 + Ns][Nx*Ny*Nz] float
 * (variables rho, rho u, rho v, rho w, E, rho_1.. ; cell a + Nx (b + Ny c)).
 * f3d_sweep_conv: split -> hybrid reconstruction -> face-flux difference of
 *   direction dir for all 5 + Ns variables; accumulate = 0 stores, 1 adds.
 * f3d_sweep_diff: viscous stresses, heat flux, species diffusion and its
 *   enthalpy flux of direction dir, added (Yi, lambda, mu, D from f3d_mixture).
 * f3d_sweep_stage_end: U += J dt rhs[0..4], rho_k += J dt rhs[5..], linComb
 *   with Un, rhon for KRK = 2, 3, then c2Prim (warm-started e -> T). */
int f3d_sweep_conv(f3d_ctx* ctx, int32_t dir, float* rhs, const double* Q, const double* rhoi,
                   const float* phi, int32_t accumulate, hipStream_t s);
int f3d_sweep_diff(f3d_ctx* ctx, int32_t dir, float* rhs, const double* Q, const float* Yi,
                   const float* lambda, const float* mu, const float* D, hipStream_t s);
int f3d_sweep_stage_end(f3d_ctx* ctx, int32_t KRK, double* U, const double* Un, double* rhoi,
                        const double* rhon, double* Q, const float* rhs, double dt,
                        hipStream_t s);
/* Chemistry of f3d_time_step's Strang half-steps. */
#define F3D_CHEM_NONE 0
#define F3D_CHEM_STIFF 1     /* f3d_chem_stiff, sub_step substeps (eval_gpu_stiff) */
#define F3D_CHEM_EXPLICIT 2  /* f3d_chem_explicit, sub_step substeps (eval_gpu) */
#define F3D_CHEM_NEWTON 3    /* f3d_chem_newton (maxiter, rtol, ttol) */
#define F3D_CHEM_SURROGATE 4 /* f3d_nn_chem; *n_fallback accumulates the fallbacks */
/* One time step as solver.jl:240-378: chemistry over dt/2 and f3d_fill,
 * the three f3d_rk3_stage calls, chemistry over dt/2 and f3d_fill. */
int f3d_time_step(f3d_ctx* ctx, double* U, double* Un, double* rhoi, double* rhon, double* Q,
                  double dt, int32_t chem, int32_t sub_step, int32_t maxiter, double rtol,
                  double ttol, int64_t* n_fallback, const int64_t* tag, hipStream_t s);
/* any(isnan, U) of solver.jl:388 over the 5 padded fields of U; *flag (host)
 * = 1 if a NaN is present, else 0.  Synchronizes the stream. */
int f3d_nan_check(f3d_ctx* ctx, const double* U, int32_t* flag, hipStream_t s);
/* Device pointer of a context-owned intermediate after f3d_rk3_stage, by
 * name ("Yi", "lambda", "mu", "D", "phi", "Fx", "Fy", "Fz", "Fvx", ...,
 * "Fhz"); NULL if unknown. */
void* f3d_workspace(f3d_ctx* ctx, const char* name);

#ifdef __cplusplus
}
#endif
#endif /* FLAME3D_H */
