// Viscous stresses, heat flux and species diffusion at faces (viscous.jl).
// One thread per face; 4th-order face-normal derivative
// 1.25(q_i - q_{i-1}) - (q_{i+1} - q_{i-2})/12 and averaged 4th-order central
// cross derivatives (2nd order for species, viscous.jl:476-479).
#include "f3d_internal.h"

namespace f3d {

template <typename R>
__device__ inline R d_normal(const double* __restrict__ q, long ic, long s) {
  const R c12 = R(1.0 / 12.0);
  return R(1.25) * ((R)q[ic] - (R)q[ic - s]) - c12 * ((R)q[ic + s] - (R)q[ic - 2 * s]);
}
template <typename R>
__device__ inline R d_cross4(const double* __restrict__ q, long ic, long s, long t) {
  const R c23 = R(2.0 / 3.0), c12 = R(1.0 / 12.0);
  return R(0.5) * (c23 * ((R)q[ic + t] + (R)q[ic - s + t] - (R)q[ic - t] - (R)q[ic - s - t]) -
                   c12 * ((R)q[ic + 2 * t] + (R)q[ic - s + 2 * t] - (R)q[ic - 2 * t] -
                          (R)q[ic - s - 2 * t]));
}
template <typename R>
__device__ inline R y_normal(const R* __restrict__ q, long ic, long s) {
  const R c12 = R(1.0 / 12.0);
  return R(1.25) * (q[ic] - q[ic - s]) - c12 * (q[ic + s] - q[ic - 2 * s]);
}
template <typename R>
__device__ inline R y_cross2(const R* __restrict__ q, long ic, long s, long t) {
  return R(0.25) * (q[ic + t] + q[ic - s + t] - q[ic - t] - q[ic - s - t]);
}

struct Metrics {
  const double* m[9];
};

// viscousFlux_x/y/z, viscous.jl:1-428 (LES branches off, run.jl:7-8)
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_viscous(R* __restrict__ Fv, const double* __restrict__ Q, Metrics M,
              const double* __restrict__ J, const R* __restrict__ lam, const R* __restrict__ mu,
              const R* __restrict__ Fh, const int64_t* __restrict__ tag, Geom g, int dir) {
  const int nx = g.Nx + (dir == 0), ny = g.Ny + (dir == 1), nz = g.Nz + (dir == 2);
  const long nf = (long)nx * ny * nz;
  const long f = gid();
  if (f >= nf) return;
  const int a = (int)(f % nx);
  const int tt = (int)(f / nx);
  const int b = tt % ny, c = tt / ny;
  const long ic = (a + NG) + (long)g.Nxt * ((b + NG) + (long)g.Nyt * (c + NG));
  if (tag && tag[ic] == 1) return;
  const long st[3] = {1, (long)g.Nxt, (long)g.Nxt * g.Nyt};
  const long s = st[dir];
  const long im = ic - s;
  const long NP = g.NP;
  R m[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) m[q] = ((R)M.m[q][im] + (R)M.m[q][ic]) * R(0.5);
  const R Jac = ((R)J[im] + (R)J[ic]) * R(0.5);
  const R mui = (mu[im] + mu[ic]) * R(0.5);
  const R lami = (lam[im] + lam[ic]) * R(0.5);
  R dq[4][3];
  const int qv[4] = {1, 2, 3, 5};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double* fq = Q + qv[q] * NP;
#pragma unroll
    for (int t = 0; t < 3; ++t)
      dq[q][t] = (t == dir) ? d_normal<R>(fq, ic, s) : d_cross4<R>(fq, ic, s, st[t]);
  }
  const R u = ((R)Q[im + NP] + (R)Q[ic + NP]) * R(0.5);
  const R v = ((R)Q[im + 2 * NP] + (R)Q[ic + 2 * NP]) * R(0.5);
  const R w = ((R)Q[im + 3 * NP] + (R)Q[ic + 3 * NP]) * R(0.5);
  R gr[4][3];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      gr[q][j] = (dq[q][0] * m[j] + dq[q][1] * m[3 + j] + dq[q][2] * m[6 + j]) * Jac;
  const R c23 = R(2.0 / 3.0);
  const R divu = gr[0][0] + gr[1][1] + gr[2][2];
  const R t11 = mui * (R(2) * gr[0][0] - c23 * divu);
  const R t12 = mui * (gr[0][1] + gr[1][0]);
  const R t13 = mui * (gr[0][2] + gr[2][0]);
  const R t22 = mui * (R(2) * gr[1][1] - c23 * divu);
  const R t23 = mui * (gr[2][1] + gr[1][2]);
  const R t33 = mui * (R(2) * gr[2][2] - c23 * divu);
  const R E1 = u * t11 + v * t12 + w * t13 + lami * gr[3][0] + Fh[f];
  const R E2 = u * t12 + v * t22 + w * t23 + lami * gr[3][1] + Fh[f + nf];
  const R E3 = u * t13 + v * t23 + w * t33 + lami * gr[3][2] + Fh[f + 2 * nf];
  const R n0 = m[3 * dir], n1 = m[3 * dir + 1], n2 = m[3 * dir + 2];
  Fv[f] = n0 * t11 + n1 * t12 + n2 * t13;
  Fv[f + nf] = n0 * t12 + n1 * t22 + n2 * t23;
  Fv[f + 2 * nf] = n0 * t13 + n1 * t23 + n2 * t33;
  Fv[f + 3 * nf] = n0 * E1 + n1 * E2 + n2 * E3;
}

// specViscousFlux_x/y/z, viscous.jl:430-665: rho D_k = (D_{i-1} + D_i) rho_face
// (no 1/2, viscous.jl:473), correction velocity, enthalpy flux Fh.
template <int NS, typename R>
__global__ void __launch_bounds__(kBlock)
    k_spec_viscous(R* __restrict__ Fd, const double* __restrict__ Q, const R* __restrict__ Yi,
                   Metrics M, const double* __restrict__ J, const R* __restrict__ D,
                   R* __restrict__ Fh, const int64_t* __restrict__ tag,
                   const DevParams* __restrict__ P, Geom g, int dir) {
  const int nx = g.Nx + (dir == 0), ny = g.Ny + (dir == 1), nz = g.Nz + (dir == 2);
  const long nf = (long)nx * ny * nz;
  const long f = gid();
  if (f >= nf) return;
  const int a = (int)(f % nx);
  const int tt = (int)(f / nx);
  const int b = tt % ny, c = tt / ny;
  const long ic = (a + NG) + (long)g.Nxt * ((b + NG) + (long)g.Nyt * (c + NG));
  if (tag && tag[ic] == 1) return;
  const long st[3] = {1, (long)g.Nxt, (long)g.Nxt * g.Nyt};
  const long s = st[dir];
  const long im = ic - s;
  const long NP = g.NP;
  R m[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) m[q] = ((R)M.m[q][im] + (R)M.m[q][ic]) * R(0.5);
  const R Jac = ((R)J[im] + (R)J[ic]) * R(0.5);
  const R rho = ((R)Q[im] + (R)Q[ic]) * R(0.5);
  const R T = ((R)Q[im + 5 * NP] + (R)Q[ic + 5 * NP]) * R(0.5);
  R hi[NS], Vk1[NS], Vk2[NS], Vk3[NS];
  h_specs<NS, R>(P, T, hi);
  R sum1 = 0, sum2 = 0, sum3 = 0;
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    const R* y = Yi + (long)n * NP;
    const R rhoDi = (D[im + (long)n * NP] + D[ic + (long)n * NP]) * rho;
    R dY[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) dY[t] = (t == dir) ? y_normal<R>(y, ic, s) : y_cross2<R>(y, ic, s, st[t]);
    const R Vx = (dY[0] * m[0] + dY[1] * m[3] + dY[2] * m[6]) * Jac * rhoDi;
    const R Vy = (dY[0] * m[1] + dY[1] * m[4] + dY[2] * m[7]) * Jac * rhoDi;
    const R Vz = (dY[0] * m[2] + dY[1] * m[5] + dY[2] * m[8]) * Jac * rhoDi;
    Vk1[n] = Vx;
    Vk2[n] = Vy;
    Vk3[n] = Vz;
    sum1 += Vx;
    sum2 += Vy;
    sum3 += Vz;
  }
  const R n0 = m[3 * dir], n1 = m[3 * dir + 1], n2 = m[3 * dir + 2];
  R fh1 = 0, fh2 = 0, fh3 = 0;
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    const R Yn = R(0.5) * (Yi[im + (long)n * NP] + Yi[ic + (long)n * NP]);
    const R V1 = Vk1[n] - sum1 * Yn;
    const R V2 = Vk2[n] - sum2 * Yn;
    const R V3 = Vk3[n] - sum3 * Yn;
    Fd[f + n * nf] = V1 * n0 + V2 * n1 + V3 * n2;
    fh1 += V1 * hi[n];
    fh2 += V2 * hi[n];
    fh3 += V3 * hi[n];
  }
  Fh[f] = fh1;
  Fh[f + nf] = fh2;
  Fh[f + 2 * nf] = fh3;
}

static Metrics mk(const double* const* M) {
  Metrics r;
  for (int q = 0; q < 9; ++q) r.m[q] = M[q];
  return r;
}

int launch_viscous(f3d_ctx* c, int dir, void* Fv, const double* Q, const double* const* M,
                   const double* J, const void* lam, const void* mu, const void* Fh,
                   const int64_t* tag, hipStream_t s) {
  const Geom& g = c->g;
  const long nf = (long)(g.Nx + (dir == 0)) * (g.Ny + (dir == 1)) * (g.Nz + (dir == 2));
  if (c->precision == F3D_FP64)
    hipLaunchKernelGGL(k_viscous<double>, dim3(nblocks(nf)), dim3(kBlock), 0, s, (double*)Fv, Q,
                       mk(M), J, (const double*)lam, (const double*)mu, (const double*)Fh, tag,
                       g, dir);
  else
    hipLaunchKernelGGL(k_viscous<float>, dim3(nblocks(nf)), dim3(kBlock), 0, s, (float*)Fv, Q,
                       mk(M), J, (const float*)lam, (const float*)mu, (const float*)Fh, tag, g,
                       dir);
  return hip_check(hipGetLastError(), "viscous_flux");
}

template <int NS>
static void spec_visc_ns(f3d_ctx* c, int dir, void* Fd, const double* Q, const void* Yi,
                         const double* const* M, const double* J, const void* D, void* Fh,
                         const int64_t* tag, hipStream_t s) {
  const Geom& g = c->g;
  const long nf = (long)(g.Nx + (dir == 0)) * (g.Ny + (dir == 1)) * (g.Nz + (dir == 2));
  if (c->precision == F3D_FP64)
    hipLaunchKernelGGL((k_spec_viscous<NS, double>), dim3(nblocks(nf)), dim3(kBlock), 0, s,
                       (double*)Fd, Q, (const double*)Yi, mk(M), J, (const double*)D,
                       (double*)Fh, tag, c->dparams, g, dir);
  else
    hipLaunchKernelGGL((k_spec_viscous<NS, float>), dim3(nblocks(nf)), dim3(kBlock), 0, s,
                       (float*)Fd, Q, (const float*)Yi, mk(M), J, (const float*)D, (float*)Fh,
                       tag, c->dparams, g, dir);
}

int launch_spec_viscous(f3d_ctx* c, int dir, void* Fd, const double* Q, const void* Yi,
                        const double* const* M, const double* J, const void* D, void* Fh,
                        const int64_t* tag, hipStream_t s) {
  switch (c->g.Ns) {
    case 1: spec_visc_ns<1>(c, dir, Fd, Q, Yi, M, J, D, Fh, tag, s); break;
    case 5: spec_visc_ns<5>(c, dir, Fd, Q, Yi, M, J, D, Fh, tag, s); break;
    case 9: spec_visc_ns<9>(c, dir, Fd, Q, Yi, M, J, D, Fh, tag, s); break;
    default: set_error("spec_viscous_flux: Ns must be 1, 5 or 9"); return F3D_ERR_UNSUPPORTED;
  }
  return hip_check(hipGetLastError(), "spec_viscous_flux");
}

}  // namespace f3d
