// Convective path: Steger-Warming splitting (split.jl), pressure shock sensor
// and hybrid UP7 / WENO5-Z / NND face reconstruction (schemes.jl).
//
// One thread per padded point (splitting, sensor) or per face
// (reconstruction), 1-D grids in i-fastest order so every wavefront reads 64
// consecutive i of one variable plane (coalesced), with XCD-aware block
// ordering so y/z stencil neighbours are reused from the same L2.
// Template parameter R is the arithmetic/storage type of the intermediates
// (double for the FP64 path, float for the FP32 path); the state Q and rho_k
// are read as double, the metrics as MT (double, or the FP32 stage's float
// copies).
#include "f3d_stencil.h"

namespace f3d {

// fluxSplit, split.jl:17-89
template <typename R, typename MT>
__global__ void __launch_bounds__(kBlock)
    k_flux_split(const double* __restrict__ Q, R* __restrict__ Fp, R* __restrict__ Fm,
                 const MT* __restrict__ Ax, const MT* __restrict__ Ay,
                 const MT* __restrict__ Az, const int64_t* __restrict__ tag,
                 const DevParams* __restrict__ P, int NP) {
  const long p = gid();
  if (p >= NP) return;
  if (tag && tag[p] == 1) return;
  R fp[5], fm[5];
  split_point<R>((R)Q[p], (R)Q[p + NP], (R)Q[p + 2L * NP], (R)Q[p + 3L * NP], (R)Q[p + 4L * NP],
                 (R)Q[p + 6L * NP], (R)Ax[p], (R)Ay[p], (R)Az[p], PT<R>::sweps2(P), fp, fm);
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    Fp[p + n * (long)NP] = fp[n];
    Fm[p + n * (long)NP] = fm[n];
  }
}

// split, split.jl:101-130
template <typename R, typename MT>
__global__ void __launch_bounds__(kBlock)
    k_species_split(const double* __restrict__ rhoi, const double* __restrict__ Q,
                    R* __restrict__ Fp, R* __restrict__ Fm, const MT* __restrict__ Ax,
                    const MT* __restrict__ Ay, const MT* __restrict__ Az,
                    const int64_t* __restrict__ tag, int NP, int Ns) {
  const long p = gid();
  if (p >= NP) return;
  if (tag && tag[p] == 1) return;
  const R u = (R)Q[p + NP], v = (R)Q[p + 2L * NP], w = (R)Q[p + 3L * NP];
  const R un = (R)Ax[p] * u + (R)Ay[p] * v + (R)Az[p] * w;
  const R Ep = R(0.5) * (un + fabs(un));
  const R Em = R(0.5) * (un - fabs(un));
  for (int n = 0; n < Ns; ++n) {
    const R r = (R)rhoi[p + (long)n * NP];
    Fp[p + (long)n * NP] = Ep * r;
    Fm[p + (long)n * NP] = Em * r;
  }
}

// shockSensor, schemes.jl:1-24 (0-based range [1, N+2NG-2] per direction).
// Evaluated in FP64 on both paths and stored in R: the second difference of p
// cancels (p ~ 1e5-1e6 Pa against differences of ~1e2), which in float
// arithmetic would move phi by ~1e-4 relative and flip the UP7/WENO/NND branch
// at faces near a threshold.
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_shock_sensor(R* __restrict__ phi, const double* __restrict__ Q, Geom g) {
  const long p = gid();
  if (p >= g.NP) return;
  const int i = (int)(p % g.Nxt);
  const int t = (int)(p / g.Nxt);
  const int j = t % g.Nyt, k = t / g.Nyt;
  if (i < 1 || i > g.Nxt - 2 || j < 1 || j > g.Nyt - 2 || k < 1 || k > g.Nzt - 2) return;
  const double* Pp = Q + 4L * g.NP;
  const long sy = g.Nxt, sz = (long)g.Nxt * g.Nyt;
  const double Px1 = Pp[p - 1], Px2 = Pp[p], Px3 = Pp[p + 1];
  const double Py1 = Pp[p - sy], Py3 = Pp[p + sy];
  const double Pz1 = Pp[p - sz], Pz3 = Pp[p + sz];
  const double px = fabs(-Px1 + 2.0 * Px2 - Px3) / (Px1 + 2.0 * Px2 + Px3);
  const double py = fabs(-Py1 + 2.0 * Px2 - Py3) / (Py1 + 2.0 * Px2 + Py3);
  const double pz = fabs(-Pz1 + 2.0 * Px2 - Pz3) / (Pz1 + 2.0 * Px2 + Pz3);
  phi[p] = (R)(px + py + pz);
}

// Ducros sensor (no reference counterpart; the reference switches on the
// pressure sensor above): theta = (div u)^2 / ((div u)^2 + |curl u|^2 + 1e-30),
// velocity gradients from 2nd-order central differences along xi, eta, zeta
// and the metrics (d xi_a / d x_j = M[3a + j] J, the J^-1-weighted cofactors),
// in FP64 on both paths, on the points shockSensor covers.
struct Met9 {
  const double* m[9];
};
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_ducros(R* __restrict__ phi, const double* __restrict__ Q, Met9 M,
             const double* __restrict__ J, Geom g) {
  const long p = gid();
  if (p >= g.NP) return;
  const int i = (int)(p % g.Nxt);
  const int t = (int)(p / g.Nxt);
  const int j = t % g.Nyt, k = t / g.Nyt;
  if (i < 1 || i > g.Nxt - 2 || j < 1 || j > g.Nyt - 2 || k < 1 || k > g.Nzt - 2) return;
  const long st[3] = {1, (long)g.Nxt, (long)g.Nxt * g.Nyt};
  const double Jp = J[p];
  double G[3][3];  // G[c][x] = d u_c / d x_x
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double* u = Q + (long)(1 + c) * g.NP;
    double dxi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) dxi[a] = 0.5 * (u[p + st[a]] - u[p - st[a]]);
#pragma unroll
    for (int x = 0; x < 3; ++x)
      G[c][x] = Jp * (M.m[x][p] * dxi[0] + M.m[3 + x][p] * dxi[1] + M.m[6 + x][p] * dxi[2]);
  }
  const double dv = G[0][0] + G[1][1] + G[2][2];
  const double wx = G[2][1] - G[1][2], wy = G[0][2] - G[2][0], wz = G[1][0] - G[0][1];
  phi[p] = (R)(dv * dv / (dv * dv + wx * wx + wy * wy + wz * wz + 1e-30));
}

// pointer form: gathers the stencil and evaluates weno_face_v
template <typename R>
__device__ inline R weno_face(const R* __restrict__ fp, const R* __restrict__ fm, long ic, long s,
                              R phif, const DevParams* __restrict__ P) {
  const int br = weno_branch<R>(phif, P);
  R p[8], m[8];
  if (br == 0 || (br == 1 && P->recon_kind == 1)) {
#pragma unroll
    for (int k = 0; k < 7; ++k) p[k] = fp[ic + (k - 4) * s];
#pragma unroll
    for (int k = 1; k < 8; ++k) m[k] = fm[ic + (k - 4) * s];
  } else if (br == 1) {
#pragma unroll
    for (int k = 1; k < 6; ++k) p[k] = fp[ic + (k - 4) * s];
#pragma unroll
    for (int k = 2; k < 7; ++k) m[k] = fm[ic + (k - 4) * s];
  } else {
#pragma unroll
    for (int k = 2; k < 5; ++k) p[k] = fp[ic + (k - 4) * s];
#pragma unroll
    for (int k = 3; k < 6; ++k) m[k] = fm[ic + (k - 4) * s];
  }
  return weno_face_v<R>(p, m, br, P);
}

template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_weno(R* __restrict__ F, const R* __restrict__ phi, const R* __restrict__ Fp,
           const R* __restrict__ Fm, int NV, const int64_t* __restrict__ tag,
           const DevParams* __restrict__ P, Geom g, int dir) {
  const int nx = g.Nx + (dir == 0), ny = g.Ny + (dir == 1), nz = g.Nz + (dir == 2);
  const long nf = (long)nx * ny * nz;
  const long f = gid();
  if (f >= nf) return;
  const int a = (int)(f % nx);
  const int t = (int)(f / nx);
  const int b = t % ny, c = t / ny;
  const long ic = (a + NG) + (long)g.Nxt * ((b + NG) + (long)g.Nyt * (c + NG));
  if (tag && tag[ic] == 1) return;
  const long s = (dir == 0) ? 1 : (dir == 1) ? (long)g.Nxt : (long)g.Nxt * g.Nyt;
  const R phif = fmax(phi[ic - s], phi[ic]);
  for (int n = 0; n < NV; ++n)
    F[f + n * nf] = weno_face<R>(Fp + (long)n * g.NP, Fm + (long)n * g.NP, ic, s, phif, P);
}

// y and z sweeps: one thread per (x, transverse) line marches through a
// segment of faces along the sweep for one variable, holding the eight-point
// stencil of Fp and Fm in registers, so each point is read once per segment
// (plus a 7-point halo) with x-coalesced loads; the face branches are found
// from phi and kept as 2-bit codes (3 = tagged face, skipped).  One wave per
// (64 x, transverse line, segment, variable): the variables of a segment are
// consecutive blocks after the XCD remap, so they share the phi rows in L2,
// and the launch has NV times the waves of a wave-per-segment march (a march
// is latency-bound: its waves' chunks of loads are the parallelism).
constexpr int kMarchSeg = 32;
template <typename R>
__global__ void __launch_bounds__(64)
    k_weno_march(R* __restrict__ F, const R* __restrict__ phi, const R* __restrict__ Fp,
                 const R* __restrict__ Fm, int NV, const int64_t* __restrict__ tag,
                 const DevParams* __restrict__ P, Geom g, int dir, int seg) {
  const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
  const unsigned L = xcd_remap(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), nb);
  const int var = (int)(L % (unsigned)NV);
  const unsigned rest = L / (unsigned)NV;
  const int a = (int)(rest % gridDim.x) * 64 + threadIdx.x;
  if (a >= g.Nx) return;
  const int tr = (int)((rest / gridDim.x) % gridDim.y);  // z index (y sweep) or y index (z sweep)
  const int sgi = (int)(rest / (gridDim.x * gridDim.y));
  const int nd = (dir == 1 ? g.Ny : g.Nz) + 1;  // faces along the sweep
  const int j0 = sgi * seg;
  const int j1 = min(j0 + seg, nd);
  if (j0 >= j1) return;
  const long s = (dir == 1) ? (long)g.Nxt : (long)g.Nxt * g.Nyt;
  const long base = (dir == 1) ? (a + NG) + (long)g.Nxt * g.Nyt * (tr + NG)
                               : (a + NG) + (long)g.Nxt * (tr + NG);
  const long nf = (dir == 1) ? (long)g.Nx * nd * g.Nz : (long)g.Nx * g.Ny * nd;
  // face j -> F index: a + Nx (j + nd tr) (y sweep) or a + Nx (tr + Ny j) (z sweep)
  const long fbase = (dir == 1) ? a + (long)g.Nx * nd * tr : a + (long)g.Nx * tr;
  const long fstr = (dir == 1) ? (long)g.Nx : (long)g.Nx * g.Ny;
  unsigned long long c0 = 0, c1 = 0;  // bit j - j0 of the 2-bit branch code
  {
    R ph_prev = phi[base + (long)(j0 + NG - 1) * s];
#pragma unroll 8
    for (int j = j0; j < j1; ++j) {
      const long ic = base + (long)(j + NG) * s;
      const R ph = phi[ic];
      int br = weno_branch<R>(fmax(ph_prev, ph), P);
      if (tag && tag[ic] == 1) br = 3;
      c0 |= (unsigned long long)(br & 1) << (j - j0);
      c1 |= (unsigned long long)(br >> 1) << (j - j0);
      ph_prev = ph;
    }
  }
  // faces in chunks of kCh: the kCh new stencil rows of a chunk are loaded
  // together (independent loads in flight), rows past the array clamped
  constexpr int kCh = 8;
  const long rmax = base + (long)((dir == 1 ? g.Nyt : g.Nzt) - 1) * s;
  {
    const int n = var;
    const R* fp = Fp + (long)n * g.NP;
    const R* fm = Fm + (long)n * g.NP;
    R pw[kCh + 7], mw[kCh + 7];  // rows j-4 .. j+kCh+2 of the chunk starting at face j
    const long ic0 = base + (long)(j0 + NG) * s;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      pw[kCh + k] = fp[ic0 + (k - 4) * s];
      mw[kCh + k] = fm[ic0 + (k - 4) * s];
    }
    for (int j = j0; j < j1; j += kCh) {
      const long ic = base + (long)(j + NG) * s;
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        pw[k] = pw[kCh + k];
        mw[k] = mw[kCh + k];
      }
#pragma unroll
      for (int k = 0; k < kCh; ++k) {
        const long q = min(ic + (long)(k + 3) * s, rmax);
        pw[7 + k] = fp[q];
        mw[7 + k] = fm[q];
      }
#pragma unroll
      for (int k = 0; k < kCh; ++k) {
        const int jj = j + k;
        if (jj < j1) {
          const int br = (int)((c0 >> (jj - j0)) & 1ull) | ((int)((c1 >> (jj - j0)) & 1ull) << 1);
          if (br != 3)
            F[fbase + (long)jj * fstr + (long)n * nf] = weno_face_v<R>(&pw[k], &mw[k], br, P);
        }
      }
    }
  }
}

// ---- launchers --------------------------------------------------------------
int launch_flux_split(f3d_ctx* c, const double* Q, void* Fp, void* Fm, const void* Ax,
                      const void* Ay, const void* Az, const int64_t* tag, hipStream_t s,
                      bool mf) {
  const int NP = c->g.NP;
  const dim3 gr(nblocks(NP)), bl(kBlock);
  if (c->precision == F3D_FP64)
    hipLaunchKernelGGL((k_flux_split<double, double>), gr, bl, 0, s, Q, (double*)Fp, (double*)Fm,
                       (const double*)Ax, (const double*)Ay, (const double*)Az, tag, c->dparams,
                       NP);
  else if (mf)
    hipLaunchKernelGGL((k_flux_split<float, float>), gr, bl, 0, s, Q, (float*)Fp, (float*)Fm,
                       (const float*)Ax, (const float*)Ay, (const float*)Az, tag, c->dparams, NP);
  else
    hipLaunchKernelGGL((k_flux_split<float, double>), gr, bl, 0, s, Q, (float*)Fp, (float*)Fm,
                       (const double*)Ax, (const double*)Ay, (const double*)Az, tag, c->dparams,
                       NP);
  return hip_check(hipGetLastError(), "flux_split");
}

int launch_species_split(f3d_ctx* c, const double* rhoi, const double* Q, void* Fp, void* Fm,
                         const void* Ax, const void* Ay, const void* Az,
                         const int64_t* tag, hipStream_t s, bool mf) {
  const int NP = c->g.NP;
  const dim3 gr(nblocks(NP)), bl(kBlock);
  if (c->precision == F3D_FP64)
    hipLaunchKernelGGL((k_species_split<double, double>), gr, bl, 0, s, rhoi, Q, (double*)Fp,
                       (double*)Fm, (const double*)Ax, (const double*)Ay, (const double*)Az, tag,
                       NP, c->g.Ns);
  else if (mf)
    hipLaunchKernelGGL((k_species_split<float, float>), gr, bl, 0, s, rhoi, Q, (float*)Fp,
                       (float*)Fm, (const float*)Ax, (const float*)Ay, (const float*)Az, tag, NP,
                       c->g.Ns);
  else
    hipLaunchKernelGGL((k_species_split<float, double>), gr, bl, 0, s, rhoi, Q, (float*)Fp,
                       (float*)Fm, (const double*)Ax, (const double*)Ay, (const double*)Az, tag,
                       NP, c->g.Ns);
  return hip_check(hipGetLastError(), "species_split");
}

__global__ void __launch_bounds__(kBlock)
    k_to_float(float* __restrict__ dst, const double* __restrict__ src, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (float)src[i];
}

int launch_to_float(f3d_ctx* c, float* dst, const double* src, long n, hipStream_t s) {
  (void)c;
  hipLaunchKernelGGL(k_to_float, dim3(nblocks(n)), dim3(kBlock), 0, s, dst, src, n);
  return hip_check(hipGetLastError(), "to_float");
}

int launch_shock_sensor(f3d_ctx* c, void* phi, const double* Q, hipStream_t s) {
  const int NP = c->g.NP;
  if (c->sensor_kind == F3D_SENSOR_DUCROS) {
    if (!c->J) {
      set_error("shock_sensor: the Ducros sensor needs the metrics (f3d_set_metrics first)");
      return F3D_ERR_ARG;
    }
    Met9 M;
    for (int q = 0; q < 9; ++q) M.m[q] = c->metrics[q];
    if (c->precision == F3D_FP64)
      hipLaunchKernelGGL(k_ducros<double>, dim3(nblocks(NP)), dim3(kBlock), 0, s, (double*)phi, Q,
                         M, c->J, c->g);
    else
      hipLaunchKernelGGL(k_ducros<float>, dim3(nblocks(NP)), dim3(kBlock), 0, s, (float*)phi, Q,
                         M, c->J, c->g);
    return hip_check(hipGetLastError(), "shock_sensor");
  }
  if (c->precision == F3D_FP64)
    hipLaunchKernelGGL(k_shock_sensor<double>, dim3(nblocks(NP)), dim3(kBlock), 0, s,
                       (double*)phi, Q, c->g);
  else
    hipLaunchKernelGGL(k_shock_sensor<float>, dim3(nblocks(NP)), dim3(kBlock), 0, s,
                       (float*)phi, Q, c->g);
  return hip_check(hipGetLastError(), "shock_sensor");
}

int launch_weno(f3d_ctx* c, int dir, void* F, const void* phi, const void* Fp, const void* Fm,
                int NV, const int64_t* tag, hipStream_t s) {
  const Geom& g = c->g;
  const long nf = (long)(g.Nx + (dir == 0)) * (g.Ny + (dir == 1)) * (g.Nz + (dir == 2));
  if (dir != 0) {
    const int nd = (dir == 1 ? g.Ny : g.Nz) + 1;
    const int nseg = (nd + kMarchSeg - 1) / kMarchSeg;
    const int seg = (nd + nseg - 1) / nseg;  // <= kMarchSeg: the branch codes fit 64 bits
    const dim3 grid((g.Nx + 63) / 64, dir == 1 ? g.Nz : g.Ny, nseg * NV);
    if (c->precision == F3D_FP64)
      hipLaunchKernelGGL(k_weno_march<double>, grid, dim3(64), 0, s, (double*)F,
                         (const double*)phi, (const double*)Fp, (const double*)Fm, NV, tag,
                         c->dparams, g, dir, seg);
    else
      hipLaunchKernelGGL(k_weno_march<float>, grid, dim3(64), 0, s, (float*)F,
                         (const float*)phi, (const float*)Fp, (const float*)Fm, NV, tag,
                         c->dparams, g, dir, seg);
  } else if (c->precision == F3D_FP64) {
    hipLaunchKernelGGL(k_weno<double>, dim3(nblocks(nf)), dim3(kBlock), 0, s, (double*)F,
                       (const double*)phi, (const double*)Fp, (const double*)Fm, NV, tag,
                       c->dparams, g, dir);
  } else {
    hipLaunchKernelGGL(k_weno<float>, dim3(nblocks(nf)), dim3(kBlock), 0, s, (float*)F,
                       (const float*)phi, (const float*)Fp, (const float*)Fm, NV, tag,
                       c->dparams, g, dir);
  }
  return hip_check(hipGetLastError(), "weno");
}

}  // namespace f3d
