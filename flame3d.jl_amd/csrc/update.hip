// Pointwise and update kernels: divergence (div.jl), mixture transport
// (thermo.jl:284-403), c2Prim / prim2c / linComb / getY (utils.jl) and the
// on-device boundary fills that replace boundary.jl's case fills and the MPI
// halo exchange on a single GPU.
#include "f3d_internal.h"
#include "f3d_stencil.h"

namespace f3d {

__device__ inline void interior_idx(long id, const Geom& g, int& a, int& b, int& c) {
  a = (int)(id % g.Nx);
  const int t = (int)(id / g.Nx);
  b = t % g.Ny;
  c = t / g.Ny;
}

// div, div.jl:16-54
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_div(double* __restrict__ U, const R* __restrict__ Fx, const R* __restrict__ Fy,
          const R* __restrict__ Fz, const R* __restrict__ Fvx, const R* __restrict__ Fvy,
          const R* __restrict__ Fvz, double dt, const double* __restrict__ J,
          const int64_t* __restrict__ tag, Geom g) {
  const long id = gid();
  if (id >= (long)g.Nx * g.Ny * g.Nz) return;
  int a, b, c;
  interior_idx(id, g, a, b, c);
  const long p = (a + NG) + (long)g.Nxt * ((b + NG) + (long)g.Nyt * (c + NG));
  if (tag && tag[p] != 0) return;
  const long NP = g.NP;
  const long nfx = (long)(g.Nx + 1) * g.Ny * g.Nz, nfy = (long)g.Nx * (g.Ny + 1) * g.Nz,
             nfz = (long)g.Nx * g.Ny * (g.Nz + 1);
  const long fx = a + (long)(g.Nx + 1) * (b + (long)g.Ny * c);
  const long fy = a + (long)g.Nx * (b + (long)(g.Ny + 1) * c);
  const long fz = a + (long)g.Nx * (b + (long)g.Ny * c);
  const long ox = 1, oy = g.Nx, oz = (long)g.Nx * g.Ny;
  const R Jact = (R)(J[p] * dt);
  R dV[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const R v1 = Fvx[fx + ox + q * nfx] - Fvx[fx + q * nfx];
    const R v2 = Fvy[fy + oy + q * nfy] - Fvy[fy + q * nfy];
    const R v3 = Fvz[fz + oz + q * nfz] - Fvz[fz + q * nfz];
    dV[q] = v1 + v2 + v3;
  }
  double u[5];
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    const R d = (Fx[fx + n * nfx] - Fx[fx + ox + n * nfx] + Fy[fy + n * nfy] -
                 Fy[fy + oy + n * nfy] + Fz[fz + n * nfz] - Fz[fz + oz + n * nfz]) *
                Jact;
    u[n] = U[p + n * NP] + (double)d;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) u[q + 1] += (double)(dV[q] * Jact);
#pragma unroll
  for (int n = 0; n < 5; ++n) U[p + n * NP] = u[n];
}

// divSpecs, div.jl:61-84
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_div_specs(double* __restrict__ rhoi, const R* __restrict__ Fx, const R* __restrict__ Fy,
                const R* __restrict__ Fz, const R* __restrict__ Fdx, const R* __restrict__ Fdy,
                const R* __restrict__ Fdz, double dt, const double* __restrict__ J,
                const int64_t* __restrict__ tag, Geom g) {
  const long id = gid();
  if (id >= (long)g.Nx * g.Ny * g.Nz) return;
  int a, b, c;
  interior_idx(id, g, a, b, c);
  const long p = (a + NG) + (long)g.Nxt * ((b + NG) + (long)g.Nyt * (c + NG));
  if (tag && tag[p] != 0) return;
  const long NP = g.NP;
  const long nfx = (long)(g.Nx + 1) * g.Ny * g.Nz, nfy = (long)g.Nx * (g.Ny + 1) * g.Nz,
             nfz = (long)g.Nx * g.Ny * (g.Nz + 1);
  const long fx = a + (long)(g.Nx + 1) * (b + (long)g.Ny * c);
  const long fy = a + (long)g.Nx * (b + (long)(g.Ny + 1) * c);
  const long fz = a + (long)g.Nx * (b + (long)g.Ny * c);
  const long ox = 1, oy = g.Nx, oz = (long)g.Nx * g.Ny;
  const R Jact = (R)(J[p] * dt);
  for (int n = 0; n < g.Ns; ++n) {
    const long nx = n * nfx, ny = n * nfy, nz = n * nfz;
    const R d = (Fx[fx + nx] - Fx[fx + ox + nx] + Fy[fy + ny] - Fy[fy + oy + ny] + Fz[fz + nz] -
                 Fz[fz + oz + nz] + Fdx[fx + ox + nx] - Fdx[fx + nx] + Fdy[fy + oy + ny] -
                 Fdy[fy + ny] + Fdz[fz + oz + nz] - Fdz[fz + nz]) *
                Jact;
    rhoi[p + n * NP] += (double)d;
  }
}

// mixtureProperties, thermo.jl:284-369, in precision R
template <int NS, typename R>
__device__ inline void mixture_props(const DevParams* __restrict__ P, R T, R Pr, const R* X,
                                     R* Diff, R& lambda, R& mu) {
  const R sqT = sqrt(T);
  const R sqsqT = sqrt(sqT);
  const R lgT = log(T);
  const R lgT2 = lgT * lgT;
  const R lgT3 = lgT * lgT2;
  const R lgT4 = lgT2 * lgT2;
  R mui[NS], sqi[NS];
  R s1 = 0, s2 = 0;
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    const R* p = PT<R>::cond(P, n);
    const R lk = sqT * (p[0] + lgT * p[1] + lgT2 * p[2] + lgT3 * p[3] + lgT4 * p[4]);
    s1 += X[n] * lk;
    s2 += qdiv(X[n], lk);
  }
  lambda = R(0.5) * (s1 + qdiv(R(1), s2));
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    const R* p = PT<R>::visc(P, n);
    const R sq = sqsqT * (p[0] + lgT * p[1] + lgT2 * p[2] + lgT3 * p[3] + lgT4 * p[4]);
    mui[n] = sq * sq;
    sqi[n] = fabs(sq);
  }
  // Wilke, Phi[n][l] (thermo.jl:321-334)
  R phi[NS][NS];
  if constexpr (sizeof(R) == 4) {
    // FP32: sqrt(mu_n / mu_l) = |sq_n| / |sq_l| (mu = sq^2) and the species-pair
    // factors precomputed, so no division or square root per pair
    R isq[NS];
#pragma unroll
    for (int n = 0; n < NS; ++n) isq[n] = qdiv(R(1), sqi[n]);
#pragma unroll
    for (int n = 0; n < NS; ++n) {
#pragma unroll
      for (int l = 0; l < n; ++l) {
        const R rs = sqi[n] * isq[l];  // sqrt(vr)
        const R f1 = R(1) + rs * P->wr4_f[n][l];
        const R tmp = f1 * f1 * P->c8_f[n][l];
        phi[n][l] = tmp;
        const R ir = sqi[l] * isq[n];  // 1 / sqrt(vr)
        phi[l][n] = tmp * (ir * ir) * P->iwr_f[n][l];
      }
      phi[n][n] = R(1);
    }
  } else {
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const R tmp1 = R(1) / PT<R>::mw(P, n);
      const R tmp2 = mui[n];
#pragma unroll
      for (int l = 0; l <= n; ++l) {
        const R wr = PT<R>::mw(P, l) * tmp1;
        const R vr = tmp2 / mui[l];
        const R f1 = R(1) + sqrt(vr * sqrt(wr));
        const R tmp = f1 * f1 / sqrt(R(8) + R(8) / wr);
        phi[n][l] = tmp;
        phi[l][n] = tmp / (vr * wr);
      }
      phi[n][n] = R(1);
    }
  }
  R m = 0;
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    R tmp = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) tmp += X[l] * phi[n][l];
    m += qdiv(X[n] * mui[n], tmp);
  }
  mu = m;
  // binary diffusion, D_kl P = T^1.5 poly (thermo.jl:349-367)
  const R T15 = T * sqT;
#pragma unroll
  for (int n = 0; n < NS; ++n) {
#pragma unroll
    for (int nn = n; nn < NS; ++nn) {
      const R* p = PT<R>::bdiff(P, n, nn);
      const R tmp = T15 * (p[0] + lgT * p[1] + lgT2 * p[2] + lgT3 * p[3] + lgT4 * p[4]);
      // FP32: the pair's reciprocal once (the sums below multiply by it)
      const R v = (sizeof(R) == 4) ? qdiv(R(1), tmp) : tmp;
      phi[nn][n] = v;
      phi[n][nn] = v;
    }
  }
  // D_k = (1 - X_k) / (P sum_{l!=k} X_l / D_kl + eps) with 1 - X_k taken as
  // sum_{l!=k} X_l (equal since sum X = 1): the difference 1 - X_k cancels
  // catastrophically in nearly pure cells, where the sum stays exact (a
  // binary mixture then gives D_1 = D_12 / P for any composition)
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    R s = 0, sx = 0;
#pragma unroll
    for (int nn = 0; nn < NS; ++nn) {
      if (nn == n) continue;
      s += (sizeof(R) == 4) ? X[nn] * phi[n][nn] : X[nn] / phi[n][nn];
      sx += X[nn];
    }
    s *= Pr;
    Diff[n] = qdiv(sx, s + PT<R>::eps());
  }
}

// mixture, thermo.jl:371-403 (every padded point)
template <int NS, typename R>
__global__ void __launch_bounds__(kBlock)
    k_mixture(const double* __restrict__ Q, const double* __restrict__ rhoi, R* __restrict__ Yi,
              R* __restrict__ lam, R* __restrict__ mu, R* __restrict__ D,
              const int64_t* __restrict__ tag, const DevParams* __restrict__ P, int NP) {
  const long p = gid();
  if (p >= NP) return;
  if (tag && tag[p] == 1) return;
  const R T = (R)Q[p + 5L * NP];
  const R Pr = (R)Q[p + 4L * NP];
  const R rinv = R(1) / (R)fmax(Q[p], kEps);
  R Y[NS], X[NS], Df[NS];
  R YOW = 0;
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    Y[n] = fmax((R)rhoi[p + (long)n * NP] * rinv, R(0));
    Yi[p + (long)n * NP] = Y[n];
    YOW += Y[n] / PT<R>::mw(P, n);
  }
  const R YOWINV = R(1) / YOW;
#pragma unroll
  for (int n = 0; n < NS; ++n) X[n] = Y[n] / PT<R>::mw(P, n) * YOWINV;
  R l, m;
  mixture_props<NS, R>(P, T, Pr, X, Df, l, m);
#pragma unroll
  for (int n = 0; n < NS; ++n) D[p + (long)n * NP] = Df[n];
  lam[p] = l;
  mu[p] = m;
}

// c2Prim, utils.jl:2-48 (FP64).  WARM (FP32-flux contexts): the e -> T
// Newton starts from the cell's previous temperature Q[5] (get_T_from).
template <int NS, bool WARM>
__global__ void __launch_bounds__(kBlock)
    k_c2prim(double* __restrict__ U, double* __restrict__ Q, double* __restrict__ rhoi,
             const int64_t* __restrict__ tag, const DevParams* __restrict__ P, Geom g) {
  const long id = gid();
  if (id >= (long)g.Nx * g.Ny * g.Nz) return;
  int a, b, c;
  interior_idx(id, g, a, b, c);
  const long p = (a + NG) + (long)g.Nxt * ((b + NG) + (long)g.Nyt * (c + NG));
  if (tag && tag[p] != 0) return;
  const long NP = g.NP;
  const double rho = fmax(U[p], kEps);
  const double rinv = 1 / rho;
  double r[NS];
  double sr = 0.0;
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    r[n] = fmax(rhoi[p + n * NP], 0.0);
    sr += r[n];
  }
#pragma unroll
  for (int n = 0; n < NS; ++n) r[n] *= rho / sr;
  const double u = U[p + NP] * rinv;
  const double v = U[p + 2 * NP] * rinv;
  const double w = U[p + 3 * NP] * rinv;
  const double ei = fmax(U[p + 4 * NP] - 0.5 * rho * (u * u + v * v + w * w), kEps);
  const double T = fmax(WARM ? get_T_from<NS>(P, ei, r, Q[p + 5 * NP]) : get_T<NS>(P, ei, r, nullptr),
                       kEps);
  const double pr = fmax(pmixture<NS>(P, T, r), kEps);
#pragma unroll
  for (int n = 0; n < NS; ++n) rhoi[p + n * NP] = r[n];
  Q[p] = rho;
  Q[p + NP] = u;
  Q[p + 2 * NP] = v;
  Q[p + 3 * NP] = w;
  Q[p + 4 * NP] = pr;
  Q[p + 5 * NP] = T;
  Q[p + 6 * NP] = ei;
}

// prim2c, utils.jl:71-92 (every padded point)
__global__ void __launch_bounds__(kBlock) k_prim2c(double* __restrict__ U,
                                                   const double* __restrict__ Q, int NP) {
  const long p = gid();
  if (p >= NP) return;
  const double rho = Q[p], u = Q[p + NP], v = Q[p + 2L * NP], w = Q[p + 3L * NP],
               ei = Q[p + 6L * NP];
  U[p] = rho;
  U[p + NP] = u * rho;
  U[p + 2L * NP] = v * rho;
  U[p + 3L * NP] = w * rho;
  U[p + 4L * NP] = ei + 0.5 * rho * (u * u + v * v + w * w);
}

// linComb, utils.jl:95-112 (interior)
__global__ void __launch_bounds__(kBlock)
    k_lincomb(double* __restrict__ U, const double* __restrict__ Un, int NV, double a, double b,
              const int64_t* __restrict__ tag, Geom g) {
  const long id = gid();
  if (id >= (long)g.Nx * g.Ny * g.Nz) return;
  int i, j, k;
  interior_idx(id, g, i, j, k);
  const long p = (i + NG) + (long)g.Nxt * ((j + NG) + (long)g.Nyt * (k + NG));
  if (tag && tag[p] != 0) return;
  for (int n = 0; n < NV; ++n) U[p + (long)n * g.NP] = U[p + (long)n * g.NP] * a + Un[p + (long)n * g.NP] * b;
}

// getY, utils.jl:51-68 (every padded point)
__global__ void __launch_bounds__(kBlock)
    k_get_y(double* __restrict__ Yi, const double* __restrict__ rhoi, const double* __restrict__ Q,
            const int64_t* __restrict__ tag, int NP, int Ns) {
  const long p = gid();
  if (p >= NP) return;
  if (tag && tag[p] != 0) return;
  const double rinv = 1 / fmax(Q[p], kEps);
  for (int n = 0; n < Ns; ++n) Yi[p + (long)n * NP] = fmax(rhoi[p + (long)n * NP] * rinv, 0.0);
}

// One face of the boundary fill.  Threads cover the ghost slab of that face
// (NG layers for periodic, NG+1 for copy/inflow whose boundary node is
// included as in boundary.jl:10, :68, :75), with the full padded extent in
// the two other directions so faces filled later also fill edges/corners.
struct FillState {
  double Q[7];
  double U[5];
  double r[MAXNS];
};

__global__ void __launch_bounds__(kBlock)
    k_fill(double* __restrict__ Q, double* __restrict__ U, double* __restrict__ rhoi, Geom g,
           int dir, int side, int kind, int depth, FillState st) {
  const int Nt[3] = {g.Nxt, g.Nyt, g.Nzt};
  const int N[3] = {g.Nx, g.Ny, g.Nz};
  const int d1 = (dir + 1) % 3, d2 = (dir + 2) % 3;
  const long n = (long)depth * Nt[d1] * Nt[d2];
  const long id = gid();
  if (id >= n) return;
  int x = (int)(id % depth);
  const int t = (int)(id / depth);
  int idx[3];
  idx[d1] = t % Nt[d1];
  idx[d2] = t / Nt[d1];
  // x counts layers outward-in from the block edge
  int pos, src = 0;
  if (side == 0) {
    pos = x;
  } else {
    pos = Nt[dir] - 1 - x;
  }
  if (kind == F3D_BC_PERIODIC) {
    src = (side == 0) ? pos + N[dir] : pos - N[dir];
  } else if (kind == F3D_BC_COPY) {
    src = (side == 0) ? NG + 1 : NG + N[dir] - 2;
  }
  idx[dir] = pos;
  const long NP = g.NP;
  const long p = idx[0] + (long)g.Nxt * (idx[1] + (long)g.Nyt * idx[2]);
  if (kind == F3D_BC_INFLOW) {
    for (int q = 0; q < 7; ++q) Q[p + q * NP] = st.Q[q];
    for (int q = 0; q < 5; ++q) U[p + q * NP] = st.U[q];
    for (int q = 0; q < g.Ns; ++q) rhoi[p + q * NP] = st.r[q];
  } else {
    idx[dir] = src;
    const long s = idx[0] + (long)g.Nxt * (idx[1] + (long)g.Nyt * idx[2]);
    for (int q = 0; q < 7; ++q) Q[p + q * NP] = Q[s + q * NP];
    for (int q = 0; q < 5; ++q) U[p + q * NP] = U[s + q * NP];
    for (int q = 0; q < g.Ns; ++q) rhoi[p + q * NP] = rhoi[s + q * NP];
  }
}

// ---- launchers --------------------------------------------------------------
// The end of an RK3 stage in one pass per interior cell: divSpecs and div
// (div.jl:16-84: the face-flux differences into rho_k and U), for stages 2
// and 3 linComb with the stage-1 copies (utils.jl:95-112, solver.jl:304-317),
// then c2Prim (utils.jl:2-48) -- the arithmetic of k_div_specs, k_div,
// k_lincomb and k_c2prim in the same order, with U and rho_k read once and
// written once instead of three times (nothing between those kernels reads
// a neighbouring cell, and flowAdvance does not read rho_k).
template <int NS, typename R, bool WARM>
__global__ void __launch_bounds__(kBlock)
    k_stage_end(double* __restrict__ U, const double* __restrict__ Un, double* __restrict__ rhoi,
                const double* __restrict__ rhon, double* __restrict__ Q, const R* __restrict__ Fx,
                const R* __restrict__ Fy, const R* __restrict__ Fz, const R* __restrict__ Fvx,
                const R* __restrict__ Fvy, const R* __restrict__ Fvz, const R* __restrict__ Fxi,
                const R* __restrict__ Fyi, const R* __restrict__ Fzi, const R* __restrict__ Fdx,
                const R* __restrict__ Fdy, const R* __restrict__ Fdz, double dt,
                const double* __restrict__ J, const int64_t* __restrict__ tag,
                const DevParams* __restrict__ P, Geom g, int lin, double la, double lb) {
  const long id = gid();
  if (id >= (long)g.Nx * g.Ny * g.Nz) return;
  int a, b, c;
  interior_idx(id, g, a, b, c);
  const long p = (a + NG) + (long)g.Nxt * ((b + NG) + (long)g.Nyt * (c + NG));
  if (tag && tag[p] != 0) return;
  const long NP = g.NP;
  const long nfx = (long)(g.Nx + 1) * g.Ny * g.Nz, nfy = (long)g.Nx * (g.Ny + 1) * g.Nz,
             nfz = (long)g.Nx * g.Ny * (g.Nz + 1);
  const long fx = a + (long)(g.Nx + 1) * (b + (long)g.Ny * c);
  const long fy = a + (long)g.Nx * (b + (long)(g.Ny + 1) * c);
  const long fz = a + (long)g.Nx * (b + (long)g.Ny * c);
  const long ox = 1, oy = g.Nx, oz = (long)g.Nx * g.Ny;
  const R Jact = (R)(J[p] * dt);
  // divSpecs
  double r[NS];
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    const long nx = n * nfx, ny = n * nfy, nz = n * nfz;
    const R d = (Fxi[fx + nx] - Fxi[fx + ox + nx] + Fyi[fy + ny] - Fyi[fy + oy + ny] +
                 Fzi[fz + nz] - Fzi[fz + oz + nz] + Fdx[fx + ox + nx] - Fdx[fx + nx] +
                 Fdy[fy + oy + ny] - Fdy[fy + ny] + Fdz[fz + oz + nz] - Fdz[fz + nz]) *
                Jact;
    r[n] = rhoi[p + n * NP] + (double)d;
  }
  // div
  R dV[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const R v1 = Fvx[fx + ox + q * nfx] - Fvx[fx + q * nfx];
    const R v2 = Fvy[fy + oy + q * nfy] - Fvy[fy + q * nfy];
    const R v3 = Fvz[fz + oz + q * nfz] - Fvz[fz + q * nfz];
    dV[q] = v1 + v2 + v3;
  }
  double u[5];
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    const R d = (Fx[fx + n * nfx] - Fx[fx + ox + n * nfx] + Fy[fy + n * nfy] -
                 Fy[fy + oy + n * nfy] + Fz[fz + n * nfz] - Fz[fz + oz + n * nfz]) *
                Jact;
    u[n] = U[p + n * NP] + (double)d;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) u[q + 1] += (double)(dV[q] * Jact);
  // linComb
  if (lin) {
#pragma unroll
    for (int n = 0; n < 5; ++n) u[n] = u[n] * la + Un[p + n * NP] * lb;
#pragma unroll
    for (int n = 0; n < NS; ++n) r[n] = r[n] * la + rhon[p + n * NP] * lb;
  }
#pragma unroll
  for (int n = 0; n < 5; ++n) U[p + n * NP] = u[n];
  // c2Prim
  const double rho = fmax(u[0], kEps);
  const double rinv = 1 / rho;
  double sr = 0.0;
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    r[n] = fmax(r[n], 0.0);
    sr += r[n];
  }
#pragma unroll
  for (int n = 0; n < NS; ++n) r[n] *= rho / sr;
  const double uu = u[1] * rinv;
  const double vv = u[2] * rinv;
  const double ww = u[3] * rinv;
  const double ei = fmax(u[4] - 0.5 * rho * (uu * uu + vv * vv + ww * ww), kEps);
  const double T = fmax(WARM ? get_T_from<NS>(P, ei, r, Q[p + 5 * NP]) : get_T<NS>(P, ei, r, nullptr),
                       kEps);
  const double pr = fmax(pmixture<NS>(P, T, r), kEps);
#pragma unroll
  for (int n = 0; n < NS; ++n) rhoi[p + n * NP] = r[n];
  Q[p] = rho;
  Q[p + NP] = uu;
  Q[p + 2 * NP] = vv;
  Q[p + 3 * NP] = ww;
  Q[p + 4 * NP] = pr;
  Q[p + 5 * NP] = T;
  Q[p + 6 * NP] = ei;
}

int launch_stage_end(f3d_ctx* c, int KRK, double* U, const double* Un, double* rhoi,
                     const double* rhon, double* Q, void* const* F, double dt, const double* J,
                     const int64_t* tag, hipStream_t s) {
  // F: Fx, Fy, Fz, Fvx, Fvy, Fvz, Fxi, Fyi, Fzi, Fdx, Fdy, Fdz
  const long n = (long)c->g.Nx * c->g.Ny * c->g.Nz;
  const int lin = KRK > 1;
  const double la = KRK == 2 ? 0.25 : 2.0 / 3.0, lb = KRK == 2 ? 0.75 : 1.0 / 3.0;
#define F3D_SE(NS_, R_, W_)                                                                      \
  hipLaunchKernelGGL((k_stage_end<NS_, R_, W_>), dim3(nblocks(n)), dim3(kBlock), 0, s, U, Un,   \
                     rhoi, rhon, Q, (const R_*)F[0], (const R_*)F[1], (const R_*)F[2],            \
                     (const R_*)F[3], (const R_*)F[4], (const R_*)F[5], (const R_*)F[6],          \
                     (const R_*)F[7], (const R_*)F[8], (const R_*)F[9], (const R_*)F[10],         \
                     (const R_*)F[11], dt, J, tag, c->dparams, c->g, lin, la, lb)
  const bool f64 = c->precision == F3D_FP64;
  switch (c->g.Ns) {
    case 1: if (f64) F3D_SE(1, double, false); else F3D_SE(1, float, true); break;
    case 5: if (f64) F3D_SE(5, double, false); else F3D_SE(5, float, true); break;
    case 9: if (f64) F3D_SE(9, double, false); else F3D_SE(9, float, true); break;
    default: set_error("stage_end: Ns must be 1, 5 or 9"); return F3D_ERR_UNSUPPORTED;
  }
#undef F3D_SE
  return hip_check(hipGetLastError(), "stage_end");
}

int launch_div(f3d_ctx* c, double* U, const void* Fx, const void* Fy, const void* Fz,
               const void* Fvx, const void* Fvy, const void* Fvz, double dt, const double* J,
               const int64_t* tag, hipStream_t s) {
  const long n = (long)c->g.Nx * c->g.Ny * c->g.Nz;
  if (c->precision == F3D_FP64)
    hipLaunchKernelGGL(k_div<double>, dim3(nblocks(n)), dim3(kBlock), 0, s, U, (const double*)Fx,
                       (const double*)Fy, (const double*)Fz, (const double*)Fvx,
                       (const double*)Fvy, (const double*)Fvz, dt, J, tag, c->g);
  else
    hipLaunchKernelGGL(k_div<float>, dim3(nblocks(n)), dim3(kBlock), 0, s, U, (const float*)Fx,
                       (const float*)Fy, (const float*)Fz, (const float*)Fvx, (const float*)Fvy,
                       (const float*)Fvz, dt, J, tag, c->g);
  return hip_check(hipGetLastError(), "div");
}

int launch_div_specs(f3d_ctx* c, double* rhoi, const void* Fx, const void* Fy, const void* Fz,
                     const void* Fdx, const void* Fdy, const void* Fdz, double dt,
                     const double* J, const int64_t* tag, hipStream_t s) {
  const long n = (long)c->g.Nx * c->g.Ny * c->g.Nz;
  if (c->precision == F3D_FP64)
    hipLaunchKernelGGL(k_div_specs<double>, dim3(nblocks(n)), dim3(kBlock), 0, s, rhoi,
                       (const double*)Fx, (const double*)Fy, (const double*)Fz,
                       (const double*)Fdx, (const double*)Fdy, (const double*)Fdz, dt, J, tag,
                       c->g);
  else
    hipLaunchKernelGGL(k_div_specs<float>, dim3(nblocks(n)), dim3(kBlock), 0, s, rhoi,
                       (const float*)Fx, (const float*)Fy, (const float*)Fz, (const float*)Fdx,
                       (const float*)Fdy, (const float*)Fdz, dt, J, tag, c->g);
  return hip_check(hipGetLastError(), "div_specs");
}

template <int NS>
static void mixture_ns(f3d_ctx* c, const double* Q, const double* rhoi, void* Yi, void* lam,
                       void* mu, void* D, const int64_t* tag, hipStream_t s) {
  const int NP = c->g.NP;
  if (c->precision == F3D_FP64)
    hipLaunchKernelGGL((k_mixture<NS, double>), dim3(nblocks(NP)), dim3(kBlock), 0, s, Q, rhoi,
                       (double*)Yi, (double*)lam, (double*)mu, (double*)D, tag, c->dparams, NP);
  else
    hipLaunchKernelGGL((k_mixture<NS, float>), dim3(nblocks(NP)), dim3(kBlock), 0, s, Q, rhoi,
                       (float*)Yi, (float*)lam, (float*)mu, (float*)D, tag, c->dparams, NP);
}

int launch_mixture(f3d_ctx* c, const double* Q, const double* rhoi, void* Yi, void* lam,
                   void* mu, void* D, const int64_t* tag, hipStream_t s) {
  switch (c->g.Ns) {
    case 1: mixture_ns<1>(c, Q, rhoi, Yi, lam, mu, D, tag, s); break;
    case 5: mixture_ns<5>(c, Q, rhoi, Yi, lam, mu, D, tag, s); break;
    case 9: mixture_ns<9>(c, Q, rhoi, Yi, lam, mu, D, tag, s); break;
    default: set_error("mixture: Ns must be 1, 5 or 9"); return F3D_ERR_UNSUPPORTED;
  }
  return hip_check(hipGetLastError(), "mixture");
}

int launch_c2prim(f3d_ctx* c, double* U, double* Q, double* rhoi, const int64_t* tag,
                  hipStream_t s) {
  const long n = (long)c->g.Nx * c->g.Ny * c->g.Nz;
  const bool warm = c->precision == F3D_FP32;
#define F3D_C2P(NS_)                                                                         \
  if (warm)                                                                                  \
    hipLaunchKernelGGL((k_c2prim<NS_, true>), dim3(nblocks(n)), dim3(kBlock), 0, s, U, Q, rhoi, \
                       tag, c->dparams, c->g);                                               \
  else                                                                                       \
    hipLaunchKernelGGL((k_c2prim<NS_, false>), dim3(nblocks(n)), dim3(kBlock), 0, s, U, Q,      \
                       rhoi, tag, c->dparams, c->g)
  switch (c->g.Ns) {
    case 1: F3D_C2P(1); break;
    case 5: F3D_C2P(5); break;
    case 9: F3D_C2P(9); break;
    default: set_error("c2prim: Ns must be 1, 5 or 9"); return F3D_ERR_UNSUPPORTED;
  }
#undef F3D_C2P
  return hip_check(hipGetLastError(), "c2prim");
}

int launch_prim2c(f3d_ctx* c, double* U, const double* Q, hipStream_t s) {
  hipLaunchKernelGGL(k_prim2c, dim3(nblocks(c->g.NP)), dim3(kBlock), 0, s, U, Q, c->g.NP);
  return hip_check(hipGetLastError(), "prim2c");
}

int launch_lincomb(f3d_ctx* c, double* U, const double* Un, int NV, double a, double b,
                   const int64_t* tag, hipStream_t s) {
  const long n = (long)c->g.Nx * c->g.Ny * c->g.Nz;
  hipLaunchKernelGGL(k_lincomb, dim3(nblocks(n)), dim3(kBlock), 0, s, U, Un, NV, a, b, tag, c->g);
  return hip_check(hipGetLastError(), "lincomb");
}

int launch_get_y(f3d_ctx* c, double* Yi, const double* rhoi, const double* Q,
                 const int64_t* tag, hipStream_t s) {
  hipLaunchKernelGGL(k_get_y, dim3(nblocks(c->g.NP)), dim3(kBlock), 0, s, Yi, rhoi, Q, tag,
                     c->g.NP, c->g.Ns);
  return hip_check(hipGetLastError(), "get_y");
}

int launch_fill(f3d_ctx* c, double* Q, double* U, double* rhoi, hipStream_t s) {
  const Geom& g = c->g;
  FillState st;
  for (int q = 0; q < 7; ++q) st.Q[q] = c->inflow_Q[q];
  for (int q = 0; q < MAXNS; ++q) st.r[q] = c->inflow_rhoi[q];
  {
    const double rho = st.Q[0], u = st.Q[1], v = st.Q[2], w = st.Q[3], ei = st.Q[6];
    st.U[0] = rho;
    st.U[1] = u * rho;
    st.U[2] = v * rho;
    st.U[3] = w * rho;
    st.U[4] = ei + 0.5 * rho * (u * u + v * v + w * w);
  }
  const int Nt[3] = {g.Nxt, g.Nyt, g.Nzt};
  for (int dir = 0; dir < 3; ++dir) {
    for (int side = 0; side < 2; ++side) {
      const int kind = c->bc[2 * dir + side];
      if (kind == F3D_BC_KEEP) continue;
      const int depth = (kind == F3D_BC_PERIODIC) ? NG : NG + 1;
      const long n = (long)depth * Nt[(dir + 1) % 3] * Nt[(dir + 2) % 3];
      hipLaunchKernelGGL(k_fill, dim3(nblocks(n)), dim3(kBlock), 0, s, Q, U, rhoi, g, dir, side,
                         kind, depth, st);
      int rc = hip_check(hipGetLastError(), "fill");
      if (rc) return rc;
    }
  }
  return F3D_OK;
}

}  // namespace f3d
