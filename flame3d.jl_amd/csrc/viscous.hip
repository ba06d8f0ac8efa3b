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

template <typename MT>
struct Metrics {
  const MT* m[9];
};

// viscousFlux_x/y/z, viscous.jl:1-428 (LES branches off, run.jl:7-8)
template <typename R, typename MT>
__global__ void __launch_bounds__(kBlock)
    k_viscous(R* __restrict__ Fv, const double* __restrict__ Q, Metrics<MT> M,
              const MT* __restrict__ J, const R* __restrict__ lam, const R* __restrict__ mu,
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
template <int NS, typename R, typename MT>
__global__ void __launch_bounds__(kBlock)
    k_spec_viscous(R* __restrict__ Fd, const double* __restrict__ Q, const R* __restrict__ Yi,
                   Metrics<MT> M, const MT* __restrict__ J, const R* __restrict__ D,
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

// ---------------------------------------------------------------------------
// FP32 y and z sweeps: one thread per (x, transverse) line marches through a
// segment of faces along the sweep.  The cross derivatives at face j-1/2 are
// averages of cell-centred central differences at cells j-1 and j
// (viscous.jl:36-52, 476-479), so each cell row is visited once: its own
// values enter a 4-row register window for the normal derivative and its
// tangential differences (x neighbours from the same cache lines, the other
// transverse direction's neighbours from the adjacent planes, which the
// XCD-ordered grid keeps in one L2) enter a 2-row window.  Same arithmetic
// as k_viscous / k_spec_viscous up to the association of the averaged sums.
// ---------------------------------------------------------------------------
struct MarchMap {
  int a, tr, j0, j1;   // x index, other transverse index, face range [j0, j1)
  long base, s, t2;    // padded index of (a, tr, row 0 - NG), sweep and transverse strides
  long fbase, fstr, nf;
};

__device__ inline bool march_map(const Geom& g, int dir, int seg, MarchMap& m) {
  const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
  const unsigned hw = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const unsigned L = xcd_remap(hw, nb);
  const int bx = (int)(L % gridDim.x);
  m.tr = (int)((L / gridDim.x) % gridDim.y);
  const int sg = (int)(L / (gridDim.x * gridDim.y));
  m.a = bx * 64 + (int)threadIdx.x;
  if (m.a >= g.Nx) return false;
  const int nd = (dir == 1 ? g.Ny : g.Nz) + 1;
  m.j0 = sg * seg;
  m.j1 = min(m.j0 + seg, nd);
  if (m.j0 >= m.j1) return false;
  const long sy = g.Nxt, sz = (long)g.Nxt * g.Nyt;
  m.s = (dir == 1) ? sy : sz;
  m.t2 = (dir == 1) ? sz : sy;
  m.base = (m.a + NG) + m.t2 * (m.tr + NG);
  m.nf = (dir == 1) ? (long)g.Nx * nd * g.Nz : (long)g.Nx * g.Ny * nd;
  m.fbase = (dir == 1) ? m.a + (long)g.Nx * nd * m.tr : m.a + (long)g.Nx * m.tr;
  m.fstr = (dir == 1) ? (long)g.Nx : (long)g.Nx * g.Ny;
  return true;
}

// 4th-order central difference (without the 1/2 of the face average)
__device__ inline float cdiff4(const double* __restrict__ q, long ic, long t) {
  const float c23 = 2.0f / 3.0f, c12 = 1.0f / 12.0f;
  return c23 * ((float)q[ic + t] - (float)q[ic - t]) -
         c12 * ((float)q[ic + 2 * t] - (float)q[ic - 2 * t]);
}

template <typename MT>
__global__ void __launch_bounds__(64)
    k_viscous_march(float* __restrict__ Fv, const double* __restrict__ Q, Metrics<MT> M,
                    const MT* __restrict__ J, const float* __restrict__ lam,
                    const float* __restrict__ mu, const float* __restrict__ Fh,
                    const int64_t* __restrict__ tag, Geom g, int dir, int seg) {
  MarchMap mm;
  if (!march_map(g, dir, seg, mm)) return;
  const long NP = g.NP;
  const int qv[4] = {1, 2, 3, 5};  // u, v, w, T
  float q[4][4];                    // rows j-2 .. j+1
  float dX[4][2], dT[4][2];         // x and other-transverse differences at rows j-1, j
  float mt[10][2], mu2[2], la2[2];  // 9 metrics + J, mu, lambda at rows j-1, j
  auto row_cross = [&](long ic, int r) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double* fq = Q + qv[k] * NP;
      dX[k][r] = cdiff4(fq, ic, 1);
      dT[k][r] = cdiff4(fq, ic, mm.t2);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) mt[k][r] = (float)M.m[k][ic];
    mt[9][r] = (float)J[ic];
    mu2[r] = mu[ic];
    la2[r] = lam[ic];
  };
  {
    const long ic0 = mm.base + (long)(mm.j0 + NG) * mm.s;
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int r = 0; r < 3; ++r) q[k][r + 1] = (float)Q[qv[k] * NP + ic0 + (r - 2) * mm.s];
    row_cross(ic0 - mm.s, 1);
  }
  const float c12 = 1.0f / 12.0f, c23 = 2.0f / 3.0f;
  for (int j = mm.j0; j < mm.j1; ++j) {
    const long ic = mm.base + (long)(j + NG) * mm.s;
    // shift the windows: row j-1 becomes slot 0, row j is loaded into slot 1
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      q[k][0] = q[k][1];
      q[k][1] = q[k][2];
      q[k][2] = q[k][3];
      q[k][3] = (float)Q[qv[k] * NP + ic + mm.s];
      dX[k][0] = dX[k][1];
      dT[k][0] = dT[k][1];
    }
#pragma unroll
    for (int k = 0; k < 10; ++k) mt[k][0] = mt[k][1];
    mu2[0] = mu2[1];
    la2[0] = la2[1];
    row_cross(ic, 1);
    if (tag && tag[ic] == 1) continue;
    float m[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) m[k] = (mt[k][0] + mt[k][1]) * 0.5f;
    const float Jac = (mt[9][0] + mt[9][1]) * 0.5f;
    const float mui = (mu2[0] + mu2[1]) * 0.5f;
    const float lami = (la2[0] + la2[1]) * 0.5f;
    float dq[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float dn = 1.25f * (q[k][2] - q[k][1]) - c12 * (q[k][3] - q[k][0]);
      const float dx = 0.5f * (dX[k][0] + dX[k][1]);
      const float dt = 0.5f * (dT[k][0] + dT[k][1]);
      dq[k][0] = dx;
      dq[k][1] = (dir == 1) ? dn : dt;
      dq[k][2] = (dir == 1) ? dt : dn;
    }
    const float u = (q[0][1] + q[0][2]) * 0.5f;
    const float v = (q[1][1] + q[1][2]) * 0.5f;
    const float w = (q[2][1] + q[2][2]) * 0.5f;
    float gr[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int jj = 0; jj < 3; ++jj)
        gr[k][jj] = (dq[k][0] * m[jj] + dq[k][1] * m[3 + jj] + dq[k][2] * m[6 + jj]) * Jac;
    const float divu = gr[0][0] + gr[1][1] + gr[2][2];
    const float t11 = mui * (2.0f * gr[0][0] - c23 * divu);
    const float t12 = mui * (gr[0][1] + gr[1][0]);
    const float t13 = mui * (gr[0][2] + gr[2][0]);
    const float t22 = mui * (2.0f * gr[1][1] - c23 * divu);
    const float t23 = mui * (gr[2][1] + gr[1][2]);
    const float t33 = mui * (2.0f * gr[2][2] - c23 * divu);
    const long f = mm.fbase + (long)j * mm.fstr;
    const long nf = mm.nf;
    const float E1 = u * t11 + v * t12 + w * t13 + lami * gr[3][0] + Fh[f];
    const float E2 = u * t12 + v * t22 + w * t23 + lami * gr[3][1] + Fh[f + nf];
    const float E3 = u * t13 + v * t23 + w * t33 + lami * gr[3][2] + Fh[f + 2 * nf];
    const float n0 = m[3 * dir], n1 = m[3 * dir + 1], n2 = m[3 * dir + 2];
    Fv[f] = n0 * t11 + n1 * t12 + n2 * t13;
    Fv[f + nf] = n0 * t12 + n1 * t22 + n2 * t23;
    Fv[f + 2 * nf] = n0 * t13 + n1 * t23 + n2 * t33;
    Fv[f + 3 * nf] = n0 * E1 + n1 * E2 + n2 * E3;
  }
}

template <int NS, typename MT>
__global__ void __launch_bounds__(64)
    k_spec_viscous_march(float* __restrict__ Fd, const double* __restrict__ Q,
                         const float* __restrict__ Yi, Metrics<MT> M, const MT* __restrict__ J,
                         const float* __restrict__ D, float* __restrict__ Fh,
                         const int64_t* __restrict__ tag, const DevParams* __restrict__ P, Geom g,
                         int dir, int seg) {
  MarchMap mm;
  if (!march_map(g, dir, seg, mm)) return;
  const long NP = g.NP;
  float y[NS][4];                    // rows j-2 .. j+1
  float cX[NS][2], cT[NS][2], d2[NS][2];
  float mt[10][2], rho2[2], T2[2];
  auto row_cross = [&](long ic, int r) {
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const float* yy = Yi + (long)n * NP;
      cX[n][r] = yy[ic + 1] - yy[ic - 1];
      cT[n][r] = yy[ic + mm.t2] - yy[ic - mm.t2];
      d2[n][r] = D[ic + (long)n * NP];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) mt[k][r] = (float)M.m[k][ic];
    mt[9][r] = (float)J[ic];
    rho2[r] = (float)Q[ic];
    T2[r] = (float)Q[ic + 5 * NP];
  };
  {
    const long ic0 = mm.base + (long)(mm.j0 + NG) * mm.s;
#pragma unroll
    for (int n = 0; n < NS; ++n)
#pragma unroll
      for (int r = 0; r < 3; ++r) y[n][r + 1] = Yi[(long)n * NP + ic0 + (r - 2) * mm.s];
    row_cross(ic0 - mm.s, 1);
  }
  const float c12 = 1.0f / 12.0f;
  for (int j = mm.j0; j < mm.j1; ++j) {
    const long ic = mm.base + (long)(j + NG) * mm.s;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      y[n][0] = y[n][1];
      y[n][1] = y[n][2];
      y[n][2] = y[n][3];
      y[n][3] = Yi[(long)n * NP + ic + mm.s];
      cX[n][0] = cX[n][1];
      cT[n][0] = cT[n][1];
      d2[n][0] = d2[n][1];
    }
#pragma unroll
    for (int k = 0; k < 10; ++k) mt[k][0] = mt[k][1];
    rho2[0] = rho2[1];
    T2[0] = T2[1];
    row_cross(ic, 1);
    if (tag && tag[ic] == 1) continue;
    float m[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) m[k] = (mt[k][0] + mt[k][1]) * 0.5f;
    const float Jac = (mt[9][0] + mt[9][1]) * 0.5f;
    const float rho = (rho2[0] + rho2[1]) * 0.5f;
    const float T = (T2[0] + T2[1]) * 0.5f;
    float hi[NS], V1[NS], V2[NS], V3[NS];
    h_specs<NS, float>(P, T, hi);
    float sum1 = 0, sum2 = 0, sum3 = 0;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const float rhoDi = (d2[n][0] + d2[n][1]) * rho;
      const float dn = 1.25f * (y[n][2] - y[n][1]) - c12 * (y[n][3] - y[n][0]);
      const float dx = 0.25f * (cX[n][0] + cX[n][1]);
      const float dt = 0.25f * (cT[n][0] + cT[n][1]);
      const float e0 = dx, e1 = (dir == 1) ? dn : dt, e2 = (dir == 1) ? dt : dn;
      const float Vx = (e0 * m[0] + e1 * m[3] + e2 * m[6]) * Jac * rhoDi;
      const float Vy = (e0 * m[1] + e1 * m[4] + e2 * m[7]) * Jac * rhoDi;
      const float Vz = (e0 * m[2] + e1 * m[5] + e2 * m[8]) * Jac * rhoDi;
      V1[n] = Vx;
      V2[n] = Vy;
      V3[n] = Vz;
      sum1 += Vx;
      sum2 += Vy;
      sum3 += Vz;
    }
    const float n0 = m[3 * dir], n1 = m[3 * dir + 1], n2 = m[3 * dir + 2];
    const long f = mm.fbase + (long)j * mm.fstr;
    const long nf = mm.nf;
    float fh1 = 0, fh2 = 0, fh3 = 0;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const float Yn = 0.5f * (y[n][1] + y[n][2]);
      const float a1 = V1[n] - sum1 * Yn;
      const float a2 = V2[n] - sum2 * Yn;
      const float a3 = V3[n] - sum3 * Yn;
      Fd[f + n * nf] = a1 * n0 + a2 * n1 + a3 * n2;
      fh1 += a1 * hi[n];
      fh2 += a2 * hi[n];
      fh3 += a3 * hi[n];
    }
    Fh[f] = fh1;
    Fh[f + nf] = fh2;
    Fh[f + 2 * nf] = fh3;
  }
}

constexpr int kViscSeg = 64;

static void march_grid(const Geom& g, int dir, dim3& grid, int& seg) {
  const int nd = (dir == 1 ? g.Ny : g.Nz) + 1;
  const int nseg = (nd + kViscSeg - 1) / kViscSeg;
  seg = (nd + nseg - 1) / nseg;
  grid = dim3((g.Nx + 63) / 64, dir == 1 ? g.Nz : g.Ny, nseg);
}

template <typename MT>
static Metrics<MT> mk(const void* const* M) {
  Metrics<MT> r;
  for (int q = 0; q < 9; ++q) r.m[q] = (const MT*)M[q];
  return r;
}

int launch_viscous(f3d_ctx* c, int dir, void* Fv, const double* Q, const void* const* M,
                   const void* J, const void* lam, const void* mu, const void* Fh,
                   const int64_t* tag, hipStream_t s, bool mf) {
  const Geom& g = c->g;
  const long nf = (long)(g.Nx + (dir == 0)) * (g.Ny + (dir == 1)) * (g.Nz + (dir == 2));
  const dim3 gr(nblocks(nf)), bl(kBlock);
  if (c->precision == F3D_FP64)
    hipLaunchKernelGGL((k_viscous<double, double>), gr, bl, 0, s, (double*)Fv, Q, mk<double>(M),
                       (const double*)J, (const double*)lam, (const double*)mu,
                       (const double*)Fh, tag, g, dir);
  else if (dir != 0) {
    dim3 mg;
    int seg;
    march_grid(g, dir, mg, seg);
    if (mf)
      hipLaunchKernelGGL((k_viscous_march<float>), mg, dim3(64), 0, s, (float*)Fv, Q,
                         mk<float>(M), (const float*)J, (const float*)lam, (const float*)mu,
                         (const float*)Fh, tag, g, dir, seg);
    else
      hipLaunchKernelGGL((k_viscous_march<double>), mg, dim3(64), 0, s, (float*)Fv, Q,
                         mk<double>(M), (const double*)J, (const float*)lam, (const float*)mu,
                         (const float*)Fh, tag, g, dir, seg);
  } else if (mf)
    hipLaunchKernelGGL((k_viscous<float, float>), gr, bl, 0, s, (float*)Fv, Q, mk<float>(M),
                       (const float*)J, (const float*)lam, (const float*)mu, (const float*)Fh,
                       tag, g, dir);
  else
    hipLaunchKernelGGL((k_viscous<float, double>), gr, bl, 0, s, (float*)Fv, Q, mk<double>(M),
                       (const double*)J, (const float*)lam, (const float*)mu, (const float*)Fh,
                       tag, g, dir);
  return hip_check(hipGetLastError(), "viscous_flux");
}

template <int NS>
static void spec_visc_ns(f3d_ctx* c, int dir, void* Fd, const double* Q, const void* Yi,
                         const void* const* M, const void* J, const void* D, void* Fh,
                         const int64_t* tag, hipStream_t s, bool mf) {
  const Geom& g = c->g;
  const long nf = (long)(g.Nx + (dir == 0)) * (g.Ny + (dir == 1)) * (g.Nz + (dir == 2));
  const dim3 gr(nblocks(nf)), bl(kBlock);
  if (c->precision == F3D_FP64)
    hipLaunchKernelGGL((k_spec_viscous<NS, double, double>), gr, bl, 0, s, (double*)Fd, Q,
                       (const double*)Yi, mk<double>(M), (const double*)J, (const double*)D,
                       (double*)Fh, tag, c->dparams, g, dir);
  else if (dir != 0) {
    dim3 mg;
    int seg;
    march_grid(g, dir, mg, seg);
    if (mf)
      hipLaunchKernelGGL((k_spec_viscous_march<NS, float>), mg, dim3(64), 0, s, (float*)Fd, Q,
                         (const float*)Yi, mk<float>(M), (const float*)J, (const float*)D,
                         (float*)Fh, tag, c->dparams, g, dir, seg);
    else
      hipLaunchKernelGGL((k_spec_viscous_march<NS, double>), mg, dim3(64), 0, s, (float*)Fd, Q,
                         (const float*)Yi, mk<double>(M), (const double*)J, (const float*)D,
                         (float*)Fh, tag, c->dparams, g, dir, seg);
  } else if (mf)
    hipLaunchKernelGGL((k_spec_viscous<NS, float, float>), gr, bl, 0, s, (float*)Fd, Q,
                       (const float*)Yi, mk<float>(M), (const float*)J, (const float*)D,
                       (float*)Fh, tag, c->dparams, g, dir);
  else
    hipLaunchKernelGGL((k_spec_viscous<NS, float, double>), gr, bl, 0, s, (float*)Fd, Q,
                       (const float*)Yi, mk<double>(M), (const double*)J, (const float*)D,
                       (float*)Fh, tag, c->dparams, g, dir);
}

int launch_spec_viscous(f3d_ctx* c, int dir, void* Fd, const double* Q, const void* Yi,
                        const void* const* M, const void* J, const void* D, void* Fh,
                        const int64_t* tag, hipStream_t s, bool mf) {
  switch (c->g.Ns) {
    case 1: spec_visc_ns<1>(c, dir, Fd, Q, Yi, M, J, D, Fh, tag, s, mf); break;
    case 5: spec_visc_ns<5>(c, dir, Fd, Q, Yi, M, J, D, Fh, tag, s, mf); break;
    case 9: spec_visc_ns<9>(c, dir, Fd, Q, Yi, M, J, D, Fh, tag, s, mf); break;
    default: set_error("spec_viscous_flux: Ns must be 1, 5 or 9"); return F3D_ERR_UNSUPPORTED;
  }
  return hip_check(hipGetLastError(), "spec_viscous_flux");
}

}  // namespace f3d
