// Fused FP32 stage sweeps.  Per direction, the convective right-hand side
// (Steger-Warming splitting -> hybrid UP7 / WENO5-Z / NND reconstruction ->
// face-flux difference, split.jl / schemes.jl) and the diffusive one
// (viscous stresses and heat flux, species diffusion and its enthalpy flux,
// viscous.jl) of every interior cell are accumulated into one float array
//   rhs[v][cell],  v = 0..4 (rho, rho u, rho v, rho w, E), 5 + k (rho_k),
//   cell = a + Nx (b + Ny c),
// so the split fluxes, face fluxes, Fv, Fd and Fh of the reference's
// decomposition (solver.jl:22-56) never reach HBM; the stage end applies
// U += J dt rhs (div.jl:16-84), linComb and c2Prim in one pass.
//
// x sweeps: one 256-thread block per segment of up to 256 cells of an x line;
// the split fluxes (convective) or the face values (diffusive) of the segment
// are staged in LDS so each point is read from HBM once and each face is
// evaluated once; the face-flux difference of a cell reads its two faces
// from LDS.  y and z sweeps: one thread per (x, transverse) line marches
// through a segment of cells with the stencil window in registers (x
// coalesced loads), the previous face's value kept for the difference.
// The convective march splits the variables over the waves of a block
// (wave 0: the five flow fluxes, waves 1..: groups of three species) so each
// wave's window fits its registers; the waves of a block march the same
// lines and share their loads of u, v, w and the metrics through L1.
//
// Arithmetic per face is the kernel-granular path's (split_point,
// weno_face_v, the viscous face formulas); the sums of the face differences
// over the three directions are associated per direction instead of per
// cell, a rounding-level difference on the float path.
#include "f3d_stencil.h"

namespace f3d {

constexpr int kXSeg = 256;  // cells per x-line block
constexpr int kConvSeg = 32;  // cells per march segment (convective)
constexpr int kViscSeg = 64;  // cells per march segment (viscous)

__device__ inline void rhs_put(float* __restrict__ r, long i, float d, int acc) {
  r[i] = acc ? r[i] + d : d;
}

// ---------------------------------------------------------------------------
// Convective x sweep
// ---------------------------------------------------------------------------
template <int NS>
__global__ void __launch_bounds__(kXSeg)
    k_sw_conv_x(float* __restrict__ rhs, const double* __restrict__ Q,
                const double* __restrict__ rhoi, const float* __restrict__ A1,
                const float* __restrict__ A2, const float* __restrict__ A3,
                const float* __restrict__ phi, const DevParams* __restrict__ P, Geom g,
                int acc) {
  constexpr int NV = 5 + NS;
  constexpr int NPT = kXSeg + 8;
  __shared__ float sp[NV][NPT];
  __shared__ float sm[NV][NPT];
  __shared__ float sphi[NPT];
  const int nseg = (g.Nx + kXSeg - 1) / kXSeg;
  const unsigned L = xcd_block();
  const int seg = (int)(L % (unsigned)nseg);
  const int line = (int)(L / (unsigned)nseg);
  const int b = line % g.Ny, c = line / g.Ny;
  const int c0 = seg * kXSeg;
  const int nc = min(kXSeg, g.Nx - c0);
  const int t = threadIdx.x;
  const long NP = g.NP;
  // point q of the segment is padded i = c0 + q (cell c0 + q - NG)
  const long row = (long)c0 + (long)g.Nxt * ((b + NG) + (long)g.Nyt * (c + NG));
  const float sw = PT<float>::sweps2(P);
  for (int q = t; q < nc + 8; q += kXSeg) {
    const long p = row + q;
    const float u = (float)Q[p + NP], v = (float)Q[p + 2 * NP], w = (float)Q[p + 3 * NP];
    const float a1 = A1[p], a2 = A2[p], a3 = A3[p];
    float fp[5], fm[5];
    split_point<float>((float)Q[p], u, v, w, (float)Q[p + 4 * NP], (float)Q[p + 6 * NP], a1, a2,
                       a3, sw, fp, fm);
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      sp[n][q] = fp[n];
      sm[n][q] = fm[n];
    }
    const float un = a1 * u + a2 * v + a3 * w;
    const float Ep = 0.5f * (un + fabsf(un));
    const float Em = 0.5f * (un - fabsf(un));
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const float r = (float)rhoi[p + (long)n * NP];
      sp[5 + n][q] = Ep * r;
      sm[5 + n][q] = Em * r;
    }
    sphi[q] = phi[p];
  }
  __syncthreads();
  // face f (0..nc) between cells c0+f-1 and c0+f: stencil points q = f .. f+7
  float fv[NV];
  if (t < nc) {
    const int br = weno_branch<float>(fmaxf(sphi[t + 3], sphi[t + 4]), P);
    weno_faces_v<float, NV>(br, P, [&](int n, float* pv, float* mv) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        pv[k] = sp[n][t + k];
        mv[k] = sm[n][t + k];
      }
    }, fv);
  }
  // the segment's last face, one variable per thread
  float fl = 0.0f;
  if (t < NV) {
    const int br = weno_branch<float>(fmaxf(sphi[nc + 3], sphi[nc + 4]), P);
    float pv[8], mv[8];
    for (int k = 0; k < 8; ++k) {
      pv[k] = sp[t][nc + k];
      mv[k] = sm[t][nc + k];
    }
    fl = weno_face_v<float>(pv, mv, br, P);
  }
  __syncthreads();
  float(*sf)[NPT] = sp;  // face values, reusing the F+ tile
  if (t < nc) {
#pragma unroll
    for (int n = 0; n < NV; ++n) sf[n][t] = fv[n];
  }
  if (t < NV) sf[t][nc] = fl;
  __syncthreads();
  if (t < nc) {
    const long NC = (long)g.Nx * g.Ny * g.Nz;
    const long cell = (c0 + t) + (long)g.Nx * (b + (long)g.Ny * c);
#pragma unroll
    for (int n = 0; n < NV; ++n) rhs_put(rhs, n * NC + cell, sf[n][t] - sf[n][t + 1], acc);
  }
}

// ---------------------------------------------------------------------------
// Convective y / z march
// ---------------------------------------------------------------------------
struct SweepLine {
  int a, tr, c0, c1;  // x index, other transverse index, cells [c0, c1) along the sweep
  long base, s;       // padded index of (a, tr) at row -NG; sweep stride
  long cell0, cstr;   // rhs cell index of row 0 and its stride along the sweep
};

__device__ inline bool sweep_line(const Geom& g, int dir, int seg, SweepLine& m) {
  const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
  const unsigned hw = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const unsigned L = xcd_remap(hw, nb);
  const int bx = (int)(L % gridDim.x);
  m.tr = (int)((L / gridDim.x) % gridDim.y);
  const int sg = (int)(L / (gridDim.x * gridDim.y));
  m.a = bx * 64 + (int)threadIdx.x;
  if (m.a >= g.Nx) return false;
  const int nd = dir == 1 ? g.Ny : g.Nz;
  m.c0 = sg * seg;
  m.c1 = min(m.c0 + seg, nd);
  if (m.c0 >= m.c1) return false;
  const long sy = g.Nxt, sz = (long)g.Nxt * g.Nyt;
  m.s = (dir == 1) ? sy : sz;
  const long t2 = (dir == 1) ? sz : sy;
  m.base = (m.a + NG) + t2 * (m.tr + NG);
  m.cell0 = (dir == 1) ? m.a + (long)g.Nx * g.Ny * m.tr : m.a + (long)g.Nx * m.tr;
  m.cstr = (dir == 1) ? (long)g.Nx : (long)g.Nx * g.Ny;
  return true;
}

// The marches keep their loads two steps ahead of their use: the inputs of
// the stencil row a step needs, its sensor value and the right-hand-side
// entries it accumulates into are fetched two steps earlier into registers,
// so each wave has three steps' loads in flight instead of waiting on each.
constexpr int kAhead = 2;

// wave 0 of a block: the five flow fluxes
struct FlowIn {
  double rho, u, v, w, p, e;
  float a1, a2, a3, phi;
};

__device__ inline void conv_march_flow(float* __restrict__ rhs, const double* __restrict__ Q,
                                       const float* __restrict__ A1, const float* __restrict__ A2,
                                       const float* __restrict__ A3,
                                       const float* __restrict__ phi,
                                       const DevParams* __restrict__ P, const Geom& g,
                                       const SweepLine& m, int acc) {
  const long NP = g.NP, NC = (long)g.Nx * g.Ny * g.Nz;
  const float sw = PT<float>::sweps2(P);
  float wp[5][8], wm[5][8];
  // row r's inputs and the sensor of row r - 3 (the face f = r - 3 needs
  // rows f - 4 .. f + 3 and the sensor of rows f - 1, f)
  auto fetch = [&](int r, FlowIn& x, bool with_phi) {
    const long p = m.base + (long)(r + NG) * m.s;
    x.rho = Q[p];
    x.u = Q[p + NP];
    x.v = Q[p + 2 * NP];
    x.w = Q[p + 3 * NP];
    x.p = Q[p + 4 * NP];
    x.e = Q[p + 6 * NP];
    x.a1 = A1[p];
    x.a2 = A2[p];
    x.a3 = A3[p];
    x.phi = with_phi ? phi[p - 3 * m.s] : 0.0f;  // rows r >= c0 + 3 only: row r - 3 >= 0
  };
  auto split = [&](const FlowIn& x, int slot) {
    float fp[5], fm[5];
    split_point<float>((float)x.rho, (float)x.u, (float)x.v, (float)x.w, (float)x.p, (float)x.e,
                       x.a1, x.a2, x.a3, sw, fp, fm);
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      wp[n][slot] = fp[n];
      wm[n][slot] = fm[n];
    }
  };
  auto fetch_rhs = [&](int f, float* o) {  // the entries of cell f - 1
    const long cell = m.cell0 + (long)(f - 1) * m.cstr;
#pragma unroll
    for (int n = 0; n < 5; ++n) o[n] = (acc && f > m.c0 && f <= m.c1) ? rhs[n * NC + cell] : 0.0f;
  };
  {
    FlowIn x;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      fetch(m.c0 - 4 + k, x, false);
      split(x, k);
    }
  }
  float phA = phi[m.base + (long)(m.c0 - 1 + NG) * m.s];
  const int last = m.c1 + 3;  // the last row the segment reads
  FlowIn in[kAhead + 1];
  float ro[kAhead + 1][5];
#pragma unroll
  for (int d = 0; d < kAhead; ++d) {
    if (m.c0 + 3 + d <= last) fetch(m.c0 + 3 + d, in[d], true);
    fetch_rhs(m.c0 + d, ro[d]);
  }
  float prev[5];
  for (int f = m.c0; f <= m.c1; ++f) {
    if (f + 3 + kAhead <= last) fetch(f + 3 + kAhead, in[kAhead], true);
    fetch_rhs(f + kAhead, ro[kAhead]);
    split(in[0], 7);
    const float phB = in[0].phi;
    const int br = weno_branch<float>(fmaxf(phA, phB), P);
    float cur[5];
    weno_faces_v<float, 5>(br, P, [&](int n, float* pv, float* mv) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        pv[k] = wp[n][k];
        mv[k] = wm[n][k];
      }
    }, cur);
    if (f > m.c0) {
      const long cell = m.cell0 + (long)(f - 1) * m.cstr;
#pragma unroll
      for (int n = 0; n < 5; ++n) rhs[n * NC + cell] = ro[0][n] + (prev[n] - cur[n]);
    }
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      prev[n] = cur[n];
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        wp[n][k] = wp[n][k + 1];
        wm[n][k] = wm[n][k + 1];
      }
    }
#pragma unroll
    for (int d = 0; d < kAhead; ++d) {
      in[d] = in[d + 1];
#pragma unroll
      for (int n = 0; n < 5; ++n) ro[d][n] = ro[d + 1][n];
    }
    phA = phB;
  }
}

// waves 1..: species k0 .. k0+2
struct SpecIn {
  double u, v, w, r[3];
  float a1, a2, a3, phi;
};

template <int NS>
__device__ inline void conv_march_species(float* __restrict__ rhs, const double* __restrict__ Q,
                                          const double* __restrict__ rhoi,
                                          const float* __restrict__ A1,
                                          const float* __restrict__ A2,
                                          const float* __restrict__ A3,
                                          const float* __restrict__ phi,
                                          const DevParams* __restrict__ P, const Geom& g,
                                          const SweepLine& m, int k0, int acc) {
  const long NP = g.NP, NC = (long)g.Nx * g.Ny * g.Nz;
  float wp[3][8], wm[3][8];
  auto fetch = [&](int r, SpecIn& x, bool with_phi) {
    const long p = m.base + (long)(r + NG) * m.s;
    x.u = Q[p + NP];
    x.v = Q[p + 2 * NP];
    x.w = Q[p + 3 * NP];
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) x.r[kk] = (k0 + kk < NS) ? rhoi[p + (long)(k0 + kk) * NP] : 0.0;
    x.a1 = A1[p];
    x.a2 = A2[p];
    x.a3 = A3[p];
    x.phi = with_phi ? phi[p - 3 * m.s] : 0.0f;  // rows r >= c0 + 3 only: row r - 3 >= 0
  };
  auto split = [&](const SpecIn& x, int slot) {
    const float un = x.a1 * (float)x.u + x.a2 * (float)x.v + x.a3 * (float)x.w;
    const float Ep = 0.5f * (un + fabsf(un));
    const float Em = 0.5f * (un - fabsf(un));
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) {
      const float r = (float)x.r[kk];
      wp[kk][slot] = Ep * r;
      wm[kk][slot] = Em * r;
    }
  };
  auto fetch_rhs = [&](int f, float* o) {
    const long cell = m.cell0 + (long)(f - 1) * m.cstr;
#pragma unroll
    for (int kk = 0; kk < 3; ++kk)
      o[kk] = (acc && k0 + kk < NS && f > m.c0 && f <= m.c1) ? rhs[(5 + k0 + kk) * NC + cell]
                                                             : 0.0f;
  };
  {
    SpecIn x;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      fetch(m.c0 - 4 + k, x, false);
      split(x, k);
    }
  }
  float phA = phi[m.base + (long)(m.c0 - 1 + NG) * m.s];
  const int last = m.c1 + 3;
  SpecIn in[kAhead + 1];
  float ro[kAhead + 1][3];
#pragma unroll
  for (int d = 0; d < kAhead; ++d) {
    if (m.c0 + 3 + d <= last) fetch(m.c0 + 3 + d, in[d], true);
    fetch_rhs(m.c0 + d, ro[d]);
  }
  float prev[3];
  for (int f = m.c0; f <= m.c1; ++f) {
    if (f + 3 + kAhead <= last) fetch(f + 3 + kAhead, in[kAhead], true);
    fetch_rhs(f + kAhead, ro[kAhead]);
    split(in[0], 7);
    const float phB = in[0].phi;
    const int br = weno_branch<float>(fmaxf(phA, phB), P);
    float cur[3];
    weno_faces_v<float, 3>(br, P, [&](int kk, float* pv, float* mv) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        pv[k] = wp[kk][k];
        mv[k] = wm[kk][k];
      }
    }, cur);
#pragma unroll
    for (int kk = 0; kk < 3; ++kk)
      if (k0 + kk >= NS) cur[kk] = 0.0f;
    if (f > m.c0) {
      const long cell = m.cell0 + (long)(f - 1) * m.cstr;
#pragma unroll
      for (int kk = 0; kk < 3; ++kk)
        if (k0 + kk < NS) rhs[(5 + k0 + kk) * NC + cell] = ro[0][kk] + (prev[kk] - cur[kk]);
    }
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) {
      prev[kk] = cur[kk];
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        wp[kk][k] = wp[kk][k + 1];
        wm[kk][k] = wm[kk][k + 1];
      }
    }
#pragma unroll
    for (int d = 0; d < kAhead; ++d) {
      in[d] = in[d + 1];
#pragma unroll
      for (int kk = 0; kk < 3; ++kk) ro[d][kk] = ro[d + 1][kk];
    }
    phA = phB;
  }
}

template <int NS>
__global__ void __launch_bounds__(64 * (1 + (NS + 2) / 3)) __attribute__((amdgpu_waves_per_eu(3)))
    k_sw_conv_m(float* __restrict__ rhs, const double* __restrict__ Q,
                const double* __restrict__ rhoi, const float* __restrict__ A1,
                const float* __restrict__ A2, const float* __restrict__ A3,
                const float* __restrict__ phi, const DevParams* __restrict__ P, Geom g, int dir,
                int seg, int acc) {
  SweepLine m;
  if (!sweep_line(g, dir, seg, m)) return;
  const int grp = threadIdx.y;
  if (grp == 0)
    conv_march_flow(rhs, Q, A1, A2, A3, phi, P, g, m, acc);
  else
    conv_march_species<NS>(rhs, Q, rhoi, A1, A2, A3, phi, P, g, m, 3 * (grp - 1), acc);
}

// ---------------------------------------------------------------------------
// Diffusive x sweep: viscous stresses and heat flux (viscous.jl:1-142) and
// species diffusion (viscous.jl:430-507) at one face, float arithmetic as the
// kernel-granular k_viscous / k_spec_viscous; the enthalpy flux Fh enters the
// energy flux through the face normal, n . Fh.
// ---------------------------------------------------------------------------
struct FMet {
  const float* m[9];
};

// x-line block of the viscous sweep with each cell's inputs read once: the
// metrics, J, mu and lambda of the cells -1 .. nc are staged in LDS, then
// variable by variable (u, v, w, T) its values on the cells -2 .. nc+1 and
// its transverse fourth-order differences (y, z) on the cells -1 .. nc (two
// buffers, one barrier per variable, the next variable's loads in flight
// across it); face f (between cells f-1 and f) takes its gradients from its
// two cells (viscous.jl:1-142: the face average of the cells' transverse
// differences, the fourth-order normal difference), then its stresses and
// heat flux; the face values go back to LDS for the difference of each
// cell's two faces (rhs 1..4).  A full segment has kXSeg + 1 faces: its last
// thread also takes the last face.
__global__ void __launch_bounds__(kXSeg)
    k_sw_visc_x(float* __restrict__ rhs, const double* __restrict__ Q, FMet M,
                const float* __restrict__ J, const float* __restrict__ lam,
                const float* __restrict__ mu, Geom g) {
  constexpr int NE = kXSeg + 2;  // cells -1 .. nc
  constexpr int NQ = kXSeg + 4;  // cells -2 .. nc + 1
  __shared__ float sm[12][NE];   // metrics, J, mu, lambda
  __shared__ float sq[2][NQ], sy[2][NE], sz[2][NE];
  __shared__ float sf[4][kXSeg + 1];
  const int nseg = (g.Nx + kXSeg - 1) / kXSeg;
  const unsigned L = xcd_block();
  const int seg = (int)(L % (unsigned)nseg);
  const int line = (int)(L / (unsigned)nseg);
  const int b = line % g.Ny, c = line / g.Ny;
  const int c0 = seg * kXSeg;
  const int nc = min(kXSeg, g.Nx - c0);
  const int t = threadIdx.x;
  const long NP = g.NP, ty = g.Nxt, tz = (long)g.Nxt * g.Nyt;
  const long row = (long)(c0 + NG) + ty * (b + NG) + tz * (c + NG);  // cell 0
  const int qv[4] = {1, 2, 3, 5};  // u, v, w, T
  // variable k's loads of entries t and t + kXSeg of the staged ranges (the
  // second exists for t < 4 in a full segment)
  double rq[2], rt[2][8];
  auto fetch = [&](int k) {
    const double* fq = Q + qv[k] * NP;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int i = t + j * kXSeg;
      if (i < nc + 4) rq[j] = fq[row + i - 2];
      if (i < nc + 2) {
        const long ic = row + i - 1;
        rt[j][0] = fq[ic + ty];
        rt[j][1] = fq[ic - ty];
        rt[j][2] = fq[ic + 2 * ty];
        rt[j][3] = fq[ic - 2 * ty];
        rt[j][4] = fq[ic + tz];
        rt[j][5] = fq[ic - tz];
        rt[j][6] = fq[ic + 2 * tz];
        rt[j][7] = fq[ic - 2 * tz];
      }
    }
  };
  fetch(0);
  for (int i = t; i < nc + 2; i += kXSeg) {
    const long ic = row + i - 1;
#pragma unroll
    for (int q = 0; q < 9; ++q) sm[q][i] = M.m[q][ic];
    sm[9][i] = J[ic];
    sm[10][i] = mu[ic];
    sm[11][i] = lam[ic];
  }
  const float c12 = 1.0f / 12.0f, c23 = 2.0f / 3.0f;
  const bool own = t <= nc;                          // face t
  const bool extra = nc == kXSeg && t == kXSeg - 1;  // and face kXSeg
  float dq[4][3], vel[3], dx[4][3], vx[3];           // face t's and the extra face's
  auto grads = [&](int f, int bb, int k, float (&d)[4][3], float (&vv)[3]) {
    d[k][0] = 1.25f * (sq[bb][f + 2] - sq[bb][f + 1]) - c12 * (sq[bb][f + 3] - sq[bb][f]);
    d[k][1] = 0.5f * (sy[bb][f] + sy[bb][f + 1]);
    d[k][2] = 0.5f * (sz[bb][f] + sz[bb][f + 1]);
    if (k < 3) vv[k] = (sq[bb][f + 1] + sq[bb][f + 2]) * 0.5f;
  };
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int bb = k & 1;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int i = t + j * kXSeg;
      if (i < nc + 4) sq[bb][i] = (float)rq[j];
      if (i < nc + 2) {
        sy[bb][i] = c23 * ((float)rt[j][0] - (float)rt[j][1]) - c12 * ((float)rt[j][2] - (float)rt[j][3]);
        sz[bb][i] = c23 * ((float)rt[j][4] - (float)rt[j][5]) - c12 * ((float)rt[j][6] - (float)rt[j][7]);
      }
    }
    if (k + 1 < 4) fetch(k + 1);
    __syncthreads();
    if (own) grads(t, bb, k, dq, vel);
    if (extra) grads(kXSeg, bb, k, dx, vx);
  }
  auto stress = [&](int f, const float (&d)[4][3], const float (&vv)[3]) {
    float m[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) m[q] = (sm[q][f] + sm[q][f + 1]) * 0.5f;
    const float Jac = (sm[9][f] + sm[9][f + 1]) * 0.5f;
    const float mui = (sm[10][f] + sm[10][f + 1]) * 0.5f;
    const float lami = (sm[11][f] + sm[11][f + 1]) * 0.5f;
    float gr[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        gr[k][j] = (d[k][0] * m[j] + d[k][1] * m[3 + j] + d[k][2] * m[6 + j]) * Jac;
    const float divu = gr[0][0] + gr[1][1] + gr[2][2];
    const float t11 = mui * (2.0f * gr[0][0] - c23 * divu);
    const float t12 = mui * (gr[0][1] + gr[1][0]);
    const float t13 = mui * (gr[0][2] + gr[2][0]);
    const float t22 = mui * (2.0f * gr[1][1] - c23 * divu);
    const float t23 = mui * (gr[2][1] + gr[1][2]);
    const float t33 = mui * (2.0f * gr[2][2] - c23 * divu);
    const float u = vv[0], v = vv[1], w = vv[2];
    const float E1 = u * t11 + v * t12 + w * t13 + lami * gr[3][0];
    const float E2 = u * t12 + v * t22 + w * t23 + lami * gr[3][1];
    const float E3 = u * t13 + v * t23 + w * t33 + lami * gr[3][2];
    const float n0 = m[0], n1 = m[1], n2 = m[2];
    sf[0][f] = n0 * t11 + n1 * t12 + n2 * t13;
    sf[1][f] = n0 * t12 + n1 * t22 + n2 * t23;
    sf[2][f] = n0 * t13 + n1 * t23 + n2 * t33;
    sf[3][f] = n0 * E1 + n1 * E2 + n2 * E3;
  };
  if (own) stress(t, dq, vel);
  if (extra) stress(kXSeg, dx, vx);
  __syncthreads();
  if (t < nc) {
    const long NC = (long)g.Nx * g.Ny * g.Nz;
    const long cell = (c0 + t) + (long)g.Nx * (b + (long)g.Ny * c);
    // diffusive fluxes enter with the opposite sign of the convective ones
#pragma unroll
    for (int q = 0; q < 4; ++q) rhs_put(rhs, (q + 1) * NC + cell, sf[q][t + 1] - sf[q][t], 1);
  }
}

// x-line block of the species-diffusion sweep with each cell's inputs read
// once: the metrics, J, rho and T of the cells -1 .. nc are staged in LDS,
// then species by species Y_k of the cells -2 .. nc+1 and the transverse
// differences and D_k of the cells -1 .. nc (two buffers, one barrier per
// species, the next species' loads in flight across it); face f (between
// cells f-1 and f) forms its gradient from them
// (viscous.jl:430-507: the face average of the cells' transverse central
// differences, the fourth-order normal difference).  A full segment has
// nc + 1 = kXSeg + 1 faces: its last thread also takes the last face.
template <int NS>
__global__ void __launch_bounds__(kXSeg)
    k_sw_spec_x(float* __restrict__ rhs, const double* __restrict__ Q,
                const float* __restrict__ Yi, FMet M, const float* __restrict__ J,
                const float* __restrict__ D, const DevParams* __restrict__ P, Geom g) {
  constexpr int NE = kXSeg + 2;  // cells -1 .. nc
  constexpr int NY = kXSeg + 4;  // cells -2 .. nc + 1
  __shared__ float sm[12][NE];   // metrics, J, rho, T
  __shared__ float sy[2][NY], scy[2][NE], scz[2][NE], sd[2][NE];
  __shared__ float sf[NS + 1][kXSeg + 1];
  __shared__ float sx[4][NS];  // the extra face's V_x, V_y, V_z and Y
  const int nseg = (g.Nx + kXSeg - 1) / kXSeg;
  const unsigned L = xcd_block();
  const int seg = (int)(L % (unsigned)nseg);
  const int line = (int)(L / (unsigned)nseg);
  const int b = line % g.Ny, c = line / g.Ny;
  const int c0 = seg * kXSeg;
  const int nc = min(kXSeg, g.Nx - c0);
  const int t = threadIdx.x;
  const long NP = g.NP, ty = g.Nxt, tz = (long)g.Nxt * g.Nyt;
  const long row = (long)(c0 + NG) + ty * (b + NG) + tz * (c + NG);  // cell 0
  for (int i = t; i < nc + 2; i += kXSeg) {
    const long ic = row + i - 1;
#pragma unroll
    for (int q = 0; q < 9; ++q) sm[q][i] = M.m[q][ic];
    sm[9][i] = J[ic];
    sm[10][i] = (float)Q[ic];
    sm[11][i] = (float)Q[ic + 5 * NP];
  }
  const bool own = t <= nc;                       // face t
  const bool extra = nc == kXSeg && t == kXSeg - 1;  // and face kXSeg
  const float c12 = 1.0f / 12.0f;
  float V[3][NS], Yn[NS];
  // V_k of face f (staged index f: cell f-1, f + 1: cell f) for species n
  auto fspec = [&](int f, int bb, float* out) {
    float m[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) m[q] = (sm[q][f] + sm[q][f + 1]) * 0.5f;
    const float Jac = (sm[9][f] + sm[9][f + 1]) * 0.5f;
    const float rho = (sm[10][f] + sm[10][f + 1]) * 0.5f;
    const float rhoDi = (sd[bb][f] + sd[bb][f + 1]) * rho;
    const float dY0 = 1.25f * (sy[bb][f + 2] - sy[bb][f + 1]) - c12 * (sy[bb][f + 3] - sy[bb][f]);
    const float dY1 = 0.25f * (scy[bb][f + 1] + scy[bb][f]);
    const float dY2 = 0.25f * (scz[bb][f + 1] + scz[bb][f]);
    out[0] = (dY0 * m[0] + dY1 * m[3] + dY2 * m[6]) * Jac * rhoDi;
    out[1] = (dY0 * m[1] + dY1 * m[4] + dY2 * m[7]) * Jac * rhoDi;
    out[2] = (dY0 * m[2] + dY1 * m[5] + dY2 * m[8]) * Jac * rhoDi;
    out[3] = 0.5f * (sy[bb][f + 1] + sy[bb][f + 2]);
  };
  // species n + 1's loads are issued before species n's barrier and faces
  // (entries t and t + kXSeg of the staged ranges; the second exists for
  // t < 4 in a full segment)
  float ry[2], rt[2][5];
  auto fetch = [&](int n) {
    const float* y = Yi + (long)n * NP;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int i = t + k * kXSeg;
      if (i < nc + 4) ry[k] = y[row + i - 2];
      if (i < nc + 2) {
        const long ic = row + i - 1;
        rt[k][0] = y[ic + ty];
        rt[k][1] = y[ic - ty];
        rt[k][2] = y[ic + tz];
        rt[k][3] = y[ic - tz];
        rt[k][4] = D[ic + (long)n * NP];
      }
    }
  };
  fetch(0);
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    const int bb = n & 1;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int i = t + k * kXSeg;
      if (i < nc + 4) sy[bb][i] = ry[k];
      if (i < nc + 2) {
        scy[bb][i] = rt[k][0] - rt[k][1];
        scz[bb][i] = rt[k][2] - rt[k][3];
        sd[bb][i] = rt[k][4];
      }
    }
    if (n + 1 < NS) fetch(n + 1);
    __syncthreads();
    float o[4];
    if (own) {
      fspec(t, bb, o);
      V[0][n] = o[0];
      V[1][n] = o[1];
      V[2][n] = o[2];
      Yn[n] = o[3];
    }
    if (extra) {
      fspec(kXSeg, bb, o);
#pragma unroll
      for (int k = 0; k < 4; ++k) sx[k][n] = o[k];
    }
  }
  // the face's corrected diffusion velocities, its species fluxes and the
  // enthalpy flux n . sum_k V_k h_k
  auto ffin = [&](int f, const float (&Vf)[3][NS], const float (&Yf)[NS]) {
    const float n0 = (sm[0][f] + sm[0][f + 1]) * 0.5f;
    const float n1 = (sm[1][f] + sm[1][f + 1]) * 0.5f;
    const float n2 = (sm[2][f] + sm[2][f + 1]) * 0.5f;
    const float T = (sm[11][f] + sm[11][f + 1]) * 0.5f;
    float hi[NS];
    h_specs<NS, float>(P, T, hi);
    float s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      s1 += Vf[0][n];
      s2 += Vf[1][n];
      s3 += Vf[2][n];
    }
    float fh1 = 0, fh2 = 0, fh3 = 0;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const float V1 = Vf[0][n] - s1 * Yf[n];
      const float V2 = Vf[1][n] - s2 * Yf[n];
      const float V3 = Vf[2][n] - s3 * Yf[n];
      sf[n][f] = V1 * n0 + V2 * n1 + V3 * n2;
      fh1 += V1 * hi[n];
      fh2 += V2 * hi[n];
      fh3 += V3 * hi[n];
    }
    sf[NS][f] = n0 * fh1 + n1 * fh2 + n2 * fh3;
  };
  if (own) ffin(t, V, Yn);
  if (extra) {
    float Vx[3][NS], Yx[NS];
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      Vx[0][n] = sx[0][n];
      Vx[1][n] = sx[1][n];
      Vx[2][n] = sx[2][n];
      Yx[n] = sx[3][n];
    }
    ffin(kXSeg, Vx, Yx);
  }
  __syncthreads();
  if (t < nc) {
    const long NC = (long)g.Nx * g.Ny * g.Nz;
    const long cell = (c0 + t) + (long)g.Nx * (b + (long)g.Ny * c);
#pragma unroll
    for (int n = 0; n < NS; ++n) rhs_put(rhs, (5 + n) * NC + cell, sf[n][t + 1] - sf[n][t], 1);
    rhs_put(rhs, 4 * NC + cell, sf[NS][t + 1] - sf[NS][t], 1);
  }
}

// ---------------------------------------------------------------------------
// Diffusive y / z marches.  The cross derivatives at face j-1/2 average the
// cell-centred central differences of cells j-1 and j (viscous.jl:36-52,
// :476-479), so each row's tangential differences are computed once and kept
// for the next face.
// ---------------------------------------------------------------------------
__device__ inline float cd4(const double* __restrict__ q, long ic, long t) {
  const float c23 = 2.0f / 3.0f, c12 = 1.0f / 12.0f;
  return c23 * ((float)q[ic + t] - (float)q[ic - t]) -
         c12 * ((float)q[ic + 2 * t] - (float)q[ic - 2 * t]);
}

__global__ void __launch_bounds__(64)
    k_sw_visc_m(float* __restrict__ rhs, const double* __restrict__ Q, FMet M,
                const float* __restrict__ J, const float* __restrict__ lam,
                const float* __restrict__ mu, Geom g, int dir, int seg) {
  SweepLine mm;
  if (!sweep_line(g, dir, seg, mm)) return;
  const long NP = g.NP, NC = (long)g.Nx * g.Ny * g.Nz;
  const long t2 = (dir == 1) ? (long)g.Nxt * g.Nyt : (long)g.Nxt;
  const int qv[4] = {1, 2, 3, 5};  // u, v, w, T
  float q[4][4];                    // rows f-2 .. f+1
  float dX[4][2], dT[4][2];         // x and other-transverse differences at rows f-1, f
  float mt[10][2], mu2[2], la2[2];
  auto row_cross = [&](long ic, int r) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double* fq = Q + qv[k] * NP;
      dX[k][r] = cd4(fq, ic, 1);
      dT[k][r] = cd4(fq, ic, t2);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) mt[k][r] = M.m[k][ic];
    mt[9][r] = J[ic];
    mu2[r] = mu[ic];
    la2[r] = lam[ic];
  };
  {
    const long ic0 = mm.base + (long)(mm.c0 + NG) * mm.s;
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int r = 0; r < 3; ++r) q[k][r + 1] = (float)Q[qv[k] * NP + ic0 + (r - 2) * mm.s];
    row_cross(ic0 - mm.s, 1);
  }
  const float c12 = 1.0f / 12.0f, c23 = 2.0f / 3.0f;
  float prev[4];
  for (int f = mm.c0; f <= mm.c1; ++f) {
    const long ic = mm.base + (long)(f + NG) * mm.s;
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
    const float E1 = u * t11 + v * t12 + w * t13 + lami * gr[3][0];
    const float E2 = u * t12 + v * t22 + w * t23 + lami * gr[3][1];
    const float E3 = u * t13 + v * t23 + w * t33 + lami * gr[3][2];
    const float n0 = m[3 * dir], n1 = m[3 * dir + 1], n2 = m[3 * dir + 2];
    float cur[4];
    cur[0] = n0 * t11 + n1 * t12 + n2 * t13;
    cur[1] = n0 * t12 + n1 * t22 + n2 * t23;
    cur[2] = n0 * t13 + n1 * t23 + n2 * t33;
    cur[3] = n0 * E1 + n1 * E2 + n2 * E3;
    if (f > mm.c0) {
      const long cell = mm.cell0 + (long)(f - 1) * mm.cstr;
#pragma unroll
      for (int k = 0; k < 4; ++k) rhs_put(rhs, (k + 1) * NC + cell, cur[k] - prev[k], 1);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) prev[k] = cur[k];
  }
}

template <int NS>
__global__ void __launch_bounds__(64)
    k_sw_spec_m(float* __restrict__ rhs, const double* __restrict__ Q,
                const float* __restrict__ Yi, FMet M, const float* __restrict__ J,
                const float* __restrict__ D, const DevParams* __restrict__ P, Geom g, int dir,
                int seg) {
  SweepLine mm;
  if (!sweep_line(g, dir, seg, mm)) return;
  const long NP = g.NP, NC = (long)g.Nx * g.Ny * g.Nz;
  const long t2 = (dir == 1) ? (long)g.Nxt * g.Nyt : (long)g.Nxt;
  float y[NS][4];  // rows f-2 .. f+1
  float cX[NS][2], cT[NS][2], d2[NS][2];
  float mt[10][2], rho2[2], T2[2];
  // the loads of one step, fetched a step ahead: row f+1 of Y, the cross
  // differences' inputs, D, metrics, J, rho and T of row f, and the
  // right-hand-side entries of cell f-1
  struct Row {
    float y[NS], xp[NS], xm[NS], tp[NS], tm[NS], d[NS], mt[10], r[NS + 1];
    double rho, T;
  };
  auto fetch = [&](int f, Row& R) {
    const long ic = mm.base + (long)(f + NG) * mm.s;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const float* yy = Yi + (long)n * NP;
      R.y[n] = yy[ic + mm.s];
      R.xp[n] = yy[ic + 1];
      R.xm[n] = yy[ic - 1];
      R.tp[n] = yy[ic + t2];
      R.tm[n] = yy[ic - t2];
      R.d[n] = D[ic + (long)n * NP];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) R.mt[k] = M.m[k][ic];
    R.mt[9] = J[ic];
    R.rho = Q[ic];
    R.T = Q[ic + 5 * NP];
    if (f > mm.c0) {
      const long cell = mm.cell0 + (long)(f - 1) * mm.cstr;
#pragma unroll
      for (int n = 0; n < NS; ++n) R.r[n] = rhs[(5 + n) * NC + cell];
      R.r[NS] = rhs[4 * NC + cell];
    }
  };
  auto take_cross = [&](const Row& R) {
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      cX[n][1] = R.xp[n] - R.xm[n];
      cT[n][1] = R.tp[n] - R.tm[n];
      d2[n][1] = R.d[n];
    }
#pragma unroll
    for (int k = 0; k < 10; ++k) mt[k][1] = R.mt[k];
    rho2[1] = (float)R.rho;
    T2[1] = (float)R.T;
  };
  Row nx;
  {
    const long ic0 = mm.base + (long)(mm.c0 + NG) * mm.s;
#pragma unroll
    for (int n = 0; n < NS; ++n)
#pragma unroll
      for (int r = 0; r < 3; ++r) y[n][r + 1] = Yi[(long)n * NP + ic0 + (r - 2) * mm.s];
    fetch(mm.c0 - 1, nx);  // row c0-1's cross inputs (its rhs entries are not read)
    take_cross(nx);
    fetch(mm.c0, nx);
  }
  const float c12 = 1.0f / 12.0f;
  float prev[NS + 1];
  for (int f = mm.c0; f <= mm.c1; ++f) {
    const Row cu = nx;
    if (f < mm.c1) fetch(f + 1, nx);
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      y[n][0] = y[n][1];
      y[n][1] = y[n][2];
      y[n][2] = y[n][3];
      y[n][3] = cu.y[n];
      cX[n][0] = cX[n][1];
      cT[n][0] = cT[n][1];
      d2[n][0] = d2[n][1];
    }
#pragma unroll
    for (int k = 0; k < 10; ++k) mt[k][0] = mt[k][1];
    rho2[0] = rho2[1];
    T2[0] = T2[1];
    take_cross(cu);
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
    float cur[NS + 1];
    float fh1 = 0, fh2 = 0, fh3 = 0;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const float Yn = 0.5f * (y[n][1] + y[n][2]);
      const float a1 = V1[n] - sum1 * Yn;
      const float a2 = V2[n] - sum2 * Yn;
      const float a3 = V3[n] - sum3 * Yn;
      cur[n] = a1 * n0 + a2 * n1 + a3 * n2;
      fh1 += a1 * hi[n];
      fh2 += a2 * hi[n];
      fh3 += a3 * hi[n];
    }
    cur[NS] = n0 * fh1 + n1 * fh2 + n2 * fh3;
    if (f > mm.c0) {
      const long cell = mm.cell0 + (long)(f - 1) * mm.cstr;
#pragma unroll
      for (int n = 0; n < NS; ++n) rhs[(5 + n) * NC + cell] = cu.r[n] + (cur[n] - prev[n]);
      rhs[4 * NC + cell] = cu.r[NS] + (cur[NS] - prev[NS]);
    }
#pragma unroll
    for (int n = 0; n <= NS; ++n) prev[n] = cur[n];
  }
}

// ---------------------------------------------------------------------------
// Stage end from the accumulated right-hand side: U += J dt rhs, rho_k += J dt
// rhs (div.jl:16-84), linComb with the stage-1 copies for stages 2 and 3
// (utils.jl:95-112), c2Prim with the warm-started e -> T Newton (utils.jl:2-48);
// the pressure with the reciprocal molar masses (no division per species).
// ---------------------------------------------------------------------------
template <int NS>
__global__ void __launch_bounds__(kBlock)
    k_sw_stage_end(double* __restrict__ U, const double* __restrict__ Un,
                   double* __restrict__ rhoi, const double* __restrict__ rhon,
                   double* __restrict__ Q, const float* __restrict__ rhs, double dt,
                   const double* __restrict__ J, const DevParams* __restrict__ P, Geom g,
                   int lin, double la, double lb) {
  const long NC = (long)g.Nx * g.Ny * g.Nz;
  const long id = gid();
  if (id >= NC) return;
  const int a = (int)(id % g.Nx);
  const int tt = (int)(id / g.Nx);
  const int b = tt % g.Ny, c = tt / g.Ny;
  const long p = (a + NG) + (long)g.Nxt * ((b + NG) + (long)g.Nyt * (c + NG));
  const long NP = g.NP;
  const float Jact = (float)(J[p] * dt);
  double r[NS], u[5];
#pragma unroll
  for (int n = 0; n < NS; ++n) r[n] = rhoi[p + n * NP] + (double)(rhs[(5 + n) * NC + id] * Jact);
#pragma unroll
  for (int n = 0; n < 5; ++n) u[n] = U[p + n * NP] + (double)(rhs[n * NC + id] * Jact);
  if (lin) {
#pragma unroll
    for (int n = 0; n < 5; ++n) u[n] = u[n] * la + Un[p + n * NP] * lb;
#pragma unroll
    for (int n = 0; n < NS; ++n) r[n] = r[n] * la + rhon[p + n * NP] * lb;
  }
#pragma unroll
  for (int n = 0; n < 5; ++n) U[p + n * NP] = u[n];
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
  const double T = fmax(get_T_from<NS>(P, ei, r, Q[p + 5 * NP]), kEps);
  const double pr = fmax(pmixture<NS, true>(P, T, r), kEps);
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

// ---- launchers --------------------------------------------------------------
// Convective sweep of direction dir into rhs (acc = 0: the first sweep of the
// stage stores, later sweeps add).  A[3]: the sweep's metrics (float copies).
template <int NS>
static void conv_ns(f3d_ctx* c, int dir, float* rhs, const double* Q, const double* rhoi,
                    const float* const* A, const float* phi, int acc, hipStream_t s) {
  const Geom& g = c->g;
  if (dir == 0) {
    const int nseg = (g.Nx + kXSeg - 1) / kXSeg;
    const long nb = (long)nseg * g.Ny * g.Nz;
    hipLaunchKernelGGL((k_sw_conv_x<NS>), dim3((unsigned)nb), dim3(kXSeg), 0, s, rhs, Q, rhoi,
                       A[0], A[1], A[2], phi, c->dparams, g, acc);
  } else {
    const int nd = dir == 1 ? g.Ny : g.Nz;
    const int nseg = (nd + kConvSeg - 1) / kConvSeg;
    const int seg = (nd + nseg - 1) / nseg;
    const dim3 grid((g.Nx + 63) / 64, dir == 1 ? g.Nz : g.Ny, nseg);
    hipLaunchKernelGGL((k_sw_conv_m<NS>), grid, dim3(64, 1 + (NS + 2) / 3), 0, s, rhs, Q, rhoi,
                       A[0], A[1], A[2], phi, c->dparams, g, dir, seg, acc);
  }
}

// Segments per line of the species-diffusion march.  Its registers hold few
// waves per SIMD (3 for Ns = 9), so the segment count is chosen for whole
// rounds of resident waves: cost ~ rounds x (segment length + 4 priming
// rows), e.g. 3 segments of a 256-cell line at 256^3 (3072 waves, one round
// on 256 CUs) instead of 4 (4096 waves, a second round of 1024).
template <int NS>
static int spec_segments(f3d_ctx* c, int nd, long lines) {
  if (c->spec_waves == 0) {
    int nb = 0, ncu = 0, dev = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void*>(&k_sw_spec_m<NS>), 64, 0) != hipSuccess || nb < 1)
      nb = 1;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        ncu < 1)
      ncu = 1;
    c->spec_waves = (long)nb * ncu;
  }
  const long cap = c->spec_waves;
  int best = 1;
  long bc = -1;
  for (int n = 1; n <= max(1, nd / 16); ++n) {
    const long rounds = (lines * n + cap - 1) / cap;
    const long cost = rounds * ((nd + n - 1) / n + 4);
    if (bc < 0 || cost < bc) {
      bc = cost;
      best = n;
    }
  }
  return best;
}

template <int NS>
static void diff_ns(f3d_ctx* c, int dir, float* rhs, const double* Q, const float* Yi,
                    const FMet& M, const float* J, const float* lam, const float* mu,
                    const float* D, hipStream_t s) {
  const Geom& g = c->g;
  if (dir == 0) {
    const int nseg = (g.Nx + kXSeg - 1) / kXSeg;
    const long nb = (long)nseg * g.Ny * g.Nz;
    hipLaunchKernelGGL(k_sw_visc_x, dim3((unsigned)nb), dim3(kXSeg), 0, s, rhs, Q, M, J, lam, mu,
                       g);
    hipLaunchKernelGGL((k_sw_spec_x<NS>), dim3((unsigned)nb), dim3(kXSeg), 0, s, rhs, Q, Yi, M,
                       J, D, c->dparams, g);
  } else {
    const int nd = dir == 1 ? g.Ny : g.Nz;
    const int nsv = (nd + kViscSeg - 1) / kViscSeg, sv = (nd + nsv - 1) / nsv;
    const unsigned nl = (g.Nx + 63) / 64, nt = dir == 1 ? g.Nz : g.Ny;
    const int nss = spec_segments<NS>(c, nd, (long)nl * nt), ss = (nd + nss - 1) / nss;
    hipLaunchKernelGGL(k_sw_visc_m, dim3(nl, nt, nsv), dim3(64), 0, s, rhs, Q, M, J, lam, mu, g,
                       dir, sv);
    hipLaunchKernelGGL((k_sw_spec_m<NS>), dim3(nl, nt, nss), dim3(64), 0, s, rhs, Q, Yi, M, J, D,
                       c->dparams, g, dir, ss);
  }
}

template <int NS>
static void stage_end_ns(f3d_ctx* c, int KRK, double* U, const double* Un, double* rhoi,
                         const double* rhon, double* Q, const float* rhs, double dt,
                         const double* J, hipStream_t s) {
  const long n = (long)c->g.Nx * c->g.Ny * c->g.Nz;
  const int lin = KRK > 1;
  const double la = KRK == 2 ? 0.25 : 2.0 / 3.0, lb = KRK == 2 ? 0.75 : 1.0 / 3.0;
  hipLaunchKernelGGL((k_sw_stage_end<NS>), dim3(nblocks(n)), dim3(kBlock), 0, s, U, Un, rhoi,
                     rhon, Q, rhs, dt, J, c->dparams, c->g, lin, la, lb);
}

int launch_sweep_conv(f3d_ctx* c, int dir, float* rhs, const double* Q, const double* rhoi,
                      const void* phi, int acc, hipStream_t s) {
  const float* A[3] = {c->mf[3 * dir], c->mf[3 * dir + 1], c->mf[3 * dir + 2]};
  switch (c->g.Ns) {
    case 1: conv_ns<1>(c, dir, rhs, Q, rhoi, A, (const float*)phi, acc, s); break;
    case 5: conv_ns<5>(c, dir, rhs, Q, rhoi, A, (const float*)phi, acc, s); break;
    case 9: conv_ns<9>(c, dir, rhs, Q, rhoi, A, (const float*)phi, acc, s); break;
    default: set_error("sweeps: Ns must be 1, 5 or 9"); return F3D_ERR_UNSUPPORTED;
  }
  return hip_check(hipGetLastError(), "convective sweep");
}

int launch_sweep_diff(f3d_ctx* c, int dir, float* rhs, const double* Q, const void* Yi,
                      const void* lam, const void* mu, const void* D, hipStream_t s) {
  FMet M;
  for (int q = 0; q < 9; ++q) M.m[q] = c->mf[q];
  const float* J = c->mf[9];
  const float *y = (const float*)Yi, *l = (const float*)lam, *m = (const float*)mu,
              *d = (const float*)D;
  switch (c->g.Ns) {
    case 1: diff_ns<1>(c, dir, rhs, Q, y, M, J, l, m, d, s); break;
    case 5: diff_ns<5>(c, dir, rhs, Q, y, M, J, l, m, d, s); break;
    case 9: diff_ns<9>(c, dir, rhs, Q, y, M, J, l, m, d, s); break;
    default: set_error("sweeps: Ns must be 1, 5 or 9"); return F3D_ERR_UNSUPPORTED;
  }
  return hip_check(hipGetLastError(), "diffusive sweep");
}

int launch_sweeps(f3d_ctx* c, float* rhs, const double* Q, const double* rhoi, const void* phi,
                  const void* Yi, const void* lam, const void* mu, const void* D,
                  hipStream_t s) {
  for (int d = 0; d < 3; ++d) {
    int rc = launch_sweep_conv(c, d, rhs, Q, rhoi, phi, d > 0, s);
    if (rc) return rc;
  }
  for (int d = 0; d < 3; ++d) {
    int rc = launch_sweep_diff(c, d, rhs, Q, Yi, lam, mu, D, s);
    if (rc) return rc;
  }
  return F3D_OK;
}

int launch_sweep_stage_end(f3d_ctx* c, int KRK, double* U, const double* Un, double* rhoi,
                           const double* rhon, double* Q, const float* rhs, double dt,
                           const double* J, hipStream_t s) {
  switch (c->g.Ns) {
    case 1: stage_end_ns<1>(c, KRK, U, Un, rhoi, rhon, Q, rhs, dt, J, s); break;
    case 5: stage_end_ns<5>(c, KRK, U, Un, rhoi, rhon, Q, rhs, dt, J, s); break;
    case 9: stage_end_ns<9>(c, KRK, U, Un, rhoi, rhon, Q, rhs, dt, J, s); break;
    default: set_error("sweep stage end: Ns must be 1, 5 or 9"); return F3D_ERR_UNSUPPORTED;
  }
  return hip_check(hipGetLastError(), "sweep stage end");
}

}  // namespace f3d
