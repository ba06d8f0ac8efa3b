// Mesh metrics from node coordinates (parse_mesh.jl:24-221): ghost nodes by
// point reflection through the boundary node, then the coordinate
// derivatives (sixth-order central differences inside, second-order
// one-sided ones in the three outermost layers of the padded block), the
// cofactor metrics (the metric terms times 1/J) and J = 1/det.
//
// Layout: x, y, z, the nine metrics and J are padded arrays of
// (Nx+8)(Ny+8)(Nz+8) doubles, x index fastest (the reference's column-major
// [i, j, k]).  The metric order is dxi/dx, dxi/dy, dxi/dz, deta/dx, ...,
// dzeta/dz, as f3d_set_metrics takes them.  One thread per point; every
// expression keeps the reference's operation order with contraction off, so
// the results equal a plain double-precision evaluation of the same formulas.
#include <string>

#include "f3d_internal.h"

// no fused multiply-adds anywhere in this file (operation order as written)
#pragma clang fp contract(off)

namespace f3d {

namespace {

// Reflection through the boundary node along one axis: ghost node g of a
// side mirrors the interior node at the same distance on the other side of
// the boundary node b (parse_mesh.jl:52-75, 102-113): v[g] = 2 v[b] - v[2b - g].
__device__ inline int mirror(int g, int b) { return 2 * b - g; }

// Phase 1 (parse_mesh.jl:52-75): the x-ghosts of the rows with interior
// (j, k) and the y-ghosts of the columns with interior (i, k).  Neither reads
// the other's output.
__global__ void __launch_bounds__(kBlock)
    k_ghost_xy(double* __restrict__ x, double* __restrict__ y, double* __restrict__ z, Geom g) {
  const long n = (long)g.Nxt * g.Nyt * g.Nz;
  const long t = gid();
  if (t >= n) return;
  const int I = (int)(t % g.Nxt);
  const long r = t / g.Nxt;
  const int J = (int)(r % g.Nyt), K = (int)(r / g.Nyt) + NG;
  const bool gi = I < NG || I >= g.Nx + NG, gj = J < NG || J >= g.Ny + NG;
  if (gi == gj) return;  // interior node, or a corner (phase 2)
  const long sx = 1, sy = g.Nxt;
  long p = I + (long)g.Nxt * (J + (long)g.Nyt * K), pb, pm;
  if (gi) {
    const int b = I < NG ? NG : g.Nx + NG - 1;
    pb = p + (long)(b - I) * sx;
    pm = p + (long)(mirror(I, b) - I) * sx;
  } else {
    const int b = J < NG ? NG : g.Ny + NG - 1;
    pb = p + (long)(b - J) * sy;
    pm = p + (long)(mirror(J, b) - J) * sy;
  }
  x[p] = 2 * x[pb] - x[pm];
  y[p] = 2 * y[pb] - y[pm];
  z[p] = 2 * z[pb] - z[pm];
}

// Phase 2 (parse_mesh.jl:77-100): the four x-y corner columns with interior
// k, v[i, j] = v[i, jb] + v[ib, j] - v[ib, jb] (ib, jb the boundary nodes of
// the corner's sides).
__global__ void __launch_bounds__(kBlock)
    k_ghost_corner(double* __restrict__ x, double* __restrict__ y, double* __restrict__ z, Geom g) {
  const long n = (long)(2 * NG) * (2 * NG) * g.Nz;
  const long t = gid();
  if (t >= n) return;
  const int a = (int)(t % (2 * NG));
  const long r = t / (2 * NG);
  const int bq = (int)(r % (2 * NG)), K = (int)(r / (2 * NG)) + NG;
  const int I = a < NG ? a : g.Nx + a;   // a = NG.. 2NG-1 -> right ghosts Nx+NG ..
  const int J = bq < NG ? bq : g.Ny + bq;
  const int ib = I < NG ? NG : g.Nx + NG - 1, jb = J < NG ? NG : g.Ny + NG - 1;
  const long row = (long)g.Nyt * K;
  const long p = I + (long)g.Nxt * (J + row);
  const long pij = I + (long)g.Nxt * (jb + row);   // (i, jb)
  const long pbj = ib + (long)g.Nxt * (J + row);   // (ib, j)
  const long pbb = ib + (long)g.Nxt * (jb + row);  // (ib, jb)
  x[p] = x[pij] + x[pbj] - x[pbb];
  y[p] = y[pij] + y[pbj] - y[pbb];
  z[p] = z[pij] + z[pbj] - z[pbb];
}

// Phase 3 (parse_mesh.jl:102-113): the z-ghost planes over the whole x-y
// extent (ghost columns included), reflected through k = NG / Nz+NG-1.
__global__ void __launch_bounds__(kBlock)
    k_ghost_z(double* __restrict__ x, double* __restrict__ y, double* __restrict__ z, Geom g) {
  const long plane = (long)g.Nxt * g.Nyt;
  const long n = plane * (2 * NG);
  const long t = gid();
  if (t >= n) return;
  const long q = t % plane;
  const int c = (int)(t / plane);
  const int K = c < NG ? c : g.Nz + c;
  const int b = K < NG ? NG : g.Nz + NG - 1;
  const long p = q + plane * K, pb = q + plane * b, pm = q + plane * mirror(K, b);
  x[p] = 2 * x[pb] - x[pm];
  y[p] = 2 * y[pb] - y[pm];
  z[p] = 2 * z[pb] - z[pm];
}

// d v / d(index) at node I of an axis of n nodes, stride s (parse_mesh.jl:116-130,
// 155-207): CD6 for 3 <= I <= n-4, CD2_L in the first three nodes, CD2_R in
// the last three.
__device__ inline double dindex(const double* __restrict__ v, long p, int I, int n, long s) {
  if (I < 3) {
    const double f1 = v[p], f2 = v[p + s], f3 = v[p + 2 * s];
    return 2 * f2 - 0.5 * f3 - 1.5 * f1;
  }
  if (I > n - 4) {
    const double f1 = v[p - 2 * s], f2 = v[p - s], f3 = v[p];
    return -2 * f2 + 0.5 * f1 + 1.5 * f3;
  }
  const double f1 = v[p - 3 * s], f2 = v[p - 2 * s], f3 = v[p - s];
  const double f5 = v[p + s], f6 = v[p + 2 * s], f7 = v[p + 3 * s];
  return 1.0 / 60 * (f7 - f1) - 3.0 / 20 * (f6 - f2) + 3.0 / 4 * (f5 - f3);
}

// parse_mesh.jl:132-221 at every padded point.
__global__ void __launch_bounds__(kBlock)
    k_metrics(const double* __restrict__ x, const double* __restrict__ y,
              const double* __restrict__ z, double* __restrict__ m0, double* __restrict__ m1,
              double* __restrict__ m2, double* __restrict__ m3, double* __restrict__ m4,
              double* __restrict__ m5, double* __restrict__ m6, double* __restrict__ m7,
              double* __restrict__ m8, double* __restrict__ Jout, Geom g) {
  const long n = (long)g.NP;
  const long p = gid();
  if (p >= n) return;
  const int I = (int)(p % g.Nxt);
  const long r = p / g.Nxt;
  const int J = (int)(r % g.Nyt), K = (int)(r / g.Nyt);
  const long sx = 1, sy = g.Nxt, sz = (long)g.Nxt * g.Nyt;
  const double xxi = dindex(x, p, I, g.Nxt, sx), yxi = dindex(y, p, I, g.Nxt, sx),
               zxi = dindex(z, p, I, g.Nxt, sx);
  const double xeta = dindex(x, p, J, g.Nyt, sy), yeta = dindex(y, p, J, g.Nyt, sy),
               zeta = dindex(z, p, J, g.Nyt, sy);
  const double xze = dindex(x, p, K, g.Nzt, sz), yze = dindex(y, p, K, g.Nzt, sz),
               zze = dindex(z, p, K, g.Nzt, sz);
  Jout[p] = 1 / (xxi * (yeta * zze - yze * zeta) - xeta * (yxi * zze - yze * zxi) +
                 xze * (yxi * zeta - yeta * zxi));
  m0[p] = yeta * zze - yze * zeta;
  m1[p] = xze * zeta - xeta * zze;
  m2[p] = xeta * yze - xze * yeta;
  m3[p] = yze * zxi - yxi * zze;
  m4[p] = xxi * zze - xze * zxi;
  m5[p] = xze * yxi - xxi * yze;
  m6[p] = yxi * zeta - yeta * zxi;
  m7[p] = xeta * zxi - xxi * zeta;
  m8[p] = xxi * yeta - xeta * yxi;
}

}  // namespace

}  // namespace f3d

using namespace f3d;

extern "C" {

int f3d_mesh_ghosts(f3d_ctx* c, double* x, double* y, double* z, hipStream_t s) {
  if (!c || !x || !y || !z) {
    set_error("f3d_mesh_ghosts: null argument");
    return F3D_ERR_ARG;
  }
  const Geom& g = c->g;
  if (g.Nx < NG + 1 || g.Ny < NG + 1 || g.Nz < NG + 1) {
    set_error("f3d_mesh_ghosts: need Nx, Ny, Nz >= 5 (the mirror of the outermost ghost is interior)");
    return F3D_ERR_ARG;
  }
  const long n1 = (long)g.Nxt * g.Nyt * g.Nz, n2 = (long)(2 * NG) * (2 * NG) * g.Nz,
             n3 = (long)g.Nxt * g.Nyt * (2 * NG);
  hipLaunchKernelGGL(k_ghost_xy, dim3(nblocks(n1)), dim3(kBlock), 0, s, x, y, z, g);
  hipLaunchKernelGGL(k_ghost_corner, dim3(nblocks(n2)), dim3(kBlock), 0, s, x, y, z, g);
  hipLaunchKernelGGL(k_ghost_z, dim3(nblocks(n3)), dim3(kBlock), 0, s, x, y, z, g);
  return hip_check(hipGetLastError(), "f3d_mesh_ghosts");
}

int f3d_mesh_metrics(f3d_ctx* c, const double* x, const double* y, const double* z,
                     double* const* metrics, double* J, hipStream_t s) {
  if (!c || !x || !y || !z || !metrics || !J) {
    set_error("f3d_mesh_metrics: null argument");
    return F3D_ERR_ARG;
  }
  for (int q = 0; q < 9; ++q)
    if (!metrics[q]) {
      set_error("f3d_mesh_metrics: null metric array");
      return F3D_ERR_ARG;
    }
  const Geom& g = c->g;
  if (g.Nxt < 7 || g.Nyt < 7 || g.Nzt < 7) {
    set_error("f3d_mesh_metrics: the padded block needs at least 7 nodes per axis");
    return F3D_ERR_ARG;
  }
  hipLaunchKernelGGL(k_metrics, dim3(nblocks(g.NP)), dim3(kBlock), 0, s, x, y, z, metrics[0],
                     metrics[1], metrics[2], metrics[3], metrics[4], metrics[5], metrics[6],
                     metrics[7], metrics[8], J, g);
  return hip_check(hipGetLastError(), "f3d_mesh_metrics");
}

}  // extern "C"
