// Reconstruction and splitting formulas shared by the kernel-granular
// launches (flux.hip) and the fused FP32 sweeps (sweep.hip): Steger-Warming
// splitting at one point (split.jl:17-89) and the hybrid UP7 / WENO5-Z /
// WENO7-Z / NND face value (schemes.jl:86-437).
#pragma once

#include "f3d_internal.h"

namespace f3d {

// a / b: IEEE division in FP64; in FP32 (the float flux path, whose
// divisors here are normal numbers: sound speeds, pressures, metric norms and
// WENO indicators floored by eps = 1e-26) the hardware reciprocal times a,
// within 1 ulp, instead of the division's denormal-safe expansion
template <typename R>
__device__ __forceinline__ R qdiv(R a, R b) {
  if constexpr (sizeof(R) == 4)
    return a * __builtin_amdgcn_rcpf(b);
  else
    return a / b;
}

// fluxSplit at one point, split.jl:17-89: F+ and F- of the five conserved
// fluxes from rho, u, v, w, p, e and the sweep's three metrics.
template <typename R>
__device__ inline void split_point(R rho, R u, R v, R w, R pr, R ei, R A1, R A2, R A3, R SWeps2,
                                   R* __restrict__ Fp, R* __restrict__ Fm) {
  const R gam = qdiv(pr, ei) + R(1);
  const R c = sqrt(qdiv(gam * pr, rho));
  R ss = sqrt(A1 * A1 + A2 * A2 + A3 * A3);
  const R E1 = A1 * u + A2 * v + A3 * w;
  const R E2 = E1 - c * ss;
  const R E3 = E1 + c * ss;
  ss = qdiv(R(1), ss);
  A1 *= ss;
  A2 *= ss;
  A3 *= ss;
  const R E1P = (E1 + sqrt(E1 * E1 + SWeps2)) * R(0.5);
  const R E2P = (E2 + sqrt(E2 * E2 + SWeps2)) * R(0.5);
  const R E3P = (E3 + sqrt(E3 * E3 + SWeps2)) * R(0.5);
  const R E1M = E1 - E1P, E2M = E2 - E2P, E3M = E3 - E3P;
  const R uc1 = u - c * A1, uc2 = u + c * A1;
  const R vc1 = v - c * A2, vc2 = v + c * A2;
  const R wc1 = w - c * A3, wc2 = w + c * A3;
  const R vvc1 = (uc1 * uc1 + vc1 * vc1 + wc1 * wc1) * R(0.5);
  const R vvc2 = (uc2 * uc2 + vc2 * vc2 + wc2 * wc2) * R(0.5);
  const R vv = (gam - R(1)) * (u * u + v * v + w * w);
  const R W2 = qdiv(R(3) - gam, R(2) * (gam - R(1))) * (c * c);
  const R tmp1 = qdiv(rho, R(2) * gam);
  const R tmp2 = R(2) * (gam - R(1));
  Fp[0] = tmp1 * (tmp2 * E1P + E2P + E3P);
  Fp[1] = tmp1 * (tmp2 * E1P * u + E2P * uc1 + E3P * uc2);
  Fp[2] = tmp1 * (tmp2 * E1P * v + E2P * vc1 + E3P * vc2);
  Fp[3] = tmp1 * (tmp2 * E1P * w + E2P * wc1 + E3P * wc2);
  Fp[4] = tmp1 * (E1P * vv + E2P * vvc1 + E3P * vvc2 + W2 * (E2P + E3P));
  Fm[0] = tmp1 * (tmp2 * E1M + E2M + E3M);
  Fm[1] = tmp1 * (tmp2 * E1M * u + E2M * uc1 + E3M * uc2);
  Fm[2] = tmp1 * (tmp2 * E1M * v + E2M * vc1 + E3M * vc2);
  Fm[3] = tmp1 * (tmp2 * E1M * w + E2M * wc1 + E3M * wc2);
  Fm[4] = tmp1 * (E1M * vv + E2M * vvc1 + E3M * vvc2 + W2 * (E2M + E3M));
}

// WENO5-Z weights alpha_k = d_k (1 + (tau / (eps + beta_k))^2), d = (1, 6, 3)
// (schemes.jl:152-155).  FP64 evaluates the reference expression as written.
// In FP32, eps = 1e-26 makes (tau/(eps+beta_k))^2 overflow float whenever a
// beta_k is ~0 (three equal stencil values) and tau > ~2e-7; the weights are
// then computed as alpha_k * u with u = 1/(1 + T^2), T = tau/min_k q_k,
// q_k = eps + beta_k:  alpha_k u = d_k (u + (1 - u) s_k^2), s_k = min q / q_k.
// Every term is in [0, d_k] and the ratios alpha_k / sum alpha are unchanged.
template <typename R>
__device__ inline void zweights(R s11, R s22, R s33, R e, R& w1, R& w2, R& w3) {
  const R tau = fabs(s11 - s33);
  if constexpr (sizeof(R) == 8) {
    R t = tau / (e + s11);
    w1 = (R(1) + t * t) * R(1);
    t = tau / (e + s22);
    w2 = (R(1) + t * t) * R(6);
    t = tau / (e + s33);
    w3 = (R(1) + t * t) * R(3);
  } else {
    const R q1 = e + s11, q2 = e + s22, q3 = e + s33;
    const R qm = fmin(q1, fmin(q2, q3));
    const R T = qdiv(tau, qm);
    const R u = qdiv(R(1), R(1) + T * T);
    const R v = R(1) - u;
    const R a = qdiv(qm, q1), b = qdiv(qm, q2), c = qdiv(qm, q3);
    w1 = (u + v * (a * a)) * R(1);
    w2 = (u + v * (b * b)) * R(6);
    w3 = (u + v * (c * c)) * R(3);
  }
}

// WENO7-Z face value from v[0..6] = v_{j-3..j+3} of the upwind cell j (no
// counterpart in the reference, whose middle branch is WENO5-Z): Balsara &
// Shu (2000) candidates and smoothness indicators, linear weights
// (1, 12, 18, 4)/35, Z weights d_k (1 + (tau7/(eps + beta_k))^2) with
// tau7 = |b0 + 3 b1 - 3 b2 - b3| (Castro, Costa & Don 2011); FP32 takes the
// overflow-safe form of zweights.
template <typename R>
__device__ inline R weno7z(const R* v, R eps) {
  // no FMA contraction: the smoothness indicators are long sums of products
  // whose rounding the (tau / beta)^2 weights amplify near a jump; kept in the
  // oracle's operation order so the FP64 path matches it to 1e-12
#pragma clang fp contract(off)
  const R a = v[0], b = v[1], c = v[2], d = v[3], e = v[4], f = v[5], g = v[6];
  // indicators as quadratic forms in first differences (constants drop out
  // exactly: no cancellation of the values' magnitude, which in float would
  // swamp the indicators of a smooth stencil)
  const R D0 = b - a, D1 = c - b, D2 = d - c, D3 = e - d, D4 = f - e, D5 = g - f;
  const R r = R(1.0 / 240.0);
  const R b0 = r * (D0 * (R(547) * D0 - R(2788) * D1 + R(1854) * D2) +
                    D1 * (R(3708) * D1 - R(5188) * D2) + R(2107) * D2 * D2);
  const R b1 = r * (D1 * (R(267) * D1 - R(1108) * D2 + R(494) * D3) +
                    D2 * (R(1468) * D2 - R(1428) * D3) + R(547) * D3 * D3);
  const R b2 = r * (D2 * (R(547) * D2 - R(1428) * D3 + R(494) * D4) +
                    D3 * (R(1468) * D3 - R(1108) * D4) + R(267) * D4 * D4);
  const R b3 = r * (D3 * (R(2107) * D3 - R(5188) * D4 + R(1854) * D5) +
                    D4 * (R(3708) * D4 - R(2788) * D5) + R(547) * D5 * D5);
  const R tau = fabs(b0 + R(3) * b1 - R(3) * b2 - b3);
  R w0, w1, w2, w3;
  if constexpr (sizeof(R) == 8) {
    R t = tau / (eps + b0);
    w0 = (R(1) + t * t) * R(1);
    t = tau / (eps + b1);
    w1 = (R(1) + t * t) * R(12);
    t = tau / (eps + b2);
    w2 = (R(1) + t * t) * R(18);
    t = tau / (eps + b3);
    w3 = (R(1) + t * t) * R(4);
  } else {
    const R q0 = eps + b0, q1 = eps + b1, q2 = eps + b2, q3 = eps + b3;
    const R qm = fmin(fmin(q0, q1), fmin(q2, q3));
    const R T = qdiv(tau, qm);
    const R u = qdiv(R(1), R(1) + T * T);
    const R vv = R(1) - u;
    const R s0 = qdiv(qm, q0), s1 = qdiv(qm, q1), s2 = qdiv(qm, q2), s3 = qdiv(qm, q3);
    w0 = (u + vv * (s0 * s0)) * R(1);
    w1 = (u + vv * (s1 * s1)) * R(12);
    w2 = (u + vv * (s2 * s2)) * R(18);
    w3 = (u + vv * (s3 * s3)) * R(4);
  }
  const R c0 = -R(3) * a + R(13) * b - R(23) * c + R(25) * d;
  const R c1 = b - R(5) * c + R(13) * d + R(3) * e;
  const R c2 = -c + R(7) * d + R(7) * e - f;
  const R c3 = R(3) * d + R(13) * e - R(5) * f + g;
  return qdiv(w0 * c0 + w1 * c1 + w2 * c2 + w3 * c3, R(12) * (w0 + w1 + w2 + w3));
}

template <typename R>
__device__ inline R minmod(R a, R b) {
  // schemes.jl:26-28
  return (a * b > R(0)) ? ((fabs(a) > fabs(b)) ? b : a) : R(0);
}

// Branch of the hybrid reconstruction at a face (schemes.jl:96-100):
// 0 upwind7, 1 WENO5-Z, 2 NND.
template <typename R>
__device__ inline int weno_branch(R phif, const DevParams* __restrict__ P) {
  return (phif < PT<R>::phi1(P)) ? 0 : (phif < PT<R>::phi2(P)) ? 1 : 2;
}

// WENO_x / WENO_y / WENO_z face value, schemes.jl:86-437, on the eight
// values v[k] = f[ic + (k - 4) s] of Fp (p) and Fm (m) around the face i-1/2
// (ic its right cell, s the sweep stride).
// BR: the branch as a compile-time constant (-1: read br at run time), so
// a caller evaluating several variables at one face tests the branch once.
template <typename R, int BR = -1>
__device__ inline R weno_face_v(const R* p, const R* m, int br, const DevParams* __restrict__ P) {
  if (BR >= 0) br = BR;
  const R* c = PT<R>::up7(P);
  if (br == 0) {
    const R fpv = c[0] * p[0] + c[1] * p[1] + c[2] * p[2] + c[3] * p[3] + c[4] * p[4] +
                  c[5] * p[5] + c[6] * p[6];
    const R fmv = c[6] * m[1] + c[5] * m[2] + c[4] * m[3] + c[3] * m[4] + c[2] * m[5] +
                  c[1] * m[6] + c[0] * m[7];
    return fpv + fmv;
  } else if (br == 1 && P->recon_kind == 1) {
    // WENO7-Z: F+ from the upwind cell ic-1 (p[0..6]), F- mirrored (m[7..1])
    const R mr[7] = {m[7], m[6], m[5], m[4], m[3], m[2], m[1]};
    return weno7z<R>(p, PT<R>::weps(P)) + weno7z<R>(mr, PT<R>::weps(P));
  } else if (br == 1) {
    const R tmp1 = R(13.0 / 12.0), tmp2 = R(1.0 / 6.0);
    const R e = PT<R>::weps(P);
    R V1 = p[1], V2 = p[2], V3 = p[3], V4 = p[4], V5 = p[5];
    R t = V1 - R(2) * V2 + V3, t2 = V1 - R(4) * V2 + R(3) * V3;
    R s11 = tmp1 * (t * t) + R(0.25) * (t2 * t2);
    t = V2 - R(2) * V3 + V4;
    t2 = V2 - V4;
    R s22 = tmp1 * (t * t) + R(0.25) * (t2 * t2);
    t = V3 - R(2) * V4 + V5;
    t2 = R(3) * V3 - R(4) * V4 + V5;
    R s33 = tmp1 * (t * t) + R(0.25) * (t2 * t2);
    zweights<R>(s11, s22, s33, e, s11, s22, s33);
    R invsum = qdiv(R(1), s11 + s22 + s33);
    R v1 = R(2) * V1 - R(7) * V2 + R(11) * V3;
    R v2 = -V2 + R(5) * V3 + R(2) * V4;
    R v3 = R(2) * V3 + R(5) * V4 - V5;
    const R fpv = tmp2 * invsum * (s11 * v1 + s22 * v2 + s33 * v3);

    V1 = m[2];
    V2 = m[3];
    V3 = m[4];
    V4 = m[5];
    V5 = m[6];
    t = V5 - R(2) * V4 + V3;
    t2 = V5 - R(4) * V4 + R(3) * V3;
    s11 = tmp1 * (t * t) + R(0.25) * (t2 * t2);
    t = V2 - R(2) * V3 + V4;
    t2 = V4 - V2;
    s22 = tmp1 * (t * t) + R(0.25) * (t2 * t2);
    t = V3 - R(2) * V2 + V1;
    t2 = R(3) * V3 - R(4) * V2 + V1;
    s33 = tmp1 * (t * t) + R(0.25) * (t2 * t2);
    zweights<R>(s11, s22, s33, e, s11, s22, s33);
    invsum = qdiv(R(1), s11 + s22 + s33);
    v1 = R(11) * V3 - R(7) * V4 + R(2) * V5;
    v2 = -V4 + R(5) * V3 + R(2) * V2;
    v3 = R(2) * V3 + R(5) * V2 - V1;
    const R fmv = tmp2 * invsum * (s11 * v1 + s22 * v2 + s33 * v3);
    return fpv + fmv;
  } else {
    const R fpv = p[3] + R(0.5) * minmod(p[4] - p[3], p[3] - p[2]);
    const R fmv = m[4] - R(0.5) * minmod(m[5] - m[4], m[4] - m[3]);
    return fpv + fmv;
  }
}

// the hybrid face values of NV variables at one face: the branch is
// tested once for all of them (get(n, pv, mv) gathers variable n's stencil)
template <typename R, int NV, class G>
__device__ inline void weno_faces_v(int br, const DevParams* __restrict__ P, G&& get, R* out) {
  if (br == 0) {
#pragma unroll
    for (int n = 0; n < NV; ++n) {
      R pv[8], mv[8];
      get(n, pv, mv);
      out[n] = weno_face_v<R, 0>(pv, mv, 0, P);
    }
  } else if (br == 1) {
#pragma unroll
    for (int n = 0; n < NV; ++n) {
      R pv[8], mv[8];
      get(n, pv, mv);
      out[n] = weno_face_v<R, 1>(pv, mv, 1, P);
    }
  } else {
#pragma unroll
    for (int n = 0; n < NV; ++n) {
      R pv[8], mv[8];
      get(n, pv, mv);
      out[n] = weno_face_v<R, 2>(pv, mv, 2, P);
    }
  }
}

}  // namespace f3d
