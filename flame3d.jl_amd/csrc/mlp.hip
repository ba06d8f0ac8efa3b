// Learned chemistry surrogate: an (Ns+2) -> 64 -> 64 -> 64 -> Ns MLP with GELU
// (tanh form), BF16 weights and activations on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16) with FP32 accumulation, trained on the GPU by the
// kernels below (forward, backward, Adam) and evaluated in a fused chemistry
// kernel that gates each cell (Mahalanobis distance of its inputs, element
// balance of the prediction) and hands rejected cells to the FP64 Newton
// integrator.  The reference pattern is reactions.jl:2-72 (pre_input,
// evalModel, post_predict); the network shape and data follow the 128^3
// H2/O2 chemistry workload (Ns = 9); the same kernels, with inputs padded to
// 32 and outputs to 32, serve 5-species air (Ns = 5).
//
// MFMA operand layout of v_mfma_f32_16x16x32_bf16 (verified on the device):
// lane l holds A[l%16][8*(l/16) + i] and B[8*(l/16) + i][l%16] (i = 0..7) and
// receives C[4*(l/16) + i][l%16] (i = 0..3).  Activations are staged in LDS in
// both [sample][feature] and [feature][sample] order so that every operand is
// a contiguous 16-byte read.
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>

#include "f3d_chem.h"
#include "f3d_internal.h"
#include "f3d_nn.h"
#include "f3d_mech_tables.h"

namespace f3d {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline f32x4 mfma(const bf16x8& a, const bf16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ inline bf16x8 ld8(const __bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

// GELU, tanh form: 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))
__device__ inline float gelu(float x, float& dg) {
  const float c = 0.7978845608028654f, k = 0.044715f;
  const float u = c * (x + k * x * x * x);
  // tanh(u) = 1 - 2 / (1 + e^2u), with the hardware reciprocal (1 ulp; the
  // activations are rounded to bf16 next): saturates cleanly at +-1
  const float t = 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + __expf(2.0f * u));
  dg = 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * c * (1.0f + 3.0f * k * x * x);
  return 0.5f * x * (1.0f + t);
}

// ---------------------------------------------------------------------------
// Training tile: 64 samples per workgroup, 4 waves (wave w owns samples
// 16w..16w+15 on the data path).  Writes this tile's weight/bias gradient of
// the loss L = sum (y - t)^2 / (B * NOUT) and its partial loss to
// partial[blockIdx.x][0 .. NPARAM].
// ---------------------------------------------------------------------------
struct TileLds {
  __bf16 X[64][NN_INP];    // [s][k], forward operand of layer 0
  __bf16 Xt[NN_INP][64];   // [k][s], weight-gradient operand of layer 0
  __bf16 At[3][NN_H][64];  // hidden activations [o][s] (weight-gradient operands)
  union {
    __bf16 A[2][64][NN_H];  // forward ping-pong [s][o]
    struct {
      __bf16 dZ[64][NN_H];   // backward: dZ [s][o] (data-gradient operand)
      __bf16 dZt[NN_H][64];  // dZ^T [o][s] (weight-gradient operand)
    } g;
  } u;
  float bsum[4][NN_H];  // per-wave bias-gradient partial sums
  float red[256];
};

// sum over the four 16-lane groups holding the same column (lanes r, r+16, r+32, r+48)
__device__ inline float sum_q(float v) {
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 32);
  return v;
}

__global__ void __launch_bounds__(256)
    k_nn_train_tile(const float* __restrict__ Xg, const float* __restrict__ Tg, int B, int nin,
                    int nout, const __bf16* __restrict__ Wb, const __bf16* __restrict__ Wtb,
                    const float* __restrict__ master, float* __restrict__ partial) {
  __shared__ __align__(16) TileLds L;
  const int tid = threadIdx.x;
  const int w = tid >> 6, lane = tid & 63;
  const int r16 = lane & 15, q = lane >> 4;
  const long s0 = (long)blockIdx.x * 64;
  for (int e = tid; e < 64 * NN_INP; e += 256) {
    const int s = e / NN_INP, k = e % NN_INP;
    const float v = (k < nin) ? Xg[(s0 + s) * nin + k] : 0.0f;
    L.X[s][k] = (__bf16)v;
    L.Xt[k][s] = (__bf16)v;
  }
  __syncthreads();
  const int sw = 16 * w;  // this wave's first sample on the data path
  float gp[3][4][4];      // GELU' of the wave's hidden pre-activations
  // ---- forward, hidden layers ----
  for (int l = 0; l < 3; ++l) {
    const int NI = nn_in_dim(l);
    const __bf16* W = Wb + nn_w_off(l);
    const float* bias = master + nn_b_off(l);
#pragma unroll
    for (int ot = 0; ot < NN_H / 16; ++ot) {
      f32x4 acc = {0, 0, 0, 0};
      for (int kk = 0; kk < NI / 32; ++kk) {
        const bf16x8 a = (l == 0) ? ld8(&L.X[sw + r16][32 * kk + 8 * q])
                                  : ld8(&L.u.A[(l - 1) & 1][sw + r16][32 * kk + 8 * q]);
        const bf16x8 b = ld8(W + (16 * ot + r16) * NI + 32 * kk + 8 * q);
        acc = mfma(a, b, acc);
      }
      const int o = 16 * ot + r16;
      const float bo = bias[o];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int s = sw + 4 * q + i;
        float d;
        const __bf16 a = (__bf16)gelu(acc[i] + bo, d);
        gp[l][ot][i] = d;
        L.u.A[l & 1][s][o] = a;
        L.At[l][o][s] = a;
      }
    }
    __syncthreads();
  }
  // ---- output layer and dL/dy (the output layer is linear: dZ4 = dL/dy) ----
  float lsum = 0.0f;
  f32x4 yv[NN_OUTP / 16];
  {
    const __bf16* W = Wb + nn_w_off(3);
#pragma unroll
    for (int ot = 0; ot < NN_OUTP / 16; ++ot) {
      f32x4 acc = {0, 0, 0, 0};
      for (int kk = 0; kk < NN_H / 32; ++kk) {
        const bf16x8 a = ld8(&L.u.A[0][sw + r16][32 * kk + 8 * q]);
        const bf16x8 b = ld8(W + (16 * ot + r16) * NN_H + 32 * kk + 8 * q);
        acc = mfma(a, b, acc);
      }
      yv[ot] = acc;
    }
  }
  __syncthreads();  // A[0] (layer-2 output) fully read before dZ overwrites the union
  {
    const float* bias = master + nn_b_off(3);
    const float scale = 2.0f / ((float)B * (float)nout);
#pragma unroll
    for (int ot = 0; ot < NN_OUTP / 16; ++ot) {
      const int o = 16 * ot + r16;
      float bs = 0.0f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int s = sw + 4 * q + i;
        float g = 0.0f;
        if (o < nout) {
          const float d = yv[ot][i] + bias[o] - Tg[(s0 + s) * nout + o];
          lsum += d * d;
          g = scale * d;
        }
        bs += g;
        L.u.g.dZ[s][o] = (__bf16)g;
        L.u.g.dZt[o][s] = (__bf16)g;
      }
      bs = sum_q(bs);
      if (q == 0) L.bsum[w][o] = bs;
    }
  }
  __syncthreads();
  float* part = partial + (long)blockIdx.x * (NN_NPARAM + 1);
  // ---- backward, layers 3..0 ----
  for (int l = 3; l >= 0; --l) {
    const int NI = nn_in_dim(l), NO = nn_out_dim(l);
    for (int o = tid; o < NO; o += 256)
      part[nn_b_off(l) + o] = L.bsum[0][o] + L.bsum[1][o] + L.bsum[2][o] + L.bsum[3][o];
    // weight gradient dW[o][k] = sum_s dZ[s][o] A_prev[s][k], output tiles over the waves
    const int nto = NO / 16, ntk = NI / 16;
    for (int t = w; t < nto * ntk; t += 4) {
      const int ot = t / ntk, kt = t % ntk;
      f32x4 acc = {0, 0, 0, 0};
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) {
        const bf16x8 a = ld8(&L.u.g.dZt[16 * ot + r16][32 * ks + 8 * q]);
        const bf16x8 b = (l == 0) ? ld8(&L.Xt[16 * kt + r16][32 * ks + 8 * q])
                                  : ld8(&L.At[l - 1][16 * kt + r16][32 * ks + 8 * q]);
        acc = mfma(a, b, acc);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
        part[nn_w_off(l) + (16 * ot + 4 * q + i) * NI + 16 * kt + r16] = acc[i];
    }
    if (l == 0) break;
    // data gradient dA_prev[s][k] = sum_o dZ[s][o] W[o][k]  (W^T copy: [k][o])
    f32x4 dA[4];
    const __bf16* Wt = Wtb + nn_w_off(l);
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) {
      f32x4 acc = {0, 0, 0, 0};
      for (int ks = 0; ks < NO / 32; ++ks) {
        const bf16x8 a = ld8(&L.u.g.dZ[sw + r16][32 * ks + 8 * q]);
        const bf16x8 b = ld8(Wt + (16 * kt + r16) * NO + 32 * ks + 8 * q);
        acc = mfma(a, b, acc);
      }
      dA[kt] = acc;
    }
    __syncthreads();  // all reads of dZ, dZ^T and bsum of layer l are done
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) {
      const int k = 16 * kt + r16;
      float bs = 0.0f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int s = sw + 4 * q + i;
        const float g = dA[kt][i] * gp[l - 1][kt][i];
        bs += g;
        L.u.g.dZ[s][k] = (__bf16)g;
        L.u.g.dZt[k][s] = (__bf16)g;
      }
      bs = sum_q(bs);
      if (q == 0) L.bsum[w][k] = bs;
    }
    __syncthreads();
  }
  L.red[tid] = lsum;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st) L.red[tid] += L.red[tid + st];
    __syncthreads();
  }
  if (tid == 0) part[NN_NPARAM] = L.red[0] / ((float)B * (float)nout);
}

// Sum of the per-tile partials (gradient and loss), one thread per slot.
__global__ void __launch_bounds__(256)
    k_nn_reduce(const float* __restrict__ partial, int ntiles, float* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > NN_NPARAM) return;
  float s = 0.0f;
  for (int t = 0; t < ntiles; ++t) s += partial[(long)t * (NN_NPARAM + 1) + p];
  out[p] = s;
}

// Adam (Kingma & Ba 2015) on the FP32 master weights; refreshes the BF16
// copies (row-major and transposed).  Padded weights stay zero: their
// gradient is zero and they start at zero.
__global__ void __launch_bounds__(256)
    k_nn_adam(const float* __restrict__ partial, int ntiles, float* __restrict__ master,
              float* __restrict__ m1, float* __restrict__ m2, __bf16* __restrict__ Wb,
              __bf16* __restrict__ Wtb, float lr, float b1, float b2, float eps, float bc1,
              float bc2, float* __restrict__ loss_out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > NN_NPARAM) return;
  float g = 0.0f;
  for (int t = 0; t < ntiles; ++t) g += partial[(long)t * (NN_NPARAM + 1) + p];
  if (p == NN_NPARAM) {
    if (loss_out) *loss_out = g;
    return;
  }
  const float a = b1 * m1[p] + (1.0f - b1) * g;
  const float v = b2 * m2[p] + (1.0f - b2) * g * g;
  m1[p] = a;
  m2[p] = v;
  const float wv = master[p] - lr * (a / bc1) / (sqrtf(v / bc2) + eps);
  master[p] = wv;
  int l, o, k;
  if (nn_locate_weight(p, l, o, k)) {
    Wb[p] = (__bf16)wv;
    Wtb[nn_w_off(l) + k * nn_out_dim(l) + o] = (__bf16)wv;
  }
}

// refresh the BF16 copies from the FP32 master (after f3d_nn_set_params)
__global__ void __launch_bounds__(256)
    k_nn_cast(const float* __restrict__ master, __bf16* __restrict__ Wb, __bf16* __restrict__ Wtb) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NN_NPARAM) return;
  int l, o, k;
  if (nn_locate_weight(p, l, o, k)) {
    Wb[p] = (__bf16)master[p];
    Wtb[nn_w_off(l) + k * nn_out_dim(l) + o] = (__bf16)master[p];
  }
}

// ---------------------------------------------------------------------------
// Forward of 64 samples by one wave (inference).  One 8 KB LDS buffer per
// wave holds, in turn, the BF16 input tile, each hidden activation tile and
// the FP32 outputs: a layer's 4x4 (or 4x2) 16x16 output tiles stay in
// registers until every lane has read the previous activations.
// ---------------------------------------------------------------------------
// Rows are padded by 16 bytes (one float for Y): a 16-byte MFMA operand read
// by the 16 lanes of a quarter-wave then spans different bank groups instead
// of the same one (an unpadded row of 64 bf16 is 128 B, a whole bank cycle).
constexpr int kRowPad = 8;
union FwdLds {
  __bf16 X[64][NN_INP + kRowPad];
  __bf16 A[64][NN_H + kRowPad];
  float Y[64][NN_OUTP + 1];
};

// one dense layer for the wave's 64 samples: NI inputs, NO outputs (compile-time)
template <int NI, int NO, bool LAST>
__device__ __forceinline__ void nn_layer_wave(FwdLds& L, const __bf16* __restrict__ W,
                                     const float* __restrict__ bias, int lane) {
  const int r16 = lane & 15, q = lane >> 4;
  f32x4 acc[4][NO / 16];
#pragma unroll
  for (int ot = 0; ot < NO / 16; ++ot) {
#pragma unroll
    for (int st = 0; st < 4; ++st) acc[st][ot] = {0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < NI / 32; ++kk) {
      const bf16x8 b = ld8(W + (16 * ot + r16) * NI + 32 * kk + 8 * q);
#pragma unroll
      for (int st = 0; st < 4; ++st) {
        const __bf16* src = (NI == NN_INP) ? &L.X[16 * st + r16][32 * kk + 8 * q]
                                           : &L.A[16 * st + r16][32 * kk + 8 * q];
        acc[st][ot] = mfma(ld8(src), b, acc[st][ot]);
      }
    }
  }
  __syncthreads();  // every lane has read this layer's input tile
#pragma unroll
  for (int ot = 0; ot < NO / 16; ++ot) {
    const int o = 16 * ot + r16;
    const float bo = bias[o];
#pragma unroll
    for (int st = 0; st < 4; ++st)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int s = 16 * st + 4 * q + i;
        if (LAST) {
          L.Y[s][o] = acc[st][ot][i] + bo;
        } else {
          float d;
          L.A[s][o] = (__bf16)gelu(acc[st][ot][i] + bo, d);
        }
      }
  }
  __syncthreads();
}

__device__ __forceinline__ void nn_forward_wave(FwdLds& L, const __bf16* __restrict__ Wb,
                                       const float* __restrict__ master, int lane) {
  nn_layer_wave<NN_INP, NN_H, false>(L, Wb + nn_w_off(0), master + nn_b_off(0), lane);
  nn_layer_wave<NN_H, NN_H, false>(L, Wb + nn_w_off(1), master + nn_b_off(1), lane);
  nn_layer_wave<NN_H, NN_H, false>(L, Wb + nn_w_off(2), master + nn_b_off(2), lane);
  nn_layer_wave<NN_H, NN_OUTP, true>(L, Wb + nn_w_off(3), master + nn_b_off(3), lane);
}

// plain forward on normalized inputs X[B][nin] -> Y[B][nout] (B % 64 == 0)
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4)))
    k_nn_forward(const float* __restrict__ Xg, float* __restrict__ Yg, int nin, int nout,
                 const __bf16* __restrict__ Wb, const float* __restrict__ master) {
  __shared__ __align__(16) FwdLds L;
  const int lane = threadIdx.x;
  const long s0 = (long)blockIdx.x * 64;
  for (int k = 0; k < NN_INP; ++k) L.X[lane][k] = (__bf16)(k < nin ? Xg[(s0 + lane) * nin + k] : 0.0f);
  __syncthreads();
  nn_forward_wave(L, Wb, master, lane);
  for (int o = 0; o < nout; ++o) Yg[(s0 + lane) * nout + o] = L.Y[lane][o];
}

// Element-conserving projection with per-cell weights S:
// dY = dY0 - S E^T (E S E^T)^-1 E dY0, E[e][k] = atoms_ek / W_k, solved as a
// 4x4 system (rows past n_elem are identity) by Gaussian elimination with
// partial pivoting in registers (compile-time indices, row swaps by selects).
// Returns false if E S E^T is singular (dY = dY0 then).
template <int NS>
__device__ __forceinline__ bool project_weighted(const NNDev* __restrict__ C,
                                                 const DevParams* __restrict__ P,
                                                 const double (&S)[NS], const double (&dY0)[NS],
                                                 double (&dY)[NS]) {
  const int ne = C->n_elem;
  double G[4][4], b[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    b[a] = 0.0;
#pragma unroll
    for (int c = 0; c < 4; ++c) G[a][c] = (a == c && a >= ne) ? 1.0 : 0.0;
  }
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    double e[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) e[a] = a < ne ? C->elem_mw[a][k] : 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      b[a] += e[a] * dY0[k];
#pragma unroll
      for (int c = 0; c < 4; ++c) G[a][c] += e[a] * S[k] * e[c];
    }
  }
  bool ok = true;
#pragma unroll
  for (int col = 0; col < 4; ++col) {
    int piv = col;
    double pv = fabs(G[col][col]);
#pragma unroll
    for (int a = col + 1; a < 4; ++a)
      if (fabs(G[a][col]) > pv) {
        pv = fabs(G[a][col]);
        piv = a;
      }
    if (!(pv > 0.0)) ok = false;
#pragma unroll
    for (int a = col + 1; a < 4; ++a) {
      const bool sw = (piv == a);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double t = G[col][c];
        G[col][c] = sw ? G[a][c] : t;
        G[a][c] = sw ? t : G[a][c];
      }
      const double t = b[col];
      b[col] = sw ? b[a] : t;
      b[a] = sw ? t : b[a];
    }
    const double inv = ok ? 1.0 / G[col][col] : 0.0;
#pragma unroll
    for (int a = col + 1; a < 4; ++a) {
      const double f = G[a][col] * inv;
#pragma unroll
      for (int c = col; c < 4; ++c) G[a][c] -= f * G[col][c];
      b[a] -= f * b[col];
    }
  }
  double z[4];
#pragma unroll
  for (int a = 3; a >= 0; --a) {
    double v = b[a];
#pragma unroll
    for (int c = a + 1; c < 4; ++c) v -= G[a][c] * z[c];
    z[a] = ok ? v / G[a][a] : 0.0;
  }
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    double v = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) v += (a < ne ? C->elem_mw[a][k] : 0.0) * z[a];
    dY[k] = dY0[k] - S[k] * v;
  }
  return ok;
}

// ---------------------------------------------------------------------------
// Surrogate chemistry step (pre_input -> network -> post_predict,
// reactions.jl:2-56) in three kernels: k_nn_net runs the network on the
// matrix cores, one wave per 64 interior cells, and stores its raw outputs;
// then the FP64 epilogue, one thread per cell: k_nn_pred (gates 2-4, the
// predicted state with the FP64 energy bookkeeping of the backward-Euler
// integrator and its temperature) and k_nn_gate (gates 5-7 on one rate
// evaluation), which either commits the prediction or marks its cell for
// the Newton pass and leaves it untouched (status = 1 + gate); the marked
// cells are then listed in cell order (k_nn_scan, k_nn_compact):
//   1. Mahalanobis distance of the normalized inputs <= d_max (out of
//      distribution; in k_nn_net, on the inputs it already holds);
//   2. element balance of the raw prediction, each element's net change against
//      the gross elemental turnover (then projected to conserve elements);
//   3. positivity, Y_k + dY_k >= -neg_tol;
//   4. predicted temperature change sum_k |h_f,k dY_k| / c_v <= dT_max (fast
//      chemistry, where a relative error of the network costs the most);
//   5. backward-Euler residual at the predicted state, max_k |R_k| / rho <= res_tol
//      and sum_k |h_f,k R_k| / (rho c_v) <= res_T_tol,
//      R_k = rho_k' - rho_k - dt W_k wdot_k(rho', T'), T' from the energy (error gate);
//   6. thermal explosiveness chi = dt (d HRR / dT) / (rho c_v) <= chi_max at the
//      predicted state: near chi ~ 1 the backward-Euler equation can have a
//      second (ignited) root that a small residual does not rule out;
//   7. chain explosiveness: every eigenvalue of dt d wdot / d c over the chain
//      carriers below chain_max (near dt lambda = 1 a hot radical-free mixture
//      has a near-root with a tiny residual while the Newton step ignites it).
// Between the network and the epilogue the outputs take 4 B per cell and
// output ([Ns][cells], 0.6 GB at 256^3 for Ns = 9, written and read once).
// ---------------------------------------------------------------------------
template <int NS>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4)))
    k_nn_net(const double* __restrict__ Q, const double* __restrict__ rhoi,
             const int64_t* __restrict__ tag, const DevParams* __restrict__ P,
             const NNDev* __restrict__ C, const __bf16* __restrict__ Wb,
             const float* __restrict__ master, Geom g, float* __restrict__ out,
             int8_t* __restrict__ status) {
  constexpr int NIN = NS + 2;  // T, log P, Box-Cox of each species
  __shared__ __align__(16) FwdLds L;
  const int lane = threadIdx.x;
  const long ncell = (long)g.Nx * g.Ny * g.Nz;
  const long id = (long)xcd_block() * 64 + lane;
  const long NP = g.NP;
  bool act = false;
  float xn[NIN];
#pragma unroll
  for (int k = 0; k < NIN; ++k) xn[k] = 0.0f;
  if (id < ncell) {
    const int a = (int)(id % g.Nx);
    const int t = (int)(id / g.Nx);
    const int b = t % g.Ny, c = t / g.Ny;
    const long p = (a + NG) + (long)g.Nxt * ((b + NG) + (long)g.Nyt * (c + NG));
    const double T = Q[p + 5 * NP];
    act = (!tag || tag[p] == 0) && (T > P->T_criteria);
    status[id] = 0;
    if (act) {
      // pre_input (reactions.jl:2-23): T, log P, Box-Cox(Y), z-score
      const double rinv = 1.0 / fmax(Q[p], kEps);
      const float lam = C->boxcox_lambda;
      const float ilam = __builtin_amdgcn_rcpf(lam);  // the inputs are rounded to bf16 next
      float x[NIN];
      x[0] = (float)T;
      x[1] = logf((float)Q[p + 4 * NP]);
#pragma unroll
      for (int n = 0; n < NS; ++n) {
        const float Y = (float)fmax(rhoi[p + n * NP] * rinv, (double)C->y_floor);
        x[2 + n] = (powf(Y, lam) - 1.0f) * ilam;
      }
#pragma unroll
      for (int k = 0; k < NIN; ++k)
        xn[k] = (x[k] - C->in_mean[k]) * __builtin_amdgcn_rcpf(C->in_std[k]);
    }
  }
  // a tile with no reacting cell (cold gas, tagged cells) skips the network;
  // the workgroup is one wave, so the exit is uniform over the workgroup
  if (!__any(act)) return;
#pragma unroll
  for (int k = 0; k < NN_INP; ++k) L.X[lane][k] = (__bf16)(k < NIN ? xn[k] : 0.0f);
  __syncthreads();
  nn_forward_wave(L, Wb, master, lane);
  if (act) {
    // 1. Mahalanobis distance of the normalized inputs
    float d2 = 0.0f;
#pragma unroll
    for (int i = 0; i < NIN; ++i) {
      float row = 0.0f;
#pragma unroll
      for (int j = 0; j < NIN; ++j) row += C->mahal_inv[i][j] * (xn[j] - C->mahal_mean[j]);
      d2 += (xn[i] - C->mahal_mean[i]) * row;
    }
    status[id] = (int8_t)((sqrtf(fmaxf(d2, 0.0f)) <= C->d_max) ? 1 : 2);
#pragma unroll
    for (int n = 0; n < NS; ++n) out[(long)n * ncell + id] = L.Y[lane][n];
  }
}

// The epilogue runs as two kernels so each keeps its registers at 4 waves
// per SIMD: k_nn_pred (gates 2-4, the predicted state and its temperature)
// stores the candidate state of every cell still accepted (rho_k', T', e',
// the energy change) to `cand` ([NS + 3][cells], FP64); k_nn_gate evaluates
// the rate-based gates 5-7 on it and commits or rejects.
template <int NS>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3)))
    k_nn_pred(const double* __restrict__ Q, const double* __restrict__ rhoi,
              const DevParams* __restrict__ P, const NNDev* __restrict__ C, Geom g,
              const float* __restrict__ out, double* __restrict__ cand,
              int8_t* __restrict__ status) {
  const long ncell = (long)g.Nx * g.Ny * g.Nz;
  const long id = gid();
  if (id >= ncell) return;
  int code = status[id];  // 0 inactive, 1 candidate, 2 Mahalanobis
  if (code != 1) return;
  const long NP = g.NP;
  const int a = (int)(id % g.Nx);
  const int t = (int)(id / g.Nx);
  const int b = t % g.Ny, c = t / g.Ny;
  const long p = (a + NG) + (long)g.Nxt * ((b + NG) + (long)g.Nyt * (c + NG));
  const double T = Q[p + 5 * NP];
  const double rho = Q[p];
  double r[NS];
#pragma unroll
  for (int n = 0; n < NS; ++n) r[n] = rhoi[p + n * NP];
  // post_predict (reactions.jl:26-56) on the de-normalized outputs
  const double rinv = 1.0 / fmax(rho, kEps);
  const double lam = (double)C->boxcox_lambda;
  double Y[NS], dY0[NS], dY[NS], S[NS];
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    Y[n] = fmax(r[n] * rinv, 0.0);
    const double o = (double)(out[(long)n * ncell + id] * C->out_std[n] + C->out_mean[n]);
    if (C->label_mode == 0) {
      dY0[n] = o;
    } else if (C->label_mode == 2) {
      dY0[n] = (double)C->label_scale[n] * sinh(o);
    } else if (C->label_mode == 3) {
      const double sk = fmin((double)C->label_scale[n],
                             fmax((double)C->label_floor, (double)C->label_rel * Y[n]));
      dY0[n] = sk * sinh(o);
      S[n] = (sk + fabs(dY0[n])) * (sk + fabs(dY0[n])) + 1e-60;
    } else {
      const double base = lam * ((pow(Y[n], lam) - 1.0) / lam + o) + 1.0;
      dY0[n] = (base > 0.0 ? pow(base, 1.0 / lam) : 0.0) - Y[n];
    }
  }
  // 2. element balance of the raw prediction (net change of each element
  //    against the gross turnover of all elements, sum_e sum_k a_ek |dY_k| / W_k,
  //    so an inert species' noise counts against the whole change), then
  //    the projection
  bool eok = true;
  {
    double gross = 0.0, worst = 0.0;
    for (int e = 0; e < C->n_elem; ++e) {
      double num = 0.0;
#pragma unroll
      for (int n = 0; n < NS; ++n) {
        const double ae = C->elem_mw[e][n];
        num += ae * dY0[n];
        gross += ae * fabs(dY0[n]);
      }
      worst = fmax(worst, fabs(num));
    }
    if (gross > 0.0 && worst > (double)C->elem_tol * gross) eok = false;
  }
  if (C->label_mode == 3) {
    // per-cell weights S_k = (s_k + |dY0_k|)^2: dY = dY0 - S E^T (E S E^T)^-1 E dY0
    if (!project_weighted<NS>(C, P, S, dY0, dY)) eok = false;
  } else {
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      double v = 0.0;
#pragma unroll
      for (int m = 0; m < NS; ++m) v += C->proj[n][m] * dY0[m];
      dY[n] = v;
    }
  }
  if (!eok) code = 3;
  bool pok = true;
#pragma unroll
  for (int n = 0; n < NS; ++n)  // 3. positivity
    if (Y[n] + dY[n] < -(double)C->neg_tol) pok = false;
  if (code == 1 && !pok) code = 4;
  // 4. fast chemistry: the temperature change the prediction implies
  if (code == 1 && C->dT_max > 0.0f) {
    double esum = 0.0;
#pragma unroll
    for (int n = 0; n < NS; ++n) esum += fabs(P->hfm[n] * dY[n]);
    if (!(esum * rho <= (double)C->dT_max * cv_mix<NS, true>(P, T, r))) code = 5;
  }
  if (code == 1) {
    double rn[NS], dei = 0.0, sr = 0.0;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      rn[n] = fmax(Y[n] + dY[n], 0.0) * rho;
      dei += -P->hfm[n] * (rn[n] - r[n]);
    }
    // the state chem_epilogue would write, evaluated before committing
    const double rhoc = fmax(rho, kEps);
#pragma unroll
    for (int n = 0; n < NS; ++n) sr += rn[n];
#pragma unroll
    for (int n = 0; n < NS; ++n) rn[n] *= rhoc / sr;
    const double ein = Q[p + 6 * NP] + dei;
    const double Tn = fmax(get_T_from<NS>(P, ein, rn, T), kEps);
#pragma unroll
    for (int n = 0; n < NS; ++n) cand[(long)n * ncell + id] = rn[n];
    cand[(long)NS * ncell + id] = Tn;
    cand[(long)(NS + 1) * ncell + id] = ein;
    cand[(long)(NS + 2) * ncell + id] = dei;
  }
  if (code != 1) status[id] = (int8_t)code;
}

template <int NS, class K = KinRT>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2)))
    k_nn_gate(double* __restrict__ U, double* __restrict__ Q, double* __restrict__ rhoi,
              const DevParams* __restrict__ P, const NNDev* __restrict__ C, Geom g, double dt,
              const double* __restrict__ cand, int32_t* __restrict__ tile_count,
              int8_t* __restrict__ status) {
  const long ncell = (long)g.Nx * g.Ny * g.Nz;
  const long id = gid();  // a wave holds the 64 cells of one tile
  const long NP = g.NP;
  int code = (id < ncell) ? status[id] : 0;
  if (code == 1) {
    const int a = (int)(id % g.Nx);
    const int t = (int)(id / g.Nx);
    const int b = t % g.Ny, c = t / g.Ny;
    const long p = (a + NG) + (long)g.Nxt * ((b + NG) + (long)g.Nyt * (c + NG));
    double rn[NS];
#pragma unroll
    for (int n = 0; n < NS; ++n) rn[n] = cand[(long)n * ncell + id];
    const double Tn = cand[(long)NS * ncell + id];
    const double rhoc = fmax(Q[p], kEps);
    // 5. backward-Euler residual at (rho', T'): its largest species entry,
    //    and the temperature error its formation energies imply
    const bool res_on = C->res_tol > 0.0f || C->res_T_tol > 0.0f;
    if (res_on || C->chi_max > 0.0f || C->chain_max > 0.0f) {
      double sc[NS], wdot[NS], Adum[NS][NS], Mc[3][3];
#pragma unroll
      for (int n = 0; n < NS; ++n) sc[n] = rn[n] * P->cmw[n];
      if (C->chain_max > 0.0f) {
        const int cidx[3] = {C->chain_idx[0], C->chain_idx[1], C->chain_idx[2]};
        production_rate_block3<NS, K>(P, sc, Tn, cidx, wdot, Mc);
      } else {
        production_rate<NS, false>(P, sc, Tn, Adum, wdot);
      }
      const double cvn = cv_mix<NS, true>(P, Tn, rn);
      if (res_on) {
        double rmax = 0.0, esum = 0.0;
#pragma unroll
        for (int n = 0; n < NS; ++n) {
          const double R = rn[n] - rhoi[p + n * NP] - wdot[n] * P->mw[n] * 1e6 * dt;
          rmax = fmax(rmax, fabs(R));
          esum += fabs(P->hfm[n] * R);
        }
        if (C->res_tol > 0.0f && !(rmax <= (double)C->res_tol * rhoc)) code = 6;
        if (C->res_T_tol > 0.0f && !(esum <= (double)C->res_T_tol * cvn)) code = 6;
      }
      // 6. thermal explosiveness: the heat-release response to T over one step
      if (code == 1 && C->chi_max > 0.0f) {
        double w2[NS];
        const double hT = 1e-4 * Tn;
        production_rate<NS, false>(P, sc, Tn + hT, Adum, w2);
        double dq = 0.0;
#pragma unroll
        for (int n = 0; n < NS; ++n)
          dq -= P->lo[n][5] * P->Ru * 1e6 * (w2[n] - wdot[n]);  // h_f,k W_k d(wdot_k)
        if (!(dt * dq / hT <= (double)C->chi_max * cvn)) code = 7;
      }
      // 7. chain explosiveness: the 3x3 block M = dt d wdot / d c over the
      //    chain carriers (from production_rate_block3 above), rejected
      //    unless every eigenvalue of M has real part < chain_max:
      //    Routh-Hurwitz on N = chain_max I - M (tr N > 0, det N > 0,
      //    tr N * m2(N) > det N, m2 = sum of the principal 2x2 minors)
      if (code == 1 && C->chain_max > 0.0f) {
        double Nm[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j)
            Nm[i][j] = (i == j ? (double)C->chain_max : 0.0) - dt * Mc[i][j];
        const double tr = Nm[0][0] + Nm[1][1] + Nm[2][2];
        const double m2 = Nm[0][0] * Nm[1][1] - Nm[0][1] * Nm[1][0] +
                          Nm[0][0] * Nm[2][2] - Nm[0][2] * Nm[2][0] +
                          Nm[1][1] * Nm[2][2] - Nm[1][2] * Nm[2][1];
        const double det = Nm[0][0] * (Nm[1][1] * Nm[2][2] - Nm[1][2] * Nm[2][1]) -
                           Nm[0][1] * (Nm[1][0] * Nm[2][2] - Nm[1][2] * Nm[2][0]) +
                           Nm[0][2] * (Nm[1][0] * Nm[2][1] - Nm[1][1] * Nm[2][0]);
        if (!(tr > 0.0 && det > 0.0 && tr * m2 > det)) code = 8;
      }
    }
    if (code == 1) {
      U[p + 4 * NP] += cand[(long)(NS + 2) * ncell + id];
      Q[p + 4 * NP] = fmax(pmixture<NS, true>(P, Tn, rn), kEps);
      Q[p + 5 * NP] = Tn;
      Q[p + 6 * NP] = cand[(long)(NS + 1) * ncell + id];
#pragma unroll
      for (int n = 0; n < NS; ++n) rhoi[p + n * NP] = rn[n];
    } else {
      status[id] = (int8_t)code;
    }
  }
  // rejected cells of this 64-cell tile (status >= 2); the ordered list the
  // Newton pass runs over is built from these counts (k_nn_scan, k_nn_compact)
  const unsigned long long bal = __ballot(code >= 2);
  if ((threadIdx.x & 63) == 0 && id < ncell) tile_count[id >> 6] = (int32_t)__popcll(bal);
}

// Exclusive scan of n counts in place (one workgroup of 1024 threads, each
// over a contiguous run of entries); *total = their sum.
__global__ void __launch_bounds__(1024) k_nn_scan(int32_t* __restrict__ cnt, long n,
                                                   int32_t* __restrict__ total) {
  __shared__ int32_t part[1024];
  const int t = threadIdx.x;
  const long chunk = (n + 1023) / 1024;
  const long b = t * chunk, e = min(b + chunk, n);
  int32_t s = 0;
  for (long i = b; i < e; ++i) s += cnt[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan
    const int32_t v = (t >= off) ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int32_t run = part[t] - s;  // exclusive prefix of this chunk
  for (long i = b; i < e; ++i) {
    const int32_t c = cnt[i];
    cnt[i] = run;
    run += c;
  }
  if (t == 1023) *total = part[1023];
}

// The tile counts are scanned in chunks of kScanChunk tiles: k_scan_sums
// writes each chunk's sum, k_nn_scan scans the chunk sums (one workgroup),
// k_scan_apply turns each chunk into exclusive offsets from its chunk's
// offset.  Three short launches in place of one workgroup walking all tiles.
constexpr int kScanThreads = 256, kScanPer = 4, kScanChunk = kScanThreads * kScanPer;

__global__ void __launch_bounds__(kScanThreads)
    k_scan_sums(const int32_t* __restrict__ cnt, long n, int32_t* __restrict__ sums) {
  __shared__ int32_t part[kScanThreads];
  const int t = threadIdx.x;
  const long b = (long)blockIdx.x * kScanChunk + (long)t * kScanPer;
  int32_t v = 0;
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) v += (b + i < n) ? cnt[b + i] : 0;
  part[t] = v;
  __syncthreads();
  for (int off = kScanThreads / 2; off > 0; off >>= 1) {
    if (t < off) part[t] += part[t + off];
    __syncthreads();
  }
  if (t == 0) sums[blockIdx.x] = part[0];
}

__global__ void __launch_bounds__(kScanThreads)
    k_scan_apply(int32_t* __restrict__ cnt, long n, const int32_t* __restrict__ offs) {
  __shared__ int32_t part[kScanThreads];
  const int t = threadIdx.x;
  const long b = (long)blockIdx.x * kScanChunk + (long)t * kScanPer;
  int32_t v[kScanPer], s = 0;
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) {
    v[i] = (b + i < n) ? cnt[b + i] : 0;
    s += v[i];
  }
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {  // inclusive Hillis-Steele scan
    const int32_t x = (t >= off) ? part[t - off] : 0;
    __syncthreads();
    part[t] += x;
    __syncthreads();
  }
  int32_t run = offs[blockIdx.x] + part[t] - s;  // exclusive prefix of this thread's tiles
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) {
    if (b + i < n) cnt[b + i] = run;
    run += v[i];
  }
}

// Rejected cell ids of each 64-cell tile at the tile's offset, in cell order.
__global__ void __launch_bounds__(64)
    k_nn_compact(const int8_t* __restrict__ status, const int32_t* __restrict__ off, long ncell,
                 int32_t* __restrict__ list) {
  const int lane = threadIdx.x;
  const long id = (long)blockIdx.x * 64 + lane;
  const bool rej = id < ncell && status[id] >= 2;
  const unsigned long long bal = __ballot(rej);
  if (rej) list[off[blockIdx.x] + __popcll(bal & ((1ull << lane) - 1ull))] = (int32_t)id;
}

// P = I - S E^T (E S E^T)^-1 E with E[e][k] = atoms_ek / W_k and S = diag(out_std^2):
// the out_std-weighted orthogonal projection onto {dY : E dY = 0}, which
// conserves every element (and so total mass).  Returns nonzero if E S E^T is
// singular.
static int nn_projection(const f3d_ctx* c, const f3d_nn_config* cfg, double (&Pm)[NN_OUT][NN_OUT]) {
  const int ne = cfg->n_elem;
  const int ns = c->g.Ns;  // species of the context (<= NN_OUT)
  for (int i = 0; i < NN_OUT; ++i)
    for (int j = 0; j < NN_OUT; ++j) Pm[i][j] = (i == j) ? 1.0 : 0.0;
  if (ne == 0) return 0;
  double E[4][NN_OUT], S[NN_OUT], G[4][8];
  for (int e = 0; e < ne; ++e)
    for (int k = 0; k < ns; ++k) E[e][k] = cfg->elem[e][k] / c->host_params.mw[k];
  for (int k = 0; k < ns; ++k) S[k] = (double)cfg->out_std[k] * cfg->out_std[k];
  // G = [E S E^T | I], Gauss-Jordan with partial pivoting -> inverse in G[:, ne:]
  for (int a = 0; a < ne; ++a)
    for (int b = 0; b < 2 * ne; ++b) {
      double v = 0.0;
      if (b < ne)
        for (int k = 0; k < ns; ++k) v += E[a][k] * S[k] * E[b][k];
      else
        v = (b - ne == a) ? 1.0 : 0.0;
      G[a][b] = v;
    }
  double scale = 0.0;
  for (int a = 0; a < ne; ++a) scale = std::fmax(scale, std::fabs(G[a][a]));
  for (int col = 0; col < ne; ++col) {
    int piv = col;
    for (int a = col + 1; a < ne; ++a)
      if (std::fabs(G[a][col]) > std::fabs(G[piv][col])) piv = a;
    if (!(std::fabs(G[piv][col]) > 1e-14 * scale)) return 1;
    for (int b = 0; b < 2 * ne; ++b) std::swap(G[col][b], G[piv][b]);
    const double d = G[col][col];
    for (int b = 0; b < 2 * ne; ++b) G[col][b] /= d;
    for (int a = 0; a < ne; ++a)
      if (a != col) {
        const double f = G[a][col];
        for (int b = 0; b < 2 * ne; ++b) G[a][b] -= f * G[col][b];
      }
  }
  // Pm -= S E^T Ginv E
  for (int i = 0; i < ns; ++i)
    for (int j = 0; j < ns; ++j) {
      double v = 0.0;
      for (int a = 0; a < ne; ++a)
        for (int b = 0; b < ne; ++b) v += S[i] * E[a][i] * G[a][ne + b] * E[b][j];
      Pm[i][j] -= v;
    }
  return 0;
}

static int nn_check(f3d_ctx* c) {
  if (!c->nn_ready) {
    set_error("call f3d_nn_init first");
    return F3D_ERR_STATE;
  }
  return F3D_OK;
}

static int nn_partials(f3d_ctx* c, long tiles) {
  if (tiles <= c->nn_partial_tiles) return F3D_OK;
  if (c->nn_partial) (void)hipFree(c->nn_partial);
  c->nn_partial = nullptr;
  c->nn_partial_tiles = 0;
  if (hipMalloc(&c->nn_partial, sizeof(float) * tiles * (NN_NPARAM + 1)) != hipSuccess) {
    set_error("f3d_nn: partial-gradient allocation failed");
    return F3D_ERR_ALLOC;
  }
  c->nn_partial_tiles = tiles;
  return F3D_OK;
}

// Scratch of f3d_nn_train_step_static: the per-tile partial gradients and
// losses of at most kStaticTiles tiles of 64 samples (batch <= 8192), 6.4 MB.
// Every tile writes all of its NN_NPARAM + 1 slots, so nothing needs
// clearing.  Each GPU holds its own copy until the process ends; host code
// uses the device address hipGetSymbolAddress gives on that device, never
// the array's name.
constexpr int kStaticTiles = 128;
__device__ __align__(256) float scratch[(size_t)kStaticTiles * (NN_NPARAM + 1)];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static float* g_scratch_ptr[kMaxDevices] = {nullptr};

// The current device's address of `scratch`, looked up on the first call on
// that device (with the device current) and kept.
static int static_scratch(float** out) {
  int dev = -1;
  int rc = hip_check(hipGetDevice(&dev), "f3d_nn_train_step_static: hipGetDevice");
  if (rc) return rc;
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("f3d_nn_train_step_static: device ordinal beyond the scratch table");
    return F3D_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    rc = hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)),
                   "f3d_nn_train_step_static: hipGetSymbolAddress");
    if (rc) return rc;
    g_scratch_ptr[dev] = (float*)ptr;
  }
  *out = g_scratch_ptr[dev];
  return F3D_OK;
}

void nn_free(f3d_ctx* c) {
  void** ptrs[] = {(void**)&c->nn_master, (void**)&c->nn_m1, (void**)&c->nn_m2,
                   (void**)&c->nn_Wb, (void**)&c->nn_Wtb, (void**)&c->nn_partial,
                   (void**)&c->nn_cfg, (void**)&c->nn_status, (void**)&c->nn_fbcount,
                   (void**)&c->nn_fblist, (void**)&c->nn_fbtotal, (void**)&c->nn_out,
                   (void**)&c->nn_cand};
  for (void** q : ptrs) {
    if (*q) (void)hipFree(*q);
    *q = nullptr;
  }
  for (hipEvent_t& e : c->nn_ev)
    if (e) {
      (void)hipEventDestroy(e);
      e = nullptr;
    }
  c->nn_timing = c->nn_timed = false;
  c->nn_ready = false;
}

}  // namespace f3d

using namespace f3d;

extern "C" {

int f3d_nn_init(f3d_ctx* c, const f3d_nn_config* cfg) {
  if (!c || !cfg) {
    set_error("f3d_nn_init: null argument");
    return F3D_ERR_ARG;
  }
  if (c->g.Ns != 9 && c->g.Ns != 5) {
    set_error("f3d_nn_init: the surrogate step is built for Ns = 9 (H2/O2) or 5 (air)");
    return F3D_ERR_UNSUPPORTED;
  }
  const int ns = c->g.Ns, nin = c->g.Ns + 2;  // outputs and inputs used (<= 9, <= 11)
  if (cfg->n_elem < 0 || cfg->n_elem > 4 || !(cfg->dt > 0) || !(cfg->boxcox_lambda > 0) ||
      cfg->label_mode < 0 || cfg->label_mode > 3 || !(cfg->y_floor >= 0)) {
    set_error("f3d_nn_init: need 0 <= n_elem <= 4, dt > 0, boxcox_lambda > 0, label_mode 0..3, "
              "y_floor >= 0");
    return F3D_ERR_ARG;
  }
  if (cfg->label_mode == 3) {
    bool okl = cfg->label_rel > 0 && cfg->label_floor > 0;
    for (int k = 0; k < ns; ++k) okl = okl && cfg->label_scale[k] >= cfg->label_floor;
    if (!okl) {
      set_error("f3d_nn_init: label_mode 3 needs label_rel > 0 and label_scale >= label_floor > 0");
      return F3D_ERR_ARG;
    }
  }
  for (int k = 0; k < nin; ++k)
    if (!(cfg->in_std[k] > 0)) {
      set_error("f3d_nn_init: in_std must be positive");
      return F3D_ERR_ARG;
    }
  NNDev h;
  std::memset(&h, 0, sizeof(h));
  for (int k = 0; k < NN_IN; ++k) {
    h.in_mean[k] = cfg->in_mean[k];
    h.in_std[k] = cfg->in_std[k];
    h.mahal_mean[k] = cfg->mahal_mean[k];
    for (int j = 0; j < NN_IN; ++j) h.mahal_inv[k][j] = cfg->mahal_inv[k * NN_IN + j];
  }
  for (int k = 0; k < NN_OUT; ++k) {
    h.out_mean[k] = cfg->out_mean[k];
    h.out_std[k] = cfg->out_std[k];
  }
  h.d_max = cfg->d_max;
  h.elem_tol = cfg->elem_tol;
  h.neg_tol = cfg->neg_tol;
  h.res_tol = cfg->res_tol;
  h.res_T_tol = cfg->res_T_tol;
  h.dT_max = cfg->dT_max;
  h.chi_max = cfg->chi_max;
  h.chain_max = cfg->chain_max;
  {
    int nc = 0;
    for (int j = 0; j < 3; ++j) h.chain_idx[j] = -1;
    for (int k = 0; k < ns; ++k)
      if ((cfg->chain_species >> k) & 1) {
        if (nc == 3) {
          set_error("f3d_nn_init: chain_species names more than 3 species");
          return F3D_ERR_ARG;
        }
        h.chain_idx[nc++] = k;
      }
    if ((cfg->chain_species >> ns) != 0 || (cfg->chain_max > 0 && nc == 0)) {
      set_error("f3d_nn_init: chain_species must name 1..3 of the context's species when chain_max > 0");
      return F3D_ERR_ARG;
    }
  }
  h.boxcox_lambda = cfg->boxcox_lambda;
  h.label_mode = cfg->label_mode;
  for (int k = 0; k < NN_OUT; ++k) h.label_scale[k] = cfg->label_scale[k];
  h.label_rel = cfg->label_rel;
  h.label_floor = cfg->label_floor;
  h.y_floor = cfg->y_floor;
  h.n_elem = cfg->n_elem;
  for (int e = 0; e < 4; ++e)
    for (int k = 0; k < NN_OUT; ++k) {
      h.elem[e][k] = cfg->elem[e][k];
      h.elem_mw[e][k] = k < c->g.Ns ? (double)h.elem[e][k] / c->host_params.mw[k] : 0.0;
    }
  h.dt = cfg->dt;
  if (nn_projection(c, cfg, h.proj)) {
    set_error("f3d_nn_init: element matrix is rank deficient (an element with no species?)");
    return F3D_ERR_ARG;
  }
  if (!c->nn_ready) {
    const long ncell = (long)c->g.Nx * c->g.Ny * c->g.Nz;
    const long ntile = (ncell + 63) / 64, nchunk = (ntile + kScanChunk - 1) / kScanChunk;
    bool okm = hipMalloc(&c->nn_master, sizeof(float) * NN_NPARAM) == hipSuccess &&
               hipMalloc(&c->nn_m1, sizeof(float) * NN_NPARAM) == hipSuccess &&
               hipMalloc(&c->nn_m2, sizeof(float) * NN_NPARAM) == hipSuccess &&
               hipMalloc(&c->nn_Wb, 2 * NN_NPARAM) == hipSuccess &&
               hipMalloc(&c->nn_Wtb, 2 * NN_NPARAM) == hipSuccess &&
               hipMalloc(&c->nn_cfg, sizeof(NNDev)) == hipSuccess &&
               hipMalloc(&c->nn_status, ncell) == hipSuccess &&
               hipMalloc(&c->nn_fbcount, sizeof(int32_t) * (ntile + nchunk)) == hipSuccess &&
               hipMalloc(&c->nn_fblist, sizeof(int32_t) * ncell) == hipSuccess &&
               hipMalloc(&c->nn_fbtotal, sizeof(int32_t)) == hipSuccess &&
               hipMalloc(&c->nn_out, sizeof(float) * ncell * ns) == hipSuccess &&
               hipMalloc(&c->nn_cand, sizeof(double) * ncell * (ns + 3)) == hipSuccess;
    if (!okm) {
      nn_free(c);
      set_error("f3d_nn_init: device allocation failed");
      return F3D_ERR_ALLOC;
    }
    (void)hipMemset(c->nn_master, 0, sizeof(float) * NN_NPARAM);
    (void)hipMemset(c->nn_m1, 0, sizeof(float) * NN_NPARAM);
    (void)hipMemset(c->nn_m2, 0, sizeof(float) * NN_NPARAM);
    (void)hipMemset(c->nn_Wb, 0, 2 * NN_NPARAM);
    (void)hipMemset(c->nn_Wtb, 0, 2 * NN_NPARAM);
    c->nn_ready = true;
  }
  c->nn_dt = cfg->dt;
  return hip_check(hipMemcpy(c->nn_cfg, &h, sizeof(NNDev), hipMemcpyHostToDevice), "f3d_nn_init");
}

int f3d_nn_set_params(f3d_ctx* c, const float* params) {
  if (!c || !params) {
    set_error("f3d_nn_set_params: null argument");
    return F3D_ERR_ARG;
  }
  int rc = nn_check(c);
  if (rc) return rc;
  rc = hip_check(hipMemcpy(c->nn_master, params, sizeof(float) * NN_NPARAM, hipMemcpyDefault),
                 "f3d_nn_set_params");
  if (rc) return rc;
  (void)hipMemset(c->nn_m1, 0, sizeof(float) * NN_NPARAM);
  (void)hipMemset(c->nn_m2, 0, sizeof(float) * NN_NPARAM);
  c->nn_step = 0;
  hipLaunchKernelGGL(k_nn_cast, dim3(nblocks(NN_NPARAM)), dim3(kBlock), 0, 0, c->nn_master,
                     (__bf16*)c->nn_Wb, (__bf16*)c->nn_Wtb);
  rc = hip_check(hipGetLastError(), "f3d_nn_set_params");
  if (rc) return rc;
  return hip_check(hipDeviceSynchronize(), "f3d_nn_set_params");
}

int f3d_nn_get_params(f3d_ctx* c, float* params) {
  if (!c || !params) {
    set_error("f3d_nn_get_params: null argument");
    return F3D_ERR_ARG;
  }
  int rc = nn_check(c);
  if (rc) return rc;
  return hip_check(hipMemcpy(params, c->nn_master, sizeof(float) * NN_NPARAM, hipMemcpyDefault),
                   "f3d_nn_get_params");
}

static int nn_batch_args(f3d_ctx* c, const void* X, const void* Y, int32_t B, const char* who) {
  if (!c || !X || !Y) {
    set_error(std::string(who) + ": null argument");
    return F3D_ERR_ARG;
  }
  if (B <= 0 || B % 64 != 0) {
    set_error(std::string(who) + ": batch size must be a positive multiple of 64");
    return F3D_ERR_ARG;
  }
  return nn_check(c);
}

int f3d_nn_grad(f3d_ctx* c, const float* X, const float* Y, int32_t B, float* grad, hipStream_t s) {
  int rc = nn_batch_args(c, X, Y, B, "f3d_nn_grad");
  if (rc) return rc;
  if (!grad) {
    set_error("f3d_nn_grad: null grad");
    return F3D_ERR_ARG;
  }
  const long tiles = B / 64;
  rc = nn_partials(c, tiles);
  if (rc) return rc;
  hipLaunchKernelGGL(k_nn_train_tile, dim3(tiles), dim3(256), 0, s, X, Y, B, c->g.Ns + 2, c->g.Ns,
                     (const __bf16*)c->nn_Wb, (const __bf16*)c->nn_Wtb, c->nn_master,
                     c->nn_partial);
  hipLaunchKernelGGL(k_nn_reduce, dim3(nblocks(NN_NPARAM + 1)), dim3(kBlock), 0, s,
                     c->nn_partial, (int)tiles, grad);
  return hip_check(hipGetLastError(), "f3d_nn_grad");
}

int f3d_nn_train_step(f3d_ctx* c, const float* X, const float* Y, int32_t B, float lr,
                      float* loss, hipStream_t s) {
  int rc = nn_batch_args(c, X, Y, B, "f3d_nn_train_step");
  if (rc) return rc;
  const long tiles = B / 64;
  rc = nn_partials(c, tiles);
  if (rc) return rc;
  c->nn_step += 1;
  const double b1 = 0.9, b2 = 0.999;
  const float bc1 = (float)(1.0 - std::pow(b1, (double)c->nn_step));
  const float bc2 = (float)(1.0 - std::pow(b2, (double)c->nn_step));
  hipLaunchKernelGGL(k_nn_train_tile, dim3(tiles), dim3(256), 0, s, X, Y, B, c->g.Ns + 2, c->g.Ns,
                     (const __bf16*)c->nn_Wb, (const __bf16*)c->nn_Wtb, c->nn_master,
                     c->nn_partial);
  hipLaunchKernelGGL(k_nn_adam, dim3(nblocks(NN_NPARAM + 1)), dim3(kBlock), 0, s, c->nn_partial,
                     (int)tiles, c->nn_master, c->nn_m1, c->nn_m2, (__bf16*)c->nn_Wb,
                     (__bf16*)c->nn_Wtb, lr, (float)b1, (float)b2, 1e-8f, bc1, bc2, loss);
  return hip_check(hipGetLastError(), "f3d_nn_train_step");
}

int f3d_nn_train_step_static(f3d_ctx* c, const float* X, const float* Y, int32_t B, float lr,
                             float* loss, hipStream_t s) {
  int rc = nn_batch_args(c, X, Y, B, "f3d_nn_train_step_static");
  if (rc) return rc;
  const long tiles = B / 64;
  if (tiles > kStaticTiles) {
    set_error("f3d_nn_train_step_static: batch size above the static scratch's 8192 samples");
    return F3D_ERR_ARG;
  }
  float* part = nullptr;
  rc = static_scratch(&part);
  if (rc) return rc;
  c->nn_step += 1;
  const double b1 = 0.9, b2 = 0.999;
  const float bc1 = (float)(1.0 - std::pow(b1, (double)c->nn_step));
  const float bc2 = (float)(1.0 - std::pow(b2, (double)c->nn_step));
  hipLaunchKernelGGL(k_nn_train_tile, dim3(tiles), dim3(256), 0, s, X, Y, B, c->g.Ns + 2, c->g.Ns,
                     (const __bf16*)c->nn_Wb, (const __bf16*)c->nn_Wtb, c->nn_master, part);
  hipLaunchKernelGGL(k_nn_adam, dim3(nblocks(NN_NPARAM + 1)), dim3(kBlock), 0, s, part,
                     (int)tiles, c->nn_master, c->nn_m1, c->nn_m2, (__bf16*)c->nn_Wb,
                     (__bf16*)c->nn_Wtb, lr, (float)b1, (float)b2, 1e-8f, bc1, bc2, loss);
  return hip_check(hipGetLastError(), "f3d_nn_train_step_static");
}

int f3d_nn_forward(f3d_ctx* c, const float* X, float* Y, int32_t B, hipStream_t s) {
  int rc = nn_batch_args(c, X, Y, B, "f3d_nn_forward");
  if (rc) return rc;
  hipLaunchKernelGGL(k_nn_forward, dim3(B / 64), dim3(64), 0, s, X, Y, c->g.Ns + 2, c->g.Ns,
                     (const __bf16*)c->nn_Wb, c->nn_master);
  return hip_check(hipGetLastError(), "f3d_nn_forward");
}

int f3d_nn_chem(f3d_ctx* c, double* U, double* Q, double* rhoi, double dt, int32_t maxiter,
                double rtol, double ttol, int8_t* status, int64_t* n_fallback, const int64_t* tag,
                hipStream_t s) {
  if (!c || !U || !Q || !rhoi || !n_fallback) {
    set_error("f3d_nn_chem: null argument");
    return F3D_ERR_ARG;
  }
  int rc = nn_check(c);
  if (rc) return rc;
  if (std::fabs(dt - c->nn_dt) > 1e-12 * c->nn_dt) {
    set_error("f3d_nn_chem: dt differs from the step the network was trained for");
    return F3D_ERR_ARG;
  }
  if (maxiter < 1 || maxiter > 1000 || !(rtol > 0) || !(ttol > 0)) {
    set_error("f3d_nn_chem: need 1 <= maxiter <= 1000, rtol > 0, ttol > 0");
    return F3D_ERR_ARG;
  }
  const long ncell = (long)c->g.Nx * c->g.Ny * c->g.Nz;
  const long ntile = (ncell + 63) / 64;
  int8_t* st = status ? status : c->nn_status;
  const bool timed = c->nn_timing;
  if (timed) (void)hipEventRecord(c->nn_ev[0], s);
  // every tile writes its count, so no clearing is needed
  const unsigned nb = nblocks(ncell);
  if (c->g.Ns == 9)
    hipLaunchKernelGGL((k_nn_net<9>), dim3((unsigned)ntile), dim3(64), 0, s, Q, rhoi, tag,
                       c->dparams, (const NNDev*)c->nn_cfg, (const __bf16*)c->nn_Wb, c->nn_master,
                       c->g, c->nn_out, st);
  else
    hipLaunchKernelGGL((k_nn_net<5>), dim3((unsigned)ntile), dim3(64), 0, s, Q, rhoi, tag,
                       c->dparams, (const NNDev*)c->nn_cfg, (const __bf16*)c->nn_Wb, c->nn_master,
                       c->g, c->nn_out, st);
  if (timed) (void)hipEventRecord(c->nn_ev[1], s);
  if (c->g.Ns == 9) {
    hipLaunchKernelGGL((k_nn_pred<9>), dim3(nb), dim3(kBlock), 0, s, Q, rhoi, c->dparams,
                       (const NNDev*)c->nn_cfg, c->g, c->nn_out, c->nn_cand, st);
    if (c->kin_table == F3D_KIN_H2_LIDRYER)
      hipLaunchKernelGGL((k_nn_gate<9, KinCT<MechH2LiDryer>>), dim3(nb), dim3(kBlock), 0, s, U, Q,
                         rhoi, c->dparams, (const NNDev*)c->nn_cfg, c->g, dt, c->nn_cand,
                         c->nn_fbcount, st);
    else
      hipLaunchKernelGGL((k_nn_gate<9>), dim3(nb), dim3(kBlock), 0, s, U, Q, rhoi, c->dparams,
                         (const NNDev*)c->nn_cfg, c->g, dt, c->nn_cand, c->nn_fbcount, st);
  } else {
    hipLaunchKernelGGL((k_nn_pred<5>), dim3(nb), dim3(kBlock), 0, s, Q, rhoi, c->dparams,
                       (const NNDev*)c->nn_cfg, c->g, c->nn_out, c->nn_cand, st);
    if (c->kin_table == F3D_KIN_AIR5)
      hipLaunchKernelGGL((k_nn_gate<5, KinCT<MechAir5>>), dim3(nb), dim3(kBlock), 0, s, U, Q,
                         rhoi, c->dparams, (const NNDev*)c->nn_cfg, c->g, dt, c->nn_cand,
                         c->nn_fbcount, st);
    else
      hipLaunchKernelGGL((k_nn_gate<5>), dim3(nb), dim3(kBlock), 0, s, U, Q, rhoi, c->dparams,
                         (const NNDev*)c->nn_cfg, c->g, dt, c->nn_cand, c->nn_fbcount, st);
  }
  if (timed) (void)hipEventRecord(c->nn_ev[2], s);
  // tile counts -> tile offsets; the chunk sums sit after the tile counts
  const long nchunk = (ntile + kScanChunk - 1) / kScanChunk;
  int32_t* sums = c->nn_fbcount + ntile;
  hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)nchunk), dim3(kScanThreads), 0, s,
                     c->nn_fbcount, ntile, sums);
  hipLaunchKernelGGL(k_nn_scan, dim3(1), dim3(1024), 0, s, sums, nchunk, c->nn_fbtotal);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nchunk), dim3(kScanThreads), 0, s,
                     c->nn_fbcount, ntile, sums);
  hipLaunchKernelGGL(k_nn_compact, dim3((unsigned)ntile), dim3(64), 0, s, st, c->nn_fbcount,
                     ncell, c->nn_fblist);
  rc = hip_check(hipGetLastError(), "f3d_nn_chem");
  if (rc) return rc;
  if (timed) (void)hipEventRecord(c->nn_ev[3], s);
  int32_t nfb = 0;
  rc = hip_check(hipMemcpyAsync(&nfb, c->nn_fbtotal, sizeof(int32_t), hipMemcpyDeviceToHost, s),
                 "f3d_nn_chem");
  if (rc) return rc;
  rc = hip_check(hipStreamSynchronize(s), "f3d_nn_chem");
  if (rc) return rc;
  *n_fallback = nfb;
  // the rejected cells, in cell order: the stiff ones stay grouped in a few
  // waves (a wave's cost is its slowest cell) instead of one in every wave
  if (timed) (void)hipEventRecord(c->nn_ev[4], s);  // after the read-back
  if (nfb > 0) {
    rc = launch_chem(c, 2, U, Q, rhoi, dt, maxiter, rtol, ttol, nullptr, tag, s, c->nn_fblist,
                     nfb);
    if (rc) return rc;
  }
  if (timed) {
    // the fallback's end replaces the read-back mark: ev[4] - ev[3] then spans
    // the 4-byte read-back and the Newton fallback launch
    (void)hipEventRecord(c->nn_ev[4], s);
    c->nn_timed = true;
  }
  return F3D_OK;
}

int f3d_nn_timing(f3d_ctx* c, int32_t enable, double* ms) {
  if (!c) {
    set_error("f3d_nn_timing: null context");
    return F3D_ERR_ARG;
  }
  if (enable && !c->nn_ev[0]) {
    for (hipEvent_t& e : c->nn_ev) {
      int rc = hip_check(hipEventCreate(&e), "f3d_nn_timing");
      if (rc) return rc;
    }
  }
  c->nn_timing = enable != 0;
  if (ms) {
    if (!c->nn_timed) {
      set_error("f3d_nn_timing: no timed f3d_nn_chem call yet");
      return F3D_ERR_STATE;
    }
    int rc = hip_check(hipEventSynchronize(c->nn_ev[4]), "f3d_nn_timing");
    if (rc) return rc;
    float t[4];
    for (int k = 0; k < 4; ++k) {
      rc = hip_check(hipEventElapsedTime(&t[k], c->nn_ev[k], c->nn_ev[k + 1]), "f3d_nn_timing");
      if (rc) return rc;
      ms[k] = t[k];
    }
  }
  return F3D_OK;
}

}  // extern "C"
