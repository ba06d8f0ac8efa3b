// Shape and packed-parameter layout of the chemistry-surrogate MLP
// 11 -> 64 -> 64 -> 64 -> 9 (inputs T, log P, Y_1..Y_9; outputs dY_1..dY_9,
// all normalized).  Input and output widths are padded to multiples of 32
// (the K of v_mfma_f32_16x16x32_bf16); padded weights are zero.
//
// FP32 parameter vector (NN_NPARAM floats):
//   W1 [64][32] | b1 [64] | W2 [64][64] | b2 [64] | W3 [64][64] | b3 [64] | W4 [32][64] | b4 [32]
// Row-major W[o][k] (output o, input k).  The BF16 copies use the same
// offsets; the transposed BF16 copy holds W^T[k][o] at the same offsets.
#pragma once

#define NN_IN 11
#define NN_INP 32
#define NN_H 64
#define NN_OUT 9
#define NN_OUTP 32
#define NN_NPARAM 12512

#ifdef __HIPCC__
#define NN_HD __host__ __device__
#else
#define NN_HD
#endif

namespace f3d {

NN_HD inline int nn_in_dim(int l) { return l == 0 ? NN_INP : NN_H; }
NN_HD inline int nn_out_dim(int l) { return l == 3 ? NN_OUTP : NN_H; }
NN_HD inline int nn_w_off(int l) {
  return l == 0 ? 0 : l == 1 ? 2112 : l == 2 ? 6272 : 10432;
}
NN_HD inline int nn_b_off(int l) { return nn_w_off(l) + nn_in_dim(l) * nn_out_dim(l); }

// map a parameter index to (layer, o, k) if it is a weight (not a bias)
NN_HD inline bool nn_locate_weight(int p, int& l, int& o, int& k) {
  for (l = 3; l >= 0; --l)
    if (p >= nn_w_off(l)) break;
  const int r = p - nn_w_off(l);
  const int ni = nn_in_dim(l), no = nn_out_dim(l);
  if (r >= ni * no) return false;  // bias
  o = r / ni;
  k = r % ni;
  return true;
}

}  // namespace f3d

namespace f3d {

// Device-side surrogate configuration (normalisation, gates).
struct NNDev {
  float in_mean[NN_IN], in_std[NN_IN], out_mean[NN_OUT], out_std[NN_OUT];
  float mahal_mean[NN_IN];
  float mahal_inv[NN_IN][NN_IN];  // inverse covariance of the normalized inputs
  float d_max;                     // Mahalanobis-distance gate
  float elem_tol;                  // element-balance gate (net / gross turnover)
  float neg_tol;                   // most negative admissible Y_k + dY_k
  float res_tol;                   // backward-Euler residual gate (<= 0: off)
  float res_T_tol;                 // temperature-weighted residual gate [K] (<= 0: off)
  float dT_max;                    // predicted chemical temperature change gate [K] (<= 0: off)
  float chi_max;                   // thermal explosiveness gate (<= 0: off)
  float chain_max;                 // chain explosiveness gate (<= 0: off)
  int chain_idx[3];                // chain-carrier species of that gate (-1: unused)
  float boxcox_lambda;             // Y -> (Y^lambda - 1)/lambda input transform
  int label_mode;                  // 0: dY; 1: d BoxCox(Y); 2: asinh(dY / label_scale);
                                   // 3: asinh(dY / s(Y)), s = min(scale, max(floor, rel Y))
  float label_scale[NN_OUT];
  float label_rel, label_floor;    // label_mode 3
  float y_floor;                   // inputs use BoxCox(max(Y, y_floor))
  int n_elem;
  float elem[4][NN_OUT];           // atoms of element e per molecule of species k
  double dt;                       // chemistry time step the network was trained for
  double proj[NN_OUT][NN_OUT];     // element-conserving projection of dY (identity if none)
  double elem_mw[4][NN_OUT];       // elem[e][k] / W_k (atoms of e per kg of species k, / 1000)
};

}  // namespace f3d
