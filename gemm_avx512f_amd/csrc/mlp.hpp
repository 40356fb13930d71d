// The configuration picker's network: a 3-layer perceptron (34 -> 64 -> 64 ->
// 20, ReLU after the two hidden layers, FP32) trained and run by the HIP
// kernels of mlp.hip.
//
// Parameters are one flat FP32 array, row-major weights [in][out]:
//   W1[34][64] b1[64] W2[64][64] b2[64] W3[64][20] b3[20]   (7700 floats)
// Input features of one GEMM problem (mlp_features):
//   [0..8]   (log2 M - 9)/5, (log2 N - 9)/5, (log2 K - 9)/5, transA, transB,
//            op one-hot (SGEMM, DGEMM, CGEMM3M), (lda - rows of op(A) storage) mod 64 / 64
//   [9]      1 when the kernels take their 16-byte vector-load path (leading
//            dimensions a multiple of 16 bytes), else 0
//   [10..29] per candidate c (tile c / 5, split-K 2^(c % 5)): how full the
//            last wave of workgroups is, w / (ceil(w / CUs) CUs) with
//            w = tiles x K partitions
//   [30..33] per tile shape: the share of the padded tile grid that lies
//            inside C, M N / (ceil(M/BM) BM ceil(N/BN) BN)
// The last two groups carry what log2 sizes cannot: a 2048^2 FP32 product
// fills the 256 CUs with 128x128 tiles exactly, a 2185^2 one needs a second,
// mostly empty wave.
// Output c is the predicted log2(time of candidate c / time of the best
// candidate) for the 20 tuning candidates; the picker takes the argmin over
// the candidates valid for the shape.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gemm_amd {

constexpr int MLP_IN = 34, MLP_HID = 64, MLP_OUT = 20;
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_IN * MLP_HID;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_HID;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_HID * MLP_HID;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_HID;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_HID * MLP_OUT;
constexpr int MLP_NPARAMS = MLP_OFF_B3 + MLP_OUT;

// `vec`: the 16-byte load path applies (gemm_vec_path); `cus`: compute units.
void mlp_features(int op, bool ta, bool tb, int m, int n, int k, int lda_pad, bool vec, int cus, float *f);

// Device-side building blocks (all pointers device memory, row-major
// [batch][features]).  Scratch `ws` holds the two hidden activations
// (2 * batch * MLP_HID floats) and, for the gradient, their gradients and the
// output gradient (another 2 * batch * MLP_HID + batch * MLP_OUT floats).
size_t mlp_forward_ws_floats(int batch);
size_t mlp_grad_ws_floats(int batch);
int mlp_forward(const float *params, const float *x, int batch, float *y, float *ws, hipStream_t s);
// loss = sum(mask * (y - t)^2) / (batch * MLP_OUT); grad = d loss / d params.
int mlp_grad(const float *params, const float *x, const float *t, const float *mask, int batch, float *grad,
             float *loss, float *ws, hipStream_t s);
int mlp_adam(float *params, const float *grad, float *m, float *v, int n, float lr, float beta1, float beta2,
             float eps, int step, hipStream_t s);
int mlp_gather(const float *src, const int *idx, int rows, int width, float *dst, hipStream_t s);

}  // namespace gemm_amd
