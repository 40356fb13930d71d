// HIP kernels of the configuration picker's perceptron: forward, backward
// (loss gradient with respect to every parameter), Adam update and the row
// gather used to form minibatches.
//
// The network is tiny (6100 parameters, batches of a few hundred rows), so
// every kernel is a plain one-thread-per-output loop with a fixed summation
// order: results are bitwise reproducible run to run, which the training
// tests rely on.  There are no atomics.
#include "mlp.hpp"
#include "gemm_kernels.hpp"
#include <cmath>

namespace gemm_amd {

void mlp_features(int op, bool ta, bool tb, int m, int n, int k, int lda_pad, bool vec, int cus, float *f) {
    f[0] = (float)((std::log2((double)(m > 0 ? m : 1)) - 9.0) / 5.0);
    f[1] = (float)((std::log2((double)(n > 0 ? n : 1)) - 9.0) / 5.0);
    f[2] = (float)((std::log2((double)(k > 0 ? k : 1)) - 9.0) / 5.0);
    f[3] = ta ? 1.0f : 0.0f;
    f[4] = tb ? 1.0f : 0.0f;
    f[5] = op == 0 ? 1.0f : 0.0f;
    f[6] = op == 1 ? 1.0f : 0.0f;
    f[7] = op == 2 ? 1.0f : 0.0f;
    f[8] = (float)(((lda_pad % 64) + 64) % 64) / 64.0f;
    f[9] = vec ? 1.0f : 0.0f;
    const int eb = op == 1 ? 8 : 4;   // the CGEMM3M products are FP32 GEMMs
    const int64_t mm = m > 0 ? m : 1, nn = n > 0 ? n : 1, cu = cus > 0 ? cus : 1;
    for (int c = 0; c < 20; ++c) {
        const int tile = c / 5, bm = tile_bm(tile, eb), bn = tile_bn(tile, eb);
        const int64_t w = ((mm + bm - 1) / bm) * ((nn + bn - 1) / bn) * gemm_k_partitions(k, 1 << (c % 5));
        const int64_t waves = (w + cu - 1) / cu;
        f[10 + c] = (float)((double)w / (double)(waves * cu));
    }
    for (int t = 0; t < 4; ++t) {
        const int64_t bm = tile_bm(t, eb), bn = tile_bn(t, eb);
        f[30 + t] = (float)((double)(mm * nn) / (double)(((mm + bm - 1) / bm) * bm * ((nn + bn - 1) / bn) * bn));
    }
}

// Z[b][o] = bias[o] + sum_i X[b][i] W[i][o]  (ReLU optional)
template <bool RELU>
__global__ __launch_bounds__(256) void linear_fwd_kernel(const float *__restrict__ X, int B, int in,
                                                         const float *__restrict__ W, const float *__restrict__ bias,
                                                         int out, float *__restrict__ Z) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= B * out) return;
    const int b = idx / out, o = idx % out;
    float acc = bias[o];
    const float *x = X + (int64_t)b * in;
    for (int i = 0; i < in; ++i) acc = __builtin_fmaf(x[i], W[i * out + o], acc);
    Z[idx] = RELU ? fmaxf(acc, 0.0f) : acc;
}

// dY = 2 mask (Y - T) / (B * out); loss = sum mask (Y - T)^2 / (B * out).  One workgroup.
__global__ __launch_bounds__(256) void mse_grad_kernel(const float *__restrict__ Y, const float *__restrict__ T,
                                                       const float *__restrict__ M, int total, float inv,
                                                       float *__restrict__ dY, float *__restrict__ loss) {
    __shared__ float part[256];
    float s = 0.0f;
    for (int i = threadIdx.x; i < total; i += 256) {
        const float d = M[i] * (Y[i] - T[i]);
        dY[i] = 2.0f * d * inv;
        s = __builtin_fmaf(d, d, s);
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = part[0] * inv;
}

// dW[i][o] = sum_b X[b][i] dZ[b][o];  db[o] = sum_b dZ[b][o]
__global__ __launch_bounds__(256) void linear_bwd_w_kernel(const float *__restrict__ X, const float *__restrict__ dZ,
                                                           int B, int in, int out, float *__restrict__ dW,
                                                           float *__restrict__ db) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx < in * out) {
        const int i = idx / out, o = idx % out;
        float acc = 0.0f;
        for (int b = 0; b < B; ++b) acc = __builtin_fmaf(X[(int64_t)b * in + i], dZ[(int64_t)b * out + o], acc);
        dW[idx] = acc;
    } else if (idx < in * out + out) {
        const int o = idx - in * out;
        float acc = 0.0f;
        for (int b = 0; b < B; ++b) acc += dZ[(int64_t)b * out + o];
        db[o] = acc;
    }
}

// dH[b][i] = [H[b][i] > 0] * sum_o dZ[b][o] W[i][o]   (H = ReLU output feeding the layer)
__global__ __launch_bounds__(256) void linear_bwd_x_kernel(const float *__restrict__ dZ, const float *__restrict__ W,
                                                           const float *__restrict__ H, int B, int in, int out,
                                                           float *__restrict__ dH) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= B * in) return;
    const int b = idx / in, i = idx % in;
    float acc = 0.0f;
    if (H[idx] > 0.0f) {
        const float *dz = dZ + (int64_t)b * out;
        for (int o = 0; o < out; ++o) acc = __builtin_fmaf(dz[o], W[i * out + o], acc);
    }
    dH[idx] = acc;
}

__global__ __launch_bounds__(256) void adam_kernel(float *__restrict__ p, const float *__restrict__ g,
                                                   float *__restrict__ m, float *__restrict__ v, int n, float lr,
                                                   float b1, float b2, float eps, float bc1, float bc2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

__global__ __launch_bounds__(256) void gather_kernel(const float *__restrict__ src, const int *__restrict__ idx,
                                                     int rows, int width, float *__restrict__ dst) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= rows * width) return;
    const int r = t / width, c = t % width;
    dst[t] = src[(int64_t)idx[r] * width + c];
}

static unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

size_t mlp_forward_ws_floats(int batch) { return (size_t)2 * batch * MLP_HID; }
size_t mlp_grad_ws_floats(int batch) { return (size_t)4 * batch * MLP_HID + (size_t)2 * batch * MLP_OUT; }

int mlp_forward(const float *P, const float *x, int B, float *y, float *ws, hipStream_t s) {
    if (B <= 0) return 0;
    float *h1 = ws, *h2 = ws + (size_t)B * MLP_HID;
    hipLaunchKernelGGL(linear_fwd_kernel<true>, dim3(blocks((int64_t)B * MLP_HID)), dim3(256), 0, s, x, B, MLP_IN,
                       P + MLP_OFF_W1, P + MLP_OFF_B1, MLP_HID, h1);
    hipLaunchKernelGGL(linear_fwd_kernel<true>, dim3(blocks((int64_t)B * MLP_HID)), dim3(256), 0, s, h1, B, MLP_HID,
                       P + MLP_OFF_W2, P + MLP_OFF_B2, MLP_HID, h2);
    hipLaunchKernelGGL(linear_fwd_kernel<false>, dim3(blocks((int64_t)B * MLP_OUT)), dim3(256), 0, s, h2, B, MLP_HID,
                       P + MLP_OFF_W3, P + MLP_OFF_B3, MLP_OUT, y);
    return (int)hipGetLastError();
}

int mlp_grad(const float *P, const float *x, const float *t, const float *mask, int B, float *G, float *loss,
             float *ws, hipStream_t s) {
    if (B <= 0) return (int)hipErrorInvalidValue;
    float *h1 = ws, *h2 = ws + (size_t)B * MLP_HID;
    float *dh1 = h2 + (size_t)B * MLP_HID, *dh2 = dh1 + (size_t)B * MLP_HID;
    float *y = dh2 + (size_t)B * MLP_HID, *dy = y + (size_t)B * MLP_OUT;
    if (int e = mlp_forward(P, x, B, y, ws, s)) return e;
    hipLaunchKernelGGL(mse_grad_kernel, dim3(1), dim3(256), 0, s, y, t, mask, B * MLP_OUT,
                       1.0f / (float)(B * MLP_OUT), dy, loss);
    hipLaunchKernelGGL(linear_bwd_w_kernel, dim3(blocks(MLP_HID * MLP_OUT + MLP_OUT)), dim3(256), 0, s, h2, dy, B,
                       MLP_HID, MLP_OUT, G + MLP_OFF_W3, G + MLP_OFF_B3);
    hipLaunchKernelGGL(linear_bwd_x_kernel, dim3(blocks((int64_t)B * MLP_HID)), dim3(256), 0, s, dy, P + MLP_OFF_W3,
                       h2, B, MLP_HID, MLP_OUT, dh2);
    hipLaunchKernelGGL(linear_bwd_w_kernel, dim3(blocks(MLP_HID * MLP_HID + MLP_HID)), dim3(256), 0, s, h1, dh2, B,
                       MLP_HID, MLP_HID, G + MLP_OFF_W2, G + MLP_OFF_B2);
    hipLaunchKernelGGL(linear_bwd_x_kernel, dim3(blocks((int64_t)B * MLP_HID)), dim3(256), 0, s, dh2, P + MLP_OFF_W2,
                       h1, B, MLP_HID, MLP_HID, dh1);
    hipLaunchKernelGGL(linear_bwd_w_kernel, dim3(blocks(MLP_IN * MLP_HID + MLP_HID)), dim3(256), 0, s, x, dh1, B,
                       MLP_IN, MLP_HID, G + MLP_OFF_W1, G + MLP_OFF_B1);
    return (int)hipGetLastError();
}

int mlp_adam(float *p, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps, int step,
             hipStream_t s) {
    if (n <= 0 || step < 1) return (int)hipErrorInvalidValue;
    const float bc1 = 1.0f - (float)std::pow((double)b1, step), bc2 = 1.0f - (float)std::pow((double)b2, step);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks(n)), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, eps, bc1, bc2);
    return (int)hipGetLastError();
}

int mlp_gather(const float *src, const int *idx, int rows, int width, float *dst, hipStream_t s) {
    if (rows <= 0) return 0;
    hipLaunchKernelGGL(gather_kernel, dim3(blocks((int64_t)rows * width)), dim3(256), 0, s, src, idx, rows, width,
                       dst);
    return (int)hipGetLastError();
}

}  // namespace gemm_amd
