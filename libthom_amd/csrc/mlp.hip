// Cost-model MLP kernels: y = W2 relu(W1 x + b1) + b2 in FP32, the masked
// mean-squared-error gradient, and an Adam step.  The model predicts the
// log-time of every tuning configuration from a problem's shape features; it
// is trained on the autotuner's timing sweeps.  Sizes are tiny (batch ~2k,
// inputs ~4, hidden 64, outputs ~16), so every reduction is a plain loop in a
// fixed order: results are deterministic.  No counterpart in the reference
// (its tuner is an exhaustive search, hgemm.cu:258-323).
#include "lt_common.hpp"

namespace lt {
namespace {

// hidden[b][h] = relu(b1[h] + sum_i W1[h][i] x[b][i])
__global__ void mlp_hidden_kernel(const float* __restrict__ x, const float* __restrict__ W1,
                                  const float* __restrict__ b1, float* __restrict__ hid, int B, int I, int H) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= B * H) return;
    const int b = idx / H, h = idx % H;
    float s = b1[h];
    for (int i = 0; i < I; i++) s = fmaf(W1[h * I + i], x[b * I + i], s);
    hid[idx] = s > 0.0f ? s : 0.0f;
}

// y[b][o] = b2[o] + sum_h W2[o][h] hidden[b][h]
__global__ void mlp_out_kernel(const float* __restrict__ hid, const float* __restrict__ W2,
                               const float* __restrict__ b2, float* __restrict__ y, int B, int H, int O) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= B * O) return;
    const int b = idx / O, o = idx % O;
    float s = b2[o];
    for (int h = 0; h < H; h++) s = fmaf(W2[o * H + h], hid[b * H + h], s);
    y[idx] = s;
}

// dy = 2 (y - t) mask / count; loss = sum mask (y - t)^2 / count.  One block.
__global__ __launch_bounds__(256) void mlp_dy_kernel(const float* __restrict__ y, const float* __restrict__ t,
                                                     const uint8_t* __restrict__ mask, float* __restrict__ dy,
                                                     float* __restrict__ loss, int n) {
    __shared__ float red[4];
    float cnt = 0.0f;
    for (int i = threadIdx.x; i < n; i += 256) cnt += mask[i] ? 1.0f : 0.0f;
    cnt = block_sum<256>(cnt, red);
    const float inv = cnt > 0.0f ? 1.0f / cnt : 0.0f;
    float l = 0.0f;
    for (int i = threadIdx.x; i < n; i += 256) {
        const float d = mask[i] ? y[i] - t[i] : 0.0f;
        dy[i] = 2.0f * d * inv;
        l = fmaf(d * d, inv, l);
    }
    l = block_sum<256>(l, red);
    if (threadIdx.x == 0) *loss = l;
}

// gW2[o][h] = sum_b dy[b][o] hid[b][h]; gb2[o] = sum_b dy[b][o]
__global__ void mlp_gw2_kernel(const float* __restrict__ dy, const float* __restrict__ hid, float* __restrict__ gW2,
                               float* __restrict__ gb2, int B, int H, int O) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= O * (H + 1)) return;
    const int o = idx / (H + 1), h = idx % (H + 1);
    float s = 0.0f;
    if (h < H) {
        for (int b = 0; b < B; b++) s = fmaf(dy[b * O + o], hid[b * H + h], s);
        gW2[o * H + h] = s;
    } else {
        for (int b = 0; b < B; b++) s += dy[b * O + o];
        gb2[o] = s;
    }
}

// dh[b][h] = [hid > 0] sum_o dy[b][o] W2[o][h]
__global__ void mlp_dh_kernel(const float* __restrict__ dy, const float* __restrict__ W2,
                              const float* __restrict__ hid, float* __restrict__ dh, int B, int H, int O) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= B * H) return;
    const int b = idx / H, h = idx % H;
    float s = 0.0f;
    if (hid[idx] > 0.0f)
        for (int o = 0; o < O; o++) s = fmaf(dy[b * O + o], W2[o * H + h], s);
    dh[idx] = s;
}

// gW1[h][i] = sum_b dh[b][h] x[b][i]; gb1[h] = sum_b dh[b][h]
__global__ void mlp_gw1_kernel(const float* __restrict__ dh, const float* __restrict__ x, float* __restrict__ gW1,
                               float* __restrict__ gb1, int B, int I, int H) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= H * (I + 1)) return;
    const int h = idx / (I + 1), i = idx % (I + 1);
    float s = 0.0f;
    if (i < I) {
        for (int b = 0; b < B; b++) s = fmaf(dh[b * H + h], x[b * I + i], s);
        gW1[h * I + i] = s;
    } else {
        for (int b = 0; b < B; b++) s += dh[b * H + h];
        gb1[h] = s;
    }
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace
}  // namespace lt

extern "C" {

int lthip_mlp_forward(const float* x, const float* W1, const float* b1, const float* W2, const float* b2, float* hidden,
                      float* y, int B, int I, int H, int O, void* stream) {
    using namespace lt;
    if (B < 0 || I <= 0 || H <= 0 || O <= 0) return fail(LTHIP_ERR_INVALID_ARG, "mlp: bad sizes");
    if (B == 0) return LTHIP_OK;
    if (!x || !W1 || !b1 || !W2 || !b2 || !hidden || !y) return fail(LTHIP_ERR_INVALID_ARG, "mlp: null pointer");
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(mlp_hidden_kernel, dim3(nblk((long)B * H)), dim3(256), 0, s, x, W1, b1, hidden, B, I, H);
    hipLaunchKernelGGL(mlp_out_kernel, dim3(nblk((long)B * O)), dim3(256), 0, s, hidden, W2, b2, y, B, H, O);
    LT_TRY_LAUNCH("mlp forward launch");
    return LTHIP_OK;
}

size_t lthip_mlp_workspace_size(int B, int H, int O) {
    return sizeof(float) * ((size_t)B * H * 2 + (size_t)B * O * 2 + 64);
}

int lthip_mlp_loss_grad(const float* x, const float* W1, const float* b1, const float* W2, const float* b2,
                        const float* target, const uint8_t* mask, int B, int I, int H, int O, float* gW1, float* gb1,
                        float* gW2, float* gb2, float* loss, void* ws, size_t ws_bytes, void* stream) {
    using namespace lt;
    if (B <= 0 || I <= 0 || H <= 0 || O <= 0) return fail(LTHIP_ERR_INVALID_ARG, "mlp: bad sizes");
    if (!x || !W1 || !b1 || !W2 || !b2 || !target || !mask || !gW1 || !gb1 || !gW2 || !gb2 || !loss)
        return fail(LTHIP_ERR_INVALID_ARG, "mlp: null pointer");
    const size_t need = lthip_mlp_workspace_size(B, H, O);
    if (!ws || ws_bytes < need) return fail(LTHIP_ERR_WORKSPACE, "mlp: need %zu workspace bytes", need);
    float* hid = reinterpret_cast<float*>(ws);
    float* dh = hid + (size_t)B * H;
    float* y = dh + (size_t)B * H;
    float* dy = y + (size_t)B * O;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(mlp_hidden_kernel, dim3(nblk((long)B * H)), dim3(256), 0, s, x, W1, b1, hid, B, I, H);
    hipLaunchKernelGGL(mlp_out_kernel, dim3(nblk((long)B * O)), dim3(256), 0, s, hid, W2, b2, y, B, H, O);
    hipLaunchKernelGGL(mlp_dy_kernel, dim3(1), dim3(256), 0, s, y, target, mask, dy, loss, B * O);
    hipLaunchKernelGGL(mlp_gw2_kernel, dim3(nblk((long)O * (H + 1))), dim3(256), 0, s, dy, hid, gW2, gb2, B, H, O);
    hipLaunchKernelGGL(mlp_dh_kernel, dim3(nblk((long)B * H)), dim3(256), 0, s, dy, W2, hid, dh, B, H, O);
    hipLaunchKernelGGL(mlp_gw1_kernel, dim3(nblk((long)H * (I + 1))), dim3(256), 0, s, dh, x, gW1, gb1, B, I, H);
    LT_TRY_LAUNCH("mlp backward launch");
    return LTHIP_OK;
}

int lthip_adam_step(float* params, const float* grads, float* m, float* v, int n, float lr, float beta1, float beta2,
                    float eps, int step, void* stream) {
    using namespace lt;
    if (n < 0 || step < 1) return fail(LTHIP_ERR_INVALID_ARG, "adam: bad n or step");
    if (n == 0) return LTHIP_OK;
    if (!params || !grads || !m || !v) return fail(LTHIP_ERR_INVALID_ARG, "adam: null pointer");
    const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(nblk(n)), dim3(256), 0, as_stream(stream), params, grads, m, v, n, lr, beta1,
                       beta2, eps, c1, c2);
    LT_TRY_LAUNCH("adam launch");
    return LTHIP_OK;
}

}  // extern "C"
