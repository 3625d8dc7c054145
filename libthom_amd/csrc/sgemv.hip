// sgemv: y = alpha * op(A) x + beta * y, FP32, HBM-bound.
//
// Row-major (the reference's sgemv_vecto layout, sgemv_vecto.cu:10-60): one
// wave per row, x staged once per block into LDS (which also absorbs incx),
// eight float4 loads of A in flight per lane, DPP wave sum.  Blocks stride over
// rows so each block stages x once for several rows.
//
// Column-major (sgemv_col3_vecto.cu:7-66 computes the same thing): each lane
// owns four consecutive rows and walks a slice of the columns; column slices
// write FP32 partials to the workspace and a second kernel adds them in slice
// order (deterministic, no atomics) and applies alpha/beta.
#include "lt_common.hpp"

namespace lt {
namespace {

constexpr int kRowWaves = 8;
constexpr int kRowBlock = kRowWaves * kWave;
constexpr int kRowUnroll = 8;
constexpr int kMaxLdsX = 16384;   // floats of x staged in LDS (64 KB)

template <bool VEC, bool XLDS>
__global__ __launch_bounds__(kRowBlock) void sgemv_row_kernel(int m, int n, const float* __restrict__ alpha_p,
                                                              const float* __restrict__ A, long lda,
                                                              const float* __restrict__ x, long incx,
                                                              const float* __restrict__ beta_p, float* __restrict__ y,
                                                              long incy) {
    extern __shared__ float4 xs4[];
    float* xs = reinterpret_cast<float*>(xs4);
    if (XLDS) {
        for (int j = threadIdx.x; j < n; j += kRowBlock) xs[j] = x[(long)j * incx];
        __syncthreads();
    }
    const float alpha = *alpha_p, beta = *beta_p;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    for (long row = (long)blockIdx.x * kRowWaves + wave; row < m; row += (long)gridDim.x * kRowWaves) {
        const float* arow = A + row * lda;
        float acc[kRowUnroll];
#pragma unroll
        for (int u = 0; u < kRowUnroll; u++) acc[u] = 0.0f;
        if (VEC) {
            const float4* a4 = reinterpret_cast<const float4*>(arow);
            const float4* xg4 = reinterpret_cast<const float4*>(x);
            const int n4 = n >> 2;
            int c = lane;
            for (; c + (kRowUnroll - 1) * kWave < n4; c += kRowUnroll * kWave) {
                float4 av[kRowUnroll];
#pragma unroll
                for (int u = 0; u < kRowUnroll; u++) av[u] = ld_nt(&a4[c + u * kWave]);
#pragma unroll
                for (int u = 0; u < kRowUnroll; u++) {
                    const float4 xv = XLDS ? xs4[c + u * kWave] : xg4[c + u * kWave];
                    acc[u] = fmaf(av[u].x, xv.x, acc[u]);
                    acc[u] = fmaf(av[u].y, xv.y, acc[u]);
                    acc[u] = fmaf(av[u].z, xv.z, acc[u]);
                    acc[u] = fmaf(av[u].w, xv.w, acc[u]);
                }
            }
            for (; c < n4; c += kWave) {
                const float4 av = a4[c];
                const float4 xv = XLDS ? xs4[c] : xg4[c];
                acc[0] = fmaf(av.x, xv.x, acc[0]);
                acc[0] = fmaf(av.y, xv.y, acc[0]);
                acc[0] = fmaf(av.z, xv.z, acc[0]);
                acc[0] = fmaf(av.w, xv.w, acc[0]);
            }
            const int t = 4 * n4 + lane;
            if (t < n) acc[1] = fmaf(arow[t], XLDS ? xs[t] : x[t], acc[1]);
        } else {
            for (int c = lane; c < n; c += kWave) acc[c & 1] = fmaf(arow[c], XLDS ? xs[c] : x[(long)c * incx], acc[c & 1]);
        }
        float s = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
        s = wave_sum(s);
        if (lane == 0) {
            float r = alpha * s;
            if (beta != 0.0f) r = fmaf(beta, y[row * incy], r);
            y[row * incy] = r;
        }
    }
}

constexpr int kColBlock = 256;
constexpr int kColRows = kColBlock * 4;   // rows per block (4 per lane)
constexpr int kColMaxSplit = 64;

int col_splits(int m, int n) {
    (void)m;
    int s = n / 64;
    if (s > kColMaxSplit) s = kColMaxSplit;
    if (s < 1) s = 1;
    return s;
}

// partial[split][row] = sum over the split's columns of A[row, j] x[j]
template <bool VEC>
__global__ __launch_bounds__(kColBlock) void sgemv_col_partial_kernel(int m, int n, const float* __restrict__ A,
                                                                      long lda, const float* __restrict__ x, long incx,
                                                                      int splits, float* __restrict__ partial) {
    const int split = blockIdx.y;
    const int j0 = (int)((long)n * split / splits), j1 = (int)((long)n * (split + 1) / splits);
    const long i0 = (long)blockIdx.x * kColRows + 4L * threadIdx.x;
    if (i0 >= m) return;
    float4 acc0 = make_float4(0.f, 0.f, 0.f, 0.f), acc1 = acc0;
    const bool full = VEC && (i0 + 3 < m);
    int j = j0;
    if (full) {
        for (; j + 1 < j1; j += 2) {
            const float xa = x[(long)j * incx], xb = x[(long)(j + 1) * incx];
            const float4 a = ld_nt(reinterpret_cast<const float4*>(A + i0 + (long)j * lda));
            const float4 b = ld_nt(reinterpret_cast<const float4*>(A + i0 + (long)(j + 1) * lda));
            acc0.x = fmaf(a.x, xa, acc0.x); acc0.y = fmaf(a.y, xa, acc0.y);
            acc0.z = fmaf(a.z, xa, acc0.z); acc0.w = fmaf(a.w, xa, acc0.w);
            acc1.x = fmaf(b.x, xb, acc1.x); acc1.y = fmaf(b.y, xb, acc1.y);
            acc1.z = fmaf(b.z, xb, acc1.z); acc1.w = fmaf(b.w, xb, acc1.w);
        }
        for (; j < j1; j++) {
            const float xa = x[(long)j * incx];
            const float4 a = *reinterpret_cast<const float4*>(A + i0 + (long)j * lda);
            acc0.x = fmaf(a.x, xa, acc0.x); acc0.y = fmaf(a.y, xa, acc0.y);
            acc0.z = fmaf(a.z, xa, acc0.z); acc0.w = fmaf(a.w, xa, acc0.w);
        }
    } else {
        for (; j < j1; j++) {
            const float xa = x[(long)j * incx];
            const float* col = A + (long)j * lda;
            acc0.x = fmaf(col[i0], xa, acc0.x);
            if (i0 + 1 < m) acc0.y = fmaf(col[i0 + 1], xa, acc0.y);
            if (i0 + 2 < m) acc0.z = fmaf(col[i0 + 2], xa, acc0.z);
            if (i0 + 3 < m) acc0.w = fmaf(col[i0 + 3], xa, acc0.w);
        }
    }
    float* out = partial + (long)split * m + i0;
    out[0] = acc0.x + acc1.x;
    if (i0 + 1 < m) out[1] = acc0.y + acc1.y;
    if (i0 + 2 < m) out[2] = acc0.z + acc1.z;
    if (i0 + 3 < m) out[3] = acc0.w + acc1.w;
}

__global__ __launch_bounds__(256) void sgemv_col_finish_kernel(int m, int splits, const float* __restrict__ partial,
                                                               const float* __restrict__ alpha_p,
                                                               const float* __restrict__ beta_p, float* __restrict__ y,
                                                               long incy) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    float s = 0.0f;
    for (int k = 0; k < splits; k++) s += partial[(long)k * m + i];
    const float alpha = *alpha_p, beta = *beta_p;
    float r = alpha * s;
    if (beta != 0.0f) r = fmaf(beta, y[i * incy], r);
    y[i * incy] = r;
}

}  // namespace
}  // namespace lt

extern "C" {

size_t lthip_sgemv_f32_workspace_size(int row_major, int m, int n) {
    if (row_major || m <= 0 || n <= 0) return 0;
    return sizeof(float) * (size_t)lt::col_splits(m, n) * (size_t)m;
}

int lthip_sgemv_f32(int row_major, int m, int n, const float* alpha, const float* A, int lda, const float* x,
                    int incx, const float* beta, float* y, int incy, void* ws, size_t ws_bytes, void* stream) {
    using namespace lt;
    if (m < 0 || n < 0) return fail(LTHIP_ERR_INVALID_ARG, "sgemv: negative size");
    if (incx <= 0 || incy <= 0) return fail(LTHIP_ERR_INVALID_ARG, "sgemv: incx and incy must be positive");
    if (m == 0) return LTHIP_OK;
    if (!alpha || !beta || !y || (n > 0 && (!A || !x))) return fail(LTHIP_ERR_INVALID_ARG, "sgemv: null pointer");
    if (row_major ? (lda < (n > 1 ? n : 1)) : (lda < m)) return fail(LTHIP_ERR_INVALID_ARG, "sgemv: lda too small");
    hipStream_t s = as_stream(stream);
    if (row_major) {
        const bool vec = (reinterpret_cast<uintptr_t>(A) % 16 == 0) && (lda % 4 == 0);
        const bool xlds = n <= kMaxLdsX;
        const bool xvec_ok = xlds || (incx == 1 && reinterpret_cast<uintptr_t>(x) % 16 == 0);
        long blocks = (m + kRowWaves - 1) / kRowWaves;
        const long cap = (long)cu_count() * 4;
        if (blocks > cap) blocks = cap;
        const size_t lds = xlds ? sizeof(float) * (size_t)((n + 3) & ~3) : 0;
        if (vec && xvec_ok) {
            if (xlds)
                hipLaunchKernelGGL((sgemv_row_kernel<true, true>), dim3((unsigned)blocks), dim3(kRowBlock), lds, s, m, n,
                                   alpha, A, (long)lda, x, (long)incx, beta, y, (long)incy);
            else
                hipLaunchKernelGGL((sgemv_row_kernel<true, false>), dim3((unsigned)blocks), dim3(kRowBlock), 0, s, m, n,
                                   alpha, A, (long)lda, x, (long)incx, beta, y, (long)incy);
        } else {
            if (xlds)
                hipLaunchKernelGGL((sgemv_row_kernel<false, true>), dim3((unsigned)blocks), dim3(kRowBlock), lds, s, m,
                                   n, alpha, A, (long)lda, x, (long)incx, beta, y, (long)incy);
            else
                hipLaunchKernelGGL((sgemv_row_kernel<false, false>), dim3((unsigned)blocks), dim3(kRowBlock), 0, s, m,
                                   n, alpha, A, (long)lda, x, (long)incx, beta, y, (long)incy);
        }
        LT_TRY_LAUNCH("sgemv row launch");
        return LTHIP_OK;
    }
    // column-major
    const int splits = col_splits(m, n);
    const size_t need = lthip_sgemv_f32_workspace_size(0, m, n);
    if (!ws || ws_bytes < need) return fail(LTHIP_ERR_WORKSPACE, "sgemv col-major: need %zu workspace bytes", need);
    float* partial = reinterpret_cast<float*>(ws);
    dim3 grid((unsigned)((m + kColRows - 1) / kColRows), (unsigned)splits);
    const bool vec = (reinterpret_cast<uintptr_t>(A) % 16 == 0) && (lda % 4 == 0);
    if (vec)
        hipLaunchKernelGGL((sgemv_col_partial_kernel<true>), grid, dim3(kColBlock), 0, s, m, n, A, (long)lda, x,
                           (long)incx, splits, partial);
    else
        hipLaunchKernelGGL((sgemv_col_partial_kernel<false>), grid, dim3(kColBlock), 0, s, m, n, A, (long)lda, x,
                           (long)incx, splits, partial);
    LT_TRY_LAUNCH("sgemv col partial launch");
    hipLaunchKernelGGL(sgemv_col_finish_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, m, splits, partial,
                       alpha, beta, y, (long)incy);
    LT_TRY_LAUNCH("sgemv col finish launch");
    return LTHIP_OK;
}

}  // extern "C"
