// sgemv: y = alpha * op(A) x + beta * y, FP32, HBM-bound.
//
// Row-major (the reference's sgemv_vecto layout, sgemv_vecto.cu:10-60): one
// wave per row, x staged once per block into LDS (which also absorbs incx),
// eight float4 loads of A in flight per lane (a ragged end of a row as one
// predicated batch), DPP wave sum.  Blocks stride over rows so each block
// stages x once for several rows; on large matrices the grid gives every wave
// the same number of rows with about 1536 waves in all (see row_blocks).
//
// Column-major (sgemv_col3_vecto.cu:7-66 computes the same thing): each lane
// owns four consecutive rows and walks a slice of the columns with eight
// column loads in flight; column slices write FP32 partials to the workspace
// and a second kernel adds them in slice order (deterministic, no atomics) and
// applies alpha/beta.  The slice count aims at about 192 blocks: enough
// streams to keep HBM busy, few enough that each block walks long runs of
// columns (more, shorter streams measured slower at every large shape).
#include "lt_common.hpp"

namespace lt {
namespace {

constexpr int kRowWaves = 8;
constexpr int kRowBlock = kRowWaves * kWave;
constexpr int kRowUnroll = 8;
constexpr int kMaxLdsX = 16384;   // floats of x staged in LDS (64 KB)

// TAIL: n/4 is not a multiple of the eight-chunk step, so each row ends with
// a predicated batch of up to seven chunks (without it, rows whose length is a
// multiple of the step skip that code: measured faster for short rows)
template <bool VEC, bool XLDS, bool TAIL>
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
            // fewer than kRowUnroll chunks left: issue them all, then use them
            if (TAIL) {
                float4 av[kRowUnroll];
#pragma unroll
                for (int u = 0; u < kRowUnroll; u++)
                    av[u] = c + u * kWave < n4 ? ld_nt(&a4[c + u * kWave]) : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                for (int u = 0; u < kRowUnroll; u++) {
                    if (c + u * kWave < n4) {
                        const float4 xv = XLDS ? xs4[c + u * kWave] : xg4[c + u * kWave];
                        acc[u] = fmaf(av[u].x, xv.x, acc[u]);
                        acc[u] = fmaf(av[u].y, xv.y, acc[u]);
                        acc[u] = fmaf(av[u].z, xv.z, acc[u]);
                        acc[u] = fmaf(av[u].w, xv.w, acc[u]);
                    }
                }
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

// Blocks for the row kernel.  Matrices of 128 MB and more, with x in LDS:
// about 1536 waves
// (12 MB of loads in flight at eight float4 per lane), every wave given the
// same number of rows, since the waves with one row more than the others set
// the time; this measured faster than a grid that fills every CU (fewer,
// longer streams).  Rows shorter than eight chunks of 64 float4 keep fewer
// loads in flight per wave and get proportionally more waves.  Smaller
// matrices are latency-bound, and rows too long for LDS reread x from L2
// beside A: one row per wave, up to four blocks per CU.
long row_blocks(int m, int n, bool xlds) {
    if (!xlds || (double)m * (double)n * 4.0 < 128.0 * 1024 * 1024) {
        const long blocks = (m + kRowWaves - 1) / kRowWaves;
        const long cap = (long)cu_count() * 4;
        return blocks < cap ? blocks : cap;
    }
    int chunks = (n / 4 + kWave - 1) / kWave;   // float4 loads per lane and row
    if (chunks > kRowUnroll) chunks = kRowUnroll;
    if (chunks < 1) chunks = 1;
    const long waves = 1536L * kRowUnroll / chunks;
    long rows = (long)((double)m / (double)waves + 0.5);
    if (rows < 1) rows = 1;
    return (m + kRowWaves * rows - 1) / (kRowWaves * rows);
}

constexpr int kColBlock = 256;
constexpr int kColRows = kColBlock * 4;   // rows per block (4 per lane)
constexpr int kColUnroll = 8;             // column loads in flight per lane
constexpr int kColTargetBlocks = 192;
constexpr int kColMinCols = 64;           // columns per slice, at least (when n allows)
constexpr int kColMaxSplit = 256;

int col_splits(int m, int n) {
    const int rowblocks = (m + kColRows - 1) / kColRows;
    int s = (kColTargetBlocks + rowblocks - 1) / rowblocks;
    const int cap = n / kColMinCols;
    if (s > cap) s = cap;
    if (s > kColMaxSplit) s = kColMaxSplit;
    if (s < 1) s = 1;
    return s;
}

// partial[split][row] = sum over the split's columns of A[row, j] x[j], in
// column order.  XU: incx == 1, so the eight x values of a step come from one
// contiguous scalar load
template <bool VEC, bool XU>
__global__ __launch_bounds__(kColBlock) void sgemv_col_partial_kernel(int m, int n, const float* __restrict__ A,
                                                                      long lda, const float* __restrict__ x, long incx,
                                                                      int splits, float* __restrict__ partial) {
    const int split = blockIdx.y;
    const int j0 = (int)((long)n * split / splits), j1 = (int)((long)n * (split + 1) / splits);
    const long i0 = (long)blockIdx.x * kColRows + 4L * threadIdx.x;
    if (i0 >= m) return;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    const bool full = VEC && (i0 + 3 < m);
    int j = j0;
    if (full) {
        for (; j + kColUnroll <= j1; j += kColUnroll) {
            float4 a[kColUnroll];
            float xs[kColUnroll];
#pragma unroll
            for (int u = 0; u < kColUnroll; u++) {
                xs[u] = XU ? x[j + u] : x[(long)(j + u) * incx];
                a[u] = ld_nt(reinterpret_cast<const float4*>(A + i0 + (long)(j + u) * lda));
            }
#pragma unroll
            for (int u = 0; u < kColUnroll; u++) {
                acc.x = fmaf(a[u].x, xs[u], acc.x); acc.y = fmaf(a[u].y, xs[u], acc.y);
                acc.z = fmaf(a[u].z, xs[u], acc.z); acc.w = fmaf(a[u].w, xs[u], acc.w);
            }
        }
        for (; j < j1; j++) {
            const float xa = x[(long)j * incx];
            const float4 a = *reinterpret_cast<const float4*>(A + i0 + (long)j * lda);
            acc.x = fmaf(a.x, xa, acc.x); acc.y = fmaf(a.y, xa, acc.y);
            acc.z = fmaf(a.z, xa, acc.z); acc.w = fmaf(a.w, xa, acc.w);
        }
    } else {
        for (; j < j1; j++) {
            const float xa = x[(long)j * incx];
            const float* col = A + (long)j * lda;
            acc.x = fmaf(col[i0], xa, acc.x);
            if (i0 + 1 < m) acc.y = fmaf(col[i0 + 1], xa, acc.y);
            if (i0 + 2 < m) acc.z = fmaf(col[i0 + 2], xa, acc.z);
            if (i0 + 3 < m) acc.w = fmaf(col[i0 + 3], xa, acc.w);
        }
    }
    float* out = partial + (long)split * m + i0;
    out[0] = acc.x;
    if (i0 + 1 < m) out[1] = acc.y;
    if (i0 + 2 < m) out[2] = acc.z;
    if (i0 + 3 < m) out[3] = acc.w;
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
    if (row_major || m <= 0 || n < 0) return 0;
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
        const long blocks = row_blocks(m, n, xlds);
        const size_t lds = xlds ? sizeof(float) * (size_t)((n + 3) & ~3) : 0;
        const bool tail = (n / 4) % (kRowUnroll * kWave) != 0;
#define LT_ROW(V, X, T)                                                                                              \
    hipLaunchKernelGGL((sgemv_row_kernel<V, X, T>), dim3((unsigned)blocks), dim3(kRowBlock), lds, s, m, n,         \
                       alpha, A, (long)lda, x, (long)incx, beta, y, (long)incy)
        if (vec && xvec_ok) {
            if (xlds) {
                if (tail) LT_ROW(true, true, true);
                else LT_ROW(true, true, false);
            } else {
                if (tail) LT_ROW(true, false, true);
                else LT_ROW(true, false, false);
            }
        } else {
            if (xlds) LT_ROW(false, true, false);
            else LT_ROW(false, false, false);
        }
#undef LT_ROW
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
    if (vec && incx == 1)
        hipLaunchKernelGGL((sgemv_col_partial_kernel<true, true>), grid, dim3(kColBlock), 0, s, m, n, A, (long)lda,
                           x, (long)incx, splits, partial);
    else if (vec)
        hipLaunchKernelGGL((sgemv_col_partial_kernel<true, false>), grid, dim3(kColBlock), 0, s, m, n, A, (long)lda,
                           x, (long)incx, splits, partial);
    else
        hipLaunchKernelGGL((sgemv_col_partial_kernel<false, false>), grid, dim3(kColBlock), 0, s, m, n, A, (long)lda,
                           x, (long)incx, splits, partial);
    LT_TRY_LAUNCH("sgemv col partial launch");
    hipLaunchKernelGGL(sgemv_col_finish_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, m, splits, partial,
                       alpha, beta, y, (long)incy);
    LT_TRY_LAUNCH("sgemv col finish launch");
    return LTHIP_OK;
}

}  // extern "C"
