// HBM-streaming kernels: input generation, copy, reduction and sdot.
//
// reduction / sdot are one launch each: a grid-stride float4 stream with four
// independent loads in flight per lane, a DPP wave sum, an LDS block sum, and a
// last-block-done finish (the block that takes the final ticket of an agent-
// scope counter sums the per-block partials in block order).  The reference
// needs two launches plus a cudaMalloc inside its timed window
// (reduction_warp.cu:94-105, sdot_warp.cu:117-136).
#include "lt_common.hpp"

#include <mutex>

namespace lt {
namespace {

constexpr int kRBlock = 256;
constexpr int kRUnroll = 4;
constexpr unsigned kRMaxBlocks = 4096;
constexpr size_t kRWorkspace = 256 + sizeof(float) * kRMaxBlocks;   // counter + partials

// ---- SplitMix64 fill ---------------------------------------------------
__device__ __forceinline__ uint64_t sm_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double uniform_at(uint64_t base, size_t i, double lo, double span) {
    const uint64_t x = sm_mix(base + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull);
    const double u = (double)(x >> 40) * 0x1p-24;
    return __dadd_rn(lo, __dmul_rn(span, u));   // no contraction: matches the host spec
}

template <typename T>
__device__ __forceinline__ T from_double(double v);
template <>
__device__ __forceinline__ float from_double<float>(double v) { return (float)v; }
template <>
__device__ __forceinline__ double from_double<double>(double v) { return v; }
template <>
__device__ __forceinline__ uint16_t from_double<uint16_t>(double v) {
    return __half_as_ushort(__float2half_rn((float)v));
}

template <typename T>
__global__ __launch_bounds__(256) void fill_uniform_kernel(T* __restrict__ out, size_t n, double lo, double span,
                                                           uint64_t base) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = from_double<T>(uniform_at(base, i, lo, span));
}

template <typename T>
int fill_uniform(T* out, size_t n, double lo, double hi, uint64_t seed, uint64_t stream_id, void* stream) {
    if (n == 0) return LTHIP_OK;
    if (!out) return fail(LTHIP_ERR_INVALID_ARG, "fill_uniform: null output");
    const uint64_t base = seed + stream_id * 0xD1B54A32D192ED03ull;
    size_t blocks = (n + 255) / 256;
    const size_t cap = (size_t)cu_count() * 16;
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(fill_uniform_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), out, n, lo,
                       hi - lo, base);
    LT_TRY_LAUNCH("fill_uniform launch");
    return LTHIP_OK;
}

// ---- copy ---------------------------------------------------------------
__global__ __launch_bounds__(256) void copy_kernel(float4* __restrict__ dst, const float4* __restrict__ src, size_t n4,
                                                   float* __restrict__ dtail, const float* __restrict__ stail,
                                                   size_t ntail) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t i = tid;
    for (; i + 3 * stride < n4; i += 4 * stride) {
        float4 a = ld_nt(&src[i]);
        float4 b = ld_nt(&src[i + stride]);
        float4 c = ld_nt(&src[i + 2 * stride]);
        float4 d = ld_nt(&src[i + 3 * stride]);
        st_nt(&dst[i], a);
        st_nt(&dst[i + stride], b);
        st_nt(&dst[i + 2 * stride], c);
        st_nt(&dst[i + 3 * stride], d);
    }
    for (; i < n4; i += stride) dst[i] = src[i];
    if (tid < ntail) dtail[tid] = stail[tid];
}

// ---- reduction / dot ---------------------------------------------------
template <bool DOT>
__device__ __forceinline__ float term4(float4 a, float4 b) {
    if (DOT) return fmaf(a.x, b.x, a.y * b.y) + fmaf(a.z, b.z, a.w * b.w);
    return (a.x + a.y) + (a.z + a.w);
}

// VEC: x (and y) 16-byte aligned, float4 path with a scalar tail.  BLOCK
// threads per block: 256 for the streaming grid, 512 for small n (fewer blocks,
// so fewer arrivals queue on the one ticket counter: those atomics serialise at
// the counter's address, ~6 ns each, which at n = 1e6 with 245 blocks was
// 1.5 of the kernel's 5.4 us).
template <bool DOT, bool VEC, int BLOCK>
__global__ __launch_bounds__(BLOCK) void sum_kernel(const float* __restrict__ x, const float* __restrict__ y, size_t n,
                                                    float* __restrict__ out, unsigned* __restrict__ counter,
                                                    float* __restrict__ partials) {
    constexpr int kRBlock = BLOCK;
    __shared__ float red[kRBlock / kWave];
    __shared__ int is_last;
    const size_t nthreads = (size_t)gridDim.x * kRBlock;
    const size_t tid = (size_t)blockIdx.x * kRBlock + threadIdx.x;
    float acc[kRUnroll] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (VEC) {
        const size_t n4 = n >> 2;
        const float4* x4 = reinterpret_cast<const float4*>(x);
        const float4* y4 = reinterpret_cast<const float4*>(y);
        size_t i = tid;
        for (; i + (kRUnroll - 1) * nthreads < n4; i += kRUnroll * nthreads) {
            float4 a[kRUnroll], b[kRUnroll];
#pragma unroll
            for (int u = 0; u < kRUnroll; u++) {
                a[u] = ld_nt(&x4[i + u * nthreads]);
                if (DOT) b[u] = ld_nt(&y4[i + u * nthreads]);
            }
#pragma unroll
            for (int u = 0; u < kRUnroll; u++) acc[u] += term4<DOT>(a[u], DOT ? b[u] : a[u]);
        }
        for (; i < n4; i += nthreads) {
            float4 a = x4[i];
            acc[0] += term4<DOT>(a, DOT ? y4[i] : a);
        }
        const size_t t = 4 * n4 + tid;   // n % 4 leftover elements, one per thread
        if (t < n) acc[1] += DOT ? x[t] * y[t] : x[t];
    } else {
        for (size_t i = tid; i < n; i += nthreads) acc[0] += DOT ? x[i] * y[i] : x[i];
    }
    float s = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    s = block_sum<kRBlock>(s, red);
    if (threadIdx.x == 0) {
        publish(&partials[blockIdx.x], s);
        drain();
        is_last = (take_ticket(counter) == gridDim.x - 1);
    }
    __syncthreads();
    if (!is_last) return;
    // this thread's block partials b = tid + kRBlock q, q ascending, all loads
    // in flight at once (kRMaxBlocks / kRBlock of them at most)
    constexpr int Q = (kRMaxBlocks + kRBlock - 1) / kRBlock;
    float v[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const unsigned b = threadIdx.x + kRBlock * q;
        v[q] = collect(&partials[b < gridDim.x ? b : 0]);
    }
    float p = 0.0f;
#pragma unroll
    for (int q = 0; q < Q; q++)
        if (threadIdx.x + kRBlock * q < gridDim.x) p += v[q];
    p = block_sum<kRBlock>(p, red);
    if (threadIdx.x == 0) {
        *out = p;
        reset_counter(counter);
    }
}

// Inputs up to this many floats take 512-thread blocks, one unrolled step
// per thread (n = 1e6: 123 blocks; 4.45 against 5.4 us from the first wave to
// the result with 245 blocks of 256).
constexpr size_t kSmallSum = (size_t)2 << 20;
constexpr int kSmallBlock = 512;

// Blocks for a sum over n floats.  Each thread's kRUnroll loads (2 kRUnroll
// for a dot) are one grid-width apart, so the grid sets the stride between
// the concurrent requests: 1.5 blocks per CU for the sum (1.5 MB apart on 256
// CUs) and 3 per CU for the dot measured 7.0-7.35 and 6.8-7.0 TB/s from 5e7
// to 1e9 floats, against 5.4-6.0 at 8 blocks per CU (8 MB apart: the
// requests pile onto the same banks).
unsigned sum_blocks(size_t n, bool dot) {
    const int block = n <= kSmallSum ? kSmallBlock : kRBlock;
    const size_t per_block = (size_t)block * 4 * kRUnroll;   // floats one block covers in one unrolled step
    size_t b = (n + per_block - 1) / per_block;
    const size_t cap = dot ? (size_t)cu_count() * 3 : (size_t)cu_count() * 3 / 2;
    if (b > cap) b = cap;
    if (b > kRMaxBlocks) b = kRMaxBlocks;
    if (b < 1) b = 1;
    return (unsigned)b;
}

template <bool DOT>
int launch_sum(const float* x, const float* y, size_t n, float* out, void* ws, size_t ws_bytes, void* stream) {
    if (!out) return fail(LTHIP_ERR_INVALID_ARG, "null output pointer");
    if (n > 0 && (!x || (DOT && !y))) return fail(LTHIP_ERR_INVALID_ARG, "null input pointer");
    if (!ws || ws_bytes < kRWorkspace) return fail(LTHIP_ERR_WORKSPACE, "need %zu workspace bytes", kRWorkspace);
    if (reinterpret_cast<uintptr_t>(ws) % 16 != 0)
        return fail(LTHIP_ERR_INVALID_ARG, "workspace must be 16-byte aligned");
    unsigned* counter = reinterpret_cast<unsigned*>(ws);
    float* partials = reinterpret_cast<float*>(reinterpret_cast<char*>(ws) + 256);
    const unsigned blocks = sum_blocks(n, DOT);
    if (256 + sizeof(float) * (size_t)blocks > ws_bytes)   // the counter head and one partial per block
        return fail(LTHIP_ERR_WORKSPACE, "%u blocks need %zu workspace bytes, %zu given", blocks,
                    256 + sizeof(float) * (size_t)blocks, ws_bytes);
    const bool vec = (reinterpret_cast<uintptr_t>(x) % 16 == 0) && (!DOT || reinterpret_cast<uintptr_t>(y) % 16 == 0);
    hipStream_t s = as_stream(stream);
    if (n <= kSmallSum) {
        if (vec)
            hipLaunchKernelGGL((sum_kernel<DOT, true, kSmallBlock>), dim3(blocks), dim3(kSmallBlock), 0, s, x, y, n,
                               out, counter, partials);
        else
            hipLaunchKernelGGL((sum_kernel<DOT, false, kSmallBlock>), dim3(blocks), dim3(kSmallBlock), 0, s, x, y, n,
                               out, counter, partials);
    } else if (vec) {
        hipLaunchKernelGGL((sum_kernel<DOT, true, kRBlock>), dim3(blocks), dim3(kRBlock), 0, s, x, y, n, out, counter,
                           partials);
    } else {
        hipLaunchKernelGGL((sum_kernel<DOT, false, kRBlock>), dim3(blocks), dim3(kRBlock), 0, s, x, y, n, out, counter,
                           partials);
    }
    LT_TRY_LAUNCH(DOT ? "sdot launch" : "reduction launch");
    return LTHIP_OK;
}

// ---- allocation-free reduction -------------------------------------------
// The arrival counter (first 256 bytes) and the block partials of sum_kernel
// for callers that cannot allocate.  Each GPU holds its own copy at its own
// address, found once per device with hipGetSymbolAddress; host code uses that
// address, never this array's name.  Static storage starts zeroed, the first
// call on a device clears the counter head again, and every launch leaves the
// counter at zero (the last block resets it).  All calls on one GPU share the
// array: calls that overlap on one GPU race.
constexpr size_t kScratchBytes = kRWorkspace;   // 256 + 4 * 4096 = 16,640 bytes per GPU
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;

__global__ void scratch_head_clear_kernel(unsigned* head) { head[threadIdx.x] = 0u; }

// This device's scratch address (set up on the first call with the device current).
int scratch_on_device(void** out, void* stream) {
    static std::mutex mu;
    static void* ptrs[kMaxDevices] = {nullptr};
    int dev = 0;
    int st = check_hip(hipGetDevice(&dev), "reduce_sum_noalloc: hipGetDevice");
    if (st != LTHIP_OK) return st;
    if (dev < 0 || dev >= kMaxDevices) return fail(LTHIP_ERR_UNSUPPORTED, "reduce_sum_noalloc: device id %d", dev);
    std::lock_guard<std::mutex> lock(mu);
    if (!ptrs[dev]) {
        void* p = nullptr;
        st = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "reduce_sum_noalloc: hipGetSymbolAddress");
        if (st != LTHIP_OK) return st;
        // a small kernel, not hipMemsetAsync, so the clear is safe under stream capture
        hipLaunchKernelGGL(scratch_head_clear_kernel, dim3(1), dim3(64), 0, as_stream(stream),
                           reinterpret_cast<unsigned*>(p));
        LT_TRY_LAUNCH("reduce_sum_noalloc: scratch clear");
        ptrs[dev] = p;
    }
    *out = ptrs[dev];
    return LTHIP_OK;
}

}  // namespace
}  // namespace lt

extern "C" {

int lthip_fill_uniform_f32(float* out, size_t n, double lo, double hi, uint64_t seed, uint64_t stream_id,
                           void* stream) {
    return lt::fill_uniform<float>(out, n, lo, hi, seed, stream_id, stream);
}
int lthip_fill_uniform_f16(uint16_t* out, size_t n, double lo, double hi, uint64_t seed, uint64_t stream_id,
                           void* stream) {
    return lt::fill_uniform<uint16_t>(out, n, lo, hi, seed, stream_id, stream);
}
int lthip_fill_uniform_f64(double* out, size_t n, double lo, double hi, uint64_t seed, uint64_t stream_id,
                           void* stream) {
    return lt::fill_uniform<double>(out, n, lo, hi, seed, stream_id, stream);
}

int lthip_copy_f32(float* dst, const float* src, size_t n, void* stream) {
    if (n == 0) return LTHIP_OK;
    if (!dst || !src) return lt::fail(LTHIP_ERR_INVALID_ARG, "copy: null pointer");
    if (reinterpret_cast<uintptr_t>(dst) % 16 || reinterpret_cast<uintptr_t>(src) % 16)
        return lt::fail(LTHIP_ERR_INVALID_ARG, "copy: pointers must be 16-byte aligned");
    const size_t n4 = n / 4, ntail = n % 4;
    size_t blocks = (n4 + 1023) / 1024;
    const size_t cap = (size_t)lt::cu_count() * 8;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(lt::copy_kernel, dim3((unsigned)blocks), dim3(256), 0, lt::as_stream(stream),
                       reinterpret_cast<float4*>(dst), reinterpret_cast<const float4*>(src), n4, dst + 4 * n4,
                       src + 4 * n4, ntail);
    LT_TRY_LAUNCH("copy launch");
    return LTHIP_OK;
}

size_t lthip_reduce_sum_f32_workspace_size(size_t) { return lt::kRWorkspace; }
size_t lthip_sdot_f32_workspace_size(size_t) { return lt::kRWorkspace; }

int lthip_reduce_sum_f32(const float* x, size_t n, float* out, void* ws, size_t ws_bytes, void* stream) {
    return lt::launch_sum<false>(x, nullptr, n, out, ws, ws_bytes, stream);
}

int lthip_sdot_f32(size_t n, const float* x, const float* y, float* result, void* ws, size_t ws_bytes,
                   void* stream) {
    return lt::launch_sum<true>(x, y, n, result, ws, ws_bytes, stream);
}

int lthip_reduce_sum_f32_noalloc(const float* x, size_t n, float* out, void* stream) {
    void* s = nullptr;
    const int st = lt::scratch_on_device(&s, stream);
    if (st != LTHIP_OK) return st;
    return lt::launch_sum<false>(x, nullptr, n, out, s, lt::kScratchBytes, stream);
}

}  // extern "C"
