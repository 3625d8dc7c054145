// Peak probes: the MFMA throughput and HBM read bandwidth this device
// actually delivers, the denominators of "fraction of measured peak".
//
// MFMA probe: every wave runs independent accumulator chains of one MFMA
// instruction for `iters` rounds on register operands (no memory
// traffic); one value per thread is written at the end so nothing is
// optimised away.  Read probe: a grid-stride float4 stream whose per-thread
// XOR of the loaded bits is written out.
#include "lt_common.hpp"

namespace lt {
namespace {

constexpr int kChains = 8;
typedef int v8i_t __attribute__((ext_vector_type(8)));

__global__ __launch_bounds__(256) void probe_f16(float* sink, int iters) {
    v16f acc[kChains];
#pragma unroll
    for (int c = 0; c < kChains; c++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[c][r] = 0.0f;
    v8h a, b;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        a[i] = (_Float16)(threadIdx.x * 1e-3f + i);
        b[i] = (_Float16)(0.5f - i * 1e-2f);
    }
    for (int it = 0; it < iters; it++)
#pragma unroll
        for (int c = 0; c < kChains; c++) acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, acc[c], 0, 0, 0);
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < kChains; c++)
#pragma unroll
        for (int r = 0; r < 16; r++) s += acc[c][r];
    sink[blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void probe_f8(float* sink, int iters) {
    v16f acc[kChains / 2];
#pragma unroll
    for (int c = 0; c < kChains / 2; c++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[c][r] = 0.0f;
    v8i_t a, b;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        a[i] = 0x38303830 ^ (int)(threadIdx.x & 7);   // finite e4m3 bytes
        b[i] = 0x30383038 ^ i;
    }
    for (int it = 0; it < iters; it++)
#pragma unroll
        for (int c = 0; c < kChains / 2; c++)
            acc[c] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, acc[c], 0, 0, 0, 0, 0, 0);
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < kChains / 2; c++)
#pragma unroll
        for (int r = 0; r < 16; r++) s += acc[c][r];
    sink[blockIdx.x * 256 + threadIdx.x] = s;
}

// FP64: a GEMM-like mix, 4 A x 4 B operand registers feeding 16 independent
// accumulators, in 512-thread blocks (8 waves).  At 512 threads the compiler
// keeps the 16 accumulators in VGPRs; with 256-thread blocks it moves them
// through AGPRs every round, and a single operand pair with 8 accumulators
// measured only ~49 of the 78.6 TFLOP/s.  This form measures ~77.5.
constexpr int kF64Threads = 512;
__global__ __launch_bounds__(kF64Threads) void probe_f64(float* sink, int iters) {
    v4d acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
    double a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        a[i] = threadIdx.x * 1e-3 + i;
        b[i] = 0.999 - i * 1e-3;
    }
    for (int it = 0; it < iters; it++)
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) s += acc[i][j][0] + acc[i][j][3];
    sink[blockIdx.x * kF64Threads + threadIdx.x] = (float)s;
}

__global__ __launch_bounds__(256) void probe_read(const float4* __restrict__ x, size_t n4, float* sink) {
    uint32_t acc = 0;
    const size_t stride = (size_t)gridDim.x * 256;
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    for (; i + 3 * stride < n4; i += 4 * stride) {
        const float4 a = ld_nt(x + i), b = ld_nt(x + i + stride), c = ld_nt(x + i + 2 * stride),
                     d = ld_nt(x + i + 3 * stride);
        acc ^= __float_as_uint(a.x) ^ __float_as_uint(a.y) ^ __float_as_uint(a.z) ^ __float_as_uint(a.w);
        acc ^= __float_as_uint(b.x) ^ __float_as_uint(b.y) ^ __float_as_uint(b.z) ^ __float_as_uint(b.w);
        acc ^= __float_as_uint(c.x) ^ __float_as_uint(c.y) ^ __float_as_uint(c.z) ^ __float_as_uint(c.w);
        acc ^= __float_as_uint(d.x) ^ __float_as_uint(d.y) ^ __float_as_uint(d.z) ^ __float_as_uint(d.w);
    }
    for (; i < n4; i += stride) {
        const float4 a = ld_nt(x + i);
        acc ^= __float_as_uint(a.x) ^ __float_as_uint(a.y) ^ __float_as_uint(a.z) ^ __float_as_uint(a.w);
    }
    sink[blockIdx.x * 256 + threadIdx.x] = __uint_as_float(acc);
}

}  // namespace
}  // namespace lt

extern "C" {

double lthip_probe_mfma_flops(int kind, int blocks, int iters) {
    const double per_wave_round = kind == 0   ? 8.0 * 32 * 32 * 16 * 2
                                  : kind == 1 ? 4.0 * 32 * 32 * 64 * 2
                                  : kind == 2 ? 16.0 * 16 * 16 * 4 * 2
                                              : 0.0;
    const double waves = kind == 2 ? lt::kF64Threads / 64 : 4.0;
    return per_wave_round * waves * (double)blocks * (double)iters;
}

int lthip_probe_mfma(int kind, int blocks, int iters, float* sink, void* stream) {
    using namespace lt;
    if (blocks <= 0 || iters <= 0 || !sink) return fail(LTHIP_ERR_INVALID_ARG, "probe_mfma: bad arguments");
    hipStream_t s = as_stream(stream);
    if (kind == 0)
        hipLaunchKernelGGL(probe_f16, dim3(blocks), dim3(256), 0, s, sink, iters);
    else if (kind == 1)
        hipLaunchKernelGGL(probe_f8, dim3(blocks), dim3(256), 0, s, sink, iters);
    else if (kind == 2)
        hipLaunchKernelGGL(probe_f64, dim3(blocks), dim3(kF64Threads), 0, s, sink, iters);
    else
        return fail(LTHIP_ERR_INVALID_ARG, "probe_mfma: kind %d (0 fp16, 1 fp8, 2 fp64)", kind);
    LT_TRY_LAUNCH("probe_mfma launch");
    return LTHIP_OK;
}

int lthip_probe_read_f32(const float* x, size_t n, float* sink, int blocks, void* stream) {
    using namespace lt;
    if (!x || !sink || blocks <= 0 || n % 4 != 0 || reinterpret_cast<uintptr_t>(x) % 16)
        return fail(LTHIP_ERR_INVALID_ARG, "probe_read: need x 16-byte aligned, n % 4 == 0, blocks > 0");
    hipLaunchKernelGGL(probe_read, dim3(blocks), dim3(256), 0, as_stream(stream), reinterpret_cast<const float4*>(x),
                       n / 4, sink);
    LT_TRY_LAUNCH("probe_read launch");
    return LTHIP_OK;
}

}  // extern "C"
