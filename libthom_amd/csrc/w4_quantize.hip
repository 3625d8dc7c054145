// The int4 quantizer of quantize.cpp:8-53 as a GPU producer: fp16 weights
// W (K x N, row-major) -> packed int4 in the GPTQ layout (quantize.cpp:42-53:
// row k of column n in bits 4*(k%8) of W_packed[k/8][n]), fp16 scales
// (K/group) x N and, for the asymmetric variant, packed zero-points
// (K/group) x (N/8) uint32 with column n in bits 4*(n%8).
//
// Symmetric (quantize.cpp:8-37): scale = fp16(2 max|W| / 15), or 1 for an
// all-zero column group; q = clamp(nearbyint(W / scale) + 8, 0, 15) with the
// fp16-rounded scale.  Asymmetric (no counterpart in the reference): lo =
// min(0, min W), hi = max(0, max W); scale = fp16((hi - lo) / 15), or 1 when
// hi == lo; z = clamp(nearbyint(-lo / scale), 0, 15); q = clamp(nearbyint(W /
// scale) + z, 0, 15).
//
// A block owns one group of rows x 64 columns: the tile is staged in LDS,
// 256 threads take the per-column extremes in four row phases (max and min
// are exact, so the order does not matter), 64 threads derive the scales and
// zero-points, then every thread quantises and packs 8-row words; the packed
// writes of a wave cover 64 consecutive columns.  Divisions are IEEE FP32
// (__fdiv_rn), rintf rounds half to even like nearbyint, __float2half rounds
// to nearest even: the output is bit-identical to the host quantizer.
#include "lt_common.hpp"

namespace lt {
namespace {

constexpr int kQCols = 64, kQThreads = 256, kQPhases = kQThreads / kQCols;

size_t quantize_lds(int group) {
    return (size_t)group * kQCols * 2 + (size_t)2 * kQPhases * kQCols * 4 + (size_t)kQCols * 8;
}

template <bool ASYM>
__global__ __launch_bounds__(kQThreads) void w4_quantize_kernel(const __half* __restrict__ W,
                                                                uint32_t* __restrict__ packed,
                                                                __half* __restrict__ scales,
                                                                uint32_t* __restrict__ qzeros, int N, int group) {
    extern __shared__ uint4 qsm[];
    __half* tile = reinterpret_cast<__half*>(qsm);                     // [group][64]
    float* red = reinterpret_cast<float*>(tile + group * kQCols);       // [2][phases][64]: lo, hi
    float* colscale = red + 2 * kQPhases * kQCols;                      // [64]
    int* colz = reinterpret_cast<int*>(colscale + kQCols);              // [64]
    const int g = blockIdx.y, n0 = blockIdx.x * kQCols;
    const int t = threadIdx.x;
    const long row0 = (long)g * group;
    for (int c = t; c < group * 8; c += kQThreads) {   // 16-byte chunks, 8 per tile row
        const int r = c >> 3, c8 = c & 7, n = n0 + 8 * c8;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (n < N) v = *reinterpret_cast<const uint4*>(W + (row0 + r) * N + n);
        *reinterpret_cast<uint4*>(tile + r * kQCols + 8 * c8) = v;
    }
    __syncthreads();
    {
        const int c = t % kQCols, ph = t / kQCols;
        float lo = 0.0f, hi = 0.0f;
        for (int r = ph; r < group; r += kQPhases) {
            const float w = __half2float(tile[r * kQCols + c]);
            if (ASYM) {
                lo = fminf(lo, w);
                hi = fmaxf(hi, w);
            } else {
                hi = fmaxf(hi, fabsf(w));
            }
        }
        red[ph * kQCols + c] = lo;
        red[(kQPhases + ph) * kQCols + c] = hi;
    }
    __syncthreads();
    if (t < kQCols) {
        float lo = red[t], hi = red[kQPhases * kQCols + t];
        for (int ph = 1; ph < kQPhases; ph++) {
            lo = fminf(lo, red[ph * kQCols + t]);
            hi = fmaxf(hi, red[(kQPhases + ph) * kQCols + t]);
        }
        __half hs;
        int z = 8;
        if (ASYM) {
            const float r = hi - lo;
            hs = __float2half(r > 0.0f ? __fdiv_rn(r, 15.0f) : 1.0f);
            const int zz = (int)rintf(__fdiv_rn(-lo, __half2float(hs)));
            z = zz < 0 ? 0 : (zz > 15 ? 15 : zz);
        } else {
            const float s = __fdiv_rn(2.0f * hi, 15.0f);   // quantize.cpp:23
            hs = __float2half(s > 0.0f ? s : 1.0f);
        }
        colscale[t] = __half2float(hs);
        colz[t] = z;
        if (n0 + t < N) scales[(long)g * N + n0 + t] = hs;
    }
    __syncthreads();
    if (ASYM && t < kQCols / 8 && n0 + 8 * t < N) {
        uint32_t v = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) v |= (uint32_t)colz[8 * t + i] << (4 * i);
        qzeros[(long)g * (N / 8) + n0 / 8 + t] = v;
    }
    const int words = (group / 8) * kQCols;
    for (int it = t; it < words; it += kQThreads) {
        const int k8 = it / kQCols, c = it % kQCols, n = n0 + c;
        if (n >= N) continue;
        const float s = colscale[c];
        const int z = colz[c];
        uint32_t v = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const float w = __half2float(tile[(8 * k8 + i) * kQCols + c]);
            int q = (int)rintf(__fdiv_rn(w, s)) + z;
            q = q < 0 ? 0 : (q > 15 ? 15 : q);
            v |= (uint32_t)q << (4 * i);
        }
        packed[(row0 / 8 + k8) * N + n] = v;
    }
}

// Fast path (group 64, 128 or 256; 16-byte aligned W, packed words and
// scales): one wave owns one group of G = 8R rows x 64 columns, entirely in
// registers -- no LDS, no barrier.  Lane (rb, co) = (lane / 8, lane % 8) holds
// rows rb R .. +R-1 of columns 8 co .. +7 (R 16-byte loads; each load
// instruction covers 8 rows x 128 contiguous bytes).  Column extremes: packed
// fp16 max/min over the lane's rows (exact, NaN ignored as fmaxf does), then
// across the 8 row blocks with three xor shuffles.  Every lane of a column
// octet derives the same 8 scales (and zero-points) with the IEEE divisions
// of the host quantizer, and packs its own R/8 words per column.
//
// The per-weight division: w and s are fp16, so the exact quotient w / s is
// either exactly a half-integer or at least ~2^-14 away from one (for
// |w / s| <= 16, against an fp32 ulp of 2^-19 there).  Hence w * rcp(s), a few
// ulp from w / s, rounds to the same integer as rintf(w / s) unless the
// quotient is an exact tie, which h s == w detects exactly for the nearest
// half-integer h (h s has at most 17 significant bits).  Beyond |w / s| = 16
// the clamp to [0, 15] decides.  (Checked against IEEE division over every
// finite fp16 w and 2583 fp16 scales: no difference.)
typedef _Float16 qh2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ qh2_t as_qh2(uint32_t v) { return __builtin_bit_cast(qh2_t, v); }
__device__ __forceinline__ uint32_t qh2_bits(qh2_t v) { return __builtin_bit_cast(uint32_t, v); }

__device__ __forceinline__ int quant_round(float w, float s, float rs) {
    const float q = w * rs;
    const float h = floorf(q) + 0.5f;
    return (int)rintf(h * s == w ? h : q);
}

template <bool ASYM, int R>
__global__ __launch_bounds__(256) void w4_quantize_wave_kernel(const __half* __restrict__ W,
                                                               uint32_t* __restrict__ packed,
                                                               __half* __restrict__ scales,
                                                               uint32_t* __restrict__ qzeros, int N, int groups) {
    static_assert(R % 8 == 0 && R >= 8 && R <= 32, "rows per lane: whole packed words");
    const int lane = threadIdx.x & 63;
    const int tiles_n = (N + 63) / 64;
    const int wv = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int g = wv / tiles_n, tn = wv % tiles_n;
    if (g >= groups) return;
    const int rb = lane >> 3, co = lane & 7;
    const int n = tn * 64 + 8 * co;
    const bool colok = n < N;
    const long row0 = (long)g * 8 * R + (long)rb * R;
    uint4 v[R];
#pragma unroll
    for (int i = 0; i < R; i++)
        v[i] = colok ? ld_nt(reinterpret_cast<const uint4*>(W + (row0 + i) * N + n)) : make_uint4(0, 0, 0, 0);
    // per column pair p (columns 2p, 2p+1): hi = max(0, |w|) (symmetric) or
    // max(0, w) and lo = min(0, w) (asymmetric); elementwise max/min follow
    // maxNum/minNum (a NaN is ignored, as fmaxf/fminf do)
    qh2_t hi[4], lo[4];
#pragma unroll
    for (int p = 0; p < 4; p++) hi[p] = lo[p] = qh2_t{(_Float16)0.0f, (_Float16)0.0f};
#pragma unroll
    for (int i = 0; i < R; i++) {
        const uint32_t wd[4] = {v[i].x, v[i].y, v[i].z, v[i].w};
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const qh2_t x = as_qh2(wd[p]);
            if (ASYM) {
                hi[p] = __builtin_elementwise_max(hi[p], x);
                lo[p] = __builtin_elementwise_min(lo[p], x);
            } else {
                hi[p] = __builtin_elementwise_max(hi[p], __builtin_elementwise_abs(x));
            }
        }
    }
#pragma unroll
    for (int off = 8; off < 64; off <<= 1)
#pragma unroll
        for (int p = 0; p < 4; p++) {
            hi[p] = __builtin_elementwise_max(hi[p], as_qh2((uint32_t)__shfl_xor((int)qh2_bits(hi[p]), off)));
            if (ASYM)
                lo[p] = __builtin_elementwise_min(lo[p], as_qh2((uint32_t)__shfl_xor((int)qh2_bits(lo[p]), off)));
        }
    float sc[8], rs[8];
    int z[8];
    uint32_t shalf[4] = {0, 0, 0, 0}, zword = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const float hj = (float)hi[j / 2][j & 1];
        __half hs;
        int zj = 8;
        if (ASYM) {
            const float lj = (float)lo[j / 2][j & 1];
            const float r = hj - lj;
            hs = __float2half(r > 0.0f ? __fdiv_rn(r, 15.0f) : 1.0f);
            const int zz = (int)rintf(__fdiv_rn(-lj, __half2float(hs)));
            zj = zz < 0 ? 0 : (zz > 15 ? 15 : zz);
        } else {
            const float sv = __fdiv_rn(2.0f * hj, 15.0f);   // quantize.cpp:23
            hs = __float2half(sv > 0.0f ? sv : 1.0f);
        }
        sc[j] = __half2float(hs);
        rs[j] = __builtin_amdgcn_rcpf(sc[j]);
        z[j] = zj;
        shalf[j / 2] |= (uint32_t)__half_as_ushort(hs) << (16 * (j & 1));
        zword |= (uint32_t)zj << (4 * j);
    }
    if (colok && rb == 0) {
        *reinterpret_cast<uint4*>(scales + (long)g * N + n) = make_uint4(shalf[0], shalf[1], shalf[2], shalf[3]);
        if (ASYM) qzeros[(long)g * (N / 8) + n / 8] = zword;
    }
    if (!colok) return;
#pragma unroll
    for (int kw = 0; kw < R / 8; kw++) {
        uint32_t word[8];
#pragma unroll
        for (int j = 0; j < 8; j++) word[j] = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint4 rv = v[8 * kw + i];
            const uint32_t wd[4] = {rv.x, rv.y, rv.z, rv.w};
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const float w = (float)as_qh2(wd[j / 2])[j & 1];
                int q = quant_round(w, sc[j], rs[j]) + z[j];
                q = q < 0 ? 0 : (q > 15 ? 15 : q);
                word[j] |= (uint32_t)q << (4 * i);
            }
        }
        uint32_t* dst = packed + (row0 / 8 + kw) * (long)N + n;
        reinterpret_cast<uint4*>(dst)[0] = make_uint4(word[0], word[1], word[2], word[3]);
        reinterpret_cast<uint4*>(dst)[1] = make_uint4(word[4], word[5], word[6], word[7]);
    }
}

}  // namespace
}  // namespace lt

extern "C" int lthip_w4_quantize(const uint16_t* W, uint32_t* W_packed, uint16_t* scales, uint32_t* qzeros,
                                 size_t K, size_t N, int group, int mode, void* stream) {
    using namespace lt;
    if (mode != 0 && mode != 1) return fail(LTHIP_ERR_INVALID_ARG, "w4_quantize: mode 0 (symmetric) or 1 (asymmetric)");
    if (group <= 0 || group % 8 != 0 || group > 1024 || K % (size_t)group != 0 || N % 8 != 0)
        return fail(LTHIP_ERR_INVALID_ARG, "w4_quantize: need K %% group == 0, group %% 8 == 0, group <= 1024, N %% 8 == 0");
    if (K > (1u << 30) || N > (1u << 30) || (K / (size_t)group) > 65535)
        return fail(LTHIP_ERR_INVALID_ARG, "w4_quantize: size too large");
    if (K == 0 || N == 0) return LTHIP_OK;
    if (!W || !W_packed || !scales || (mode == 1 && !qzeros))
        return fail(LTHIP_ERR_INVALID_ARG, "w4_quantize: null pointer");
    if (reinterpret_cast<uintptr_t>(W) % 16 || reinterpret_cast<uintptr_t>(W_packed) % 4 ||
        reinterpret_cast<uintptr_t>(scales) % 2 || reinterpret_cast<uintptr_t>(qzeros) % 4)
        return fail(LTHIP_ERR_INVALID_ARG, "w4_quantize: misaligned pointer");
    // the register-resident fast path
    const bool aligned = reinterpret_cast<uintptr_t>(W_packed) % 16 == 0 && reinterpret_cast<uintptr_t>(scales) % 16 == 0;
    if (aligned && (group == 64 || group == 128 || group == 256)) {
        const long waves = (long)(K / (size_t)group) * (long)((N + 63) / 64);
        const dim3 grid((unsigned)((waves + 3) / 4));
        const __half* w = reinterpret_cast<const __half*>(W);
        __half* sp = reinterpret_cast<__half*>(scales);
        const int groups = (int)(K / (size_t)group);
#define LT_WQ(A, R_)                                                                                    \
    hipLaunchKernelGGL((w4_quantize_wave_kernel<A, R_>), grid, dim3(256), 0, as_stream(stream), w, W_packed, sp, \
                       qzeros, (int)N, groups)
        if (mode == 1) {
            if (group == 64) LT_WQ(true, 8); else if (group == 128) LT_WQ(true, 16); else LT_WQ(true, 32);
        } else {
            if (group == 64) LT_WQ(false, 8); else if (group == 128) LT_WQ(false, 16); else LT_WQ(false, 32);
        }
#undef LT_WQ
        LT_TRY_LAUNCH("w4_quantize launch");
        return LTHIP_OK;
    }
    const size_t lds = quantize_lds(group);
    const void* fn = mode == 1 ? reinterpret_cast<const void*>(w4_quantize_kernel<true>)
                               : reinterpret_cast<const void*>(w4_quantize_kernel<false>);
    if (lds > 65536) {   // above the default dynamic LDS limit
        const int st = check_hip(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                                 "w4_quantize attribute");
        if (st != LTHIP_OK) return st;
    }
    const dim3 grid((unsigned)((N + kQCols - 1) / kQCols), (unsigned)(K / (size_t)group));
    const __half* w = reinterpret_cast<const __half*>(W);
    __half* s = reinterpret_cast<__half*>(scales);
    if (mode == 1)
        hipLaunchKernelGGL(w4_quantize_kernel<true>, grid, dim3(kQThreads), lds, as_stream(stream), w, W_packed, s,
                           qzeros, (int)N, group);
    else
        hipLaunchKernelGGL(w4_quantize_kernel<false>, grid, dim3(kQThreads), lds, as_stream(stream), w, W_packed, s,
                           qzeros, (int)N, group);
    LT_TRY_LAUNCH("w4_quantize launch");
    return LTHIP_OK;
}
