// w4a16: C (M x N fp16) = A (M x K fp16) * dequant(W int4), FP32 accumulation.
//
// Weight format (quantize.cpp:42-65, generalised with zero-points):
//   W_packed[k/8][n] bits 4*(k%8) = q[k][n]; scales[k/group][n] fp16;
//   qzeros[k/group][n/8] bits 4*(n%8) = z (null: z = 8);
//   W[k][n] = fp16((q - z) * scale), rounded once.
//
// Because 8 consecutive k of one column live in one uint32, a dequantised
// uint32 is exactly one MFMA operand fragment (8 halves along K).
//
// kind 1, small M (decode, M <= 128): MFMA 16x16x32 on C^T = W^T A^T.  A wave
// owns 64 columns; lane l loads the uint4 W[kp][n0 + 4(l%16) .. +3] (kp = k-block
// l/16), so MFMA tile j covers columns n0 + 4r + j.  The activations are the
// B operand (16 rows of A per MFMA).  The four waves of a block split the
// block's K range; K is further split across blocks, whose FP32 partials are
// summed in split order by the last block to finish (deterministic).
//
// kind 2, large M: the shared tiled core (gemm_core.hpp) with the int4 loader,
// which dequantises each K-slab of W into LDS once per block.
//
// kind 3, large M: the pipelined tile of w4_core2.hpp (A, packed W, scales and
// zero-points reach LDS by DMA; W is dequantised in registers, one uint32 per
// MFMA fragment).  Needs group % 64 == 0, N % 32 == 0 and 16-byte aligned
// scales/zero-points; the automatic choice falls back to kind 2 otherwise.
//
// kind 4, decode (M <= 4): v_dot2 on the dequantised weights fp16((q - z) s),
// each rounded once; no split-K (described at the kernel).
//
// kind 5, skinny M (<= 64): the kind-1 lane mapping on MFMA with kind 4's
// exactly dequantised weights, 64- or 32-column blocks; no split-K.
//
// Every kind forms the reference's dequantised weight fp16((q - z) s) with one
// rounding before it multiplies (quantize.cpp:58-65); none takes the
// scale-after-the-sum shortcut of w4a16_naive.cu:12-18.
#include "w4_core2.hpp"

namespace lt {
namespace {

__global__ __launch_bounds__(256) void dequant_kernel(const uint32_t* __restrict__ W, const __half* __restrict__ scales,
                                                      const uint32_t* __restrict__ qz, __half* __restrict__ out, int K,
                                                      int N, int group) {
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    const long total = (long)(K / 8) * N;
    if (idx >= total) return;
    const int kp = (int)(idx / N), n = (int)(idx % N);
    const uint32_t w = W[(long)kp * N + n];
    const int g = kp * 8 / group;
    const uint32_t z = qz ? (qz[(long)g * (N / 8) + n / 8] >> (4 * (n & 7))) & 0xFu : 8u;
    const uint32_t s = __half_as_ushort(scales[(long)g * N + n]);
    const uint32_t zz = (0x6400u + z) * 0x10001u;
    const uint4 d = dequant8(w, zz, s | (s << 16));   // k order [0,4,1,5,2,6,3,7]
    const uint32_t dv[4] = {d.x, d.y, d.z, d.w};
#pragma unroll
    for (int j = 0; j < 4; j++) {
        out[(long)(8 * kp + j) * N + n] = __ushort_as_half((unsigned short)(dv[j] & 0xFFFFu));
        out[(long)(8 * kp + j + 4) * N + n] = __ushort_as_half((unsigned short)(dv[j] >> 16));
    }
}

constexpr int kStreamCols = 64;

// U k-steps per wave are loaded before any is used: every load is
// unconditional (out-of-range indices are clamped to a valid address and the
// activation operand is zeroed instead), so all of them are in flight at once.
// CPL columns per lane (4 or 2): a block covers 16 CPL columns.
template <int MT, int U, int WAVES, int CPL>
__global__ __launch_bounds__(64 * WAVES) void w4a16_stream_kernel(const __half* __restrict__ A, BW4Params p,
                                                                  __half* __restrict__ C, int M, int N, int K,
                                                                  int splits, float* __restrict__ partial,
                                                                  unsigned* __restrict__ counters) {
    constexpr int CT = 16 * CPL;   // columns per block
    typedef uint32_t wvec __attribute__((ext_vector_type(CPL)));
    __shared__ float red[WAVES - 1][4 * CPL][64];
    __shared__ int last_flag;
    const int ntile = blockIdx.x, split = blockIdx.y;
    const int n0 = ntile * CT;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, kb = lane >> 4;
    const int ncol = n0 + CPL * r;
    const int ncolc = ncol < N ? ncol : N - CPL;        // clamped: loads stay in bounds, results unused
    const int KP = K / 8;
    const int ksteps = (KP + 3) / 4;   // 32 k per MFMA step
    const int ks0 = (int)((long)ksteps * split / splits), ks1 = (int)((long)ksteps * (split + 1) / splits);
    const long zrow = N / 8;

    v4f acc[MT][CPL];
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int j = 0; j < CPL; j++) acc[mt][j] = v4f{0.f, 0.f, 0.f, 0.f};

    for (int ks = ks0 + wave; ks < ks1; ks += WAVES * U) {
        wvec w[U];
        uint32_t s[U][CPL / 2];
        uint32_t z[U];
        uint4 a[U][MT];
        bool kv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int kp = (ks + WAVES * u) * 4 + kb;
            kv[u] = (ks + WAVES * u < ks1) && (kp < KP);
            const int kpc = kv[u] ? kp : 0;
            const int g = kpc * 8 / p.group;
            w[u] = __builtin_nontemporal_load(reinterpret_cast<const wvec*>(p.W + (long)kpc * N + ncolc));
            const uint32_t* sp = reinterpret_cast<const uint32_t*>(p.scales + (long)g * N + ncolc);
#pragma unroll
            for (int h = 0; h < CPL / 2; h++) s[u][h] = sp[h];
            z[u] = p.qzeros ? p.qzeros[(long)g * zrow + ncolc / 8] : 0x88888888u;
#pragma unroll
            for (int mt = 0; mt < MT; mt++) {
                const int m = 16 * mt + r;
                const int mc = m < M ? m : M - 1;
                a[u][mt] = *reinterpret_cast<const uint4*>(A + (long)mc * K + 8 * kpc);
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t zs = z[u] >> (4 * (ncolc & 7));
            v8h wf[CPL];
#pragma unroll
            for (int j = 0; j < CPL; j++) {
                const uint32_t zj = 0x6400u + ((zs >> (4 * j)) & 0xFu);
                const uint32_t sj = (j & 1) ? s[u][j / 2] >> 16 : s[u][j / 2] & 0xFFFFu;
                const uint4 d = dequant8(w[u][j], zj | (zj << 16), sj | (sj << 16));
                __builtin_memcpy(&wf[j], &d, 16);
            }
#pragma unroll
            for (int mt = 0; mt < MT; mt++) {
                const bool use = kv[u] && (16 * mt + r < M);
                uint4 av = perm_k8(a[u][mt]);
                if (!use) av = make_uint4(0, 0, 0, 0);
                v8h af;
                __builtin_memcpy(&af, &av, 16);
#pragma unroll
                for (int j = 0; j < CPL; j++)
                    acc[mt][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(wf[j], af, acc[mt][j], 0, 0, 0);
            }
        }
    }

    // cross-wave sum in wave order (deterministic), one 16-row m-tile at a time
#pragma unroll
    for (int mt = 0; mt < MT; mt++) {
        if (wave > 0) {
#pragma unroll
            for (int j = 0; j < CPL; j++)
#pragma unroll
                for (int i = 0; i < 4; i++) red[wave - 1][j * 4 + i][lane] = acc[mt][j][i];
        }
        __syncthreads();
        if (wave == 0) {
#pragma unroll
            for (int j = 0; j < CPL; j++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    float v = acc[mt][j][i];
#pragma unroll
                    for (int w = 0; w < WAVES - 1; w++) v += red[w][j * 4 + i][lane];
                    acc[mt][j][i] = v;
                }
        }
        __syncthreads();
    }
    // acc[mt][j][i] of lane l is C[m = 16 mt + l%16][n = n0 + CPL (4 (l/16) + i) + j]
    if (splits <= 1) {
        if (wave != 0) return;
#pragma unroll
        for (int mt = 0; mt < MT; mt++) {
            const int m = 16 * mt + r;
            if (m >= M) continue;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int n = n0 + CPL * (4 * kb + i);
                if (n >= N) continue;
                uint32_t pk[CPL / 2];
#pragma unroll
                for (int h = 0; h < CPL / 2; h++)
                    pk[h] = (uint32_t)__half_as_ushort(__float2half_rn(acc[mt][2 * h][i])) |
                            ((uint32_t)__half_as_ushort(__float2half_rn(acc[mt][2 * h + 1][i])) << 16);
#pragma unroll
                for (int h = 0; h < CPL / 2; h++) reinterpret_cast<uint32_t*>(C + (long)m * N + n)[h] = pk[h];
            }
        }
        return;
    }
    if (wave == 0) {
        float* mine = partial + (long)split * M * N;
#pragma unroll
        for (int mt = 0; mt < MT; mt++) {
            const int m = 16 * mt + r;
            if (m >= M) continue;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int n = n0 + CPL * (4 * kb + i);
                if (n >= N) continue;
#pragma unroll
                for (int j = 0; j < CPL; j++) publish(mine + (long)m * N + n + j, acc[mt][j][i]);
            }
        }
        drain();
    }
    __syncthreads();
    if (threadIdx.x == 0) last_flag = (take_ticket(&counters[ntile]) == (unsigned)splits - 1);
    __syncthreads();
    if (!last_flag) return;
    // the last block sums the CT-column x min(M, 16 MT)-row tile, 4 columns per thread
    const int rows = M < 16 * MT ? M : 16 * MT;
    for (int e = threadIdx.x; e < rows * (CT / 4); e += 64 * WAVES) {
        const int m = e / (CT / 4), n = n0 + 4 * (e % (CT / 4));
        if (n >= N) continue;
        // splits summed in order, four splits' loads in flight at a time
        float v[4] = {0.f, 0.f, 0.f, 0.f};
        for (int sp = 0; sp < splits; sp += 4) {
            float x[4][4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int q = sp + u < splits ? sp + u : splits - 1;
                const float* src = partial + ((long)q * M + m) * N + n;
#pragma unroll
                for (int j = 0; j < 4; j++) x[u][j] = collect(src + j);
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (sp + u < splits) {
#pragma unroll
                    for (int j = 0; j < 4; j++) v[j] += x[u][j];
                }
        }
        const uint32_t lo = (uint32_t)__half_as_ushort(__float2half_rn(v[0])) |
                            ((uint32_t)__half_as_ushort(__float2half_rn(v[1])) << 16);
        const uint32_t hi = (uint32_t)__half_as_ushort(__float2half_rn(v[2])) |
                            ((uint32_t)__half_as_ushort(__float2half_rn(v[3])) << 16);
        *reinterpret_cast<uint2*>(C + (long)m * N + n) = make_uint2(lo, hi);
    }
    if (threadIdx.x == 0) reset_counter(&counters[ntile]);
}

// ---- kind 4: decode (M <= 4) on the vector ALUs -------------------------
// At M <= 4 a 16x16 MFMA tile wastes 15/16 of its rows; here the products run
// as v_dot2_f32_f16 on the vector ALUs.  Each weight is first rounded once to
// fp16((q - z) * s), the reference's dequantised weight (quantize.cpp:58-65),
// so the kernel sums exactly the products of w4a16.cpp:178-184.  One packed
// uint32 (8 k of one column) costs 1 shift, 4 v_and_or_b32, 4 packed fp16
// subtractions, 4 packed fp16 multiplications and 4 v_dot2 per activation row:
//   (x & 0x000F000F) | 0x64006400 = fp16 pair (1024 + q_k, 1024 + q_k+4)
//   (x & 0x00F000F0) | 0x54005400 = fp16 pair (64 + q_k+1, 64 + q_k+5)
// (a mantissa field of 16 q under the exponent of 64, whose ulp is 1/16), and
// the same on x >> 8 for k+2/k+6, k+3/k+7.  Subtracting 1024 + z (resp. 64 + z)
// leaves (q - z) exactly; the multiplication by the pair (s, s) rounds the
// exact product once to fp16.
//
// Layout: RG weight rows per wave load, 64 / RG lanes per row, 4 columns per
// lane (one uint4): a block covers 256 / RG columns and all of K.  The
// block's A rows (M <= MR) sit in LDS in the (k, k + 4) pair order.  Waves
// take the K / group groups round robin, QG = group / (8 RG) row quads each,
// in batches of UQ quads; the next batch's loads are issued before the
// current one is consumed.  Weight, scale and zero-point loads are buffer
// loads with the nontemporal hint (each byte is used once).  The RG row
// groups of a wave are summed by lane exchange, the waves in LDS in wave
// order: the result is deterministic.
typedef _Float16 h2v_t __attribute__((ext_vector_type(2)));
typedef uint32_t u4v_t __attribute__((ext_vector_type(4)));
typedef uint32_t u2v_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ h2v_t as_h2(uint32_t v) {
    h2v_t h;
    __builtin_memcpy(&h, &v, 4);
    return h;
}

constexpr int kDecodeWaves = 8;
constexpr int kDecodeThreads = 64 * kDecodeWaves;
constexpr int kDecodeAStage = 8;   // A elements (uint4) staged per thread: MR * K / 8 <= 4096
constexpr int kDecodeMaxMK = kDecodeAStage * kDecodeThreads * 8;

template <int MR, int RG, int QG, int UQ>
__global__ __launch_bounds__(kDecodeThreads) void w4a16_decode_kernel(const __half* __restrict__ A, BW4Params p,
                                                                      __half* __restrict__ C, int M, int N, int K) {
    static_assert(UQ % QG == 0, "a batch holds whole groups");
    constexpr int WAVES = kDecodeWaves, NT = kDecodeThreads, ASTG = kDecodeAStage;
    constexpr int LPR = 64 / RG, CT = 4 * LPR, GB = UQ / QG;
    constexpr int AUX = 2;   // nontemporal
    extern __shared__ uint4 dsm[];
    const int KP = K / 8;
    const int NG = KP / (RG * QG);
    uint4* As = dsm;                                         // [MR][KP], (k, k + 4) pair order
    float* red = reinterpret_cast<float*>(dsm + MR * KP);    // [WAVES][MR][CT]
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int cl = lane % LPR, rl = lane / LPR;
    const int n0 = blockIdx.x * CT;
    const int ncol = n0 + 4 * cl;
    const int colc = ncol < N ? ncol : N - 4;
    const __amdgpu_buffer_rsrc_t wr = __builtin_amdgcn_make_buffer_rsrc((void*)p.W, 0, KP * N * 4, 0x00020000);
    const __amdgpu_buffer_rsrc_t sr = __builtin_amdgcn_make_buffer_rsrc((void*)p.scales, 0, NG * N * 2, 0x00020000);
    const __amdgpu_buffer_rsrc_t zr =
        __builtin_amdgcn_make_buffer_rsrc((void*)(p.qzeros ? p.qzeros : p.W), 0, NG * (N / 8) * 4, 0x00020000);
    const bool has_z = p.qzeros != nullptr;
    const int wv_off = (rl * N + colc) * 4, sv_off = colc * 2, zv_off = (colc / 8) * 4;
    const int zsh = 4 * (colc & 7);
    // A loads first: the first weight batch is issued before their LDS writes wait
    uint4 av[ASTG];
    bool aval[ASTG];
#pragma unroll
    for (int i = 0; i < ASTG; i++) {
        const int e = threadIdx.x + i * NT;
        const int m = e / KP, kp = e % KP;
        aval[i] = e < MR * KP && m < M;
        av[i] = *reinterpret_cast<const uint4*>(aval[i] ? A + (long)m * K + 8 * kp : A);
    }
    __builtin_amdgcn_sched_barrier(0);
    const uint32_t MASK_LO = 0x000F000Fu, MASK_HI = 0x00F000F0u, MAG_LO = 0x64006400u, MAG_HI = 0x54005400u;
    // two FP32 sums per (row, column): even and odd nibble pairs, added at the end
    float acc[MR][4], acc2[MR][4];
#pragma unroll
    for (int m = 0; m < MR; m++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[m][j] = acc2[m][j] = 0.f;
    // this wave's groups: wave, wave + WAVES, ... (ngw of them), in batches of GB
    const int ngw = wave < NG ? (NG - 1 - wave) / WAVES + 1 : 0;
    const int nb = (ngw + GB - 1) / GB;
    const int glast = wave + WAVES * (ngw - 1);
    struct Batch {
        u4v_t w[UQ];
        u2v_t s[GB];
        uint32_t z[GB];
    };
    auto load = [&](Batch& bt, int b) {
#pragma unroll
        for (int gi = 0; gi < GB; gi++) {
            int g = wave + WAVES * (b * GB + gi);
            g = g <= glast ? g : glast;   // past the wave's last group: a repeat, not used
            bt.s[gi] = __builtin_amdgcn_raw_buffer_load_b64(sr, sv_off, g * N * 2, AUX);
            bt.z[gi] = has_z ? __builtin_amdgcn_raw_buffer_load_b32(zr, zv_off, g * (N / 8) * 4, AUX) : 0x88888888u;
#pragma unroll
            for (int qi = 0; qi < QG; qi++)
                bt.w[gi * QG + qi] = __builtin_amdgcn_raw_buffer_load_b128(wr, wv_off, (g * QG + qi) * RG * N * 4, AUX);
        }
    };
    auto compute = [&](const Batch& bt, int b) {
#pragma unroll
        for (int gi = 0; gi < GB; gi++) {
            const int g = wave + WAVES * (b * GB + gi);
            if (g > glast) break;
            const uint32_t zs = bt.z[gi] >> zsh;
            const uint32_t sv[2] = {bt.s[gi].x, bt.s[gi].y};
            uint32_t zze[4], zzo[4];
            h2v_t ss[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t zj = (zs >> (4 * j)) & 0xFu;
                const uint32_t e = 0x6400u | zj, o = 0x5400u | (zj << 4);
                zze[j] = e | (e << 16);
                zzo[j] = o | (o << 16);
                ss[j] = as_h2(__builtin_amdgcn_perm(sv[j / 2], sv[j / 2], (j & 1) ? 0x03020302u : 0x01000100u));
            }
#pragma unroll
            for (int qi = 0; qi < QG; qi++) {
                const int row = (g * QG + qi) * RG + rl;
                uint4 a[MR];
#pragma unroll
                for (int m = 0; m < MR; m++) a[m] = As[m * KP + row];
                const u4v_t wq = bt.w[gi * QG + qi];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const uint32_t x = wq[j];
                    const uint32_t t = x >> 8;
                    const h2v_t w0 = (as_h2(and_or(x, MASK_LO, MAG_LO)) - as_h2(zze[j])) * ss[j];
                    const h2v_t w1 = (as_h2(and_or(x, MASK_HI, MAG_HI)) - as_h2(zzo[j])) * ss[j];
                    const h2v_t w2 = (as_h2(and_or(t, MASK_LO, MAG_LO)) - as_h2(zze[j])) * ss[j];
                    const h2v_t w3 = (as_h2(and_or(t, MASK_HI, MAG_HI)) - as_h2(zzo[j])) * ss[j];
#pragma unroll
                    for (int m = 0; m < MR; m++) {
                        acc[m][j] = __builtin_amdgcn_fdot2(w0, as_h2(a[m].x), acc[m][j], false);
                        acc2[m][j] = __builtin_amdgcn_fdot2(w1, as_h2(a[m].y), acc2[m][j], false);
                        acc[m][j] = __builtin_amdgcn_fdot2(w2, as_h2(a[m].z), acc[m][j], false);
                        acc2[m][j] = __builtin_amdgcn_fdot2(w3, as_h2(a[m].w), acc2[m][j], false);
                    }
                }
            }
        }
    };
    Batch cur;
    if (nb > 0) load(cur, 0);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < ASTG; i++) {
        const int e = threadIdx.x + i * NT;
        if (e < MR * KP) As[e] = aval[i] ? perm_k8(av[i]) : make_uint4(0, 0, 0, 0);
    }
    // LDS writes visible; the weight loads stay in flight (no vmcnt wait)
    asm volatile("s_waitcnt lgkmcnt(0)\n s_barrier" ::: "memory");
    for (int b = 0; b + 1 < nb; b++) {
        Batch nxt;
        load(nxt, b + 1);
        compute(cur, b);
        cur = nxt;
    }
    if (nb > 0) compute(cur, nb - 1);
#pragma unroll
    for (int m = 0; m < MR; m++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[m][j] += acc2[m][j];
#pragma unroll
    for (int off = LPR; off < 64; off <<= 1)
#pragma unroll
        for (int m = 0; m < MR; m++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[m][j] += __shfl_xor(acc[m][j], off);
    if (rl == 0) {
#pragma unroll
        for (int m = 0; m < MR; m++)
#pragma unroll
            for (int j = 0; j < 4; j++) red[(wave * MR + m) * CT + 4 * cl + j] = acc[m][j];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < MR * CT; t += NT) {
        const int m = t / CT, c = t % CT;
        float v = 0.f;
        for (int w2 = 0; w2 < WAVES; w2++) v += red[(w2 * MR + m) * CT + c];
        if (m < M && n0 + c < N) C[(long)m * N + n0 + c] = __float2half_rn(v);
    }
}

typedef void (*decode_fn)(const __half*, BW4Params, __half*, int, int, int);
struct DecodeEntry {
    int mr, rg, qg;
    decode_fn fn;
};
#define LT_W4D(mr, rg, qg) {mr, rg, qg, w4a16_decode_kernel<mr, rg, qg, (qg > 4 ? qg : 4)>},
#define LT_W4D_QG(mr, rg) LT_W4D(mr, rg, 1) LT_W4D(mr, rg, 2) LT_W4D(mr, rg, 4) LT_W4D(mr, rg, 8)
const DecodeEntry kDecode[] = {LT_W4D_QG(1, 4) LT_W4D_QG(1, 8) LT_W4D_QG(2, 4) LT_W4D_QG(2, 8) LT_W4D_QG(4, 4)
                                   LT_W4D_QG(4, 8)};
#undef LT_W4D_QG
#undef LT_W4D

// Kind-4 launch shape for M x N x K (group): rows per wave load RG = 256 /
// tile_n, MR = 1, 2 or 4 >= M, QG = group / (8 RG).  Null when the shape is
// outside the kernel's domain.
const DecodeEntry* decode_entry(long M, long N, long K, int group, int tile_n) {
    if (tile_n != 32 && tile_n != 64) return nullptr;
    const int rg = 256 / tile_n;
    const int mr = M <= 1 ? 1 : (M <= 2 ? 2 : 4);
    if (M < 1 || M > 4 || K <= 0 || N < 4 || group % (8 * rg) != 0) return nullptr;
    if ((long)mr * K > kDecodeMaxMK) return nullptr;          // A staging bound (LDS <= 64 KiB)
    if ((K / 8) * N * 4 >= (1L << 31)) return nullptr;        // 32-bit buffer offsets
    const int qg = group / (8 * rg);
    for (const DecodeEntry& e : kDecode)
        if (e.mr == mr && e.rg == rg && e.qg == qg) return &e;
    return nullptr;
}
size_t decode_lds(const DecodeEntry& e, long K) {
    return (size_t)e.mr * (size_t)(K / 8) * 16 + (size_t)kDecodeWaves * e.mr * (4 * (64 / e.rg)) * 4;
}

// ---- kind 5: skinny M (<= 64) on MFMA with exactly dequantised weights ----
// C^T = W^T A^T on v_mfma_f32_16x16x32_f16: lane l loads CPL consecutive
// columns n0 + CPL (l%16) .. of weight row 4 ks + l/16 (one uint4 or uint2),
// MFMA j takes columns n0 + CPL r + j (r = l%16) as its A operand and 16
// activation rows as its B operand.  Each uint32 becomes the fragment of
// fp16((q - z) * s) of its 8 k (kind 4's exact assembly: 1 shift, 4
// v_and_or_b32, 4 packed subtractions, 4 packed multiplications), so the MFMAs
// sum exactly the products of the reference's dequantised weights
// (quantize.cpp:58-65, w4a16.cpp:178-184).  A block covers 16 CPL columns and
// all of K; its waves take the groups round robin, batches of UQ k-steps with
// the next batch's loads in flight; weights, scales and zero-points are
// nontemporal buffer loads.  Waves are summed in LDS in wave order
// (deterministic).
// Output lane l: C[16 mt + l%16][n0 + CPL (4 (l/16) + i) + j] = acc[mt][j][i].
template <int MT, int QS, int WAVES, int UQ, int CPL>
__global__ __launch_bounds__(64 * WAVES) void w4a16_skinny_kernel(const __half* __restrict__ A, BW4Params p,
                                                                  __half* __restrict__ C, int M, int N, int K) {
    static_assert(UQ % QS == 0, "a batch holds whole groups");
    static_assert(CPL == 2 || CPL == 4, "two or four columns per lane");
    constexpr int GB = UQ / QS, AUX = 2, CT = 16 * CPL;
    typedef uint32_t wvec __attribute__((ext_vector_type(CPL)));
    extern __shared__ float sred[];   // [WAVES][MT][4 i][CPL j][64]
    const int KP = K / 8;
    const int NG = KP / (4 * QS);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r = lane & 15, kb = lane >> 4;
    const int n0 = blockIdx.x * CT;
    const int colw = n0 + CPL * r < N ? n0 + CPL * r : N - CPL;   // clamped: loads in range, results unused
    const __amdgpu_buffer_rsrc_t wr = __builtin_amdgcn_make_buffer_rsrc((void*)p.W, 0, KP * N * 4, 0x00020000);
    const __amdgpu_buffer_rsrc_t sr = __builtin_amdgcn_make_buffer_rsrc((void*)p.scales, 0, NG * N * 2, 0x00020000);
    const __amdgpu_buffer_rsrc_t zr =
        __builtin_amdgcn_make_buffer_rsrc((void*)(p.qzeros ? p.qzeros : p.W), 0, NG * (N / 8) * 4, 0x00020000);
    const __amdgpu_buffer_rsrc_t ar = __builtin_amdgcn_make_buffer_rsrc((void*)A, 0, M * K * 2, 0x00020000);
    const bool has_z = p.qzeros != nullptr;
    const int wv_off = (kb * N + colw) * 4, sv_off = colw * 2, zv_off = (colw / 8) * 4;
    const int zsh = 4 * (colw & 7);
    int av_off[MT];
    bool mv[MT];
#pragma unroll
    for (int mt = 0; mt < MT; mt++) {
        const int m = 16 * mt + r;
        mv[mt] = m < M;
        av_off[mt] = ((mv[mt] ? m : M - 1) * K + 8 * kb) * 2;
    }
    const uint32_t MASK_LO = 0x000F000Fu, MASK_HI = 0x00F000F0u, MAG_LO = 0x64006400u, MAG_HI = 0x54005400u;
    v4f acc[MT][CPL];
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int j = 0; j < CPL; j++) acc[mt][j] = v4f{0.f, 0.f, 0.f, 0.f};
    const int ngw = wave < NG ? (NG - 1 - wave) / WAVES + 1 : 0;
    const int nb = (ngw + GB - 1) / GB;
    const int glast = wave + WAVES * (ngw - 1);
    struct Batch {
        wvec w[UQ];
        u4v_t a[UQ][MT];
        uint32_t s[GB][CPL / 2];
        uint32_t z[GB];
    };
    auto load = [&](Batch& bt, int b) {
#pragma unroll
        for (int gi = 0; gi < GB; gi++) {
            int g = wave + WAVES * (b * GB + gi);
            g = g <= glast ? g : glast;   // past the wave's last group: a repeat, not used
            bt.z[gi] = has_z ? __builtin_amdgcn_raw_buffer_load_b32(zr, zv_off, g * (N / 8) * 4, AUX) : 0x88888888u;
            if constexpr (CPL == 4) {
                const u2v_t sv = __builtin_amdgcn_raw_buffer_load_b64(sr, sv_off, g * N * 2, AUX);
                bt.s[gi][0] = sv.x;
                bt.s[gi][1] = sv.y;
            } else {
                bt.s[gi][0] = __builtin_amdgcn_raw_buffer_load_b32(sr, sv_off, g * N * 2, AUX);
            }
#pragma unroll
            for (int qi = 0; qi < QS; qi++) {
                const int ks = g * QS + qi;
                if constexpr (CPL == 4) {
                    const u4v_t wq = __builtin_amdgcn_raw_buffer_load_b128(wr, wv_off, ks * 4 * N * 4, AUX);
                    bt.w[gi * QS + qi] = wvec{wq.x, wq.y, wq.z, wq.w};
                } else {
                    const u2v_t wq = __builtin_amdgcn_raw_buffer_load_b64(wr, wv_off, ks * 4 * N * 4, AUX);
                    bt.w[gi * QS + qi] = wvec{wq.x, wq.y};
                }
#pragma unroll
                for (int mt = 0; mt < MT; mt++)
                    bt.a[gi * QS + qi][mt] = __builtin_amdgcn_raw_buffer_load_b128(ar, av_off[mt], ks * 64, 0);
            }
        }
    };
    auto compute = [&](const Batch& bt, int b) {
#pragma unroll
        for (int gi = 0; gi < GB; gi++) {
            const int g = wave + WAVES * (b * GB + gi);
            if (g > glast) break;
            const uint32_t zs = bt.z[gi] >> zsh;
            uint32_t zze[CPL], zzo[CPL];
            h2v_t ss[CPL];
#pragma unroll
            for (int j = 0; j < CPL; j++) {
                const uint32_t zj = (zs >> (4 * j)) & 0xFu;
                const uint32_t e = 0x6400u | zj, o = 0x5400u | (zj << 4);
                zze[j] = e | (e << 16);
                zzo[j] = o | (o << 16);
                const uint32_t sw = bt.s[gi][j / 2];
                ss[j] = as_h2(__builtin_amdgcn_perm(sw, sw, (j & 1) ? 0x03020302u : 0x01000100u));
            }
#pragma unroll
            for (int qi = 0; qi < QS; qi++) {
                v8h af[MT];
#pragma unroll
                for (int mt = 0; mt < MT; mt++) {
                    const u4v_t a = bt.a[gi * QS + qi][mt];
                    uint4 pa = perm_k8(make_uint4(a.x, a.y, a.z, a.w));
                    if (!mv[mt]) pa = make_uint4(0, 0, 0, 0);
                    __builtin_memcpy(&af[mt], &pa, 16);
                }
                const wvec wq = bt.w[gi * QS + qi];
#pragma unroll
                for (int j = 0; j < CPL; j++) {
                    const uint32_t x = wq[j];
                    const uint32_t t = x >> 8;
                    // fp16 pairs (k, k + 4) of fp16((q - z) s), in the perm_k8 order
                    const h2v_t e0 = (as_h2(and_or(x, MASK_LO, MAG_LO)) - as_h2(zze[j])) * ss[j];
                    const h2v_t e1 = (as_h2(and_or(x, MASK_HI, MAG_HI)) - as_h2(zzo[j])) * ss[j];
                    const h2v_t e2 = (as_h2(and_or(t, MASK_LO, MAG_LO)) - as_h2(zze[j])) * ss[j];
                    const h2v_t e3 = (as_h2(and_or(t, MASK_HI, MAG_HI)) - as_h2(zzo[j])) * ss[j];
                    v8h wf;
                    wf[0] = e0[0];
                    wf[1] = e0[1];
                    wf[2] = e1[0];
                    wf[3] = e1[1];
                    wf[4] = e2[0];
                    wf[5] = e2[1];
                    wf[6] = e3[0];
                    wf[7] = e3[1];
#pragma unroll
                    for (int mt = 0; mt < MT; mt++)
                        acc[mt][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(wf, af[mt], acc[mt][j], 0, 0, 0);
                }
            }
        }
    };
    Batch cur;
    if (nb > 0) load(cur, 0);
    for (int b = 0; b + 1 < nb; b++) {
        Batch nxt;
        load(nxt, b + 1);
        compute(cur, b);
        cur = nxt;
    }
    if (nb > 0) compute(cur, nb - 1);
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < CPL; j++) sred[(((wave * MT + mt) * 4 + i) * CPL + j) * 64 + lane] = acc[mt][j][i];
    __syncthreads();
    // thread (mt, i, lane): CPL consecutive columns, waves summed in order
    for (int t = threadIdx.x; t < MT * 4 * 64; t += 64 * WAVES) {
        const int l = t & 63, i = (t >> 6) & 3, mt = t >> 8;
        const int m = 16 * mt + (l & 15), n = n0 + CPL * (4 * (l >> 4) + i);
        float v[CPL];
#pragma unroll
        for (int j = 0; j < CPL; j++) v[j] = 0.f;
        for (int w2 = 0; w2 < WAVES; w2++)
#pragma unroll
            for (int j = 0; j < CPL; j++) v[j] += sred[(((w2 * MT + mt) * 4 + i) * CPL + j) * 64 + l];
        if (m < M && n < N) {
            uint32_t pk[CPL / 2];
#pragma unroll
            for (int h = 0; h < CPL / 2; h++)
                pk[h] = (uint32_t)__half_as_ushort(__float2half_rn(v[2 * h])) |
                        ((uint32_t)__half_as_ushort(__float2half_rn(v[2 * h + 1])) << 16);
            if constexpr (CPL == 4)
                *reinterpret_cast<uint2*>(C + (long)m * N + n) = make_uint2(pk[0], pk[1]);
            else
                *reinterpret_cast<uint32_t*>(C + (long)m * N + n) = pk[0];
        }
    }
}

struct SkinnyEntry {
    int mt, qs, waves, cpl;
    decode_fn fn;
};
#define LT_W4K(mt, qs, w, cpl) {mt, qs, w, cpl, w4a16_skinny_kernel<mt, qs, w, (qs > 4 ? qs : 4), cpl>},
#define LT_W4K_QS(mt, w, cpl) LT_W4K(mt, 1, w, cpl) LT_W4K(mt, 2, w, cpl) LT_W4K(mt, 4, w, cpl) LT_W4K(mt, 8, w, cpl)
// 64-column blocks (CPL 4) and 32-column blocks (CPL 2): one row tile with 16
// waves where the batch fits 128 VGPRs (group 128), else 8
const SkinnyEntry kSkinny[] = {LT_W4K(1, 1, 8, 4) LT_W4K(1, 2, 8, 4) LT_W4K(1, 4, 16, 4) LT_W4K(1, 8, 8, 4)
                                   LT_W4K_QS(2, 4, 4) LT_W4K_QS(4, 4, 4)
                                   LT_W4K(1, 1, 8, 2) LT_W4K(1, 2, 8, 2) LT_W4K(1, 4, 16, 2) LT_W4K(1, 8, 8, 2)
                                   LT_W4K_QS(2, 4, 2) LT_W4K_QS(4, 4, 2)};
#undef LT_W4K_QS
#undef LT_W4K

// Kind-5 launch entry for M x N x K (group) and 64- or 32-column blocks, or
// null outside its domain.
const SkinnyEntry* skinny_entry(long M, long N, long K, int group, int tile_n) {
    if (tile_n != 64 && tile_n != 32) return nullptr;
    if (M < 1 || M > 64 || K <= 0 || N < 8 || group % 32 != 0) return nullptr;
    if ((K / 8) * N * 4 >= (1L << 31) || M * K * 2 >= (1L << 31)) return nullptr;   // 32-bit buffer offsets
    const int mt = M <= 16 ? 1 : (M <= 32 ? 2 : 4), qs = group / 32, cpl = tile_n / 16;
    for (const SkinnyEntry& e : kSkinny)
        if (e.mt == mt && e.qs == qs && e.cpl == cpl) return &e;
    return nullptr;
}

template <int BM, int BN, int BK>
__global__ __launch_bounds__(256) void w4a16_tiled_kernel(const __half* __restrict__ A, BW4Params p,
                                                          __half* __restrict__ C, int M, int N, int K, int tiles_m,
                                                          int tiles_n, int splits, float* __restrict__ partial,
                                                          unsigned* __restrict__ counters) {
    __shared__ __align__(16) __half smem[2 * (BM + BN) * BK];
    // tiles in column-major order of (tm, tn) after the XCD remap: consecutive
    // blocks share the same weight columns
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int tm = t % tiles_m, tn = t / tiles_m;
    const int split = blockIdx.y;
    TileOut o;
    o.C = C;
    o.ldc = N;
    o.alpha = 1.0f;
    o.beta = 0.0f;
    o.partial = partial;
    o.counter = counters ? counters + t : nullptr;
    const int slabs = (K + BK - 1) / BK;
    const int s0 = (int)((long)slabs * split / splits), s1 = (int)((long)slabs * (split + 1) / splits);
    gemm_tile<BM, BN, BK>(smem, A, (long)K, p, M, N, K, tm * BM, tn * BN, s0, s1, split, splits, o);
}

typedef void (*tiled_fn)(const __half*, BW4Params, __half*, int, int, int, int, int, int, float*, unsigned*);
struct TiledEntry {
    int bm, bn;
    tiled_fn fn;
};
const TiledEntry kTiled[] = {
    {64, 64, w4a16_tiled_kernel<64, 64, 64>},
    {64, 128, w4a16_tiled_kernel<64, 128, 64>},
    {128, 128, w4a16_tiled_kernel<128, 128, 64>},
    {128, 256, w4a16_tiled_kernel<128, 256, 64>},
    {256, 128, w4a16_tiled_kernel<256, 128, 64>},
    {256, 256, w4a16_tiled_kernel<256, 256, 64>},
};
constexpr int kNumTiled = sizeof(kTiled) / sizeof(kTiled[0]);

// Blocks [0, full) run whole tiles (split blockIdx.y of gridDim.y); with a
// tail split, block full + j runs split j % tail of tile full + j / tail, so
// the tiles left over after the last full round of launch slots are shared
// by all slots instead of leaving most of them idle.
template <class C, bool LA>
__global__ __launch_bounds__(C::THREADS) void w4a16_tiled2_kernel(const __half* __restrict__ A, BW4Params p,
                                                                   __half* __restrict__ Cm, int M, int N, int K,
                                                                   int tiles_m, int full, int tail,
                                                                   float* __restrict__ partial,
                                                                   unsigned* __restrict__ counters) {
    __shared__ __align__(16) char smem[C::SMEM];
    int t, split, splits;
    if ((int)blockIdx.x < full) {
        t = xcd_remap(blockIdx.x, full);
        split = blockIdx.y;
        splits = gridDim.y;
    } else {
        const int j = blockIdx.x - full;
        t = full + j / tail;
        split = j % tail;
        splits = tail;
    }
    const int tm = t % tiles_m, tn = t / tiles_m;
    TileOut o;
    o.C = Cm;
    o.ldc = N;
    o.alpha = 1.0f;
    o.beta = 0.0f;
    o.partial = partial;
    o.counter = counters ? counters + t : nullptr;
    const int slabs = K / C::BK;
    const int s0 = (int)((long)slabs * split / splits), s1 = (int)((long)slabs * (split + 1) / splits);
    w4_tile2<C, LA>(smem, A, (long)K, p, M, N, tm * C::BM, tn * C::BN, s0, s1, split, splits, o);
}

typedef void (*tiled2_fn)(const __half*, BW4Params, __half*, int, int, int, int, int, int, float*, unsigned*);
// fn_la: the look-ahead loop with buffer-load DMA (A and W below 2^31 bytes)
struct Tiled2Entry {
    int bm, bn, smem;
    tiled2_fn fn, fn_la;
    int threads;
};
#define LT_W4T2(bm, bn, st, wm, wn)                                                                           \
    {bm, bn, Core2W4<bm, bn, st, wm, wn>::SMEM, w4a16_tiled2_kernel<Core2W4<bm, bn, st, wm, wn>, false>,   \
     w4a16_tiled2_kernel<Core2W4<bm, bn, st, wm, wn>, true>, Core2W4<bm, bn, st, wm, wn>::THREADS},
// Every tile's waves hold all of its rows (one row of waves): each weight
// word is dequantised once per block and feeds BM / 32 MFMAs.  At N = 11008,
// K = 4096 with the look-ahead loop: 256 x 256 (eight waves of 256 x 32) at
// M = 2048 in 193 against 206 us for 2 x 4 waves of 128 x 64; 128 x 128
// (four of 128 x 32) at M = 300 / 512 / 1024 in 57 / 63 / 119 against
// 67 / 73 / 131 us for 2 x 2 waves of 64 x 64; 256 x 128 (four of 256 x 32)
// 55 / 59 / 114 against 69 / 69 / 127 us; 128 x 256 as eight waves of
// 128 x 32 level with 2 x 4 waves of 64 x 64
const Tiled2Entry kTiled2[] = {LT_W4T2(256, 256, 3, 1, 8) LT_W4T2(256, 128, 3, 1, 4) LT_W4T2(128, 256, 3, 1, 8)
                                   LT_W4T2(128, 128, 3, 1, 4) LT_W4T2(64, 256, 3, 1, 8) LT_W4T2(64, 128, 3, 1, 4)
                                       LT_W4T2(64, 64, 3, 1, 2)};
#undef LT_W4T2
constexpr int kNumTiled2 = sizeof(kTiled2) / sizeof(kTiled2[0]);

const Tiled2Entry* find_tiled2(int bm, int bn) {
    for (int i = 0; i < kNumTiled2; i++)
        if (kTiled2[i].bm == bm && kTiled2[i].bn == bn) return &kTiled2[i];
    return nullptr;
}

// ---- kind 6: stream-K over the pipelined tile ---------------------------
// The launch's (tile, K-slab) iterations, tiles x slabs of them in tile-major
// order, are cut into gridDim.x contiguous ranges of equal length (+-1), one
// per block, so every block does the same MFMA and dequantisation work
// whatever the tile count.  A range covers the end of one tile, whole tiles
// and the start of another; each piece is one w4_tile2 call whose split index
// is the block's place among the blocks that share the tile.  The block that
// holds a shared tile's last slab owns it: the other pieces publish FP32
// partials in fragment order and the owner adds them to its own in split
// order (w4_owner_collect; deterministic); a tile covered by one block is
// stored directly.  Workspace: the counter head, then (most blocks sharing a
// tile - 1) planes of tiles x BM x BN floats.
// Block b covers iterations [floor(b T / G), floor((b + 1) T / G)), so the
// block holding iteration i is floor(((i + 1) G - 1) / T).
__device__ __forceinline__ int streamk_owner(long i, long T, int G) { return (int)(((i + 1) * G - 1) / T); }

template <class C>
__global__ __launch_bounds__(C::THREADS) void w4a16_streamk_kernel(const __half* __restrict__ A, BW4Params p,
                                                                    __half* __restrict__ Cm, int M, int N, int K,
                                                                    int tiles_m, int tiles, float* __restrict__ partial,
                                                                    unsigned* __restrict__ counters) {
    __shared__ __align__(16) char smem[C::SMEM];
    const int S = K / C::BK;
    const long T = (long)tiles * S;
    const int G = gridDim.x, b = blockIdx.x;
    auto piece = [&](long i0, long i1) {   // iterations [i0, i1) of one tile
        const int t = (int)(i0 / S);
        const long tbeg = (long)t * S;
        const int first = streamk_owner(tbeg, T, G), last = streamk_owner(tbeg + S - 1, T, G);
        TileOut o;
        o.C = Cm;
        o.ldc = N;
        o.alpha = 1.0f;
        o.beta = 0.0f;
        o.partial = nullptr;
        o.counter = counters + t;
        o.fpart = partial + (long)t * (C::BM * C::BN);
        o.fstride = (long)tiles * (C::BM * C::BN);
        const int tm = t % tiles_m, tn = t / tiles_m;
        w4_tile2<C, true>(smem, A, (long)K, p, M, N, tm * C::BM, tn * C::BN, (int)(i0 - tbeg), (int)(i1 - tbeg), b - first,
                    last - first + 1, o);
        // the next piece's DMA must not overwrite LDS words (the last-split
        // flag) that a slower wave of this piece still reads
        __syncthreads();
    };
    long i = (long)b * T / G, iend = (long)(b + 1) * T / G;
    // The range's last piece, when it starts a tile without finishing it, goes
    // first: its partial is then published early, and the piece that finishes
    // a tile (the range's first) comes last, so that block usually arrives
    // last at its tile and collects it without a second collect behind it.
    const long lbeg = iend > i ? ((iend - 1) / S) * S : iend;
    if (lbeg > i && iend < lbeg + S) {
        piece(lbeg, iend);
        iend = lbeg;
    }
    while (i < iend) {
        const long tend = (i / S + 1) * S;
        const long pend = iend < tend ? iend : tend;
        piece(i, pend);
        i = pend;
    }
}

typedef void (*streamk_fn)(const __half*, BW4Params, __half*, int, int, int, int, int, float*, unsigned*);
struct StreamKEntry {
    int bm, bn, smem;
    streamk_fn fn;
    int threads;
};
#define LT_W4SK(bm, bn, st, wm, wn)                                                                     \
    {bm, bn, Core2W4<bm, bn, st, wm, wn>::SMEM, w4a16_streamk_kernel<Core2W4<bm, bn, st, wm, wn>>, \
     Core2W4<bm, bn, st, wm, wn>::THREADS},
// four waves of 128 x 32 (all the tile's rows: a dequantised weight fragment
// feeds four MFMAs), one per SIMD, 96 KiB of ring: one block per CU
const StreamKEntry kStreamK[] = {LT_W4SK(128, 128, 4, 1, 4)};
#undef LT_W4SK
constexpr int kNumStreamK = sizeof(kStreamK) / sizeof(kStreamK[0]);

const StreamKEntry* find_streamk(int bm, int bn) {
    for (int i = 0; i < kNumStreamK; i++)
        if (kStreamK[i].bm == bm && kStreamK[i].bn == bn) return &kStreamK[i];
    return nullptr;
}

// Kind-6 launch geometry: G blocks (as many as fit on the CUs at once, at
// most one per 4 iterations) and the most blocks sharing one tile (the
// partial planes the workspace holds).
struct StreamKPlan {
    int tiles_m, tiles, grid, max_splits;
};
StreamKPlan streamk_plan(const StreamKEntry& e, long M, long N, long K) {
    StreamKPlan pl;
    pl.tiles_m = (int)((M + e.bm - 1) / e.bm);
    pl.tiles = pl.tiles_m * (int)((N + e.bn - 1) / e.bn);
    const long S = K / 64, T = (long)pl.tiles * S;
    long per_cu = 163840 / (e.smem + 64);
    if (per_cu < 1) per_cu = 1;
    long G = (long)cu_count() * per_cu;
    if (G > T / 4) G = T / 4;
    if (G < 1) G = 1;
    pl.grid = (int)G;
    pl.max_splits = 1;
    for (long t = 0; t < pl.tiles; t++) {
        const long first = ((t * S + 1) * G - 1) / T, last = (((t + 1) * S) * G - 1) / T;
        if (last - first + 1 > pl.max_splits) pl.max_splits = (int)(last - first + 1);
    }
    return pl;
}

// Tail split of a kind-3 launch with split_k == 1: `tail` splits for the
// tiles after the last full round of launch slots (1: none).
int tiled2_tail(const Tiled2Entry& e, long M, long N, long K, int* full) {
    const long tiles = ((M + e.bm - 1) / e.bm) * ((N + e.bn - 1) / e.bn);
    const long per_cu = 163840 / (e.smem + 64);
    const long slots = (long)cu_count() * (per_cu >= 1 ? per_cu : 1);
    const long rem = tiles % slots;
    *full = (int)tiles;
    if (rem == 0 || tiles > kMaxCounters) return 1;
    long tail = slots / rem;
    if (tail > 8) tail = 8;
    if (tail > K / 64 / 4) tail = K / 64 / 4;
    if (tail < 2) return 1;
    *full = (int)(tiles - rem);
    return (int)tail;
}

// Automatic choice where the pipelined tiles do not take the shape (group not
// a multiple of 64, scales or zero-points not 16-byte aligned): the
// LDS-dequantising tile, 64 rows up to M = 1024 (64 columns while 64 x 128
// tiles would cover less than half the CUs), else 128 x 128, with K split
// until about three blocks per CU run (at most 8, 4 for the 128 x 128 tile,
// 4 slabs per split).  Over the tuner's sweep at group 32, M > 64 this runs
// 1.07x the best configuration in geometric mean, against 1.59x for the
// pipelined choice's tile at split 1.
lthip_w4a16_config kind2_fallback(size_t M, size_t N, size_t K) {
    lthip_w4a16_config c{2, 0, 0, 1};
    c.tile_m = M <= 1024 ? 64 : 128;
    const long wide_tiles = (long)((M + c.tile_m - 1) / c.tile_m) * (long)((N + 127) / 128);
    c.tile_n = (c.tile_m == 64 && 2 * wide_tiles < cu_count()) ? 64 : 128;
    const long tiles = (long)((M + c.tile_m - 1) / c.tile_m) * (long)((N + c.tile_n - 1) / c.tile_n);
    const long slabs = (long)((K + 63) / 64);
    long split = 1;
    while (split < 8 && tiles * split * 2 <= 3L * cu_count() && split * 2 <= slabs / 4) split *= 2;
    if (c.tile_m == 128 && split > 4) split = 4;
    c.split_k = (int)split;
    return c;
}

void resolve(size_t M, size_t N, size_t K, const lthip_w4a16_config* cfg, lthip_w4a16_config* out) {
    lthip_w4a16_config c = cfg ? *cfg : lthip_w4a16_config{0, 0, 0, 0};
    if (c.kind == 0) {
        if (M <= 2) {
            // decode on the vector ALUs; 64 columns per block (RG = 4) once
            // N / 64 fills most CUs, 32 otherwise.  At M = 3, 4 the skinny
            // MFMA kernel is faster (measured at N = 4096 to 28672: 7.9 vs
            // 9.8 us at 4096 x 4096, 13.4 vs 15.4 at 8192 x 8192).  The launch
            // falls back to kind 5, then kind 1, where the conditions do not
            // hold.
            c.kind = 4;
            c.tile_n = N >= 10240 ? 64 : 32;
            c.tile_m = 0;
            c.split_k = 1;
        } else if (M <= 64) {
            c.kind = 5;
            c.tile_n = 64;
            c.tile_m = 0;
            c.split_k = 1;
        } else if (N % 32 == 0 && M <= 256 &&
                   ((M + 127) / 128) * ((N + 127) / 128) * (K / 64) >= 8 * (size_t)cu_count()) {
            // stream-K over 128 x 128 tiles whose waves hold all rows (each
            // weight dequantised once, every CU busy): at N = 11008, K = 4096,
            // M = 128 in 27 against 32 us for the 64 x 128 pipelined tile,
            // M = 256 in 40 against 46 us.  Over the tuner's sweep it beats
            // that tile on 9 shapes in 10 where each block gets at least 8
            // tile-slabs, on 6 in 10 below (tiles shared by more blocks, more
            // partials to collect): those keep the 64 x 128 tile.  The launch
            // falls back to that tile where kind 6 does not take the shape
            c.kind = 6;
            c.tile_m = 128;
            c.tile_n = 128;
            c.split_k = 1;
        } else if (N % 32 == 0) {
            // pipelined tiles: 64 x 128 up to M = 256 where kind 6 gives a
            // block fewer than 8 tile-slabs, then 128 x 128, or 256 x 256
            // (8 waves) once those cover half the CUs (256 x 256 faster from
            // M = 1024); the launch shares the last partial round of tiles
            // out by splitting K
            c.kind = 3;
            const long tiles256 = (long)((M + 255) / 256) * (long)((N + 255) / 256);
            const bool big = M > 512 && 2 * tiles256 >= cu_count();
            c.tile_m = big ? 256 : (M <= 256 ? 64 : 128);
            c.tile_n = big ? 256 : 128;
            c.split_k = 1;
        } else {
            c.kind = 2;
            c.tile_m = M >= 1024 ? 256 : (M > 128 ? 128 : 64);
            c.tile_n = (M >= 1024 && N >= 4096) ? 256 : 128;
            // split K until the tile grid holds about two blocks per CU
            const long tiles = (long)((M + c.tile_m - 1) / c.tile_m) * (long)((N + c.tile_n - 1) / c.tile_n);
            const long slabs = (long)((K + 63) / 64);
            long split = 1;
            while (tiles * split * 2 <= 2L * cu_count() && split * 2 <= slabs / 4) split *= 2;
            c.split_k = (int)split;
        }
    }
    if (c.split_k < 1) c.split_k = 1;
    *out = c;
}

}  // namespace
}  // namespace lt

extern "C" {

namespace lt {
namespace {

// Workspace of one resolved configuration.
size_t ws_need(const lthip_w4a16_config& c, size_t M, size_t N, size_t K) {
    int splits = c.split_k;
    if (c.kind == 6) {
        const StreamKEntry* e = find_streamk(c.tile_m, c.tile_n);
        if (!e || K == 0) return 0;
        const StreamKPlan pl = streamk_plan(*e, (long)M, (long)N, (long)K);
        if (pl.max_splits <= 1) return 0;
        return kCounterBytes + sizeof(float) * (size_t)(pl.max_splits - 1) * pl.tiles * e->bm * e->bn;
    }
    if (c.kind == 3 && splits <= 1) {
        const Tiled2Entry* e = find_tiled2(c.tile_m, c.tile_n);
        int full = 0;
        if (e) splits = tiled2_tail(*e, (long)M, (long)N, (long)K, &full);
    }
    if (splits <= 1) return 0;
    return kCounterBytes + sizeof(float) * (size_t)splits * M * N;
}

}  // namespace
}  // namespace lt

size_t lthip_w4a16_workspace_size(size_t M, size_t N, size_t K, const lthip_w4a16_config* cfg) {
    lthip_w4a16_config c;
    lt::resolve(M, N, K, cfg, &c);
    size_t need = lt::ws_need(c, M, N, K);
    if ((!cfg || cfg->kind == 0) && (c.kind == 3 || c.kind == 6)) {
        // the automatic choice may fall back (kind 6 -> the 64 x 128 pipelined
        // tile -> kind 2) on what the query does not see: group, alignment
        if (c.kind == 6) {
            const lthip_w4a16_config c3{3, 128, 64, 1};
            const size_t n3 = lt::ws_need(c3, M, N, K);
            if (n3 > need) need = n3;
        }
        const size_t n2 = lt::ws_need(lt::kind2_fallback(M, N, K), M, N, K);
        if (n2 > need) need = n2;
    }
    return need;
}

int lthip_w4a16_gemm(const uint16_t* A, const uint32_t* W_packed, const uint16_t* scales, const uint32_t* qzeros,
                     uint16_t* C, size_t M, size_t N, size_t K, int group, const lthip_w4a16_config* cfg, void* ws,
                     size_t ws_bytes, void* stream) {
    using namespace lt;
    if (M > (1u << 30) || N > (1u << 30) || K > (1u << 30)) return fail(LTHIP_ERR_INVALID_ARG, "w4a16: size too large");
    if (group <= 0 || group % 8 != 0 || K % (size_t)group != 0)
        return fail(LTHIP_ERR_INVALID_ARG, "w4a16: K must be a multiple of group, group a multiple of 8");
    if (N % 8 != 0) return fail(LTHIP_ERR_INVALID_ARG, "w4a16: N must be a multiple of 8");
    if (M == 0 || N == 0) return LTHIP_OK;
    if (!C || (K > 0 && (!A || !W_packed || !scales))) return fail(LTHIP_ERR_INVALID_ARG, "w4a16: null pointer");
    if (reinterpret_cast<uintptr_t>(A) % 16 || reinterpret_cast<uintptr_t>(W_packed) % 16 ||
        reinterpret_cast<uintptr_t>(scales) % 8 || reinterpret_cast<uintptr_t>(C) % 8)
        return fail(LTHIP_ERR_INVALID_ARG, "w4a16: misaligned pointer");
    lthip_w4a16_config c;
    resolve(M, N, K, cfg, &c);
    BW4Params p{W_packed, reinterpret_cast<const __half*>(scales), qzeros, group};
    const int m = (int)M, n = (int)N, k = (int)K;
    hipStream_t s = as_stream(stream);
    if ((c.kind == 4 || c.kind == 5 || c.kind == 6) && K == 0) c.kind = 1;   // nothing to accumulate: the stream kernel writes zeros
    if (c.kind == 4) {
        const bool explicit4 = cfg && cfg->kind == 4;
        const DecodeEntry* e = decode_entry((long)M, (long)N, (long)K, group, c.tile_n);
        if (explicit4 && c.split_k != 1) return fail(LTHIP_ERR_INVALID_ARG, "w4a16: kind 4 takes split_k = 1");
        if (e) {
            const size_t lds = decode_lds(*e, (long)K);
            if (lds > 65536) {   // above the default dynamic LDS limit
                const int st = check_hip(hipFuncSetAttribute(reinterpret_cast<const void*>(e->fn),
                                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                                         "w4a16 decode attribute");
                if (st != LTHIP_OK) return st;
            }
            const int ct = 4 * (64 / e->rg);
            hipLaunchKernelGGL(e->fn, dim3((unsigned)((n + ct - 1) / ct)), dim3(kDecodeThreads), lds, s,
                               reinterpret_cast<const __half*>(A), p, reinterpret_cast<__half*>(C), m, n, k);
            LT_TRY_LAUNCH("w4a16 decode launch");
            return LTHIP_OK;
        }
        if (explicit4)
            return fail(LTHIP_ERR_UNSUPPORTED,
                        "w4a16: kind 4 needs M <= 4, K > 0, tile_n 32 or 64, group %% (2048 / tile_n) == 0, "
                        "M * K <= %d and K * N / 2 < 2^31",
                        kDecodeMaxMK);
        c.kind = 5;   // automatic choice outside the decode kernel's domain
        c.split_k = 1;
        c.tile_n = 64;
        c.tile_m = 0;
    }
    if (c.kind == 5) {
        const bool explicit5 = cfg && cfg->kind == 5;
        if (explicit5 && (c.split_k != 1 || (c.tile_n != 64 && c.tile_n != 32)))
            return fail(LTHIP_ERR_INVALID_ARG, "w4a16: kind 5 takes tile_n = 64 or 32 and split_k = 1");
        const SkinnyEntry* e = skinny_entry((long)M, (long)N, (long)K, group, c.tile_n);
        if (e) {
            const int ct = 16 * e->cpl;
            const size_t lds = (size_t)e->waves * e->mt * 4 * e->cpl * 64 * 4;
            hipLaunchKernelGGL(e->fn, dim3((unsigned)((n + ct - 1) / ct)), dim3((unsigned)(64 * e->waves)), lds, s,
                               reinterpret_cast<const __half*>(A), p, reinterpret_cast<__half*>(C), m, n, k);
            LT_TRY_LAUNCH("w4a16 skinny launch");
            return LTHIP_OK;
        }
        if (explicit5 && K > 0)
            return fail(LTHIP_ERR_UNSUPPORTED,
                        "w4a16: kind 5 needs M <= 64, N >= 8, group %% 32 == 0, group <= 256 and 32-bit offsets");
        c.kind = 1;   // outside the skinny kernel's domain (or K == 0): the stream kernel
        c.split_k = 1;
        c.tile_n = kStreamCols;
        c.tile_m = 16;
    }
    if (c.kind == 1) {
        if (M > 128) return fail(LTHIP_ERR_INVALID_ARG, "w4a16: kind 1 supports M <= 128");
        const int ntiles = (n + kStreamCols - 1) / kStreamCols;
        const int ksteps = (k / 8 + 3) / 4;
        int splits = c.split_k;
        if (splits > ksteps) splits = ksteps > 0 ? ksteps : 1;
        if ((size_t)ntiles > kMaxCounters) splits = 1;
        float* partial = nullptr;
        unsigned* counters = nullptr;
        if (splits > 1) {
            lthip_w4a16_config c2 = c;
            c2.split_k = splits;
            const size_t need = lthip_w4a16_workspace_size(M, N, K, &c2);
            if (!ws || ws_bytes < need) return fail(LTHIP_ERR_WORKSPACE, "w4a16: need %zu workspace bytes", need);
            if (reinterpret_cast<uintptr_t>(ws) % 16) return fail(LTHIP_ERR_INVALID_ARG, "w4a16: workspace alignment");
            counters = reinterpret_cast<unsigned*>(ws);
            partial = reinterpret_cast<float*>(reinterpret_cast<char*>(ws) + kCounterBytes);
        }
        dim3 grid((unsigned)ntiles, (unsigned)splits);
        const __half* a = reinterpret_cast<const __half*>(A);
        __half* cc = reinterpret_cast<__half*>(C);
        const int mt = (m + 15) / 16;
        // waves per block: enough that one unrolled round of loads covers the
        // split's k-steps (more bytes in flight per CU), 4 to 16
        const int unroll = mt <= 1 ? 8 : (mt <= 2 ? 4 : 2);
        const int per_split = (ksteps + splits - 1) / splits;
        const int want = (per_split + unroll - 1) / unroll;
        const int waves = want >= 16 && mt <= 2 ? 16 : (want >= 8 ? 8 : 4);
#define LT_W4S(MT_, U_, W_, CPL_)                                                                          \
    hipLaunchKernelGGL((w4a16_stream_kernel<MT_, U_, W_, CPL_>), grid, dim3(64 * W_), 0, s, a, p, cc, m, n, k, \
                       splits, partial, counters)
        // (two columns per lane, 32-column tiles, measured slower at M = 1:
        // half-width loads leave fewer bytes in flight than the extra CUs add)
        if (mt <= 1) {
            if (waves == 16) LT_W4S(1, 8, 16, 4); else if (waves == 8) LT_W4S(1, 8, 8, 4); else LT_W4S(1, 8, 4, 4);
        } else if (mt <= 2) {
            if (waves == 16) LT_W4S(2, 4, 16, 4); else if (waves == 8) LT_W4S(2, 4, 8, 4); else LT_W4S(2, 4, 4, 4);
        } else if (mt <= 4) {
            if (waves == 8) LT_W4S(4, 2, 8, 4); else LT_W4S(4, 2, 4, 4);
        } else {
            if (waves == 8) LT_W4S(8, 2, 8, 4); else LT_W4S(8, 2, 4, 4);
        }
#undef LT_W4S
        LT_TRY_LAUNCH("w4a16 stream launch");
        return LTHIP_OK;
    }
    if (c.kind == 6) {
        const bool explicit6 = cfg && cfg->kind == 6;
        const StreamKEntry* e = find_streamk(c.tile_m, c.tile_n);
        const bool ok = e && group % 64 == 0 && N % 32 == 0 && K > 0 &&
                        reinterpret_cast<uintptr_t>(scales) % 16 == 0 &&
                        reinterpret_cast<uintptr_t>(qzeros) % 16 == 0 &&
                        ((M + e->bm - 1) / e->bm) * ((N + e->bn - 1) / e->bn) <= kMaxCounters &&
                        M * K * 2 < (1u << 31) && (K / 8) * N * 4 < (1u << 31);   // 32-bit buffer offsets
        if (!ok && explicit6)
            return fail(LTHIP_ERR_UNSUPPORTED,
                        "w4a16: kind 6 needs a built tile (128x128), K > 0, group %% 64 == 0, N %% 32 == 0, "
                        "16-byte aligned scales/zero-points, at most %zu tiles, A and W below 2 GiB",
                        kMaxCounters);
        if (explicit6 && c.split_k != 1) return fail(LTHIP_ERR_INVALID_ARG, "w4a16: kind 6 takes split_k = 1");
        if (ok) {
            const StreamKPlan pl = streamk_plan(*e, (long)M, (long)N, (long)K);
            float* partial = nullptr;
            unsigned* counters = nullptr;
            if (pl.max_splits > 1) {
                const size_t need =
                    kCounterBytes + sizeof(float) * (size_t)(pl.max_splits - 1) * pl.tiles * e->bm * e->bn;
                if (!ws || ws_bytes < need) return fail(LTHIP_ERR_WORKSPACE, "w4a16: need %zu workspace bytes", need);
                if (reinterpret_cast<uintptr_t>(ws) % 16)
                    return fail(LTHIP_ERR_INVALID_ARG, "w4a16: workspace alignment");
                counters = reinterpret_cast<unsigned*>(ws);
                partial = reinterpret_cast<float*>(reinterpret_cast<char*>(ws) + kCounterBytes);
            }
            hipLaunchKernelGGL(e->fn, dim3((unsigned)pl.grid), dim3((unsigned)e->threads), 0, s,
                               reinterpret_cast<const __half*>(A), p, reinterpret_cast<__half*>(C), m, n, k,
                               pl.tiles_m, pl.tiles, partial, counters);
            LT_TRY_LAUNCH("w4a16 stream-K launch");
            return LTHIP_OK;
        }
        // automatic choice on a shape kind 6 does not take: the 64 x 128
        // pipelined tile (which falls back to kind 2 in turn)
        c.kind = 3;
        c.tile_m = 64;
        c.tile_n = 128;
        c.split_k = 1;
    }
    if (c.kind == 3) {
        int idx2 = -1;
        for (int i = 0; i < kNumTiled2; i++)
            if (kTiled2[i].bm == c.tile_m && kTiled2[i].bn == c.tile_n) idx2 = i;
        const bool ok = idx2 >= 0 && group % 64 == 0 && N % 32 == 0 &&
                        reinterpret_cast<uintptr_t>(scales) % 16 == 0 &&
                        reinterpret_cast<uintptr_t>(qzeros) % 16 == 0;
        if (!ok && cfg && cfg->kind == 3)
            return fail(LTHIP_ERR_UNSUPPORTED,
                        "w4a16: kind 3 needs a built tile, group %% 64 == 0, N %% 32 == 0 and 16-byte aligned "
                        "scales/zero-points");
        if (ok) {
            const Tiled2Entry& e = kTiled2[idx2];
            const int tiles_m = (m + c.tile_m - 1) / c.tile_m, tiles_n = (n + c.tile_n - 1) / c.tile_n;
            const int tiles = tiles_m * tiles_n;
            int splits = c.split_k;
            const int slabs = k / 64;
            if (splits > slabs) splits = slabs > 0 ? slabs : 1;
            if ((size_t)tiles > kMaxCounters) splits = 1;
            int full = tiles, tail = 1;
            if (splits <= 1) {
                tail = tiled2_tail(e, m, n, k, &full);
                // the tail split needs the workspace; without one the launch runs whole tiles
                if (tail > 1 && (!ws || ws_bytes < kCounterBytes + sizeof(float) * (size_t)tail * M * N ||
                                 reinterpret_cast<uintptr_t>(ws) % 16)) {
                    tail = 1;
                    full = tiles;
                }
            }
            float* partial = nullptr;
            unsigned* counters = nullptr;
            if (splits > 1 || tail > 1) {
                const size_t need = kCounterBytes + sizeof(float) * (size_t)(splits > tail ? splits : tail) * M * N;
                if (!ws || ws_bytes < need) return fail(LTHIP_ERR_WORKSPACE, "w4a16: need %zu workspace bytes", need);
                if (reinterpret_cast<uintptr_t>(ws) % 16)
                    return fail(LTHIP_ERR_INVALID_ARG, "w4a16: workspace alignment");
                counters = reinterpret_cast<unsigned*>(ws);
                partial = reinterpret_cast<float*>(reinterpret_cast<char*>(ws) + kCounterBytes);
            }
            const unsigned blocks = (unsigned)(full + (tiles - full) * tail);
            const bool la = M * K * 2 < (1u << 31) && (K / 8) * N * 4 < (1u << 31);   // 32-bit buffer offsets
            hipLaunchKernelGGL(la ? e.fn_la : e.fn, dim3(blocks, (unsigned)splits), dim3((unsigned)e.threads), 0, s,
                               reinterpret_cast<const __half*>(A), p, reinterpret_cast<__half*>(C), m, n, k, tiles_m,
                               full, tail, partial, counters);
            LT_TRY_LAUNCH("w4a16 tiled2 launch");
            return LTHIP_OK;
        }
        // automatic choice on a shape kind 3 does not take (an explicit kind 3
        // failed above); the split shrinks to the workspace given, which the
        // query sizes for it
        c = kind2_fallback(M, N, K);
        while (c.split_k > 1 && (!ws || reinterpret_cast<uintptr_t>(ws) % 16 ||
                                 ws_bytes < kCounterBytes + sizeof(float) * (size_t)c.split_k * M * N))
            c.split_k /= 2;
    }
    if (c.kind != 2) return fail(LTHIP_ERR_INVALID_ARG, "w4a16: unknown kind %d", c.kind);
    int idx = -1;
    for (int i = 0; i < kNumTiled; i++)
        if (kTiled[i].bm == c.tile_m && kTiled[i].bn == c.tile_n) idx = i;
    if (idx < 0) return fail(LTHIP_ERR_UNSUPPORTED, "w4a16: tile %dx%d is not built", c.tile_m, c.tile_n);
    const int tiles_m = (m + c.tile_m - 1) / c.tile_m, tiles_n = (n + c.tile_n - 1) / c.tile_n;
    int splits = c.split_k;
    const int slabs = (k + 63) / 64;
    if (splits > slabs) splits = slabs > 0 ? slabs : 1;
    if ((size_t)tiles_m * tiles_n > kMaxCounters) splits = 1;
    float* partial = nullptr;
    unsigned* counters = nullptr;
    if (splits > 1) {
        lthip_w4a16_config c2 = c;
        c2.split_k = splits;
        const size_t need = lthip_w4a16_workspace_size(M, N, K, &c2);
        if (!ws || ws_bytes < need) return fail(LTHIP_ERR_WORKSPACE, "w4a16: need %zu workspace bytes", need);
        if (reinterpret_cast<uintptr_t>(ws) % 16) return fail(LTHIP_ERR_INVALID_ARG, "w4a16: workspace alignment");
        counters = reinterpret_cast<unsigned*>(ws);
        partial = reinterpret_cast<float*>(reinterpret_cast<char*>(ws) + kCounterBytes);
    }
    hipLaunchKernelGGL(kTiled[idx].fn, dim3((unsigned)(tiles_m * tiles_n), (unsigned)splits), dim3(256), 0, s,
                       reinterpret_cast<const __half*>(A), p, reinterpret_cast<__half*>(C), m, n, k, tiles_m, tiles_n,
                       splits, partial, counters);
    LT_TRY_LAUNCH("w4a16 tiled launch");
    return LTHIP_OK;
}

int lthip_w4_dequant(const uint32_t* W_packed, const uint16_t* scales, const uint32_t* qzeros, uint16_t* W, size_t K,
                     size_t N, int group, void* stream) {
    using namespace lt;
    if (group <= 0 || group % 8 != 0 || K % (size_t)group != 0 || N % 8 != 0)
        return fail(LTHIP_ERR_INVALID_ARG, "w4_dequant: bad K/N/group");
    if (K == 0 || N == 0) return LTHIP_OK;
    if (!W_packed || !scales || !W) return fail(LTHIP_ERR_INVALID_ARG, "w4_dequant: null pointer");
    if ((K / 8) * N > (size_t)1 << 40) return fail(LTHIP_ERR_INVALID_ARG, "w4_dequant: too large");
    const long total = (long)(K / 8) * (long)N;
    hipLaunchKernelGGL(dequant_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, as_stream(stream), W_packed,
                       reinterpret_cast<const __half*>(scales), qzeros, reinterpret_cast<__half*>(W), (int)K, (int)N,
                       group);
    LT_TRY_LAUNCH("w4_dequant launch");
    return LTHIP_OK;
}

}  // extern "C"
