// Tiled FP16 MFMA GEMM core shared by hgemm (row-major B) and w4a16 (packed
// int4 B, dequantised into LDS).
//
// Block = 256 threads = 2x2 waves; block tile BM x BN, K-slab BK (32 or 64).
// Both operands sit in LDS as [rows][BK] fp16 with K contiguous, in 16-byte
// chunks XOR-swizzled per row (chunk kc of row r lives at kc ^ swz(r)), so the
// MFMA fragment reads (32 rows x one chunk) and the tile writes are both free of
// bank conflicts.  Each wave owns a (BM/2) x (BN/2) sub-tile of
// v_mfma_f32_32x32x16_f16 tiles with FP32 accumulators.  The next slab is
// fetched into registers while the current one is multiplied; two LDS buffers
// give one barrier per slab.
//
// The K order inside each 8-element chunk is an implementation detail: a dot
// product is invariant under a permutation of k applied to both operands, so
// the int4 loader emits k as [0,4,1,5,2,6,3,7] (cheapest nibble unpacking) and
// the A loader applies the same permutation (PERM).
#pragma once

#include "lt_common.hpp"

namespace lt {

template <int BK>
__device__ __forceinline__ int swz(int row) {
    if constexpr (BK == 64) return (row ^ (row >> 3)) & 7;
    else return ((row >> 1) ^ (row >> 3)) & 3;
}

// offset (in halves) of chunk kc of row `row` in a swizzled [rows][BK] tile
template <int BK>
__device__ __forceinline__ int soff(int row, int kc) {
    return row * BK + ((kc ^ swz<BK>(row)) << 3);
}

// halves [0..7] -> [0,4,1,5,2,6,3,7]
__device__ __forceinline__ uint4 perm_k8(uint4 a) {
    // one v_perm_b32 per word: byte i of the result is byte sel[i] of
    // {first operand (bytes 4-7), second operand (bytes 0-3)}
    uint4 r;
    r.x = __builtin_amdgcn_perm(a.z, a.x, 0x05040100u);   // lo(a.x) | lo(a.z) << 16
    r.y = __builtin_amdgcn_perm(a.z, a.x, 0x07060302u);   // hi(a.x) | hi(a.z) << 16
    r.z = __builtin_amdgcn_perm(a.w, a.y, 0x05040100u);
    r.w = __builtin_amdgcn_perm(a.w, a.y, 0x07060302u);
    return r;
}

__device__ __forceinline__ __half2 u32_as_h2(uint32_t v) {
    __half2 h;
    __builtin_memcpy(&h, &v, 4);
    return h;
}
__device__ __forceinline__ uint32_t h2_as_u32(__half2 h) {
    uint32_t v;
    __builtin_memcpy(&v, &h, 4);
    return v;
}

// v_and_or_b32 with both constants in registers: one VOP3 instruction (with
// literal constants the compiler splits (x & m) | c into v_and + v_or)
__device__ __forceinline__ uint32_t and_or(uint32_t x, uint32_t m, uint32_t c) {
    uint32_t r;
    asm("v_and_or_b32 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(m), "v"(c));
    return r;
}

// Eight int4 weights (k = i at bits 4i of w) -> fp16((q - z) * s), emitted in
// k order [0,4,1,5,2,6,3,7].  zz holds fp16(1024 + z) in both halves, ss the
// scale in both halves.  0x6400|q is fp16(1024 + q); the subtraction is exact
// and the product is rounded once, which is dequantize_w4 bit for bit
// (quantize.cpp:58-65).
__device__ __forceinline__ uint4 dequant8(uint32_t w, uint32_t zz, uint32_t ss) {
    const __half2 z2 = u32_as_h2(zz), s2 = u32_as_h2(ss);
    uint4 r;
    const uint32_t m = 0x000F000Fu, c = 0x64006400u;
    r.x = h2_as_u32(__hmul2(__hsub2(u32_as_h2(and_or(w, m, c)), z2), s2));
    r.y = h2_as_u32(__hmul2(__hsub2(u32_as_h2(and_or(w >> 4, m, c)), z2), s2));
    r.z = h2_as_u32(__hmul2(__hsub2(u32_as_h2(and_or(w >> 8, m, c)), z2), s2));
    r.w = h2_as_u32(__hmul2(__hsub2(u32_as_h2(and_or(w >> 12, m, c)), z2), s2));
    return r;
}

// The same eight weights in natural k order [0..7] (MFMA fragment order, so
// the A operand needs no permutation): pair p = (k 2p, k 2p + 1) comes from
// byte p of w, copied into bytes 0 and 2 (one v_perm_b32), then
// (y & 0x00F0000F) | 0x54006400 = fp16 (1024 + q_2p, 64 + q_2p+1): the high
// nibble sits under the exponent of 64, whose ulp is 1/16.  zz = the pair
// (0x6400 | z, 0x5400 | 16 z) = fp16 (1024 + z, 64 + z), so the subtraction
// leaves (q - z) exactly and the multiplication by (s, s) rounds once.
__device__ __forceinline__ uint4 dequant8_nat(uint32_t w, uint32_t zz, uint32_t ss) {
    const __half2 z2 = u32_as_h2(zz), s2 = u32_as_h2(ss);
    const uint32_t m = 0x00F0000Fu, c = 0x54006400u;
    uint4 r;
    r.x = h2_as_u32(__hmul2(__hsub2(u32_as_h2(and_or(__builtin_amdgcn_perm(w, w, 0x0c000c00u), m, c)), z2), s2));
    r.y = h2_as_u32(__hmul2(__hsub2(u32_as_h2(and_or(__builtin_amdgcn_perm(w, w, 0x0c010c01u), m, c)), z2), s2));
    r.z = h2_as_u32(__hmul2(__hsub2(u32_as_h2(and_or(__builtin_amdgcn_perm(w, w, 0x0c020c02u), m, c)), z2), s2));
    r.w = h2_as_u32(__hmul2(__hsub2(u32_as_h2(and_or(__builtin_amdgcn_perm(w, w, 0x0c030c03u), m, c)), z2), s2));
    return r;
}
// zz operand of dequant8_nat for zero-point z
__device__ __forceinline__ uint32_t zz_nat(uint32_t z) { return (0x6400u | z) | ((0x5400u | (z << 4)) << 16); }

// ---- A operand: row-major fp16 M x K (lda), 16-byte chunks ---------------
template <int BM, int BK, bool PERM>
struct ALoader {
    static constexpr int KC = BK / 8, CHUNKS = BM * KC, ITERS = (CHUNKS + 255) / 256;
    uint4 r[ITERS];
    __device__ __forceinline__ void load(const __half* __restrict__ A, long lda, int M, int K, int m0, int k0) {
#pragma unroll
        for (int i = 0; i < ITERS; i++) {
            const int c = threadIdx.x + 256 * i;
            const int row = c / KC, kc = c % KC;
            const int gm = m0 + row, gk = k0 + kc * 8;
            if ((CHUNKS % 256 == 0 || c < CHUNKS) && gm < M && gk < K)
                r[i] = *reinterpret_cast<const uint4*>(A + (long)gm * lda + gk);
            else
                r[i] = make_uint4(0, 0, 0, 0);
        }
    }
    __device__ __forceinline__ void store(__half* As) const {
#pragma unroll
        for (int i = 0; i < ITERS; i++) {
            const int c = threadIdx.x + 256 * i;
            if (CHUNKS % 256 == 0 || c < CHUNKS)
                *reinterpret_cast<uint4*>(As + soff<BK>(c / KC, c % KC)) = PERM ? perm_k8(r[i]) : r[i];
        }
    }
};

// ---- B operand, row-major fp16 K x N (ldb): 8x8 register transpose --------
struct BRowParams {
    const __half* B;
    long ldb;
};

template <int BN, int BK>
struct BRowLoader {
    static constexpr bool kPerm = false;
    static constexpr int NB = BN / 8, KB = BK / 8, ITEMS = NB * KB, ITERS = (ITEMS + 255) / 256;
    uint4 r[ITERS][8];
    __device__ __forceinline__ void load(const BRowParams& p, int N, int K, int n0, int k0) {
#pragma unroll
        for (int it = 0; it < ITERS; it++) {
            const int item = threadIdx.x + 256 * it;
            const int nb = item % NB, kb = item / NB;
            const int gn = n0 + 8 * nb;
#pragma unroll
            for (int rr = 0; rr < 8; rr++) {
                const int gk = k0 + 8 * kb + rr;
                if ((ITEMS % 256 == 0 || item < ITEMS) && gk < K && gn < N)
                    r[it][rr] = *reinterpret_cast<const uint4*>(p.B + (long)gk * p.ldb + gn);
                else
                    r[it][rr] = make_uint4(0, 0, 0, 0);
            }
        }
    }
    __device__ __forceinline__ static uint32_t sel(const uint4& v, int c) {
        const uint32_t d = (c >> 1) == 0 ? v.x : (c >> 1) == 1 ? v.y : (c >> 1) == 2 ? v.z : v.w;
        return (c & 1) ? (d >> 16) : (d & 0xFFFFu);
    }
    __device__ __forceinline__ void store(__half* Bs) const {
#pragma unroll
        for (int it = 0; it < ITERS; it++) {
            const int item = threadIdx.x + 256 * it;
            if (!(ITEMS % 256 == 0 || item < ITEMS)) continue;
            const int nb = item % NB, kb = item / NB;
#pragma unroll
            for (int c = 0; c < 8; c++) {
                uint4 o;
                o.x = sel(r[it][0], c) | (sel(r[it][1], c) << 16);
                o.y = sel(r[it][2], c) | (sel(r[it][3], c) << 16);
                o.z = sel(r[it][4], c) | (sel(r[it][5], c) << 16);
                o.w = sel(r[it][6], c) | (sel(r[it][7], c) << 16);
                *reinterpret_cast<uint4*>(Bs + soff<BK>(8 * nb + c, kb)) = o;
            }
        }
    }
};

// ---- B operand, packed int4 (K/8) x N + scales + zero-points ---------------
struct BW4Params {
    const uint32_t* W;        // (K/8) x N
    const __half* scales;     // (K/group) x N
    const uint32_t* qzeros;   // (K/group) x (N/8), or null for z = 8
    int group;
};

template <int BN, int BK>
struct BW4Loader {
    static constexpr bool kPerm = true;
    static constexpr int KP = BK / 8, NQ = BN / 4, ITEMS = KP * NQ, ITERS = (ITEMS + 255) / 256;
    uint4 w[ITERS];
    uint2 s[ITERS];
    uint32_t z[ITERS];
    __device__ __forceinline__ void load(const BW4Params& p, int N, int K, int n0, int k0) {
#pragma unroll
        for (int it = 0; it < ITERS; it++) {
            const int item = threadIdx.x + 256 * it;
            const int nq = item % NQ, kp = item / NQ;
            const int gkp = k0 / 8 + kp, gn = n0 + 4 * nq;
            if ((ITEMS % 256 == 0 || item < ITEMS) && gkp < K / 8 && gn < N) {
                const int g = gkp * 8 / p.group;
                w[it] = *reinterpret_cast<const uint4*>(p.W + (long)gkp * N + gn);
                s[it] = *reinterpret_cast<const uint2*>(p.scales + (long)g * N + gn);
                z[it] = p.qzeros ? (p.qzeros[(long)g * (N / 8) + gn / 8] >> (4 * (gn & 7))) & 0xFFFFu : 0x8888u;
            } else {
                w[it] = make_uint4(0, 0, 0, 0);
                s[it] = make_uint2(0, 0);
                z[it] = 0;
            }
        }
    }
    __device__ __forceinline__ void store(__half* Bs) const {
#pragma unroll
        for (int it = 0; it < ITERS; it++) {
            const int item = threadIdx.x + 256 * it;
            if (!(ITEMS % 256 == 0 || item < ITEMS)) continue;
            const int nq = item % NQ, kp = item / NQ;
            const uint32_t wv[4] = {w[it].x, w[it].y, w[it].z, w[it].w};
            const uint32_t sv[4] = {s[it].x & 0xFFFFu, s[it].x >> 16, s[it].y & 0xFFFFu, s[it].y >> 16};
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t zj = 0x6400u + ((z[it] >> (4 * j)) & 0xFu);
                *reinterpret_cast<uint4*>(Bs + soff<BK>(4 * nq + j, kp)) =
                    dequant8(wv[j], zj | (zj << 16), sv[j] | (sv[j] << 16));
            }
        }
    }
};

template <int BN, int BK, class P>
struct BLoaderFor;
template <int BN, int BK>
struct BLoaderFor<BN, BK, BRowParams> {
    using type = BRowLoader<BN, BK>;
};
template <int BN, int BK>
struct BLoaderFor<BN, BK, BW4Params> {
    using type = BW4Loader<BN, BK>;
};

// ---- one output tile ------------------------------------------------------
struct TileOut {
    __half* C;
    long ldc;
    float alpha, beta;   // C = alpha*acc + beta*C (beta == 0: C not read)
    float* partial;      // split-K: [splits][M][N] FP32 partials (or null)
    unsigned* counter;   // split-K: this tile's arrival counter
    // Owner-collects split protocol (w4_tile2 only): when set, split s of the
    // tile keeps its partial at fpart + s * fstride in the tile's MFMA
    // fragment order, and the last split (the owner) waits for the others.
    float* fpart = nullptr;
    long fstride = 0;
};

// The last block of a split tile writes C[m0:m0+BM, n0:n0+BN] = alpha * (sum
// of the splits' partials, in split order, starting from 0) (+ beta * C),
// element by element over the tile.  Nothing of the MFMA accumulator layout
// stays live here, so the epilogue adds no registers to the tile's main loop
// (a read-back per accumulator fragment made the 256 x 256 tiles spill 100-200
// VGPRs).  Partials are read with the cache policy of an agent-scope atomic
// load (sc1, the encoding the compiler gives collect()), 16 bytes at a time:
// each thread takes four column quads of the tile and issues the loads of four
// splits of all of them (16 x 16 bytes) before adding, so the read-back costs
// a few memory round trips, not one per element.  N % 4 != 0 (or partials past
// 2 GiB) falls back to one float per load.
template <int BM, int BN, int THREADS>
__device__ __forceinline__ void collect_tile(const TileOut& out, int M, int N, int splits, int m0, int n0) {
    static_assert(BN % 4 == 0, "column quads");
    constexpr int Q = BN / 4, QUADS = BM * Q, QU = 4, SB = 4, SC1 = 16;
    const long slice = (long)M * N;
    auto store4 = [&](int row, int col, const float* v, int cols) {
#pragma unroll
        for (int c = 0; c < 4; c++) {
            if (c >= cols) break;
            __half* cp = out.C + (long)row * out.ldc + col + c;
            float x = out.alpha * v[c];
            if (out.beta != 0.0f) x = fmaf(out.beta, __half2float(*cp), x);
            *cp = __float2half_rn(x);
        }
    };
    if (N % 4 == 0 && (long)splits * slice * 4 <= 0x7FFFFFFFL) {
        const __amdgpu_buffer_rsrc_t pr =
            __builtin_amdgcn_make_buffer_rsrc((void*)out.partial, 0, (int)((long)splits * slice * 4), 0x00020000);
        for (int e0 = threadIdx.x; e0 < QUADS; e0 += THREADS * QU) {
            int off[QU], rows[QU], cols[QU];
            bool ok[QU];
#pragma unroll
            for (int u = 0; u < QU; u++) {
                const int e = e0 + u * THREADS;
                rows[u] = m0 + e / Q;
                cols[u] = n0 + 4 * (e % Q);
                ok[u] = e < QUADS && rows[u] < M && cols[u] < N;
                off[u] = ok[u] ? (rows[u] * N + cols[u]) * 4 : 0;
            }
            v4f sum[QU];
#pragma unroll
            for (int u = 0; u < QU; u++) sum[u] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
            for (int sp = 0; sp < splits; sp += SB) {
                v4f v[SB][QU];
#pragma unroll
                for (int b = 0; b < SB; b++) {
                    const int q = sp + b < splits ? sp + b : splits - 1;
#pragma unroll
                    for (int u = 0; u < QU; u++)
                        v[b][u] = __builtin_amdgcn_raw_buffer_load_b128(pr, off[u], (int)(q * slice * 4), SC1);
                }
#pragma unroll
                for (int b = 0; b < SB; b++)
                    if (sp + b < splits) {
#pragma unroll
                        for (int u = 0; u < QU; u++) sum[u] += v[b][u];
                    }
            }
#pragma unroll
            for (int u = 0; u < QU; u++)
                if (ok[u]) {
                    const float w[4] = {sum[u][0], sum[u][1], sum[u][2], sum[u][3]};
                    store4(rows[u], cols[u], w, 4);
                }
        }
        return;
    }
    for (int e = threadIdx.x; e < QUADS; e += THREADS) {
        const int row = m0 + e / Q, col0 = n0 + 4 * (e % Q);
        if (row >= M || col0 >= N) continue;
        const long off = (long)row * N + col0;
        const int ncols = N - col0 < 4 ? N - col0 : 4;
        float sum[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int sp = 0; sp < splits; sp += SB) {
            float v[SB][4];
#pragma unroll
            for (int b = 0; b < SB; b++) {
                const int q = sp + b < splits ? sp + b : splits - 1;
#pragma unroll
                for (int c = 0; c < 4; c++) v[b][c] = collect(out.partial + q * slice + off + (c < ncols ? c : 0));
            }
#pragma unroll
            for (int b = 0; b < SB; b++)
                if (sp + b < splits) {
#pragma unroll
                    for (int c = 0; c < 4; c++) sum[c] += v[b][c];
                }
        }
        store4(row, col0, sum, ncols);
    }
}

// Computes C[m0:m0+BM, n0:n0+BN] over K-slabs [s0, s1) and writes it (or its
// split-K partial).  `smem` holds 2*(BM+BN)*BK halves.
template <int BM, int BN, int BK, class BP>
__device__ __forceinline__ void gemm_tile(__half* smem, const __half* __restrict__ A, long lda, const BP& bp, int M,
                                          int N, int K, int m0, int n0, int s0, int s1, int split, int splits,
                                          const TileOut& out) {
    using BL = typename BLoaderFor<BN, BK, BP>::type;
    constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 32, TN = WN / 32;
    static_assert(TM >= 1 && TN >= 1, "tile too small");
    __half* As0 = smem;
    __half* As1 = smem + BM * BK;
    __half* Bs0 = smem + 2 * BM * BK;
    __half* Bs1 = Bs0 + BN * BK;

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * WM, wn = (wave & 1) * WN;

    ALoader<BM, BK, BL::kPerm> al;
    BL bl;
    v16f acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[i][j][r] = 0.0f;

    if (s0 < s1) {
        al.load(A, lda, M, K, m0, s0 * BK);
        bl.load(bp, N, K, n0, s0 * BK);
        al.store(As0);
        bl.store(Bs0);
    }
    __syncthreads();
    const int frow = lane & 31, fchunk = lane >> 5;
    for (int s = s0; s < s1; s++) {
        const bool odd = ((s - s0) & 1) != 0;
        const __half* As = odd ? As1 : As0;
        const __half* Bs = odd ? Bs1 : Bs0;
        if (s + 1 < s1) {
            al.load(A, lda, M, K, m0, (s + 1) * BK);
            bl.load(bp, N, K, n0, (s + 1) * BK);
        }
#pragma unroll
        for (int kk = 0; kk < BK / 16; kk++) {
            v8h af[TM], bf[TN];
#pragma unroll
            for (int i = 0; i < TM; i++)
                af[i] = *reinterpret_cast<const v8h*>(As + soff<BK>(wm + 32 * i + frow, 2 * kk + fchunk));
#pragma unroll
            for (int j = 0; j < TN; j++)
                bf[j] = *reinterpret_cast<const v8h*>(Bs + soff<BK>(wn + 32 * j + frow, 2 * kk + fchunk));
#pragma unroll
            for (int i = 0; i < TM; i++)
#pragma unroll
                for (int j = 0; j < TN; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
        if (s + 1 < s1) {
            al.store(odd ? As0 : As1);
            bl.store(odd ? Bs0 : Bs1);
        }
        __syncthreads();
    }

    // C fragment of a 32x32 tile: lane holds rows 8*(r/4) + 4*(lane/32) + r%4, column lane%32
    const int fcol = lane & 31, frow4 = 4 * (lane >> 5);
    if (splits <= 1) {
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int j = 0; j < TN; j++)
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const int row = m0 + wm + 32 * i + 8 * (r >> 2) + frow4 + (r & 3);
                    const int col = n0 + wn + 32 * j + fcol;
                    if (row < M && col < N) {
                        __half* cp = out.C + (long)row * out.ldc + col;
                        float v = out.alpha * acc[i][j][r];
                        if (out.beta != 0.0f) v = fmaf(out.beta, __half2float(*cp), v);
                        *cp = __float2half_rn(v);
                    }
                }
        return;
    }
    // split-K: publish the FP32 partial; the last split to arrive sums all
    // partials in split order and writes C.
    float* mine = out.partial + (long)split * M * N;
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int row = m0 + wm + 32 * i + 8 * (r >> 2) + frow4 + (r & 3);
                const int col = n0 + wn + 32 * j + fcol;
                if (row < M && col < N) publish(mine + (long)row * N + col, acc[i][j][r]);
            }
    __shared__ int last_flag;
    drain();
    __syncthreads();
    if (threadIdx.x == 0) last_flag = (take_ticket(out.counter) == (unsigned)splits - 1);
    __syncthreads();
    if (!last_flag) return;
    collect_tile<BM, BN, 256>(out, M, N, splits, m0, n0);
    if (threadIdx.x == 0) reset_counter(out.counter);
}

}  // namespace lt
