// Pipelined FP16 MFMA GEMM core for row-major A (M x K) and B (K x N).
//
// Both operand tiles go global -> LDS with global_load_lds (16 bytes per lane,
// written by the hardware to consecutive LDS addresses, no registers), into a
// ring of STAGES LDS buffers: while slab s is multiplied, slabs s+1 .. s+STAGES-1
// are in flight.  One barrier per slab.
//
//   A tile [BM][BK] fp16, K contiguous; 16-byte chunk kc of row r sits at
//   kc ^ swzA(r).  MFMA A fragments (8 consecutive k of one row) are ds_read_b128.
//   B tile [BK][BN] fp16, N contiguous (as in memory); 16-byte chunk nc of row k
//   sits at nc ^ 4(k & 3).  MFMA B fragments (8 consecutive k of one column) are
//   two ds_read_b64_tr_b16.  That instruction transposes within each 16-lane
//   group: treating lane s's four halves as row s/4, columns 4(s%4)..+3 of a
//   4 x 16 block, lane i receives column i.  So lane s addresses row s/4 and
//   column 4(s%4) of the 4 x 16 block that its group needs.
//
// The swizzles keep the fragment reads free of LDS bank conflicts (below); the DMA
// fills are contiguous by construction, and the lane that fills an LDS chunk
// fetches the global chunk that belongs there, so the swizzle costs nothing.
//
// Block = WGM x WGN waves, each owning a (BM/WGM) x (BN/WGN) sub-tile of
// v_mfma_f32_32x32x16_f16 tiles with FP32 accumulators.
//
// Requirements (checked by the caller): K % 8 == 0, N % 8 == 0, lda and ldb
// multiples of 8, A and B 16-byte aligned.  Rows past M and columns past N are
// read from clamped addresses and never stored.
#pragma once

#include "gemm_core.hpp"

namespace lt {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(3))) v4s lds_v4s_t;
typedef __attribute__((address_space(3))) char lds_char_t;

// LDS bank model: 64 banks x 4 bytes; a ds_read_b128 is served 16 lanes
// (256 bytes) at a time, a ds_read_b64_tr_b16 32 lanes at a time.
//  A: 16 consecutive rows, same logical chunk -> 16 distinct 16-byte slots
//     of a 256-byte bank row (rows are 128 or 64 bytes long).
//  B: rows k..k+3 x 32 columns (4 chunks each) -> 16 distinct slots (rows are
//     a multiple of 256 bytes long).
template <int BK>
__device__ __forceinline__ int swzA2(int row) {
    if constexpr (BK == 64) return (row >> 1) & 7;
    else return (row >> 2) & 3;
}
__device__ __forceinline__ int swzB2(int k) { return 4 * (k & 3); }

// s_waitcnt with vmcnt <= n, other counters untouched (gfx9 encoding)
template <int n>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(n >= 0 && n < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((n & 0xF) | ((n >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

// Ring hand-over: this wave's DMA for the slab about to be read has landed
// (at most n newer DMA instructions outstanding), its own LDS reads are done,
// and every wave got here.  Unlike __syncthreads() this does not wait for the
// newer slabs still in flight.  The memory clobber keeps the compiler from
// moving LDS accesses across it.
template <int n>
__device__ __forceinline__ void ring_barrier() {
    asm volatile("s_waitcnt vmcnt(%0)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::"n"(n) : "memory");
}

template <int BM_, int BN_, int BK_, int STAGES_, int WGM_, int WGN_>
struct Core2 {
    static constexpr int BM = BM_, BN = BN_, BK = BK_, STAGES = STAGES_, WGM = WGM_, WGN = WGN_;
    static constexpr int WAVES = WGM * WGN, THREADS = 64 * WAVES;
    static constexpr int WM = BM / WGM, WN = BN / WGN, TM = WM / 32, TN = WN / 32;
    static constexpr int A_BYTES = BM * BK * 2, B_BYTES = BK * BN * 2;
    static constexpr int STAGE_BYTES = A_BYTES + B_BYTES;
    static constexpr int A_INSTR = A_BYTES / (1024 * WAVES);   // 1 KiB DMA instructions per wave per slab
    static constexpr int B_INSTR = B_BYTES / (1024 * WAVES);
    static constexpr int LOADS = A_INSTR + B_INSTR;
    static constexpr int KC = BK / 8, NC = BN / 8;
    static constexpr int A_ROWS_PER = 1024 / (BK * 2), B_ROWS_PER = 1024 / (BN * 2);
    static constexpr int SMEM = STAGES * STAGE_BYTES;
    static_assert(TM >= 1 && TN >= 1 && WM % 32 == 0 && WN % 32 == 0, "wave tile");
    static_assert(A_BYTES % (1024 * WAVES) == 0 && B_BYTES % (1024 * WAVES) == 0, "tile too small for the DMA split");
    static_assert(BK == 32 || BK == 64, "BK");
    static_assert(TM % 2 == 0, "the last k-step splits the M tiles in halves");
    static_assert(BN % 128 == 0 && BN * 2 <= 1024, "BN");
    static_assert(STAGES >= 3 && SMEM <= 160 * 1024, "LDS budget (the ring needs 3 stages)");
    static_assert(LOADS * (STAGES - 1) < 64, "vmcnt range");
};

template <int V>
struct int_c {
    static constexpr int value = V;
};
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(int_c<B>{});
        static_for<B + 1, E>(f);
    }
}

// Scheduling hint for one k-step region: MF MFMAs with DS LDS reads and VM
// DMA instructions spread evenly between them (reads first where they do not
// divide evenly), so the MFMA pipe never waits behind a burst of either.
template <int MF, int DS, int VM, int Q = 0>
__device__ __forceinline__ void interleave() {
    if constexpr (Q < MF) {
        constexpr int ds = (DS * (Q + 1)) / MF - (DS * Q) / MF;
        constexpr int vm = (VM * (Q + 1)) / MF - (VM * Q) / MF;
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        if constexpr (ds > 0) __builtin_amdgcn_sched_group_barrier(0x100, ds, 0);
        if constexpr (vm > 0) __builtin_amdgcn_sched_group_barrier(0x020, vm, 0);
        interleave<MF, DS, VM, Q + 1>();
    }
}

// Per-lane DMA source pointers for one tile; advanced by one slab per issue.
template <class C>
struct Core2Src {
    const __half* a[C::A_INSTR];
    const __half* b[C::B_INSTR];
    long b_step;   // elements between consecutive slabs of B

    __device__ __forceinline__ void init(const __half* __restrict__ A, long lda, const __half* __restrict__ B,
                                         long ldb, int M, int N, int m0, int n0, int k0) {
        const int lane = threadIdx.x & 63;
        const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
#pragma unroll
        for (int i = 0; i < C::A_INSTR; i++) {
            const int inst = wave * C::A_INSTR + i;   // 1 KiB block of the A tile
            const int row = inst * C::A_ROWS_PER + lane / C::KC;
            const int kc = (lane % C::KC) ^ swzA2<C::BK>(row);
            int gm = m0 + row;
            gm = gm < M ? gm : M - 1;
            a[i] = A + (long)gm * lda + k0 + kc * 8;
        }
#pragma unroll
        for (int i = 0; i < C::B_INSTR; i++) {
            const int inst = wave * C::B_INSTR + i;
            const int row = inst * C::B_ROWS_PER + lane / C::NC;
            const int nc = (lane % C::NC) ^ swzB2(row);
            int gn = n0 + nc * 8;
            gn = gn < N ? gn : N - 8;
            b[i] = B + (long)(k0 + row) * ldb + gn;
        }
        b_step = (long)C::BK * ldb;
    }

    // DMA instructions [Q0, Q1) of slab `slab` (relative to k0) into LDS
    // stage `stage`; instruction q < A_INSTR is a 1 KiB block of the A tile,
    // the rest blocks of the B tile
    template <int Q0, int Q1>
    __device__ __forceinline__ void issue(char* smem, int stage, int slab) const {
        const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        char* sa = smem + stage * C::STAGE_BYTES;
        char* sb = sa + C::A_BYTES;
#pragma unroll
        for (int q = Q0; q < Q1; q++) {
            if (q < C::A_INSTR)
                __builtin_amdgcn_global_load_lds(a[q] + (long)slab * C::BK,
                                                 (lds_void_t*)(lds_char_t*)(sa + (wave * C::A_INSTR + q) * 1024), 16,
                                                 0, 0);
            else
                __builtin_amdgcn_global_load_lds(
                    b[q - C::A_INSTR] + (long)slab * b_step,
                    (lds_void_t*)(lds_char_t*)(sb + (wave * C::B_INSTR + q - C::A_INSTR) * 1024), 16, 0, 0);
        }
    }
};

template <class C>
struct Core2Frags {
    v8h a[C::TM], b[C::TN];

    __device__ __forceinline__ void load(const char* sa, const char* sb, int kk, int wm, int wn) {
        const int lane = threadIdx.x & 63;
        const int frow = lane & 31, fkc = lane >> 5;   // A: row within a 32-row tile, 8-k chunk within 16
        const int g = lane >> 4, gi = lane & 15;       // B: 16-lane group, lane within it
        const int bcol = 16 * (g & 1) + 4 * (gi & 3);  // first column of this lane's 8-byte read
        const int brow = 8 * (g >> 1) + (gi >> 2);     // k row of the first read (+4 for the second)
#pragma unroll
        for (int i = 0; i < C::TM; i++) {
            const int row = wm + 32 * i + frow;
            const int kc = (2 * kk + fkc) ^ swzA2<C::BK>(row);
            a[i] = *reinterpret_cast<const v8h*>(sa + row * (C::BK * 2) + kc * 16);
        }
#pragma unroll
        for (int j = 0; j < C::TN; j++) {
            const int col = wn + 32 * j + bcol;
            const int k1 = 16 * kk + brow, k2 = k1 + 4;
            const int off1 = k1 * (C::BN * 2) + (((col >> 3) ^ swzB2(k1)) << 4) + (col & 7) * 2;
            const int off2 = k2 * (C::BN * 2) + (((col >> 3) ^ swzB2(k2)) << 4) + (col & 7) * 2;
            const v4s lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s_t*)(lds_char_t*)(sb + off1));
            const v4s hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s_t*)(lds_char_t*)(sb + off2));
            b[j] = __builtin_bit_cast(v8h, (v8s)__builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
        }
    }
};

// One slab at k0 into stage 0, every address clamped into the matrices (rows
// to M-1 and K-1, columns to N-8, k chunks to K-8): the partial last slab.
// What lands beyond K is masked out by the caller.
template <class C>
__device__ __forceinline__ void core2_issue_clamped(char* smem, const __half* __restrict__ A, long lda,
                                                    const __half* __restrict__ B, long ldb, int M, int N, int K,
                                                    int m0, int n0, int k0) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
#pragma unroll
    for (int i = 0; i < C::A_INSTR; i++) {
        const int inst = wave * C::A_INSTR + i;
        const int row = inst * C::A_ROWS_PER + lane / C::KC;
        const int kc = (lane % C::KC) ^ swzA2<C::BK>(row);
        int gm = m0 + row, gk = k0 + kc * 8;
        gm = gm < M ? gm : M - 1;
        gk = gk < K ? gk : K - 8;
        __builtin_amdgcn_global_load_lds(A + (long)gm * lda + gk, (lds_void_t*)(lds_char_t*)(smem + inst * 1024), 16,
                                         0, 0);
    }
#pragma unroll
    for (int i = 0; i < C::B_INSTR; i++) {
        const int inst = wave * C::B_INSTR + i;
        const int row = inst * C::B_ROWS_PER + lane / C::NC;
        const int nc = (lane % C::NC) ^ swzB2(row);
        int gn = n0 + nc * 8, gk = k0 + row;
        gn = gn < N ? gn : N - 8;
        gk = gk < K ? gk : K - 1;
        __builtin_amdgcn_global_load_lds(B + (long)gk * ldb + gn,
                                         (lds_void_t*)(lds_char_t*)(smem + C::A_BYTES + inst * 1024), 16, 0, 0);
    }
}

// Computes the BM x BN tile of C at (m0, n0) over K-slabs [s0, s1) of BK
// (the last slab of the matrix may be partial: K % 8 == 0 suffices); with
// splits > 1 the partial sums go through out.partial and the last split to
// finish sums them in split order (same protocol as gemm_tile).
template <class C>
__device__ __forceinline__ void gemm_tile2(char* smem, const __half* __restrict__ A, long lda,
                                           const __half* __restrict__ B, long ldb, int M, int N, int K, int m0,
                                           int n0, int s0, int s1, int split, int splits, const TileOut& out) {
    constexpr int TM = C::TM, TN = C::TN, BK = C::BK, STAGES = C::STAGES, KSTEPS = BK / 16;
    constexpr int MFMAS = TM * TN, READS = TM + 2 * TN;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = (wave / C::WGN) * C::WM, wn = (wave % C::WGN) * C::WN;

    v16f acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[i][j][r] = 0.0f;

    // full slabs go through the ring; a partial last slab (K % BK != 0) is
    // multiplied afterwards with the k beyond K masked out
    const int rem = K - (s1 - 1) * BK;   // valid k in the last slab of the range
    const bool partial = s1 > s0 && rem < BK;
    const int nslabs = s1 - s0 - (partial ? 1 : 0);
    if (nslabs > 0) {
        Core2Src<C> src;
        src.init(A, lda, B, ldb, M, N, m0, n0, s0 * BK);
        wait_vmcnt<0>();   // nothing older may be counted against the ring below
        // Ring of STAGES slabs.  While slab s is multiplied, slab s+STAGES-1 is
        // fetched into the stage slab s-1 used, its DMA instructions spread over
        // the k-steps of slab s (chunk kk during k-step kk).  The hand-over in the
        // last k-step waits for slab s+1, with slabs s+2 .. s+STAGES-1 (all issued
        // by then) still in flight.  Past the end the last slab is fetched again
        // into a stage nobody reads, so the vmcnt arithmetic never changes.
        // DMA chunk kk = instructions [kk * LOADS / KSTEPS, (kk + 1) * LOADS / KSTEPS)
#pragma unroll
        for (int p = 0; p < STAGES - 1; p++) src.template issue<0, C::LOADS>(smem, p, p < nslabs ? p : nslabs - 1);
        ring_barrier<C::LOADS * (STAGES - 2)>();

        Core2Frags<C> fr[2];
        fr[0].load(smem, smem + C::A_BYTES, 0, wm, wn);
        int stage = 0;                  // s % STAGES
        int fill_stage = STAGES - 1;    // (s + STAGES - 1) % STAGES
        for (int s = 0; s < nslabs; s++) {
            const char* sa = smem + stage * C::STAGE_BYTES;
            const char* sb = sa + C::A_BYTES;
            const int fill = s + STAGES - 1 < nslabs ? s + STAGES - 1 : nslabs - 1;
            __builtin_amdgcn_sched_barrier(0);
            // k-steps 0 .. KSTEPS-2: MFMAs of step kk, fragment reads of step
            // kk+1 and DMA chunk kk; each step is its own scheduling region
            static_for<0, KSTEPS - 1>([&](auto kkc) {
                constexpr int kk = decltype(kkc)::value;
                fr[(kk + 1) & 1].load(sa, sb, kk + 1, wm, wn);
                src.template issue<kk * C::LOADS / KSTEPS, (kk + 1) * C::LOADS / KSTEPS>(smem, fill_stage, fill);
                const Core2Frags<C>& f = fr[kk & 1];
#pragma unroll
                for (int i = 0; i < TM; i++)
#pragma unroll
                    for (int j = 0; j < TN; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(f.a[i], f.b[j], acc[i][j], 0, 0, 0);
                interleave<MFMAS, READS, (kk + 1) * C::LOADS / KSTEPS - kk * C::LOADS / KSTEPS>();
                __builtin_amdgcn_sched_barrier(0);
            });
            // last k-step: half the MFMAs with the last DMA chunk, hand-over,
            // first fragments of slab s+1 with the other half
            {
                constexpr int KL = KSTEPS - 1;
                const Core2Frags<C>& f = fr[KL & 1];
                src.template issue<KL * C::LOADS / KSTEPS, C::LOADS>(smem, fill_stage, fill);
#pragma unroll
                for (int i = 0; i < TM / 2; i++)
#pragma unroll
                    for (int j = 0; j < TN; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(f.a[i], f.b[j], acc[i][j], 0, 0, 0);
                interleave<MFMAS / 2, 0, C::LOADS - KL * C::LOADS / KSTEPS>();
                __builtin_amdgcn_sched_barrier(0);
                ring_barrier<C::LOADS * (STAGES - 2)>();
                stage = stage + 1 == STAGES ? 0 : stage + 1;
                fill_stage = fill_stage + 1 == STAGES ? 0 : fill_stage + 1;
                const char* na = smem + stage * C::STAGE_BYTES;
                fr[0].load(na, na + C::A_BYTES, 0, wm, wn);   // KSTEPS is even
#pragma unroll
                for (int i = TM / 2; i < TM; i++)
#pragma unroll
                    for (int j = 0; j < TN; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(f.a[i], f.b[j], acc[i][j], 0, 0, 0);
                interleave<MFMAS / 2, READS, 0>();
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        wait_vmcnt<0>();   // the redundant tail DMA must land before the LDS is reused
    }
    if (partial) {
        const int k0 = (s1 - 1) * BK;
        ring_barrier<0>();   // every wave is done with the ring
        core2_issue_clamped<C>(smem, A, lda, B, ldb, M, N, K, m0, n0, k0);
        ring_barrier<0>();
        const int kl = 8 * (lane >> 5);   // first k of this lane's 8 in a k-step
#pragma unroll
        for (int kk = 0; kk < KSTEPS; kk++) {
            Core2Frags<C> f;
            f.load(smem, smem + C::A_BYTES, kk, wm, wn);
            if (16 * kk + kl >= rem) {
#pragma unroll
                for (int i = 0; i < TM; i++) f.a[i] = v8h{};
#pragma unroll
                for (int j = 0; j < TN; j++) f.b[j] = v8h{};
            }
#pragma unroll
            for (int i = 0; i < TM; i++)
#pragma unroll
                for (int j = 0; j < TN; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(f.a[i], f.b[j], acc[i][j], 0, 0, 0);
        }
    }
    __syncthreads();   // the caller may refill the ring for its next tile

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
    int* last_split = reinterpret_cast<int*>(smem);   // the ring is idle by now
    drain();
    __syncthreads();
    if (threadIdx.x == 0) *last_split = (take_ticket(out.counter) == (unsigned)splits - 1);
    __syncthreads();
    if (!*last_split) return;
    collect_tile<C::BM, C::BN, C::THREADS>(out, M, N, splits, m0, n0);
    if (threadIdx.x == 0) reset_counter(out.counter);
}

}  // namespace lt
