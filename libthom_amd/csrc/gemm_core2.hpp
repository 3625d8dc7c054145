// Pipelined FP16 MFMA GEMM core for row-major A (M x K) and B (K x N).
//
// Both operand tiles go global -> LDS with global_load_lds (16 bytes per lane,
// written by the hardware to consecutive LDS addresses, no registers), into a
// ring of STAGES LDS buffers: while slab s is multiplied, slabs s+1 .. s+STAGES-1
// are in flight.  One barrier per slab.
//
// The product is built from v_mfma_f32_16x16x32_f16 (16 x 16 outputs, 32 k per
// instruction): it sustains a higher clock than the 32 x 32 x 16 shape at equal
// flops (half the accumulator bytes per flop; measured with MFMA-only loops:
// about 2.0 against 1.7 GHz under the board power limit), which is worth more
// than its extra instructions.
// Operand layout: lane l holds A[l % 16][8 (l / 16) .. +7] and
// B[8 (l / 16) .. +7][l % 16].  The instruction is given B's fragment as its
// first operand and A's as its second, so it computes the transposed tile and
// lane l receives C[l % 16][4 (l / 16) .. +3]: four consecutive columns of one
// row, stored with one 8-byte write (the natural order would leave each lane
// one column of four rows, four 2-byte writes).
//
//   A tile [BM][BK] fp16, K contiguous; 16-byte chunk kc of row r sits at
//   kc ^ swzA(r).  An A fragment (8 consecutive k of one row) is one ds_read_b128.
//   B tile [BK][BN] fp16, N contiguous (as in memory); 16-byte chunk nc of row k
//   sits at nc ^ swzB(k).  A B fragment (8 consecutive k of one column) is two
//   ds_read_b64_tr_b16.  That instruction transposes within each 16-lane group:
//   treating lane s's four halves as row s/4, columns 4(s%4)..+3 of a 4 x 16
//   block, lane i receives column i.  So lane s addresses row s/4 and column
//   4(s%4) of the 4 x 16 block (k rows 8g..8g+3, then 8g+4..8g+7 for group g).
//
// The swizzles keep the fragment reads free of LDS bank conflicts (below); the DMA
// fills are contiguous by construction, and the lane that fills an LDS chunk
// fetches the global chunk that belongs there, so the swizzle costs nothing.
//
// Block = WGM x WGN waves, each owning a (BM/WGM) x (BN/WGN) sub-tile of
// 16 x 16 MFMA tiles with FP32 accumulators.
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
//  B: two lane groups read k rows {k0..k0+3} and {k0+8..k0+11} x 16 columns
//     (2 chunks) -> the 8 rows differ in (k & 3) or bit 3 of k, so chunk ^ swzB
//     gives 16 distinct slots (rows are a multiple of 256 bytes long).
template <int BK>
__device__ __forceinline__ int swzA2(int row) {
    if constexpr (BK == 64) return (row >> 1) & 7;
    else return (row >> 2) & 3;
}
__device__ __forceinline__ int swzB2(int k) { return (4 * (k & 3)) ^ (2 * ((k >> 3) & 1)); }

// s_waitcnt with vmcnt <= n, other counters untouched (gfx9 encoding)
template <int n>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(n >= 0 && n < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((n & 0xF) | ((n >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

// Ring hand-over: this wave's DMA for the slab about to be read has landed
// (at most n newer DMA instructions outstanding), its own LDS reads are done,
// and every wave got here.  Unlike __syncthreads() this does not wait for the
// newer slabs still in flight.  The waits are an s_waitcnt builtin, so the
// compiler's own wait insertion knows that no LDS read is outstanding after
// it (after a wait written as inline asm it assumed every earlier read might
// be, and waited for all of them, the next fragments' reads included, before
// the first MFMA of every slab).  The empty asm before the wait and the
// barrier's own asm carry memory clobbers, so no memory access moves across.
template <int n>
__device__ __forceinline__ void ring_barrier() {
    static_assert(n >= 0 && n < 64, "vmcnt range");
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt((n & 0xF) | ((n >> 4) << 14) | (0x7 << 4));   // vmcnt(n) lgkmcnt(0)
    asm volatile("s_barrier" ::: "memory");
}

template <int BM_, int BN_, int BK_, int STAGES_, int WGM_, int WGN_>
struct Core2 {
    static constexpr int BM = BM_, BN = BN_, BK = BK_, STAGES = STAGES_, WGM = WGM_, WGN = WGN_;
    static constexpr int WAVES = WGM * WGN, THREADS = 64 * WAVES;
    static constexpr int WM = BM / WGM, WN = BN / WGN, TM = WM / 16, TN = WN / 16;
    static constexpr int A_BYTES = BM * BK * 2, B_BYTES = BK * BN * 2;
    static constexpr int STAGE_BYTES = A_BYTES + B_BYTES;
    static constexpr int A_INSTR = A_BYTES / (1024 * WAVES);   // 1 KiB DMA instructions per wave per slab
    static constexpr int B_INSTR = B_BYTES / (1024 * WAVES);
    static constexpr int LOADS = A_INSTR + B_INSTR;
    static constexpr int KC = BK / 8, NC = BN / 8;
    static constexpr int A_ROWS_PER = 1024 / (BK * 2), B_ROWS_PER = 1024 / (BN * 2);
    static constexpr int SMEM = STAGES * STAGE_BYTES;
    static_assert(TM >= 1 && TN >= 1 && WM % 16 == 0 && WN % 16 == 0, "wave tile");
    static_assert(A_BYTES % (1024 * WAVES) == 0 && B_BYTES % (1024 * WAVES) == 0, "tile too small for the DMA split");
    static_assert(BK == 32 || BK == 64, "BK");
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

// interleave() with VA vector-ALU instructions spread between the MFMAs too
// (the dequantisation of the next k-step in the look-ahead w4 tile).
template <int MF, int DS, int VM, int VA, int Q = 0>
__device__ __forceinline__ void interleave_va() {
    if constexpr (Q < MF) {
        constexpr int ds = (DS * (Q + 1)) / MF - (DS * Q) / MF;
        constexpr int vm = (VM * (Q + 1)) / MF - (VM * Q) / MF;
        constexpr int va = (VA * (Q + 1)) / MF - (VA * Q) / MF;
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        if constexpr (ds > 0) __builtin_amdgcn_sched_group_barrier(0x100, ds, 0);
        if constexpr (va > 0) __builtin_amdgcn_sched_group_barrier(0x002, va, 0);
        if constexpr (vm > 0) __builtin_amdgcn_sched_group_barrier(0x020, vm, 0);
        interleave_va<MF, DS, VM, VA, Q + 1>();
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

// Core2Src::issue as a free function, for the calls from inside generic
// lambdas (the member-template call fails to resolve there in the host pass)
template <int Q0, int Q1, class S>
__device__ __forceinline__ void core2_issue(const S& src, char* smem, int stage, int slab) {
    src.template issue<Q0, Q1>(smem, stage, slab);
}

template <class C>
struct Core2Frags {
    v8h a[C::TM], b[C::TN];

    // fragment F of k-step kk (32 k): F < TM is A fragment F, else B fragment F - TM
    template <int F>
    __device__ __forceinline__ void load_one(const char* sa, const char* sb, int kk, int wm, int wn) {
        const int lane = threadIdx.x & 63;
        const int g = lane >> 4, gi = lane & 15;   // 16-lane group: k 8g..8g+7
        if constexpr (F < C::TM) {
            const int row = wm + 16 * F + gi;
            const int kc = (4 * kk + g) ^ swzA2<C::BK>(row);
            a[F] = *reinterpret_cast<const v8h*>(sa + row * (C::BK * 2) + kc * 16);
        } else {
            constexpr int j = F - C::TM;
            const int col = wn + 16 * j + 4 * (gi & 3);
            const int k1 = 32 * kk + 8 * g + (gi >> 2), k2 = k1 + 4;
            const int off1 = k1 * (C::BN * 2) + (((col >> 3) ^ swzB2(k1)) << 4) + (col & 7) * 2;
            const int off2 = k2 * (C::BN * 2) + (((col >> 3) ^ swzB2(k2)) << 4) + (col & 7) * 2;
            const v4s lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s_t*)(lds_char_t*)(sb + off1));
            const v4s hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s_t*)(lds_char_t*)(sb + off2));
            b[j] = __builtin_bit_cast(v8h, (v8s)__builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
        }
    }

    __device__ __forceinline__ void load(const char* sa, const char* sb, int kk, int wm, int wn) {
        static_for<0, C::TM + C::TN>([&](auto fc) { load_one<decltype(fc)::value>(sa, sb, kk, wm, wn); });
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
// Positions within a k-step of MF MFMAs: fragment load f (of NF) goes after
// MFMA floor(f MF / NF), DMA instruction d (of ND) after MFMA
// 1 + floor(d MF / ND), so both spread evenly and the DMA sits between reads.
template <int NF, int MF>
constexpr int frag_at(int q) {
    for (int f = 0; f < NF; f++)
        if ((f * MF) / NF == q) return f;
    return -1;
}
template <int ND, int MF>
constexpr int dma_at(int q) {
    for (int d = 0; d < ND; d++)
        if ((1 + (d * MF) / ND) % MF == q) return d;
    return -1;
}

// four fp16 (rounded to nearest) to an 8-byte-aligned global address; the
// global (not flat) store leaves the LDS counter alone
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) u32x2 global_u32x2_t;
typedef __attribute__((address_space(1))) u32x4 global_u32x4_t;
__device__ __forceinline__ void store_h4(__half* p, float a, float b, float c, float d) {
    const unsigned lo = (unsigned)__half_as_ushort(__float2half_rn(a)) |
                        ((unsigned)__half_as_ushort(__float2half_rn(b)) << 16);
    const unsigned hi = (unsigned)__half_as_ushort(__float2half_rn(c)) |
                        ((unsigned)__half_as_ushort(__float2half_rn(d)) << 16);
    *(global_u32x2_t*)(p) = u32x2{lo, hi};
}

// Stores a wave's WM x WN tile (alpha * acc, fp16) at c (row stride ldc,
// 16-byte aligned, ldc % 8 == 0) through its own slice of the idle ring:
// 8-byte LDS writes in the MFMA layout, 16-byte reads of whole row segments,
// 16-byte global stores (8 lanes cover 128 contiguous bytes of a row, so a
// store instruction writes whole cache lines instead of 32-byte pieces of 16).
// Rows go in passes of PR so that every wave's slice fits in the ring.
// Chunk q of row r sits at q ^ f(r), f(r) = (r / (16 / CH)) % CH, which keeps
// both the writes (32 lanes: 16 rows x 2 half-chunks) and the reads (16 lanes:
// 256 contiguous bytes) free of bank conflicts for rows of 64, 128 or 256 bytes.  LDS accesses of one wave
// execute in order, so no barrier is needed between the passes.
template <class C>
__device__ __forceinline__ void store_tile_staged(char* smem, const v4f (&acc)[C::TM][C::TN], __half* c, long ldc,
                                                  float alpha) {
    constexpr int WM = C::WM, WN = C::WN, CH = WN / 8;   // 16-byte chunks per staged row
    static_assert(CH == 4 || CH == 8 || CH == 16, "staged rows of 64, 128 or 256 bytes");
    constexpr int LIM = C::SMEM / (C::WAVES * WN * 2);
    constexpr int PR = WM <= LIM ? WM : (WM / 2 <= LIM ? WM / 2 : (WM / 4 <= LIM ? WM / 4 : 16));
    static_assert(PR % 16 == 0 && WM % PR == 0 && PR * WN * 2 * C::WAVES <= C::SMEM, "staging passes");
    constexpr int RPI = 64 / CH;   // rows per 16-byte read instruction
    // the lane index through an opaque copy: a persistent caller's loop would
    // otherwise hoist every lane-dependent address below out of the loop and
    // keep them all live (spilled) through the main loop
    int lane = threadIdx.x & 63;
    asm volatile("" : "+v"(lane));
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    char* st = smem + wave * (PR * WN * 2);
    const int m = lane & 15, g = lane >> 4;
#pragma unroll
    for (int pass = 0; pass < WM / PR; pass++) {
#pragma unroll
        for (int ii = 0; ii < PR / 16; ii++)
#pragma unroll
            for (int j = 0; j < C::TN; j++) {
                const int i = pass * (PR / 16) + ii;
                const int r = 16 * ii + m;                       // row within the pass
                const int q = 2 * j + (g >> 1), h = g & 1;       // chunk, 8-byte half
                const int qs = q ^ ((r / (16 / CH)) % CH);
                const unsigned lo = (unsigned)__half_as_ushort(__float2half_rn(alpha * acc[i][j][0])) |
                                    ((unsigned)__half_as_ushort(__float2half_rn(alpha * acc[i][j][1])) << 16);
                const unsigned hi = (unsigned)__half_as_ushort(__float2half_rn(alpha * acc[i][j][2])) |
                                    ((unsigned)__half_as_ushort(__float2half_rn(alpha * acc[i][j][3])) << 16);
                *reinterpret_cast<u32x2*>(st + r * (WN * 2) + qs * 16 + h * 8) = u32x2{lo, hi};
            }
#pragma unroll
        for (int t = 0; t < PR / RPI; t++) {
            const int r = RPI * t + lane / CH, q = lane % CH;
            const int qs = q ^ ((r / (16 / CH)) % CH);
            const u32x4 v = *reinterpret_cast<const u32x4*>(st + r * (WN * 2) + qs * 16);
            *(global_u32x4_t*)(c + (long)(pass * PR + r) * ldc + q * 8) = v;
        }
    }
}

struct NoHook {
    __device__ void operator()() const {}
};

// loop_done() runs once the products are in registers and the ring is idle,
// before the stores (a persistent caller takes its next work item there, so
// the ticket's latency hides behind the stores).
template <class C, bool SKIP = false, class Hook = NoHook>
__device__ __forceinline__ void gemm_tile2(char* smem, const __half* __restrict__ A, long lda,
                                           const __half* __restrict__ B, long ldb, int M, int N, int K, int m0,
                                           int n0, int s0, int s1, int split, int splits, const TileOut& out,
                                           const Hook& loop_done = Hook{}) {
    constexpr int TM = C::TM, TN = C::TN, BK = C::BK, STAGES = C::STAGES, KSTEPS = BK / 32;
    constexpr int MFMAS = TM * TN, NF = TM + TN;   // per k-step: MFMAs, fragment loads
    constexpr int D = C::LOADS / KSTEPS;            // DMA instructions per k-step
    static_assert(C::LOADS % KSTEPS == 0, "DMA chunks");
    static_assert(NF <= MFMAS && D <= MFMAS, "at most one fragment load and one DMA instruction per MFMA");
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // SKIP (for many small problems, where edge tiles are common): waves w
    // and w + 4 share a SIMD, and with eight waves (WGM even) the upper half
    // takes the wave tiles of the lower half's partners in the other row half
    // and, reversed, the other column half; so a tile that crosses the
    // problem's last row or column in its first half leaves every SIMD one
    // wave with work instead of two on some SIMDs and none on others.  A wave
    // whose whole WM x WN sub-tile lies outside the problem then skips its
    // fragment reads and MFMAs (it still issues its DMA share and takes every
    // barrier).  Without SKIP the roles are in wave order and every wave
    // multiplies (about 1 % faster on whole tiles).
    int wr = wave / C::WGN, wc = wave % C::WGN;
    if constexpr (SKIP && C::WAVES == 8 && C::WGM % 2 == 0) {
        if (wave >= 4) {
            wr = (wave - 4) / C::WGN + C::WGM / 2;
            wc = C::WGN - 1 - (wave - 4) % C::WGN;
        }
    }
    const int wm = wr * C::WM, wn = wc * C::WN;
    const bool active = !SKIP || (m0 + wm < M && n0 + wn < N);

    v4f acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) acc[i][j] = v4f{0.0f, 0.0f, 0.0f, 0.0f};

    // full slabs go through the ring; a partial last slab (K % BK != 0) is
    // multiplied afterwards with the k beyond K masked out
    const int rem = K - (s1 - 1) * BK;   // valid k in the last slab of the range
    const bool partial = s1 > s0 && rem < BK;
    const int nslabs = s1 - s0 - (partial ? 1 : 0);
    if (nslabs > 0) {
        Core2Src<C> src;
        src.init(A, lda, B, ldb, M, N, m0, n0, s0 * BK);
        // Older memory operations still in flight (a persistent caller's stores
        // of its previous tile) only make the ring's vmcnt waits stricter: a wait
        // for at most n outstanding leaves at most n DMA instructions outstanding.
        // So they drain while this tile's first slabs are fetched.
        // Ring of STAGES slabs.  During slab s, slab s+STAGES-1 is fetched into the
        // stage slab s-1 used (chunk kk of its DMA during k-step kk).  Every k-step
        // multiplies fragments loaded during the k-step before and loads the next
        // k-step's fragments; before the first read of a new stage (at the start of
        // the last k-step of each slab) a barrier makes sure that slab landed for
        // every wave, and that every wave is done with the stage the next DMA
        // overwrites.  Past the end the last slab is fetched again into a stage
        // nobody reads, so the vmcnt arithmetic never changes.
#pragma unroll
        for (int p = 0; p < STAGES - 1; p++) src.template issue<0, C::LOADS>(smem, p, p < nslabs ? p : nslabs - 1);
        ring_barrier<C::LOADS * (STAGES - 2)>();

        Core2Frags<C> fr[2];
        if (active) fr[0].load(smem, smem + C::A_BYTES, 0, wm, wn);
        int stage = 0;                  // s % STAGES
        int fill_stage = STAGES - 1;    // (s + STAGES - 1) % STAGES
        // one k-step on fr[CUR], loading fr[CUR ^ 1] from (la, lb, lk) and
        // issuing DMA chunk KK of slab `fill`; the order of MFMAs, fragment loads
        // and DMA instructions is fixed by a scheduling fence after each MFMA
        // (ACT false: the DMA instructions and barriers only)
        auto kstep = [&](auto actc, auto curc, auto kkc, const char* la, const char* lb, int lk, int fill) {
            constexpr bool ACT = decltype(actc)::value != 0;
            constexpr int CUR = decltype(curc)::value, KK = decltype(kkc)::value;
            if constexpr (KK == KSTEPS - 1)   // slab s+1 landed, newest DMA: chunks 0..KK-1 of slab s+STAGES-1
                ring_barrier<C::LOADS * (STAGES - 3) + KK * D>();
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (ACT) {
                static_for<0, MFMAS>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    constexpr int i = q / TN, j = q % TN;
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fr[CUR].b[j], fr[CUR].a[i], acc[i][j], 0, 0, 0);
                    // loads in the order of first use by the next k-step: the B
                    // fragments and A fragment 0 feed its first TN MFMAs
                    constexpr int f = frag_at<NF, MFMAS>(q);
                    if constexpr (f >= 0)
                        fr[CUR ^ 1].template load_one<(f < TN ? TM + f : f - TN)>(la, lb, lk, wm, wn);
                    constexpr int dd = dma_at<D, MFMAS>(q);
                    if constexpr (dd >= 0) {
                        constexpr int d = KK * D + dd;
                        core2_issue<d, d + 1>(src, smem, fill_stage, fill);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                });
            } else {
                static_for<0, D>([&](auto dc) {
                    constexpr int d = KK * D + decltype(dc)::value;
                    core2_issue<d, d + 1>(src, smem, fill_stage, fill);
                });
                __builtin_amdgcn_sched_barrier(0);
            }
        };
        auto slab = [&](auto actc, auto c0, int s) {
            constexpr int C0 = decltype(c0)::value;   // fragment buffer of the slab's first k-step
            const char* sa = smem + stage * C::STAGE_BYTES;
            const int fill = s + STAGES - 1 < nslabs ? s + STAGES - 1 : nslabs - 1;
            const int nstage = stage + 1 == STAGES ? 0 : stage + 1;
            const char* na = smem + nstage * C::STAGE_BYTES;
            if constexpr (KSTEPS == 1) {
                kstep(actc, int_c<C0>{}, int_c<0>{}, na, na + C::A_BYTES, 0, fill);
            } else {
                kstep(actc, int_c<C0>{}, int_c<0>{}, sa, sa + C::A_BYTES, 1, fill);
                kstep(actc, int_c<C0 ^ 1>{}, int_c<1>{}, na, na + C::A_BYTES, 0, fill);
            }
            stage = nstage;
            fill_stage = fill_stage + 1 == STAGES ? 0 : fill_stage + 1;
        };
        auto slabs = [&](auto actc) {
            if constexpr (KSTEPS == 2) {
                for (int s = 0; s < nslabs; s++) slab(actc, int_c<0>{}, s);
            } else {
                int s = 0;
                for (; s + 1 < nslabs; s += 2) {
                    slab(actc, int_c<0>{}, s);
                    slab(actc, int_c<1>{}, s + 1);
                }
                if (s < nslabs) slab(actc, int_c<0>{}, s);
            }
        };
        if (active)
            slabs(int_c<1>{});
        else
            slabs(int_c<0>{});
        wait_vmcnt<0>();   // the redundant tail DMA must land before the LDS is reused
    }
    if (partial) {
        const int k0 = (s1 - 1) * BK;
        ring_barrier<0>();   // every wave is done with the ring
        core2_issue_clamped<C>(smem, A, lda, B, ldb, M, N, K, m0, n0, k0);
        ring_barrier<0>();
        const int kl = 8 * (lane >> 4);   // first k of this lane's 8 in a k-step
#pragma unroll
        for (int kk = 0; kk < KSTEPS && active; kk++) {
            Core2Frags<C> f;
            f.load(smem, smem + C::A_BYTES, kk, wm, wn);
            if (32 * kk + kl >= rem) {
#pragma unroll
                for (int i = 0; i < TM; i++) f.a[i] = v8h{};
#pragma unroll
                for (int j = 0; j < TN; j++) f.b[j] = v8h{};
            }
#pragma unroll
            for (int i = 0; i < TM; i++)
#pragma unroll
                for (int j = 0; j < TN; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(f.b[j], f.a[i], acc[i][j], 0, 0, 0);
        }
    }
    loop_done();
    __syncthreads();   // the caller may refill the ring for its next tile

    // lane l: row m0 + wm + 16 i + l % 16, columns n0 + wn + 16 j + 4 (l / 16) .. +3
    const int frow = lane & 15, fcol4 = 4 * (lane >> 4);
    if (splits <= 1) {
        if (!active) return;   // nothing of this wave's sub-tile is in the problem
        // 8-byte stores need C 8-byte aligned and ldc % 4 == 0 (N % 8 == 0 keeps
        // every group of four columns inside the matrix)
        const bool vec = reinterpret_cast<uintptr_t>(out.C) % 8 == 0 && out.ldc % 4 == 0;
        const bool whole = out.beta == 0.0f && m0 + C::BM <= M && n0 + C::BN <= N;
        if (whole && reinterpret_cast<uintptr_t>(out.C) % 16 == 0 && out.ldc % 8 == 0) {
            store_tile_staged<C>(smem, acc, out.C + (long)(m0 + wm) * out.ldc + n0 + wn, out.ldc, out.alpha);
            return;
        }
        if (whole && vec) {   // no branches
#pragma unroll
            for (int i = 0; i < TM; i++)
#pragma unroll
                for (int j = 0; j < TN; j++) {
                    const long row = m0 + wm + 16 * i + frow;
                    const int col = n0 + wn + 16 * j + fcol4;
                    store_h4(out.C + row * out.ldc + col, out.alpha * acc[i][j][0], out.alpha * acc[i][j][1],
                             out.alpha * acc[i][j][2], out.alpha * acc[i][j][3]);
                }
            return;
        }
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int j = 0; j < TN; j++) {
                const int row = m0 + wm + 16 * i + frow;
                const int col = n0 + wn + 16 * j + fcol4;
                if (row >= M || col >= N) continue;
                __half* cp = out.C + (long)row * out.ldc + col;
                float v[4];
#pragma unroll
                for (int r = 0; r < 4; r++) v[r] = out.alpha * acc[i][j][r];
                if (vec && col + 3 < N) {
                    if (out.beta != 0.0f) {
                        const uint2 old = *reinterpret_cast<const uint2*>(cp);
                        v[0] = fmaf(out.beta, __half2float(__ushort_as_half((unsigned short)(old.x & 0xFFFF))), v[0]);
                        v[1] = fmaf(out.beta, __half2float(__ushort_as_half((unsigned short)(old.x >> 16))), v[1]);
                        v[2] = fmaf(out.beta, __half2float(__ushort_as_half((unsigned short)(old.y & 0xFFFF))), v[2]);
                        v[3] = fmaf(out.beta, __half2float(__ushort_as_half((unsigned short)(old.y >> 16))), v[3]);
                    }
                    store_h4(cp, v[0], v[1], v[2], v[3]);
                } else {
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        if (col + r < N) {
                            float x = v[r];
                            if (out.beta != 0.0f) x = fmaf(out.beta, __half2float(cp[r]), x);
                            cp[r] = __float2half_rn(x);
                        }
                }
            }
        return;
    }
    float* mine = out.partial + (long)split * M * N;
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) {
            const int row = m0 + wm + 16 * i + frow;
            const int col = n0 + wn + 16 * j + fcol4;
            if (row >= M) continue;
#pragma unroll
            for (int r = 0; r < 4; r++)
                if (col + r < N) publish(mine + (long)row * N + col + r, acc[i][j][r]);
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
