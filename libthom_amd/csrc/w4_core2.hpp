// Pipelined w4a16 tile: C (BM x BN) = A (row-major fp16) * dequant(W int4).
//
// The same DMA ring as gemm_core2.hpp, with the weight side replaced:
//   A tile   [BM][64] fp16, swizzled as in gemm_core2.hpp;
//   W tile   [8][BN] packed uint32 (8 consecutive k of one column each), as in
//            memory, 16-byte chunk c of row r at c ^ 8(r & 1) so that the two
//            rows one wave reads together fall in different LDS banks;
//   SZ block, 1 KiB per wave column group: the BN/WGN scales of the slab's
//            group (fp16) followed by their packed zero-points.
// Every wave issues one DMA instruction per slab for its SZ block (a DMA
// instruction writes 16 bytes for each of the 64 lanes, hence 1 KiB; waves of
// the same columns write identical bytes), so all waves count the same number
// of DMA instructions and the vmcnt arithmetic of the ring is unchanged.
//
// One uint32 of W is one MFMA B fragment: lane l of n-tile j reads row
// 2 kk + l/32, column wn + 32 j + l%32, and dequantises it in registers to
// fp16((q - z) * s) in natural k order (dequant8_nat, one rounding), so the A
// fragment goes from LDS to the MFMA as it is.
//
// Requirements (checked by the caller): group % 64 == 0, K % group == 0,
// N % 32 == 0, M >= 1; A, W, scales and qzeros 16-byte aligned.
#pragma once

#include "gemm_core2.hpp"

namespace lt {

__device__ __forceinline__ int swzW2(int r) { return 8 * (r & 1); }

template <int BM_, int BN_, int STAGES_, int WGM_, int WGN_>
struct Core2W4 {
    static constexpr int BM = BM_, BN = BN_, BK = 64, STAGES = STAGES_, WGM = WGM_, WGN = WGN_;
    static constexpr int WAVES = WGM * WGN, THREADS = 64 * WAVES;
    static constexpr int WM = BM / WGM, WN = BN / WGN, TM = WM / 32, TN = WN / 32;
    static constexpr int A_BYTES = BM * BK * 2, B_BYTES = (BK / 8) * BN * 4;
    static constexpr int SZ_SCALE = 2 * WN, SZ_ZERO = WN / 2, SZ_STRIDE = 1024;
    static constexpr int SZ_BYTES = WGN * SZ_STRIDE;
    static constexpr int STAGE_BYTES = A_BYTES + B_BYTES + SZ_BYTES;
    static constexpr int A_INSTR = A_BYTES / (1024 * WAVES), B_INSTR = B_BYTES / (1024 * WAVES);
    static constexpr int LOADS = A_INSTR + B_INSTR + 1;
    static constexpr int KC = BK / 8;
    static constexpr int A_ROWS_PER = 1024 / (BK * 2);
    static constexpr int NC = BN * 4 / 16;                  // 16-byte chunks per packed row
    static constexpr int B_ROWS_PER = 1024 / (BN * 4);
    static constexpr int SZ_SLANES = SZ_SCALE / 16, SZ_ZLANES = SZ_ZERO / 16;
    static constexpr int SMEM = STAGES * STAGE_BYTES;
    static_assert(TM >= 2 && TM % 2 == 0 && TN >= 1, "wave tile");
    static_assert(A_BYTES % (1024 * WAVES) == 0 && B_BYTES % (1024 * WAVES) == 0, "DMA split");
    static_assert(BN >= 64 && BN <= 256 && B_ROWS_PER >= 1, "BN");
    static_assert(SZ_ZERO % 16 == 0 && SZ_SLANES + SZ_ZLANES <= 64, "SZ block");
    static_assert(STAGES >= 3 && SMEM <= 160 * 1024, "LDS budget");
    static_assert(LOADS * (STAGES - 1) < 64, "vmcnt range");
};

template <class C>
struct W4Src {
    const __half* a[C::A_INSTR];
    const uint32_t* w[C::B_INSTR];
    const char* sz;          // this lane's SZ source in group 0 (scales or zero-points)
    long sz_group_step;      // bytes between consecutive groups for that source
    long w_step;             // uint32 between consecutive slabs of W
    int group_slabs;         // slabs per quantisation group
    int slab0;               // absolute index of slab 0 of this range

    __device__ __forceinline__ void init(const __half* __restrict__ A, long lda, const BW4Params& p, int M, int N,
                                         int m0, int n0, int k0) {
        const int lane = threadIdx.x & 63;
        const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
#pragma unroll
        for (int i = 0; i < C::A_INSTR; i++) {
            const int inst = wave * C::A_INSTR + i;
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
            const int ch = (lane % C::NC) ^ swzW2(row);
            int gn = n0 + ch * 4;
            gn = gn < N ? gn : N - 4;
            w[i] = p.W + (long)(k0 / 8 + row) * N + gn;
        }
        w_step = (long)(C::BK / 8) * N;
        group_slabs = p.group / C::BK;
        slab0 = k0 / C::BK;
        // SZ: lanes [0, SZ_SLANES) fetch 8 scales each, the next SZ_ZLANES 4
        // packed zero-point words each (32 columns); the rest repeat lane 0
        const int wn = (wave % C::WGN) * C::WN;
        const int sl = lane < C::SZ_SLANES + C::SZ_ZLANES ? lane : 0;
        if (sl < C::SZ_SLANES || !p.qzeros) {
            int gn = n0 + wn + 8 * (sl < C::SZ_SLANES ? sl : 0);
            gn = gn < N ? gn : N - 8;
            sz = reinterpret_cast<const char*>(p.scales + gn);
            sz_group_step = (long)N * 2;
        } else {
            int zw = (n0 + wn) / 8 + 4 * (sl - C::SZ_SLANES);
            zw = zw < N / 8 ? zw : N / 8 - 4;
            sz = reinterpret_cast<const char*>(p.qzeros + zw);
            sz_group_step = (long)(N / 8) * 4;
        }
    }

    template <int Q0, int Q1>
    __device__ __forceinline__ void issue(char* smem, int stage, int slab) const {
        const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        char* sa = smem + stage * C::STAGE_BYTES;
        char* sb = sa + C::A_BYTES;
        char* ss = sb + C::B_BYTES;
#pragma unroll
        for (int q = Q0; q < Q1; q++) {
            if (q < C::A_INSTR) {
                __builtin_amdgcn_global_load_lds(a[q] + (long)slab * C::BK,
                                                 (lds_void_t*)(lds_char_t*)(sa + (wave * C::A_INSTR + q) * 1024), 16,
                                                 0, 0);
            } else if (q < C::A_INSTR + C::B_INSTR) {
                const int i = q - C::A_INSTR;
                __builtin_amdgcn_global_load_lds(w[i] + (long)slab * w_step,
                                                 (lds_void_t*)(lds_char_t*)(sb + (wave * C::B_INSTR + i) * 1024), 16,
                                                 0, 0);
            } else {
                const int cg = wave % C::WGN;
                __builtin_amdgcn_global_load_lds(sz + (long)((slab0 + slab) / group_slabs) * sz_group_step,
                                                 (lds_void_t*)(lds_char_t*)(ss + cg * C::SZ_STRIDE), 16, 0, 0);
            }
        }
    }
};

// W4Src for the look-ahead tiles: A and W through buffer loads to LDS, the
// per-lane offsets fixed at init and the slab's offset a scalar (soffset), so
// advancing a slab costs no vector instructions.  Needs A's rows of the tile
// and W below 2^31 bytes (checked by the launch).
template <class C>
struct W4SrcB {
    __amdgpu_buffer_rsrc_t ra, rw;
    int va[C::A_INSTR], vw[C::B_INSTR];
    int w_slab_bytes;
    W4Src<C> base;   // the scale / zero-point source (one pointer load a slab)

    __device__ __forceinline__ void init(const __half* __restrict__ A, long lda, const BW4Params& p, int M, int N,
                                         int m0, int n0, int k0) {
        const int lane = threadIdx.x & 63;
        const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        base.init(A, lda, p, M, N, m0, n0, k0);
        ra = __builtin_amdgcn_make_buffer_rsrc((void*)(A + (long)m0 * lda + k0), 0, 0x7FFFFFFF, 0x00020000);
        rw = __builtin_amdgcn_make_buffer_rsrc((void*)(p.W + (long)(k0 / 8) * N + n0), 0, 0x7FFFFFFF, 0x00020000);
#pragma unroll
        for (int i = 0; i < C::A_INSTR; i++) {
            const int inst = wave * C::A_INSTR + i;
            const int row = inst * C::A_ROWS_PER + lane / C::KC;
            const int kc = (lane % C::KC) ^ swzA2<C::BK>(row);
            int gm = m0 + row;
            gm = gm < M ? gm : M - 1;
            va[i] = (int)(((long)(gm - m0) * lda + kc * 8) * 2);
        }
#pragma unroll
        for (int i = 0; i < C::B_INSTR; i++) {
            const int inst = wave * C::B_INSTR + i;
            const int row = inst * C::B_ROWS_PER + lane / C::NC;
            const int ch = (lane % C::NC) ^ swzW2(row);
            int gn = n0 + ch * 4;
            gn = gn < N ? gn : N - 4;
            vw[i] = (int)(((long)row * N + (gn - n0)) * 4);
        }
        w_slab_bytes = (C::BK / 8) * N * 4;
    }

    template <int Q0, int Q1>
    __device__ __forceinline__ void issue(char* smem, int stage, int slab) const {
        const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        char* sa = smem + stage * C::STAGE_BYTES;
        char* sb = sa + C::A_BYTES;
        char* ss = sb + C::B_BYTES;
#pragma unroll
        for (int q = Q0; q < Q1; q++) {
            if (q < C::A_INSTR) {
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    ra, (lds_void_t*)(lds_char_t*)(sa + (wave * C::A_INSTR + q) * 1024), 16, va[q], slab * C::BK * 2,
                    0, 0);
            } else if (q < C::A_INSTR + C::B_INSTR) {
                const int i = q - C::A_INSTR;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rw, (lds_void_t*)(lds_char_t*)(sb + (wave * C::B_INSTR + i) * 1024), 16, vw[i],
                    slab * w_slab_bytes, 0, 0);
            } else {
                const int cg = wave % C::WGN;
                __builtin_amdgcn_global_load_lds(base.sz + (long)((base.slab0 + slab) / base.group_slabs) *
                                                               base.sz_group_step,
                                                 (lds_void_t*)(lds_char_t*)(ss + cg * C::SZ_STRIDE), 16, 0, 0);
            }
        }
    }
};

template <class C>
struct W4Frags {
    v8h a[C::TM];
    uint32_t w[C::TN];   // packed int4 words, dequantised where they are used (dq)

    __device__ __forceinline__ void load(const char* sa, const char* sb, int kk, int wm, int wn) {
        const int lane = threadIdx.x & 63;
        const int frow = lane & 31, fkc = lane >> 5;
#pragma unroll
        for (int i = 0; i < C::TM; i++) {
            const int row = wm + 32 * i + frow;
            const int kc = (2 * kk + fkc) ^ swzA2<C::BK>(row);
            const uint4 raw = *reinterpret_cast<const uint4*>(sa + row * (C::BK * 2) + kc * 16);
            __builtin_memcpy(&a[i], &raw, 16);
        }
        const int r = 2 * kk + fkc;
#pragma unroll
        for (int j = 0; j < C::TN; j++) {
            const int c = wn + 32 * j + frow;
            w[j] = *reinterpret_cast<const uint32_t*>(sb + r * (C::BN * 4) + (((c >> 2) ^ swzW2(r)) << 4) + (c & 3) * 4);
        }
    }
};

// B fragment j of f: fp16((q - z) * s) in natural k order; zz = zz_nat(z),
// s2 = the scale in both halves
template <class C>
__device__ __forceinline__ v8h dq(const W4Frags<C>& f, int j, const uint32_t* zz, const uint32_t* s2) {
    const uint4 d = dequant8_nat(f.w[j], zz[j], s2[j]);
    v8h b;
    __builtin_memcpy(&b, &d, 16);
    return b;
}

// scales and zero-points of this lane's columns for the slab in stage `ss`
// (without zero-points, zmask 0 and zor 8 make every z 8: selected
// arithmetically, since a branch here splits the k-loop into blocks whose
// joins cost the compiler's wait insertion its knowledge of the LDS reads)
template <class C>
__device__ __forceinline__ void w4_read_sz(const char* ss, int cg, uint32_t zmask, uint32_t zor, uint32_t* zz,
                                           uint32_t* s2) {
    const int lane = threadIdx.x & 63;
    const char* blk = ss + cg * C::SZ_STRIDE;
#pragma unroll
    for (int j = 0; j < C::TN; j++) {
        const int c = 32 * j + (lane & 31);   // column within the wave's WN
        const uint32_t s = *reinterpret_cast<const uint16_t*>(blk + 2 * c);
        const uint32_t zw = *reinterpret_cast<const uint32_t*>(blk + C::SZ_SCALE + 4 * (c >> 3));
        zz[j] = zz_nat(((zw >> (4 * (c & 7))) & zmask) | zor);
        s2[j] = s | (s << 16);
    }
}

// Owner-collects hand-off of one split tile (out.fpart set).  Split s of the
// tile writes its accumulators to plane s in fragment order: 16 floats per
// lane and MFMA tile, 16-byte stores with the cache policy of an agent-scope
// atomic store (sc1: written through past the XCD's L2; each instruction
// writes 1 KiB in one piece), then drains them and
// adds one to the tile's arrival counter (the ordering argument of publish /
// drain / ticket).  The last split is the owner: it keeps its own sums in
// registers, waits until the counter reads splits - 1, reads the other planes
// (sc1 loads, issued only after that value came back and passed a barrier)
// and adds them in split order, its own last: ((p0 + p1) + ...) + own, so the
// result does not depend on arrival order.  It resets the counter for the next
// call.  Returns true in the owner (acc then holds the tile's sum).
// The wait is for pieces that run before the owner's own piece in their
// blocks, in blocks dispatched before the owner's; it gives up after ~2^24
// polls, marking the tile with NaN instead of hanging.
template <class C>
__device__ __forceinline__ bool w4_owner_collect(char* smem, v16f (&acc)[C::TM][C::TN], const TileOut& out,
                                                 int split, int splits) {
    constexpr int TM = C::TM, TN = C::TN, SC1 = 16;
    constexpr int PLANE = C::BM * C::BN * 4;   // bytes
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // 16-byte piece q of fragment f (= wave TM TN + i TN + j) of lane l at
    // ((4 f + q) 64 + l) 16 bytes: each store / load instruction covers 1 KiB
    const int voff = (wave * TM * TN * 256 + lane) * 16;   // + ((i TN + j) 4 + q) * 1024
    if (split < splits - 1) {
        const __amdgpu_buffer_rsrc_t pr =
            __builtin_amdgcn_make_buffer_rsrc((void*)(out.fpart + (long)split * out.fstride), 0, PLANE, 0x00020000);
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int j = 0; j < TN; j++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const v4f v = {acc[i][j][4 * q], acc[i][j][4 * q + 1], acc[i][j][4 * q + 2], acc[i][j][4 * q + 3]};
                    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4i, v), pr,
                                                           voff + ((i * TN + j) * 4 + q) * 1024, 0, SC1);
                }
        drain();
        __syncthreads();
        if (threadIdx.x == 0) take_ticket(out.counter);
        return false;
    }
    // the flag lives in the idle ring: a second LDS object would cost the
    // compiler's wait insertion its tracking of the ring's DMA in the k-loop
    int* ok_flag = reinterpret_cast<int*>(smem);
    if (threadIdx.x == 0) {
        int ok = 0;
        for (int poll = 0; poll < (1 << 24); poll++) {
            if (__hip_atomic_load(out.counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == (unsigned)(splits - 1)) {
                ok = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        if (ok) reset_counter(out.counter);
        *ok_flag = ok;
    }
    __syncthreads();
    if (!*ok_flag) {
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int j = 0; j < TN; j++)
#pragma unroll
                for (int r = 0; r < 16; r++) acc[i][j][r] = __builtin_nanf("");
        return true;
    }
    v16f sum[TM][TN];
    for (int sp = 0; sp < splits - 1; sp++) {
        const __amdgpu_buffer_rsrc_t pr =
            __builtin_amdgcn_make_buffer_rsrc((void*)(out.fpart + (long)sp * out.fstride), 0, PLANE, 0x00020000);
        v4f v[TM][TN][4];
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int j = 0; j < TN; j++)
#pragma unroll
                for (int q = 0; q < 4; q++)
                    v[i][j][q] = __builtin_bit_cast(
                        v4f, __builtin_amdgcn_raw_buffer_load_b128(pr, voff + ((i * TN + j) * 4 + q) * 1024, 0, SC1));
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int j = 0; j < TN; j++)
#pragma unroll
                for (int q = 0; q < 4; q++)
#pragma unroll
                    for (int e = 0; e < 4; e++)
                        sum[i][j][4 * q + e] = sp == 0 ? v[i][j][q][e] : sum[i][j][4 * q + e] + v[i][j][q][e];
    }
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[i][j][r] = sum[i][j][r] + acc[i][j][r];
    return true;
}

// Main loop of the look-ahead tile (w4_tile2 with LA): nslabs >= 1 slabs
// from slab s0, accumulated into acc.
template <class C>
__device__ __forceinline__ void w4_mainloop_la(char* smem, const __half* __restrict__ A, long lda, const BW4Params& p,
                                               int M, int N, int m0, int n0, int s0, int nslabs, uint32_t zmask,
                                               uint32_t zor, v16f (&acc)[C::TM][C::TN]) {
    constexpr int TM = C::TM, TN = C::TN, BK = C::BK, STAGES = C::STAGES, KSTEPS = BK / 16;
    constexpr int MFMAS = TM * TN;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int cg = wave % C::WGN;
    const int wm = (wave / C::WGN) * C::WM, wn = cg * C::WN;
    W4SrcB<C> src;
    src.init(A, lda, p, M, N, m0, n0, s0 * BK);
    wait_vmcnt<0>();
#pragma unroll
    for (int q = 0; q < STAGES - 1; q++) src.template issue<0, C::LOADS>(smem, q, q < nslabs ? q : nslabs - 1);
    ring_barrier<C::LOADS * (STAGES - 2)>();
    const int frow = lane & 31, fkc = lane >> 5;
    // A fragments of k-step kk (two buffers) and weight words (two
    // buffers), each read straight from the stage as W4Frags::load does
    auto load_a = [&](v8h* a, const char* sa, int kk) {
#pragma unroll
        for (int i = 0; i < TM; i++) {
            const int row = wm + 32 * i + frow;
            const int kc = (2 * kk + fkc) ^ swzA2<C::BK>(row);
            const uint4 raw = *reinterpret_cast<const uint4*>(sa + row * (C::BK * 2) + kc * 16);
            __builtin_memcpy(&a[i], &raw, 16);
        }
    };
    auto load_w = [&](uint32_t* w, const char* sb, int kk) {
        const int r = 2 * kk + fkc;
#pragma unroll
        for (int j = 0; j < TN; j++) {
            const int c = wn + 32 * j + frow;
            w[j] = *reinterpret_cast<const uint32_t*>(sb + r * (C::BN * 4) + (((c >> 2) ^ swzW2(r)) << 4) +
                                                      (c & 3) * 4);
        }
    };
    auto dq_all = [&](v8h* b, const uint32_t* w, const uint32_t* zz, const uint32_t* s2) {
#pragma unroll
        for (int j = 0; j < TN; j++) {
            const uint4 d = dequant8_nat(w[j], zz[j], s2[j]);
            __builtin_memcpy(&b[j], &d, 16);
        }
    };
    uint32_t zz[TN], s2[TN];
    v8h a[2][TM], bq[2][TN];
    uint32_t wr[2][TN];
    w4_read_sz<C>(smem + C::A_BYTES + C::B_BYTES, cg, zmask, zor, zz, s2);
    load_a(a[0], smem, 0);
    load_w(wr[0], smem + C::A_BYTES, 0);
    load_w(wr[1], smem + C::A_BYTES, 1);
    dq_all(bq[0], wr[0], zz, s2);
    int stage = 0, fill_stage = STAGES - 1;
    for (int s = 0; s < nslabs; s++) {
        const char* sa = smem + stage * C::STAGE_BYTES;
        const char* sb = sa + C::A_BYTES;
        const int fill = s + STAGES - 1 < nslabs ? s + STAGES - 1 : nslabs - 1;
        __builtin_amdgcn_sched_barrier(0);
        // k-steps 0 .. KSTEPS-2: read A(kk+1) and W(kk+2), dequantise
        // W(kk+1), MFMAs of kk
        static_for<0, KSTEPS - 1>([&](auto kkl) {
            constexpr int kk = decltype(kkl)::value;
            load_a(a[(kk + 1) & 1], sa, kk + 1);
            if constexpr (kk + 2 < KSTEPS) load_w(wr[kk & 1], sb, kk + 2);
            src.template issue<kk * C::LOADS / KSTEPS, (kk + 1) * C::LOADS / KSTEPS>(smem, fill_stage, fill);
            dq_all(bq[(kk + 1) & 1], wr[(kk + 1) & 1], zz, s2);
#pragma unroll
            for (int j = 0; j < TN; j++)
#pragma unroll
                for (int i = 0; i < TM; i++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[kk & 1][i], bq[kk & 1][j], acc[i][j], 0,
                                                                       0, 0);
            interleave_va<MFMAS, TM + (kk + 2 < KSTEPS ? TN : 0),
                          (kk + 1) * C::LOADS / KSTEPS - kk * C::LOADS / KSTEPS, 16 * TN>();
            __builtin_amdgcn_sched_barrier(0);
        });
        {
            constexpr int KL = KSTEPS - 1;
            src.template issue<KL * C::LOADS / KSTEPS, C::LOADS>(smem, fill_stage, fill);
#pragma unroll
            for (int i = 0; i < TM / 2; i++)
#pragma unroll
                for (int j = 0; j < TN; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[KL & 1][i], bq[KL & 1][j], acc[i][j], 0,
                                                                       0, 0);
            interleave<MFMAS / 2, 0, C::LOADS - KL * C::LOADS / KSTEPS>();
            __builtin_amdgcn_sched_barrier(0);
            ring_barrier<C::LOADS * (STAGES - 2)>();
            stage = stage + 1 == STAGES ? 0 : stage + 1;
            fill_stage = fill_stage + 1 == STAGES ? 0 : fill_stage + 1;
            const char* na = smem + stage * C::STAGE_BYTES;
            w4_read_sz<C>(na + C::A_BYTES + C::B_BYTES, cg, zmask, zor, zz, s2);
            load_a(a[0], na, 0);   // KSTEPS is even: k-step 0 uses buffer 0
            load_w(wr[0], na + C::A_BYTES, 0);
            load_w(wr[1], na + C::A_BYTES, 1);
#pragma unroll
            for (int i = TM / 2; i < TM; i++)
#pragma unroll
                for (int j = 0; j < TN; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[KL & 1][i], bq[KL & 1][j], acc[i][j], 0,
                                                                       0, 0);
            interleave<MFMAS / 2, TM + 2 * TN, 0>();
            __builtin_amdgcn_sched_barrier(0);
            dq_all(bq[0], wr[0], zz, s2);
        }
    }
    wait_vmcnt<0>();
}

// C tile at (m0, n0) over K-slabs [s0, s1) of 64 (K % 64 == 0); split-K
// protocol as gemm_tile2 (alpha 1, beta 0), or the owner-collects hand-off
// above when out.fpart is set.
//
// LA (look-ahead): the weight words are read two k-steps ahead and
// dequantised one k-step ahead, between the MFMAs of the current k-step, so
// the vector work of a wave runs under its own MFMAs (with one wave per SIMD
// nothing else fills the MFMA pipe while it dequantises).  Only the first
// k-step of each slab, whose words and scales arrive after the ring barrier,
// is dequantised after its reads.
template <class C, bool LA = false>
__device__ __forceinline__ void w4_tile2(char* smem, const __half* __restrict__ A, long lda, const BW4Params& p, int M,
                                         int N, int m0, int n0, int s0, int s1, int split, int splits,
                                         const TileOut& out) {
    constexpr int TM = C::TM, TN = C::TN, BK = C::BK, STAGES = C::STAGES, KSTEPS = BK / 16;
    constexpr int MFMAS = TM * TN, READS = TM + TN;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int cg = wave % C::WGN;
    const int wm = (wave / C::WGN) * C::WM, wn = cg * C::WN;
    const uint32_t zmask = p.qzeros ? 0xFu : 0u, zor = p.qzeros ? 0u : 8u;

    v16f acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[i][j][r] = 0.0f;

    const int nslabs = s1 - s0;
    if constexpr (LA) {
        if (nslabs > 0) w4_mainloop_la<C>(smem, A, lda, p, M, N, m0, n0, s0, nslabs, zmask, zor, acc);
    } else if (nslabs > 0) {
        W4Src<C> src;
        src.init(A, lda, p, M, N, m0, n0, s0 * BK);
        wait_vmcnt<0>();
#pragma unroll
        for (int q = 0; q < STAGES - 1; q++) src.template issue<0, C::LOADS>(smem, q, q < nslabs ? q : nslabs - 1);
        ring_barrier<C::LOADS * (STAGES - 2)>();

        uint32_t zz[TN], s2[TN];
        W4Frags<C> fr[2];
        w4_read_sz<C>(smem + C::A_BYTES + C::B_BYTES, cg, zmask, zor, zz, s2);
        fr[0].load(smem, smem + C::A_BYTES, 0, wm, wn);
        int stage = 0, fill_stage = STAGES - 1;
        for (int s = 0; s < nslabs; s++) {
            const char* sa = smem + stage * C::STAGE_BYTES;
            const char* sb = sa + C::A_BYTES;
            const int fill = s + STAGES - 1 < nslabs ? s + STAGES - 1 : nslabs - 1;
            __builtin_amdgcn_sched_barrier(0);
            static_for<0, KSTEPS - 1>([&](auto kkc) {
                constexpr int kk = decltype(kkc)::value;
                fr[(kk + 1) & 1].load(sa, sb, kk + 1, wm, wn);
                src.template issue<kk * C::LOADS / KSTEPS, (kk + 1) * C::LOADS / KSTEPS>(smem, fill_stage, fill);
                const W4Frags<C>& f = fr[kk & 1];
#pragma unroll
                for (int j = 0; j < TN; j++) {
                    const v8h bj = dq(f, j, zz, s2);
#pragma unroll
                    for (int i = 0; i < TM; i++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(f.a[i], bj, acc[i][j], 0, 0, 0);
                }
                interleave<MFMAS, READS, (kk + 1) * C::LOADS / KSTEPS - kk * C::LOADS / KSTEPS>();
                __builtin_amdgcn_sched_barrier(0);
            });
            {
                constexpr int KL = KSTEPS - 1;
                const W4Frags<C>& f = fr[KL & 1];
                // this slab's scales and zero-points are replaced below, so the
                // last k-step's B fragments are dequantised first
                v8h bl[TN];
#pragma unroll
                for (int j = 0; j < TN; j++) bl[j] = dq(f, j, zz, s2);
                src.template issue<KL * C::LOADS / KSTEPS, C::LOADS>(smem, fill_stage, fill);
#pragma unroll
                for (int i = 0; i < TM / 2; i++)
#pragma unroll
                    for (int j = 0; j < TN; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(f.a[i], bl[j], acc[i][j], 0, 0, 0);
                interleave<MFMAS / 2, 0, C::LOADS - KL * C::LOADS / KSTEPS>();
                __builtin_amdgcn_sched_barrier(0);
                ring_barrier<C::LOADS * (STAGES - 2)>();
                stage = stage + 1 == STAGES ? 0 : stage + 1;
                fill_stage = fill_stage + 1 == STAGES ? 0 : fill_stage + 1;
                const char* na = smem + stage * C::STAGE_BYTES;
                w4_read_sz<C>(na + C::A_BYTES + C::B_BYTES, cg, zmask, zor, zz, s2);
                fr[0].load(na, na + C::A_BYTES, 0, wm, wn);   // KSTEPS is even
#pragma unroll
                for (int i = TM / 2; i < TM; i++)
#pragma unroll
                    for (int j = 0; j < TN; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(f.a[i], bl[j], acc[i][j], 0, 0, 0);
                interleave<MFMAS / 2, READS, 0>();
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        wait_vmcnt<0>();
    }
    __syncthreads();

    const int fcol = lane & 31, frow4 = 4 * (lane >> 5);
    if (splits > 1 && out.fpart) {
        if (!w4_owner_collect<C>(smem, acc, out, split, splits)) return;
        splits = 1;   // the owner now holds the tile's sum: stored below
    }
    // Look-ahead tiles store C through the idle ring: the fp16 tile is
    // assembled in LDS (rows padded by 8 halves against bank conflicts), then
    // written row by row in 16-byte pieces.  The others store element-wise.
    constexpr int LDT = C::BN + 8;
    constexpr bool STAGED = LA && C::BM * LDT * 2 <= C::SMEM;
    if (STAGED && splits <= 1 && reinterpret_cast<uintptr_t>(out.C) % 16 == 0 && out.ldc % 8 == 0) {
        __half* ct = reinterpret_cast<__half*>(smem);
        __syncthreads();   // every wave is past its last use of the ring
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int j = 0; j < TN; j++)
#pragma unroll
                for (int r = 0; r < 16; r++)
                    ct[(wm + 32 * i + 8 * (r >> 2) + frow4 + (r & 3)) * LDT + wn + 32 * j + fcol] =
                        __float2half_rn(acc[i][j][r]);
        __syncthreads();
        constexpr int CH = C::BN / 8;
        for (int e = threadIdx.x; e < C::BM * CH; e += C::THREADS) {
            const int row = e / CH, ch = e % CH;
            const int grow = m0 + row, gcol = n0 + 8 * ch;
            if (grow < M && gcol < N)
                *reinterpret_cast<uint4*>(out.C + (long)grow * out.ldc + gcol) =
                    *reinterpret_cast<const uint4*>(ct + row * LDT + 8 * ch);
        }
        return;
    }
    if (splits <= 1) {
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int j = 0; j < TN; j++)
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const int row = m0 + wm + 32 * i + 8 * (r >> 2) + frow4 + (r & 3);
                    const int col = n0 + wn + 32 * j + fcol;
                    if (row < M && col < N) out.C[(long)row * out.ldc + col] = __float2half_rn(acc[i][j][r]);
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
    int* last_split = reinterpret_cast<int*>(smem);
    drain();
    __syncthreads();
    if (threadIdx.x == 0) *last_split = (take_ticket(out.counter) == (unsigned)splits - 1);
    __syncthreads();
    if (!*last_split) return;
    collect_tile<C::BM, C::BN, C::THREADS>(out, M, N, splits, m0, n0);
    if (threadIdx.x == 0) reset_counter(out.counter);
}

}  // namespace lt
