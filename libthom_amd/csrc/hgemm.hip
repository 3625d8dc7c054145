// hgemm: C = alpha*A*B + beta*C, FP16 in/out, FP32 accumulation on
// v_mfma_f32_32x32x16_f16 (the reference accumulates in FP16,
// hgemm_better_reuse.cu:261; its CPU check accumulates in FP32, hgemm.cu:143-147).
//
// Tile configurations are compile-time instances of gemm_tile (gemm_core.hpp,
// register-staged, 2 LDS buffers: config.stages == 2) and gemm_tile2
// (gemm_core2.hpp, DMA ring of 3-5 LDS stages: config.stages >= 3);
// split-K is a launch-time choice.  The two tables below are the search space
// of the autotuner (the reference's 192-instance template list,
// hgemm_better_reuse.cu:378-454, plays the same role).
#include "gemm_core2.hpp"
#include "hgemm_table.hpp"

namespace lt {
namespace {

__device__ __forceinline__ void group_coords(int t, int tiles_m, int tiles_n, int& tm, int& tn) {
    constexpr int G = 8;
    const int per_group = G * tiles_n;
    const int first_m = (t / per_group) * G;
    int gsize = tiles_m - first_m;
    if (gsize > G) gsize = G;
    const int r = t % per_group;
    tm = first_m + r % gsize;
    tn = r / gsize;
}

template <int BM, int BN, int BK>
__global__ __launch_bounds__(256) void hgemm_kernel(const __half* __restrict__ A, long lda, BRowParams bp, int M, int N,
                                                    int K, int tiles_m, int tiles_n, int splits, TileOut out,
                                                    unsigned* counters) {
    __shared__ __align__(16) __half smem[2 * (BM + BN) * BK];
    int tm, tn;
    group_coords(xcd_remap(blockIdx.x, gridDim.x), tiles_m, tiles_n, tm, tn);
    const int split = blockIdx.y;
    const int slabs = (K + BK - 1) / BK;
    const int s0 = (int)((long)slabs * split / splits), s1 = (int)((long)slabs * (split + 1) / splits);
    TileOut o = out;
    if (counters) o.counter = counters + (long)tm * tiles_n + tn;
    gemm_tile<BM, BN, BK>(smem, A, lda, bp, M, N, K, tm * BM, tn * BN, s0, s1, split, splits, o);
}

typedef void (*hgemm_fn)(const __half*, long, BRowParams, int, int, int, int, int, int, TileOut, unsigned*);

template <int BM, int BN, int BK>
constexpr hgemm_fn fn_of() {
    return hgemm_kernel<BM, BN, BK>;
}

#define LT_HG(bm, bn, bk) fn_of<bm, bn, bk>(),
const hgemm_fn kTable[] = {LT_HGEMM_TILES(LT_HG)};
#undef LT_HG
#define LT_HG(bm, bn, bk) {bm, bn, bk},
const int kTileDims[][3] = {LT_HGEMM_TILES(LT_HG)};
#undef LT_HG
constexpr int kNumTiles = sizeof(kTileDims) / sizeof(kTileDims[0]);

template <class C>
__global__ __launch_bounds__(C::THREADS) void hgemm2_kernel(const __half* __restrict__ A, long lda,
                                                             const __half* __restrict__ B, long ldb, int M, int N,
                                                             int K, int tiles_m, int tiles_n, int splits, TileOut out,
                                                             unsigned* counters) {
    __shared__ __align__(16) char smem[C::SMEM];
    int tm, tn;
    group_coords(xcd_remap(blockIdx.x, gridDim.x), tiles_m, tiles_n, tm, tn);
    const int split = blockIdx.y;
    const int slabs = (K + C::BK - 1) / C::BK;
    const int s0 = (int)((long)slabs * split / splits), s1 = (int)((long)slabs * (split + 1) / splits);
    TileOut o = out;
    if (counters) o.counter = counters + (long)tm * tiles_n + tn;
    gemm_tile2<C>(smem, A, lda, B, ldb, M, N, K, tm * C::BM, tn * C::BN, s0, s1, split, splits, o);
}

typedef void (*hgemm2_fn)(const __half*, long, const __half*, long, int, int, int, int, int, int, TileOut, unsigned*);
struct Tile2 {
    hgemm2_fn fn;
    int bm, bn, bk, stages, threads;
};
#define LT_HG2(bm, bn, bk, st, wm, wn) \
    {hgemm2_kernel<Core2<bm, bn, bk, st, wm, wn>>, bm, bn, bk, st, Core2<bm, bn, bk, st, wm, wn>::THREADS},
const Tile2 kTable2[] = {LT_HGEMM2_TILES(LT_HG2)};
#undef LT_HG2
constexpr int kNumTiles2 = sizeof(kTable2) / sizeof(kTable2[0]);

int find_tile(int bm, int bn, int bk) {
    for (int i = 0; i < kNumTiles; i++)
        if (kTileDims[i][0] == bm && kTileDims[i][1] == bn && kTileDims[i][2] == bk) return i;
    return -1;
}

int find_tile2(int bm, int bn, int bk, int stages) {
    for (int i = 0; i < kNumTiles2; i++)
        if (kTable2[i].bm == bm && kTable2[i].bn == bn && kTable2[i].bk == bk && kTable2[i].stages == stages) return i;
    return -1;
}

// default (no tuner): large tiles for large problems, split-K when the tile
// grid cannot fill the chip and K is long
// The automatic tile: the largest DMA-ring tile that still gives about half
// of the CUs a tile of their own, whole tiles (no split-K).  Measured with
// settled clocks over square sizes 512-6144: 256 x 256 x 32 from 3072 up,
// 256 x 128 x 64 at 2560, 128 x 128 x 64 at 2048 and below, each 1.3-3.6x
// the split-K launches of the larger tiles.  M or N below 512: the 64 x 64
// register-staged tile, split over K until the CUs are about half full.
void default_config(long M, long N, long K, lthip_hgemm_config* c) {
    const long cus = cu_count();
    const long need = cus * 9 / 16;
    auto tiles = [&](long bm, long bn) { return ((M + bm - 1) / bm) * ((N + bn - 1) / bn); };
    c->split_k = 1;
    if (M >= 512 && N >= 512) {
        if (tiles(256, 256) >= need) {
            c->bm = 256; c->bn = 256; c->bk = 32; c->stages = 4;
        } else if (tiles(256, 128) >= need) {
            c->bm = M >= N ? 256 : 128; c->bn = M >= N ? 128 : 256; c->bk = 64; c->stages = 3;
        } else {
            c->bm = 128; c->bn = 128; c->bk = 64; c->stages = 3;
        }
        return;
    }
    c->bm = 64; c->bn = 64; c->bk = 64; c->stages = 2;
    const long t = tiles(64, 64);
    int split = 1;
    while (split < 16 && t * split * 2 <= cus && K / (c->bk * split * 2) >= 4) split *= 2;
    c->split_k = split;
}

}  // namespace

int hgemm_check_config(const lthip_hgemm_config* c, int* idx) {
    if (c->stages != 2 && c->stages != 0)
        return fail(LTHIP_ERR_UNSUPPORTED, "hgemm: this entry point takes only the 2-stage tiles");
    const int t = find_tile(c->bm, c->bn, c->bk);
    if (t < 0) return fail(LTHIP_ERR_UNSUPPORTED, "hgemm: tile %dx%dx%d is not built", c->bm, c->bn, c->bk);
    if (c->split_k < 1 || c->split_k > 64) return fail(LTHIP_ERR_INVALID_ARG, "hgemm: split_k out of range");
    *idx = t;
    return LTHIP_OK;
}

}  // namespace lt

extern "C" {

int lthip_hgemm_config_count(void) { return lt::kNumTiles + lt::kNumTiles2; }

int lthip_hgemm_config_get(int index, lthip_hgemm_config* cfg) {
    if (index < 0 || index >= lt::kNumTiles + lt::kNumTiles2 || !cfg)
        return lt::fail(LTHIP_ERR_INVALID_ARG, "config index");
    cfg->split_k = 1;
    if (index < lt::kNumTiles) {
        cfg->bm = lt::kTileDims[index][0];
        cfg->bn = lt::kTileDims[index][1];
        cfg->bk = lt::kTileDims[index][2];
        cfg->stages = 2;
    } else {
        const lt::Tile2& t = lt::kTable2[index - lt::kNumTiles];
        cfg->bm = t.bm;
        cfg->bn = t.bn;
        cfg->bk = t.bk;
        cfg->stages = t.stages;
    }
    return LTHIP_OK;
}

int lthip_hgemm_default_config(size_t M, size_t N, size_t K, lthip_hgemm_config* cfg) {
    if (!cfg) return lt::fail(LTHIP_ERR_INVALID_ARG, "null config");
    lt::default_config((long)M, (long)N, (long)K, cfg);
    return LTHIP_OK;
}

size_t lthip_hgemm_workspace_size(size_t M, size_t N, size_t K, const lthip_hgemm_config* cfg) {
    lthip_hgemm_config c;
    if (cfg) c = *cfg; else lt::default_config((long)M, (long)N, (long)K, &c);
    if (c.split_k <= 1 || c.bm <= 0 || c.bn <= 0) return 0;
    return lt::kCounterBytes + sizeof(float) * (size_t)c.split_k * M * N;
}

int lthip_hgemm_f16(const uint16_t* A, const uint16_t* B, uint16_t* C, size_t M, size_t N, size_t K, uint16_t alpha,
                    uint16_t beta, const lthip_hgemm_config* cfg, void* ws, size_t ws_bytes, void* stream) {
    using namespace lt;
    if (M > (1u << 30) || N > (1u << 30) || K > (1u << 30)) return fail(LTHIP_ERR_INVALID_ARG, "hgemm: size too large");
    if (M == 0 || N == 0) return LTHIP_OK;
    if (!C || (K > 0 && (!A || !B))) return fail(LTHIP_ERR_INVALID_ARG, "hgemm: null pointer");
    if (K % 8 != 0 || N % 8 != 0) return fail(LTHIP_ERR_INVALID_ARG, "hgemm: K and N must be multiples of 8");
    if (reinterpret_cast<uintptr_t>(A) % 16 || reinterpret_cast<uintptr_t>(B) % 16)
        return fail(LTHIP_ERR_INVALID_ARG, "hgemm: A and B must be 16-byte aligned");
    lthip_hgemm_config c;
    if (cfg) c = *cfg; else default_config((long)M, (long)N, (long)K, &c);
    int t = 0, t2 = -1;
    if (c.stages >= 3) {
        t2 = find_tile2(c.bm, c.bn, c.bk, c.stages);
        if (t2 < 0)
            return fail(LTHIP_ERR_UNSUPPORTED, "hgemm: tile %dx%dx%d with %d stages is not built", c.bm, c.bn, c.bk,
                        c.stages);
        if (c.split_k < 1 || c.split_k > 64) return fail(LTHIP_ERR_INVALID_ARG, "hgemm: split_k out of range");
    } else {
        int st = hgemm_check_config(&c, &t);
        if (st != LTHIP_OK) return st;
    }
    const int m = (int)M, n = (int)N, k = (int)K;
    const int tiles_m = (m + c.bm - 1) / c.bm, tiles_n = (n + c.bn - 1) / c.bn;
    const int slabs = (k + c.bk - 1) / c.bk;
    int splits = c.split_k;
    if (splits > slabs) splits = slabs > 0 ? slabs : 1;
    TileOut out;
    out.C = reinterpret_cast<__half*>(C);
    out.ldc = n;
    out.alpha = __half2float(__ushort_as_half(alpha));
    out.beta = __half2float(__ushort_as_half(beta));
    out.partial = nullptr;
    out.counter = nullptr;
    unsigned* counters = nullptr;
    if (splits > 1 && (size_t)tiles_m * tiles_n > kMaxCounters) splits = 1;   // too many tiles to need split-K
    if (splits > 1) {
        const size_t need = kCounterBytes + sizeof(float) * (size_t)splits * M * N;
        if (!ws || ws_bytes < need) return fail(LTHIP_ERR_WORKSPACE, "hgemm: split-K needs %zu workspace bytes", need);
        if (reinterpret_cast<uintptr_t>(ws) % 16) return fail(LTHIP_ERR_INVALID_ARG, "hgemm: workspace alignment");
        counters = reinterpret_cast<unsigned*>(ws);
        out.partial = reinterpret_cast<float*>(reinterpret_cast<char*>(ws) + kCounterBytes);
    }
    dim3 grid((unsigned)(tiles_m * tiles_n), (unsigned)splits);
    if (t2 >= 0) {
        hipLaunchKernelGGL(kTable2[t2].fn, grid, dim3(kTable2[t2].threads), 0, as_stream(stream),
                           reinterpret_cast<const __half*>(A), (long)k, reinterpret_cast<const __half*>(B), (long)n, m,
                           n, k, tiles_m, tiles_n, splits, out, counters);
    } else {
        BRowParams bp{reinterpret_cast<const __half*>(B), (long)n};
        hipLaunchKernelGGL(kTable[t], grid, dim3(256), 0, as_stream(stream), reinterpret_cast<const __half*>(A),
                           (long)k, bp, m, n, k, tiles_m, tiles_n, splits, out, counters);
    }
    LT_TRY_LAUNCH("hgemm launch");
    return LTHIP_OK;
}

}  // extern "C"
