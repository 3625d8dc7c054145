// Grouped (batched) hgemm: many independent FP16 GEMMs of uneven shapes in one
// persistent launch, optionally captured in a hipGraph.
//
// The plan uploads a problem table to the workspace: per problem its operands,
// shape, tile grid, split-K factor and first work item.  Problems are ordered by
// the cost of one work item (K / split_k), longest first, and the resident
// blocks pull work items from an agent-scope atomic queue, so the long items
// start first and the short ones fill the tail (a greedy longest-processing-
// time schedule).  The last block to find the queue empty resets both
// counters, so the same launch (or graph) can be replayed as is.  Split-K
// problems publish FP32 partials and the last split of a tile sums them in
// split order (gemm_core.hpp), as in the single hgemm.
//
// No counterpart in the reference, whose hgemm runs one GEMM per launch
// (hgemm_better_reuse.cu:321-340).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "gemm_core2.hpp"
#include "hgemm_table.hpp"

namespace lt {

int hgemm_check_config(const lthip_hgemm_config* c, int* idx);

namespace {

struct GProb {
    const __half* A;
    const __half* B;
    __half* C;
    float* partial;      // split-K partials of this problem (null if splits == 1)
    unsigned* counters;  // split-K tile counters (null if splits == 1)
    int M, N, K, lda, ldb, ldc;
    int tiles_m, tiles_n, splits, work_begin;
};

template <int BM, int BN, int BK>
__global__ __launch_bounds__(256) void grouped_kernel(const GProb* __restrict__ probs,
                                                      const int* __restrict__ item_prob, int count, int total,
                                                      unsigned* __restrict__ queue, float alpha, float beta) {
    __shared__ __align__(16) __half smem[2 * (BM + BN) * BK];
    __shared__ int item_s;
    for (;;) {
        if (threadIdx.x == 0) item_s = (int)take_ticket(&queue[0]);
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(item_s);
        __syncthreads();
        if (item >= total) break;
        (void)count;
        const GProb& p = probs[item_prob[item]];
        const int local = item - p.work_begin;
        const int tile = local / p.splits, split = local % p.splits;
        // 8-row groups of tiles (L2 reuse of B panels)
        constexpr int G = 8;
        const int per_group = G * p.tiles_n;
        const int first_m = (tile / per_group) * G;
        int gsize = p.tiles_m - first_m;
        if (gsize > G) gsize = G;
        const int tm = first_m + (tile % per_group) % gsize, tn = (tile % per_group) / gsize;
        const int slabs = (p.K + BK - 1) / BK;
        const int s0 = (int)((long)slabs * split / p.splits), s1 = (int)((long)slabs * (split + 1) / p.splits);
        TileOut o;
        o.C = p.C;
        o.ldc = p.ldc;
        o.alpha = alpha;
        o.beta = beta;
        o.partial = p.partial;
        o.counter = p.counters ? p.counters + tile : nullptr;
        BRowParams bp{p.B, (long)p.ldb};
        gemm_tile<BM, BN, BK>(smem, p.A, (long)p.lda, bp, p.M, p.N, p.K, tm * BM, tn * BN, s0, s1, split, p.splits, o);
        __syncthreads();
    }
    // the last block out resets the queue for the next launch / graph replay
    if (threadIdx.x == 0) {
        const unsigned done = take_ticket(&queue[1]);
        if (done == gridDim.x - 1) {
            reset_counter(&queue[0]);
            reset_counter(&queue[1]);
        }
    }
}

// Same queue protocol with the DMA-ring tile (config.stages >= 3), two
// refinements: item_prob maps a work item straight to its problem (one load
// instead of a dependent binary search), and a block takes its next ticket
// when its products are in registers, so the atomic's latency hides behind
// the stores.  The item index travels through the first word of the ring,
// which is idle between tiles.
template <class C>
__global__ __launch_bounds__(C::THREADS) void grouped2_kernel(const GProb* __restrict__ probs,
                                                               const int* __restrict__ item_prob, int count, int total,
                                                               unsigned* __restrict__ queue, float alpha, float beta) {
    __shared__ __align__(16) char smem[C::SMEM];
    int* item_s = reinterpret_cast<int*>(smem);
    (void)count;
    if (threadIdx.x == 0) *item_s = (int)take_ticket(&queue[0]);
    __syncthreads();
    // uniform across the block: the table lookups below become scalar loads
    int item = __builtin_amdgcn_readfirstlane(*item_s);
    __syncthreads();
    while (item < total) {
        const GProb& p = probs[item_prob[item]];
        const int local = item - p.work_begin;
        const int tile = local / p.splits, split = local % p.splits;
        constexpr int G = 8;
        const int per_group = G * p.tiles_n;
        const int first_m = (tile / per_group) * G;
        int gsize = p.tiles_m - first_m;
        if (gsize > G) gsize = G;
        const int tm = first_m + (tile % per_group) % gsize, tn = (tile % per_group) / gsize;
        const int slabs = (p.K + C::BK - 1) / C::BK;
        const int s0 = (int)((long)slabs * split / p.splits), s1 = (int)((long)slabs * (split + 1) / p.splits);
        TileOut o;
        o.C = p.C;
        o.ldc = p.ldc;
        o.alpha = alpha;
        o.beta = beta;
        o.partial = p.partial;
        o.counter = p.counters ? p.counters + tile : nullptr;
        unsigned next = 0;
        const auto take_next = [&]() {
            if (threadIdx.x == 0) next = take_ticket(&queue[0]);
        };
        gemm_tile2<C, true>(smem, p.A, (long)p.lda, p.B, (long)p.ldb, p.M, p.N, p.K, tm * C::BM, tn * C::BN, s0, s1,
                            split, p.splits, o, take_next);
        // LDS-only barriers: the tile's stores drain during the next prologue
        lds_barrier();   // everyone is done with the ring's first word
        if (threadIdx.x == 0) *item_s = (int)next;
        lds_barrier();
        item = __builtin_amdgcn_readfirstlane(*item_s);
        lds_barrier();   // read before the next prologue's DMA overwrites it
    }
    if (threadIdx.x == 0) {
        const unsigned done = take_ticket(&queue[1]);
        if (done == gridDim.x - 1) {
            reset_counter(&queue[0]);
            reset_counter(&queue[1]);
        }
    }
}

typedef void (*grouped_fn)(const GProb*, const int*, int, int, unsigned*, float, float);
struct GTile2 {
    grouped_fn fn;
    int bm, bn, bk, stages, threads, smem;
};
#define LT_HG2(bm, bn, bk, st, wm, wn)                                                   \
    {grouped2_kernel<Core2<bm, bn, bk, st, wm, wn>>, bm, bn, bk, st,                     \
     Core2<bm, bn, bk, st, wm, wn>::THREADS, Core2<bm, bn, bk, st, wm, wn>::SMEM},
const GTile2 kGTable2[] = {LT_HGEMM2_TILES(LT_HG2)};
#undef LT_HG2
constexpr int kNumGTiles2 = sizeof(kGTable2) / sizeof(kGTable2[0]);
template <int BM, int BN, int BK>
constexpr grouped_fn gfn_of() {
    return grouped_kernel<BM, BN, BK>;
}
#define LT_HG(bm, bn, bk) gfn_of<bm, bn, bk>(),
const grouped_fn kGTable[] = {LT_HGEMM_TILES(LT_HG)};
#undef LT_HG

// blocks that fit on a CU at once (LDS bytes per block)
int resident_per_cu(int lds) {
    int by_lds = 163840 / (lds + 64);
    if (by_lds < 1) by_lds = 1;
    return by_lds < 8 ? by_lds : 8;
}

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

}  // namespace
}  // namespace lt

struct lthip_grouped_plan {
    lt::grouped_fn fn;
    int threads;
    int tile_index;
    int count;
    int total;
    int grid;
    float alpha, beta;
    lt::GProb* table;       // device (inside the workspace)
    int* item_prob;         // device: work item -> table row
    unsigned* queue;        // device, 2 counters
    hipGraph_t graph;
    hipGraphExec_t exec;
};

extern "C" {

static int grouped_layout(const lthip_hgemm_problem* probs, int count, const lthip_hgemm_config* cfg,
                          std::vector<lt::GProb>* table, std::vector<int>* order, size_t* ws_total) {
    using namespace lt;
    size_t off_queue = 0;
    size_t off_table = 256;
    size_t off_counters = off_table + align256(sizeof(GProb) * (size_t)count);
    size_t off_partial = off_counters + kCounterBytes;
    size_t counters_used = 0, partial_bytes = 0;
    table->assign(count, GProb{});
    order->resize(count);
    for (int i = 0; i < count; i++) (*order)[i] = i;
    // longest work item first
    std::stable_sort(order->begin(), order->end(), [&](int a, int b) {
        const int sa = probs[a].split_k > 0 ? probs[a].split_k : 1, sb = probs[b].split_k > 0 ? probs[b].split_k : 1;
        return (long)probs[a].K * sb > (long)probs[b].K * sa;
    });
    int work = 0;
    for (int r = 0; r < count; r++) {
        const lthip_hgemm_problem& q = probs[(*order)[r]];
        GProb& g = (*table)[r];
        g.M = q.M; g.N = q.N; g.K = q.K;
        g.lda = q.lda > 0 ? q.lda : q.K;
        g.ldb = q.ldb > 0 ? q.ldb : q.N;
        g.ldc = q.ldc > 0 ? q.ldc : q.N;
        g.tiles_m = (q.M + cfg->bm - 1) / cfg->bm;
        g.tiles_n = (q.N + cfg->bn - 1) / cfg->bn;
        const int slabs = (q.K + cfg->bk - 1) / cfg->bk;
        int splits = q.split_k > 0 ? q.split_k : 1;
        if (splits > slabs) splits = slabs > 0 ? slabs : 1;
        const size_t tiles = (size_t)g.tiles_m * g.tiles_n;
        if (splits > 1 && counters_used + tiles > kMaxCounters) splits = 1;
        g.splits = splits;
        g.A = reinterpret_cast<const __half*>(q.A);
        g.B = reinterpret_cast<const __half*>(q.B);
        g.C = reinterpret_cast<__half*>(q.C);
        g.work_begin = work;
        g.partial = nullptr;
        g.counters = nullptr;
        if (splits > 1) {
            g.counters = reinterpret_cast<unsigned*>(off_counters + counters_used * sizeof(unsigned));   // offset for now
            counters_used += tiles;
            g.partial = reinterpret_cast<float*>(off_partial + partial_bytes);
            partial_bytes += align256(sizeof(float) * (size_t)splits * q.M * q.N);
        }
        const long items = (long)tiles * splits;
        if ((long)work + items > 0x7fffffffL) return lt::fail(LTHIP_ERR_INVALID_ARG, "grouped: too many work items");
        work += (int)items;
    }
    (void)off_queue;
    // work item -> table row, after the partials
    *ws_total = off_partial + align256(partial_bytes) + align256(sizeof(int) * (size_t)work);
    return LTHIP_OK;
}

size_t lthip_hgemm_grouped_workspace_size(const lthip_hgemm_problem* probs, int count, const lthip_hgemm_config* cfg) {
    if (!probs || count <= 0 || !cfg) return 0;
    std::vector<lt::GProb> table;
    std::vector<int> order;
    size_t total = 0;
    if (grouped_layout(probs, count, cfg, &table, &order, &total) != LTHIP_OK) return 0;
    return total;
}

int lthip_hgemm_grouped_plan_create(const lthip_hgemm_problem* probs, int count, const lthip_hgemm_config* cfg,
                                    uint16_t alpha, uint16_t beta, void* ws, size_t ws_bytes, void* stream,
                                    lthip_grouped_plan** plan) {
    using namespace lt;
    if (!plan) return fail(LTHIP_ERR_INVALID_ARG, "grouped: null plan pointer");
    *plan = nullptr;
    if (!probs || count <= 0 || !cfg) return fail(LTHIP_ERR_INVALID_ARG, "grouped: no problems or config");
    int t = 0, st = LTHIP_OK;
    grouped_fn fn = nullptr;
    int threads = 256, lds = 0;
    if (cfg->stages >= 3) {
        for (int i = 0; i < kNumGTiles2 && !fn; i++) {
            const GTile2& g = kGTable2[i];
            if (g.bm == cfg->bm && g.bn == cfg->bn && g.bk == cfg->bk && g.stages == cfg->stages) {
                fn = g.fn;
                threads = g.threads;
                lds = g.smem;
                t = i;
            }
        }
        if (!fn)
            return fail(LTHIP_ERR_UNSUPPORTED, "grouped: tile %dx%dx%d with %d stages is not built", cfg->bm, cfg->bn,
                        cfg->bk, cfg->stages);
        if (cfg->split_k < 1 || cfg->split_k > 64) return fail(LTHIP_ERR_INVALID_ARG, "grouped: split_k out of range");
    } else {
        st = hgemm_check_config(cfg, &t);
        if (st != LTHIP_OK) return st;
        fn = kGTable[t];
        lds = 2 * (cfg->bm + cfg->bn) * cfg->bk * 2;
    }
    for (int i = 0; i < count; i++) {
        const lthip_hgemm_problem& q = probs[i];
        if (q.M < 0 || q.N < 0 || q.K < 0) return fail(LTHIP_ERR_INVALID_ARG, "grouped: problem %d negative size", i);
        if (q.M == 0 || q.N == 0) continue;
        if (!q.C || (q.K > 0 && (!q.A || !q.B))) return fail(LTHIP_ERR_INVALID_ARG, "grouped: problem %d null pointer", i);
        const int lda = q.lda > 0 ? q.lda : q.K, ldb = q.ldb > 0 ? q.ldb : q.N;
        if (q.K % 8 || q.N % 8 || lda % 8 || ldb % 8 || lda < q.K || ldb < q.N)
            return fail(LTHIP_ERR_INVALID_ARG, "grouped: problem %d needs K, N, lda, ldb multiples of 8", i);
        if (reinterpret_cast<uintptr_t>(q.A) % 16 || reinterpret_cast<uintptr_t>(q.B) % 16)
            return fail(LTHIP_ERR_INVALID_ARG, "grouped: problem %d operands not 16-byte aligned", i);
    }
    std::vector<GProb> table;
    std::vector<int> order;
    size_t total = 0;
    st = grouped_layout(probs, count, cfg, &table, &order, &total);
    if (st != LTHIP_OK) return st;
    if (!ws || ws_bytes < total) return fail(LTHIP_ERR_WORKSPACE, "grouped: need %zu workspace bytes", total);
    if (reinterpret_cast<uintptr_t>(ws) % 256) return fail(LTHIP_ERR_INVALID_ARG, "grouped: workspace alignment");
    char* base = reinterpret_cast<char*>(ws);
    int work = 0;
    for (auto& g : table) {
        if (g.counters) g.counters = reinterpret_cast<unsigned*>(base + reinterpret_cast<uintptr_t>(g.counters));
        if (g.partial) g.partial = reinterpret_cast<float*>(base + reinterpret_cast<uintptr_t>(g.partial));
        work = g.work_begin + g.tiles_m * g.tiles_n * g.splits;
    }
    hipStream_t s = as_stream(stream);
    GProb* dtable = reinterpret_cast<GProb*>(base + 256);
    st = check_hip(hipMemcpyAsync(dtable, table.data(), sizeof(GProb) * table.size(), hipMemcpyHostToDevice, s),
                   "grouped: table upload");
    if (st != LTHIP_OK) return st;
    std::vector<int> map((size_t)work);
    for (size_t r = 0; r < table.size(); r++) {
        const GProb& g = table[r];
        const int end = g.work_begin + g.tiles_m * g.tiles_n * g.splits;
        for (int it = g.work_begin; it < end; it++) map[(size_t)it] = (int)r;
    }
    int* dmap = reinterpret_cast<int*>(base + total - align256(sizeof(int) * (size_t)work));
    if (work > 0) {
        st = check_hip(hipMemcpyAsync(dmap, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice, s),
                       "grouped: item map upload");
        if (st != LTHIP_OK) return st;
    }
    st = check_hip(hipStreamSynchronize(s), "grouped: table upload sync");
    if (st != LTHIP_OK) return st;
    lthip_grouped_plan* p = new (std::nothrow) lthip_grouped_plan;
    if (!p) return fail(LTHIP_ERR_INVALID_ARG, "grouped: out of host memory");
    p->fn = fn;
    p->threads = threads;
    p->tile_index = t;
    p->count = count;
    p->total = work;
    const int cap = cu_count() * resident_per_cu(lds);
    p->grid = work < cap ? (work > 0 ? work : 1) : cap;
    p->alpha = __half2float(__ushort_as_half(alpha));
    p->beta = __half2float(__ushort_as_half(beta));
    p->table = dtable;
    p->item_prob = dmap;
    p->queue = reinterpret_cast<unsigned*>(base);
    p->graph = nullptr;
    p->exec = nullptr;
    *plan = p;
    return LTHIP_OK;
}

int lthip_hgemm_grouped_run(lthip_grouped_plan* plan, void* stream) {
    using namespace lt;
    if (!plan) return fail(LTHIP_ERR_INVALID_ARG, "grouped: null plan");
    hipLaunchKernelGGL(plan->fn, dim3((unsigned)plan->grid), dim3((unsigned)plan->threads), 0, as_stream(stream),
                       plan->table, plan->item_prob, plan->count, plan->total, plan->queue, plan->alpha, plan->beta);
    LT_TRY_LAUNCH("grouped hgemm launch");
    return LTHIP_OK;
}

int lthip_hgemm_grouped_capture(lthip_grouped_plan* plan, void* stream) {
    using namespace lt;
    if (!plan) return fail(LTHIP_ERR_INVALID_ARG, "grouped: null plan");
    if (!stream) return fail(LTHIP_ERR_INVALID_ARG, "grouped: graph capture needs a non-null stream");
    hipStream_t s = as_stream(stream);
    if (plan->exec) {
        (void)hipGraphExecDestroy(plan->exec);
        plan->exec = nullptr;
    }
    if (plan->graph) {
        (void)hipGraphDestroy(plan->graph);
        plan->graph = nullptr;
    }
    int st = check_hip(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "grouped: begin capture");
    if (st != LTHIP_OK) return st;
    hipLaunchKernelGGL(plan->fn, dim3((unsigned)plan->grid), dim3((unsigned)plan->threads), 0, s, plan->table,
                       plan->item_prob, plan->count, plan->total, plan->queue, plan->alpha, plan->beta);
    hipError_t le = hipGetLastError();
    hipGraph_t g = nullptr;
    st = check_hip(hipStreamEndCapture(s, &g), "grouped: end capture");
    if (le != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return check_hip(le, "grouped: launch during capture");
    }
    if (st != LTHIP_OK) return st;
    plan->graph = g;
    st = check_hip(hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0), "grouped: instantiate");
    return st;
}

int lthip_hgemm_grouped_graph_launch(lthip_grouped_plan* plan, void* stream) {
    using namespace lt;
    if (!plan || !plan->exec) return fail(LTHIP_ERR_INVALID_ARG, "grouped: no captured graph");
    return check_hip(hipGraphLaunch(plan->exec, as_stream(stream)), "grouped: graph launch");
}

int lthip_hgemm_grouped_plan_destroy(lthip_grouped_plan* plan) {
    if (!plan) return LTHIP_OK;
    if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
    if (plan->graph) (void)hipGraphDestroy(plan->graph);
    delete plan;
    return LTHIP_OK;
}

int lthip_hgemm_grouped_plan_info(const lthip_grouped_plan* plan, int* work_items, int* grid) {
    if (!plan) return lt::fail(LTHIP_ERR_INVALID_ARG, "grouped: null plan");
    if (work_items) *work_items = plan->total;
    if (grid) *grid = plan->grid;
    return LTHIP_OK;
}

}  // extern "C"
