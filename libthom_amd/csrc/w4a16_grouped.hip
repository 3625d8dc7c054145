// Grouped (batched) w4a16: many independent int4-weight GEMMs of uneven
// shapes in one persistent launch, optionally captured in a hipGraph.
//
// Every problem brings its own activations, packed weights, scales and
// zero-points (a mixture of experts, or the layers of a model).  Each problem
// takes one of three pipelined tiles (w4_core2.hpp) that share one block
// shape, eight waves: 256 x 256, 128 x 256 or 64 x 256 (rows x columns),
// chosen by its M unless the config forces one, and a single kernel holds all
// three, so one launch mixes them.  Problems are ordered by the cost of one
// work item (K x tile rows), longest first, and the resident blocks pull work
// items from an agent-scope atomic queue (a greedy longest-processing-time
// schedule): the small tiles fill the CUs the large ones leave idle in their
// last round.  The last block out resets the queue, so the launch and its
// graph replay as is.  Whole tiles only (no split-K).  Within a problem the
// row tiles of one column panel are consecutive work items, so the blocks
// running at the same time share the panel's weights in L2.
//
// No counterpart in the reference, whose w4a16 implementations take one
// problem per call (w4a16.cpp:247).
#include <algorithm>
#include <new>
#include <vector>

#include "w4_core2.hpp"

namespace lt {
namespace {

// the big tile as in the single launch: eight waves of 256 x 32
using GW4Big = Core2W4<256, 256, 3, 1, 8>;
using GW4Mid = Core2W4<128, 256, 3, 2, 4>;
using GW4Small = Core2W4<64, 256, 3, 1, 8>;
static_assert(GW4Big::THREADS == GW4Mid::THREADS && GW4Mid::THREADS == GW4Small::THREADS, "one block shape");
constexpr int kGW4Threads = GW4Big::THREADS;
constexpr int kGW4Smem = GW4Big::SMEM > GW4Mid::SMEM ? (GW4Big::SMEM > GW4Small::SMEM ? GW4Big::SMEM : GW4Small::SMEM)
                                                     : (GW4Mid::SMEM > GW4Small::SMEM ? GW4Mid::SMEM : GW4Small::SMEM);
constexpr int kGW4TileN = 256;
constexpr int kGW4TileM[3] = {256, 128, 64};

struct GW4Prob {
    const __half* A;
    BW4Params p;
    __half* C;
    int M, N, K, cls, tiles_m, work_begin;
};

// LA: the look-ahead tile loop with buffer-load DMA (every problem's A and W
// below 2^31 bytes)
template <bool LA>
__global__ __launch_bounds__(kGW4Threads) void w4a16_grouped_kernel(const GW4Prob* __restrict__ probs, int count,
                                                                    int total, unsigned* __restrict__ queue) {
    __shared__ __align__(16) char smem[kGW4Smem];
    int* item_s = reinterpret_cast<int*>(smem);   // the ring's first word is idle between tiles
    for (;;) {
        if (threadIdx.x == 0) *item_s = (int)take_ticket(&queue[0]);
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(*item_s);
        __syncthreads();
        if (item >= total) break;
        int lo = 0, hi = count - 1;   // the problem holding this item: last p with work_begin <= item
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (probs[mid].work_begin <= item) lo = mid; else hi = mid - 1;
        }
        const GW4Prob& q = probs[lo];
        const int t = item - q.work_begin;
        const int tm = t % q.tiles_m, tn = t / q.tiles_m;
        TileOut o;
        o.C = q.C;
        o.ldc = q.N;
        o.alpha = 1.0f;
        o.beta = 0.0f;
        o.partial = nullptr;
        o.counter = nullptr;
        const int slabs = q.K / 64;
        if (q.cls == 0)
            w4_tile2<GW4Big, LA>(smem, q.A, (long)q.K, q.p, q.M, q.N, tm * 256, tn * kGW4TileN, 0, slabs, 0, 1, o);
        else if (q.cls == 1)
            w4_tile2<GW4Mid, LA>(smem, q.A, (long)q.K, q.p, q.M, q.N, tm * 128, tn * kGW4TileN, 0, slabs, 0, 1, o);
        else
            w4_tile2<GW4Small, LA>(smem, q.A, (long)q.K, q.p, q.M, q.N, tm * 64, tn * kGW4TileN, 0, slabs, 0, 1, o);
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

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// tile class of one problem: the config's tile, or by M
int gw4_class(const lthip_w4a16_problem& q, const lthip_w4a16_config* cfg, int* cls) {
    if (cfg && cfg->kind != 0) {
        if (cfg->kind != 3 || cfg->split_k > 1 || cfg->tile_n != kGW4TileN)
            return fail(LTHIP_ERR_UNSUPPORTED, "grouped w4a16: config kind 0 (auto) or 3 with tile_n 256, split_k 1");
        for (int c = 0; c < 3; c++)
            if (cfg->tile_m == kGW4TileM[c]) {
                *cls = c;
                return LTHIP_OK;
            }
        return fail(LTHIP_ERR_UNSUPPORTED, "grouped w4a16: tile_m 256, 128 or 64");
    }
    *cls = q.M > 128 ? 0 : (q.M > 64 ? 1 : 2);
    return LTHIP_OK;
}

int gw4_check(const lthip_w4a16_problem& q, int i) {
    if (q.M < 0 || q.N < 0 || q.K < 0) return fail(LTHIP_ERR_INVALID_ARG, "grouped w4a16: problem %d negative size", i);
    if (q.M == 0 || q.N == 0) return LTHIP_OK;
    if (q.M >= (1 << 30) || q.N >= (1 << 30) || q.K >= (1 << 30))
        return fail(LTHIP_ERR_INVALID_ARG, "grouped w4a16: problem %d size too large", i);
    if (q.K == 0 || q.group <= 0 || q.group % 64 != 0 || q.K % q.group != 0 || q.N % 32 != 0)
        return fail(LTHIP_ERR_INVALID_ARG,
                    "grouped w4a16: problem %d needs K > 0, group %% 64 == 0, K %% group == 0, N %% 32 == 0", i);
    if ((long)(q.K / 8) * q.N >= (1L << 31)) return fail(LTHIP_ERR_INVALID_ARG, "grouped w4a16: problem %d too large", i);
    if (!q.A || !q.W_packed || !q.scales || !q.C)
        return fail(LTHIP_ERR_INVALID_ARG, "grouped w4a16: problem %d null pointer", i);
    if (reinterpret_cast<uintptr_t>(q.A) % 16 || reinterpret_cast<uintptr_t>(q.W_packed) % 16 ||
        reinterpret_cast<uintptr_t>(q.scales) % 16 || reinterpret_cast<uintptr_t>(q.qzeros) % 16 ||
        reinterpret_cast<uintptr_t>(q.C) % 8)
        return fail(LTHIP_ERR_INVALID_ARG,
                    "grouped w4a16: problem %d needs A, W, scales, zero-points 16-byte and C 8-byte aligned", i);
    return LTHIP_OK;
}

// problem table in launch order (longest work item first) and the workspace size
int gw4_layout(const lthip_w4a16_problem* probs, int count, const lthip_w4a16_config* cfg,
               std::vector<GW4Prob>* table, size_t* ws_total, int* total, bool* la = nullptr) {
    std::vector<int> cls(count, 0), order;
    for (int i = 0; i < count; i++) {
        int st = gw4_check(probs[i], i);
        if (st != LTHIP_OK) return st;
        st = gw4_class(probs[i], cfg, &cls[i]);
        if (st != LTHIP_OK) return st;
        if (probs[i].M > 0 && probs[i].N > 0) order.push_back(i);
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return (long)probs[a].K * kGW4TileM[cls[a]] > (long)probs[b].K * kGW4TileM[cls[b]];
    });
    table->clear();
    long work = 0;
    if (la) *la = true;
    for (int i : order) {
        const lthip_w4a16_problem& q = probs[i];
        GW4Prob g;
        g.A = reinterpret_cast<const __half*>(q.A);
        g.p = BW4Params{q.W_packed, reinterpret_cast<const __half*>(q.scales), q.qzeros, q.group};
        g.C = reinterpret_cast<__half*>(q.C);
        g.M = q.M;
        g.N = q.N;
        g.K = q.K;
        g.cls = cls[i];
        g.tiles_m = (q.M + kGW4TileM[g.cls] - 1) / kGW4TileM[g.cls];
        g.work_begin = (int)work;
        if (la && ((size_t)q.M * q.K * 2 >= (1u << 31) || (size_t)(q.K / 8) * q.N * 4 >= (1u << 31))) *la = false;
        work += (long)g.tiles_m * ((q.N + kGW4TileN - 1) / kGW4TileN);
        if (work > 0x7fffffffL) return fail(LTHIP_ERR_INVALID_ARG, "grouped w4a16: too many work items");
        table->push_back(g);
    }
    *total = (int)work;
    *ws_total = 256 + align256(sizeof(GW4Prob) * (table->empty() ? 1 : table->size()));
    return LTHIP_OK;
}

}  // namespace
}  // namespace lt

struct lthip_w4a16_grouped_plan {
    int count;   // problems with work
    int total;   // work items
    int grid;
    bool la;              // every problem fits the look-ahead loop's 32-bit offsets
    lt::GW4Prob* table;   // device (inside the workspace)
    unsigned* queue;      // device, 2 counters
    hipGraph_t graph;
    hipGraphExec_t exec;
};

extern "C" {

size_t lthip_w4a16_grouped_workspace_size(const lthip_w4a16_problem* probs, int count,
                                          const lthip_w4a16_config* cfg) {
    if (!probs || count <= 0) return 0;
    std::vector<lt::GW4Prob> table;
    size_t total = 0;
    int work = 0;
    if (lt::gw4_layout(probs, count, cfg, &table, &total, &work) != LTHIP_OK) return 0;
    return total;
}

int lthip_w4a16_grouped_plan_create(const lthip_w4a16_problem* probs, int count, const lthip_w4a16_config* cfg,
                                    void* ws, size_t ws_bytes, void* stream, lthip_w4a16_grouped_plan** plan) {
    using namespace lt;
    if (!plan) return fail(LTHIP_ERR_INVALID_ARG, "grouped w4a16: null plan pointer");
    *plan = nullptr;
    if (!probs || count <= 0) return fail(LTHIP_ERR_INVALID_ARG, "grouped w4a16: no problems");
    std::vector<GW4Prob> table;
    size_t need = 0;
    int work = 0;
    bool la = true;
    int st = gw4_layout(probs, count, cfg, &table, &need, &work, &la);
    if (st != LTHIP_OK) return st;
    if (!ws || ws_bytes < need) return fail(LTHIP_ERR_WORKSPACE, "grouped w4a16: need %zu workspace bytes", need);
    if (reinterpret_cast<uintptr_t>(ws) % 256) return fail(LTHIP_ERR_INVALID_ARG, "grouped w4a16: workspace alignment");
    char* base = reinterpret_cast<char*>(ws);
    hipStream_t s = as_stream(stream);
    GW4Prob* dtable = reinterpret_cast<GW4Prob*>(base + 256);
    if (!table.empty()) {
        st = check_hip(hipMemcpyAsync(dtable, table.data(), sizeof(GW4Prob) * table.size(), hipMemcpyHostToDevice, s),
                       "grouped w4a16: table upload");
        if (st != LTHIP_OK) return st;
        st = check_hip(hipStreamSynchronize(s), "grouped w4a16: table upload sync");
        if (st != LTHIP_OK) return st;
    }
    lthip_w4a16_grouped_plan* p = new (std::nothrow) lthip_w4a16_grouped_plan;
    if (!p) return fail(LTHIP_ERR_INVALID_ARG, "grouped w4a16: out of host memory");
    p->count = (int)table.size();
    p->total = work;
    const int per_cu = 163840 / (kGW4Smem + 64);
    const int cap = cu_count() * (per_cu >= 1 ? per_cu : 1);
    p->grid = work < cap ? (work > 0 ? work : 1) : cap;
    p->la = la;
    p->table = dtable;
    p->queue = reinterpret_cast<unsigned*>(base);
    p->graph = nullptr;
    p->exec = nullptr;
    *plan = p;
    return LTHIP_OK;
}

int lthip_w4a16_grouped_run(lthip_w4a16_grouped_plan* plan, void* stream) {
    using namespace lt;
    if (!plan) return fail(LTHIP_ERR_INVALID_ARG, "grouped w4a16: null plan");
    if (plan->total == 0) return LTHIP_OK;
    hipLaunchKernelGGL(plan->la ? w4a16_grouped_kernel<true> : w4a16_grouped_kernel<false>,
                       dim3((unsigned)plan->grid), dim3(kGW4Threads), 0, as_stream(stream), plan->table, plan->count,
                       plan->total, plan->queue);
    LT_TRY_LAUNCH("grouped w4a16 launch");
    return LTHIP_OK;
}

int lthip_w4a16_grouped_capture(lthip_w4a16_grouped_plan* plan, void* stream) {
    using namespace lt;
    if (!plan) return fail(LTHIP_ERR_INVALID_ARG, "grouped w4a16: null plan");
    if (!stream) return fail(LTHIP_ERR_INVALID_ARG, "grouped w4a16: graph capture needs a non-null stream");
    if (plan->total == 0) return fail(LTHIP_ERR_INVALID_ARG, "grouped w4a16: nothing to capture (no work)");
    hipStream_t s = as_stream(stream);
    if (plan->exec) {
        (void)hipGraphExecDestroy(plan->exec);
        plan->exec = nullptr;
    }
    if (plan->graph) {
        (void)hipGraphDestroy(plan->graph);
        plan->graph = nullptr;
    }
    int st = check_hip(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "grouped w4a16: begin capture");
    if (st != LTHIP_OK) return st;
    hipLaunchKernelGGL(plan->la ? w4a16_grouped_kernel<true> : w4a16_grouped_kernel<false>,
                       dim3((unsigned)plan->grid), dim3(kGW4Threads), 0, s, plan->table, plan->count, plan->total,
                       plan->queue);
    hipError_t le = hipGetLastError();
    hipGraph_t g = nullptr;
    st = check_hip(hipStreamEndCapture(s, &g), "grouped w4a16: end capture");
    if (le != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return check_hip(le, "grouped w4a16: launch during capture");
    }
    if (st != LTHIP_OK) return st;
    plan->graph = g;
    return check_hip(hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0), "grouped w4a16: instantiate");
}

int lthip_w4a16_grouped_graph_launch(lthip_w4a16_grouped_plan* plan, void* stream) {
    using namespace lt;
    if (!plan || !plan->exec) return fail(LTHIP_ERR_INVALID_ARG, "grouped w4a16: no captured graph");
    return check_hip(hipGraphLaunch(plan->exec, as_stream(stream)), "grouped w4a16: graph launch");
}

int lthip_w4a16_grouped_plan_info(const lthip_w4a16_grouped_plan* plan, int* work_items, int* grid) {
    if (!plan) return lt::fail(LTHIP_ERR_INVALID_ARG, "grouped w4a16: null plan");
    if (work_items) *work_items = plan->total;
    if (grid) *grid = plan->grid;
    return LTHIP_OK;
}

int lthip_w4a16_grouped_plan_destroy(lthip_w4a16_grouped_plan* plan) {
    if (!plan) return LTHIP_OK;
    if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
    if (plan->graph) (void)hipGraphDestroy(plan->graph);
    delete plan;
    return LTHIP_OK;
}

}  // extern "C"
