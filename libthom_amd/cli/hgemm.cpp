// hgemm routine driver (the reference's routines/hgemm/hgemm.cu): FP16 GEMM with
// FP32 accumulation, test/bench, and the exhaustive `tune` mode
// (hgemm.cu:258-323).  Also the hgemm_batched routine: 256 uneven problems in
// one persistent launch ("persistent") or its captured hipGraph ("graph").
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <fstream>
#include <random>

#include "host_math.h"
#include "routines.h"
#include "tuning.h"

namespace hgemm {

#define BENCH_RUNS 30
#define WARMUP_RUNS 5
// hgemm.cu:17-27
static const std::vector<std::vector<int>> SIZES = {{512, 512, 512},    {1024, 1024, 1024}, {1536, 1536, 1536},
                                                    {2048, 2048, 2048}, {2560, 2560, 2560}, {3072, 3072, 3072},
                                                    {3584, 3584, 3584}, {4096, 4096, 4096}, {8192, 8192, 8192}};

using HgemmImpl = float (*)(const uint16_t* A, const uint16_t* B, uint16_t* C, size_t M, size_t N, size_t K,
                            uint16_t alpha, uint16_t beta);   // hgemm.h:6-7 (fp16 as bit patterns)

static float run_cfg(const lthip_hgemm_config* cfg, const uint16_t* A, const uint16_t* B, uint16_t* C, size_t M,
                     size_t N, size_t K, uint16_t alpha, uint16_t beta) {
    const size_t need = lthip_hgemm_workspace_size(M, N, K, cfg);
    void* ws = need ? cli::workspace("hgemm", need) : nullptr;
    return cli::time_ms(
        [&] { LTHIP_CHECK(lthip_hgemm_f16(A, B, C, M, N, K, alpha, beta, cfg, ws, need, cli::stream())); });
}

static float hgemm_hip(const uint16_t* A, const uint16_t* B, uint16_t* C, size_t M, size_t N, size_t K, uint16_t alpha,
                       uint16_t beta) {
    return run_cfg(nullptr, A, B, C, M, N, K, alpha, beta);
}

// the configuration the tuning database holds for this exact shape, if any
static float hgemm_tuned(const uint16_t* A, const uint16_t* B, uint16_t* C, size_t M, size_t N, size_t K,
                         uint16_t alpha, uint16_t beta) {
    lthip_hgemm_config c;
    if (!tuning::lookup_hgemm((int)M, (int)N, (int)K, &c)) return run_cfg(nullptr, A, B, C, M, N, K, alpha, beta);
    return run_cfg(&c, A, B, C, M, N, K, alpha, beta);
}

// Host implementations: the reference's CPU product (hgemm.cu:137-152, an FP32
// sum of fp16 products, C = fp16(alpha sum + beta C)) on one thread ("cpu") or
// all host cores ("cpu_mt"), timed with std::chrono; copies to and from the
// device are outside the timed region.
static float hgemm_host(const uint16_t* A, const uint16_t* B, uint16_t* C, size_t M, size_t N, size_t K,
                        uint16_t alpha, uint16_t beta, int threads) {
    const std::vector<uint16_t> hA = cli::to_host(A, M * K), hB = cli::to_host(B, K * N);
    std::vector<uint16_t> hC = cli::to_host(C, M * N);
    const float al = cli::f16_to_f32(alpha), be = cli::f16_to_f32(beta);
    const float ms = cli::cpu_time_ms(threads, [&] {
        std::vector<float> fb(K * N);
        cli::parallel_for(K * N, [&](size_t lo, size_t hi) {
            for (size_t i = lo; i < hi; i++) fb[i] = cli::f16_to_f32(hB[i]);
        });
        cli::parallel_for(M, [&](size_t lo, size_t hi) {
            std::vector<float> acc(N);
            for (size_t i = lo; i < hi; i++) {
                std::fill(acc.begin(), acc.end(), 0.0f);
                for (size_t k = 0; k < K; k++) {
                    const float a = cli::f16_to_f32(hA[i * K + k]);
                    const float* b = &fb[k * N];
                    for (size_t j = 0; j < N; j++) acc[j] += a * b[j];
                }
                for (size_t j = 0; j < N; j++)
                    hC[i * N + j] = cli::f32_to_f16(al * acc[j] + be * cli::f16_to_f32(hC[i * N + j]));
            }
        });
    });
    cli::to_device(C, hC);
    return ms;
}
static float hgemm_cpu(const uint16_t* A, const uint16_t* B, uint16_t* C, size_t M, size_t N, size_t K, uint16_t alpha,
                       uint16_t beta) {
    return hgemm_host(A, B, C, M, N, K, alpha, beta, 1);
}
static float hgemm_cpu_mt(const uint16_t* A, const uint16_t* B, uint16_t* C, size_t M, size_t N, size_t K,
                          uint16_t alpha, uint16_t beta) {
    return hgemm_host(A, B, C, M, N, K, alpha, beta, cli::host_cores());
}

static const std::unordered_map<std::string, HgemmImpl> hgemmImpls = {
#if LIB_HAS(HGEMM_HIP)
    {"hip", hgemm_hip},
    {"tuned", hgemm_tuned},
#endif
    {"cpu", hgemm_cpu},
    {"cpu_mt", hgemm_cpu_mt},
};

static cli::DevBuf<uint16_t> gpuA, gpuB, gpuC;
static std::vector<uint16_t> cpuA, cpuB;
static const uint16_t alphaH = 0x3C00, betaH = 0x0000;   // alpha = 1, beta = 0 (hgemm.cu:121-122)

static void resetVectors(int M, int N, int K, bool needHost) {
    LTHIP_CHECK(lthip_fill_uniform_f16(gpuA.p, (size_t)M * K, -1.0, 1.0, cli::kSeed, 31, cli::stream()));
    LTHIP_CHECK(lthip_fill_uniform_f16(gpuB.p, (size_t)K * N, -1.0, 1.0, cli::kSeed, 32, cli::stream()));
    HIP_CHECK(hipStreamSynchronize(cli::stream()));
    if (needHost) {
        cpuA.resize((size_t)M * K);
        cpuB.resize((size_t)K * N);
        gpuA.download(cpuA.data());
        gpuB.download(cpuB.data());
    }
}

// hgemm.cu:137-182: FP32 sum of fp16 products, C = fp16(sum); an element is
// wrong iff rel > 1e-2 (denominator max(|ref|, 1e-3, 1e-2 (|A||B|)_ij for
// signed inputs)) and abs > 1e-3.  Checks up to 256 rows (sampled for big M).
static bool compareToReference(int M, int N, int K) {
    std::vector<uint16_t> got((size_t)M * N);
    gpuC.download(got.data());
    std::vector<int> rows;
    const int step = M > 256 ? M / 256 : 1;
    for (int r = 0; r < M; r += step) rows.push_back(r);
    std::vector<float> fb((size_t)K * N);
    for (size_t i = 0; i < fb.size(); i++) fb[i] = cli::f16_to_f32(cpuB[i]);
    std::vector<int> bad(rows.size(), 0);
    cli::parallel_for(rows.size(), [&](size_t lo, size_t hi) {
        std::vector<float> acc(N), mag(N);
        for (size_t t = lo; t < hi; t++) {
            const int i = rows[t];
            std::fill(acc.begin(), acc.end(), 0.0f);
            std::fill(mag.begin(), mag.end(), 0.0f);
            for (int k = 0; k < K; k++) {
                const float a = cli::f16_to_f32(cpuA[(size_t)i * K + k]);
                const float* b = &fb[(size_t)k * N];
                for (int j = 0; j < N; j++) {
                    acc[j] += a * b[j];
                    mag[j] += std::fabs(a * b[j]);
                }
            }
            for (int j = 0; j < N; j++) {
                const float ref = cli::f16_to_f32(cli::f32_to_f16(acc[j]));
                const float g = cli::f16_to_f32(got[(size_t)i * N + j]);
                const double diff = std::fabs((double)g - ref);
                const double den = std::max({(double)std::fabs(ref), 1e-3, 1e-2 * mag[j]});
                if (diff / den > 1e-2 && diff > 1e-3) bad[t]++;
            }
        }
    });
    int fails = 0;
    for (int b : bad) fails += b;
    if (fails) {
        printf("Number of different results: %d\n", fails);
        cli::g_check_failed = true;
    }
    return fails == 0;
}

static float getGflops(int M, int N, int K, float ms) { return (float)(2.0 * M * N * K / (ms * 1e6)); }

static float runner(HgemmImpl impl, std::vector<int> sizes, int nb_runs, int nb_warmup_runs, bool checkResult,
                    bool printEach, int, char**) {
    const int M = sizes[0], N = sizes[1], K = sizes[2];
    gpuA.alloc((size_t)M * K);
    gpuB.alloc((size_t)K * N);
    gpuC.alloc((size_t)M * N);
    resetVectors(M, N, K, checkResult);
    float sum_gflops = 0.0f;
    for (int i = -nb_warmup_runs; i < nb_runs; i++) {
        const float ms = impl(gpuA.p, gpuB.p, gpuC.p, M, N, K, alphaH, betaH);
        if (checkResult) compareToReference(M, N, K);
        if (i >= 0) {
            const float g = getGflops(M, N, K, ms);
            if (printEach)
                printf("[internal][iter %d]: %f GFLOP/s (%.1f%% of FP16 MFMA peak)\n", i, g,
                       100.0 * g / 1000.0 / cli::kFp16MfmaTflops);
            sum_gflops += g;
        }
    }
    gpuA.free();
    gpuB.free();
    gpuC.free();
    return sum_gflops / nb_runs;
}

static void exportCSV(const std::map<std::vector<int>, float>& results, const std::string& filename) {
    std::ofstream file(filename);
    file << "M,N,K,gflops\n";
    for (const auto& [s, g] : results) file << s[0] << "," << s[1] << "," << s[2] << "," << g << "\n";
}

// ./libthom tune hgemm [M N K [runs]]: every built tile x split-K in
// {1, 2, 4, 8}: validate against the CPU check, time `runs` launches after
// WARMUP_RUNS, print the reference's [result]/[best] lines and append every
// timing to the tuning database (tuning.h).
static void tune(int argc, char* argv[]) {
    int M = 2048, N = 2048, K = 2048, runs = BENCH_RUNS;
    if (argc != 3 && argc != 6 && argc != 7) {
        fprintf(stderr, "Usage: ./libthom tune hgemm [M N K [runs]]\n");
        exit(1);
    }
    if (argc >= 6) {
        M = atoi(argv[3]);
        N = atoi(argv[4]);
        K = atoi(argv[5]);
    }
    if (argc == 7) runs = atoi(argv[6]);
    if (M <= 0 || N <= 0 || K <= 0 || runs <= 0 || K % 8 || N % 8) {
        fprintf(stderr, "M, N, K and runs must be positive; K and N multiples of 8.\n");
        exit(1);
    }
    printf("-- Tuning hgemm for M=%d N=%d K=%d (%d measured runs) --\n", M, N, K, runs);
    gpuA.alloc((size_t)M * K);
    gpuB.alloc((size_t)K * N);
    gpuC.alloc((size_t)M * N);
    resetVectors(M, N, K, true);
    float best = -1.0f;
    lthip_hgemm_config bestc{};
    const int ntiles = lthip_hgemm_config_count();
    for (int t = 0; t < ntiles; t++) {
        for (int split : {1, 2, 4, 8}) {
            lthip_hgemm_config c;
            LTHIP_CHECK(lthip_hgemm_config_get(t, &c));
            c.split_k = split;
            char name[128];
            snprintf(name, sizeof(name), "BM=%d BN=%d BK=%d split_k=%d stages=%d", c.bm, c.bn, c.bk, c.split_k, c.stages);
            if ((K + c.bk - 1) / c.bk < split) {
                printf("[skip] %s (fewer K-slabs than splits)\n", name);
                continue;
            }
            HIP_CHECK(hipMemset(gpuC.p, 0, sizeof(uint16_t) * gpuC.n));
            run_cfg(&c, gpuA.p, gpuB.p, gpuC.p, M, N, K, alphaH, betaH);
            if (!compareToReference(M, N, K)) {
                printf("[failed] %s\n", name);
                continue;
            }
            for (int w = 0; w < WARMUP_RUNS; w++) run_cfg(&c, gpuA.p, gpuB.p, gpuC.p, M, N, K, alphaH, betaH);
            std::vector<float> ms;
            for (int r = 0; r < runs; r++) ms.push_back(run_cfg(&c, gpuA.p, gpuB.p, gpuC.p, M, N, K, alphaH, betaH));
            std::sort(ms.begin(), ms.end());
            const float med = ms[ms.size() / 2];
            const float g = getGflops(M, N, K, med);
            printf("[result] %s: %.2f GFLOP/s\n", name, g);
            tuning::append_hgemm(M, N, K, c, med);
            if (g > best) {
                best = g;
                bestc = c;
            }
        }
    }
    if (best < 0) {
        fprintf(stderr, "No candidate supports these dimensions.\n");
        exit(1);
    }
    printf("[best] BM=%d BN=%d BK=%d split_k=%d stages=%d: %.2f GFLOP/s\n", bestc.bm, bestc.bn, bestc.bk, bestc.split_k,
           bestc.stages, best);
    gpuA.free();
    gpuB.free();
    gpuC.free();
}

}  // namespace hgemm

void hgemm_handler(RunMode runMode, int argc, char* argv[]) {
    switch (runMode) {
        case RunMode::TEST:
            Handlers<hgemm::HgemmImpl, 3, BENCH_RUNS, WARMUP_RUNS>::test_handler(argc, argv, hgemm::hgemmImpls,
                                                                                 hgemm::runner);
            break;
        case RunMode::BENCH:
            Handlers<hgemm::HgemmImpl, 3, BENCH_RUNS, WARMUP_RUNS>::bench_handler(
                argc, argv, hgemm::hgemmImpls, hgemm::runner, hgemm::SIZES, hgemm::exportCSV);
            break;
        case RunMode::TUNE:
            hgemm::tune(argc, argv);
            break;
    }
}

// ---------------------------------------------------------------------------
// hgemm_batched: `test hgemm_batched <impl> <count> <nb_iter>`; problem shapes
// log-uniform in [64, 4096] (multiples of 16) from the seeded generator.
namespace hgemm_batched {

struct Batch {
    std::vector<cli::DevBuf<uint16_t>*> bufs;
    std::vector<lthip_hgemm_problem> probs;
    double flops = 0.0;
    ~Batch() {
        for (auto* b : bufs) delete b;
    }
};

using BatchedImpl = float (*)(Batch&, const lthip_hgemm_config&, int runs, bool graph);

static std::vector<std::array<int, 3>> dims_of(int count) {
    std::vector<std::array<int, 3>> d(count);
    for (int i = 0; i < count; i++)
        for (int j = 0; j < 3; j++) {
            const double u = cli::rng_unit(cli::kSeed, 8, (uint64_t)(3 * i + j));
            const double v = std::exp(std::log(64.0) + (std::log(4096.0) - std::log(64.0)) * u);
            int x = (int)std::nearbyint(v / 16.0) * 16;
            d[i][j] = std::min(4096, std::max(64, x));
        }
    return d;
}

static float run_batch(Batch& b, const lthip_hgemm_config& cfg, int runs, bool graph) {
    const size_t need = lthip_hgemm_grouped_workspace_size(b.probs.data(), (int)b.probs.size(), &cfg);
    void* ws = cli::workspace("grouped", need);
    lthip_grouped_plan* plan = nullptr;
    LTHIP_CHECK(lthip_hgemm_grouped_plan_create(b.probs.data(), (int)b.probs.size(), &cfg, 0x3C00, 0, ws, need,
                                                cli::stream(), &plan));
    if (graph) LTHIP_CHECK(lthip_hgemm_grouped_capture(plan, cli::stream()));
    auto once = [&] {
        if (graph)
            LTHIP_CHECK(lthip_hgemm_grouped_graph_launch(plan, cli::stream()));
        else
            LTHIP_CHECK(lthip_hgemm_grouped_run(plan, cli::stream()));
    };
    once();
    const float ms = cli::time_ms([&] {
        for (int r = 0; r < runs; r++) once();
    });
    LTHIP_CHECK(lthip_hgemm_grouped_plan_destroy(plan));
    return ms / runs;
}

static const std::unordered_map<std::string, bool> batchedImpls = {{"persistent", false}, {"graph", true}};

}  // namespace hgemm_batched

void hgemm_batched_handler(RunMode runMode, int argc, char* argv[]) {
    using namespace hgemm_batched;
    if (runMode == RunMode::TUNE) {
        fprintf(stderr, "Tune the batched hgemm with libthom_amd/tuner.py (tile x split-K policy)\n");
        exit(1);
    }
    if (argc < 4 || !batchedImpls.count(argv[3])) {
        fprintf(stderr, "Usage: ./libthom <test|bench> hgemm_batched <persistent|graph> [count=256 [nb_iter=10]]\n");
        exit(1);
    }
    const bool graph = batchedImpls.at(argv[3]);
    const int count = argc > 4 ? atoi(argv[4]) : 256;
    const int iters = argc > 5 ? atoi(argv[5]) : 10;
    if (count <= 0 || iters <= 0) {
        fprintf(stderr, "count and nb_iter must be positive\n");
        exit(1);
    }
    // tile and split-K policy: the tuned choice (tuning/batched.csv), else
    // the 256x256x32 4-stage tile without split-K
    lthip_hgemm_config cfg{256, 256, 32, 1, 4};
    int policy = 0;
    tuning::lookup_batched(count, &cfg, &policy);
    Batch b;
    const auto dims = dims_of(count);
    for (int i = 0; i < count; i++) {
        const int M = dims[i][0], N = dims[i][1], K = dims[i][2];
        auto* A = new cli::DevBuf<uint16_t>((size_t)M * K);
        auto* B = new cli::DevBuf<uint16_t>((size_t)K * N);
        auto* C = new cli::DevBuf<uint16_t>((size_t)M * N);
        b.bufs.insert(b.bufs.end(), {A, B, C});
        LTHIP_CHECK(lthip_fill_uniform_f16(A->p, A->n, -1.0, 1.0, cli::kSeed, 1000 + 2 * i, cli::stream()));
        LTHIP_CHECK(lthip_fill_uniform_f16(B->p, B->n, -1.0, 1.0, cli::kSeed, 1001 + 2 * i, cli::stream()));
        lthip_hgemm_problem p{A->p, B->p, C->p, M, N, K, 0, 0, 0, 1};
        // policy 1: split-K for problems whose tile grid is small next to their K
        const long tiles = (long)((M + cfg.bm - 1) / cfg.bm) * ((N + cfg.bn - 1) / cfg.bn);
        if (policy == 1 && tiles < 16 && K >= 1024) p.split_k = K >= 2048 ? 4 : 2;
        b.probs.push_back(p);
        b.flops += 2.0 * M * N * K;
    }
    printf("[config] BM=%d BN=%d BK=%d stages=%d split policy %d\n", cfg.bm, cfg.bn, cfg.bk, cfg.stages, policy);
    const float ms = run_batch(b, cfg, iters, graph);
    if (runMode == RunMode::TEST) {
        // every problem against the single-problem kernel with the same tile/split
        int bad = 0;
        for (int i = 0; i < count; i++) {
            const auto& p = b.probs[i];
            cli::DevBuf<uint16_t> ref((size_t)p.M * p.N);
            lthip_hgemm_config c = cfg;
            c.split_k = p.split_k;
            hgemm::run_cfg(&c, p.A, p.B, ref.p, p.M, p.N, p.K, 0x3C00, 0);
            std::vector<uint16_t> x(ref.n), y(ref.n);
            ref.download(x.data());
            HIP_CHECK(hipMemcpy(y.data(), p.C, sizeof(uint16_t) * y.size(), hipMemcpyDeviceToHost));
            bad += memcmp(x.data(), y.data(), sizeof(uint16_t) * x.size()) != 0;
        }
        if (bad) {
            printf("ERROR: %d of %d problems differ from the single-problem kernel\n", bad, count);
            cli::g_check_failed = true;
        }
    }
    const double tf = b.flops / (ms * 1e9);
    printf("# %d problems, %.3e flop: %.3f ms, %.1f TFLOP/s (%.1f%% of FP16 MFMA peak), tile %dx%dx%d\n", count, b.flops,
           ms, tf, 100.0 * tf / cli::kFp16MfmaTflops, cfg.bm, cfg.bn, cfg.bk);
    printf("# GFLOP/s: %f\n", tf * 1000.0);
}
#undef BENCH_RUNS
#undef WARMUP_RUNS
