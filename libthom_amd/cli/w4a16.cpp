// w4a16 routine driver (the reference's routines/w4a16/w4a16.cpp):
// C = A * dequant(W) with int4 weights.  Inputs follow the configured format
// (uniform nibbles, FP16 scales in [0.005, 0.02], per-(group, column) int4
// zero-points); a trailing "sym" token instead quantizes fp16 weights with the
// reference's symmetric quantizer (quantize.cpp:8-53, zero-point 8).
// Implementations: "hip" (size-based choice), "stream" (split-K MFMA stream
// kernel), "decode" (vector-ALU kernel, M <= 4), "skinny" (MFMA with exact
// (q - z) operands, M <= 64), "tiled" (MFMA tiles, dequantized into LDS),
// "pipelined" (DMA-ring MFMA tiles, dequantized in registers), "tuned"
// (tuning database, else "hip").  Shapes outside an implementation's domain
// run the size-based choice.
// `tune w4a16 M N K` sweeps every configuration into the database.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>

#include "host_math.h"
#include "routines.h"
#include "tuning.h"

namespace w4a16 {

#define BENCH_RUNS 20
#define WARMUP_RUNS 5
#define GROUP_SIZE 128
// w4a16.cpp:17-27 (M sweep at K = N = 8192) and the configured shapes
static const std::vector<std::vector<int>> SIZES = {
    {1, 8192, 8192},  {2, 8192, 8192},  {4, 8192, 8192},   {8, 8192, 8192},    {16, 8192, 8192},
    {32, 8192, 8192}, {64, 8192, 8192}, {128, 8192, 8192}, {256, 8192, 8192},  {1, 11008, 4096},
    {16, 11008, 4096}, {128, 11008, 4096}, {2048, 11008, 4096}};

struct Problem {   // the reference's W4A16Problem (w4a16.h:23-32) plus zero-points
    const uint16_t* A;
    const uint32_t* W;
    const uint16_t* scales;
    const uint32_t* qzeros;
    uint16_t* C;
    size_t M, N, K;
};
using W4A16Function = float (*)(const Problem&);

static float run_cfg(const Problem& p, const lthip_w4a16_config* cfg) {
    const size_t need = lthip_w4a16_workspace_size(p.M, p.N, p.K, cfg);
    void* ws = need ? cli::workspace("w4a16", need) : nullptr;
    return cli::time_ms([&] {
        LTHIP_CHECK(lthip_w4a16_gemm(p.A, p.W, p.scales, p.qzeros, p.C, p.M, p.N, p.K, GROUP_SIZE, cfg, ws, need,
                                     cli::stream()));
    });
}
static float w4a16_hip(const Problem& p) { return run_cfg(p, nullptr); }
static float w4a16_stream(const Problem& p) {
    lthip_w4a16_config c{1, 64, 16, 1};
    if (p.M > 128) return run_cfg(p, nullptr);
    return run_cfg(p, &c);
}
static float w4a16_decode(const Problem& p) {
    if (p.M > 4) return run_cfg(p, nullptr);
    lthip_w4a16_config c{4, p.N >= 10240 ? 64 : 32, 0, 1};
    return run_cfg(p, &c);
}
static float w4a16_skinny(const Problem& p) {
    if (p.M > 64) return run_cfg(p, nullptr);
    lthip_w4a16_config c{5, 64, 0, 1};
    return run_cfg(p, &c);
}
static float w4a16_tiled(const Problem& p) {
    lthip_w4a16_config c{2, 128, 128, 1};
    return run_cfg(p, &c);
}
static float w4a16_pipelined(const Problem& p) {
    if (p.N % 32 != 0) return run_cfg(p, nullptr);   // the pipelined tiles need N % 32 == 0
    // the library's choice: 256 x 256 once those tiles cover half the CUs, else 128 x 128
    const size_t tiles256 = ((p.M + 255) / 256) * ((p.N + 255) / 256);
    const int t = p.M > 512 && 2 * tiles256 >= 256 ? 256 : 128;
    lthip_w4a16_config c{3, t, t, 1};
    return run_cfg(p, &c);
}
static float w4a16_tuned(const Problem& p) {
    lthip_w4a16_config c;
    if (!tuning::lookup_w4a16((int)p.M, (int)p.N, (int)p.K, GROUP_SIZE, &c)) return run_cfg(p, nullptr);
    return run_cfg(p, &c);
}

// Host implementations: the reference's CPU path (quantize.cpp:58-65, W =
// fp16((q - z) s) rounded once, then w4a16.cpp:178-184's product with an FP32
// sum, C = fp16(sum)) on one thread ("cpu") or all host cores ("cpu_mt"),
// timed with std::chrono; copies to and from the device are outside the timed
// region.
static float w4a16_host(const Problem& p, int threads) {
    const size_t M = p.M, N = p.N, K = p.K, G = K / GROUP_SIZE;
    const std::vector<uint16_t> hA = cli::to_host(p.A, M * K), hS = cli::to_host(p.scales, G * N);
    const std::vector<uint32_t> hW = cli::to_host(p.W, (K / 8) * N);
    const std::vector<uint32_t> hZ = p.qzeros ? cli::to_host(p.qzeros, G * (N / 8)) : std::vector<uint32_t>();
    std::vector<uint16_t> hC(M * N);
    const float ms = cli::cpu_time_ms(threads, [&] {
        std::vector<float> w(K * N);
        cli::parallel_for(K, [&](size_t lo, size_t hi) {
            for (size_t k = lo; k < hi; k++) {
                const size_t g = k / GROUP_SIZE;
                for (size_t n = 0; n < N; n++) {
                    const int q = (int)((hW[(k / 8) * N + n] >> (4 * (k % 8))) & 0xFu);
                    const int z = p.qzeros ? (int)((hZ[g * (N / 8) + n / 8] >> (4 * (n % 8))) & 0xFu) : 8;
                    w[k * N + n] = cli::f16_to_f32(cli::f32_to_f16((float)(q - z) * cli::f16_to_f32(hS[g * N + n])));
                }
            }
        });
        cli::parallel_for(M, [&](size_t lo, size_t hi) {
            std::vector<float> acc(N);
            for (size_t i = lo; i < hi; i++) {
                std::fill(acc.begin(), acc.end(), 0.0f);
                for (size_t k = 0; k < K; k++) {
                    const float a = cli::f16_to_f32(hA[i * K + k]);
                    const float* wk = &w[k * N];
                    for (size_t n = 0; n < N; n++) acc[n] += a * wk[n];
                }
                for (size_t n = 0; n < N; n++) hC[i * N + n] = cli::f32_to_f16(acc[n]);
            }
        });
    });
    cli::to_device(p.C, hC);
    return ms;
}
static cli::DevBuf<uint16_t> gpuWdq;   // the dequantised weights of the baseline

// The baseline the reference times beside every implementation (w4a16.cpp:306,
// cuBLAS fp16 on the dequantised weights): here the library's own FP16 GEMM on
// the weights dequantised once into gpuWdq (outside the timed region).
static float dense_fp16_on_dequantised(const Problem& p) {
    const uint16_t one = 0x3C00, zero = 0x0000;
    const size_t need = lthip_hgemm_workspace_size(p.M, p.N, p.K, nullptr);
    void* ws = need ? cli::workspace("w4a16_dense", need) : nullptr;
    return cli::time_ms([&] {
        LTHIP_CHECK(lthip_hgemm_f16(p.A, gpuWdq.p, p.C, p.M, p.N, p.K, one, zero, nullptr, ws, need, cli::stream()));
    });
}

static float w4a16_cpu(const Problem& p) { return w4a16_host(p, 1); }
static float w4a16_cpu_mt(const Problem& p) { return w4a16_host(p, cli::host_cores()); }

static const std::unordered_map<std::string, W4A16Function> w4a16Impls = {
#if LIB_HAS(W4A16_HIP)
    {"hip", w4a16_hip},       {"stream", w4a16_stream}, {"decode", w4a16_decode}, {"skinny", w4a16_skinny},
    {"tiled", w4a16_tiled},   {"pipelined", w4a16_pipelined}, {"tuned", w4a16_tuned},
#endif
    {"cpu", w4a16_cpu},       {"cpu_mt", w4a16_cpu_mt},
};

static cli::DevBuf<uint16_t> gpuA, gpuScales, gpuC;
static cli::DevBuf<uint32_t> gpuW, gpuZ;
static std::vector<uint16_t> cpuA, cpuScales, cpuWdq, cpuCref;
static bool symmetric = false;

// quantize.cpp:8-37: symmetric per-(group, column) quantization of fp16 W
static void quantize_w4(const std::vector<uint16_t>& W, std::vector<uint8_t>& q, std::vector<uint16_t>& scales, int K,
                        int N) {
    const int groups = K / GROUP_SIZE;
    std::vector<float> mx(N);
    for (int g = 0; g < groups; g++) {
        std::fill(mx.begin(), mx.end(), 0.0f);
        for (int k = g * GROUP_SIZE; k < (g + 1) * GROUP_SIZE; k++)
            for (int n = 0; n < N; n++) mx[n] = std::max(mx[n], std::fabs(cli::f16_to_f32(W[(size_t)k * N + n])));
        for (int n = 0; n < N; n++) {
            const float s = 2.0f * mx[n] / 15.0f;
            scales[(size_t)g * N + n] = cli::f32_to_f16(s > 0.0f ? s : 1.0f);
        }
        for (int k = g * GROUP_SIZE; k < (g + 1) * GROUP_SIZE; k++)
            for (int n = 0; n < N; n++) {
                const float s = cli::f16_to_f32(scales[(size_t)g * N + n]);
                const int v = (int)std::nearbyint(cli::f16_to_f32(W[(size_t)k * N + n]) / s) + 8;
                q[(size_t)k * N + n] = (uint8_t)std::min(std::max(v, 0), 15);
            }
    }
}

// builds the problem on the host; the CPU check uses the dequantized weights
static void initProblem(int M, int N, int K) {
    const int G = K / GROUP_SIZE;
    cpuA.resize((size_t)M * K);
    for (size_t i = 0; i < cpuA.size(); i++) cpuA[i] = cli::f32_to_f16((float)cli::rng_normal(cli::kSeed, 9, i));
    std::vector<uint8_t> q((size_t)K * N), z((size_t)G * N);
    cpuScales.resize((size_t)G * N);
    if (symmetric) {
        std::vector<uint16_t> W((size_t)K * N);
        for (size_t i = 0; i < W.size(); i++)
            W[i] = cli::f32_to_f16((float)(2.0 * cli::rng_unit(cli::kSeed, 17, i) - 1.0));
        quantize_w4(W, q, cpuScales, K, N);
        std::fill(z.begin(), z.end(), 8);
    } else {
        cli::parallel_for(q.size(), [&](size_t lo, size_t hi) {
            for (size_t i = lo; i < hi; i++) q[i] = (uint8_t)cli::rng_int(cli::kSeed, 10, i, 0, 15);
        });
        for (size_t i = 0; i < cpuScales.size(); i++)
            cpuScales[i] = cli::f32_to_f16((float)(0.005 + 0.015 * cli::rng_unit(cli::kSeed, 11, i)));
        for (size_t i = 0; i < z.size(); i++) z[i] = (uint8_t)cli::rng_int(cli::kSeed, 12, i, 0, 15);
    }
    // pack (quantize.cpp:42-53) and the zero-points, 8 per uint32 along N
    std::vector<uint32_t> packed((size_t)(K / 8) * N, 0u), qz((size_t)G * (N / 8), 0u);
    for (int k = 0; k < K; k++)
        for (int n = 0; n < N; n++) packed[(size_t)(k / 8) * N + n] |= (uint32_t)q[(size_t)k * N + n] << (4 * (k % 8));
    for (int g = 0; g < G; g++)
        for (int n = 0; n < N; n++) qz[(size_t)g * (N / 8) + n / 8] |= (uint32_t)z[(size_t)g * N + n] << (4 * (n % 8));
    // dequantize (quantize.cpp:58-65 with zero-points): one rounding
    cpuWdq.resize((size_t)K * N);
    for (int k = 0; k < K; k++)
        for (int n = 0; n < N; n++) {
            const int g = k / GROUP_SIZE;
            cpuWdq[(size_t)k * N + n] = cli::f32_to_f16((float)((int)q[(size_t)k * N + n] - (int)z[(size_t)g * N + n]) *
                                                        cli::f16_to_f32(cpuScales[(size_t)g * N + n]));
        }
    gpuA.upload(cpuA.data());
    gpuW.upload(packed.data());
    gpuScales.upload(cpuScales.data());
    gpuZ.upload(qz.data());
}

// C_ref rows (FP32 k-ascending sum over the dequantized weights), up to 64 rows
static std::vector<int> ref_rows(int M) {
    std::vector<int> r;
    const int step = M > 64 ? M / 64 : 1;
    for (int i = 0; i < M; i += step) r.push_back(i);
    return r;
}

static void computeReference(int M, int N, int K) {
    const auto rows = ref_rows(M);
    cpuCref.assign(rows.size() * (size_t)N, 0);
    std::vector<float> fw((size_t)K * N);
    for (size_t i = 0; i < fw.size(); i++) fw[i] = cli::f16_to_f32(cpuWdq[i]);
    cli::parallel_for(rows.size(), [&](size_t lo, size_t hi) {
        std::vector<float> acc(N);
        for (size_t t = lo; t < hi; t++) {
            std::fill(acc.begin(), acc.end(), 0.0f);
            for (int k = 0; k < K; k++) {
                const float a = cli::f16_to_f32(cpuA[(size_t)rows[t] * K + k]);
                const float* w = &fw[(size_t)k * N];
                for (int n = 0; n < N; n++) acc[n] += a * w[n];
            }
            for (int n = 0; n < N; n++) cpuCref[t * N + n] = cli::f32_to_f16(acc[n]);
        }
    });
}

// w4a16.cpp:194-233: |r - ref| <= 1e-2 |ref| + 1e-3 RMS(C_ref)
static bool compareToReference(int M, int N) {
    const auto rows = ref_rows(M);
    std::vector<uint16_t> got((size_t)M * N);
    gpuC.download(got.data());
    double ss = 0.0;
    for (uint16_t v : cpuCref) ss += (double)cli::f16_to_f32(v) * cli::f16_to_f32(v);
    const float atol = 1e-3f * (float)std::sqrt(ss / (double)cpuCref.size());
    int bad = 0;
    for (size_t t = 0; t < rows.size(); t++)
        for (int n = 0; n < N; n++) {
            const float r = cli::f16_to_f32(got[(size_t)rows[t] * N + n]);
            const float ref = cli::f16_to_f32(cpuCref[t * N + n]);
            if (std::fabs(r - ref) > 1e-2f * std::fabs(ref) + atol) {
                if (++bad <= 8) printf("row %d col %d | GPU: %f | reference: %f\n", rows[t], n, r, ref);
            }
        }
    if (bad) {
        printf("Number of different results: %d (tolerance: 1%% + %f)\n", bad, atol);
        cli::g_check_failed = true;
    }
    return bad == 0;
}

struct Metrics {
    // cublas_* (the reference's column names): the dense FP16 GEMM on the
    // dequantised weights, the reference's baseline, run by this library
    float time_us, gflops, gbps, pct_peak_bw, cublas_time_us, speedup_vs_cublas;
};
static std::map<std::vector<int>, Metrics> metrics;

// minimal bytes between HBM and the CUs (w4a16.cpp:268-280 plus zero-points)
static double bytesMoved(int M, int N, int K) {
    return 2.0 * M * K + 4.0 * (K / 8) * N + 2.0 * (K / GROUP_SIZE) * N + 4.0 * (K / GROUP_SIZE) * (N / 8) +
           2.0 * M * N;
}

static float median(std::vector<float> v) {
    std::sort(v.begin(), v.end());
    const size_t m = v.size() / 2;
    return v.size() % 2 ? v[m] : 0.5f * (v[m - 1] + v[m]);
}

static std::vector<float> measure(W4A16Function f, const Problem& p, int runs, int warmup, bool check) {
    std::vector<float> t;
    for (int i = -warmup; i < runs; i++) {
        if (check) HIP_CHECK(hipMemset(gpuC.p, 0, sizeof(uint16_t) * gpuC.n));
        cli::flush_caches();
        const float ms = f(p);
        if (check) compareToReference((int)p.M, (int)p.N);
        if (i >= 0) t.push_back(ms);
    }
    return t;
}

static float runner(W4A16Function impl, std::vector<int> sizes, int nb_runs, int nb_warmup_runs, bool checkResult,
                    bool printEach, int argc, char* argv[]) {
    const int M = sizes[0], N = sizes[1], K = sizes[2];
    if (K % GROUP_SIZE != 0 || N % 8 != 0) {
        fprintf(stderr, "K must be a multiple of the group size (%d) and N a multiple of 8\n", GROUP_SIZE);
        exit(1);
    }
    symmetric = strcmp(argv[argc - 1], "sym") == 0;
    nb_runs = std::max(nb_runs, 1);
    nb_warmup_runs = std::max(nb_warmup_runs, 1);
    gpuA.alloc((size_t)M * K);
    gpuW.alloc((size_t)(K / 8) * N);
    gpuScales.alloc((size_t)(K / GROUP_SIZE) * N);
    gpuZ.alloc((size_t)(K / GROUP_SIZE) * (N / 8));
    gpuC.alloc((size_t)M * N);
    initProblem(M, N, K);
    if (checkResult) computeReference(M, N, K);
    const Problem p{gpuA.p, gpuW.p, gpuScales.p, symmetric ? nullptr : gpuZ.p, gpuC.p, (size_t)M, (size_t)N,
                    (size_t)K};
    const auto times = measure(impl, p, nb_runs, nb_warmup_runs, checkResult);
    // baseline: a dense FP16 GEMM on the dequantised weights (w4a16.cpp:306)
    gpuWdq.alloc((size_t)K * N);
    LTHIP_CHECK(lthip_w4_dequant(p.W, p.scales, p.qzeros, gpuWdq.p, (size_t)K, (size_t)N, GROUP_SIZE, cli::stream()));
    const auto base = measure(dense_fp16_on_dequantised, p, nb_runs, nb_warmup_runs, false);
    if (printEach)
        for (int i = 0; i < nb_runs; i++)
            printf("[internal][iter %d]: %f GFLOP/s | %.2f us (dense fp16 on dequantised W: %.2f us)\n", i,
                   2.0 * M * N * K / (times[i] * 1e6), times[i] * 1e3f, base[i] * 1e3f);
    const float ms = median(times), bms = median(base);
    Metrics m;
    m.time_us = ms * 1e3f;
    m.gflops = (float)(2.0 * M * N * K / (ms * 1e6));
    m.gbps = (float)(bytesMoved(M, N, K) / (ms * 1e6));
    m.pct_peak_bw = 100.0f * m.gbps / (float)cli::kHbmPeakGBs;
    m.cublas_time_us = bms * 1e3f;
    m.speedup_vs_cublas = bms / ms;
    metrics[sizes] = m;
    printf("# median: %.2f us | %.1f GB/s (%.1f%% of %.0f GB/s) | %.1f TFLOP/s (%.1f%% of FP16 MFMA) | dense fp16 on dequantised W: %.2f us | speedup: %.2fx\n",
           m.time_us, m.gbps, m.pct_peak_bw, cli::kHbmPeakGBs, m.gflops / 1000.0,
           100.0 * m.gflops / 1000.0 / cli::kFp16MfmaTflops, m.cublas_time_us, m.speedup_vs_cublas);
    gpuA.free();
    gpuW.free();
    gpuScales.free();
    gpuZ.free();
    gpuC.free();
    gpuWdq.free();
    return m.gflops;
}

static void exportCSV(const std::map<std::vector<int>, float>& results, const std::string& filename) {
    std::ofstream file(filename);
    file << "M,N,K,time_us,gflops,gbps,pct_peak_bw,cublas_time_us,speedup_vs_cublas\n";   // w4a16.cpp:343
    for (const auto& [s, g] : results) {
        const Metrics& m = metrics.at(s);
        file << s[0] << "," << s[1] << "," << s[2] << "," << m.time_us << "," << m.gflops << "," << m.gbps << ","
             << m.pct_peak_bw << "," << m.cublas_time_us << "," << m.speedup_vs_cublas << "\n";
    }
}

// ./libthom tune w4a16 M N K [runs]: every configuration of the search space,
// validated, timed (median, caches flushed), appended to the tuning database.
static void tune(int argc, char* argv[]) {
    if (argc < 6) {
        fprintf(stderr, "Usage: ./libthom tune w4a16 M N K [runs]\n");
        exit(1);
    }
    const int M = atoi(argv[3]), N = atoi(argv[4]), K = atoi(argv[5]);
    const int runs = argc > 6 ? atoi(argv[6]) : 10;
    if (M <= 0 || N <= 0 || K <= 0 || K % GROUP_SIZE || N % 8 || runs <= 0) {
        fprintf(stderr, "bad sizes\n");
        exit(1);
    }
    gpuA.alloc((size_t)M * K);
    gpuW.alloc((size_t)(K / 8) * N);
    gpuScales.alloc((size_t)(K / GROUP_SIZE) * N);
    gpuZ.alloc((size_t)(K / GROUP_SIZE) * (N / 8));
    gpuC.alloc((size_t)M * N);
    initProblem(M, N, K);
    computeReference(M, N, K);
    const Problem p{gpuA.p, gpuW.p, gpuScales.p, gpuZ.p, gpuC.p, (size_t)M, (size_t)N, (size_t)K};
    std::vector<lthip_w4a16_config> cands;
    if (M <= 4) {
        cands.push_back({4, 64, 0, 1});
        cands.push_back({4, 32, 0, 1});
    }
    if (M <= 64) cands.push_back({5, 64, 0, 1});
    if (M <= 128)
        for (int s : {1, 2, 3, 4, 6, 8, 12, 16}) cands.push_back({1, 64, 16, s});
    for (auto bm_bn : {std::pair<int, int>{64, 64}, {64, 128}, {128, 128}, {128, 256}, {256, 128}, {256, 256}})
        cands.push_back({2, bm_bn.second, bm_bn.first, 1});
    if (N % 32 == 0) {
        for (auto bm_bn : {std::pair<int, int>{128, 128}, {128, 256}, {256, 128}, {256, 256}})
            cands.push_back({3, bm_bn.second, bm_bn.first, 1});
        cands.push_back({6, 128, 128, 1});
    }
    float best = 1e30f;
    lthip_w4a16_config bestc{};
    for (const auto& c : cands) {
        char name[96];
        snprintf(name, sizeof(name), "kind=%d tile_n=%d tile_m=%d split_k=%d", c.kind, c.tile_n, c.tile_m, c.split_k);
        HIP_CHECK(hipMemset(gpuC.p, 0, sizeof(uint16_t) * gpuC.n));
        run_cfg(p, &c);
        if (!compareToReference(M, N)) {
            printf("[failed] %s\n", name);
            continue;
        }
        std::vector<float> t;
        for (int r = 0; r < runs + 2; r++) {
            cli::flush_caches();
            const float ms = run_cfg(p, &c);
            if (r >= 2) t.push_back(ms);
        }
        const float med = median(t);
        printf("[result] %s: %.2f us\n", name, med * 1e3f);
        tuning::append_w4a16(M, N, K, GROUP_SIZE, c, med);
        if (med < best) {
            best = med;
            bestc = c;
        }
    }
    printf("[best] kind=%d tile_n=%d tile_m=%d split_k=%d: %.2f us\n", bestc.kind, bestc.tile_n, bestc.tile_m,
           bestc.split_k, best * 1e3f);
}

}  // namespace w4a16

// ---- grouped w4a16 (no counterpart in the reference): `experts` problems of
// N x K int4 weights each, M per problem log-uniform in [1, 512] from the
// seeded generator, all in one persistent launch (or its hipGraph).  Inputs
// come from the device generator: the check compares every problem bit for
// bit with the single-problem pipelined tile of the same shape run alone.
void w4a16_grouped_handler(RunMode runMode, int argc, char* argv[]) {
    static const std::unordered_map<std::string, bool> impls = {{"persistent", false}, {"graph", true}};
    if (runMode == RunMode::TUNE || argc < 4 || !impls.count(argv[3])) {
        fprintf(stderr,
                "Usage: ./libthom <test|bench> w4a16_grouped <persistent|graph> [experts=32 [N=2048 [K=4096 "
                "[nb_iter=10]]]]\n");
        exit(1);
    }
    const bool graph = impls.at(argv[3]);
    const int count = argc > 4 ? atoi(argv[4]) : 32;
    const int N = argc > 5 ? atoi(argv[5]) : 2048;
    const int K = argc > 6 ? atoi(argv[6]) : 4096;
    const int iters = argc > 7 ? atoi(argv[7]) : 10;
    const int group = GROUP_SIZE;
    if (count <= 0 || iters <= 0 || N <= 0 || K <= 0 || N % 32 || K % group) {
        fprintf(stderr, "need positive counts, N %% 32 == 0 and K %% %d == 0\n", group);
        exit(1);
    }
    std::vector<cli::DevBuf<uint16_t>*> h16;
    std::vector<cli::DevBuf<uint32_t>*> h32;
    std::vector<lthip_w4a16_problem> probs;
    double flops = 0.0;
    for (int i = 0; i < count; i++) {
        const double u = cli::rng_unit(cli::kSeed, 23, (uint64_t)i);
        const int M = std::max(1, (int)std::nearbyint(std::exp(std::log(512.0) * u)));
        auto* A = new cli::DevBuf<uint16_t>((size_t)M * K);
        auto* S = new cli::DevBuf<uint16_t>((size_t)(K / group) * N);
        auto* C = new cli::DevBuf<uint16_t>((size_t)M * N);
        auto* W = new cli::DevBuf<uint32_t>((size_t)(K / 8) * N);
        auto* Z = new cli::DevBuf<uint32_t>((size_t)(K / group) * (N / 8));
        h16.insert(h16.end(), {A, S, C});
        h32.insert(h32.end(), {W, Z});
        LTHIP_CHECK(lthip_fill_uniform_f16(A->p, A->n, -1.0, 1.0, cli::kSeed, 3000 + 4 * i, cli::stream()));
        LTHIP_CHECK(lthip_fill_uniform_f16(S->p, S->n, 0.005, 0.02, cli::kSeed, 3001 + 4 * i, cli::stream()));
        // any 32-bit pattern is 8 valid nibbles: the weights and zero-points
        // take the bits of uniform FP32 values
        LTHIP_CHECK(lthip_fill_uniform_f32(reinterpret_cast<float*>(W->p), W->n, 0.0, 1.0, cli::kSeed, 3002 + 4 * i,
                                           cli::stream()));
        LTHIP_CHECK(lthip_fill_uniform_f32(reinterpret_cast<float*>(Z->p), Z->n, 0.0, 1.0, cli::kSeed, 3003 + 4 * i,
                                           cli::stream()));
        probs.push_back(lthip_w4a16_problem{A->p, W->p, S->p, Z->p, C->p, M, N, K, group});
        flops += 2.0 * M * N * K;
    }
    const size_t need = lthip_w4a16_grouped_workspace_size(probs.data(), count, nullptr);
    void* ws = cli::workspace("w4a16_grouped", need);
    lthip_w4a16_grouped_plan* plan = nullptr;
    LTHIP_CHECK(lthip_w4a16_grouped_plan_create(probs.data(), count, nullptr, ws, need, cli::stream(), &plan));
    if (graph) LTHIP_CHECK(lthip_w4a16_grouped_capture(plan, cli::stream()));
    auto once = [&] {
        if (graph)
            LTHIP_CHECK(lthip_w4a16_grouped_graph_launch(plan, cli::stream()));
        else
            LTHIP_CHECK(lthip_w4a16_grouped_run(plan, cli::stream()));
    };
    once();
    const float ms = cli::time_ms([&] {
        for (int r = 0; r < iters; r++) once();
    }) / iters;
    int work = 0, grid = 0;
    LTHIP_CHECK(lthip_w4a16_grouped_plan_info(plan, &work, &grid));
    LTHIP_CHECK(lthip_w4a16_grouped_plan_destroy(plan));
    if (runMode == RunMode::TEST) {
        int bad = 0;
        for (const auto& p : probs) {
            const lthip_w4a16_config c{3, 256, p.M > 128 ? 256 : (p.M > 64 ? 128 : 64), 1};
            cli::DevBuf<uint16_t> ref((size_t)p.M * N);
            // no workspace: whole tiles, the same work as a grouped item
            LTHIP_CHECK(lthip_w4a16_gemm(p.A, p.W_packed, p.scales, p.qzeros, ref.p, p.M, N, K, group, &c, nullptr,
                                         0, cli::stream()));
            std::vector<uint16_t> x(ref.n), y(ref.n);
            ref.download(x.data());
            HIP_CHECK(hipMemcpy(y.data(), p.C, sizeof(uint16_t) * y.size(), hipMemcpyDeviceToHost));
            bad += memcmp(x.data(), y.data(), sizeof(uint16_t) * x.size()) != 0;
        }
        if (bad) {
            printf("ERROR: %d of %d problems differ from the single-problem tile\n", bad, count);
            cli::g_check_failed = true;
        }
    }
    for (auto* b : h16) delete b;
    for (auto* b : h32) delete b;
    const double tf = flops / (ms * 1e9);
    printf("# %d problems (N %d, K %d, M log-uniform 1-512), %d work items on %d blocks: %.3f ms, %.1f TFLOP/s "
           "(%.1f%% of FP16 MFMA peak)\n",
           count, N, K, work, grid, ms, tf, 100.0 * tf / cli::kFp16MfmaTflops);
    printf("# GFLOP/s: %f\n", tf * 1000.0);
}

void w4a16_handler(RunMode runMode, int argc, char* argv[]) {
    switch (runMode) {
        case RunMode::TEST:
            Handlers<w4a16::W4A16Function, 3, BENCH_RUNS, WARMUP_RUNS>::test_handler(argc, argv, w4a16::w4a16Impls,
                                                                                     w4a16::runner);
            break;
        case RunMode::BENCH:
            Handlers<w4a16::W4A16Function, 3, BENCH_RUNS, WARMUP_RUNS>::bench_handler(
                argc, argv, w4a16::w4a16Impls, w4a16::runner, w4a16::SIZES, w4a16::exportCSV);
            break;
        case RunMode::TUNE:
            w4a16::tune(argc, argv);
            break;
    }
}
#undef BENCH_RUNS
#undef WARMUP_RUNS
