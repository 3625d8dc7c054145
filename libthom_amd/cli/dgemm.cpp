// dgemm routine driver (the reference's routines/dgemm/dgemm.cpp): C = alpha*A*B
// + beta*C in FP64.  Implementations: "hip" (native FP64 MFMA) and
// "ozaki<S>" (split-precision on FP8 MFMA with S digits, S = 4..13; past 10
// only for shapes with M % 256 == N % 128 == K % 256 == 0).
// Inputs: N(0,1), or with a trailing "scaled" token rows of A and columns of B
// multiplied by 2^u, u uniform in [-20, 20].  The check compares sampled
// entries with a compensated (Dot2) FP64 dot product on the host and reports
// the error relative to (|A||B|)_ij.
#include <cmath>
#include <cstring>
#include <fstream>
#include <random>

#include "host_math.h"
#include "routines.h"

namespace dgemm {

#define BENCH_RUNS 10
#define WARMUP_RUNS 3
// dgemm.cpp:17 plus the configured 4096^3
static const std::vector<std::vector<int>> SIZES = {{512, 512, 512}, {1024, 1024, 1024}, {2048, 2048, 2048},
                                                    {4096, 4096, 4096}};

using DgemmImpl = float (*)(const double* A, const double* B, double* C, size_t M, size_t N, size_t K, double alpha,
                            double beta);   // dgemm.cpp:36

static float dgemm_hip(const double* A, const double* B, double* C, size_t M, size_t N, size_t K, double alpha,
                       double beta) {
    return cli::time_ms([&] { LTHIP_CHECK(lthip_dgemm_f64(A, B, C, M, N, K, alpha, beta, cli::stream())); });
}

template <int S>
static float dgemm_ozaki(const double* A, const double* B, double* C, size_t M, size_t N, size_t K, double alpha,
                         double beta) {
    const size_t need = lthip_dgemm_ozaki_workspace_size(M, N, K, S);
    void* ws = cli::workspace("ozaki", need);
    return cli::time_ms(
        [&] { LTHIP_CHECK(lthip_dgemm_ozaki(A, B, C, M, N, K, alpha, beta, S, ws, need, cli::stream())); });
}

static int slices_of(const std::string& name) { return name.rfind("ozaki", 0) == 0 ? atoi(name.c_str() + 5) : 0; }

// Host implementations: the reference's CPU product (dgemm.cpp:122-133,
// C = alpha A B + beta C in FP64) with the k loop outside the j loop (B read
// along rows), on one thread ("cpu") or all host cores ("cpu_mt"), timed with
// std::chrono; copies to and from the device are outside the timed region.
static float dgemm_host(const double* A, const double* B, double* C, size_t M, size_t N, size_t K, double alpha,
                        double beta, int threads) {
    const std::vector<double> hA = cli::to_host(A, M * K), hB = cli::to_host(B, K * N);
    std::vector<double> hC = cli::to_host(C, M * N);
    const float ms = cli::cpu_time_ms(threads, [&] {
        cli::parallel_for(M, [&](size_t lo, size_t hi) {
            std::vector<double> acc(N);
            for (size_t i = lo; i < hi; i++) {
                std::fill(acc.begin(), acc.end(), 0.0);
                for (size_t k = 0; k < K; k++) {
                    const double a = hA[i * K + k];
                    const double* b = &hB[k * N];
                    for (size_t j = 0; j < N; j++) acc[j] += a * b[j];
                }
                for (size_t j = 0; j < N; j++) hC[i * N + j] = alpha * acc[j] + beta * hC[i * N + j];
            }
        });
    });
    cli::to_device(C, hC);
    return ms;
}
static float dgemm_cpu(const double* A, const double* B, double* C, size_t M, size_t N, size_t K, double alpha,
                       double beta) {
    return dgemm_host(A, B, C, M, N, K, alpha, beta, 1);
}
static float dgemm_cpu_mt(const double* A, const double* B, double* C, size_t M, size_t N, size_t K, double alpha,
                          double beta) {
    return dgemm_host(A, B, C, M, N, K, alpha, beta, cli::host_cores());
}

static const std::unordered_map<std::string, DgemmImpl> dgemmImpls = {
    {"cpu", dgemm_cpu},
    {"cpu_mt", dgemm_cpu_mt},
#if LIB_HAS(DGEMM_HIP)
    {"hip", dgemm_hip},
#endif
#if LIB_HAS(DGEMM_OZAKI)
    {"ozaki4", dgemm_ozaki<4>}, {"ozaki5", dgemm_ozaki<5>}, {"ozaki6", dgemm_ozaki<6>}, {"ozaki7", dgemm_ozaki<7>},
    {"ozaki8", dgemm_ozaki<8>}, {"ozaki9", dgemm_ozaki<9>}, {"ozaki10", dgemm_ozaki<10>},
    {"ozaki11", dgemm_ozaki<11>}, {"ozaki12", dgemm_ozaki<12>}, {"ozaki13", dgemm_ozaki<13>},
#endif
};

static cli::DevBuf<double> gpuA, gpuB, gpuC;
static std::vector<double> cpuA, cpuB, cpuC0;
static const double cpuAlpha = 1.0, cpuBeta = 1.0;   // dgemm.cpp:106-107

static void resetVectors(int M, int N, int K, bool scaled) {
    cpuA.resize((size_t)M * K);
    cpuB.resize((size_t)K * N);
    cpuC0.resize((size_t)M * N);
    cli::parallel_for(cpuA.size(), [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; i++) cpuA[i] = cli::rng_normal(cli::kSeed, 4, i);
    });
    cli::parallel_for(cpuB.size(), [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; i++) cpuB[i] = cli::rng_normal(cli::kSeed, 5, i);
    });
    cli::parallel_for(cpuC0.size(), [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; i++) cpuC0[i] = cli::rng_normal(cli::kSeed, 16, i);
    });
    if (scaled) {
        for (int r = 0; r < M; r++) {
            const double s = std::ldexp(1.0, cli::rng_int(cli::kSeed, 6, r, -20, 20));
            for (int k = 0; k < K; k++) cpuA[(size_t)r * K + k] *= s;
        }
        for (int c = 0; c < N; c++) {
            const double s = std::ldexp(1.0, cli::rng_int(cli::kSeed, 7, c, -20, 20));
            for (int k = 0; k < K; k++) cpuB[(size_t)k * N + c] *= s;
        }
    }
    gpuA.upload(cpuA.data());
    gpuB.upload(cpuB.data());
    gpuC.upload(cpuC0.data());
}

// Sampled entries against Dot2 (Ogita-Rump-Oishi), error relative to
// (|alpha||A||B| + |beta||C0|)_ij.  Returns the max relative error.
static double compareToReference(int M, int N, int K, double tol) {
    std::vector<double> got((size_t)M * N);
    gpuC.download(got.data());
    const int samples = 4096;
    std::mt19937_64 gen(99);
    std::vector<double> errs(samples, 0.0);
    std::vector<int> rows(samples), cols(samples);
    for (int t = 0; t < samples; t++) {
        rows[t] = (int)(gen() % (uint64_t)M);
        cols[t] = (int)(gen() % (uint64_t)N);
    }
    cli::parallel_for(samples, [&](size_t lo, size_t hi) {
        for (size_t t = lo; t < hi; t++) {
            const int i = rows[t], j = cols[t];
            double s = 0.0, c = 0.0, mag = 0.0;
            for (int k = 0; k < K; k++) {
                const double x = cpuA[(size_t)i * K + k], y = cpuB[(size_t)k * N + j];
                const double p = x * y, pe = std::fma(x, y, -p);
                const double u = s + p, z = u - s, e = (s - (u - z)) + (p - z);
                s = u;
                c += e + pe;
                mag += std::fabs(p);
            }
            const double ref = cpuAlpha * (s + c) + cpuBeta * cpuC0[(size_t)i * N + j];
            const double scale = std::fabs(cpuAlpha) * mag + std::fabs(cpuBeta * cpuC0[(size_t)i * N + j]);
            errs[t] = std::fabs(got[(size_t)i * N + j] - ref) / (scale > 0 ? scale : 1.0);
        }
    });
    double mx = 0.0;
    int bad = 0;
    for (double e : errs) {
        mx = std::max(mx, e);
        bad += e > tol;
    }
    if (bad) {
        printf("Number of different results: %d of %d sampled (max rel err %.3e > tol %.1e)\n", bad, samples, mx, tol);
        cli::g_check_failed = true;
    }
    return mx;
}

static float getGflops(int M, int N, int K, float ms) { return (float)(2.0 * M * N * K / (ms * 1e6)); }

static float runner(DgemmImpl impl, std::vector<int> sizes, int nb_runs, int nb_warmup_runs, bool checkResult,
                    bool printEach, int argc, char* argv[]) {
    const int M = sizes[0], N = sizes[1], K = sizes[2];
    const bool scaled = strcmp(argv[argc - 1], "scaled") == 0;
    const int S = slices_of(argv[3]);
    // native: FP64 error level; emulated: the S-digit truncation (4 + 5(S-1) bits, plus margin)
    const double tol = S ? std::ldexp(1.0, -(5 * S - 3)) : 1e-13;
    gpuA.alloc((size_t)M * K);
    gpuB.alloc((size_t)K * N);
    gpuC.alloc((size_t)M * N);
    resetVectors(M, N, K, scaled);
    float sum_gflops = 0.0f;
    for (int i = -nb_warmup_runs; i < nb_runs; i++) {
        if (checkResult) gpuC.upload(cpuC0.data());
        const float ms = impl(gpuA.p, gpuB.p, gpuC.p, M, N, K, cpuAlpha, cpuBeta);
        double err = -1.0;
        if (checkResult) err = compareToReference(M, N, K, tol);
        if (i >= 0) {
            const float g = getGflops(M, N, K, ms);
            if (printEach) {
                printf("[internal][iter %d]: %f GFLOP/s (%.1f%% of FP64 MFMA peak)", i, g,
                       100.0 * g / 1000.0 / cli::kFp64MfmaTflops);
                if (err >= 0) printf(" | max rel err vs Dot2 %.3e", err);
                printf("\n");
            }
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

}  // namespace dgemm

void dgemm_handler(RunMode runMode, int argc, char* argv[]) {
    switch (runMode) {
        case RunMode::TEST:
            Handlers<dgemm::DgemmImpl, 3, BENCH_RUNS, WARMUP_RUNS>::test_handler(argc, argv, dgemm::dgemmImpls,
                                                                                 dgemm::runner);
            break;
        case RunMode::BENCH:
            Handlers<dgemm::DgemmImpl, 3, BENCH_RUNS, WARMUP_RUNS>::bench_handler(
                argc, argv, dgemm::dgemmImpls, dgemm::runner, dgemm::SIZES, dgemm::exportCSV);
            break;
        default:
            fprintf(stderr, "Mode not supported by dgemm\n");
            exit(1);
    }
}
#undef BENCH_RUNS
#undef WARMUP_RUNS
