// sgemv routine driver (the reference's routines/sgemv/sgemv.cpp):
// y = alpha*op(A)x + beta*y with device-pointer alpha/beta, row- or column-major
// selected by a trailing "row"/"col" token (sgemv.cpp:180).
#include <cmath>
#include <cstring>
#include <fstream>

#include "host_math.h"
#include "routines.h"

namespace sgemv {

#define BENCH_RUNS 10
#define WARMUP_RUNS 5
// sgemv.cpp:18 plus the configured 16384 x 8192 shape
static const std::vector<std::vector<int>> SIZES = {{1000, 1000}, {4096, 4096}, {16384, 8192}, {18432, 18432}};

using SgemvImpl = float (*)(int m, int n, const float* alpha, const float* A, int lda, const float* x, int incx,
                            const float* beta, float* y, int incy, bool rowMajor);   // sgemv.cpp:38

static float sgemv_hip(int m, int n, const float* alpha, const float* A, int lda, const float* x, int incx,
                       const float* beta, float* y, int incy, bool rowMajor) {
    const size_t need = lthip_sgemv_f32_workspace_size(rowMajor ? 1 : 0, m, n);
    void* ws = need ? cli::workspace("sgemv", need) : nullptr;
    return cli::time_ms([&] {
        LTHIP_CHECK(lthip_sgemv_f32(rowMajor ? 1 : 0, m, n, alpha, A, lda, x, incx, beta, y, incy, ws, need,
                                    cli::stream()));
    });
}

// Host implementations: the reference's CPU check (sgemv.cpp:132-147, an FP64
// dot per row, then (float)(alpha dot + beta y)) on one thread ("cpu") or all
// host cores ("cpu_mt"), timed with std::chrono; copies to and from the device
// are outside the timed region.
static float sgemv_host(int m, int n, const float* alpha, const float* A, int lda, const float* x, int incx,
                        const float* beta, float* y, int incy, bool rowMajor, int threads) {
    if (m <= 0) return 0.0f;
    const size_t a_elems = n <= 0 ? 0 : (rowMajor ? (size_t)(m - 1) * lda + n : (size_t)(n - 1) * lda + m);
    const std::vector<float> hA = cli::to_host(A, a_elems);
    const std::vector<float> hx = cli::to_host(x, n <= 0 ? 0 : (size_t)(n - 1) * incx + 1);
    std::vector<float> hy = cli::to_host(y, (size_t)(m - 1) * incy + 1);
    const double al = cli::to_host(alpha, 1)[0], be = cli::to_host(beta, 1)[0];
    const float ms = cli::cpu_time_ms(threads, [&] {
        cli::parallel_for((size_t)m, [&](size_t lo, size_t hi) {
            for (size_t row = lo; row < hi; row++) {
                double dot = 0.0;
                for (int col = 0; col < n; col++) {
                    const double a = rowMajor ? hA[row * lda + col] : hA[(size_t)col * lda + row];
                    dot += a * (double)hx[(size_t)col * incx];
                }
                float& yr = hy[row * incy];
                yr = (float)(al * dot + be * (double)yr);
            }
        });
    });
    cli::to_device(y, hy);
    return ms;
}
static float sgemv_cpu(int m, int n, const float* alpha, const float* A, int lda, const float* x, int incx,
                       const float* beta, float* y, int incy, bool rowMajor) {
    return sgemv_host(m, n, alpha, A, lda, x, incx, beta, y, incy, rowMajor, 1);
}
static float sgemv_cpu_mt(int m, int n, const float* alpha, const float* A, int lda, const float* x, int incx,
                          const float* beta, float* y, int incy, bool rowMajor) {
    return sgemv_host(m, n, alpha, A, lda, x, incx, beta, y, incy, rowMajor, cli::host_cores());
}

static const std::unordered_map<std::string, SgemvImpl> sgemvImpls = {
#if LIB_HAS(SGEMV_HIP)
    {"hip", sgemv_hip},
#endif
    {"cpu", sgemv_cpu},
    {"cpu_mt", sgemv_cpu_mt},
};

static cli::DevBuf<float> gpuA, gpuX, gpuY, gpuAlpha, gpuBeta;
static std::vector<float> cpuA, cpuX, cpuY0;
static const float cpuAlpha = 1.0f, cpuBeta = 0.0f;   // y = A x (the configuration)

static void resetVectors(int M, int N, bool needHost) {
    LTHIP_CHECK(lthip_fill_uniform_f32(gpuA.p, (size_t)M * N, -1.0, 1.0, cli::kSeed, 2, cli::stream()));
    LTHIP_CHECK(lthip_fill_uniform_f32(gpuX.p, (size_t)N, -1.0, 1.0, cli::kSeed, 3, cli::stream()));
    LTHIP_CHECK(lthip_fill_uniform_f32(gpuY.p, (size_t)M, -1.0, 1.0, cli::kSeed, 15, cli::stream()));
    HIP_CHECK(hipStreamSynchronize(cli::stream()));
    gpuAlpha.upload(&cpuAlpha);
    gpuBeta.upload(&cpuBeta);
    if (needHost) {
        cpuA.resize((size_t)M * N);
        cpuX.resize(N);
        cpuY0.resize(M);
        gpuA.download(cpuA.data());
        gpuX.download(cpuX.data());
        gpuY.download(cpuY0.data());
    }
}

// sgemv.cpp:132-147 (FP64 dot per row, then (float)(alpha*dot + beta*y)), and
// the componentwise criterion |y - ref| <= 1e-4 (|alpha| |A||x| + |beta||y|)_i
// for signed inputs.
static void compareToReference(int M, int N, bool rowMajor, int lda) {
    std::vector<float> got(M);
    gpuY.download(got.data());
    std::vector<int> bad(cli::host_cores() + 1, 0);
    int fails = 0;
    std::vector<int> counts((size_t)M, 0);
    cli::parallel_for((size_t)M, [&](size_t lo, size_t hi) {
        for (size_t row = lo; row < hi; row++) {
            double dot = 0.0, mag = 0.0;
            for (int col = 0; col < N; col++) {
                const double a = rowMajor ? cpuA[row * lda + col] : cpuA[(size_t)col * lda + row];
                dot += a * (double)cpuX[col];
                mag += std::fabs(a * (double)cpuX[col]);
            }
            const float ref = (float)((double)cpuAlpha * dot + (double)cpuBeta * (double)cpuY0[row]);
            const double tol = 1e-4 * (std::fabs((double)cpuAlpha) * mag + std::fabs((double)cpuBeta * cpuY0[row]));
            if (std::fabs((double)got[row] - ref) > tol) counts[row] = 1;
        }
    });
    for (int c : counts) fails += c;
    if (fails > 0) {
        printf("Number of different results: %d\n", fails);
        cli::g_check_failed = true;
    }
}

static float getGflops(int M, int N, float ms) { return (float)(2.0 * M * N / (ms * 1e6)); }   // sgemv.cpp:168-172

static float runner(SgemvImpl impl, std::vector<int> sizes, int nb_runs, int nb_warmup_runs, bool checkResult,
                    bool printEach, int argc, char* argv[]) {
    const int M = sizes[0], N = sizes[1];
    const bool rowMajor = strcmp(argv[argc - 1], "col") != 0;   // row-major unless the last token is "col"
    const int lda = rowMajor ? N : M;
    gpuA.alloc((size_t)M * N);
    gpuX.alloc(N);
    gpuY.alloc(M);
    gpuAlpha.alloc(1);
    gpuBeta.alloc(1);
    resetVectors(M, N, checkResult);
    const double bytes = 4.0 * ((double)M * N + N + M);
    float sum_gflops = 0.0f;
    for (int i = -nb_warmup_runs; i < nb_runs; i++) {
        const float ms = impl(M, N, gpuAlpha.p, gpuA.p, lda, gpuX.p, 1, gpuBeta.p, gpuY.p, 1, rowMajor);
        if (checkResult) compareToReference(M, N, rowMajor, lda);
        if (i >= 0) {
            const float g = getGflops(M, N, ms);
            if (printEach)
                printf("[internal][iter %d]: %f GFLOP/s | %.1f GB/s (%.1f%% of HBM)\n", i, g, bytes / (ms * 1e6),
                       100.0 * bytes / (ms * 1e6) / cli::kHbmPeakGBs);
            sum_gflops += g;
        }
    }
    gpuA.free();
    gpuX.free();
    gpuY.free();
    return sum_gflops / nb_runs;
}

static void exportCSV(const std::map<std::vector<int>, float>& results, const std::string& filename) {
    std::ofstream file(filename);
    file << "M,N,gflops\n";
    for (const auto& [sizes, gflops] : results) file << sizes[0] << "," << sizes[1] << "," << gflops << "\n";
}

}  // namespace sgemv

void sgemv_handler(RunMode runMode, int argc, char* argv[]) {
    switch (runMode) {
        case RunMode::TEST:
            Handlers<sgemv::SgemvImpl, 2, BENCH_RUNS, WARMUP_RUNS>::test_handler(argc, argv, sgemv::sgemvImpls,
                                                                                 sgemv::runner);
            break;
        case RunMode::BENCH:
            Handlers<sgemv::SgemvImpl, 2, BENCH_RUNS, WARMUP_RUNS>::bench_handler(
                argc, argv, sgemv::sgemvImpls, sgemv::runner, sgemv::SIZES, sgemv::exportCSV);
            break;
        default:
            fprintf(stderr, "Mode not supported by sgemv\n");
            exit(1);
    }
}
#undef BENCH_RUNS
#undef WARMUP_RUNS
