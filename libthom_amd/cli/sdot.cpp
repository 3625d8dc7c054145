// sdot routine driver (the reference's routines/sdot/sdot.cpp): FP32 dot
// product, checked against an FP64 CPU dot product.
#include <cmath>
#include <fstream>
#include <mutex>

#include "host_math.h"
#include "routines.h"

namespace sdot {

#define BENCH_RUNS 10
#define WARMUP_RUNS 5
// sdot.cpp:19
static const std::vector<std::vector<int>> SIZES = {{10000},      {100'000},     {1'000'000},   {10'000'000},
                                                    {100'000'000}, {200'000'000}, {300'000'000}, {500'000'000}};

using SdotImpl = float (*)(int, const float*, const float*, float*);   // sdot.cpp:32

static float sdot_hip(int n, const float* x, const float* y, float* result) {
    const size_t need = lthip_sdot_f32_workspace_size((size_t)n);
    void* ws = cli::workspace("sdot", need);
    return cli::time_ms([&] { LTHIP_CHECK(lthip_sdot_f32((size_t)n, x, y, result, ws, need, cli::stream())); });
}

// Host implementations: the reference's CPU check (sdot.cpp:96-102, an FP64
// dot) on one thread ("cpu") or all host cores ("cpu_mt"), timed with
// std::chrono; copies to and from the device are outside the timed region.
static float sdot_host(int n, const float* x, const float* y, float* result, int threads) {
    const std::vector<float> hx = cli::to_host(x, (size_t)n), hy = cli::to_host(y, (size_t)n);
    double total = 0.0;
    const float ms = cli::cpu_time_ms(threads, [&] {
        std::mutex mu;
        cli::parallel_for((size_t)n, [&](size_t lo, size_t hi) {
            double s = 0.0;
            for (size_t i = lo; i < hi; i++) s += (double)hx[i] * (double)hy[i];
            std::lock_guard<std::mutex> g(mu);
            total += s;
        });
    });
    cli::to_device(result, std::vector<float>{(float)total});
    return ms;
}
static float sdot_cpu(int n, const float* x, const float* y, float* r) { return sdot_host(n, x, y, r, 1); }
static float sdot_cpu_mt(int n, const float* x, const float* y, float* r) {
    return sdot_host(n, x, y, r, cli::host_cores());
}

static const std::unordered_map<std::string, SdotImpl> sdotImpls = {
#if LIB_HAS(SDOT_HIP)
    {"hip", sdot_hip},
#endif
    {"cpu", sdot_cpu},
    {"cpu_mt", sdot_cpu_mt},
};

static cli::DevBuf<float> gpuX, gpuY, gpuOut;
static std::vector<float> cpuX, cpuY;

static void resetVectors(int N, bool needHost) {
    // U[-1,1): signed inputs, so the check below scales by sum |x y|
    LTHIP_CHECK(lthip_fill_uniform_f32(gpuX.p, (size_t)N, -1.0, 1.0, cli::kSeed, 13, cli::stream()));
    LTHIP_CHECK(lthip_fill_uniform_f32(gpuY.p, (size_t)N, -1.0, 1.0, cli::kSeed, 14, cli::stream()));
    HIP_CHECK(hipStreamSynchronize(cli::stream()));
    if (needHost) {
        cpuX.resize(N);
        cpuY.resize(N);
        gpuX.download(cpuX.data());
        gpuY.download(cpuY.data());
    }
}

static void compareToReference(int N) {
    float got = 0.0f;
    HIP_CHECK(hipMemcpy(&got, gpuOut.p, sizeof(float), hipMemcpyDeviceToHost));
    std::mutex mu;
    double ref = 0.0, mag = 0.0;   // sdot.cpp:96-102, plus sum |x y| for the signed-input criterion
    cli::parallel_for((size_t)N, [&](size_t lo, size_t hi) {
        double s = 0.0, a = 0.0;
        for (size_t i = lo; i < hi; i++) {
            const double p = (double)cpuX[i] * (double)cpuY[i];
            s += p;
            a += std::fabs(p);
        }
        std::lock_guard<std::mutex> g(mu);
        ref += s;
        mag += a;
    });
    if (std::fabs(got - ref) > 1e-5 * mag) {
        printf("ERROR: |gpu - cpu| = %g > 1e-5 * sum|xy| (%g), gpu=%f, cpu=%f\n", std::fabs(got - ref), mag, got, ref);
        cli::g_check_failed = true;
    }
}

static float getGflops(int N, float ms) { return (float)(2.0 * N / (ms * 1e6)); }   // sdot.cpp:116-120

static float runner(SdotImpl impl, std::vector<int> sizes, int nb_runs, int nb_warmup_runs, bool checkResult,
                    bool printEach, int, char**) {
    const int N = sizes[0];
    gpuX.alloc(N);
    gpuY.alloc(N);
    gpuOut.alloc(1);
    resetVectors(N, checkResult);
    float sum_gflops = 0.0f;
    for (int i = -nb_warmup_runs; i < nb_runs; i++) {
        cli::flush_caches();
        const float ms = impl(N, gpuX.p, gpuY.p, gpuOut.p);
        if (checkResult) compareToReference(N);
        if (i >= 0) {
            const float g = getGflops(N, ms);
            if (printEach)
                printf("[internal][iter %d]: %f GFLOP/s | %.2f GB/s (%.1f%% of HBM)\n", i, g, 8.0 * N / (ms * 1e6),
                       100.0 * 8.0 * N / (ms * 1e6) / cli::kHbmPeakGBs);
            sum_gflops += g;
        }
    }
    gpuX.free();
    gpuY.free();
    gpuOut.free();
    return sum_gflops / nb_runs;
}

static void exportCSV(const std::map<std::vector<int>, float>& results, const std::string& filename) {
    std::ofstream file(filename);
    file << "size,gflops\n";
    for (const auto& [sizes, gflops] : results) file << sizes[0] << "," << gflops << "\n";
}

}  // namespace sdot

void sdot_handler(RunMode runMode, int argc, char* argv[]) {
    switch (runMode) {
        case RunMode::TEST:
            Handlers<sdot::SdotImpl, 1, BENCH_RUNS, WARMUP_RUNS>::test_handler(argc, argv, sdot::sdotImpls, sdot::runner);
            break;
        case RunMode::BENCH:
            Handlers<sdot::SdotImpl, 1, BENCH_RUNS, WARMUP_RUNS>::bench_handler(argc, argv, sdot::sdotImpls, sdot::runner,
                                                                                sdot::SIZES, sdot::exportCSV);
            break;
        default:
            fprintf(stderr, "Mode not supported by sdot\n");
            exit(1);
    }
}
#undef BENCH_RUNS
#undef WARMUP_RUNS
