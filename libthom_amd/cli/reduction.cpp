// reduction routine driver (the reference's routines/reduction/reduction.cpp):
// FP32 sum, checked against an FP64 CPU sum.
#include <cmath>
#include <fstream>
#include <mutex>

#include "host_math.h"
#include "routines.h"

namespace reduction {

#define BENCH_RUNS 10
#define WARMUP_RUNS 5
// reduction.cpp:9
static const std::vector<std::vector<int>> SIZES = {{10000},      {100'000},     {1'000'000},   {10'000'000},
                                                    {100'000'000}, {200'000'000}, {300'000'000}, {500'000'000}};

using ReductionImpl = float (*)(const float*, float*, int);   // reduction.cpp:20

static float reduction_hip(const float* in, float* out, int N) {
    const size_t need = lthip_reduce_sum_f32_workspace_size((size_t)N);
    void* ws = cli::workspace("reduction", need);
    return cli::time_ms([&] { LTHIP_CHECK(lthip_reduce_sum_f32(in, (size_t)N, out, ws, need, cli::stream())); });
}

// Host implementations: the reference's CPU check (reduction.cpp:61-67, an
// FP64 sum) on one thread ("cpu") or all host cores ("cpu_mt"), timed with
// std::chrono; the input is copied to the host and the result back outside
// the timed region.
static float reduction_host(const float* in, float* out, int N, int threads) {
    const std::vector<float> x = cli::to_host(in, (size_t)N);
    double total = 0.0;
    const float ms = cli::cpu_time_ms(threads, [&] {
        std::mutex mu;
        cli::parallel_for((size_t)N, [&](size_t lo, size_t hi) {
            double s = 0.0;
            for (size_t i = lo; i < hi; i++) s += (double)x[i];
            std::lock_guard<std::mutex> g(mu);
            total += s;
        });
    });
    cli::to_device(out, std::vector<float>{(float)total});
    return ms;
}
static float reduction_cpu(const float* in, float* out, int N) { return reduction_host(in, out, N, 1); }
static float reduction_cpu_mt(const float* in, float* out, int N) {
    return reduction_host(in, out, N, cli::host_cores());
}

static const std::unordered_map<std::string, ReductionImpl> reductionImpls = {
#if LIB_HAS(REDUCTION_HIP)
    {"hip", reduction_hip},
#endif
    {"cpu", reduction_cpu},
    {"cpu_mt", reduction_cpu_mt},
};

static cli::DevBuf<float> gpuX, gpuOut;
static std::vector<float> cpuX;

static void resetVectors(int N, bool needHost) {
    // U[0,1) from the seeded generator (stream 1), bit-identical to the host spec
    LTHIP_CHECK(lthip_fill_uniform_f32(gpuX.p, (size_t)N, 0.0, 1.0, cli::kSeed, 1, cli::stream()));
    HIP_CHECK(hipStreamSynchronize(cli::stream()));
    if (needHost) {
        cpuX.resize(N);
        gpuX.download(cpuX.data());
    }
}

// reduction.cpp:61-67, the FP64 sum (split over host threads; FP64 order
// effects are far below the FP32 tolerance)
static double getReferenceResults(int N) {
    std::mutex mu;
    double total = 0.0;
    cli::parallel_for((size_t)N, [&](size_t lo, size_t hi) {
        double s = 0.0;
        for (size_t i = lo; i < hi; i++) s += (double)cpuX[i];
        std::lock_guard<std::mutex> g(mu);
        total += s;
    });
    return total;
}

static void compareToReference(int N) {
    float got = 0.0f;
    HIP_CHECK(hipMemcpy(&got, gpuOut.p, sizeof(float), hipMemcpyDeviceToHost));
    const double ref = getReferenceResults(N);
    const double rel = ref != 0.0 ? std::fabs(got - ref) / std::fabs(ref) : std::fabs((double)got);
    if (rel > 1e-5) {   // the configuration's tolerance (the reference uses 1e-4)
        printf("ERROR: relativeError = %g, gpu=%f, cpu=%f\n", rel, got, ref);
        cli::g_check_failed = true;
    }
}

static float getGflops(int N, float ms) { return (float)((N - 1) / (ms * 1e6)); }   // reduction.cpp:82-86

static float runner(ReductionImpl impl, std::vector<int> sizes, int nb_runs, int nb_warmup_runs, bool checkResult,
                    bool printEach, int, char**) {
    const int N = sizes[0];
    gpuX.alloc(N);
    gpuOut.alloc(1);
    resetVectors(N, checkResult);
    float sum_gflops = 0.0f;
    for (int i = -nb_warmup_runs; i < nb_runs; i++) {
        cli::flush_caches();
        const float ms = impl(gpuX.p, gpuOut.p, N);
        if (checkResult) compareToReference(N);
        if (i >= 0) {
            const float g = getGflops(N, ms);
            if (printEach)
                printf("[internal][iter %d]: %f GFLOP/s | %.2f GB/s (%.1f%% of HBM)\n", i, g, 4.0 * N / (ms * 1e6),
                       100.0 * 4.0 * N / (ms * 1e6) / cli::kHbmPeakGBs);
            sum_gflops += g;
        }
    }
    gpuX.free();
    gpuOut.free();
    return sum_gflops / nb_runs;
}

static void exportCSV(const std::map<std::vector<int>, float>& results, const std::string& filename) {
    std::ofstream file(filename);
    file << "size,gflops\n";
    for (const auto& [sizes, gflops] : results) file << sizes[0] << "," << gflops << "\n";
}

}  // namespace reduction

void reduction_handler(RunMode runMode, int argc, char* argv[]) {
    switch (runMode) {
        case RunMode::TEST:
            Handlers<reduction::ReductionImpl, 1, BENCH_RUNS, WARMUP_RUNS>::test_handler(argc, argv, reduction::reductionImpls,
                                                                                        reduction::runner);
            break;
        case RunMode::BENCH:
            Handlers<reduction::ReductionImpl, 1, BENCH_RUNS, WARMUP_RUNS>::bench_handler(
                argc, argv, reduction::reductionImpls, reduction::runner, reduction::SIZES, reduction::exportCSV);
            break;
        default:
            fprintf(stderr, "Mode not supported by reduction\n");
            exit(1);
    }
}
#undef BENCH_RUNS
#undef WARMUP_RUNS
