// Harness of the libthom CLI: the same test/bench handlers, usage strings and
// CSV naming as the reference (routines/routines.h:88-169), over the gfx950
// C ABI (include/lthip.h).  Differences on purpose:
//   - a result mismatch makes the process exit non-zero (the reference prints
//     and exits 0, routines/reduction/reduction.cpp:77-79);
//   - arguments are validated before any HIP call;
//   - inputs come from the seeded SplitMix64 generator (rng.h), not cuRAND.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <functional>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "enums.h"
#include "lthip.h"

#define LIB_HAS(name) LIB_HAS_##name

#define HIP_CHECK(call)                                                                                  \
    do {                                                                                                 \
        hipError_t e__ = (call);                                                                         \
        if (e__ != hipSuccess) {                                                                         \
            fprintf(stderr, "HIP error at %s:%d: %s\n", __FILE__, __LINE__, hipGetErrorString(e__));     \
            exit(EXIT_FAILURE);                                                                          \
        }                                                                                                \
    } while (0)

#define LTHIP_CHECK(call)                                                                                \
    do {                                                                                                 \
        int s__ = (call);                                                                                \
        if (s__ != LTHIP_OK) {                                                                           \
            fprintf(stderr, "lthip error at %s:%d: %s (%s)\n", __FILE__, __LINE__,                       \
                    lthip_status_string(s__), lthip_last_error());                                       \
            exit(EXIT_FAILURE);                                                                          \
        }                                                                                                \
    } while (0)

void reduction_handler(RunMode runMode, int argc, char* argv[]);
void sdot_handler(RunMode runMode, int argc, char* argv[]);
void sgemv_handler(RunMode runMode, int argc, char* argv[]);
void dgemm_handler(RunMode runMode, int argc, char* argv[]);
void hgemm_handler(RunMode runMode, int argc, char* argv[]);
void hgemm_batched_handler(RunMode runMode, int argc, char* argv[]);
void w4a16_handler(RunMode runMode, int argc, char* argv[]);
void w4a16_grouped_handler(RunMode runMode, int argc, char* argv[]);

namespace cli {

// Set by a runner when any check fails; main() turns it into the exit code.
extern bool g_check_failed;
// The stream every implementation runs on.
hipStream_t stream();
// Writes >= 512 MiB so the next kernel reads its inputs from HBM, not from
// the 256 MB Infinity Cache (the reference's flushL2 sizes itself from the L2
// attribute, common.cpp:12-24, which is far too small on MI355X).
void flush_caches();
// Host cores and model, printed with every CPU-timed number.
int host_cores();
std::string host_cpu_model();
// Published MI355X peaks used as roofline denominators.
constexpr double kHbmPeakGBs = 8000.0;
constexpr double kFp64MfmaTflops = 78.6;
constexpr double kFp16MfmaTflops = 2500.0;
constexpr double kFp8MfmaTflops = 5000.0;

// Times `fn` (which enqueues on stream()) with HIP events; milliseconds.
float time_ms(const std::function<void()>& fn);

// Device buffer owned by the harness (the reference's allocate/freeAll).
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    void alloc(size_t count) {
        free();
        n = count;
        if (count) HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * count));
    }
    void free() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { free(); }
    void upload(const T* h) { HIP_CHECK(hipMemcpy(p, h, sizeof(T) * n, hipMemcpyHostToDevice)); }
    void download(T* h) const { HIP_CHECK(hipMemcpy(h, p, sizeof(T) * n, hipMemcpyDeviceToHost)); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

// A zero-initialised workspace for one entry point, grown on demand.
void* workspace(const std::string& name, size_t bytes);
size_t workspace_size(const std::string& name);

// Parallel loop over [0, n) on cpu_threads() host threads.
void parallel_for(size_t n, const std::function<void(size_t, size_t)>& body);
// Threads parallel_for uses: all host cores, or the count a host
// implementation ("cpu": 1, "cpu_mt": all cores) set for its timed region.
int cpu_threads();
// Runs fn with parallel_for on `threads` threads and returns its wall time in
// milliseconds (std::chrono::steady_clock): the host implementations' timer.
float cpu_time_ms(int threads, const std::function<void()>& fn);
// Copies between device buffers and host vectors (outside any timed region).
template <typename T>
std::vector<T> to_host(const T* dev, size_t n) {
    std::vector<T> h(n);
    if (n) HIP_CHECK(hipMemcpy(h.data(), dev, sizeof(T) * n, hipMemcpyDeviceToHost));
    return h;
}
template <typename T>
void to_device(T* dev, const std::vector<T>& h) {
    if (!h.empty()) HIP_CHECK(hipMemcpy(dev, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
}
// The line printed before a host implementation runs: threads, host cores, CPU model.
void print_cpu_impl(const std::string& impl);

}  // namespace cli

/* Implementation will be ReductionImpl, SdotImpl, ...; sizeDims is the number
 * of sizes on the command line (reduction 1, sgemv 2, GEMMs 3). */
template <typename Implementation, int sizeDims, int BENCH_RUNS, int WARMUP_RUNS>
struct Handlers {
    using RunnerFunc = float (*)(Implementation, std::vector<int> sizes, int nb_runs, int nb_warmup_runs,
                                 bool checkResult, bool printEach, int argc, char* argv[]);

    static void list_impls(const std::unordered_map<std::string, Implementation>& impls) {
        fprintf(stderr, "Available implementations:");
        std::map<std::string, int> sorted;
        for (const auto& e : impls) sorted[e.first] = 0;
        for (const auto& e : sorted) fprintf(stderr, " %s", e.first.c_str());
        fprintf(stderr, "\n");
    }

    static void test_handler(int argc, char* argv[], const std::unordered_map<std::string, Implementation>& impls,
                             RunnerFunc runnerFunc) {
        /* ./libthom test <routine> <impl> <size...> <nb_iter> [check_result=1] */
        const int supposedArgc = 5 + sizeDims;
        if (argc < supposedArgc) {
            fprintf(stderr, "Missing arguments.\nUsage: ./libthom test <routine> <impl>");
            for (int i = 0; i < sizeDims; i++) fprintf(stderr, " <size%d>", i + 1);
            fprintf(stderr, " <nb_iter> <check_result=1>\n");
            list_impls(impls);
            exit(1);
        }
        auto it = impls.find(argv[3]);
        if (it == impls.end()) {
            fprintf(stderr, "Invalid implementation: %s\n", argv[3]);
            list_impls(impls);
            exit(1);
        }
        std::vector<int> sizes;
        for (int i = 0; i < sizeDims; i++) {
            const int v = atoi(argv[4 + i]);
            if (v < 0) {
                fprintf(stderr, "Sizes must be non-negative\n");
                exit(1);
            }
            sizes.push_back(v);
        }
        const int nb_iter = atoi(argv[4 + sizeDims]);
        if (nb_iter <= 0) {
            fprintf(stderr, "nb_iter must be positive\n");
            exit(1);
        }
        bool checkResult = true;
        if (argc >= supposedArgc + 1) {
            const std::string c = argv[4 + sizeDims + 1];
            if (c == "0" || c == "1") checkResult = c == "1";
        }
        cli::print_cpu_impl(argv[3]);
        const float gflops = runnerFunc(it->second, sizes, nb_iter, 0, checkResult, true, argc, argv);
        printf("# GFLOP/s: %f\n", gflops);
    }

    static void bench_handler(int argc, char* argv[], const std::unordered_map<std::string, Implementation>& impls,
                              RunnerFunc runnerFunc, const std::vector<std::vector<int>>& SIZES,
                              void (*exportFunc)(const std::map<std::vector<int>, float>&, const std::string&)) {
        /* ./libthom bench <routine> <impl> */
        if (argc < 4) {
            fprintf(stderr, "Missing arguments.\nUsage: ./libthom bench <routine> <impl>\n");
            list_impls(impls);
            exit(1);
        }
        auto it = impls.find(argv[3]);
        if (it == impls.end()) {
            fprintf(stderr, "Invalid implementation: %s\n", argv[3]);
            list_impls(impls);
            exit(1);
        }
        std::map<std::vector<int>, float> results;
        printf("-- Running %s -- %s\n", argv[2], argv[3]);
        cli::print_cpu_impl(argv[3]);
        for (const std::vector<int>& sizes : SIZES) {
            printf("\nStarting benchmark for size %d...\n", sizes[0]);
            const float gflops = runnerFunc(it->second, sizes, BENCH_RUNS, WARMUP_RUNS, false, true, argc, argv);
            results[sizes] = gflops;
            printf("[size: %d] ==> %f\n", sizes[0], gflops);
        }
        const std::string filename = std::string("bench_") + argv[2] + "_" + argv[3] + ".csv";
        exportFunc(results, filename);
    }
};
