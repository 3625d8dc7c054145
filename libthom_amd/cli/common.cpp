// Shared helpers of the CLI harness (the reference's routines/common.cpp).
#include <chrono>
#include <fstream>
#include <thread>

#include "routines.h"

namespace cli {

bool g_check_failed = false;

hipStream_t stream() {
    static hipStream_t s = nullptr;
    if (!s) HIP_CHECK(hipStreamCreate(&s));   // default priority
    return s;
}

void flush_caches() {
    static float* buf = nullptr;
    static const size_t n = (size_t)512 * 1024 * 1024 / sizeof(float);
    if (!buf) {
        HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&buf), 2 * n * sizeof(float)));
        LTHIP_CHECK(lthip_fill_uniform_f32(buf, 2 * n, 0.0, 1.0, 1, 1, stream()));
    }
    LTHIP_CHECK(lthip_copy_f32(buf + n, buf, n, stream()));
}

int host_cores() {
    const unsigned n = std::thread::hardware_concurrency();
    return n ? (int)n : 1;
}

std::string host_cpu_model() {
    std::ifstream f("/proc/cpuinfo");
    std::string line;
    while (std::getline(f, line))
        if (line.rfind("model name", 0) == 0) {
            const size_t p = line.find(':');
            return p == std::string::npos ? line : line.substr(p + 2);
        }
    return "unknown";
}

float time_ms(const std::function<void()>& fn) {
    hipEvent_t a, b;
    HIP_CHECK(hipEventCreate(&a));
    HIP_CHECK(hipEventCreate(&b));
    HIP_CHECK(hipEventRecord(a, stream()));
    fn();
    HIP_CHECK(hipEventRecord(b, stream()));
    HIP_CHECK(hipEventSynchronize(b));
    float ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    HIP_CHECK(hipEventDestroy(a));
    HIP_CHECK(hipEventDestroy(b));
    return ms;
}

namespace {
struct Ws {
    void* p = nullptr;
    size_t bytes = 0;
};
std::map<std::string, Ws>& ws_map() {
    static std::map<std::string, Ws> m;
    return m;
}
}  // namespace

void* workspace(const std::string& name, size_t bytes) {
    Ws& w = ws_map()[name];
    if (bytes == 0) bytes = 256;
    if (w.bytes < bytes) {
        if (w.p) HIP_CHECK(hipFree(w.p));
        HIP_CHECK(hipMalloc(&w.p, bytes));
        w.bytes = bytes;
        LTHIP_CHECK(lthip_workspace_clear(w.p, bytes, stream()));
        HIP_CHECK(hipStreamSynchronize(stream()));
    }
    return w.p;
}

size_t workspace_size(const std::string& name) { return ws_map()[name].bytes; }

static int g_cpu_threads = 0;   // 0: all host cores

int cpu_threads() { return g_cpu_threads > 0 ? g_cpu_threads : host_cores(); }

float cpu_time_ms(int threads, const std::function<void()>& fn) {
    const int saved = g_cpu_threads;
    g_cpu_threads = threads;
    const auto t0 = std::chrono::steady_clock::now();
    fn();
    const auto t1 = std::chrono::steady_clock::now();
    g_cpu_threads = saved;
    return std::chrono::duration<float, std::milli>(t1 - t0).count();
}

void print_cpu_impl(const std::string& impl) {
    if (impl == "cpu")
        printf("# host implementation: 1 thread (host cores: %d, CPU: %s)\n", host_cores(), host_cpu_model().c_str());
    else if (impl == "cpu_mt")
        printf("# host implementation: %d threads (host cores: %d, CPU: %s)\n", host_cores(), host_cores(),
               host_cpu_model().c_str());
}

void parallel_for(size_t n, const std::function<void(size_t, size_t)>& body) {
    const size_t t = std::min<size_t>((size_t)cpu_threads(), n ? n : 1);
    if (t <= 1) {
        body(0, n);
        return;
    }
    std::vector<std::thread> th;
    for (size_t i = 0; i < t; i++) th.emplace_back(body, n * i / t, n * (i + 1) / t);
    for (auto& x : th) x.join();
}

}  // namespace cli
