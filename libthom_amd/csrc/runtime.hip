// Status reporting, device queries and the workspace-clear kernel.
#include "lt_common.hpp"

#include <cstdarg>
#include <cstdio>
#include <mutex>

namespace lt {

static thread_local char g_err[512] = "";

int fail(int status, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return status;
}

int check_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return LTHIP_OK;
    return fail(LTHIP_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

int cu_count() {
    static std::mutex mu;
    static int cache[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    std::lock_guard<std::mutex> lock(mu);
    if (cache[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;
        cache[dev] = n;
    }
    return cache[dev];
}

__global__ void clear_kernel(uint4* __restrict__ p16, size_t n16, uint8_t* __restrict__ tail, size_t ntail) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t j = i; j < n16; j += stride) p16[j] = make_uint4(0, 0, 0, 0);
    if (i < ntail) tail[i] = 0;
}

}  // namespace lt

extern "C" {

const char* lthip_status_string(int status) {
    switch (status) {
        case LTHIP_OK: return "ok";
        case LTHIP_ERR_INVALID_ARG: return "invalid argument";
        case LTHIP_ERR_UNSUPPORTED: return "unsupported";
        case LTHIP_ERR_WORKSPACE: return "workspace missing or too small";
        case LTHIP_ERR_HIP: return "HIP runtime error";
        default: return "unknown status";
    }
}

const char* lthip_last_error(void) { return lt::g_err; }

const char* lthip_version(void) { return "0.1.0 gfx950"; }

int lthip_workspace_clear(void* ws, size_t bytes, void* stream) {
    if (bytes == 0) return LTHIP_OK;
    if (!ws) return lt::fail(LTHIP_ERR_INVALID_ARG, "workspace_clear: null pointer");
    uintptr_t addr = reinterpret_cast<uintptr_t>(ws);
    if (addr % 16 != 0) return lt::fail(LTHIP_ERR_INVALID_ARG, "workspace_clear: pointer not 16-byte aligned");
    size_t n16 = bytes / 16, ntail = bytes % 16;
    size_t blocks = (n16 + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(lt::clear_kernel, dim3((unsigned)blocks), dim3(256), 0, lt::as_stream(stream),
                       reinterpret_cast<uint4*>(ws), n16, reinterpret_cast<uint8_t*>(ws) + n16 * 16, ntail);
    LT_TRY_LAUNCH("workspace_clear launch");
    return LTHIP_OK;
}

}  // extern "C"
