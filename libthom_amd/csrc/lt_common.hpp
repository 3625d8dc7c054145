// Shared device helpers and host plumbing for the gfx950 kernels behind lthip.h.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstddef>
#include <cstdint>

#include "lthip.h"

namespace lt {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kWave = 64;

// Split-K workspaces: a fixed head of arrival counters (zero between calls),
// then the FP32 partials.  The head never holds partials, so a later call with
// more tiles still finds zeroed counters.
constexpr size_t kCounterBytes = 65536;
constexpr size_t kMaxCounters = kCounterBytes / sizeof(unsigned);

// ---- host side --------------------------------------------------------
// Records a message for lthip_last_error() and returns the status.
int fail(int status, const char* fmt, ...);
// Maps a HIP error to LTHIP_ERR_HIP (recording the HIP message) or LTHIP_OK.
int check_hip(hipError_t e, const char* what);
// Compute units of the current device (cached per device id).
int cu_count();

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

#define LT_TRY_LAUNCH(what)                                  \
    do {                                                     \
        int st__ = ::lt::check_hip(hipGetLastError(), what); \
        if (st__ != LTHIP_OK) return st__;                   \
    } while (0)

// ---- device side ------------------------------------------------------
// Streaming (non-temporal) 16-byte loads/stores: the data is used once.
__device__ __forceinline__ float4 ld_nt(const float4* p) {
    const v4f v = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(p));
    return make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ uint4 ld_nt(const uint4* p) {
    const v4i v = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(p));
    return make_uint4((unsigned)v[0], (unsigned)v[1], (unsigned)v[2], (unsigned)v[3]);
}
__device__ __forceinline__ void st_nt(float4* p, float4 a) {
    const v4f v = {a.x, a.y, a.z, a.w};
    __builtin_nontemporal_store(v, reinterpret_cast<v4f*>(p));
}

// ---- handing data from one block to another -----------------------------
// Each XCD of an MI355X has its own L2, and the L2s are not coherent with one
// another, so an agent-scope release/acquire fence compiles to a write-back /
// invalidate of the whole L2 (buffer_wbl2 / buffer_inv sc1) -- per block, that
// is microseconds.  Cross-block data instead moves only through agent-scope
// atomics, which are coherent across XCDs on their own:
//   publish  = relaxed agent-scope atomic store (global_store ... sc1: written
//              through past the XCD's L2),
//   drain    = s_waitcnt vmcnt(0) lgkmcnt(0) with a memory clobber: the
//              hardware waits until this thread's stores are acknowledged
//              (gfx9 counts stores in vmcnt), and the compiler may not move
//              any memory access across it,
//   ticket   = relaxed agent-scope atomic add, issued after the drain (and,
//              where several threads publish, after a barrier behind all
//              their drains),
//   collect  = relaxed agent-scope atomic load (sc1), issued only by the block
//              that received the last ticket, after the ticket value has come
//              back and been passed through LDS and a barrier.
// No data crosses blocks outside these atomics, so no release/acquire is
// needed to order other memory traffic; the only orders required are
// publish-before-ticket (drain) and ticket-before-collect (the control
// dependency and the barrier).
__device__ __forceinline__ void publish(float* p, float v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float collect(const float* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// all of this thread's memory operations have completed; a compiler barrier too
__device__ __forceinline__ void drain() { asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory"); }
// Workgroup barrier that orders LDS accesses only: global stores issued before
// it may still be in flight afterwards (unlike __syncthreads(), whose release
// fence waits for them).  For handing LDS words between threads.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }
__device__ __forceinline__ unsigned take_ticket(unsigned* counter) {
    return __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void reset_counter(unsigned* counter) {
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// DPP move within rows of 16 lanes; every control used here reads a valid lane.
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

__device__ __forceinline__ float readlane_f32(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Sum over the 64 lanes of a wave, returned to every lane.  Rows of 16 are
// reduced with DPP (quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror,
// row_mirror), then the four row sums are combined in a fixed order, so the
// result does not depend on scheduling.
__device__ __forceinline__ float wave_sum(float v) {
    v += dpp_f32<0xB1>(v);
    v += dpp_f32<0x4E>(v);
    v += dpp_f32<0x141>(v);
    v += dpp_f32<0x140>(v);
    return (readlane_f32(v, 0) + readlane_f32(v, 16)) + (readlane_f32(v, 32) + readlane_f32(v, 48));
}

// Block sum for BLOCK threads (a multiple of 64); valid in every thread.
// `scratch` must hold BLOCK/64 floats of shared memory.
template <int BLOCK>
__device__ __forceinline__ float block_sum(float v, float* scratch) {
    constexpr int W = BLOCK / kWave;
    v = wave_sum(v);
    if (W == 1) return v;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    float s = 0.0f;
#pragma unroll
    for (int w = 0; w < W; w++) s += scratch[w];
    __syncthreads();
    return s;
}

// Remap a linear block id so that the blocks the hardware sends to one XCD
// (ids congruent mod 8) cover a contiguous range of logical tiles: neighbouring
// tiles then share operand panels in the same L2.  Ids past the last full
// group of 8 are left unchanged.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
    constexpr int kXcd = 8;
    const int full = (nblocks / kXcd) * kXcd;
    if (bid >= full) return bid;
    const int per = full / kXcd;
    return (bid % kXcd) * per + bid / kXcd;
}

}  // namespace lt
