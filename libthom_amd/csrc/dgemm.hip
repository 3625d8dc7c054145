// Native dgemm on FP64 MFMA (v_mfma_f64_16x16x4f64).
//
// Block tile 128x128x16, four waves in a 2x2 grid, each wave a 64x64 sub-tile
// = 4x4 MFMA tiles (64 FP64 accumulators per lane).  The next K-slab is
// fetched into registers while the current one is multiplied out of LDS; two
// LDS buffers give one barrier per slab.  A is kept row-major in LDS with a
// 17-double row stride (MFMA A-fragment reads: 16 lanes, 16 rows, conflict
// free); B is kept K-major.  Tiles are visited in groups of 8 tile-rows and the
// block ids are remapped so each XCD works on a contiguous run of tiles.
//
// Shapes that are tile multiples (M, N % 128, K % 16, 16-byte aligned) take
// the pipelined kernel below instead: both operands reach LDS by DMA through
// a ring of 4 slabs, as in gemm_core2.hpp.
//
// Reference: dgemm_final.cu:65-213 (64x128x16 tiles, mma.m8n8k4, cp.async).
#include "gemm_core2.hpp"

namespace lt {
namespace {

constexpr int kBM = 128, kBN = 128, kBK = 16;
constexpr int kAStride = kBK + 1;
constexpr int kGroupM = 8;

__device__ __forceinline__ void tile_coords(int t, int tiles_m, int tiles_n, int& tm, int& tn) {
    const int per_group = kGroupM * tiles_n;
    const int group = t / per_group;
    const int first_m = group * kGroupM;
    int gsize = tiles_m - first_m;
    if (gsize > kGroupM) gsize = kGroupM;
    const int r = t % per_group;
    tm = first_m + r % gsize;
    tn = r / gsize;
}

// FAST: M, N, K multiples of the tile and 16-byte aligned rows; otherwise
// every element load/store is predicated.
template <bool FAST>
__global__ __launch_bounds__(256, 2) void dgemm_kernel(const double* __restrict__ A, const double* __restrict__ B,
                                                       double* __restrict__ C, int M, int N, int K, double alpha,
                                                       double beta, int tiles_m, int tiles_n) {
    __shared__ double As[2][kBM][kAStride];
    __shared__ double Bs[2][kBK][kBN];

    int tm, tn;
    tile_coords(xcd_remap(blockIdx.x, gridDim.x), tiles_m, tiles_n, tm, tn);
    const int row0 = tm * kBM, col0 = tn * kBN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;

    // A slab: 128 rows x 16 k; thread -> row ar + 32 i, k pair akp
    const int ar = tid >> 3, akp = (tid & 7) * 2;
    // B slab: 16 k x 128 cols; thread -> k bk, cols bc + 2 i
    const int bk = tid >> 4, bc = (tid & 15) * 8;

    double2 ra[4], rb[4];

    auto load_slab = [&](int k0) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int r = row0 + ar + 32 * i, k = k0 + akp;
            if (FAST) {
                ra[i] = *reinterpret_cast<const double2*>(A + (size_t)r * K + k);
            } else {
                ra[i].x = (r < M && k < K) ? A[(size_t)r * K + k] : 0.0;
                ra[i].y = (r < M && k + 1 < K) ? A[(size_t)r * K + k + 1] : 0.0;
            }
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int k = k0 + bk, c = col0 + bc + 2 * i;
            if (FAST) {
                rb[i] = *reinterpret_cast<const double2*>(B + (size_t)k * N + c);
            } else {
                rb[i].x = (k < K && c < N) ? B[(size_t)k * N + c] : 0.0;
                rb[i].y = (k < K && c + 1 < N) ? B[(size_t)k * N + c + 1] : 0.0;
            }
        }
    };
    auto store_slab = [&](int buf) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            As[buf][ar + 32 * i][akp] = ra[i].x;
            As[buf][ar + 32 * i][akp + 1] = ra[i].y;
        }
#pragma unroll
        for (int i = 0; i < 4; i++) *reinterpret_cast<double2*>(&Bs[buf][bk][bc + 2 * i]) = rb[i];
    };

    v4d acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

    const int kslabs = (K + kBK - 1) / kBK;
    if (kslabs > 0) {
        load_slab(0);
        store_slab(0);
    }
    __syncthreads();
    const int fr = lane & 15, fk = lane >> 4;
    for (int ks = 0; ks < kslabs; ks++) {
        const int buf = ks & 1;
        if (ks + 1 < kslabs) load_slab((ks + 1) * kBK);
#pragma unroll
        for (int kk = 0; kk < kBK; kk += 4) {
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = As[buf][wm + 16 * i + fr][kk + fk];
#pragma unroll
            for (int j = 0; j < 4; j++) b[j] = Bs[buf][kk + fk][wn + 16 * j + fr];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (ks + 1 < kslabs) store_slab(buf ^ 1);
        __syncthreads();
    }

    // C fragment of a 16x16 f64 tile: register r of lane l holds row 4*r + l/16,
    // column l%16 (measured; differs from the f32 16x16 layout)
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = row0 + wm + 16 * i + 4 * r + fk, col = col0 + wn + 16 * j + fr;
                if (FAST || (row < M && col < N)) {
                    double* cp = C + (size_t)row * N + col;
                    double v = alpha * acc[i][j][r];
                    if (beta != 0.0) v = fma(beta, *cp, v);
                    *cp = v;
                }
            }
}

// ---- pipelined kernel --------------------------------------------------------
// A tile [BM][16] doubles: 16-byte chunk c (k 2c, 2c+1) of row r at
// c ^ ((r >> 1) & 7); B tile [16][128] doubles: chunk c of row k at
// c ^ 8 (k & 1).  A fragment reads (16 rows, one k, 16 lanes) and B fragment
// reads (16 columns of two rows, 32 lanes) then hit 64 distinct banks.
// Eight waves: 2 x 4 of 64 x 32 each for the 128 x 128 tile (4 slabs in the
// ring), 4 x 2 of 64 x 64 each for 256 x 128 (3 slabs): twice the MFMAs per
// fragment read and per hand-over barrier.
template <int BM_, int STAGES_>
struct D2 {
    static constexpr int BM = BM_, BN = 128, STAGES = STAGES_, WAVES = 8;
    static constexpr int WGM = BM / 64, WGN = WAVES / WGM, WN = BN / WGN, TN = WN / 16;
    static constexpr int A_BYTES = BM * kBK * 8, B_BYTES = kBK * BN * 8, STAGE_BYTES = A_BYTES + B_BYTES;
    static constexpr int A_INSTR = A_BYTES / (1024 * WAVES), B_INSTR = B_BYTES / (1024 * WAVES);
    static constexpr int LOADS = A_INSTR + B_INSTR;
    static_assert(WGM * WGN == WAVES && TN >= 1 && A_INSTR >= 1 && B_INSTR >= 1, "wave grid");
    static_assert(STAGES * STAGE_BYTES <= 160 * 1024, "LDS budget");
    static_assert(LOADS * (STAGES - 2) < 64, "vmcnt range");
};

__device__ __forceinline__ int swzA64(int r) { return (r >> 1) & 7; }
__device__ __forceinline__ int swzB64(int k) { return 8 * (k & 1); }

// the tile as a device function: with the whole tile in the kernel template's
// body the host pass emitted no launch handle for its instances (an undefined
// symbol at link time), so the kernel below only forwards
template <class D>
__device__ __forceinline__ void dgemm2_tile(const double* __restrict__ A, const double* __restrict__ B,
                                            double* __restrict__ C, int M, int N, int K, double alpha, double beta,
                                            int tiles_m, int tiles_n) {
    constexpr int TN = D::TN, LOADS = D::LOADS, STAGES = D::STAGES;
    __shared__ __align__(16) char smem[STAGES * D::STAGE_BYTES];
    int tm, tn;
    tile_coords(xcd_remap(blockIdx.x, gridDim.x), tiles_m, tiles_n, tm, tn);
    const int row0 = tm * D::BM, col0 = tn * D::BN;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = (wave / D::WGN) * 64, wn = (wave % D::WGN) * D::WN;

    // per-lane DMA sources: A instruction i covers rows 8 inst .. +7 (8 chunks
    // each), B instruction i one row of 64 chunks
    const double* asrc[D::A_INSTR];
    const double* bsrc[D::B_INSTR];
#pragma unroll
    for (int i = 0; i < D::A_INSTR; i++) {
        const int inst = wave * D::A_INSTR + i;
        const int r = inst * 8 + lane / 8;
        const int c = (lane % 8) ^ swzA64(r);
        asrc[i] = A + (size_t)(row0 + r) * K + 2 * c;
    }
#pragma unroll
    for (int i = 0; i < D::B_INSTR; i++) {
        const int k = wave * D::B_INSTR + i;
        const int c = lane ^ swzB64(k);
        bsrc[i] = B + (size_t)k * N + col0 + 2 * c;
    }
    const size_t b_step = (size_t)kBK * N;
    auto issue = [&](auto q0c, auto q1c, int stage, int slab) {
        constexpr int Q0 = decltype(q0c)::value, Q1 = decltype(q1c)::value;
        char* sa = smem + stage * D::STAGE_BYTES;
        char* sb = sa + D::A_BYTES;
#pragma unroll
        for (int q = Q0; q < Q1; q++) {
            if (q < D::A_INSTR)
                __builtin_amdgcn_global_load_lds(asrc[q] + (size_t)slab * kBK,
                                                 (lds_void_t*)(lds_char_t*)(sa + (wave * D::A_INSTR + q) * 1024), 16,
                                                 0, 0);
            else
                __builtin_amdgcn_global_load_lds(
                    bsrc[q - D::A_INSTR] + (size_t)slab * b_step,
                    (lds_void_t*)(lds_char_t*)(sb + (wave * D::B_INSTR + q - D::A_INSTR) * 1024), 16, 0, 0);
        }
    };
    const int fr = lane & 15, fk = lane >> 4;
    struct Frag {
        double a[4], b[TN];
    };
    auto load_frag = [&](Frag& f, const char* sa, const char* sb, int kk) {
        const int k = 4 * kk + fk;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int r = wm + 16 * i + fr;
            f.a[i] = *reinterpret_cast<const double*>(sa + r * (kBK * 8) + (((k >> 1) ^ swzA64(r)) << 4) + (k & 1) * 8);
        }
#pragma unroll
        for (int j = 0; j < TN; j++) {
            const int c = wn + 16 * j + fr;
            f.b[j] = *reinterpret_cast<const double*>(sb + k * (D::BN * 8) + (((c >> 1) ^ swzB64(k)) << 4) + (c & 1) * 8);
        }
    };

    v4d acc[4][TN];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

    constexpr int KSTEPS = kBK / 4;
    const int nslabs = K / kBK;
    if (nslabs > 0) {
#pragma unroll
        for (int p = 0; p < STAGES - 1; p++) issue(int_c<0>{}, int_c<LOADS>{}, p, p < nslabs ? p : nslabs - 1);
        ring_barrier<LOADS * (STAGES - 2)>();
        Frag fr2[2];
        load_frag(fr2[0], smem, smem + D::A_BYTES, 0);
        int stage = 0, fill_stage = STAGES - 1;
        for (int s = 0; s < nslabs; s++) {
            const char* sa = smem + stage * D::STAGE_BYTES;
            const char* sb = sa + D::A_BYTES;
            const int fill = s + STAGES - 1 < nslabs ? s + STAGES - 1 : nslabs - 1;
            __builtin_amdgcn_sched_barrier(0);
            static_for<0, KSTEPS - 1>([&](auto kkc) {
                constexpr int kk = decltype(kkc)::value;
                load_frag(fr2[(kk + 1) & 1], sa, sb, kk + 1);
                issue(int_c<kk * LOADS / KSTEPS>{}, int_c<(kk + 1) * LOADS / KSTEPS>{}, fill_stage, fill);
                const Frag& f = fr2[kk & 1];
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < TN; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(f.a[i], f.b[j], acc[i][j], 0, 0, 0);
                interleave<4 * TN, 4 + TN, (kk + 1) * LOADS / KSTEPS - kk * LOADS / KSTEPS>();
                __builtin_amdgcn_sched_barrier(0);
            });
            {
                constexpr int KL = KSTEPS - 1;
                const Frag& f = fr2[KL & 1];
                issue(int_c<KL * LOADS / KSTEPS>{}, int_c<LOADS>{}, fill_stage, fill);
#pragma unroll
                for (int i = 0; i < 2; i++)
#pragma unroll
                    for (int j = 0; j < TN; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(f.a[i], f.b[j], acc[i][j], 0, 0, 0);
                interleave<2 * TN, 0, LOADS - KL * LOADS / KSTEPS>();
                __builtin_amdgcn_sched_barrier(0);
                ring_barrier<LOADS * (STAGES - 2)>();
                stage = stage + 1 == STAGES ? 0 : stage + 1;
                fill_stage = fill_stage + 1 == STAGES ? 0 : fill_stage + 1;
                const char* na = smem + stage * D::STAGE_BYTES;
                load_frag(fr2[0], na, na + D::A_BYTES, 0);   // KSTEPS is even
#pragma unroll
                for (int i = 2; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < TN; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(f.a[i], f.b[j], acc[i][j], 0, 0, 0);
                interleave<2 * TN, 4 + TN, 0>();
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        wait_vmcnt<0>();
    }

#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = row0 + wm + 16 * i + 4 * r + fk, col = col0 + wn + 16 * j + fr;
                double* cp = C + (size_t)row * N + col;
                double v = alpha * acc[i][j][r];
                if (beta != 0.0) v = fma(beta, *cp, v);
                *cp = v;
            }
}

template <class D>
__global__ __launch_bounds__(64 * D::WAVES) void dgemm2_kernel(const double* __restrict__ A,
                                                               const double* __restrict__ B, double* __restrict__ C,
                                                               int M, int N, int K, double alpha, double beta,
                                                               int tiles_m, int tiles_n) {
    dgemm2_tile<D>(A, B, C, M, N, K, alpha, beta, tiles_m, tiles_n);
}

using D2Small = D2<128, 4>;
using D2Big = D2<256, 3>;
typedef void (*dgemm2_fn)(const double*, const double*, double*, int, int, int, double, double, int, int);
const dgemm2_fn kDgemm2Small = dgemm2_kernel<D2Small>;
const dgemm2_fn kDgemm2Big = dgemm2_kernel<D2Big>;

}  // namespace
}  // namespace lt

extern "C" int lthip_dgemm_f64(const double* A, const double* B, double* C, size_t M, size_t N, size_t K,
                               double alpha, double beta, void* stream) {
    using namespace lt;
    if (M > (1u << 30) || N > (1u << 30) || K > (1u << 30)) return fail(LTHIP_ERR_INVALID_ARG, "dgemm: size too large");
    if (M == 0 || N == 0) return LTHIP_OK;
    if (!C || (K > 0 && (!A || !B))) return fail(LTHIP_ERR_INVALID_ARG, "dgemm: null pointer");
    const int m = (int)M, n = (int)N, k = (int)K;
    const int tiles_m = (m + kBM - 1) / kBM, tiles_n = (n + kBN - 1) / kBN;
    const long blocks = (long)tiles_m * tiles_n;
    if (blocks > 0x7fffffffL) return fail(LTHIP_ERR_INVALID_ARG, "dgemm: too many tiles");
    const bool fast = (m % kBM == 0) && (n % kBN == 0) && (k % kBK == 0) && (reinterpret_cast<uintptr_t>(A) % 16 == 0) &&
                      (reinterpret_cast<uintptr_t>(B) % 16 == 0);
    // the 256-row tile where it needs fewer rounds of the CUs: one of its
    // tiles takes twice as long as a 128-row tile (about 5 % less per flop),
    // so it wins when 2 x its rounds x 20 < the small tile's rounds x 21
    // (4096^2 and 8192^2; not 3072^2, where 288 tiles make 2 rounds, 32 in the second)
    const long cus = cu_count();
    const long big_tiles = (long)(m / D2Big::BM) * tiles_n;
    const bool big = fast && m % D2Big::BM == 0 &&
                     2 * ((big_tiles + cus - 1) / cus) * 20 < ((blocks + cus - 1) / cus) * 21;
    if (big)
        hipLaunchKernelGGL(kDgemm2Big, dim3((unsigned)((m / D2Big::BM) * tiles_n)), dim3(64 * D2Big::WAVES),
                           0, as_stream(stream), A, B, C, m, n, k, alpha, beta, m / D2Big::BM, tiles_n);
    else if (fast)
        hipLaunchKernelGGL(kDgemm2Small, dim3((unsigned)blocks), dim3(64 * D2Small::WAVES), 0,
                           as_stream(stream), A, B, C, m, n, k, alpha, beta, tiles_m, tiles_n);
    else
        hipLaunchKernelGGL((dgemm_kernel<false>), dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), A, B, C, m,
                           n, k, alpha, beta, tiles_m, tiles_n);
    LT_TRY_LAUNCH("dgemm launch");
    return LTHIP_OK;
}
