// Split-precision (Ozaki-style) dgemm on FP8 MFMA, plus the compensated
// FP64 GEMM used as the accuracy reference.
//
// Split: every row of A (and every column of B, i.e. row of B^T) gets an
// exponent e with max|x| = f 2^e, f in [0.5, 1).  t = x 2^-e is cut into
// signed digits d_1 = rint(16 t), d_p = rint(32 r_{p-1}) (r = remainder),
// all in [-16, 16] and exact in FP8 E4M3.  Digits are stored as E4M3 bytes,
// [slice][row][k] with k contiguous.
//
// Product: for a 64x64 C tile, each 128-k chunk of all S slices of A and B is
// staged in LDS (16-byte chunks XOR-swizzled by row), and for every slice pair
// (p, q) with p + q <= S - 1 (zero-based) one v_mfma_scale_f32_16x16x128_f8f6f4
// accumulates into the FP32 accumulator of level p + q.  Products are at most
// 256 and at most S pairs share a level, so while S * 256 * K < 2^24 every
// level sum is an exact integer.  The levels are recombined in FP64 by Horner
// from the last level, scaled by 2^(e_i + f_j - 8), then alpha/beta are
// applied -- exactly the order of the oracle's orc_ozaki_dgemm.
//
// Level GEMMs (M % 256 == N % 128 == K % 256 == 0): level l is one FP8 GEMM
// over K_eff = (l + 1) K, the slices concatenated along k (A_0..A_l against
// B_l..B_0), on a 256x128 tile fed by a DMA ring (gemm_core2.hpp's scheme);
// its FP32 sums (the same exact integers) go to the workspace and a combine
// kernel applies the same FP64 Horner recombination.  Other shapes use the
// all-levels-in-registers kernel below.
//
// Non-finite inputs: a row of A (column of B) holding a NaN or an infinity has
// no exponent to scale by.  The split marks it (exponent kOzNonFinite, all
// digits 0), the combine steps leave its entries of C alone, and a fixup
// kernel computes exactly those entries as a plain FP64 dot product, k
// ascending, then alpha and beta as in the combine: NaN and infinities
// propagate as in an FP64 product.  Other entries are unaffected.
//
// This routine has no counterpart in the reference; it sits beside the
// native FP64 path (dgemm_final.cu) as the emulated alternative.
#include "gemm_core2.hpp"

namespace lt {
namespace {

constexpr int kOzTile = 64;
constexpr int kOzBK = 128;
constexpr int32_t kOzNonFinite = INT32_MIN;   // split exponent of a row with a NaN or an infinity

__device__ __forceinline__ uint32_t e4m3_of_digit(int d) {
    if (d == 0) return 0u;
    const uint32_t s = d < 0 ? 0x80u : 0u;
    const int a = d < 0 ? -d : d;   // 1..16
    const int e = 31 - __clz(a);    // floor(log2 a)
    const int m = e >= 3 ? (a >> (e - 3)) & 7 : (a << (3 - e)) & 7;
    return s | (uint32_t)((e + 7) << 3) | (uint32_t)m;
}

// one block per row of X (row-major, row stride ld); cols % 4 == 0
template <int S>
__global__ __launch_bounds__(256) void ozaki_split_kernel(const double* __restrict__ X, int cols, long ld,
                                                          uint8_t* __restrict__ digits, long slice_stride,
                                                          int32_t* __restrict__ exps) {
    __shared__ double red[4];
    __shared__ int nonfinite[4];
    const int row = blockIdx.x;
    const double* xr = X + (long)row * ld;
    double mx = 0.0;
    int nf = 0;
    for (int c = threadIdx.x; c < cols; c += 256) {
        const double v = xr[c];
        nf |= !isfinite(v);
        mx = fmax(mx, fabs(v));
    }
    // block max (exact: max is order independent)
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
    const bool wave_nf = __any(nf);
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = mx;
        nonfinite[threadIdx.x >> 6] = wave_nf;
    }
    __syncthreads();
    mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    const bool bad = nonfinite[0] | nonfinite[1] | nonfinite[2] | nonfinite[3];
    int e = 0;
    if (!bad && mx > 0.0) frexp(mx, &e);
    if (threadIdx.x == 0) exps[row] = bad ? kOzNonFinite : e;
    uint8_t* drow = digits + (long)row * cols;
    for (int c4 = threadIdx.x * 4; c4 < cols; c4 += 1024) {
        double t[4];
#pragma unroll
        for (int u = 0; u < 4; u++) t[u] = bad ? 0.0 : ldexp(xr[c4 + u], -e);
#pragma unroll
        for (int p = 0; p < S; p++) {
            uint32_t word = 0;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const double v = ldexp(t[u], p == 0 ? 4 : 5);
                const double d = rint(v);
                t[u] = v - d;   // exact
                word |= e4m3_of_digit((int)d) << (8 * u);
            }
            *reinterpret_cast<uint32_t*>(drow + (long)p * slice_stride + c4) = word;
        }
    }
}

// Entries of C whose row of A or column of B (row of Bt) holds a NaN or an
// infinity: alpha * sum_k A[i][k] Bt[j][k], k ascending, (+ beta C).  One
// 64 x 64 tile of C per block; a tile with no such row or column returns at
// once.
__global__ __launch_bounds__(256) void ozaki_nonfinite_kernel(const double* __restrict__ A, const double* __restrict__ Bt,
                                                              const int32_t* __restrict__ ea,
                                                              const int32_t* __restrict__ eb, double* __restrict__ C,
                                                              int M, int N, int K, double alpha, double beta) {
    __shared__ int flag_r[64], flag_c[64];
    __shared__ int any;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    if (threadIdx.x < 64) {
        const int r = m0 + threadIdx.x;
        flag_r[threadIdx.x] = r < M && ea[r] == kOzNonFinite;
        if (flag_r[threadIdx.x]) any = 1;
    } else if (threadIdx.x < 128) {
        const int c = n0 + threadIdx.x - 64;
        flag_c[threadIdx.x - 64] = c < N && eb[c] == kOzNonFinite;
        if (flag_c[threadIdx.x - 64]) any = 1;
    }
    __syncthreads();
    if (!any) return;
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int i = e / 64, j = e % 64, row = m0 + i, col = n0 + j;
        if (row >= M || col >= N || !(flag_r[i] | flag_c[j])) continue;
        const double* a = A + (long)row * K;
        const double* b = Bt + (long)col * K;
        double dot = 0.0;
        for (int k = 0; k < K; k++) dot = __dadd_rn(dot, __dmul_rn(a[k], b[k]));
        double v = __dmul_rn(alpha, dot);
        double* cp = C + (long)row * N + col;
        if (beta != 0.0) v = __dadd_rn(v, __dmul_rn(beta, *cp));
        *cp = v;
    }
}

// out[c][r] = in[r][c]; 64x64 tiles through LDS
__global__ __launch_bounds__(256) void transpose_f64_kernel(const double* __restrict__ in, double* __restrict__ out,
                                                            int rows, int cols) {
    __shared__ double tile[64][65];
    const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int i = ty; i < 64; i += 4) {
        const int r = r0 + i, c = c0 + tx;
        tile[i][tx] = (r < rows && c < cols) ? in[(long)r * cols + c] : 0.0;
    }
    __syncthreads();
    for (int i = ty; i < 64; i += 4) {
        const int c = c0 + i, r = r0 + tx;
        if (c < cols && r < rows) out[(long)c * rows + r] = tile[tx][i];
    }
}

__device__ __forceinline__ void oz_tile_coords(int t, int tiles_m, int tiles_n, int& tm, int& tn) {
    constexpr int G = 8;
    const int per_group = G * tiles_n;
    const int first_m = (t / per_group) * G;
    int gsize = tiles_m - first_m;
    if (gsize > G) gsize = G;
    const int r = t % per_group;
    tm = first_m + r % gsize;
    tn = r / gsize;
}

// byte offset of 16-byte chunk ch of row r in a [rows][128] swizzled slice
__device__ __forceinline__ int oz_off(int r, int ch) { return r * kOzBK + ((ch ^ (r & 7)) << 4); }

template <int S>
__global__ __launch_bounds__(256) void ozaki_fp8_kernel(const uint8_t* __restrict__ Ad, const uint8_t* __restrict__ Bd,
                                                        const int32_t* __restrict__ ea, const int32_t* __restrict__ eb,
                                                        double* __restrict__ C, int M, int N, int K, double alpha,
                                                        double beta, int tiles_m, int tiles_n) {
    __shared__ __align__(16) uint8_t As[S * kOzTile * kOzBK];
    __shared__ __align__(16) uint8_t Bs[S * kOzTile * kOzBK];
    int tm, tn;
    oz_tile_coords(xcd_remap(blockIdx.x, gridDim.x), tiles_m, tiles_n, tm, tn);
    const int row0 = tm * kOzTile, col0 = tn * kOzTile;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const long sa = (long)M * K, sb = (long)N * K;

    v4f acc[S][2][2];
#pragma unroll
    for (int l = 0; l < S; l++)
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[l][i][j] = v4f{0.f, 0.f, 0.f, 0.f};

    const int fr = lane & 15, fk = lane >> 4;   // fragment row, 32-byte k block
    for (int k0 = 0; k0 < K; k0 += kOzBK) {
        // stage S slices of A (64 rows) and B^T (64 rows), 128 bytes of k each
        for (int c = tid; c < S * 1024; c += 256) {
            const int p = c >> 10, rem = c & 1023;
            const bool isb = rem >= 512;
            const int r = (rem & 511) >> 3, ch = rem & 7;
            const int grow = (isb ? col0 : row0) + r, lim = isb ? N : M;
            const int gk = k0 + ch * 16;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (grow < lim && gk < K)
                v = *reinterpret_cast<const uint4*>((isb ? Bd + p * sb : Ad + p * sa) + (long)grow * K + gk);
            *reinterpret_cast<uint4*>((isb ? Bs : As) + p * (kOzTile * kOzBK) + oz_off(r, ch)) = v;
        }
        __syncthreads();
#pragma unroll
        for (int p = 0; p < S; p++) {
            v8i af[2];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const int r = wm + 16 * i + fr;
                const uint8_t* base = As + p * (kOzTile * kOzBK);
                const uint4 lo = *reinterpret_cast<const uint4*>(base + oz_off(r, 2 * fk));
                const uint4 hi = *reinterpret_cast<const uint4*>(base + oz_off(r, 2 * fk + 1));
                af[i] = v8i{(int)lo.x, (int)lo.y, (int)lo.z, (int)lo.w, (int)hi.x, (int)hi.y, (int)hi.z, (int)hi.w};
            }
#pragma unroll
            for (int q = 0; q + p < S; q++) {
                v8i bf[2];
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    const int r = wn + 16 * j + fr;
                    const uint8_t* base = Bs + q * (kOzTile * kOzBK);
                    const uint4 lo = *reinterpret_cast<const uint4*>(base + oz_off(r, 2 * fk));
                    const uint4 hi = *reinterpret_cast<const uint4*>(base + oz_off(r, 2 * fk + 1));
                    bf[j] = v8i{(int)lo.x, (int)lo.y, (int)lo.z, (int)lo.w, (int)hi.x, (int)hi.y, (int)hi.z, (int)hi.w};
                }
#pragma unroll
                for (int i = 0; i < 2; i++)
#pragma unroll
                    for (int j = 0; j < 2; j++)
                        acc[p + q][i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(af[i], bf[j], acc[p + q][i][j],
                                                                                          0, 0, 0, 0, 0, 0);
            }
        }
        __syncthreads();
    }

    // C fragment of a 16x16 f32 tile: register r of lane l is row 4*(l/16) + r, column l%16
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = row0 + wm + 16 * i + 4 * fk + r, col = col0 + wn + 16 * j + fr;
                if (row >= M || col >= N || ea[row] == kOzNonFinite || eb[col] == kOzNonFinite) continue;
                double h = (double)acc[S - 1][i][j][r];
#pragma unroll
                for (int l = S - 2; l >= 0; l--) h = __dadd_rn((double)acc[l][i][j][r], __dmul_rn(h, 0x1p-5));
                double v = __dmul_rn(alpha, ldexp(h, ea[row] + eb[col] - 8));
                double* cp = C + (long)row * N + col;
                if (beta != 0.0) v = __dadd_rn(v, __dmul_rn(beta, *cp));
                *cp = v;
            }
}

// ---- level GEMMs ------------------------------------------------------------
constexpr int kLvBM = 256, kLvBN = 128, kLvBK = 128, kLvStages = 3;
// eight waves, two per SIMD (one wave's fragment reads and DMA run beside the
// other's MFMAs), on a 4 x 2 grid of 64 x 64 wave tiles: 64 accumulators and
// two fragment sets of 64 registers fit the 256 registers of a wave
constexpr int kLvWaves = 8, kLvThreads = 64 * kLvWaves;
constexpr int kLvA = kLvBM * kLvBK, kLvB = kLvBN * kLvBK, kLvStage = kLvA + kLvB;
constexpr int kLvAInstr = kLvA / (1024 * kLvWaves), kLvBInstr = kLvB / (1024 * kLvWaves);
constexpr int kLvLoads = kLvAInstr + kLvBInstr;
constexpr int kLvTM = 4, kLvTN = 4;   // 16x16 MFMA tiles per wave: 64 x 64

// 128-byte rows, 16-byte chunk c of row r at c ^ ((r >> 1) & 7): the 16 rows a
// fragment read touches fall on 64 distinct banks
__device__ __forceinline__ int lv_swz(int r) { return (r >> 1) & 7; }

struct LvFrag {
    v8i a[kLvTM], b[kLvTN];
};

__device__ __forceinline__ void lv_load(LvFrag& f, const char* sa, const char* sb, int wm, int wn) {
    const int lane = threadIdx.x & 63, fr = lane & 15, fk = lane >> 4;
#pragma unroll
    for (int i = 0; i < kLvTM; i++) {
        const int r = wm + 16 * i + fr;
        const uint4 lo = *reinterpret_cast<const uint4*>(sa + r * kLvBK + (((2 * fk) ^ lv_swz(r)) << 4));
        const uint4 hi = *reinterpret_cast<const uint4*>(sa + r * kLvBK + (((2 * fk + 1) ^ lv_swz(r)) << 4));
        f.a[i] = v8i{(int)lo.x, (int)lo.y, (int)lo.z, (int)lo.w, (int)hi.x, (int)hi.y, (int)hi.z, (int)hi.w};
    }
#pragma unroll
    for (int j = 0; j < kLvTN; j++) {
        const int r = wn + 16 * j + fr;
        const uint4 lo = *reinterpret_cast<const uint4*>(sb + r * kLvBK + (((2 * fk) ^ lv_swz(r)) << 4));
        const uint4 hi = *reinterpret_cast<const uint4*>(sb + r * kLvBK + (((2 * fk + 1) ^ lv_swz(r)) << 4));
        f.b[j] = v8i{(int)lo.x, (int)lo.y, (int)lo.z, (int)lo.w, (int)hi.x, (int)hi.y, (int)hi.z, (int)hi.w};
    }
}

template <int I0, int I1>
__device__ __forceinline__ void lv_mfma(v4f (&acc)[kLvTM][kLvTN], const LvFrag& f) {
#pragma unroll
    for (int i = I0; i < I1; i++)
#pragma unroll
        for (int j = 0; j < kLvTN; j++)
            acc[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(f.a[i], f.b[j], acc[i][j], 0, 0, 0, 0, 0, 0);
}

// blockIdx.y = S-1-l: the longest levels start first.  lv[l][M][N] = level sums.
__global__ __launch_bounds__(kLvThreads) void ozaki_level_kernel(const uint8_t* __restrict__ Ad,
                                                          const uint8_t* __restrict__ Bd, float* __restrict__ lv,
                                                          int M, int N, int K, int S, int tiles_m, int tiles_n) {
    __shared__ __align__(16) char smem[kLvStages * kLvStage];
    const int l = S - 1 - (int)blockIdx.y;
    int tm, tn;
    oz_tile_coords(xcd_remap(blockIdx.x, gridDim.x), tiles_m, tiles_n, tm, tn);
    const int row0 = tm * kLvBM, col0 = tn * kLvBN;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const long sa = (long)M * K, sb = (long)N * K;
    const int kc = K / kLvBK;           // slabs per slice pair
    const int n = (l + 1) * kc;         // slabs of this level

    // per-lane source offsets within a slice (row * K + chunk), 8 rows per instruction
    long aoff[kLvAInstr], boff[kLvBInstr];
#pragma unroll
    for (int i = 0; i < kLvAInstr; i++) {
        const int r = (wave * kLvAInstr + i) * 8 + lane / 8;
        aoff[i] = (long)(row0 + r) * K + (((lane % 8) ^ lv_swz(r)) << 4);
    }
#pragma unroll
    for (int i = 0; i < kLvBInstr; i++) {
        const int r = (wave * kLvBInstr + i) * 8 + lane / 8;
        boff[i] = (long)(col0 + r) * K + (((lane % 8) ^ lv_swz(r)) << 4);
    }
    // slab t: pair p = t / kc (A slice p, B slice l - p), k offset (t % kc) * 128
    auto issue = [&](auto q0c, auto q1c, int stage, int t) {
        constexpr int Q0 = decltype(q0c)::value, Q1 = decltype(q1c)::value;
        const int p = t / kc;
        const long koff = (long)(t - p * kc) * kLvBK;
        const uint8_t* ab = Ad + p * sa + koff;
        const uint8_t* bb = Bd + (l - p) * sb + koff;
        char* ds = smem + stage * kLvStage;
#pragma unroll
        for (int q = Q0; q < Q1; q++) {
            if (q < kLvAInstr)
                __builtin_amdgcn_global_load_lds(ab + aoff[q],
                                                 (lds_void_t*)(lds_char_t*)(ds + (wave * kLvAInstr + q) * 1024), 16,
                                                 0, 0);
            else
                __builtin_amdgcn_global_load_lds(
                    bb + boff[q - kLvAInstr],
                    (lds_void_t*)(lds_char_t*)(ds + kLvA + (wave * kLvBInstr + q - kLvAInstr) * 1024), 16, 0, 0);
        }
    };

    v4f acc[kLvTM][kLvTN];
#pragma unroll
    for (int i = 0; i < kLvTM; i++)
#pragma unroll
        for (int j = 0; j < kLvTN; j++) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

    // ring of 3 slabs; one slab = one k-step, so the fragments alternate
    // between two register sets by slab parity (the loop runs slab pairs)
#pragma unroll
    for (int q = 0; q < kLvStages - 1; q++) issue(int_c<0>{}, int_c<kLvLoads>{}, q, q < n ? q : n - 1);
    ring_barrier<kLvLoads * (kLvStages - 2)>();
    LvFrag f0, f1;
    lv_load(f0, smem, smem + kLvA, wm, wn);
    int stage = 0, fill_stage = kLvStages - 1;
    auto step = [&](int t, const LvFrag& cur, LvFrag& nxt) {
        const int fill = t + kLvStages - 1 < n ? t + kLvStages - 1 : n - 1;
        __builtin_amdgcn_sched_barrier(0);
        issue(int_c<0>{}, int_c<kLvLoads>{}, fill_stage, fill);
        lv_mfma<0, kLvTM / 2>(acc, cur);
        interleave<kLvTM / 2 * kLvTN, 0, kLvLoads>();
        __builtin_amdgcn_sched_barrier(0);
        ring_barrier<kLvLoads * (kLvStages - 2)>();
        stage = stage + 1 == kLvStages ? 0 : stage + 1;
        fill_stage = fill_stage + 1 == kLvStages ? 0 : fill_stage + 1;
        const char* na = smem + stage * kLvStage;
        lv_load(nxt, na, na + kLvA, wm, wn);
        lv_mfma<kLvTM / 2, kLvTM>(acc, cur);
        interleave<kLvTM / 2 * kLvTN, kLvTM * 2 + kLvTN * 2, 0>();
        __builtin_amdgcn_sched_barrier(0);
    };
    for (int t = 0; t < n; t += 2) {   // n is even (K % 256 == 0)
        step(t, f0, f1);
        step(t + 1, f1, f0);
    }
    wait_vmcnt<0>();

    // C fragment of a 16x16 f32 tile: register r of lane x is row 4(x/16) + r, column x%16
    float* out = lv + (long)l * M * N;
    const int fr = lane & 15, fk = lane >> 4;
#pragma unroll
    for (int i = 0; i < kLvTM; i++)
#pragma unroll
        for (int j = 0; j < kLvTN; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = row0 + wm + 16 * i + 4 * fk + r, col = col0 + wn + 16 * j + fr;
                out[(long)row * N + col] = acc[i][j][r];
            }
}

// C = alpha 2^(e_i + f_j - 8) Horner_l(level sums, 2^-5) + beta C, in the
// operation order of the all-levels kernel and of the oracle
__global__ __launch_bounds__(256) void ozaki_combine_kernel(const float* __restrict__ lv, const int32_t* __restrict__ ea,
                                                            const int32_t* __restrict__ eb, double* __restrict__ C,
                                                            int M, int N, int S, double alpha, double beta) {
    const long total = (long)M * N, slice = total;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const int row = (int)(e / N), col = (int)(e % N);
        if (ea[row] == kOzNonFinite || eb[col] == kOzNonFinite) continue;   // the fixup kernel's entries
        double h = (double)lv[(long)(S - 1) * slice + e];
        for (int l = S - 2; l >= 0; l--) h = __dadd_rn((double)lv[(long)l * slice + e], __dmul_rn(h, 0x1p-5));
        double v = __dmul_rn(alpha, ldexp(h, ea[row] + eb[col] - 8));
        if (beta != 0.0) v = __dadd_rn(v, __dmul_rn(beta, C[e]));
        C[e] = v;
    }
}

// ---- compensated FP64 GEMM (Dot2: TwoProduct by fma + TwoSum) ---------------
// One output per thread, k ascending, exactly the algorithm of the oracle's
// orc_dgemm_dot2_entries.  A 16x16 thread block shares 16x16 operand tiles.
__global__ __launch_bounds__(256) void dot2_kernel(const double* __restrict__ A, const double* __restrict__ B,
                                                   double* __restrict__ hi_out, double* __restrict__ lo_out, int M,
                                                   int N, int K) {
    __shared__ double as[16][17], bs[16][17];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int row = blockIdx.y * 16 + ty, col = blockIdx.x * 16 + tx;
    double s = 0.0, c = 0.0;
    for (int k0 = 0; k0 < K; k0 += 16) {
        as[ty][tx] = (row < M && k0 + tx < K) ? A[(long)row * K + k0 + tx] : 0.0;
        const int br = k0 + ty;
        bs[ty][tx] = (br < K && col < N) ? B[(long)br * N + col] : 0.0;
        __syncthreads();
        const int kn = K - k0 < 16 ? K - k0 : 16;
        for (int kk = 0; kk < kn; kk++) {
            const double x = as[ty][kk], y = bs[kk][tx];
            const double p = __dmul_rn(x, y);
            const double pe = __fma_rn(x, y, -p);
            const double t = __dadd_rn(s, p);
            const double z = __dsub_rn(t, s);
            const double e = __dadd_rn(__dsub_rn(s, __dsub_rn(t, z)), __dsub_rn(p, z));
            s = t;
            c = __dadd_rn(c, __dadd_rn(e, pe));
        }
        __syncthreads();
    }
    if (row < M && col < N) {
        const double t = __dadd_rn(s, c);
        const double z = __dsub_rn(t, s);
        hi_out[(long)row * N + col] = t;
        lo_out[(long)row * N + col] = __dadd_rn(__dsub_rn(s, __dsub_rn(t, z)), __dsub_rn(c, z));
    }
}

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

struct OzLayout {
    size_t ad, bd, ea, eb, bt, lv, total;
};

bool oz_levels_fit(size_t M, size_t N, size_t K) {
    return M % kLvBM == 0 && N % kLvBN == 0 && K % (2 * kLvBK) == 0 && K > 0;   // an even slab count
}

// level sums [S][M][N] FP32 only when the level GEMMs run
OzLayout oz_layout(size_t M, size_t N, size_t K, int S) {
    OzLayout l;
    l.ad = 0;
    l.bd = l.ad + align256((size_t)S * M * K);
    l.ea = l.bd + align256((size_t)S * N * K);
    l.eb = l.ea + align256(4 * M);
    l.bt = l.eb + align256(4 * N);
    l.lv = l.bt + align256(8 * N * K);
    l.total = l.lv + (oz_levels_fit(M, N, K) ? align256(4 * (size_t)S * M * N) : 0);
    return l;
}

typedef void (*split_fn)(const double*, int, long, uint8_t*, long, int32_t*);
typedef void (*prod_fn)(const uint8_t*, const uint8_t*, const int32_t*, const int32_t*, double*, int, int, int, double,
                        double, int, int);

#define LT_OZ_SPLIT(S) ozaki_split_kernel<S>
#define LT_OZ_PROD(S) ozaki_fp8_kernel<S>
const split_fn kSplit[] = {nullptr,         nullptr,         LT_OZ_SPLIT(2),  LT_OZ_SPLIT(3),
                           LT_OZ_SPLIT(4),  LT_OZ_SPLIT(5),  LT_OZ_SPLIT(6),  LT_OZ_SPLIT(7),
                           LT_OZ_SPLIT(8),  LT_OZ_SPLIT(9),  LT_OZ_SPLIT(10), LT_OZ_SPLIT(11),
                           LT_OZ_SPLIT(12), LT_OZ_SPLIT(13), LT_OZ_SPLIT(14), LT_OZ_SPLIT(15)};
const prod_fn kProd[] = {nullptr, nullptr, LT_OZ_PROD(2), LT_OZ_PROD(3), LT_OZ_PROD(4), LT_OZ_PROD(5),
                         LT_OZ_PROD(6), LT_OZ_PROD(7), LT_OZ_PROD(8), LT_OZ_PROD(9), LT_OZ_PROD(10)};
// the all-levels-in-registers kernel stops at 10 slices; the level GEMMs take
// up to 15 (the exactness bound slices * 256 * K < 2^24 still applies)
constexpr int kMinSlices = 2, kMaxSlices = 10, kMaxSlicesLevels = 15;

}  // namespace
}  // namespace lt

extern "C" {

int lthip_ozaki_max_slices(size_t K) {
    for (int s = lt::kMaxSlices; s >= lt::kMinSlices; s--)
        if ((double)s * 256.0 * (double)K < 16777216.0) return s;
    return 0;
}

size_t lthip_dgemm_ozaki_workspace_size(size_t M, size_t N, size_t K, int slices) {
    const int top = lt::oz_levels_fit(M, N, K) ? lt::kMaxSlicesLevels : lt::kMaxSlices;
    if (slices < lt::kMinSlices || slices > top) return 0;
    return lt::oz_layout(M, N, K, slices).total;
}

int lthip_ozaki_split_rows(const double* X, size_t rows, size_t cols, size_t ld, int slices, uint8_t* digits,
                           int32_t* exps, void* stream) {
    using namespace lt;
    if (slices < kMinSlices || slices > kMaxSlicesLevels)
        return fail(LTHIP_ERR_INVALID_ARG, "ozaki: slices must be 2..15");
    if (cols % 4 != 0 || ld < cols || rows > (1u << 30) || cols > (1u << 30))
        return fail(LTHIP_ERR_INVALID_ARG, "ozaki split: cols must be a multiple of 4 and ld >= cols");
    if (rows == 0 || cols == 0) return LTHIP_OK;
    if (!X || !digits || !exps) return fail(LTHIP_ERR_INVALID_ARG, "ozaki split: null pointer");
    hipLaunchKernelGGL(kSplit[slices], dim3((unsigned)rows), dim3(256), 0, as_stream(stream), X, (int)cols, (long)ld,
                       digits, (long)(rows * cols), exps);
    LT_TRY_LAUNCH("ozaki split launch");
    return LTHIP_OK;
}

int lthip_dgemm_ozaki(const double* A, const double* B, double* C, size_t M, size_t N, size_t K, double alpha,
                      double beta, int slices, void* ws, size_t ws_bytes, void* stream) {
    using namespace lt;
    if (M > (1u << 30) || N > (1u << 30) || K > (1u << 30)) return fail(LTHIP_ERR_INVALID_ARG, "ozaki: size too large");
    const int top = oz_levels_fit(M, N, K) ? kMaxSlicesLevels : kMaxSlices;
    if (slices < kMinSlices || slices > top)
        return fail(LTHIP_ERR_INVALID_ARG, "ozaki: slices must be 2..%d for this shape", top);
    if (K % 16 != 0) return fail(LTHIP_ERR_INVALID_ARG, "ozaki: K must be a multiple of 16");
    if ((double)slices * 256.0 * (double)K >= 16777216.0)
        return fail(LTHIP_ERR_UNSUPPORTED, "ozaki: slices*256*K must stay below 2^24 for exact FP32 level sums");
    if (M == 0 || N == 0) return LTHIP_OK;
    if (!C || (K > 0 && (!A || !B))) return fail(LTHIP_ERR_INVALID_ARG, "ozaki: null pointer");
    const OzLayout l = oz_layout(M, N, K, slices);
    if (!ws || ws_bytes < l.total) return fail(LTHIP_ERR_WORKSPACE, "ozaki: need %zu workspace bytes", l.total);
    if (reinterpret_cast<uintptr_t>(ws) % 256) return fail(LTHIP_ERR_INVALID_ARG, "ozaki: workspace alignment");
    char* w = reinterpret_cast<char*>(ws);
    uint8_t* ad = reinterpret_cast<uint8_t*>(w + l.ad);
    uint8_t* bd = reinterpret_cast<uint8_t*>(w + l.bd);
    int32_t* ea = reinterpret_cast<int32_t*>(w + l.ea);
    int32_t* eb = reinterpret_cast<int32_t*>(w + l.eb);
    double* bt = reinterpret_cast<double*>(w + l.bt);
    hipStream_t s = as_stream(stream);
    const int m = (int)M, n = (int)N, k = (int)K;
    if (k > 0) {
        hipLaunchKernelGGL(kSplit[slices], dim3((unsigned)m), dim3(256), 0, s, A, k, (long)k, ad, (long)M * K, ea);
        LT_TRY_LAUNCH("ozaki split A launch");
        hipLaunchKernelGGL(transpose_f64_kernel, dim3((unsigned)((n + 63) / 64), (unsigned)((k + 63) / 64)), dim3(256), 0,
                           s, B, bt, k, n);
        LT_TRY_LAUNCH("ozaki transpose B launch");
        hipLaunchKernelGGL(kSplit[slices], dim3((unsigned)n), dim3(256), 0, s, bt, k, (long)k, bd, (long)N * K, eb);
        LT_TRY_LAUNCH("ozaki split B launch");
    }  // K == 0: every level sum is 0, whatever the exponents hold
    if (oz_levels_fit(M, N, K)) {
        float* lv = reinterpret_cast<float*>(w + l.lv);
        const int tiles_m = m / kLvBM, tiles_n = n / kLvBN;
        hipLaunchKernelGGL(ozaki_level_kernel, dim3((unsigned)(tiles_m * tiles_n), (unsigned)slices), dim3(kLvThreads), 0, s,
                           ad, bd, lv, m, n, k, slices, tiles_m, tiles_n);
        LT_TRY_LAUNCH("ozaki level launch");
        const long total = (long)m * n;
        long blocks = (total + 255) / 256;
        if (blocks > (long)cu_count() * 16) blocks = (long)cu_count() * 16;
        hipLaunchKernelGGL(ozaki_combine_kernel, dim3((unsigned)blocks), dim3(256), 0, s, lv, ea, eb, C, m, n, slices,
                           alpha, beta);
        LT_TRY_LAUNCH("ozaki combine launch");
    } else {
        const int tiles_m = (m + kOzTile - 1) / kOzTile, tiles_n = (n + kOzTile - 1) / kOzTile;
        hipLaunchKernelGGL(kProd[slices], dim3((unsigned)(tiles_m * tiles_n)), dim3(256), 0, s, ad, bd, ea, eb, C, m,
                           n, k, alpha, beta, tiles_m, tiles_n);
        LT_TRY_LAUNCH("ozaki product launch");
    }
    if (k > 0) {   // entries whose row or column holds a NaN or an infinity
        hipLaunchKernelGGL(ozaki_nonfinite_kernel, dim3((unsigned)((n + 63) / 64), (unsigned)((m + 63) / 64)), dim3(256),
                           0, s, A, bt, ea, eb, C, m, n, k, alpha, beta);
        LT_TRY_LAUNCH("ozaki non-finite fixup launch");
    }
    return LTHIP_OK;
}

int lthip_dgemm_compensated(const double* A, const double* B, double* C_hi, double* C_lo, size_t M, size_t N,
                            size_t K, void* stream) {
    using namespace lt;
    if (M > (1u << 30) || N > (1u << 30) || K > (1u << 30)) return fail(LTHIP_ERR_INVALID_ARG, "dot2: size too large");
    if (M == 0 || N == 0) return LTHIP_OK;
    if (!C_hi || !C_lo || (K > 0 && (!A || !B))) return fail(LTHIP_ERR_INVALID_ARG, "dot2: null pointer");
    hipLaunchKernelGGL(dot2_kernel, dim3((unsigned)((N + 15) / 16), (unsigned)((M + 15) / 16)), dim3(256), 0,
                       as_stream(stream), A, B, C_hi, C_lo, (int)M, (int)N, (int)K);
    LT_TRY_LAUNCH("dot2 launch");
    return LTHIP_OK;
}

}  // extern "C"
