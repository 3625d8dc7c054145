/*
 * oracle.h -- CPU restatement of libthom's result-check path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library or CLI links this
 * file: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it, and only as the checker (or as the timed CPU baseline), never as
 * the code path that is measured or returned to a caller.
 *
 * Every function cites the reference file:line whose arithmetic it restates.
 * Functions marked "(new)" have no reference counterpart; their check is named
 * next to them.  Pinning of each function (golden vectors, closed forms,
 * brute force) lives in tests/test_oracle_*.py.
 */
#ifndef LT_ORACLE_H
#define LT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- number formats ---------------------------------------------------- */
/* IEEE binary16 <-> binary32, round-to-nearest-even (the __float2half /
 * __half2float pair the reference uses, e.g. quantize.cpp:64). */
uint16_t orc_f32_to_f16(float f);
float orc_f16_to_f32(uint16_t h);
/* OCP FP8 E4M3 (fn) encode of a small integer in [-16, 16] (exact), decode of
 * any byte.  (new: Ozaki digits; pinned exhaustively against the format
 * definition.) */
uint8_t orc_int_to_e4m3(int d);
float orc_e4m3_to_f32(uint8_t b);

/* ---- input generation (new; the reference uses cuRAND, seed 1234ULL) ---- */
/* SplitMix64 step (Vigna); pinned by the published seed-1234567 vector. */
uint64_t orc_splitmix64(uint64_t *state);
/* Element i of stream `stream` under `seed` is output i of a SplitMix64 run
 * started at state seed + stream * 0xD1B54A32D192ED03. */
uint64_t orc_rng_u64(uint64_t seed, uint64_t stream, uint64_t i);
/* u = (x >> 40) * 2^-24 in [0,1); value = (float)(lo + (hi - lo) * u) in double. */
void orc_fill_uniform_f32(float *out, size_t n, double lo, double hi, uint64_t seed, uint64_t stream);
void orc_fill_uniform_f16(uint16_t *out, size_t n, double lo, double hi, uint64_t seed, uint64_t stream);
void orc_fill_uniform_f64(double *out, size_t n, double lo, double hi, uint64_t seed, uint64_t stream);
/* Box-Muller on pairs (u1 = ((x>>11)+1) 2^-53, u2 = (x>>11) 2^-53). */
void orc_fill_normal_f64(double *out, size_t n, uint64_t seed, uint64_t stream);
/* integers uniform in [lo, hi]: lo + ((x >> 32) * (hi - lo + 1)) >> 32 */
void orc_fill_int_u8(uint8_t *out, size_t n, int lo, int hi, uint64_t seed, uint64_t stream);

/* ---- reduction / sdot ------------------------------------------------- */
/* reduction.cpp:61-67: sequential sum in FP64 */
double orc_sum_f64(const float *x, size_t n);
/* sdot.cpp:96-102: sequential sum of (double)x*(double)y */
double orc_dot_f64(const float *x, const float *y, size_t n);

/* ---- sgemv ------------------------------------------------------------ */
/* sgemv.cpp:132-147: FP64 dot per row, y = (float)(alpha*dot + beta*y).
 * row_major selects A[row*lda+col] vs A[col*lda+row]; incx/incy > 0. */
void orc_sgemv(int row_major, int m, int n, float alpha, const float *A, int lda,
               const float *x, int incx, float beta, float *y, int incy);

/* ---- dgemm ------------------------------------------------------------ */
/* dgemm.cpp:122-133: C = alpha*sum_k A[i,k]B[k,j] + beta*C, k ascending, all
 * row-major.  Multithreaded over rows; per-element arithmetic unchanged. */
void orc_dgemm(const double *A, const double *B, double *C, int M, int N, int K,
               double alpha, double beta, int threads);
/* (new) compensated Dot2 (Ogita-Rump-Oishi 2005) for the listed entries:
 * hi[t] + lo[t] ~ sum_k A[i_t,k] B[k,j_t] with error ~ u^2 cond.  This is the
 * "FP64 oracle" against which native and emulated dgemm errors are measured. */
void orc_dgemm_dot2_entries(const double *A, const double *B, int M, int N, int K,
                            const int *rows, const int *cols, size_t count,
                            double *hi, double *lo, int threads);
/* (new) Ozaki split of the rows of X (ld = row stride): per row r an exponent
 * e_r with max|X[r,:]| = f 2^e_r, f in [0.5,1) (0 for a zero row), and
 * digits d_1 = rint(16 t), d_p = rint(32 r_{p-1}) with t = X 2^-e_r,
 * r_p = (previous) - d_p.  digits[p][r][c] in [-16,16].  A row holding a NaN
 * or an infinity gets e_r = ORC_OZAKI_NONFINITE and all-zero digits. */
#define ORC_OZAKI_NONFINITE ((int32_t)0x80000000)
void orc_ozaki_split_rows(const double *X, int rows, int cols, int ld, int slices,
                          int8_t *digits, int32_t *exps);
/* (new) emulated dgemm with exactly the GPU's combination order:
 * S_L = sum_{p+q=L} D^A_p (D^B_q)^T  (exact integers), acc = S_{s+1};
 * acc = S_L + acc*2^-5 for L = s..2;  r = ldexp(acc, e_i + f_j - 8);
 * C = alpha*r + beta*C (beta term only if beta != 0).  No FMA contraction.
 * Entries whose row of A or column of B holds a NaN or an infinity:
 * alpha * (sum_k A_ik B_kj, k ascending) (+ beta*C) in plain FP64, so NaN and
 * infinities propagate as in an FP64 product. */
void orc_ozaki_dgemm(const double *A, const double *B, double *C, int M, int N, int K,
                     double alpha, double beta, int slices, int threads);

/* ---- hgemm ------------------------------------------------------------ */
/* hgemm.cu:137-152: FP32 sum of fp16 products, k ascending, no FMA;
 * C = fp16(alpha*sum + beta*C). Operands are fp16 bit patterns, row-major. */
void orc_hgemm(const uint16_t *A, const uint16_t *B, uint16_t *C, int M, int N, int K,
               uint16_t alpha, uint16_t beta, int threads);

/* ---- w4a16 ------------------------------------------------------------ */
/* quantize.cpp:8-37 (symmetric, ZERO_POINT 8, GROUP_SIZE = group) */
void orc_w4_quantize_sym(const uint16_t *W, uint8_t *q, uint16_t *scales, int K, int N, int group);
/* Asymmetric min/max variant of quantize.cpp:8-37 (no counterpart in the
 * reference; the zero-point layout is GPTQ's qzeros, z stored as-is).  Per
 * (group, column): lo = min(0, min W), hi = max(0, max W) in FP32;
 * scale = fp16((hi - lo) / 15), or 1 when hi == lo; z = clamp(nearbyint(-lo /
 * scale), 0, 15); q = clamp(nearbyint(W / scale) + z, 0, 15).  Every division
 * is one IEEE FP32 division and nearbyint rounds half to even, as in
 * quantize.cpp:31-33.  z: (K/group) x N bytes. */
void orc_w4_quantize_asym(const uint16_t *W, uint8_t *q, uint8_t *z, uint16_t *scales, int K, int N, int group);
/* quantize.cpp:42-53: packed[k/8][n] holds row k in bits 4*(k%8) */
void orc_w4_pack(const uint8_t *q, uint32_t *packed, int K, int N);
void orc_w4_unpack(const uint32_t *packed, uint8_t *q, int K, int N);
/* (new) zero-points z[g][n] packed 8 per uint32 along N: qz[g][n/8], bits 4*(n%8) */
void orc_w4_pack_zeros(const uint8_t *z, uint32_t *qz, int groups, int N);
/* quantize.cpp:58-65 generalised with per-(group,column) zero-points:
 * W[k][n] = fp16((float)(q - z) * scale)  (one rounding; z = 8 if qz == NULL) */
void orc_w4_dequant(const uint32_t *packed, const uint16_t *scales, const uint32_t *qz,
                    uint16_t *W, int K, int N, int group);
/* w4a16.cpp:178-184 restated on the CPU (cuBLAS COMPUTE_32F on dequantised W):
 * C[m][n] = fp16(sum_k (float)A[m][k] * (float)W[k][n]), k ascending.  Rows
 * [row0, row0+rows) of C only (so large M can be sampled). */
void orc_w4a16_gemm_rows(const uint16_t *A, const uint16_t *Wdq, uint16_t *C, int M, int N, int K,
                         int row0, int rows, int threads);

/* ---- MLP cost model (new) --------------------------------------------- */
/* y = W2 relu(W1 x + b1) + b2; x [B][I], W1 [H][I], W2 [O][H], y [B][O].
 * loss = mean over valid (mask != 0) entries of (y - t)^2.  Doubles. */
double orc_mlp_forward(const double *x, const double *W1, const double *b1, const double *W2,
                       const double *b2, int B, int I, int H, int O, double *hidden, double *y);
double orc_mlp_loss_grad(const double *x, const double *W1, const double *b1, const double *W2,
                         const double *b2, const double *t, const uint8_t *mask, int B, int I, int H,
                         int O, double *gW1, double *gb1, double *gW2, double *gb2);

#ifdef __cplusplus
}
#endif
#endif
