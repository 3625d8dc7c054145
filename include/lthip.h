/*
 * lthip.h -- C ABI of the gfx950 (MI355X) implementations of libthom's routines.
 *
 * Conventions (every entry point):
 *   - plain device pointers, sizes and scalars; `stream` is a hipStream_t passed
 *     as void* (NULL = the null stream);
 *   - returns LTHIP_OK (0) or an lthip_status; never exits, never prints;
 *     lthip_last_error() gives a message for the calling thread;
 *   - work is enqueued on `stream` and NOT synchronised: outputs are ready once
 *     the stream is;
 *   - entry points that need scratch take a caller-owned device workspace whose
 *     size comes from the matching *_workspace_size() query.  Workspaces must be
 *     zero-filled once (lthip_workspace_clear) before their first use; every call
 *     leaves the counters they hold at zero again.  Calls that share one
 *     workspace must not overlap in time.
 *   - nothing here allocates device memory on the call path.
 *
 * Each entry point names the reference interface it replaces (file:line in
 * thom-gg/libthom).  The reference's implementations are host functions
 * returning elapsed milliseconds; the C++ adapters in libthom_amd/cli keep that
 * signature and call these.
 */
#ifndef LTHIP_H
#define LTHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum lthip_status {
    LTHIP_OK = 0,
    LTHIP_ERR_INVALID_ARG = 1,   /* bad size, stride, pointer or alignment */
    LTHIP_ERR_UNSUPPORTED = 2,   /* valid request this build does not implement */
    LTHIP_ERR_WORKSPACE = 3,     /* workspace missing or too small */
    LTHIP_ERR_HIP = 4            /* the HIP runtime reported an error */
};

const char *lthip_status_string(int status);
const char *lthip_last_error(void);
/* Library version string (e.g. "0.1.0 gfx950"). */
const char *lthip_version(void);

/* Zero-fills `bytes` of a device workspace with a kernel (graph-capturable). */
int lthip_workspace_clear(void *ws, size_t bytes, void *stream);

/* ---- synthetic inputs (replace curandGenerateUniform, e.g. reduction.cpp:53,
 *      sgemv.cpp:108-110, hgemm.cu:103-111) ----------------------------------
 * Element i of (seed, stream_id) is x = SplitMix64 output i started at
 * state seed + stream_id * 0xD1B54A32D192ED03; u = (x >> 40) * 2^-24 in [0,1);
 * value = lo + (hi - lo) * u computed in double, then rounded to the output
 * type (f16: double -> f32 -> f16, both round-to-nearest-even). */
int lthip_fill_uniform_f32(float *out, size_t n, double lo, double hi, uint64_t seed,
                           uint64_t stream_id, void *stream);
int lthip_fill_uniform_f16(uint16_t *out, size_t n, double lo, double hi, uint64_t seed,
                           uint64_t stream_id, void *stream);
int lthip_fill_uniform_f64(double *out, size_t n, double lo, double hi, uint64_t seed,
                           uint64_t stream_id, void *stream);

/* Plain streaming copy dst[i] = src[i] (bandwidth probe and cache flush). */
int lthip_copy_f32(float *dst, const float *src, size_t n, void *stream);

/* ---- peak probes (no counterpart in the reference, which divides by the
 *      theoretical bandwidth of common.cpp:28-34): the rates this device
 *      delivers, for "fraction of measured peak".
 *      lthip_probe_mfma: `blocks` blocks of 4 waves (kind 2: 8 waves), each
 *      wave `iters` rounds of independent MFMAs on registers; kind 0 =
 *      v_mfma_f32_32x32x16_f16, 1 = v_mfma_scale_f32_32x32x64_f8f6f4 (fp8),
 *      2 = v_mfma_f64_16x16x4_f64 (4 x 4 operand registers, 16 accumulators).
 *      sink holds blocks * 256 floats (kind 2: blocks * 512).  lthip_probe_mfma_flops gives the flops
 *      of one such launch.  lthip_probe_read_f32: read-only stream over x
 *      (16-byte aligned, n % 4 == 0); sink holds blocks * 256 floats. ------- */
double lthip_probe_mfma_flops(int kind, int blocks, int iters);
int lthip_probe_mfma(int kind, int blocks, int iters, float *sink, void *stream);
int lthip_probe_read_f32(const float *x, size_t n, float *sink, int blocks, void *stream);

/* ---- reduction: replaces ReductionImpl float(*)(const float*, float*, int)
 *      (reduction.cpp:20; reduction_warp.cu:83).  *out = sum x[0..n) in FP32,
 *      out is a device pointer.  Deterministic for a given n. -------------- */
size_t lthip_reduce_sum_f32_workspace_size(size_t n);
int lthip_reduce_sum_f32(const float *x, size_t n, float *out, void *ws, size_t ws_bytes,
                         void *stream);
/* ---- allocation-free reduction: no counterpart in the reference.  For callers
 *      that cannot allocate (code under HIP graph capture, tight serving loops,
 *      harnesses that forbid hipMalloc): the same kernel and result as
 *      lthip_reduce_sum_f32, with the arrival counter and block partials in a
 *      static per-GPU scratch array of 16,640 bytes (held until the process
 *      ends) instead of a caller workspace.  Any n fits (the grid is capped at
 *      4096 blocks); a launch that would need more scratch is rejected with
 *      LTHIP_ERR_WORKSPACE.  Calls no allocator.  All calls on one GPU share the
 *      array: calls that overlap on one GPU (different streams or threads) race. */
int lthip_reduce_sum_f32_noalloc(const float *x, size_t n, float *out, void *stream);

/* ---- sdot: replaces SdotImpl float(*)(int, const float*, const float*, float*)
 *      (sdot.cpp:32; sdot_warp.cu:101).  *result = sum x[i]*y[i] in FP32. -- */
size_t lthip_sdot_f32_workspace_size(size_t n);
int lthip_sdot_f32(size_t n, const float *x, const float *y, float *result, void *ws,
                   size_t ws_bytes, void *stream);

/* ---- sgemv: replaces SgemvImpl (sgemv.cpp:38; sgemv_vecto.cu:63,
 *      sgemv_col3_vecto.cu:70).  y = alpha*op(A)x + beta*y, op(A) m x n,
 *      stored row-major (row_major != 0, lda >= n) or column-major (lda >= m).
 *      alpha and beta are DEVICE pointers, as in the reference (sgemv.cpp:112-118).
 *      incx, incy > 0.  beta == 0 means y is not read.  Column-major needs
 *      the workspace (partial sums); row-major ignores it. ---------------- */
size_t lthip_sgemv_f32_workspace_size(int row_major, int m, int n);
int lthip_sgemv_f32(int row_major, int m, int n, const float *alpha, const float *A, int lda,
                    const float *x, int incx, const float *beta, float *y, int incy, void *ws,
                    size_t ws_bytes, void *stream);

/* ---- dgemm: replaces DgemmImpl (dgemm.cpp:36; dgemm_final.cu:175).  Row-major
 *      A (M x K), B (K x N), C (M x N); C = alpha*A*B + beta*C on FP64 MFMA
 *      (v_mfma_f64_16x16x4f64).  Any M, N, K >= 0.  beta == 0: C not read. --- */
int lthip_dgemm_f64(const double *A, const double *B, double *C, size_t M, size_t N, size_t K,
                    double alpha, double beta, void *stream);

/* ---- split-precision (Ozaki-style) dgemm on FP8 MFMA.  No counterpart in the
 *      reference: it is the emulated alternative to DgemmImpl (dgemm.cpp:36),
 *      same operands and C = alpha*A*B + beta*C.  Rows of A and columns of B are
 *      scaled by powers of two and cut into `slices` signed base-32 digits
 *      (E4M3-exact, |d| <= 16); slice products run on
 *      v_mfma_scale_f32_16x16x128_f8f6f4 with exact FP32 level sums and are
 *      recombined in FP64.  2 <= slices <= 10 (<= 15 when M % 256 == 0,
 *      N % 128 == 0 and K % 256 == 0: one FP8 GEMM per level), K % 16 == 0 and
 *      slices * 256 * K < 2^24 (lthip_ozaki_max_slices(K) gives the limit for
 *      any shape, at most 10). */
int lthip_ozaki_max_slices(size_t K);
size_t lthip_dgemm_ozaki_workspace_size(size_t M, size_t N, size_t K, int slices);
int lthip_dgemm_ozaki(const double *A, const double *B, double *C, size_t M, size_t N, size_t K,
                      double alpha, double beta, int slices, void *ws, size_t ws_bytes, void *stream);
/* The split itself (format check): digits[p][r][c] as E4M3 bytes, exps[r]. */
int lthip_ozaki_split_rows(const double *X, size_t rows, size_t cols, size_t ld, int slices,
                           uint8_t *digits, int32_t *exps, void *stream);
/* Compensated (Dot2) FP64 GEMM, C_hi + C_lo ~ A*B to about twice FP64
 * precision: the "FP64 oracle" on the GPU for measuring dgemm errors. */
int lthip_dgemm_compensated(const double *A, const double *B, double *C_hi, double *C_lo, size_t M,
                            size_t N, size_t K, void *stream);

/* ---- hgemm: replaces HgemmFunction (hgemm.h:6-7; hgemm_better_reuse.cu:357).
 *      Row-major A (M x K), B (K x N), C (M x N) fp16 (bit patterns);
 *      C = fp16(alpha*A*B + beta*C) with FP32 accumulation on
 *      v_mfma_f32_32x32x16_f16.  alpha/beta are fp16 bit patterns (the
 *      reference passes __half by value).  K % 8 == 0, N % 8 == 0, A and B
 *      16-byte aligned.  cfg NULL = built-in default; split_k > 1 needs the
 *      workspace (FP32 partials + per-tile counters). ----------------------- */
typedef struct lthip_hgemm_config {
    int bm, bn, bk;   /* block tile; must be one of the built tiles */
    int split_k;      /* >= 1 */
    int stages;       /* LDS buffers: 2 (0 = default) */
} lthip_hgemm_config;

/* Enumerate the built tiles (the autotuner's search space; the reference's
 * hgemm_better_reuse_tuning_candidates(), hgemm_better_reuse.cu:450-454). */
int lthip_hgemm_config_count(void);
int lthip_hgemm_config_get(int index, lthip_hgemm_config *cfg);
/* The untuned choice (the reference's size-keyed default, hgemm_better_reuse.cu:357-376). */
int lthip_hgemm_default_config(size_t M, size_t N, size_t K, lthip_hgemm_config *cfg);
size_t lthip_hgemm_workspace_size(size_t M, size_t N, size_t K, const lthip_hgemm_config *cfg);
int lthip_hgemm_f16(const uint16_t *A, const uint16_t *B, uint16_t *C, size_t M, size_t N, size_t K,
                    uint16_t alpha, uint16_t beta, const lthip_hgemm_config *cfg, void *ws,
                    size_t ws_bytes, void *stream);

/* ---- grouped / batched hgemm (no counterpart in the reference, which runs
 *      one GEMM per launch): `count` independent row-major problems of any
 *      shapes (K, N, lda, ldb multiples of 8) in ONE persistent launch of one
 *      tile shape `cfg`; each problem picks its own split_k (0 or 1 = none).
 *      plan_create uploads the problem table into the workspace (blocking,
 *      once); run launches it; capture records that launch into a hipGraph on a
 *      non-null stream and graph_launch replays it.  The launch leaves its work
 *      queue and split-K counters at zero, so plans and graphs replay as is.
 *      Problems sharing an output must not be in one plan. */
typedef struct lthip_hgemm_problem {
    const uint16_t *A, *B;
    uint16_t *C;
    int M, N, K;
    int lda, ldb, ldc;   /* 0 = packed (K, N, N) */
    int split_k;         /* 0/1 = no split */
} lthip_hgemm_problem;
typedef struct lthip_grouped_plan lthip_grouped_plan;

size_t lthip_hgemm_grouped_workspace_size(const lthip_hgemm_problem *probs, int count,
                                          const lthip_hgemm_config *cfg);
int lthip_hgemm_grouped_plan_create(const lthip_hgemm_problem *probs, int count,
                                    const lthip_hgemm_config *cfg, uint16_t alpha, uint16_t beta,
                                    void *ws, size_t ws_bytes, void *stream,
                                    lthip_grouped_plan **plan);
int lthip_hgemm_grouped_run(lthip_grouped_plan *plan, void *stream);
int lthip_hgemm_grouped_capture(lthip_grouped_plan *plan, void *stream);
int lthip_hgemm_grouped_graph_launch(lthip_grouped_plan *plan, void *stream);
int lthip_hgemm_grouped_plan_info(const lthip_grouped_plan *plan, int *work_items, int *grid);
int lthip_hgemm_grouped_plan_destroy(lthip_grouped_plan *plan);

/* ---- w4a16: replaces W4A16Function float(*)(const W4A16Problem&) (w4a16.h:23-34;
 *      w4a16_naive.cu:90) with per-(group, column) zero-points added.
 *      A: M x K fp16 row-major.  W_packed: (K/8) x N uint32, row k of column n
 *      in bits 4*(k%8) of W_packed[k/8][n] (quantize.cpp:42-53).
 *      scales: (K/group) x N fp16.  qzeros: (K/group) x (N/8) uint32, z of
 *      column n in bits 4*(n%8) of qzeros[g][n/8]; NULL means the reference's
 *      fixed zero-point 8 (w4a16.h:19).  C: M x N fp16.
 *      W[k][n] = fp16((q - z) * scale) exactly as quantize.cpp:58-65 (one
 *      rounding); C = fp16(sum_k A[m][k] W[k][n]) with FP32 accumulation.
 *      K % group == 0, group % 8 == 0, N % 8 == 0. ------------------------- */
typedef struct lthip_w4a16_config {
    int kind;        /* 0 = auto, 1 = small-M split-K stream, 2 = tiled MFMA, 3 = pipelined tiled MFMA
                        (DMA ring; group % 64 == 0, N % 32 == 0, 16-byte aligned scales/zero-points),
                        4 = decode on the vector ALUs (M <= 4, group % (2048 / tile_n) == 0,
                        M * K <= 32768, K * N / 2 < 2^31; no workspace),
                        5 = skinny-M MFMA (M <= 64, N >= 8, group % 32 == 0, group <= 256,
                        K * N / 2 < 2^31; no workspace),
                        6 = stream-K over the pipelined 128 x 128 tile (kind 3's conditions; one
                        block per CU, the tile x K-slab iterations cut evenly between blocks;
                        pieces of a shared tile meet in the workspace, summed in a fixed
                        order: deterministic) */
    int tile_n;      /* kind 1: 64 columns per block (fixed); kinds 2, 3: BN; kind 4: 64 or 32;
                        kind 5: 64; kind 6: 128 */
    int tile_m;      /* kinds 2, 3: BM; kinds 4, 5: ignored (row tiles from M); kind 6: 128 */
    int split_k;     /* K splits (>= 1); > 1 needs the workspace; kinds 4, 5, 6: 1 */
} lthip_w4a16_config;

/* With cfg NULL or kind 0 the size also covers the fallbacks the launch takes
 * on what the query does not see (group, alignment): the 64 x 128 pipelined
 * tile's tail split and kind 2's K split.  Given less, those launches split
 * K less (other summation order, same tolerance). */
size_t lthip_w4a16_workspace_size(size_t M, size_t N, size_t K, const lthip_w4a16_config *cfg);
int lthip_w4a16_gemm(const uint16_t *A, const uint32_t *W_packed, const uint16_t *scales,
                     const uint32_t *qzeros, uint16_t *C, size_t M, size_t N, size_t K, int group,
                     const lthip_w4a16_config *cfg, void *ws, size_t ws_bytes, void *stream);
/* ---- grouped / batched w4a16 (no counterpart in the reference, whose
 *      w4a16 takes one problem per call, w4a16.cpp:247): `count` independent
 *      problems, each with its own A, weights, scales and zero-points in the
 *      layout above, in ONE persistent launch.  Each problem takes a pipelined
 *      tile of 256 columns and 256, 128 or 64 rows: by its M (> 128, > 64,
 *      else) when cfg is NULL or kind 0, or the one tile of a kind-3 cfg
 *      (tile_n 256, tile_m 256/128/64, split_k 1) for every problem.  Per
 *      problem: K > 0, group % 64 == 0, K % group == 0, N % 32 == 0; A, W,
 *      scales, zero-points 16-byte and C 8-byte aligned; M == 0 or N == 0 is
 *      skipped.  plan_create uploads the problem table into the workspace
 *      (blocking, once); run launches it; capture records that launch into a
 *      hipGraph on a non-null stream and graph_launch replays it.  The launch
 *      leaves its work queue at zero, so plans and graphs replay as is.
 *      Problems sharing an output must not be in one plan. */
typedef struct lthip_w4a16_problem {
    const uint16_t *A;          /* M x K fp16 */
    const uint32_t *W_packed;   /* (K/8) x N */
    const uint16_t *scales;     /* (K/group) x N fp16 */
    const uint32_t *qzeros;     /* (K/group) x (N/8), or NULL: zero-point 8 */
    uint16_t *C;                /* M x N fp16 */
    int M, N, K, group;
} lthip_w4a16_problem;
typedef struct lthip_w4a16_grouped_plan lthip_w4a16_grouped_plan;

size_t lthip_w4a16_grouped_workspace_size(const lthip_w4a16_problem *probs, int count,
                                          const lthip_w4a16_config *cfg);
int lthip_w4a16_grouped_plan_create(const lthip_w4a16_problem *probs, int count,
                                    const lthip_w4a16_config *cfg, void *ws, size_t ws_bytes,
                                    void *stream, lthip_w4a16_grouped_plan **plan);
int lthip_w4a16_grouped_run(lthip_w4a16_grouped_plan *plan, void *stream);
int lthip_w4a16_grouped_capture(lthip_w4a16_grouped_plan *plan, void *stream);
int lthip_w4a16_grouped_graph_launch(lthip_w4a16_grouped_plan *plan, void *stream);
int lthip_w4a16_grouped_plan_info(const lthip_w4a16_grouped_plan *plan, int *work_items, int *grid);
int lthip_w4a16_grouped_plan_destroy(lthip_w4a16_grouped_plan *plan);

/* Dequantise the whole weight matrix: W (K x N fp16) -- the bit-exact format
 * check of dequantize_w4 (quantize.cpp:58-65) on the GPU. */
int lthip_w4_dequant(const uint32_t *W_packed, const uint16_t *scales, const uint32_t *qzeros,
                     uint16_t *W, size_t K, size_t N, int group, void *stream);
/* The quantizer on the GPU: W (K x N fp16, row-major, 16-byte aligned) ->
 * W_packed ((K/8) x N, pack_w4 of quantize.cpp:42-53), scales ((K/group) x N
 * fp16) and, for mode 1, qzeros ((K/group) x (N/8) uint32, column n in bits
 * 4*(n%8), z stored as-is).  mode 0 replaces quantize_w4 (quantize.cpp:8-37):
 * scale = fp16(2 max|W| / 15) (1 for an all-zero group), q = clamp(
 * nearbyint(W / scale) + 8, 0, 15); qzeros is not written and may be null.
 * mode 1 is the asymmetric min/max variant (no counterpart in the
 * reference): lo = min(0, min W), hi = max(0, max W), scale = fp16((hi - lo)
 * / 15) (1 when hi == lo), z = clamp(nearbyint(-lo / scale), 0, 15), q =
 * clamp(nearbyint(W / scale) + z, 0, 15).  IEEE FP32 divisions, ties to even:
 * bit-identical to the host quantizer for finite W.  K % group == 0,
 * group % 8 == 0, group <= 1024, N % 8 == 0. */
int lthip_w4_quantize(const uint16_t *W, uint32_t *W_packed, uint16_t *scales, uint32_t *qzeros,
                      size_t K, size_t N, int group, int mode, void *stream);

/* ---- learned tuner: the cost-model MLP (no counterpart in the reference; its
 *      tuner is the exhaustive search of hgemm.cu:258-323).  FP32, row-major:
 *      x [B][I], W1 [H][I], b1 [H], W2 [O][H], b2 [O], hidden [B][H], y [B][O];
 *      y = W2 relu(W1 x + b1) + b2.  loss_grad computes the masked mean squared
 *      error against target [B][O] (mask != 0 marks measured entries) and the
 *      gradients of all four parameters; reductions run in a fixed order. */
int lthip_mlp_forward(const float *x, const float *W1, const float *b1, const float *W2,
                      const float *b2, float *hidden, float *y, int B, int I, int H, int O,
                      void *stream);
size_t lthip_mlp_workspace_size(int B, int H, int O);
int lthip_mlp_loss_grad(const float *x, const float *W1, const float *b1, const float *W2,
                        const float *b2, const float *target, const uint8_t *mask, int B, int I,
                        int H, int O, float *gW1, float *gb1, float *gW2, float *gb2, float *loss,
                        void *ws, size_t ws_bytes, void *stream);
/* Adam (Kingma & Ba 2015) on n parameters, bias-corrected for step >= 1. */
int lthip_adam_step(float *params, const float *grads, float *m, float *v, int n, float lr,
                    float beta1, float beta2, float eps, int step, void *stream);

#ifdef __cplusplus
}
#endif
#endif
