/* Runs every oracle function once on small, ragged inputs (odd sizes, strides,
 * zero and non-finite rows, several threads) so that a build with
 * -fsanitize=address,undefined reports out-of-bounds accesses, leaks and
 * undefined behaviour in the oracle.  Test infrastructure: `make -C oracle
 * check-sanitize` builds and runs it; it prints "sanitize ok" at the end. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NEW(T, n) ((T *)calloc((size_t)(n) > 0 ? (size_t)(n) : 1, sizeof(T)))

static void fill_f64(double *x, size_t n, uint64_t stream) { orc_fill_normal_f64(x, n, 7, stream); }

int main(void) {
    /* number formats and the generator */
    for (int b = 0; b < 65536; b++) (void)orc_f32_to_f16(orc_f16_to_f32((uint16_t)b));
    for (int d = -16; d <= 16; d++) (void)orc_e4m3_to_f32(orc_int_to_e4m3(d));
    uint64_t st = 1234;
    for (int i = 0; i < 16; i++) (void)orc_splitmix64(&st);
    (void)orc_rng_u64(1, 2, 3);

    /* streams */
    const size_t n = 1003;
    float *x = NEW(float, n), *y = NEW(float, n);
    orc_fill_uniform_f32(x, n, 0.0, 1.0, 1, 1);
    orc_fill_uniform_f32(y, n, -1.0, 1.0, 1, 2);
    double s = orc_sum_f64(x, n) + orc_dot_f64(x, y, n) + orc_sum_f64(x, 0);

    /* sgemv, both layouts, strides */
    const int m = 13, k = 11, lda = 17;
    float *A = NEW(float, lda * (m > k ? m : k)), *xv = NEW(float, 3 * k), *yv = NEW(float, 2 * m);
    orc_fill_uniform_f32(A, (size_t)lda * (m > k ? m : k), -1.0, 1.0, 1, 3);
    orc_fill_uniform_f32(xv, 3 * k, -1.0, 1.0, 1, 4);
    orc_sgemv(1, m, k, 1.5f, A, lda, xv, 3, 0.5f, yv, 2);
    orc_sgemv(0, m, k, 1.0f, A, lda, xv, 1, 0.0f, yv, 1);

    /* dgemm, Dot2, Ozaki (with a zero row and non-finite entries) */
    const int M = 19, N = 21, K = 36;
    double *Ad = NEW(double, M * K), *Bd = NEW(double, K * N), *Cd = NEW(double, M * N);
    fill_f64(Ad, (size_t)M * K, 5);
    fill_f64(Bd, (size_t)K * N, 6);
    orc_dgemm(Ad, Bd, Cd, M, N, K, 1.0, 0.5, 3);
    int rows[5] = {0, 3, 18, 7, 7}, cols[5] = {0, 20, 1, 7, 19};
    double hi[5], lo[5];
    orc_dgemm_dot2_entries(Ad, Bd, M, N, K, rows, cols, 5, hi, lo, 2);
    for (int c = 0; c < K; c++) Ad[2 * K + c] = 0.0;
    Ad[4 * K + 3] = NAN;
    Bd[5 * N + 6] = INFINITY;
    int8_t *dg = NEW(int8_t, 6 * M * K);
    int32_t *ex = NEW(int32_t, M);
    orc_ozaki_split_rows(Ad, M, K, K, 6, dg, ex);
    orc_ozaki_dgemm(Ad, Bd, Cd, M, N, K, 2.0, -1.0, 6, 4);

    /* hgemm */
    uint16_t *Ah = NEW(uint16_t, M * K), *Bh = NEW(uint16_t, K * N), *Ch = NEW(uint16_t, M * N);
    orc_fill_uniform_f16(Ah, (size_t)M * K, -1.0, 1.0, 1, 7);
    orc_fill_uniform_f16(Bh, (size_t)K * N, -1.0, 1.0, 1, 8);
    orc_hgemm(Ah, Bh, Ch, M, N, K, 0x3C00, 0x3800, 3);

    /* w4a16: quantizers, packing, dequantisation, product */
    const int Kw = 64, Nw = 24, G = 32, Mw = 5;
    uint16_t *W = NEW(uint16_t, Kw * Nw), *sc = NEW(uint16_t, (Kw / G) * Nw), *Wdq = NEW(uint16_t, Kw * Nw);
    uint8_t *q = NEW(uint8_t, Kw * Nw), *z = NEW(uint8_t, (Kw / G) * Nw), *q2 = NEW(uint8_t, Kw * Nw);
    uint32_t *pk = NEW(uint32_t, (Kw / 8) * Nw), *qz = NEW(uint32_t, (Kw / G) * (Nw / 8));
    orc_fill_uniform_f16(W, (size_t)Kw * Nw, -1.0, 1.0, 1, 9);
    for (int c = 0; c < Nw; c++) W[c] = 0;   /* an all-zero column start */
    orc_w4_quantize_sym(W, q, sc, Kw, Nw, G);
    orc_w4_quantize_asym(W, q, z, sc, Kw, Nw, G);
    orc_w4_pack(q, pk, Kw, Nw);
    orc_w4_unpack(pk, q2, Kw, Nw);
    orc_w4_pack_zeros(z, qz, Kw / G, Nw);
    orc_w4_dequant(pk, sc, qz, Wdq, Kw, Nw, G);
    orc_w4_dequant(pk, sc, NULL, Wdq, Kw, Nw, G);
    uint16_t *Aw = NEW(uint16_t, Mw * Kw), *Cw = NEW(uint16_t, Mw * Nw);
    orc_fill_uniform_f16(Aw, (size_t)Mw * Kw, -1.0, 1.0, 1, 10);
    orc_w4a16_gemm_rows(Aw, Wdq, Cw, Mw, Nw, Kw, 1, 3, 2);

    /* MLP forward and gradients */
    const int B = 7, I = 4, H = 9, O = 5;
    double *xm = NEW(double, B * I), *W1 = NEW(double, H * I), *b1 = NEW(double, H), *W2 = NEW(double, O * H);
    double *b2 = NEW(double, O), *tt = NEW(double, B * O), *hid = NEW(double, B * H), *ym = NEW(double, B * O);
    double *g1 = NEW(double, H * I), *gb1 = NEW(double, H), *g2 = NEW(double, O * H), *gb2 = NEW(double, O);
    uint8_t *mask = NEW(uint8_t, B * O);
    fill_f64(xm, (size_t)B * I, 11);
    fill_f64(W1, (size_t)H * I, 12);
    fill_f64(W2, (size_t)O * H, 13);
    fill_f64(tt, (size_t)B * O, 14);
    for (int i = 0; i < B * O; i++) mask[i] = (uint8_t)(i % 3 != 0);
    s += orc_mlp_forward(xm, W1, b1, W2, b2, B, I, H, O, hid, ym);
    s += orc_mlp_loss_grad(xm, W1, b1, W2, b2, tt, mask, B, I, H, O, g1, gb1, g2, gb2);

    free(x); free(y); free(A); free(xv); free(yv); free(Ad); free(Bd); free(Cd); free(dg); free(ex);
    free(Ah); free(Bh); free(Ch); free(W); free(sc); free(Wdq); free(q); free(z); free(q2); free(pk); free(qz);
    free(Aw); free(Cw); free(xm); free(W1); free(b1); free(W2); free(b2); free(tt); free(hid); free(ym);
    free(g1); free(gb1); free(g2); free(gb2); free(mask);
    printf("sanitize ok (%g)\n", isfinite(s) ? 1.0 : 0.0);
    return 0;
}
