/*
 * oracle.c -- CPU restatement of libthom's result-check path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Built with -ffp-contract=off so
 * every a*b+c below is two roundings, as in the reference's host loops.
 *
 * Parity of each function is pinned in tests/test_oracle_*.py by golden
 * vectors, closed forms and brute force; see the header comments there.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* threading helper: run fn(arg, lo, hi) over [0, n) split into `threads` parts */
typedef void (*range_fn)(void *arg, int lo, int hi);
typedef struct { range_fn fn; void *arg; int lo, hi; } range_job;

static void *range_tramp(void *p) {
    range_job *j = (range_job *)p;
    j->fn(j->arg, j->lo, j->hi);
    return NULL;
}

static void parallel_for(int n, int threads, range_fn fn, void *arg) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    if (threads > n) threads = n > 0 ? n : 1;
    if (threads == 1) { fn(arg, 0, n); return; }
    pthread_t tid[256];
    range_job jobs[256];
    for (int t = 0; t < threads; t++) {
        jobs[t].fn = fn;
        jobs[t].arg = arg;
        jobs[t].lo = (int)((long long)n * t / threads);
        jobs[t].hi = (int)((long long)n * (t + 1) / threads);
        pthread_create(&tid[t], NULL, range_tramp, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
}

/* ------------------------------------------------------------------------ */
/* binary16 conversions (IEEE 754-2008), round to nearest even */
uint16_t orc_f32_to_f16(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t exp = (x >> 23) & 0xffu;
    uint32_t man = x & 0x7fffffu;
    if (exp == 0xffu) {                       /* inf / nan */
        if (man == 0) return (uint16_t)(sign | 0x7c00u);
        return (uint16_t)(sign | 0x7e00u | (man >> 13));
    }
    int e = (int)exp - 127 + 15;              /* biased half exponent */
    if (e >= 31) return (uint16_t)(sign | 0x7c00u);   /* overflow -> inf */
    if (e <= 0) {                             /* subnormal or zero in half */
        if (e < -10) return (uint16_t)sign;   /* rounds to zero (below half of min subnormal) */
        uint32_t m = man | 0x800000u;         /* 24-bit significand */
        int shift = 14 - e;                   /* bits to drop to get a 10-bit subnormal field */
        uint32_t keep = m >> shift;
        uint32_t rem = m & ((1u << shift) - 1u);
        uint32_t half = 1u << (shift - 1);
        if (rem > half || (rem == half && (keep & 1u))) keep++;
        return (uint16_t)(sign | keep);       /* a carry into bit 10 gives the min normal: correct */
    }
    uint32_t keep = man >> 13;
    uint32_t rem = man & 0x1fffu;
    uint32_t out = sign | ((uint32_t)e << 10) | keep;
    if (rem > 0x1000u || (rem == 0x1000u && (keep & 1u))) out++;   /* carry may reach inf: correct */
    return (uint16_t)out;
}

float orc_f16_to_f32(uint16_t h) {
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t exp = (h >> 10) & 0x1fu;
    uint32_t man = h & 0x3ffu;
    uint32_t x;
    if (exp == 0) {
        if (man == 0) {
            x = sign;
        } else {                              /* subnormal: normalise */
            int e = -1;
            do { man <<= 1; e++; } while (!(man & 0x400u));
            man &= 0x3ffu;
            x = sign | ((uint32_t)(127 - 15 - e) << 23) | (man << 13);
        }
    } else if (exp == 0x1fu) {
        x = sign | 0x7f800000u | (man << 13);
    } else {
        x = sign | ((exp - 15 + 127) << 23) | (man << 13);
    }
    float f;
    memcpy(&f, &x, 4);
    return f;
}

uint8_t orc_int_to_e4m3(int d) {
    if (d == 0) return 0;
    int s = d < 0;
    int a = s ? -d : d;                       /* 1..16 */
    int e = 0;
    while ((a >> (e + 1)) != 0) e++;          /* floor(log2 a) */
    int m = e >= 3 ? (a >> (e - 3)) & 7 : (a << (3 - e)) & 7;
    return (uint8_t)((s << 7) | ((e + 7) << 3) | m);
}

float orc_e4m3_to_f32(uint8_t b) {
    int s = b >> 7, e = (b >> 3) & 15, m = b & 7;
    float v;
    if (e == 15 && m == 7) return NAN;        /* E4M3FN: only S.1111.111 is NaN, no inf */
    if (e == 0) v = ldexpf((float)m, -9);     /* subnormal: m * 2^-6 / 8 */
    else v = ldexpf(1.0f + m / 8.0f, e - 7);
    return s ? -v : v;
}

/* ------------------------------------------------------------------------ */
/* SplitMix64 (Steele, Lea, Flood 2014; Vigna's public-domain code) */
#define SM_GAMMA 0x9E3779B97F4A7C15ull
#define SM_STREAM 0xD1B54A32D192ED03ull

static inline uint64_t sm_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t orc_splitmix64(uint64_t *state) {
    *state += SM_GAMMA;
    return sm_mix(*state);
}

uint64_t orc_rng_u64(uint64_t seed, uint64_t stream, uint64_t i) {
    return sm_mix(seed + stream * SM_STREAM + (i + 1) * SM_GAMMA);
}

static inline double u24(uint64_t x) { return (double)(x >> 40) * 0x1p-24; }

void orc_fill_uniform_f32(float *out, size_t n, double lo, double hi, uint64_t seed, uint64_t stream) {
    for (size_t i = 0; i < n; i++) out[i] = (float)(lo + (hi - lo) * u24(orc_rng_u64(seed, stream, i)));
}

void orc_fill_uniform_f16(uint16_t *out, size_t n, double lo, double hi, uint64_t seed, uint64_t stream) {
    for (size_t i = 0; i < n; i++)
        out[i] = orc_f32_to_f16((float)(lo + (hi - lo) * u24(orc_rng_u64(seed, stream, i))));
}

void orc_fill_uniform_f64(double *out, size_t n, double lo, double hi, uint64_t seed, uint64_t stream) {
    for (size_t i = 0; i < n; i++) out[i] = lo + (hi - lo) * u24(orc_rng_u64(seed, stream, i));
}

void orc_fill_normal_f64(double *out, size_t n, uint64_t seed, uint64_t stream) {
    const double two_pi = 6.283185307179586476925286766559;
    for (size_t i = 0; i < n; i += 2) {
        uint64_t a = orc_rng_u64(seed, stream, i), b = orc_rng_u64(seed, stream, i + 1);
        double u1 = (double)((a >> 11) + 1) * 0x1p-53;
        double u2 = (double)(b >> 11) * 0x1p-53;
        double r = sqrt(-2.0 * log(u1));
        out[i] = r * cos(two_pi * u2);
        if (i + 1 < n) out[i + 1] = r * sin(two_pi * u2);
    }
}

void orc_fill_int_u8(uint8_t *out, size_t n, int lo, int hi, uint64_t seed, uint64_t stream) {
    uint64_t span = (uint64_t)(hi - lo + 1);
    for (size_t i = 0; i < n; i++)
        out[i] = (uint8_t)(lo + (int)(((orc_rng_u64(seed, stream, i) >> 32) * span) >> 32));
}

/* ------------------------------------------------------------------------ */
double orc_sum_f64(const float *x, size_t n) {
    double acc = 0.0;                         /* reduction.cpp:62-66 */
    for (size_t i = 0; i < n; i++) acc += (double)x[i];
    return acc;
}

double orc_dot_f64(const float *x, const float *y, size_t n) {
    double acc = 0.0;                         /* sdot.cpp:97-101 */
    for (size_t i = 0; i < n; i++) acc += (double)x[i] * (double)y[i];
    return acc;
}

void orc_sgemv(int row_major, int m, int n, float alpha, const float *A, int lda,
               const float *x, int incx, float beta, float *y, int incy) {
    for (int row = 0; row < m; row++) {       /* sgemv.cpp:134-146 */
        double dot = 0.0;
        for (int col = 0; col < n; col++) {
            double a = row_major ? (double)A[(size_t)row * lda + col] : (double)A[(size_t)col * lda + row];
            dot += a * (double)x[(size_t)col * incx];
        }
        double yv = (double)y[(size_t)row * incy];
        y[(size_t)row * incy] = (float)((double)alpha * dot + (double)beta * yv);
    }
}

/* ------------------------------------------------------------------------ */
typedef struct {
    const double *A, *B;
    double *C;
    int M, N, K;
    double alpha, beta;
} dgemm_args;

/* per row: acc[j] accumulates A[i,k]*B[k,j] over k ascending, exactly the
 * per-element operation sequence of dgemm.cpp:127-131, vectorisable over j */
static void dgemm_rows(void *p, int lo, int hi) {
    dgemm_args *a = (dgemm_args *)p;
    double *acc = (double *)malloc(sizeof(double) * (size_t)a->N);
    for (int i = lo; i < hi; i++) {
        for (int j = 0; j < a->N; j++) acc[j] = 0.0;
        for (int k = 0; k < a->K; k++) {
            double aik = a->A[(size_t)i * a->K + k];
            const double *brow = a->B + (size_t)k * a->N;
            for (int j = 0; j < a->N; j++) acc[j] += aik * brow[j];
        }
        double *crow = a->C + (size_t)i * a->N;
        for (int j = 0; j < a->N; j++) crow[j] = a->alpha * acc[j] + a->beta * crow[j];
    }
    free(acc);
}

void orc_dgemm(const double *A, const double *B, double *C, int M, int N, int K,
               double alpha, double beta, int threads) {
    dgemm_args a = {A, B, C, M, N, K, alpha, beta};
    parallel_for(M, threads, dgemm_rows, &a);
}

/* TwoSum (Knuth) and TwoProduct via fma: error-free transformations */
static inline void two_sum(double a, double b, double *s, double *e) {
    double x = a + b;
    double z = x - a;
    *e = (a - (x - z)) + (b - z);
    *s = x;
}

typedef struct {
    const double *A, *B;
    int N, K;
    const int *rows, *cols;
    double *hi, *lo;
} dot2_args;

static void dot2_range(void *p, int lo, int hi) {
    dot2_args *a = (dot2_args *)p;
    for (int t = lo; t < hi; t++) {
        int i = a->rows[t], j = a->cols[t];
        double s = 0.0, c = 0.0;
        for (int k = 0; k < a->K; k++) {
            double x = a->A[(size_t)i * a->K + k], y = a->B[(size_t)k * a->N + j];
            double p = x * y;
            double pe = fma(x, y, -p);        /* exact product error */
            double e;
            two_sum(s, p, &s, &e);
            c += e + pe;
        }
        double h, l;
        two_sum(s, c, &h, &l);
        a->hi[t] = h;
        a->lo[t] = l;
    }
}

void orc_dgemm_dot2_entries(const double *A, const double *B, int M, int N, int K,
                            const int *rows, const int *cols, size_t count,
                            double *hi, double *lo, int threads) {
    (void)M;
    dot2_args a = {A, B, N, K, rows, cols, hi, lo};
    parallel_for((int)count, threads, dot2_range, &a);
}

/* ------------------------------------------------------------------------ */
void orc_ozaki_split_rows(const double *X, int rows, int cols, int ld, int slices,
                          int8_t *digits, int32_t *exps) {
    for (int r = 0; r < rows; r++) {
        const double *xr = X + (size_t)r * ld;
        double mx = 0.0;
        int bad = 0;
        for (int c = 0; c < cols; c++) {
            bad |= !isfinite(xr[c]);
            mx = fmax(mx, fabs(xr[c]));
        }
        int e = 0;
        if (!bad && mx > 0.0) frexp(mx, &e);  /* mx = f 2^e, f in [0.5, 1) */
        exps[r] = bad ? ORC_OZAKI_NONFINITE : e;
        for (int c = 0; c < cols; c++) {
            double t = bad ? 0.0 : ldexp(xr[c], -e);   /* |t| < 1 */
            for (int p = 0; p < slices; p++) {
                double v = ldexp(t, p == 0 ? 4 : 5);   /* 16 t, then 32 r */
                double d = nearbyint(v);
                digits[((size_t)p * rows + r) * cols + c] = (int8_t)d;
                t = v - d;                    /* exact, |t| <= 1/2 */
            }
        }
    }
}

typedef struct {
    const int8_t *da, *db;   /* [s][M][K], [s][N][K] */
    const double *A, *Bt;    /* for entries with a non-finite row or column */
    const int32_t *ea, *eb;
    double *C;
    int M, N, K, s;
    double alpha, beta;
} ozaki_args;

static void ozaki_rows(void *p, int lo, int hi) {
    ozaki_args *a = (ozaki_args *)p;
    int s = a->s;
    int64_t S[64];
    for (int i = lo; i < hi; i++) {
        for (int j = 0; j < a->N; j++) {
            if (a->ea[i] == ORC_OZAKI_NONFINITE || a->eb[j] == ORC_OZAKI_NONFINITE) {
                /* a plain FP64 dot product, k ascending, then alpha and beta */
                double dot = 0.0;
                for (int k = 0; k < a->K; k++) dot = dot + a->A[(size_t)i * a->K + k] * a->Bt[(size_t)j * a->K + k];
                double *c = a->C + (size_t)i * a->N + j;
                double out = a->alpha * dot;
                if (a->beta != 0.0) out = out + a->beta * *c;
                *c = out;
                continue;
            }
            for (int L = 2; L <= s + 1; L++) S[L] = 0;
            for (int pa = 1; pa <= s; pa++) {
                const int8_t *ra = a->da + ((size_t)(pa - 1) * a->M + i) * a->K;
                for (int pb = 1; pa + pb <= s + 1; pb++) {
                    const int8_t *rb = a->db + ((size_t)(pb - 1) * a->N + j) * a->K;
                    int64_t acc = 0;
                    for (int k = 0; k < a->K; k++) acc += (int64_t)ra[k] * (int64_t)rb[k];
                    S[pa + pb] += acc;
                }
            }
            double acc = (double)S[s + 1];
            for (int L = s; L >= 2; L--) acc = (double)S[L] + acc * 0x1p-5;
            double r = ldexp(acc, a->ea[i] + a->eb[j] - 8);
            double *c = a->C + (size_t)i * a->N + j;
            double out = a->alpha * r;
            if (a->beta != 0.0) out = out + a->beta * *c;
            *c = out;
        }
    }
}

void orc_ozaki_dgemm(const double *A, const double *B, double *C, int M, int N, int K,
                     double alpha, double beta, int slices, int threads) {
    int8_t *da = (int8_t *)malloc((size_t)slices * M * K);
    int8_t *db = (int8_t *)malloc((size_t)slices * N * K);
    int32_t *ea = (int32_t *)malloc(sizeof(int32_t) * (size_t)M);
    int32_t *eb = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    double *Bt = (double *)malloc(sizeof(double) * (size_t)N * K);
    for (int k = 0; k < K; k++)
        for (int j = 0; j < N; j++) Bt[(size_t)j * K + k] = B[(size_t)k * N + j];
    orc_ozaki_split_rows(A, M, K, K, slices, da, ea);
    orc_ozaki_split_rows(Bt, N, K, K, slices, db, eb);   /* columns of B */
    ozaki_args a = {da, db, A, Bt, ea, eb, C, M, N, K, slices, alpha, beta};
    parallel_for(M, threads, ozaki_rows, &a);
    free(da); free(db); free(ea); free(eb); free(Bt);
}

/* ------------------------------------------------------------------------ */
typedef struct {
    const float *A, *B;      /* converted operands */
    const uint16_t *Cin;
    uint16_t *C;
    int N, K;
    float alpha, beta;
} hgemm_args;

static void hgemm_rows(void *p, int lo, int hi) {
    hgemm_args *a = (hgemm_args *)p;
    float *acc = (float *)malloc(sizeof(float) * (size_t)a->N);
    for (int i = lo; i < hi; i++) {
        for (int j = 0; j < a->N; j++) acc[j] = 0.0f;
        for (int k = 0; k < a->K; k++) {      /* hgemm.cu:143-147, k ascending */
            float aik = a->A[(size_t)i * a->K + k];
            const float *brow = a->B + (size_t)k * a->N;
            for (int j = 0; j < a->N; j++) acc[j] += aik * brow[j];
        }
        for (int j = 0; j < a->N; j++) {      /* hgemm.cu:148-149 */
            float res = a->alpha * acc[j] + a->beta * orc_f16_to_f32(a->Cin[(size_t)i * a->N + j]);
            a->C[(size_t)i * a->N + j] = orc_f32_to_f16(res);
        }
    }
    free(acc);
}

void orc_hgemm(const uint16_t *A, const uint16_t *B, uint16_t *C, int M, int N, int K,
               uint16_t alpha, uint16_t beta, int threads) {
    float *fa = (float *)malloc(sizeof(float) * (size_t)M * K);
    float *fb = (float *)malloc(sizeof(float) * (size_t)K * N);
    for (size_t i = 0; i < (size_t)M * K; i++) fa[i] = orc_f16_to_f32(A[i]);
    for (size_t i = 0; i < (size_t)K * N; i++) fb[i] = orc_f16_to_f32(B[i]);
    uint16_t *cin = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)M * N);
    memcpy(cin, C, sizeof(uint16_t) * (size_t)M * N);
    hgemm_args a = {fa, fb, cin, C, N, K, orc_f16_to_f32(alpha), orc_f16_to_f32(beta)};
    parallel_for(M, threads, hgemm_rows, &a);
    free(fa); free(fb); free(cin);
}

/* ------------------------------------------------------------------------ */
void orc_w4_quantize_sym(const uint16_t *W, uint8_t *q, uint16_t *scales, int K, int N, int group) {
    int groups = K / group;
    float *mx = (float *)malloc(sizeof(float) * (size_t)N);
    for (int g = 0; g < groups; g++) {
        for (int n = 0; n < N; n++) mx[n] = 0.0f;                 /* quantize.cpp:14-19 */
        for (int k = g * group; k < (g + 1) * group; k++)
            for (int n = 0; n < N; n++) mx[n] = fmaxf(mx[n], fabsf(orc_f16_to_f32(W[(size_t)k * N + n])));
        for (int n = 0; n < N; n++) {                             /* quantize.cpp:22-25 */
            float s = 2.0f * mx[n] / 15.0f;
            scales[(size_t)g * N + n] = orc_f32_to_f16(s > 0.0f ? s : 1.0f);
        }
        for (int k = g * group; k < (g + 1) * group; k++)       /* quantize.cpp:28-35 */
            for (int n = 0; n < N; n++) {
                float s = orc_f16_to_f32(scales[(size_t)g * N + n]);
                int v = (int)nearbyint(orc_f16_to_f32(W[(size_t)k * N + n]) / s) + 8;
                q[(size_t)k * N + n] = (uint8_t)(v < 0 ? 0 : (v > 15 ? 15 : v));
            }
    }
    free(mx);
}

void orc_w4_quantize_asym(const uint16_t *W, uint8_t *q, uint8_t *z, uint16_t *scales, int K, int N, int group) {
    int groups = K / group;
    float *lo = (float *)malloc(sizeof(float) * (size_t)N);
    float *hi = (float *)malloc(sizeof(float) * (size_t)N);
    for (int g = 0; g < groups; g++) {
        for (int n = 0; n < N; n++) lo[n] = hi[n] = 0.0f;     /* the range always holds 0 */
        for (int k = g * group; k < (g + 1) * group; k++)
            for (int n = 0; n < N; n++) {
                float w = orc_f16_to_f32(W[(size_t)k * N + n]);
                lo[n] = fminf(lo[n], w);
                hi[n] = fmaxf(hi[n], w);
            }
        for (int n = 0; n < N; n++) {
            float r = hi[n] - lo[n];
            uint16_t sb = orc_f32_to_f16(r > 0.0f ? r / 15.0f : 1.0f);
            scales[(size_t)g * N + n] = sb;
            float s = orc_f16_to_f32(sb);
            int zz = (int)nearbyint(-lo[n] / s);
            z[(size_t)g * N + n] = (uint8_t)(zz < 0 ? 0 : (zz > 15 ? 15 : zz));
        }
        for (int k = g * group; k < (g + 1) * group; k++)
            for (int n = 0; n < N; n++) {
                float s = orc_f16_to_f32(scales[(size_t)g * N + n]);
                int v = (int)nearbyint(orc_f16_to_f32(W[(size_t)k * N + n]) / s) + z[(size_t)g * N + n];
                q[(size_t)k * N + n] = (uint8_t)(v < 0 ? 0 : (v > 15 ? 15 : v));
            }
    }
    free(lo);
    free(hi);
}

void orc_w4_pack(const uint8_t *q, uint32_t *packed, int K, int N) {
    for (int pr = 0; pr < K / 8; pr++)                            /* quantize.cpp:43-52 */
        for (int n = 0; n < N; n++) {
            uint32_t v = 0;
            for (int i = 0; i < 8; i++) v |= (uint32_t)(q[(size_t)(pr * 8 + i) * N + n] & 15u) << (4 * i);
            packed[(size_t)pr * N + n] = v;
        }
}

void orc_w4_unpack(const uint32_t *packed, uint8_t *q, int K, int N) {
    for (int k = 0; k < K; k++)
        for (int n = 0; n < N; n++) q[(size_t)k * N + n] = (packed[(size_t)(k / 8) * N + n] >> (4 * (k % 8))) & 15u;
}

void orc_w4_pack_zeros(const uint8_t *z, uint32_t *qz, int groups, int N) {
    for (int g = 0; g < groups; g++)
        for (int n8 = 0; n8 < N / 8; n8++) {
            uint32_t v = 0;
            for (int i = 0; i < 8; i++) v |= (uint32_t)(z[(size_t)g * N + n8 * 8 + i] & 15u) << (4 * i);
            qz[(size_t)g * (N / 8) + n8] = v;
        }
}

void orc_w4_dequant(const uint32_t *packed, const uint16_t *scales, const uint32_t *qz,
                    uint16_t *W, int K, int N, int group) {
    for (int k = 0; k < K; k++) {                                 /* quantize.cpp:59-64 */
        int g = k / group;
        for (int n = 0; n < N; n++) {
            int q = (int)((packed[(size_t)(k / 8) * N + n] >> (4 * (k % 8))) & 15u);
            int z = qz ? (int)((qz[(size_t)g * (N / 8) + n / 8] >> (4 * (n % 8))) & 15u) : 8;
            float s = orc_f16_to_f32(scales[(size_t)g * N + n]);
            W[(size_t)k * N + n] = orc_f32_to_f16((float)(q - z) * s);
        }
    }
}

typedef struct {
    const float *A, *W;
    uint16_t *C;
    int N, K, row0;
} w4_args;

static void w4_rows(void *p, int lo, int hi) {
    w4_args *a = (w4_args *)p;
    float *acc = (float *)malloc(sizeof(float) * (size_t)a->N);
    for (int r = lo; r < hi; r++) {
        int m = a->row0 + r;
        for (int n = 0; n < a->N; n++) acc[n] = 0.0f;
        for (int k = 0; k < a->K; k++) {
            float am = a->A[(size_t)m * a->K + k];
            const float *wrow = a->W + (size_t)k * a->N;
            for (int n = 0; n < a->N; n++) acc[n] += am * wrow[n];
        }
        for (int n = 0; n < a->N; n++) a->C[(size_t)m * a->N + n] = orc_f32_to_f16(acc[n]);
    }
    free(acc);
}

void orc_w4a16_gemm_rows(const uint16_t *A, const uint16_t *Wdq, uint16_t *C, int M, int N, int K,
                         int row0, int rows, int threads) {
    float *fa = (float *)malloc(sizeof(float) * (size_t)M * K);
    float *fw = (float *)malloc(sizeof(float) * (size_t)K * N);
    for (size_t i = (size_t)row0 * K; i < (size_t)(row0 + rows) * K; i++) fa[i] = orc_f16_to_f32(A[i]);
    for (size_t i = 0; i < (size_t)K * N; i++) fw[i] = orc_f16_to_f32(Wdq[i]);
    w4_args a = {fa, fw, C, N, K, row0};
    parallel_for(rows, threads, w4_rows, &a);
    free(fa); free(fw);
}

/* ------------------------------------------------------------------------ */
double orc_mlp_forward(const double *x, const double *W1, const double *b1, const double *W2,
                       const double *b2, int B, int I, int H, int O, double *hidden, double *y) {
    for (int b = 0; b < B; b++) {
        for (int h = 0; h < H; h++) {
            double s = b1[h];
            for (int i = 0; i < I; i++) s += W1[(size_t)h * I + i] * x[(size_t)b * I + i];
            hidden[(size_t)b * H + h] = s > 0.0 ? s : 0.0;
        }
        for (int o = 0; o < O; o++) {
            double s = b2[o];
            for (int h = 0; h < H; h++) s += W2[(size_t)o * H + h] * hidden[(size_t)b * H + h];
            y[(size_t)b * O + o] = s;
        }
    }
    return 0.0;
}

double orc_mlp_loss_grad(const double *x, const double *W1, const double *b1, const double *W2,
                         const double *b2, const double *t, const uint8_t *mask, int B, int I, int H,
                         int O, double *gW1, double *gb1, double *gW2, double *gb2) {
    double *hid = (double *)malloc(sizeof(double) * (size_t)B * H);
    double *y = (double *)malloc(sizeof(double) * (size_t)B * O);
    double *dh = (double *)malloc(sizeof(double) * (size_t)H);
    orc_mlp_forward(x, W1, b1, W2, b2, B, I, H, O, hid, y);
    size_t cnt = 0;
    for (size_t i = 0; i < (size_t)B * O; i++) cnt += mask[i] != 0;
    double inv = cnt ? 1.0 / (double)cnt : 0.0;
    double loss = 0.0;
    memset(gW1, 0, sizeof(double) * (size_t)H * I);
    memset(gb1, 0, sizeof(double) * (size_t)H);
    memset(gW2, 0, sizeof(double) * (size_t)O * H);
    memset(gb2, 0, sizeof(double) * (size_t)O);
    for (int b = 0; b < B; b++) {
        for (int h = 0; h < H; h++) dh[h] = 0.0;
        for (int o = 0; o < O; o++) {
            size_t idx = (size_t)b * O + o;
            if (!mask[idx]) continue;
            double d = y[idx] - t[idx];
            loss += d * d * inv;
            double dy = 2.0 * d * inv;
            gb2[o] += dy;
            for (int h = 0; h < H; h++) {
                gW2[(size_t)o * H + h] += dy * hid[(size_t)b * H + h];
                dh[h] += dy * W2[(size_t)o * H + h];
            }
        }
        for (int h = 0; h < H; h++) {
            if (hid[(size_t)b * H + h] <= 0.0) continue;   /* relu'(0) = 0 */
            gb1[h] += dh[h];
            for (int i = 0; i < I; i++) gW1[(size_t)h * I + i] += dh[h] * x[(size_t)b * I + i];
        }
    }
    free(hid); free(y); free(dh);
    return loss;
}
