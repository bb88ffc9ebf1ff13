/*
 * oracle.c — TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * CPU restatement of Leela's batched-CNN leaf evaluation, written from the
 * reference's behaviour:
 *   Random.cpp:27-57        xoroshiro128+ and its LCG seeding
 *   Network.cpp:1341-1379   rotate_nn_idx / rev_rotate_nn_idx
 *   Im2Col.h:9-50           im2col, layout [c][kr][kc][row][col], zero pad
 *   Network.cpp:344-393     convolve<>: sgemm RowMajor NN, bias, ELU
 *   Network.cpp:395-423     innerproduct<>: sgemv, bias, ELU iff outputs > 1
 *   Network.cpp:450-469     softmax with temperature
 *   Network.cpp:676-740     value head: (1 + tanh(x)) / 2
 *   Network.cpp:742-832     policy head: softmax(T) over the last conv
 * Only tests, the smoke check and the benchmark's CPU-baseline leg use it.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <dlfcn.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ------------------------------------------------------------------ RNG -- */

static inline uint64_t rotl64(uint64_t x, int k) {
    return (x << k) | (x >> (64 - k));
}

/* Random::seedrandom (Random.cpp:51-57): `741103597 * seed` multiplies an
 * int literal by a uint32, so it wraps at 32 bits before it is stored in
 * the uint64 state word; the second product is 64-bit. */
void orc_rng_seed(uint64_t st[2], uint32_t seed) {
    st[0] = (uint64_t)(uint32_t)(741103597u * seed);
    st[1] = (uint64_t)741103597u * st[0];
}

/* Random::random (Random.cpp:27-37). */
uint64_t orc_rng_next(uint64_t st[2]) {
    const uint64_t s0 = st[0];
    uint64_t s1 = st[1];
    const uint64_t result = s0 + s1;
    s1 ^= s0;
    st[0] = rotl64(s0, 55) ^ s1 ^ (s1 << 14);
    st[1] = rotl64(s1, 36);
    return result;
}

/* Random::randuint32(max) (Random.cpp:43-45). */
uint32_t orc_rng_randuint32_max(uint64_t st[2], uint32_t max) {
    return (uint32_t)(((orc_rng_next(st) >> 32) * (uint64_t)max) >> 32);
}

/* Random::randuint16(max) (Random.cpp:39-41). */
uint16_t orc_rng_randuint16(uint64_t st[2], uint16_t max) {
    return (uint16_t)(((orc_rng_next(st) >> 48) * max) >> 16);
}

/* Random::randflt (Random.cpp:59-65): top 24 bits scaled by 2^-24. */
float orc_rng_randflt(uint64_t st[2]) {
    const float umax = 1.0f / (float)(1u << 24);
    uint32_t rnd = (uint32_t)(orc_rng_next(st) >> 40);
    return (float)rnd * umax;
}

void orc_rng_stream(uint32_t seed, int n, uint64_t *out) {
    uint64_t st[2];
    orc_rng_seed(st, seed);
    for (int i = 0; i < n; i++) out[i] = orc_rng_next(st);
}

/* ----------------------------------------------------------- symmetries -- */

/* Network::rotate_nn_idx (Network.cpp:1348-1379): symmetry >= 4 swaps x/y
 * first, then symmetry % 4 selects identity, flip-y, flip-x or both. */
int orc_rotate_nn_idx(int vertex, int symmetry) {
    int x = vertex % 19, y = vertex / 19;
    if (symmetry >= 4) {
        int t = x; x = y; y = t;
        symmetry -= 4;
    }
    int nx = x, ny = y;
    if (symmetry == 1) {
        ny = 18 - y;
    } else if (symmetry == 2) {
        nx = 18 - x;
    } else if (symmetry == 3) {
        nx = 18 - x;
        ny = 18 - y;
    }
    return ny * 19 + nx;
}

/* Network::rev_rotate_nn_idx (Network.cpp:1341-1346): inverse table. */
int orc_rev_rotate_nn_idx(int vertex, int symmetry) {
    static const int inv[8] = {0, 1, 2, 3, 4, 6, 5, 7};
    return orc_rotate_nn_idx(vertex, inv[symmetry]);
}

/* ----------------------------------------------------------------- BLAS -- */

typedef void (*sgemm64_fn)(int, int, int, int64_t, int64_t, int64_t, float,
                           const float *, int64_t, const float *, int64_t,
                           float, float *, int64_t);
typedef void (*sgemv64_fn)(int, int, int64_t, int64_t, float, const float *,
                           int64_t, const float *, int64_t, float, float *,
                           int64_t);
typedef void (*sgemm32_fn)(int, int, int, int, int, int, float, const float *,
                           int, const float *, int, float, float *, int);
typedef void (*sgemv32_fn)(int, int, int, int, float, const float *, int,
                           const float *, int, float, float *, int);
typedef void (*setthreads_fn)(int);
typedef void (*dgemm64_fn)(int, int, int, int64_t, int64_t, int64_t, double,
                           const double *, int64_t, const double *, int64_t,
                           double, double *, int64_t);
typedef void (*dgemm32_fn)(int, int, int, int, int, int, double,
                           const double *, int, const double *, int, double,
                           double *, int);

static sgemm64_fn g_sgemm64;
static sgemv64_fn g_sgemv64;
static sgemm32_fn g_sgemm32;
static sgemv32_fn g_sgemv32;
static dgemm64_fn g_dgemm64;
static dgemm32_fn g_dgemm32;
static char g_blas_name[256] = "oracle loop sgemm (no BLAS loaded)";

enum { CblasRowMajor = 101, CblasNoTrans = 111 };

int orc_load_blas(const char *path) {
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return 0;
    static const char *gemm64[] = {"scipy_cblas_sgemm64_", "cblas_sgemm64_"};
    static const char *gemv64[] = {"scipy_cblas_sgemv64_", "cblas_sgemv64_"};
    static const char *thr[] = {"scipy_openblas_set_num_threads64_",
                                "openblas_set_num_threads64_",
                                "openblas_set_num_threads"};
    for (int i = 0; i < 2; i++) {
        sgemm64_fn a = (sgemm64_fn)dlsym(h, gemm64[i]);
        sgemv64_fn b = (sgemv64_fn)dlsym(h, gemv64[i]);
        if (a && b) {
            g_sgemm64 = a;
            g_sgemv64 = b;
            break;
        }
    }
    static const char *dgemm64[] = {"scipy_cblas_dgemm64_", "cblas_dgemm64_"};
    for (int i = 0; i < 2 && !g_dgemm64; i++)
        g_dgemm64 = (dgemm64_fn)dlsym(h, dgemm64[i]);
    if (!g_dgemm64) g_dgemm32 = (dgemm32_fn)dlsym(h, "cblas_dgemm");
    if (!g_sgemm64) {
        g_sgemm32 = (sgemm32_fn)dlsym(h, "cblas_sgemm");
        g_sgemv32 = (sgemv32_fn)dlsym(h, "cblas_sgemv");
        if (!g_sgemm32 || !g_sgemv32) {
            g_sgemm32 = NULL;
            g_sgemv32 = NULL;
            return 0;
        }
    }
    /* One BLAS thread per worker, as Network::initialize does
     * (Network.cpp:236-240). */
    for (int i = 0; i < 3; i++) {
        setthreads_fn f = (setthreads_fn)dlsym(h, thr[i]);
        if (f) {
            f(1);
            break;
        }
    }
    snprintf(g_blas_name, sizeof g_blas_name, "OpenBLAS via dlopen: %s", path);
    return 1;
}

const char *orc_blas_name(void) { return g_blas_name; }

void orc_sgemm_rm_nn(int M, int N, int K, const float *A, int lda,
                     const float *B, int ldb, float *C, int ldc) {
    if (g_sgemm64) {
        g_sgemm64(CblasRowMajor, CblasNoTrans, CblasNoTrans, M, N, K, 1.0f, A,
                  lda, B, ldb, 0.0f, C, ldc);
        return;
    }
    if (g_sgemm32) {
        g_sgemm32(CblasRowMajor, CblasNoTrans, CblasNoTrans, M, N, K, 1.0f, A,
                  lda, B, ldb, 0.0f, C, ldc);
        return;
    }
    for (int i = 0; i < M; i++) {
        float *c = C + (size_t)i * ldc;
        for (int j = 0; j < N; j++) c[j] = 0.0f;
        for (int k = 0; k < K; k++) {
            const float a = A[(size_t)i * lda + k];
            const float *b = B + (size_t)k * ldb;
            for (int j = 0; j < N; j++) c[j] += a * b[j];
        }
    }
}

/* C = A B in FP64, row-major, no transposes (BLAS dgemm when loaded). */
void orc_dgemm_rm_nn(int M, int N, int K, const double *A, int lda,
                     const double *B, int ldb, double *C, int ldc) {
    if (g_dgemm64) {
        g_dgemm64(CblasRowMajor, CblasNoTrans, CblasNoTrans, M, N, K, 1.0, A,
                  lda, B, ldb, 0.0, C, ldc);
        return;
    }
    if (g_dgemm32) {
        g_dgemm32(CblasRowMajor, CblasNoTrans, CblasNoTrans, M, N, K, 1.0, A,
                  lda, B, ldb, 0.0, C, ldc);
        return;
    }
    for (int i = 0; i < M; i++) {
        double *c = C + (size_t)i * ldc;
        for (int j = 0; j < N; j++) c[j] = 0.0;
        for (int k = 0; k < K; k++) {
            const double a = A[(size_t)i * lda + k];
            const double *b = B + (size_t)k * ldb;
            for (int j = 0; j < N; j++) c[j] += a * b[j];
        }
    }
}

static void sgemv_rm_n(int M, int N, const float *A, const float *x,
                       float *y) {
    if (g_sgemv64) {
        g_sgemv64(CblasRowMajor, CblasNoTrans, M, N, 1.0f, A, N, x, 1, 0.0f,
                  y, 1);
        return;
    }
    if (g_sgemv32) {
        g_sgemv32(CblasRowMajor, CblasNoTrans, M, N, 1.0f, A, N, x, 1, 0.0f,
                  y, 1);
        return;
    }
    for (int i = 0; i < M; i++) {
        float acc = 0.0f;
        for (int j = 0; j < N; j++) acc += A[(size_t)i * N + j] * x[j];
        y[i] = acc;
    }
}

/* --------------------------------------------------- reference CPU path -- */

/* Im2Col.h:9-50 generalised to an S x S board: for every input channel,
 * every kernel row and column, one S x S plane of the shifted input with
 * zeros outside the board. */
void orc_im2col(int channels, int filter, int S, const float *in,
                float *col) {
    const int pad = filter / 2;
    for (int c = 0; c < channels; c++) {
        const float *plane = in + (size_t)c * S * S;
        for (int kr = 0; kr < filter; kr++) {
            for (int kc = 0; kc < filter; kc++) {
                for (int r = 0; r < S; r++) {
                    const int ir = r - pad + kr;
                    for (int q = 0; q < S; q++) {
                        const int ic = q - pad + kc;
                        *col++ = (ir >= 0 && ir < S && ic >= 0 && ic < S)
                                     ? plane[ir * S + ic]
                                     : 0.0f;
                    }
                }
            }
        }
    }
}

/* The same column layout in FP64. */
static void im2col_f64(int channels, int filter, int S, const double *in,
                       double *col) {
    const int pad = filter / 2;
    for (int c = 0; c < channels; c++) {
        const double *plane = in + (size_t)c * S * S;
        for (int kr = 0; kr < filter; kr++)
            for (int kc = 0; kc < filter; kc++)
                for (int r = 0; r < S; r++) {
                    const int ir = r - pad + kr;
                    for (int q = 0; q < S; q++) {
                        const int ic = q - pad + kc;
                        *col++ = (ir >= 0 && ir < S && ic >= 0 && ic < S)
                                     ? plane[ir * S + ic]
                                     : 0.0;
                    }
                }
    }
}

static inline float elu_f32(float v) {
    return v > 0.0f ? v : (expf(v) - 1.0f);
}

static inline float act_f32(float v, int act) {
    if (act == ORC_ACT_ELU) return elu_f32(v);
    /* Caffe's ReLU: max(x, 0) (+ 0 * min(x, 0) for slope 0). */
    if (act == ORC_ACT_RELU) return v > 0.0f ? v : 0.0f;
    return v;
}

static inline double act_f64(double v, int act) {
    if (act == ORC_ACT_ELU) return v <= 0.0 ? expm1(v) : v;
    if (act == ORC_ACT_RELU) return v > 0.0 ? v : 0.0;
    return v;
}

/* convolve<filter, channels, outputs> (Network.cpp:344-393), with the
 * layer's activation in place of the hard-coded ELU. */
static void convolve_act_f32(int filter, int cin, int cout, int S,
                             const float *in, const float *w, const float *b,
                             float *out, float *col, int act) {
    const int P = S * S;
    const int K = cin * filter * filter;
    orc_im2col(cin, filter, S, in, col);
    orc_sgemm_rm_nn(cout, P, K, w, K, col, P, out, P);
    for (int o = 0; o < cout; o++)
        for (int p = 0; p < P; p++)
            out[o * P + p] = act_f32(b[o] + out[o * P + p], act);
}

void orc_convolve_f32(int filter, int cin, int cout, int S, const float *in,
                      const float *w, const float *b, float *out,
                      float *col) {
    convolve_act_f32(filter, cin, cout, S, in, w, b, out, col, ORC_ACT_ELU);
}

/* batchnorm<channels, spatial_size> (Network.cpp:427-448), in place. */
static void batchnorm_f32(const orc_layer *L, int P, float *x) {
    const float epsilon = 1e-5f;
    for (int c = 0; c < L->cout; c++) {
        const float mean = L->bn_means[c] / L->bn_scale;
        float variance = L->bn_vars[c] / L->bn_scale;
        variance += epsilon;
        const float scale_stddiv = 1.0f / sqrtf(variance);
        for (int p = 0; p < P; p++)
            x[c * P + p] = scale_stddiv * (x[c * P + p] - mean);
    }
}

/* The same in FP64 (epsilon the FP32 constant's value). */
static void batchnorm_f64(const orc_layer *L, int P, double *x) {
    const double epsilon = (double)1e-5f;
    for (int c = 0; c < L->cout; c++) {
        const double mean = (double)L->bn_means[c] / L->bn_scale;
        const double variance =
            (double)L->bn_vars[c] / L->bn_scale + epsilon;
        const double inv = 1.0 / sqrt(variance);
        for (int p = 0; p < P; p++)
            x[(size_t)c * P + p] = inv * (x[(size_t)c * P + p] - mean);
    }
}

/* innerproduct<inputs, outputs> (Network.cpp:395-423). */
void orc_innerproduct_f32(int inputs, int outputs, const float *in,
                          const float *w, const float *b, float *out) {
    sgemv_rm_n(outputs, inputs, w, in, out);
    for (int o = 0; o < outputs; o++) {
        float v = b[o] + out[o];
        if (outputs > 1) v = elu_f32(v);
        out[o] = v;
    }
}

/* Network::softmax (Network.cpp:450-469). */
void orc_softmax_f32(int n, const float *in, float *out, float temperature) {
    float alpha = in[0];
    for (int i = 1; i < n; i++)
        if (in[i] > alpha) alpha = in[i];
    alpha /= temperature;
    float denom = 0.0f;
    for (int i = 0; i < n; i++) {
        float v = expf(in[i] / temperature - alpha);
        out[i] = v;
        denom += v;
    }
    for (int i = 0; i < n; i++) out[i] = out[i] / denom;
}

/* ---------------------------------------------------- whole-net drivers -- */

typedef struct fwd_job {
    int nlayers;
    const orc_layer *L;
    int S, head, n, tid, nthreads;
    double temperature;
    const uint8_t *planes;
    void *raw, *out;
    int f64;             /* 0 FP32 reference path, 1 FP64 direct, 2 FP64 GEMM */
    double *const *wd;   /* FP64 copies of the conv weights (mode 2) */
    int status;
} fwd_job;

static int max_act(int nlayers, const orc_layer *L, int S) {
    int m = L[0].cin * S * S;
    for (int i = 0; i < nlayers; i++) {
        int a = L[i].kind == ORC_CONV ? L[i].cout * S * S : L[i].cout;
        if (a > m) m = a;
    }
    return m;
}

static int max_col(int nlayers, const orc_layer *L, int S) {
    int m = 1;
    for (int i = 0; i < nlayers; i++)
        if (L[i].kind == ORC_CONV) {
            int c = L[i].cin * L[i].filter * L[i].filter * S * S;
            if (c > m) m = c;
        }
    return m;
}

/* The reference nets are 19x19 only (Network.cpp:541-543). For a smaller
 * board the value head's 361-input inner product sees the S x S plane at
 * index y*19+x with zeros elsewhere (the library's definition); `a` holds
 * the S*S plane and is rewritten in place as the 361-vector. */
static void embed_plane_f32(float *a, int S) {
    float tmp[361];
    for (int i = 0; i < 361; i++) {
        const int y = i / 19, x = i % 19;
        tmp[i] = (y < S && x < S) ? a[y * S + x] : 0.0f;
    }
    memcpy(a, tmp, sizeof tmp);
}

static void embed_plane_f64(double *a, int S) {
    double tmp[361];
    for (int i = 0; i < 361; i++) {
        const int y = i / 19, x = i % 19;
        tmp[i] = (y < S && x < S) ? a[y * S + x] : 0.0;
    }
    memcpy(a, tmp, sizeof tmp);
}

static void fwd_one_f32(const fwd_job *j, int pos, float *a, float *b,
                        float *col) {
    const int S = j->S, P = S * S, cin0 = j->L[0].cin;
    const uint8_t *pl = j->planes + (size_t)pos * cin0 * P;
    for (int i = 0; i < cin0 * P; i++) a[i] = (float)pl[i];
    int len = cin0 * P;
    for (int l = 0; l < j->nlayers; l++) {
        const orc_layer *L = &j->L[l];
        if (L->kind == ORC_CONV) {
            convolve_act_f32(L->filter, L->cin, L->cout, S, a, L->w, L->b, b,
                             col, L->act);
            if (L->bn_means) batchnorm_f32(L, P, b);
            len = L->cout * P;
        } else {
            if (L->cin == 361 && len == S * S && S < 19) embed_plane_f32(a, S);
            orc_innerproduct_f32(L->cin, L->cout, a, L->w, L->b, b);
            len = L->cout;
        }
        float *t = a; a = b; b = t;
    }
    float *raw = (float *)j->raw, *out = (float *)j->out;
    if (j->head == ORC_HEAD_POLICY) {
        memcpy(raw + (size_t)pos * P, a, sizeof(float) * P);
        orc_softmax_f32(P, a, out + (size_t)pos * P, (float)j->temperature);
    } else {
        (void)len;
        raw[pos] = a[0];
        out[pos] = (1.0f + tanhf(a[0])) / 2.0f;
    }
}

void orc_conv_direct_f64(int filter, int cin, int cout, int S,
                         const double *in, const float *w, const float *b,
                         double *out, int elu) {
    const int pad = filter / 2, SP = S + 2 * pad, P = S * S;
    double *padded = (double *)calloc((size_t)cin * SP * SP, sizeof(double));
    for (int c = 0; c < cin; c++)
        for (int y = 0; y < S; y++)
            memcpy(padded + ((size_t)c * SP + y + pad) * SP + pad,
                   in + (size_t)c * P + y * S, sizeof(double) * S);
    for (int o = 0; o < cout; o++) {
        double *acc = out + (size_t)o * P;
        for (int p = 0; p < P; p++) acc[p] = (double)b[o];
        for (int c = 0; c < cin; c++) {
            const double *src = padded + (size_t)c * SP * SP;
            for (int ky = 0; ky < filter; ky++)
                for (int kx = 0; kx < filter; kx++) {
                    const double wv =
                        (double)w[(((size_t)o * cin + c) * filter + ky) *
                                      filter + kx];
                    for (int y = 0; y < S; y++) {
                        const double *row = src + (y + ky) * SP + kx;
                        double *dst = acc + y * S;
                        for (int x = 0; x < S; x++) dst[x] += wv * row[x];
                    }
                }
        }
        /* elu: 1 = ELU; otherwise 1 + an ORC_ACT_* kind. */
        const int act = elu == 1 ? ORC_ACT_ELU
                        : elu > 1 ? elu - 1 : ORC_ACT_NONE;
        if (elu)
            for (int p = 0; p < P; p++) acc[p] = act_f64(acc[p], act);
    }
    free(padded);
}

/* FP64 conv as im2col + dgemm + bias + ELU (mode 2): the same arithmetic
 * as orc_conv_direct_f64 up to FP64 summation order. */
static void conv_gemm_f64(const orc_layer *L, const double *wd, int S,
                          const double *in, double *out, double *col) {
    const int P = S * S, K = L->cin * L->filter * L->filter;
    im2col_f64(L->cin, L->filter, S, in, col);
    orc_dgemm_rm_nn(L->cout, P, K, wd, K, col, P, out, P);
    for (int o = 0; o < L->cout; o++)
        for (int p = 0; p < P; p++) {
            double v = (double)L->b[o] + out[(size_t)o * P + p];
            out[(size_t)o * P + p] = act_f64(v, L->act);
        }
}

static void fwd_one_f64(const fwd_job *j, int pos, double *a, double *b,
                        double *col) {
    const int S = j->S, P = S * S, cin0 = j->L[0].cin;
    const uint8_t *pl = j->planes + (size_t)pos * cin0 * P;
    for (int i = 0; i < cin0 * P; i++) a[i] = (double)pl[i];
    int len = cin0 * P;
    for (int l = 0; l < j->nlayers; l++) {
        const orc_layer *L = &j->L[l];
        if (L->kind == ORC_CONV) {
            if (j->f64 == 2)
                conv_gemm_f64(L, j->wd[l], S, a, b, col);
            else
                orc_conv_direct_f64(L->filter, L->cin, L->cout, S, a, L->w,
                                    L->b, b, 1 + L->act);
            if (L->bn_means) batchnorm_f64(L, P, b);
            len = L->cout * P;
        } else {
            if (L->cin == 361 && len == P && S < 19) embed_plane_f64(a, S);
            len = L->cout;
            for (int o = 0; o < L->cout; o++) {
                double acc = (double)L->b[o];
                for (int i = 0; i < L->cin; i++)
                    acc += (double)L->w[(size_t)o * L->cin + i] * a[i];
                if (L->cout > 1 && acc <= 0.0) acc = expm1(acc);
                b[o] = acc;
            }
        }
        double *t = a; a = b; b = t;
    }
    double *raw = (double *)j->raw, *out = (double *)j->out;
    if (j->head == ORC_HEAD_POLICY) {
        double mx = a[0];
        for (int i = 1; i < P; i++)
            if (a[i] > mx) mx = a[i];
        double den = 0.0;
        double *o = out + (size_t)pos * P;
        for (int i = 0; i < P; i++) {
            o[i] = exp((a[i] - mx) / j->temperature);
            den += o[i];
        }
        for (int i = 0; i < P; i++) o[i] /= den;
        memcpy(raw + (size_t)pos * P, a, sizeof(double) * P);
    } else {
        raw[pos] = a[0];
        out[pos] = (1.0 + tanh(a[0])) / 2.0;
    }
}

static void *fwd_worker(void *arg) {
    fwd_job *j = (fwd_job *)arg;
    const int act = max_act(j->nlayers, j->L, j->S);
    if (j->f64) {
        double *a = (double *)malloc(sizeof(double) * act);
        double *b = (double *)malloc(sizeof(double) * act);
        double *col =
            j->f64 == 2 ? (double *)malloc(sizeof(double) *
                                           max_col(j->nlayers, j->L, j->S))
                        : NULL;
        if (!a || !b || (j->f64 == 2 && !col)) {
            j->status = -1;
        } else {
            for (int p = j->tid; p < j->n; p += j->nthreads)
                fwd_one_f64(j, p, a, b, col);
        }
        free(a);
        free(b);
        free(col);
    } else {
        float *a = (float *)malloc(sizeof(float) * act);
        float *b = (float *)malloc(sizeof(float) * act);
        float *col =
            (float *)malloc(sizeof(float) * max_col(j->nlayers, j->L, j->S));
        if (!a || !b || !col) {
            j->status = -1;
        } else {
            for (int p = j->tid; p < j->n; p += j->nthreads)
                fwd_one_f32(j, p, a, b, col);
        }
        free(a);
        free(b);
        free(col);
    }
    return NULL;
}

static int check_layers(int nlayers, const orc_layer *L, int S, int head) {
    if (nlayers <= 0 || !L || S <= 0) return -1;
    int width = L[0].cin * S * S;
    int spatial = 1;
    for (int i = 0; i < nlayers; i++) {
        if (L[i].kind == ORC_CONV) {
            if (!spatial || L[i].cin * S * S != width) return -2;
            if (L[i].filter != 1 && L[i].filter != 3 && L[i].filter != 5)
                return -3;
            if (L[i].act < ORC_ACT_ELU || L[i].act > ORC_ACT_NONE) return -8;
            if (L[i].bn_means && (!L[i].bn_vars || L[i].bn_scale == 0.0f))
                return -9;
            width = L[i].cout * S * S;
        } else {
            /* A 1-channel S x S plane may feed a 361-input layer when S < 19
             * (build-defined embedding, see embed_plane). */
            const int embedded = spatial && width == S * S && S < 19 &&
                                 L[i].cin == 361;
            if (L[i].cin != width && !embedded) return -4;
            spatial = 0;
            width = L[i].cout;
        }
    }
    if (head == ORC_HEAD_POLICY && !(spatial && width == S * S)) return -5;
    if (head == ORC_HEAD_VALUE && width != 1) return -6;
    return 0;
}

static int run_forward(int nlayers, const orc_layer *L, int S, int head,
                       double T, int n, const uint8_t *planes, void *raw,
                       void *out, int threads, int f64) {
    int rc = check_layers(nlayers, L, S, head);
    if (rc) return rc;
    if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (threads < 1) threads = 1;
    if (threads > n) threads = n > 0 ? n : 1;
    /* Mode 2: FP64 copies of the conv weights, shared by the workers. */
    double **wd = (double **)calloc(nlayers, sizeof(double *));
    if (!wd) return -7;
    for (int l = 0; f64 == 2 && l < nlayers; l++) {
        if (L[l].kind != ORC_CONV) continue;
        const size_t cnt = (size_t)L[l].cout * L[l].cin * L[l].filter *
                           L[l].filter;
        wd[l] = (double *)malloc(sizeof(double) * cnt);
        if (!wd[l]) {
            for (int q = 0; q < l; q++) free(wd[q]);
            free(wd);
            return -7;
        }
        for (size_t i = 0; i < cnt; i++) wd[l][i] = (double)L[l].w[i];
    }
    pthread_t *th = (pthread_t *)calloc(threads, sizeof(pthread_t));
    fwd_job *jobs = (fwd_job *)calloc(threads, sizeof(fwd_job));
    for (int t = 0; t < threads; t++) {
        fwd_job j = {nlayers, L, S, head, n, t, threads, T, planes, raw, out,
                     f64, wd, 0};
        jobs[t] = j;
        pthread_create(&th[t], NULL, fwd_worker, &jobs[t]);
    }
    int status = 0;
    for (int t = 0; t < threads; t++) {
        pthread_join(th[t], NULL);
        if (jobs[t].status) status = jobs[t].status;
    }
    free(th);
    free(jobs);
    for (int l = 0; l < nlayers; l++) free(wd[l]);
    free(wd);
    return status;
}

int orc_forward_f32(int nlayers, const orc_layer *layers, int S, int head,
                    float temperature, int n, const uint8_t *planes,
                    float *raw, float *out, int threads) {
    return run_forward(nlayers, layers, S, head, temperature, n, planes, raw,
                       out, threads, 0);
}

int orc_forward_f64(int nlayers, const orc_layer *layers, int S, int head,
                    double temperature, int n, const uint8_t *planes,
                    double *raw, double *out, int threads) {
    return run_forward(nlayers, layers, S, head, temperature, n, planes, raw,
                       out, threads, 1);
}

int orc_forward_f64_gemm(int nlayers, const orc_layer *layers, int S,
                         int head, double temperature, int n,
                         const uint8_t *planes, double *raw, double *out,
                         int threads) {
    return run_forward(nlayers, layers, S, head, temperature, n, planes, raw,
                       out, threads, 2);
}
