/*
 * oracle.h — TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement of Leela's CNN leaf-evaluation algorithm, used by the
 * tests, the smoke check and the CPU-baseline leg of the benchmark as the
 * checker for the HIP path. Nothing in the shipped library links, loads or
 * calls this code.
 *
 * Two forms are restated:
 *   - the reference CPU path (Network.cpp:344-423 + Im2Col.h:9-50): im2col
 *     into a column buffer, RowMajor NN sgemm (OpenBLAS when one was loaded
 *     with orc_load_blas, otherwise the oracle's own loop sgemm), then bias
 *     and ELU, all in FP32;
 *   - an FP64 direct "same" cross-correlation of the same layer list, the
 *     precision reference for BF16/FP8 error reports.
 */
#ifndef LEELA_ORACLE_H
#define LEELA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_CONV = 0, ORC_FC = 1 };
enum { ORC_HEAD_POLICY = 0, ORC_HEAD_VALUE = 1 };
/* Conv activations: the reference's ELU (Network.cpp:382-392), and the
 * ReLU / none of a Caffe net's layer list. */
enum { ORC_ACT_ELU = 0, ORC_ACT_RELU = 1, ORC_ACT_NONE = 2 };

typedef struct orc_layer {
    int kind;        /* ORC_CONV or ORC_FC */
    int filter;      /* conv kernel side (1, 3, 5); 0 for FC */
    int cin;         /* input channels (conv) or inputs (FC) */
    int cout;        /* output channels (conv) or outputs (FC) */
    const float *w;  /* conv: OIHW [cout][cin][k][k]; FC: [cout][cin] */
    const float *b;  /* [cout] */
    int act;         /* conv activation (ORC_ACT_*) */
    /* Batch normalisation of the activated conv output (Network.cpp:
     * 427-448) when bn_means is not NULL: [cout] means and variances and
     * the stored scale factor. */
    const float *bn_means;
    const float *bn_vars;
    float bn_scale;
} orc_layer;

/* xoroshiro128+ with the reference's LCG seeding (Random.cpp:27-57). */
void     orc_rng_seed(uint64_t st[2], uint32_t seed);
uint64_t orc_rng_next(uint64_t st[2]);
uint32_t orc_rng_randuint32_max(uint64_t st[2], uint32_t max);
uint16_t orc_rng_randuint16(uint64_t st[2], uint16_t max);
float    orc_rng_randflt(uint64_t st[2]);
void     orc_rng_stream(uint32_t seed, int n, uint64_t *out);

/* Dihedral symmetries on a 19x19 index (Network.cpp:1341-1379). */
int orc_rotate_nn_idx(int vertex, int symmetry);
int orc_rev_rotate_nn_idx(int vertex, int symmetry);

/* BLAS: dlopen an OpenBLAS (ILP64 scipy/numpy builds or LP64 libopenblas).
 * Returns 1 when sgemm/sgemv were resolved, 0 otherwise. */
int         orc_load_blas(const char *path);
const char *orc_blas_name(void);

/* Building blocks of the reference CPU path. */
void orc_im2col(int channels, int filter, int S, const float *in, float *col);
void orc_sgemm_rm_nn(int M, int N, int K, const float *A, int lda,
                     const float *B, int ldb, float *C, int ldc);
void orc_convolve_f32(int filter, int cin, int cout, int S, const float *in,
                      const float *w, const float *b, float *out, float *col);
void orc_innerproduct_f32(int inputs, int outputs, const float *in,
                          const float *w, const float *b, float *out);
void orc_softmax_f32(int n, const float *in, float *out, float temperature);

/* Whole-net forwards over n positions of S x S, planes uint8 [n][cin0][S*S].
 * Each conv: bias, its activation, then its batch normalisation if any.
 *   policy head: raw = n x S*S (the last conv's output), head =
 *   softmax(raw, T)
 *   value  head: raw = n x 1 (last FC, no ELU),   head = (1 + tanh(raw)) / 2
 * threads <= 0 means one thread per online CPU. Returns 0 on success. */
int orc_forward_f32(int nlayers, const orc_layer *layers, int S, int head,
                    float temperature, int n, const uint8_t *planes,
                    float *raw, float *out, int threads);
int orc_forward_f64(int nlayers, const orc_layer *layers, int S, int head,
                    double temperature, int n, const uint8_t *planes,
                    double *raw, double *out, int threads);

/* The FP64 forward with each conv as im2col + dgemm (OpenBLAS when loaded;
 * one BLAS thread per worker): the FP64 reference at batch sizes where the
 * direct form is slow. Pinned against orc_forward_f64 by the tests. */
int orc_forward_f64_gemm(int nlayers, const orc_layer *layers, int S,
                         int head, double temperature, int n,
                         const uint8_t *planes, double *raw, double *out,
                         int threads);
void orc_dgemm_rm_nn(int M, int N, int K, const double *A, int lda,
                     const double *B, int ldb, double *C, int ldc);

/* FP64 direct same-padded cross-correlation + bias, then ELU when elu is
 * 1, the activation ORC_ACT_* elu - 1 when elu > 1, none when 0. */
void orc_conv_direct_f64(int filter, int cin, int cout, int S,
                         const double *in, const float *w, const float *b,
                         double *out, int elu);

#ifdef __cplusplus
}
#endif
#endif
