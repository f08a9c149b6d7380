/* pach_denoise.h -- C ABI of the learned histogram denoiser in libpach.so.
 *
 * No counterpart in the reference: Pachyderm's ray tracer converges by ray
 * count alone.  This network maps a low-ray-count energy histogram of one
 * receiver (8 octave bands x `bins` 1 ms bins, normalised log energy) plus a
 * few room descriptors to the histogram a high-ray-count trace converges to.
 * It is trained on pairs the tracer itself generates.
 *
 * Network (all matrix products are BF16 MFMA GEMMs with FP32 accumulation;
 * master weights, activations of the residual stream and gradients are FP32):
 *   patchify   : the bins are cut into P = bins / patch patches; patch p of a
 *                sample is the feature row [band-major bins (8 * patch),
 *                descriptors (n_desc), zeros] of width KE (a multiple of 32)
 *   embed      : H = X We + be + Pos[p]                         [P, width]
 *   block x n  : U = gelu(im2col3(H) Wc + bc)   (kernel-3 conv over patches)
 *                V = gelu(U W1 + b1)
 *                H = H + V W2 + b2
 *   head       : Y = H Wh + bh                                  [P, 8 * patch]
 *   output     : out = x + unpatchify(Y)       (residual on the input)
 *   loss       : mean over samples, bands and bins of w[bin] (out - target)^2
 *                (w = 1 until pach_dn_set_loss_weights)
 *
 * Parameter vector (FP32, this order): We [KE, width], be [width],
 * Pos [P, width], per block { Wc [3 width, width], bc [width], W1 [width,
 * 2 width], b1 [2 width], W2 [2 width, width], b2 [width] }, Wh [width,
 * 8 patch], bh [8 patch].  Matrices are row-major [in, out].
 *
 * Conventions as pach.h: 0 = success, negative PACH_E* codes, host buffers
 * are copied during the call, one host thread per handle.
 */
#ifndef PACH_DENOISE_H
#define PACH_DENOISE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pach_dn pach_dn;

typedef struct pach_dn_config {
    int bins;       /* histogram bins per band (multiple of patch) */
    int patch;      /* bins per patch */
    int width;      /* channels of the residual stream (multiple of 64) */
    int blocks;     /* residual blocks */
    int n_desc;     /* room descriptors per sample */
    int max_batch;  /* largest batch passed to any call */
    float beta1, beta2, eps, weight_decay;  /* AdamW */
} pach_dn_config;

int pach_dn_create(pach_dn **out, int device, const pach_dn_config *cfg);
void pach_dn_destroy(pach_dn *dn);
const char *pach_dn_last_error(pach_dn *dn);
int64_t pach_dn_param_count(pach_dn *dn);

/* FP32 parameter vector in and out (the layout above); set also resets the
 * Adam moments and step count. */
int pach_dn_set_params(pach_dn *dn, const float *params);
int pach_dn_get_params(pach_dn *dn, float *params);

/* x: [batch, 8, bins] normalised log energy, desc: [batch, n_desc];
 * out: [batch, 8, bins]. */
int pach_dn_forward(pach_dn *dn, const float *x, const float *desc, int batch, float *out);

/* Loss and its gradient (the parameter-vector layout) for one batch, without
 * an update; grad may be NULL. */
int pach_dn_loss_grad(pach_dn *dn, const float *x, const float *desc, const float *target, int batch,
                      double *loss, float *grad);

/* Per-bin loss weights w[bins] (>= 0) for later loss / gradient / training
 * calls; NULL restores w = 1. */
int pach_dn_set_loss_weights(pach_dn *dn, const float *w);

/* Training set kept in device memory: x, target [n, 8, bins], desc [n, n_desc]. */
int pach_dn_dataset(pach_dn *dn, const float *x, const float *desc, const float *target, int64_t n);

/* One AdamW step on the dataset rows idx[0..batch-1]; loss before the step. */
int pach_dn_train_step(pach_dn *dn, const int64_t *idx, int batch, float lr, double *loss);

/* C[M,N] = op(A) op(B) in BF16 with FP32 accumulation, through the same GEMM
 * kernel the network uses (A, B as BF16 bit patterns; ta: A is given as
 * [K, M]; tb: B is given as [N, K]); host buffers.  For tests. */
int pach_dn_gemm(pach_dn *dn, const uint16_t *a, const uint16_t *b, int M, int N, int K, int ta, int tb,
                 float *c);

#ifdef __cplusplus
}
#endif
#endif
