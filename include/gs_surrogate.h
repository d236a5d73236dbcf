/*
 * gs_surrogate.h - C ABI of the learned surrogate: an ensemble of MLPs that
 * maps simulation parameters to the modulation spectrum, trained and
 * evaluated by gfx950 kernels (BF16 MFMA GEMMs with FP32 accumulation,
 * FP32 master weights, Adam).
 *
 * No counterpart in the reference, which only has the Monte Carlo path; it
 * is for callers that need spectra at rates Monte Carlo cannot give (fits,
 * parameter scans) and take Monte Carlo where the ensemble disagrees or the
 * query leaves the training domain (gs_surrogate_predict's fallback flags).
 *
 * Network (per member): n_inputs -> hidden x n_hidden_layers (SiLU) ->
 * n_outputs, inputs standardised with the training set's mean and standard
 * deviation.  Loss: sum w (y - t)^2 / sum w over a minibatch, per member.
 * Members differ in initialisation and minibatch order.
 */
#ifndef GS_SURROGATE_H
#define GS_SURROGATE_H

#include <stdint.h>

#include "gs_api.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GS_SG_MAX_INPUTS 32
#define GS_SG_MAX_OUTPUTS 32

typedef struct gs_surrogate gs_surrogate;

typedef struct gs_surrogate_config {
    uint32_t n_members;        /* ensemble size (default 4)                  */
    uint32_t n_inputs;         /* <= GS_SG_MAX_INPUTS (default 8)            */
    uint32_t n_outputs;        /* <= GS_SG_MAX_OUTPUTS (default 30)          */
    uint32_t hidden;           /* multiple of 64 (default 256)               */
    uint32_t n_hidden_layers;  /* >= 1 (default 4)                           */
    uint32_t batch;            /* training minibatch, multiple of 64 (1024)  */
    uint64_t seed;             /* initialisation and minibatch order         */
    float lr;                  /* Adam step (default 1e-3)                   */
    float beta1, beta2, eps;   /* Adam (0.9, 0.999, 1e-8)                    */
    float weight_decay;        /* decoupled (default 0)                      */
    float std_limit;           /* fallback when an output's ensemble std
                                  exceeds this (default 0.01)                */
    uint32_t std_outputs;      /* bit i set: output i takes part in the std
                                  test (default: all outputs)                */
} gs_surrogate_config;

/* Fallback flags of gs_surrogate_predict */
#define GS_SG_FALLBACK_STD 0x1u     /* ensemble disagreement above std_limit */
#define GS_SG_FALLBACK_DOMAIN 0x2u  /* an input outside the training box     */

gs_status gs_surrogate_default_config(gs_surrogate_config *cfg);
gs_status gs_surrogate_create(int device, const gs_surrogate_config *cfg, gs_surrogate **out);
void gs_surrogate_destroy(gs_surrogate *s);
const char *gs_surrogate_last_error(const gs_surrogate *s);

/* Train for `epochs` passes over n samples (rows of x: n_inputs floats;
 * rows of y and w: n_outputs floats; w >= 0 weights each output).  The first
 * call fixes the input standardisation and the training box.  *last_loss
 * (may be NULL) receives the mean minibatch loss of the last epoch averaged
 * over members. */
gs_status gs_surrogate_train(gs_surrogate *s, const float *x, const float *y, const float *w, uint32_t n,
                             uint32_t epochs, float *last_loss);

/* Ensemble mean and standard deviation of every output for n queries, and
 * per query the fallback flags (NULL allowed for any output). */
gs_status gs_surrogate_predict(gs_surrogate *s, const float *x, uint32_t n, float *mean, float *sd,
                               uint8_t *fallback);

/* gs_surrogate_predict with every buffer in device memory of the
 * surrogate's device (x: n x n_inputs floats; mean, sd: n x n_outputs; fallback:
 * n bytes; NULL allowed for the outputs): no host copies.  The inputs must be
 * complete when the call is made (it runs on the surrogate's own stream);
 * the outputs are complete on return. */
gs_status gs_surrogate_predict_device(gs_surrogate *s, const float *x, uint32_t n, float *mean, float *sd,
                                      uint8_t *fallback);

/* One forward and backward pass on the given rows for every member, no
 * update: grads receives d loss / d params in the layout of
 * gs_surrogate_get_params, loss one value per member. */
gs_status gs_surrogate_gradients(gs_surrogate *s, const float *x, const float *y, const float *w, uint32_t n,
                                 float *grads, float *loss);

/* Parameters: n_members x per-member block; per member and layer l
 * (0 .. n_hidden_layers) a row-major weight [out_l][in_l] then a bias
 * [out_l], with in_0 = 32 and out_last = 32 (zero padding beyond n_inputs
 * and n_outputs).  layout receives 4 values per layer: offset of W, offset
 * of b, out_l, in_l (per-member offsets, in floats). */
gs_status gs_surrogate_layout(const gs_surrogate *s, uint64_t *per_member, uint64_t *layout, uint32_t layout_cap);
gs_status gs_surrogate_get_params(gs_surrogate *s, float *p, uint64_t count);
gs_status gs_surrogate_set_params(gs_surrogate *s, const float *p, uint64_t count);
/* Input standardisation (mean, sd) and training box (lo, hi), n_inputs each. */
gs_status gs_surrogate_get_normalization(const gs_surrogate *s, float *mean, float *sd, float *lo, float *hi);
gs_status gs_surrogate_set_normalization(gs_surrogate *s, const float *mean, const float *sd, const float *lo,
                                         const float *hi);

#ifdef __cplusplus
}
#endif

#endif
