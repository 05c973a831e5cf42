/*
 * atmo_surrogate.h — C ABI of the learned multiple-scattering surrogate.
 *
 * No counterpart in the reference (src/atmo/sources/proland/preprocess/atmo/
 * computes every scattering order with its GL passes).  A small MLP predicts
 * the contribution of the higher scattering orders (>= 2 by default) to the
 * inscatter and irradiance tables; transmittance and the lower orders stay
 * exact (the plain passes).  The plain multi-order path is the judge and the
 * fallback.
 *
 * Network: 14 inputs (4 texel coordinates in [-1, 1], 7 normalised
 * atmosphere features, and the log-scaled rgb of the exactly computed orders
 * at the texel) -> 3 x 256 (activation: z times a smoothstep gate of z) -> 4
 * outputs (log-scaled rgb of the contribution of the orders above the exactly
 * computed ones, standardised; the 4th output is trained to 0 because those
 * orders add nothing to alpha).
 * Training runs the two 256x256 layers in BF16 on the matrix cores with FP32
 * accumulation; master weights, Adam state, the first and last layers are
 * FP32.  The fused inference kernel runs the hidden and output layers in FP16
 * on the matrix cores (FP32 accumulation, FP16 activations) and the first
 * layer as a BF16 high + low pair.
 *
 * Parameter vector layout (FP32): W1[14][256], b1[256], then for l = 2..3
 * Wl[256][256] (row = input), bl[256], then the output layer W5[256][4],
 * b5[4] (named layer 5 as in a four-hidden-layer network).
 */
#ifndef ATMO_SURROGATE_H_
#define ATMO_SURROGATE_H_

#include "atmo.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct atmo_surrogate atmo_surrogate;

/* Creates a surrogate on the engine's device with seeded initial weights. */
ATMO_API atmo_status atmo_surrogate_create(atmo_surrogate **out, int device, uint64_t seed);
ATMO_API atmo_status atmo_surrogate_destroy(atmo_surrogate *s);

/* Number of FP32 parameters (136452) and host copies of them. */
ATMO_API size_t atmo_surrogate_param_count(void);
ATMO_API atmo_status atmo_surrogate_get_params(atmo_surrogate *s, float *host, size_t n);
ATMO_API atmo_status atmo_surrogate_set_params(atmo_surrogate *s, const float *host, size_t n);

/* Scattering orders computed exactly by the plain passes (1: transmittance and
 * single scattering, the default; 2: also order 2, ~half the plain path's
 * cost); the network predicts the contribution of the orders above.  Set it
 * before atmo_surrogate_build_dataset: the dataset's targets, training and
 * atmo_surrogate_precompute all use it. */
ATMO_API atmo_status atmo_surrogate_set_exact_orders(atmo_surrogate *s, int orders);

/* Output standardisation: target_c = (log(max(v_c, 0) + eps) - mu_c) / sd_c. */
ATMO_API atmo_status atmo_surrogate_get_norm(atmo_surrogate *s, float mu[3], float sd[3],
                                             float *eps);
ATMO_API atmo_status atmo_surrogate_set_norm(atmo_surrogate *s, const float mu[3],
                                             const float sd[3], float eps);

/* Training data from the plain method: each set is precomputed with `opts`
 * and with the exact orders only; `samples_per_set` seeded inscatter texels
 * and every irradiance texel of each set become rows.  Replaces any earlier
 * dataset; sets the output standardisation from the data. */
ATMO_API atmo_status atmo_surrogate_build_dataset(atmo_surrogate *s, const atmo_params *sets,
                                                  int nsets, const atmo_options *opts,
                                                  int samples_per_set, uint64_t seed,
                                                  size_t *rows);

/* Adam steps on random batches of `batch` rows (multiple of 128) with a
 * cosine learning-rate schedule from lr to lr_end; losses[steps] (host, may be
 * NULL) receives the mean-square loss of every step. */
ATMO_API atmo_status atmo_surrogate_train(atmo_surrogate *s, int steps, int batch, float lr,
                                          float lr_end, uint64_t seed, float *losses);

/* One forward + backward pass on caller rows (device pointers X[B][16] with
 * 11 used columns, T[B][4] targets, B a multiple of 128): mean-square loss
 * and the gradient of every parameter (host, param_count floats). */
ATMO_API atmo_status atmo_surrogate_loss_grad(atmo_surrogate *s, const float *X, const float *T,
                                              int B, float *loss, float *grad_host);

/* Network outputs for device rows X[B][16] into device Y[B][4] (any B > 0). */
ATMO_API atmo_status atmo_surrogate_predict(atmo_surrogate *s, const float *X, int B, float *Y);

/* Features of texels of one set (codes < n4: inscatter texel index; codes with
 * bit 31 set: irradiance texel index) into device X[n][16].  S_exact, E_exact:
 * the set's device inscatter (float4) and irradiance (float3) tables of the
 * exactly computed orders (the plain path with that many orders). */
ATMO_API atmo_status atmo_surrogate_features(const atmo_params *p, const float *S_exact,
                                             const float *E_exact, const uint32_t *codes, int n,
                                             float *X);

typedef struct atmo_surrogate_report {
    int nsets;            /* sets in the batch */
    int judged;           /* sets recomputed by the plain method as judge */
    int accepted;         /* 1 if every judged set met the tolerance */
    int fallback_sets;    /* sets whose tables come from the plain method */
    float max_rel_err_S;  /* worst judged inscatter error (floored at 1e-4 of the table max) */
    float max_rel_err_E;  /* worst judged irradiance error */
    float ms_exact;       /* exact transmittance + the exactly computed orders */
    float ms_mlp;         /* network inference + combination */
    float ms_judge;       /* plain recompute of the judged sets + comparison */
    float ms_fallback;    /* plain recompute of the rejected batch */
} atmo_surrogate_report;

/* Surrogate-accelerated precompute of nsets sets into caller device tables:
 * exact transmittance and lower orders, network-predicted higher orders,
 * then a judge: ceil(judge_fraction * nsets) seeded sets are recomputed with
 * the plain multi-order path and compared.  If every judged set is within
 * `tolerance` (max relative error, floored at 1e-4 of the table maximum) the
 * batch is accepted (judged sets keep their plain tables); otherwise every set
 * falls back to the plain path.  judge_fraction = 0 accepts without judging.
 * Plain results are bit-identical to atmo_engine_precompute. */
ATMO_API atmo_status atmo_surrogate_precompute(atmo_surrogate *s, const atmo_params *sets,
                                               int nsets, const atmo_options *opts,
                                               const atmo_tables *out, float judge_fraction,
                                               float tolerance, uint64_t seed,
                                               atmo_surrogate_report *report, void *stream);

#ifdef __cplusplus
}
#endif

#endif
