/*
 * dacti_mlp.h -- C ABI of the refinement-predictor MLP (training and
 * inference on one MI355X, hand-written BF16 MFMA kernels, FP32
 * accumulation).
 *
 * The reference has no learned component: its refinement is decided every
 * substep by the weighted-Laplacian estimator (integrator/error_estimators.hpp:10-47,
 * impl/adaptive_refinement.hpp:16-94).  This network is an addition that
 * forecasts, from a leaf's neighbourhood state, whether the estimator-driven
 * mesh will have refined, kept or coarsened that region some steps ahead;
 * the estimator stays the reference and judge.  No entry point here has a
 * reference counterpart.
 *
 * Network: n_in -> n_hidden (ReLU) x n_layers -> n_out logits; loss: mean
 * softmax cross-entropy.  Master parameters are FP32; every GEMM rounds its
 * operands to BF16 (round to nearest even) and accumulates in FP32.
 * Parameter layout (FP32, flat): for each layer l in order, W_l
 * (out_l x in_l, row-major) then b_l (out_l).  All arrays are host arrays.
 */
#ifndef DACTI_MLP_H
#define DACTI_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dacti_mlp dacti_mlp;

/* weights drawn from U(-1/sqrt(in), 1/sqrt(in)) with the given seed, biases 0 */
int dacti_mlp_create(int32_t n_in, int32_t n_hidden, int32_t n_layers, int32_t n_out, uint64_t seed,
                     dacti_mlp **out);
int dacti_mlp_n_params(dacti_mlp *m, int64_t *n);
int dacti_mlp_get_params(dacti_mlp *m, float *params);
int dacti_mlp_set_params(dacti_mlp *m, const float *params);

/* logits (n x n_out) of n samples x (n x n_in) */
int dacti_mlp_forward(dacti_mlp *m, const float *x, int64_t n, float *logits);

/* dacti_mlp_forward that allocates nothing, for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation).  No counterpart in the reference.  x_dev (n x
 * n_in) and logits_dev (n x n_out) are FP32 device arrays on the network's
 * device, which must be current; the activations live in a fixed 32 MiB
 * device scratch array per GPU (held until the process ends), processed in
 * chunks of up to 8192 rows.  Results equal dacti_mlp_forward's bit for
 * bit.  Calls that overlap on one GPU share that scratch and race: issue
 * them from one thread.  Synchronous (returns when the logits are written).
 * A network whose rows do not fit 16 at a time in the scratch is rejected
 * with an error. */
int dacti_mlp_forward_noalloc(dacti_mlp *m, const float *x_dev, int64_t n, float *logits_dev);

/* mean cross-entropy of n labelled samples and its gradient with respect to
 * the parameters (n_params values; may be NULL) */
int dacti_mlp_loss_grad(dacti_mlp *m, const float *x, const int32_t *y, int64_t n, float *loss, float *grad);

/* Adam (lr, beta1 0.9, beta2 0.999, eps 1e-8) over `epochs` passes of
 * shuffled mini-batches of `batch` samples (a multiple of 16); the data stay
 * on the GPU for the whole run.  *loss: mean loss of the last epoch. */
int dacti_mlp_train(dacti_mlp *m, const float *x, const int32_t *y, int64_t n, int32_t epochs, int32_t batch,
                    float lr, uint64_t seed, float *loss);

/* kernel time of the last train call (ms, HIP events) and its GEMM flops */
int dacti_mlp_train_stats(dacti_mlp *m, double *ms, double *flops);

int dacti_mlp_destroy(dacti_mlp *m);
const char *dacti_mlp_last_error(dacti_mlp *m);

#ifdef __cplusplus
}
#endif
#endif
