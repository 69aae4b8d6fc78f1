/* rf_mlp.h -- C ABI of the learned mtask cost model (part of librf_verilated.so).
 *
 * A 3-layer MLP, F -> 64 -> 64 -> 1 with ReLU, FP32, trained and evaluated by
 * hand-written HIP kernels. It stands in for V3InstrCount's instruction-count
 * estimate when the partitioner merges logic into mtasks (the reference's
 * LogicMTask cost, src/V3Partition.cpp:456, and finalizeCosts,
 * src/V3Partition.cpp:2481-2545); it has no counterpart in the reference
 * otherwise. Features are per logic vertex / mtask (instruction count,
 * fan-in, fan-out, state bytes touched, branch count, wide-op count), the
 * target is the measured kernel time.
 *
 * Parameter layout (FP32, contiguous): W1[64][F], b1[64], W2[64][64], b2[64],
 * W3[64], b3[1]. Host buffers throughout; errors as in rf_verilated.h.
 */
#ifndef RF_MLP_H
#define RF_MLP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RF_MLP_HIDDEN 64

typedef struct rf_mlp rf_mlp;

/* Number of parameters for F input features: 64F + 64 + 64*64 + 64 + 64 + 1. */
uint32_t rf_mlp_param_count(uint32_t n_features);

/* Create with the given parameters (NULL: seeded He-uniform init). */
int rf_mlp_create(uint32_t n_features, const float* params, uint64_t seed, rf_mlp** out);
int rf_mlp_destroy(rf_mlp* m);
int rf_mlp_get_params(rf_mlp* m, float* out);
int rf_mlp_set_params(rf_mlp* m, const float* params);

/* y[n] = f(x[n][F]) */
int rf_mlp_forward(rf_mlp* m, const float* x, uint32_t n, float* y);

/* Mean squared error loss and its gradient with respect to every parameter
 * (same layout as the parameters) for one batch; does not update. */
int rf_mlp_gradients(rf_mlp* m, const float* x, const float* target, uint32_t n, float* loss,
                     float* grads);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on one batch; returns the
 * batch loss before the step. */
int rf_mlp_train_step(rf_mlp* m, const float* x, const float* target, uint32_t n, float lr,
                      float* loss);

/* Message of the last failing rf_mlp_* call on this thread. */
const char* rf_mlp_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
