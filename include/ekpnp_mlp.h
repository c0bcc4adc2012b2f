/*
 * ekpnp_mlp.h -- C ABI of the warm-start surrogate: a multilayer perceptron
 * (n_layers hidden layers of n_hidden units, GELU, linear output) trained and
 * evaluated on the GPU by hand-written BF16 MFMA kernels with FP32
 * accumulation, FP32 master weights and an Adam optimiser.
 *
 * The reference has no counterpart: its only warm start is the PB-Picard
 * iteration (LBM.cu:106-127).  The surrogate maps channel parameters
 * (zeta, kappa h, E_x, n_y) to wall-normal profiles of psi, c+, c- and u_x;
 * the plain solver always finishes the run (ekpnp_run_to_residual) and
 * decides whether a prediction is accepted.
 *
 * Parameter vector (FP32, ekpnp_mlp_get_params/set_params): for each layer
 * l = 1..n_layers+1 the weights W_l [out][in_padded] row-major, then the
 * bias b_l [out]; in_padded of the first layer is n_in rounded up to 32
 * (the padding columns stay zero), n_hidden elsewhere.
 */
#ifndef EKPNP_MLP_H
#define EKPNP_MLP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ekpnp_mlp_config {
  int32_t n_in;      /* input features (<= 32) */
  int32_t n_hidden;  /* units per hidden layer (multiple of 32) */
  int32_t n_layers;  /* hidden layers */
  int32_t n_out;     /* outputs (multiple of 32) */
  int32_t max_batch; /* largest batch passed to forward/train calls */
  int32_t device;    /* HIP device ordinal */
  uint64_t seed;     /* weight initialisation: N(0, 2/fan_in), biases 0 */
} ekpnp_mlp_config;

typedef struct ekpnp_mlp ekpnp_mlp;

int ekpnp_mlp_create(const ekpnp_mlp_config* cfg, ekpnp_mlp** out);
void ekpnp_mlp_destroy(ekpnp_mlp* m);
const char* ekpnp_mlp_last_error(const ekpnp_mlp* m);
/* number of FP32 values in the parameter vector */
long long ekpnp_mlp_num_params(const ekpnp_mlp* m);
int ekpnp_mlp_get_params(ekpnp_mlp* m, float* out);
int ekpnp_mlp_set_params(ekpnp_mlp* m, const float* in);
/* y[batch][n_out] = MLP(x[batch][n_in]) with BF16 weights and activations */
int ekpnp_mlp_forward(ekpnp_mlp* m, const float* x, int batch, float* y);
/* mean squared error over batch x n_out and its gradient with respect to
 * the parameter vector (no update); grad may be NULL */
int ekpnp_mlp_loss_grad(ekpnp_mlp* m, const float* x, const float* t, int batch, float* loss,
                        float* grad);
/* one Adam step on the mean squared error; returns the loss before the step */
int ekpnp_mlp_train_step(ekpnp_mlp* m, const float* x, const float* t, int batch, float lr,
                         float* loss);
int ekpnp_mlp_set_adam(ekpnp_mlp* m, float beta1, float beta2, float eps, float weight_decay);

#ifdef __cplusplus
}
#endif

#endif /* EKPNP_MLP_H */
