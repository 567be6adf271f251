// ft_mlp.h -- the constitutive surrogate network shared by the engine and
// the stand-alone network entry points (no counterpart in the reference).
//
// Network: 12 input features (zero padded to 16) -> 4 hidden layers of 128
// SiLU units -> 6 outputs (padded to 8).  BF16 weights and activations on the
// matrix cores (v_mfma_f32_32x32x16_bf16), FP32 accumulation, FP32 master
// weights and Adam moments.
#pragma once

#include <hip/hip_runtime.h>

namespace ft {

constexpr int kMlpIn = 16;      // features per sample (12 used)
constexpr int kMlpHid = 128;
constexpr int kMlpOut = 8;      // outputs per sample (6 used)
constexpr int kMlpUsedIn = 12;
constexpr int kMlpUsedOut = 6;
constexpr int kMlpLayers = 5;

struct MlpNet;

// All functions return 0 or an FT_ERR_* code (message via the engine's
// error slot).  Device pointers unless stated.
int mlp_create(unsigned long long seed, MlpNet **out);
void mlp_destroy(MlpNet *net);
int mlp_param_count();
int mlp_get_params(MlpNet *net, float *host);
int mlp_set_params(MlpNet *net, const float *host);
// y[n][8] (FP32) = network(x[n][16] BF16, normalised features), one fused
// launch (all layers on chip).
int mlp_infer(MlpNet *net, const __bf16 *x, long n, float *y, hipStream_t st);
// One forward and backward pass over a batch (x BF16 [n][16], t FP32 [n][8]
// normalised targets; mean squared error over the 6 outputs).  Leaves the
// gradient in the net's gradient vector and the loss in *loss (host).
int mlp_gradients(MlpNet *net, const __bf16 *x, const float *t, int n, hipStream_t st, double *loss);
int mlp_copy_gradients(MlpNet *net, float *host);
// One Adam step with the current gradient (lr, beta1 0.9, beta2 0.999).
int mlp_adam(MlpNet *net, float lr, hipStream_t st);
// Batches drawn on the device from a data set (x [n_data][16] BF16,
// t [n_data][8]); `steps` Adam steps with a cosine-decayed rate.
int mlp_train(MlpNet *net, const __bf16 *x, const float *t, long n_data, int steps, int batch, float lr,
              unsigned long long seed, hipStream_t st, double *final_loss);

}  // namespace ft
