/*
 * rest_lg.h — C ABI of the learned initial guess's network: a per-atom-pair
 * MLP (inputs: pair geometry descriptors; outputs: the pair's density-matrix
 * block in its local frame) trained and evaluated on the device.
 *
 * REST has no learned guess; this is the network behind a new `initial_guess`
 * branch beside "sad" (src/initial_guess/mod.rs:15-72, SAD at :54-64), whose
 * prediction is McWeeny-purified (rscf_mcweeny) and kept only when its guess
 * energy is below SAD's (rscf_guess_step).
 *
 * Network: dims n_in → n_hidden (× n_hidden_layers, SiLU) → n_out (linear).
 * Parameters, in rlg_get_params / rlg_set_params order: for each layer l,
 * W_l [in][out] row-major, then b_l [out]; FP32 master copy on the device.
 * Every GEMM rounds its operands to BF16 (nearest even) and accumulates in
 * FP32 on the BF16 matrix cores.  Loss: masked mean squared error
 * L = Σ mask (y − t)² / Σ mask.  Device pointers (x_dev, t_dev, mask_dev,
 * y_dev) are row-major [batch][features] FP32.  Return codes: 0 success.
 */
#ifndef REST_LG_H
#define REST_LG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rlg_net rlg_net;

enum { RLG_OK = 0, RLG_ERR_ARG = 1, RLG_ERR_HIP = 2, RLG_ERR_NOMEM = 3 };

int rlg_create(int device, size_t n_in, size_t n_hidden, size_t n_hidden_layers, size_t n_out, size_t max_batch,
               rlg_net** out);
int rlg_destroy(rlg_net* net);
const char* rlg_last_error(const rlg_net* net);
size_t rlg_num_params(const rlg_net* net);
/* Run on an external HIP stream (void* hipStream_t, default priority); NULL = own stream. */
int rlg_set_stream(rlg_net* net, void* stream);
int rlg_set_params(rlg_net* net, const float* params_host);
int rlg_get_params(rlg_net* net, float* params_host);
/* y_dev[batch][n_out] = net(x_dev[batch][n_in]); batch <= max_batch. */
int rlg_forward(rlg_net* net, const float* x_dev, size_t batch, float* y_dev);
/* Forward, masked-MSE loss and backward; the gradient stays on the device for
 * rlg_adam_step and is also copied to grad_host (rlg_num_params floats) when
 * non-NULL.  loss may be NULL. */
int rlg_loss_grad(rlg_net* net, const float* x_dev, const float* t_dev, const float* mask_dev, size_t batch,
                  double* loss, float* grad_host);
/* One Adam step with the last gradient (bias-corrected moments, FP32). */
int rlg_adam_step(rlg_net* net, double lr, double beta1, double beta2, double eps);
int rlg_reset_optimizer(rlg_net* net);
/* The network's GEMM on its own: C[m][n] = Σ_k A(m,k) B(k,n) with
 * A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn] (device pointers,
 * operands rounded to BF16, FP32 accumulation). */
int rlg_gemm_bf16(rlg_net* net, const float* A, long sam, long sak, const float* B, long sbk, long sbn, float* C,
                  long ldc, int M, int N, int K);

#ifdef __cplusplus
}
#endif
#endif /* REST_LG_H */
