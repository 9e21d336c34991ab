/*
 * mcpdlp_mlp.h -- C-ABI of the learned warm start's network (libmcpdlp_hip.so).
 *
 * No counterpart in the reference: its PDLP always starts cold at
 * x = proj(0), y = 0 (src/lp/first_order/pdlp.cpp:186-191).  This extension
 * trains and runs a BF16 multilayer perceptron on MFMA (hand-written kernels
 * in markov-zip1_amd/csrc/mlp.hip) that maps instance features to the
 * coefficients of a reduced solution basis; the caller turns the prediction
 * into PdlpOptions::warm_primal / warm_dual.  Cold start remains the fallback
 * and the KKT score decides convergence either way.
 *
 * Network: in -> (hidden, ReLU) x hidden_layers -> out, widths multiples of
 * 32; FP32 master weights (He-uniform init from `seed`, zero biases), BF16
 * copies for the MFMA products, FP32 accumulation; loss = mean squared error;
 * Adam.  Parameters are laid out layer by layer as W (in x out, row-major)
 * then b.  Every call returns MC_OK or an MC_ERR_* code (mcpdlp.h).
 */
#ifndef MCPDLP_MLP_H_
#define MCPDLP_MLP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mcp_mlp mcp_mlp;

int mcp_mlp_create(int32_t device, int32_t in_dim, int32_t hidden, int32_t hidden_layers, int32_t out_dim,
                   int32_t batch_capacity, uint64_t seed, mcp_mlp** out);
void mcp_mlp_destroy(mcp_mlp* m);
const char* mcp_mlp_last_error(void);
int mcp_mlp_param_count(const mcp_mlp* m, int64_t* n);
int mcp_mlp_get_params(mcp_mlp* m, float* params);
int mcp_mlp_set_params(mcp_mlp* m, const float* params);
/* y[n x out] = net(x[n x in]) (row-major host arrays). */
int mcp_mlp_forward(mcp_mlp* m, const float* x, int32_t n, float* y);
/* Allocation-free forward pass.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (tight serving loops, harnesses
 * that forbid hipMalloc).  x_dev (n x in) and y_dev (n x out) are row-major
 * FP32 arrays in device memory on the network's device, ready when the call
 * is made; the call returns once y_dev is written and gives the same values
 * as mcp_mlp_forward.  It calls no hipMalloc/hipFree: the padded input, the
 * pre-activations and the activations live in a static 16 MiB device array
 * of the library (one per device, kept until the process ends), which bounds
 * n to 16 MiB / (8 x widest layer) rounded down to a multiple of 32 (8,192
 * samples for the 256-wide network); a larger n is refused with
 * MC_ERR_INVALID_ARGUMENT.  All calls on one device share that array, so
 * calls that overlap on one device (from several threads) race. */
int mcp_mlp_forward_noalloc(mcp_mlp* m, const float* x_dev, int32_t n, float* y_dev);
/* Loss of one batch (n <= capacity) and, if grad != NULL, the gradient
 * (param_count floats) -- the device backward pass, for gradient checks. */
int mcp_mlp_loss_grad(mcp_mlp* m, const float* x, const float* target, int32_t n, float* loss, float* grad);
/* Adam training: `epochs` passes over n samples in shuffled mini-batches of
 * `batch` (the tail that does not fill a batch is skipped per epoch). */
int mcp_mlp_train(mcp_mlp* m, const float* x, const float* target, int32_t n, int32_t batch, int32_t epochs,
                  float lr, uint64_t seed, float* final_loss, double* train_ms);
/* Test entry for the MFMA GEMM: C[M x N] = op(A) op(B) with A, B rounded to
 * BF16 (A given M x K, or K x M if transpose_a; B given K x N, or N x K). */
int mcp_gemm_bf16_test(int32_t device, int32_t M, int32_t N, int32_t K, const float* A, const float* B,
                       int32_t transpose_a, int32_t transpose_b, float* C);

#ifdef __cplusplus
}
#endif
#endif /* MCPDLP_MLP_H_ */
