/*
 * clo_mlp.h - FP32 multilayer perceptron (n_in -> 32 -> 32 -> 1, ReLU) with
 * GPU training (full-batch Adam on the mean squared error) and batched
 * inference.  No counterpart in CL_Ops: it is the learned cost model behind
 * the dispatcher (clo_dispatch.h).
 *
 * Parameter layout (one float array of clo_mlp_num_params() values):
 *   W1[32][n_in], b1[32], W2[32][32], b2[32], w3[32], b3
 * y = w3 . relu(W2 relu(W1 x + b1) + b2) + b3;  loss = mean 0.5 (y - t)^2.
 * All host arrays are row-major float32: X[rows][n_in], t[rows].
 */
#ifndef CLO_MLP_H_
#define CLO_MLP_H_

#include "clo_ccl.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct clo_mlp CloMlp;

/* New network with n_in inputs (1..48): He-uniform weights drawn from a
 * splitmix64 stream of `seed`, zero biases. */
CloMlp* clo_mlp_new(CCLContext* ctx, cl_uint n_in, cl_ulong seed, GError** err);
void clo_mlp_destroy(CloMlp* mlp);
cl_uint clo_mlp_num_params(CloMlp* mlp);
cl_uint clo_mlp_num_inputs(CloMlp* mlp);
cl_uint clo_mlp_num_hidden(CloMlp* mlp);
cl_bool clo_mlp_get_params(CloMlp* mlp, float* params, GError** err);
/* Replaces the parameters and resets the optimiser state. */
cl_bool clo_mlp_set_params(CloMlp* mlp, const float* params, GError** err);
/* Mean loss and its gradient w.r.t. every parameter, computed on the GPU
 * (blocking). */
cl_bool clo_mlp_gradient(CloMlp* mlp, CCLQueue* cq, const float* X, const float* t,
	size_t rows, float* grad, float* loss, GError** err);
/* `steps` full-batch Adam steps (beta1 0.9, beta2 0.999, eps 1e-8) at
 * learning rate lr; *final_loss (optional) = loss after the last step. */
cl_bool clo_mlp_train(CloMlp* mlp, CCLQueue* cq, const float* X, const float* t, size_t rows,
	cl_uint steps, float lr, float* final_loss, GError** err);
/* Batched inference, host arrays (blocking). */
cl_bool clo_mlp_predict(CloMlp* mlp, CCLQueue* cq, const float* X, size_t rows, float* out,
	GError** err);
/* Batched inference on device buffers (asynchronous). */
CCLEvent* clo_mlp_predict_device(CloMlp* mlp, CCLQueue* cq, CCLBuffer* X, size_t rows,
	CCLBuffer* out, GError** err);
/* Versioned weight file: "CLOMLP01", u32 n_in, u32 hidden, u32 count, floats. */
cl_bool clo_mlp_save(CloMlp* mlp, const char* path, GError** err);
CloMlp* clo_mlp_load(CCLContext* ctx, const char* path, GError** err);

#ifdef __cplusplus
}
#endif
#endif
