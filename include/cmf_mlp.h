/*
 * cmf_mlp.h -- C-ABI of the spin-up surrogate network: a per-cell multilayer
 * perceptron n_in -> H -> H -> H -> n_out (ReLU), trained and evaluated on the
 * GPU with BF16 matrix-core products, FP32 accumulation and FP32 master weights.
 *
 * No counterpart in the reference: CaMa-Flood spins up by repeating the forcing
 * year from the cold start of STORAGE_SEA_SURFACE (src/cmf_ctrl_vars_mod.F90:129-168;
 * gosh/test1-glb_15min.sh:311-345). The network predicts the equilibrium river and
 * floodplain storage that this spin-up converges to, as a warm start.
 *
 * Parameters are one packed FP32 vector: for layer l = 0..3, W_l (d_in x d_out,
 * row-major, so Z = X W_l) followed by b_l (d_out). Rows of X and Y are cells.
 * Every function returns 0 (or a count) on success, a negative CMF_MLP_E* code on
 * failure; cmf_mlp_last_error() gives the message.
 */
#ifndef CMF_MLP_H
#define CMF_MLP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define CMF_MLP_EINVAL  -1
#define CMF_MLP_EHIP    -2
#define CMF_MLP_ESTATE  -3

typedef struct cmf_mlp* cmf_mlp_handle;

int cmf_mlp_create(int32_t n_in, int32_t hidden, int32_t n_out, int32_t max_batch, cmf_mlp_handle* out);
void cmf_mlp_destroy(cmf_mlp_handle h);
const char* cmf_mlp_last_error(void);
int64_t cmf_mlp_param_count(cmf_mlp_handle h);

/* Loss of cmf_mlp_grad / cmf_mlp_train: mean of w (y - t)^2 with w = w_over where
 * y > t and 1 elsewhere (w_over = 1, the default, is the mean squared error; a
 * larger w_over trains a low-biased predictor). */
int cmf_mlp_set_loss(cmf_mlp_handle h, float w_over);

/* Load FP32 parameters (resets the Adam moments); read parameters and/or the
 * gradient of the last cmf_mlp_grad / training batch (either pointer may be NULL). */
int cmf_mlp_set_params(cmf_mlp_handle h, const float* params);
int cmf_mlp_get_params(cmf_mlp_handle h, float* params, float* grads);

/* Inference: y[n][n_out] = net(x[n][n_in]), host arrays, processed in batches. */
int cmf_mlp_forward(cmf_mlp_handle h, const float* x, int64_t n, float* y);

/* Forward and backward pass of one batch (rows <= max_batch) against targets t
 * under the mean squared error; the gradient stays on the device (read it with
 * cmf_mlp_get_params), *loss receives the batch loss. */
int cmf_mlp_grad(cmf_mlp_handle h, const float* x, const float* t, int32_t rows, float* loss);

/* Keep a training set resident in HBM, then run Adam over the rows listed in
 * order[n_order] in consecutive batches of `batch` rows (a final partial batch is
 * skipped). step0 = Adam steps taken before. Returns the number of batches run;
 * *loss_out = mean batch loss. */
int cmf_mlp_set_data(cmf_mlp_handle h, const float* x, const float* t, int64_t n);
int cmf_mlp_train(cmf_mlp_handle h, const int64_t* order, int64_t n_order, int32_t batch, float lr,
                  float beta1, float beta2, float eps, int32_t step0, float* loss_out);

/* The matrix-core product on its own (for tests): C[M][N] = op(A) op(B) with BF16
 * bit patterns in A and B; op(A)(m,k) = ta ? A[k*lda+m] : A[m*lda+k],
 * op(B)(k,n) = tb ? B[n*ldb+k] : B[k*ldb+n]. */
int cmf_mlp_gemm_test(int32_t M, int32_t N, int32_t K, const uint16_t* A, int32_t lda, int32_t ta,
                      const uint16_t* B, int32_t ldb, int32_t tb, float* C);

#ifdef __cplusplus
}
#endif
#endif
