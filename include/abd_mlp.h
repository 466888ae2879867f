/* abd_mlp.h — C ABI of the per-body MLP used by the learned warm start of the ABD Newton step.
 *
 * None of these entry points has a counterpart in the reference (Luke-Skycrawler/colb has no
 * learned component); they serve the learned warm start of the north-star configuration.  The
 * model is fully connected with ReLU between layers, BF16 weights read by the matrix cores with
 * FP32 accumulation, and FP32 master weights for Adam.  Flat parameter layout: for each layer,
 * W (d_l x d_{l+1}, row-major) then b (d_{l+1}).  All calls are synchronous; status codes as in
 * abd_hip.h. */
#ifndef ABD_MLP_H
#define ABD_MLP_H

#ifdef __cplusplus
extern "C" {
#endif

#ifndef ABD_MLP_TYPEDEF
#define ABD_MLP_TYPEDEF
typedef struct AbdMlp AbdMlp;
#endif

/* dims[0..n_layers]: input width, hidden widths, output width (each <= 4096). */
int abd_mlp_create(const int* dims, int n_layers, AbdMlp** out);
void abd_mlp_destroy(AbdMlp* m);
const char* abd_mlp_last_error(const AbdMlp* m);
int abd_mlp_num_params(const AbdMlp* m);
/* Sets the FP32 masters (and their BF16 copies) and resets the Adam moments. */
int abd_mlp_set_params(AbdMlp* m, const float* flat);
int abd_mlp_get_params(const AbdMlp* m, float* flat);
/* Per-output weights w of the loss sum_ij w_j (Y_ij - T_ij)^2 / n. */
int abd_mlp_set_loss_weights(AbdMlp* m, const float* w);
/* Y (n x d_L) = model(X (n x d_0)), host arrays. */
int abd_mlp_forward(AbdMlp* m, const float* X, int n, float* Y);
/* No counterpart in the reference: abd_mlp_forward without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * it).  Same arguments and result as abd_mlp_forward.  The staged input rows and every layer's
 * activations live in one static device array of 64 MiB per GPU (held until the process ends), so n
 * is limited by the model's widths: per row 4 d_0 bytes of input plus 2 d_l (l = 0..L) and 4 d_l
 * (l = 1..L) bytes of activations, widths padded to 16 and each buffer to 256 bytes (about 24 000 rows
 * for 49-128-128-128-12); larger n returns ABD_ERR_CAPACITY with the limit in the message.  Every call
 * on one GPU shares that array: calls that overlap on one GPU race.  The current device must be the
 * one the model was created on. */
int abd_mlp_forward_static(AbdMlp* m, const float* X, int n, float* Y);
/* Loss of one batch and its gradient in the flat layout (no update). */
int abd_mlp_gradient(AbdMlp* m, const float* X, const float* T, int n, float* loss, float* grad_flat);
/* Adam over `steps` minibatches of `batch` rows given by perm[steps*batch]; losses[steps] gets
 * each batch loss before its update. */
int abd_mlp_train(AbdMlp* m, const float* X, const float* T, int n, const int* perm, int steps, int batch, float lr,
                  float beta1, float beta2, float eps, float* losses);
/* C (M x N) = A (M x K) B (K x N) with BF16-rounded inputs on the matrix cores; M, N, K multiples
 * of 16 (checks the operand layout of the matrix instruction). */
int abd_mlp_gemm_test(AbdMlp* m, int M, int N, int K, const float* A, const float* B, float* C);

#ifdef __cplusplus
}
#endif

#endif
