/* sdpa_mlp.h -- C ABI of the learned initial-point scale (lambda*) model:
 * a small MLP (nin features -> nh ReLU -> nh ReLU -> 1) whose training and
 * inference run in this library's own HIP kernels, in FP64.
 *
 * No counterpart in the reference: it feeds the reference's initial-point
 * hook (Solutions::initialize, sdpa_dataset.cpp:47-74, X0 = Z0 = lambda* I;
 * per-block values go in through the init point, sdpa_solve.cpp:61-72).  The
 * default lambda* of param.sdpa stays the fallback and the prediction never
 * changes the stopping rules.
 *
 * Parameters, packed in this order (theta, length sdpa_mlp_nparams):
 *   W1 (nh x nin, row-major), b1 (nh), W2 (nh x nh, row-major), b2 (nh),
 *   W3 (nh), b3 (1).
 * Loss: mean squared error over the samples, (1/n) sum (y_s - t_s)^2.
 * All functions return 0 on success, < 0 on error. */
#ifndef SDPA_MLP_H
#define SDPA_MLP_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct sdpa_mlp sdpa_mlp;

int sdpa_mlp_create(int device, int nin, int nh, sdpa_mlp** out);
int sdpa_mlp_destroy(sdpa_mlp* h);
long sdpa_mlp_nparams(const sdpa_mlp* h);
int sdpa_mlp_set_params(sdpa_mlp* h, const double* theta, long n);
int sdpa_mlp_get_params(sdpa_mlp* h, double* theta, long n);
/* Training data on the device: X (n x nin, row-major), targets t (n). */
int sdpa_mlp_set_data(sdpa_mlp* h, const double* X, const double* t, int n);
/* Loss and its gradient (length nparams) at the current parameters over the
 * data set by sdpa_mlp_set_data. */
int sdpa_mlp_loss_grad(sdpa_mlp* h, double* loss, double* grad);
/* `steps` full-batch Adam steps (beta1 0.9, beta2 0.999, eps 1e-8; the
 * moments persist across calls until sdpa_mlp_set_params); *loss receives
 * the loss before the last step. */
int sdpa_mlp_train(sdpa_mlp* h, int steps, double lr, double* loss);
/* y = f(X) for n samples (X row-major, host arrays). */
int sdpa_mlp_predict(sdpa_mlp* h, const double* X, int n, double* y);

#ifdef __cplusplus
}
#endif
#endif /* SDPA_MLP_H */
