// mlp.hip -- training and inference kernels of the lambda* model (a
// two-hidden-layer ReLU MLP, FP64).  The data are small (a few thousand
// samples, 64 hidden units): one thread per output element, sums in a fixed
// order, so results repeat bit for bit and follow the CPU restatement's
// order of operations up to the reassociation of each dot product.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "../../include/sdpa_mlp.h"

#define MLP_CHECK(x)                              \
  do {                                            \
    if ((x) != hipSuccess) return -100;           \
  } while (0)

struct sdpa_mlp {
  int device = 0, nin = 0, nh = 0;
  long np = 0;
  double* theta = nullptr;  // parameters
  double* grad = nullptr;
  double* m1 = nullptr;     // Adam moments
  double* m2 = nullptr;
  long adam_t = 0;
  int n = 0, cap = 0;       // samples
  double* X = nullptr;      // n x nin
  double* t = nullptr;      // n
  double* Z1 = nullptr;     // n x nh pre-activations, layer 1
  double* Z2 = nullptr;     // n x nh, layer 2
  double* y = nullptr;      // n
  double* dZ = nullptr;     // n x nh (layer-2 then layer-1 deltas)
  double* dZ1 = nullptr;    // n x nh
  double* dy = nullptr;     // n
  double* loss = nullptr;   // 1
  hipStream_t s = nullptr;
};

namespace {

// offsets of the packed parameters
struct Lay {
  long W1, b1, W2, b2, W3, b3, np;
};
Lay layout(int nin, int nh) {
  Lay L;
  L.W1 = 0;
  L.b1 = L.W1 + (long)nh * nin;
  L.W2 = L.b1 + nh;
  L.b2 = L.W2 + (long)nh * nh;
  L.W3 = L.b2 + nh;
  L.b3 = L.W3 + nh;
  L.np = L.b3 + 1;
  return L;
}

// Z[s, o] = b[o] + sum_i W[o, i] A[s, i], A = relu(Ain) if relu_in
__global__ void k_layer_fwd(const double* __restrict__ A, int nin, int relu_in,
                            const double* __restrict__ W, const double* __restrict__ b,
                            double* __restrict__ Z, int n, int nout) {
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (long)n * nout) return;
  const int s = (int)(id / nout), o = (int)(id % nout);
  double z = b[o];
  for (int i = 0; i < nin; ++i) {
    double a = A[(long)s * nin + i];
    if (relu_in) a = a > 0.0 ? a : 0.0;
    z += W[(long)o * nin + i] * a;
  }
  Z[id] = z;
}

// dy = 2 (y - t) / n; loss = (1/n) sum (y - t)^2 (one work-group, fixed order)
__global__ void k_out_grad(const double* __restrict__ y, const double* __restrict__ t, int n,
                           double* __restrict__ dy, double* __restrict__ loss) {
  __shared__ double part[256];
  double acc = 0.0;
  for (int s = threadIdx.x; s < n; s += 256) {
    const double e = y[s] - t[s];
    dy[s] = 2.0 * e / n;
    acc += e * e;
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double l = 0.0;
    for (int k = 0; k < 256; ++k) l += part[k];
    *loss = l / n;
  }
}

// dA[s, i] = (Aprev[s, i] > 0) * sum_o dZ[s, o] W[o, i]   (ReLU of the layer below)
__global__ void k_layer_bwd(const double* __restrict__ dZ, int nout, const double* __restrict__ W,
                            const double* __restrict__ Zprev, double* __restrict__ dA, int n,
                            int nin) {
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (long)n * nin) return;
  const int s = (int)(id / nin), i = (int)(id % nin);
  double g = 0.0;
  for (int o = 0; o < nout; ++o) g += dZ[(long)s * nout + o] * W[(long)o * nin + i];
  dA[id] = Zprev[id] > 0.0 ? g : 0.0;
}

// gW[o, i] = sum_s dZ[s, o] A[s, i] (A = relu(Ain) if relu_in); gb[o] = sum_s dZ[s, o]
// thread (o, i), i == nin computes gb[o]
__global__ void k_wgrad(const double* __restrict__ dZ, int nout, const double* __restrict__ A,
                        int nin, int relu_in, int n, double* __restrict__ gW,
                        double* __restrict__ gb) {
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (long)nout * (nin + 1)) return;
  const int o = (int)(id / (nin + 1)), i = (int)(id % (nin + 1));
  double g = 0.0;
  if (i < nin) {
    for (int s = 0; s < n; ++s) {
      double a = A[(long)s * nin + i];
      if (relu_in) a = a > 0.0 ? a : 0.0;
      g += dZ[(long)s * nout + o] * a;
    }
    gW[(long)o * nin + i] = g;
  } else {
    for (int s = 0; s < n; ++s) g += dZ[(long)s * nout + o];
    gb[o] = g;
  }
}

__global__ void k_adam(double* __restrict__ th, const double* __restrict__ g,
                       double* __restrict__ m1, double* __restrict__ m2, long np, double lr,
                       double c1, double c2) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const double b1 = 0.9, b2 = 0.999, eps = 1e-8;
  const double a = b1 * m1[k] + (1.0 - b1) * g[k];
  const double v = b2 * m2[k] + (1.0 - b2) * g[k] * g[k];
  m1[k] = a;
  m2[k] = v;
  th[k] -= lr * (a / c1) / (sqrt(v / c2) + eps);
}

inline dim3 grid_for(long n) { return dim3((unsigned)((n + 255) / 256)); }

int ensure_samples(sdpa_mlp* h, int n) {
  if (n <= h->cap) return 0;
  for (double** p : {&h->X, &h->t, &h->Z1, &h->Z2, &h->y, &h->dZ, &h->dZ1, &h->dy})
    if (*p) (void)hipFree(*p), *p = nullptr;
  const size_t nh = h->nh;
  MLP_CHECK(hipMalloc(&h->X, sizeof(double) * n * h->nin));
  MLP_CHECK(hipMalloc(&h->t, sizeof(double) * n));
  MLP_CHECK(hipMalloc(&h->Z1, sizeof(double) * n * nh));
  MLP_CHECK(hipMalloc(&h->Z2, sizeof(double) * n * nh));
  MLP_CHECK(hipMalloc(&h->y, sizeof(double) * n));
  MLP_CHECK(hipMalloc(&h->dZ, sizeof(double) * n * nh));
  MLP_CHECK(hipMalloc(&h->dZ1, sizeof(double) * n * nh));
  MLP_CHECK(hipMalloc(&h->dy, sizeof(double) * n));
  h->cap = n;
  return 0;
}

void forward(sdpa_mlp* h, int n) {
  const Lay L = layout(h->nin, h->nh);
  const int nh = h->nh;
  hipLaunchKernelGGL(k_layer_fwd, grid_for((long)n * nh), dim3(256), 0, h->s, h->X, h->nin, 0,
                     h->theta + L.W1, h->theta + L.b1, h->Z1, n, nh);
  hipLaunchKernelGGL(k_layer_fwd, grid_for((long)n * nh), dim3(256), 0, h->s, h->Z1, nh, 1,
                     h->theta + L.W2, h->theta + L.b2, h->Z2, n, nh);
  hipLaunchKernelGGL(k_layer_fwd, grid_for(n), dim3(256), 0, h->s, h->Z2, nh, 1, h->theta + L.W3,
                     h->theta + L.b3, h->y, n, 1);
}

void backward(sdpa_mlp* h, int n) {
  const Lay L = layout(h->nin, h->nh);
  const int nh = h->nh;
  hipLaunchKernelGGL(k_out_grad, dim3(1), dim3(256), 0, h->s, h->y, h->t, n, h->dy, h->loss);
  // layer 3: gW3, gb3 from dy and relu(Z2)
  hipLaunchKernelGGL(k_wgrad, grid_for(1L * (nh + 1)), dim3(256), 0, h->s, h->dy, 1, h->Z2, nh, 1,
                     n, h->grad + L.W3, h->grad + L.b3);
  // dZ2 = relu'(Z2) * (dy W3)
  hipLaunchKernelGGL(k_layer_bwd, grid_for((long)n * nh), dim3(256), 0, h->s, h->dy, 1,
                     h->theta + L.W3, h->Z2, h->dZ, n, nh);
  hipLaunchKernelGGL(k_wgrad, grid_for((long)nh * (nh + 1)), dim3(256), 0, h->s, h->dZ, nh, h->Z1,
                     nh, 1, n, h->grad + L.W2, h->grad + L.b2);
  // dZ1 = relu'(Z1) * (dZ2 W2)
  hipLaunchKernelGGL(k_layer_bwd, grid_for((long)n * nh), dim3(256), 0, h->s, h->dZ, nh,
                     h->theta + L.W2, h->Z1, h->dZ1, n, nh);
  hipLaunchKernelGGL(k_wgrad, grid_for((long)nh * (h->nin + 1)), dim3(256), 0, h->s, h->dZ1, nh,
                     h->X, h->nin, 0, n, h->grad + L.W1, h->grad + L.b1);
}

}  // namespace

extern "C" {

int sdpa_mlp_create(int device, int nin, int nh, sdpa_mlp** out) {
  if (!out || nin <= 0 || nh <= 0) return -1;
  *out = nullptr;
  MLP_CHECK(hipSetDevice(device));
  sdpa_mlp* h = new sdpa_mlp();
  h->device = device;
  h->nin = nin;
  h->nh = nh;
  h->np = layout(nin, nh).np;
  if (hipStreamCreate(&h->s) != hipSuccess || hipMalloc(&h->theta, sizeof(double) * h->np) ||
      hipMalloc(&h->grad, sizeof(double) * h->np) || hipMalloc(&h->m1, sizeof(double) * h->np) ||
      hipMalloc(&h->m2, sizeof(double) * h->np) || hipMalloc(&h->loss, sizeof(double))) {
    sdpa_mlp_destroy(h);
    return -100;
  }
  (void)hipMemset(h->theta, 0, sizeof(double) * h->np);
  (void)hipMemset(h->m1, 0, sizeof(double) * h->np);
  (void)hipMemset(h->m2, 0, sizeof(double) * h->np);
  *out = h;
  return 0;
}

int sdpa_mlp_destroy(sdpa_mlp* h) {
  if (!h) return 0;
  if (h->s) (void)hipStreamSynchronize(h->s);
  for (double* p : {h->theta, h->grad, h->m1, h->m2, h->X, h->t, h->Z1, h->Z2, h->y, h->dZ, h->dZ1,
                    h->dy, h->loss})
    if (p) (void)hipFree(p);
  if (h->s) (void)hipStreamDestroy(h->s);
  delete h;
  return 0;
}

long sdpa_mlp_nparams(const sdpa_mlp* h) { return h ? h->np : -1; }

int sdpa_mlp_set_params(sdpa_mlp* h, const double* theta, long n) {
  if (!h || !theta || n != h->np) return -1;
  MLP_CHECK(hipMemcpy(h->theta, theta, sizeof(double) * n, hipMemcpyHostToDevice));
  MLP_CHECK(hipMemset(h->m1, 0, sizeof(double) * n));
  MLP_CHECK(hipMemset(h->m2, 0, sizeof(double) * n));
  h->adam_t = 0;
  return 0;
}

int sdpa_mlp_get_params(sdpa_mlp* h, double* theta, long n) {
  if (!h || !theta || n != h->np) return -1;
  MLP_CHECK(hipStreamSynchronize(h->s));
  MLP_CHECK(hipMemcpy(theta, h->theta, sizeof(double) * n, hipMemcpyDeviceToHost));
  return 0;
}

int sdpa_mlp_set_data(sdpa_mlp* h, const double* X, const double* t, int n) {
  if (!h || !X || !t || n <= 0) return -1;
  if (int e = ensure_samples(h, n)) return e;
  MLP_CHECK(hipMemcpy(h->X, X, sizeof(double) * n * h->nin, hipMemcpyHostToDevice));
  MLP_CHECK(hipMemcpy(h->t, t, sizeof(double) * n, hipMemcpyHostToDevice));
  h->n = n;
  return 0;
}

int sdpa_mlp_loss_grad(sdpa_mlp* h, double* loss, double* grad) {
  if (!h || h->n <= 0) return -1;
  forward(h, h->n);
  backward(h, h->n);
  MLP_CHECK(hipGetLastError());
  MLP_CHECK(hipStreamSynchronize(h->s));
  if (loss) MLP_CHECK(hipMemcpy(loss, h->loss, sizeof(double), hipMemcpyDeviceToHost));
  if (grad) MLP_CHECK(hipMemcpy(grad, h->grad, sizeof(double) * h->np, hipMemcpyDeviceToHost));
  return 0;
}

int sdpa_mlp_train(sdpa_mlp* h, int steps, double lr, double* loss) {
  if (!h || h->n <= 0 || steps < 0) return -1;
  for (int k = 0; k < steps; ++k) {
    forward(h, h->n);
    backward(h, h->n);
    ++h->adam_t;
    const double c1 = 1.0 - std::pow(0.9, (double)h->adam_t);
    const double c2 = 1.0 - std::pow(0.999, (double)h->adam_t);
    hipLaunchKernelGGL(k_adam, grid_for(h->np), dim3(256), 0, h->s, h->theta, h->grad, h->m1,
                       h->m2, h->np, lr, c1, c2);
  }
  MLP_CHECK(hipGetLastError());
  MLP_CHECK(hipStreamSynchronize(h->s));
  if (loss) MLP_CHECK(hipMemcpy(loss, h->loss, sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int sdpa_mlp_predict(sdpa_mlp* h, const double* X, int n, double* y) {
  if (!h || !X || !y || n <= 0) return -1;
  // prediction uses its own copy of the sample buffers: keep the training
  // set by restoring n afterwards (the arrays are reallocated only to grow)
  const int keep_n = h->n;
  std::vector<double> keepX, keepT;
  if (keep_n > 0 && n > 0) {
    keepX.resize((size_t)keep_n * h->nin);
    keepT.resize(keep_n);
    MLP_CHECK(hipMemcpy(keepX.data(), h->X, sizeof(double) * keepX.size(), hipMemcpyDeviceToHost));
    MLP_CHECK(hipMemcpy(keepT.data(), h->t, sizeof(double) * keep_n, hipMemcpyDeviceToHost));
  }
  if (int e = ensure_samples(h, n)) return e;
  MLP_CHECK(hipMemcpy(h->X, X, sizeof(double) * n * h->nin, hipMemcpyHostToDevice));
  forward(h, n);
  MLP_CHECK(hipGetLastError());
  MLP_CHECK(hipStreamSynchronize(h->s));
  MLP_CHECK(hipMemcpy(y, h->y, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (keep_n > 0) {
    if (int e = ensure_samples(h, keep_n)) return e;
    MLP_CHECK(hipMemcpy(h->X, keepX.data(), sizeof(double) * keepX.size(), hipMemcpyHostToDevice));
    MLP_CHECK(hipMemcpy(h->t, keepT.data(), sizeof(double) * keep_n, hipMemcpyHostToDevice));
  }
  h->n = keep_n;
  return 0;
}

}  // extern "C"
