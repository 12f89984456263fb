// C ABI of the surrogate network and its training samples (include/ppkmhd_amd.h, ppk_mlp_*,
// ppk_faceset_*).  Every call is synchronous: it returns once its GPU work has finished, so a
// solver that uses the network on its own stream always sees a complete parameter set.  A call
// that rewrites the parameters first waits for all work on the device (solver steps on other
// streams may still be reading the packed weights it is about to overwrite).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "mlp_kernels.hpp"
#include "ppkmhd_amd.h"

namespace {

template <class H>
int nfail(H *h, int code, const std::string &msg) {
  if (h)
    h->err = msg;
  else
    ppk::set_thread_error(msg);
  return code;
}

#define NHIP(h, expr)                                                                      \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess)                                                                  \
      return nfail((h), PPK_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));   \
  } while (0)

unsigned long long splitmix(unsigned long long &s) {
  s += 0x9E3779B97F4A7C15ull;
  unsigned long long z = s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double uniform(unsigned long long &s) { return (splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

void free_net(ppk_mlp_t *n) {
  for (void *p : {(void *)n->P, (void *)n->m, (void *)n->v, (void *)n->packed,
                  (void *)n->scratch, (void *)n->partials})
    if (p) (void)hipFree(p);
  if (n->stream) (void)hipStreamDestroy(n->stream);
}

// wait for every stream of the device: a solver step in flight may read the packed weights
int quiesce_readers(ppk_mlp_t *n) {
  NHIP(n, hipDeviceSynchronize());
  return PPK_OK;
}

int upload_params(ppk_mlp_t *n, const float *P) {
  int rc = quiesce_readers(n);
  if (rc != PPK_OK) return rc;
  NHIP(n, hipMemcpyAsync(n->P, P, sizeof(float) * ppk::kMlpParams, hipMemcpyHostToDevice,
                         n->stream));
  NHIP(n, ppk::mlp::pack(n->P, n->packed, n->stream));
  NHIP(n, hipStreamSynchronize(n->stream));
  return PPK_OK;
}

int ensure_scratch(ppk_mlp_t *n, long long B) {
  if (n->scratch_batch >= B) return PPK_OK;
  if (n->scratch) (void)hipFree(n->scratch);
  if (n->partials) (void)hipFree(n->partials);
  n->scratch = nullptr;
  n->partials = nullptr;
  n->scratch_batch = 0;
  NHIP(n, hipMalloc(&n->scratch, sizeof(uint16_t) * ppk::mlp::fb_scratch_elems(B)));
  NHIP(n, hipMalloc(&n->partials, sizeof(float) * (size_t)(B / 256) * ppk::kMlpTrained));
  n->scratch_batch = B;
  return PPK_OK;
}

}  // namespace

extern "C" {

int ppk_mlp_param_count(void) { return ppk::kMlpParams; }

int ppk_mlp_create(uint64_t seed, ppk_mlp_t **net) {
  if (!net) return nfail<ppk_mlp_t>(nullptr, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  *net = nullptr;
  ppk_mlp_t *n = new (std::nothrow) ppk_mlp_t();
  if (!n) return nfail<ppk_mlp_t>(nullptr, PPK_ERR_OUT_OF_MEMORY, "host allocation failed");
  auto bail = [&](int code, const std::string &msg) {
    free_net(n);
    delete n;
    return nfail<ppk_mlp_t>(nullptr, code, msg);
  };
  if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(PPK_ERR_HIP, "hipStreamCreateWithFlags failed");
  if (hipMalloc(&n->P, sizeof(float) * ppk::kMlpParams) != hipSuccess ||
      hipMalloc(&n->m, sizeof(float) * ppk::kMlpTrained) != hipSuccess ||
      hipMalloc(&n->v, sizeof(float) * ppk::kMlpTrained) != hipSuccess ||
      hipMalloc(&n->packed, sizeof(ppk::MlpPacked)) != hipSuccess)
    return bail(PPK_ERR_OUT_OF_MEMORY, "device allocation failed");
  if (hipMemsetAsync(n->m, 0, sizeof(float) * ppk::kMlpTrained, n->stream) != hipSuccess ||
      hipMemsetAsync(n->v, 0, sizeof(float) * ppk::kMlpTrained, n->stream) != hipSuccess)
    return bail(PPK_ERR_HIP, "hipMemsetAsync failed");
  // He-uniform weights (bound sqrt(6 / fan_in)), zero biases, identity standardisation
  std::vector<float> P(ppk::kMlpParams, 0.f);
  unsigned long long s = seed;
  auto fill = [&](int off, int count, int fan_in, double gain) {
    const double bound = gain * std::sqrt(6.0 / fan_in);
    for (int i = 0; i < count; ++i) P[off + i] = (float)((2.0 * uniform(s) - 1.0) * bound);
  };
  fill(ppk::kMlpW1, 128 * 16, 16, 1.0);
  fill(ppk::kMlpW2, 128 * 128, 128, 1.0);
  fill(ppk::kMlpW3, 8 * 128, 128, 0.1);
  for (int i = 0; i < 16; ++i) P[ppk::kMlpIsigIn + i] = 1.f;
  for (int i = 0; i < 8; ++i) P[ppk::kMlpSigOut + i] = 1.f;
  int rc = upload_params(n, P.data());
  if (rc != PPK_OK) return bail(rc, n->err);
  *net = n;
  return PPK_OK;
}

void ppk_mlp_destroy(ppk_mlp_t *net) {
  if (!net) return;
  (void)hipStreamSynchronize(net->stream);
  free_net(net);
  delete net;
}

const char *ppk_mlp_last_error(const ppk_mlp_t *net) {
  return net ? net->err.c_str() : ppk::thread_error();
}

int ppk_mlp_get_params(ppk_mlp_t *net, float *P, size_t count) {
  if (!net || !P) return nfail(net, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (count != (size_t)ppk::kMlpParams)
    return nfail(net, PPK_ERR_INVALID_ARGUMENT, "count must be ppk_mlp_param_count()");
  NHIP(net, hipMemcpyAsync(P, net->P, sizeof(float) * count, hipMemcpyDeviceToHost, net->stream));
  NHIP(net, hipStreamSynchronize(net->stream));
  return PPK_OK;
}

int ppk_mlp_set_params(ppk_mlp_t *net, const float *P, size_t count) {
  if (!net || !P) return nfail(net, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (count != (size_t)ppk::kMlpParams)
    return nfail(net, PPK_ERR_INVALID_ARGUMENT, "count must be ppk_mlp_param_count()");
  for (int i = 0; i < 16; ++i)
    if (!(P[ppk::kMlpIsigIn + i] > 0.f))
      return nfail(net, PPK_ERR_INVALID_ARGUMENT, "isig_in must be positive");
  for (int i = 0; i < 8; ++i)
    if (!(P[ppk::kMlpSigOut + i] > 0.f))
      return nfail(net, PPK_ERR_INVALID_ARGUMENT, "sig_out must be positive");
  return upload_params(net, P);
}

int ppk_faceset_create(size_t capacity, ppk_faceset_t **set) {
  if (!set) return nfail<ppk_faceset_t>(nullptr, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  *set = nullptr;
  if (capacity < 1) return nfail<ppk_faceset_t>(nullptr, PPK_ERR_INVALID_ARGUMENT, "capacity must be >= 1");
  ppk_faceset_t *f = new (std::nothrow) ppk_faceset_t();
  if (!f) return nfail<ppk_faceset_t>(nullptr, PPK_ERR_OUT_OF_MEMORY, "host allocation failed");
  if (hipMalloc(&f->X, sizeof(float) * ppk::kMlpIn * capacity) != hipSuccess ||
      hipMalloc(&f->T, sizeof(float) * ppk::kMlpOut * capacity) != hipSuccess) {
    if (f->X) (void)hipFree(f->X);
    if (f->T) (void)hipFree(f->T);
    delete f;
    return nfail<ppk_faceset_t>(nullptr, PPK_ERR_OUT_OF_MEMORY, "device allocation failed");
  }
  f->capacity = capacity;
  *set = f;
  return PPK_OK;
}

void ppk_faceset_destroy(ppk_faceset_t *set) {
  if (!set) return;
  (void)hipDeviceSynchronize();
  if (set->X) (void)hipFree(set->X);
  if (set->T) (void)hipFree(set->T);
  delete set;
}

const char *ppk_faceset_last_error(const ppk_faceset_t *set) {
  return set ? set->err.c_str() : ppk::thread_error();
}

int ppk_faceset_size(const ppk_faceset_t *set, size_t *n) {
  if (!set || !n) return nfail<ppk_faceset_t>(nullptr, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  *n = set->n;
  return PPK_OK;
}

int ppk_faceset_clear(ppk_faceset_t *set) {
  if (!set) return nfail<ppk_faceset_t>(nullptr, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  set->n = 0;
  return PPK_OK;
}

int ppk_faceset_append(ppk_faceset_t *set, const float *X, const float *T, size_t n) {
  if (!set || !X || !T) return nfail(set, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (n > set->capacity - set->n) return nfail(set, PPK_ERR_INVALID_ARGUMENT, "face set is full");
  NHIP(set, hipMemcpy(set->X + set->n * ppk::kMlpIn, X, sizeof(float) * ppk::kMlpIn * n,
                      hipMemcpyHostToDevice));
  NHIP(set, hipMemcpy(set->T + set->n * ppk::kMlpOut, T, sizeof(float) * ppk::kMlpOut * n,
                      hipMemcpyHostToDevice));
  set->n += n;
  return PPK_OK;
}

int ppk_faceset_read(ppk_faceset_t *set, size_t first, size_t n, float *X, float *T) {
  if (!set) return nfail<ppk_faceset_t>(nullptr, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (first > set->n || n > set->n - first)
    return nfail(set, PPK_ERR_INVALID_ARGUMENT, "range past the end of the face set");
  if (X)
    NHIP(set, hipMemcpy(X, set->X + first * ppk::kMlpIn, sizeof(float) * ppk::kMlpIn * n,
                        hipMemcpyDeviceToHost));
  if (T)
    NHIP(set, hipMemcpy(T, set->T + first * ppk::kMlpOut, sizeof(float) * ppk::kMlpOut * n,
                        hipMemcpyDeviceToHost));
  return PPK_OK;
}

int ppk_mlp_standardize(ppk_mlp_t *net, const ppk_faceset_t *set) {
  if (!net || !set) return nfail(net, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (set->n < 2) return nfail(net, PPK_ERR_INVALID_ARGUMENT, "face set needs >= 2 samples");
  const int nb = 256;
  double *parts = nullptr;
  NHIP(net, hipMalloc(&parts, sizeof(double) * nb * 48));
  std::vector<double> h(nb * 48);
  hipError_t e = ppk::mlp::moments(set->X, set->T, (long long)set->n, parts, nb, net->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.data(), parts, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                       net->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(net->stream);
  (void)hipFree(parts);
  if (e != hipSuccess) return nfail(net, PPK_ERR_HIP, std::string("moments: ") + hipGetErrorString(e));
  std::vector<float> P(ppk::kMlpParams);
  int rc = ppk_mlp_get_params(net, P.data(), P.size());
  if (rc != PPK_OK) return rc;
  const double n = (double)set->n;
  for (int c = 0; c < 24; ++c) {
    double s = 0, s2 = 0;
    for (int b = 0; b < nb; ++b) {
      s += h[(b * 24 + c) * 2];
      s2 += h[(b * 24 + c) * 2 + 1];
    }
    const double mu = s / n, var = std::max(0.0, s2 / n - mu * mu);
    const double sd = std::sqrt(var);
    const bool flat = !(sd > 1e-6 * std::max(1.0, std::fabs(mu)));
    if (c < 16) {
      P[ppk::kMlpMuIn + c] = (float)mu;
      P[ppk::kMlpIsigIn + c] = flat ? 1.f : (float)(1.0 / sd);
    } else {
      P[ppk::kMlpMuOut + c - 16] = (float)mu;
      P[ppk::kMlpSigOut + c - 16] = flat ? 1.f : (float)sd;
    }
  }
  return upload_params(net, P.data());
}

int ppk_mlp_train(ppk_mlp_t *net, const ppk_faceset_t *set, int steps, int batch, float lr,
                  uint64_t seed) {
  if (!net || !set) return nfail(net, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (steps < 0 || batch < 256 || batch % 256 || !(lr > 0.f))
    return nfail(net, PPK_ERR_INVALID_ARGUMENT,
                 "steps >= 0, batch a positive multiple of 256 and lr > 0 are required");
  if (set->n < 1) return nfail(net, PPK_ERR_INVALID_ARGUMENT, "face set is empty");
  int rc = ensure_scratch(net, batch);
  if (rc == PPK_OK) rc = quiesce_readers(net);
  if (rc != PPK_OK) return rc;
  for (int i = 0; i < steps; ++i) {
    NHIP(net, ppk::mlp::train_step(net->packed, net->P, net->m, net->v, set->X, set->T, nullptr,
                                   (long long)set->n, seed, net->adam_step, batch, net->scratch,
                                   net->partials, nullptr, lr, net->stream));
    ++net->adam_step;
  }
  NHIP(net, hipStreamSynchronize(net->stream));
  return PPK_OK;
}

int ppk_mlp_gradient(ppk_mlp_t *net, const float *X, const float *T, size_t n, float *grad,
                     size_t count) {
  if (!net || !X || !T || !grad) return nfail(net, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (n < 256 || n % 256) return nfail(net, PPK_ERR_INVALID_ARGUMENT, "n must be a multiple of 256");
  if (count != (size_t)ppk::kMlpTrained)
    return nfail(net, PPK_ERR_INVALID_ARGUMENT, "count must be the trained-parameter count");
  int rc = ensure_scratch(net, (long long)n);
  if (rc == PPK_OK) rc = quiesce_readers(net);
  if (rc != PPK_OK) return rc;
  float *dX = nullptr, *dT = nullptr, *dG = nullptr;
  long long *idx = nullptr;
  std::vector<long long> hidx(n);
  for (size_t i = 0; i < n; ++i) hidx[i] = (long long)i;
  hipError_t e = hipMalloc(&dX, sizeof(float) * 16 * n);
  if (e == hipSuccess) e = hipMalloc(&dT, sizeof(float) * 8 * n);
  if (e == hipSuccess) e = hipMalloc(&dG, sizeof(float) * ppk::kMlpTrained);
  if (e == hipSuccess) e = hipMalloc(&idx, sizeof(long long) * n);
  if (e == hipSuccess) e = hipMemcpy(dX, X, sizeof(float) * 16 * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dT, T, sizeof(float) * 8 * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(idx, hidx.data(), sizeof(long long) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = ppk::mlp::train_step(net->packed, net->P, net->m, net->v, dX, dT, idx, (long long)n, 0, 0,
                             (long long)n, net->scratch, net->partials, dG, 0.f, net->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(net->stream);
  if (e == hipSuccess) e = hipMemcpy(grad, dG, sizeof(float) * count, hipMemcpyDeviceToHost);
  for (void *p : {(void *)dX, (void *)dT, (void *)dG, (void *)idx})
    if (p) (void)hipFree(p);
  if (e != hipSuccess) return nfail(net, PPK_ERR_HIP, std::string("gradient: ") + hipGetErrorString(e));
  return PPK_OK;
}

int ppk_mlp_infer(ppk_mlp_t *net, const float *X, size_t n, float *Y) {
  if (!net || !X || !Y) return nfail(net, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (n == 0) return PPK_OK;
  float *dX = nullptr, *dY = nullptr;
  hipError_t e = hipMalloc(&dX, sizeof(float) * 16 * n);
  if (e == hipSuccess) e = hipMalloc(&dY, sizeof(float) * 8 * n);
  if (e == hipSuccess) e = hipMemcpy(dX, X, sizeof(float) * 16 * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = ppk::mlp::infer(net->packed, dX, nullptr, (long long)n, dY, nullptr, net->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(net->stream);
  if (e == hipSuccess) e = hipMemcpy(Y, dY, sizeof(float) * 8 * n, hipMemcpyDeviceToHost);
  if (dX) (void)hipFree(dX);
  if (dY) (void)hipFree(dY);
  if (e != hipSuccess) return nfail(net, PPK_ERR_HIP, std::string("infer: ") + hipGetErrorString(e));
  return PPK_OK;
}

int ppk_mlp_evaluate(ppk_mlp_t *net, const ppk_faceset_t *set, double *mse) {
  if (!net || !set || !mse) return nfail(net, PPK_ERR_INVALID_ARGUMENT, "NULL argument");
  if (set->n < 1) return nfail(net, PPK_ERR_INVALID_ARGUMENT, "face set is empty");
  const int nb = ppk::mlp::infer_blocks((long long)set->n);
  double *parts = nullptr;
  NHIP(net, hipMalloc(&parts, sizeof(double) * 8 * nb));
  std::vector<double> h((size_t)8 * nb);
  hipError_t e = ppk::mlp::infer(net->packed, set->X, set->T, (long long)set->n, nullptr, parts,
                                 net->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.data(), parts, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                       net->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(net->stream);
  (void)hipFree(parts);
  if (e != hipSuccess) return nfail(net, PPK_ERR_HIP, std::string("evaluate: ") + hipGetErrorString(e));
  for (int o = 0; o < 8; ++o) {
    double s = 0;
    for (int b = 0; b < nb; ++b) s += h[(size_t)b * 8 + o];
    mse[o] = s / (double)set->n;
  }
  return PPK_OK;
}

}  // extern "C"
