// C ABI of the planner's cost / error model (include/caps_amd.h): device
// buffers, deterministic initialisation, forward, gradient and Adam training
// driven from the host, all arithmetic in the HIP kernels of mlp.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <utility>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/caps_amd.h"
#include "mlp.h"
#include "planner.h"

namespace caps {
int set_error(int code, const char *msg);
}

struct caps_mlp {
  float *p = nullptr, *m1 = nullptr, *m2 = nullptr, *g = nullptr;
  float *X = nullptr, *T = nullptr, *M = nullptr, *Y = nullptr, *A1 = nullptr, *A2 = nullptr;
  float *partial = nullptr, *loss = nullptr;
  int cap = 0;
  int step = 0;
  hipStream_t s = nullptr;
};

namespace {

using caps::MLP_HIDDEN;
using caps::MLP_IN;
using caps::MLP_NPARAM;
using caps::MLP_OUT;

int herr(hipError_t e, const char *what) {
  if (e == hipSuccess) return CAPS_OK;
  char buf[256];
  snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
  return caps::set_error(CAPS_EHIP, buf);
}

#define TRY(expr)                                   \
  do {                                              \
    int rc_ = herr((expr), #expr);                  \
    if (rc_) return rc_;                            \
  } while (0)

uint64_t splitmix(uint64_t &x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double uniform01(uint64_t &x) { return (double)(splitmix(x) >> 11) * 0x1p-53; }

int reserve(caps_mlp *h, int B) {
  if (B <= h->cap) return CAPS_OK;
  for (float **q : {&h->X, &h->T, &h->M, &h->Y, &h->A1, &h->A2, &h->partial})
    if (*q) {
      hipFree(*q);
      *q = nullptr;
    }
  const size_t b = (size_t)B;
  TRY(hipMalloc(&h->X, b * MLP_IN * sizeof(float)));
  TRY(hipMalloc(&h->T, b * MLP_OUT * sizeof(float)));
  TRY(hipMalloc(&h->M, b * MLP_OUT * sizeof(float)));
  TRY(hipMalloc(&h->Y, b * MLP_OUT * sizeof(float)));
  TRY(hipMalloc(&h->A1, b * MLP_HIDDEN * sizeof(float)));
  TRY(hipMalloc(&h->A2, b * MLP_HIDDEN * sizeof(float)));
  TRY(hipMalloc(&h->partial, (size_t)caps::mlp_backward_blocks(B) * MLP_NPARAM * sizeof(float)));
  h->cap = B;
  return CAPS_OK;
}

int upload(caps_mlp *h, const float *X, const float *T, const float *M, int B) {
  int rc = reserve(h, B);
  if (rc) return rc;
  TRY(hipMemcpyAsync(h->X, X, (size_t)B * MLP_IN * sizeof(float), hipMemcpyHostToDevice, h->s));
  if (T)
    TRY(hipMemcpyAsync(h->T, T, (size_t)B * MLP_OUT * sizeof(float), hipMemcpyHostToDevice, h->s));
  if (M)
    TRY(hipMemcpyAsync(h->M, M, (size_t)B * MLP_OUT * sizeof(float), hipMemcpyHostToDevice, h->s));
  return CAPS_OK;
}

// one forward + backward over device rows [off, off + B) of X, T, M
int grad_on(caps_mlp *h, int off, int B) {
  const float *X = h->X + (size_t)off * MLP_IN, *T = h->T + (size_t)off * MLP_OUT,
              *M = h->M + (size_t)off * MLP_OUT;
  float *Y = h->Y + (size_t)off * MLP_OUT;
  TRY(caps::mlp_forward(h->p, X, B, h->A1, h->A2, Y, h->s));
  TRY(caps::mlp_backward(h->p, X, h->A1, h->A2, Y, T, M, B, h->partial, h->g, h->s));
  return CAPS_OK;
}

}  // namespace

extern "C" {

int caps_mlp_nparams(void) { return MLP_NPARAM; }

int caps_mlp_create(caps_mlp **out, unsigned long long seed) {
  if (!out) return caps::set_error(CAPS_EINVAL, "caps_mlp_create: NULL out");
  caps_mlp *h = new caps_mlp();
  *out = nullptr;
  TRY(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
  for (float **q : {&h->p, &h->m1, &h->m2, &h->g}) TRY(hipMalloc(q, MLP_NPARAM * sizeof(float)));
  TRY(hipMalloc(&h->loss, sizeof(float)));
  // He-uniform weights, zero biases, from a splitmix64 stream
  std::vector<float> w(MLP_NPARAM, 0.f);
  uint64_t st = seed;
  auto fill = [&](int off, int rows, int cols) {
    const double a = std::sqrt(6.0 / cols);
    for (int i = 0; i < rows * cols; i++) w[off + i] = (float)((2 * uniform01(st) - 1) * a);
  };
  fill(caps::MLP_OFF_W1, MLP_HIDDEN, MLP_IN);
  fill(caps::MLP_OFF_W2, MLP_HIDDEN, MLP_HIDDEN);
  fill(caps::MLP_OFF_W3, MLP_OUT, MLP_HIDDEN);
  TRY(hipMemcpy(h->p, w.data(), MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice));
  TRY(hipMemset(h->m1, 0, MLP_NPARAM * sizeof(float)));
  TRY(hipMemset(h->m2, 0, MLP_NPARAM * sizeof(float)));
  *out = h;
  return CAPS_OK;
}

int caps_mlp_destroy(caps_mlp *h) {
  if (!h) return CAPS_OK;
  caps::planner_forget_model(h);
  hipStreamSynchronize(h->s);
  for (float *q : {h->p, h->m1, h->m2, h->g, h->X, h->T, h->M, h->Y, h->A1, h->A2, h->partial,
                   h->loss})
    if (q) hipFree(q);
  hipStreamDestroy(h->s);
  delete h;
  return CAPS_OK;
}

int caps_mlp_get_params(caps_mlp *h, float *out, int n) {
  if (!h || !out || n != MLP_NPARAM) return caps::set_error(CAPS_EINVAL, "caps_mlp_get_params");
  TRY(hipMemcpyAsync(out, h->p, n * sizeof(float), hipMemcpyDeviceToHost, h->s));
  TRY(hipStreamSynchronize(h->s));
  return CAPS_OK;
}

int caps_mlp_set_params(caps_mlp *h, const float *in, int n) {
  if (!h || !in || n != MLP_NPARAM) return caps::set_error(CAPS_EINVAL, "caps_mlp_set_params");
  TRY(hipMemcpyAsync(h->p, in, n * sizeof(float), hipMemcpyHostToDevice, h->s));
  TRY(hipMemsetAsync(h->m1, 0, n * sizeof(float), h->s));
  TRY(hipMemsetAsync(h->m2, 0, n * sizeof(float), h->s));
  TRY(hipStreamSynchronize(h->s));
  h->step = 0;
  return CAPS_OK;
}

int caps_mlp_forward(caps_mlp *h, const float *X, int B, float *Y) {
  if (!h || !X || !Y || B <= 0) return caps::set_error(CAPS_EINVAL, "caps_mlp_forward");
  int rc = upload(h, X, nullptr, nullptr, B);
  if (rc) return rc;
  TRY(caps::mlp_forward(h->p, h->X, B, nullptr, nullptr, h->Y, h->s));
  TRY(hipMemcpyAsync(Y, h->Y, (size_t)B * MLP_OUT * sizeof(float), hipMemcpyDeviceToHost, h->s));
  TRY(hipStreamSynchronize(h->s));
  return CAPS_OK;
}

int caps_mlp_gradient(caps_mlp *h, const float *X, const float *T, const float *M, int B,
                      float *grad, float *loss) {
  if (!h || !X || !T || !M || !grad || B <= 0)
    return caps::set_error(CAPS_EINVAL, "caps_mlp_gradient");
  int rc = upload(h, X, T, M, B);
  if (rc) return rc;
  rc = grad_on(h, 0, B);
  if (rc) return rc;
  TRY(hipMemcpyAsync(grad, h->g, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, h->s));
  if (loss) {
    TRY(caps::mlp_loss(h->Y, h->T, h->M, B, h->loss, h->s));
    TRY(hipMemcpyAsync(loss, h->loss, sizeof(float), hipMemcpyDeviceToHost, h->s));
  }
  TRY(hipStreamSynchronize(h->s));
  return CAPS_OK;
}

int caps_mlp_train(caps_mlp *h, const float *X, const float *T, const float *M, int B,
                   int epochs, int batch, float lr, unsigned long long seed, float *final_loss) {
  if (!h || !X || !T || !M || B <= 0 || epochs < 0 || batch <= 0)
    return caps::set_error(CAPS_EINVAL, "caps_mlp_train");
  std::vector<float> Xs((size_t)B * MLP_IN), Ts((size_t)B * MLP_OUT), Ms((size_t)B * MLP_OUT);
  std::vector<int> order(B);
  for (int i = 0; i < B; i++) order[i] = i;
  uint64_t st = seed;
  for (int ep = 0; ep < epochs; ep++) {
    for (int i = B - 1; i > 0; i--) {  // Fisher-Yates with the seeded stream
      int j = (int)(splitmix(st) % (uint64_t)(i + 1));
      std::swap(order[i], order[j]);
    }
    for (int i = 0; i < B; i++) {
      memcpy(&Xs[(size_t)i * MLP_IN], X + (size_t)order[i] * MLP_IN, MLP_IN * sizeof(float));
      memcpy(&Ts[(size_t)i * MLP_OUT], T + (size_t)order[i] * MLP_OUT, MLP_OUT * sizeof(float));
      memcpy(&Ms[(size_t)i * MLP_OUT], M + (size_t)order[i] * MLP_OUT, MLP_OUT * sizeof(float));
    }
    int rc = upload(h, Xs.data(), Ts.data(), Ms.data(), B);
    if (rc) return rc;
    for (int off = 0; off < B; off += batch) {
      const int b = std::min(batch, B - off);
      rc = grad_on(h, off, b);
      if (rc) return rc;
      h->step++;
      TRY(caps::mlp_adam(h->p, h->g, h->m1, h->m2, lr, 0.9f, 0.999f, 1e-8f, h->step, h->s));
    }
  }
  if (final_loss) {
    int rc = upload(h, X, T, M, B);
    if (rc) return rc;
    TRY(caps::mlp_forward(h->p, h->X, B, nullptr, nullptr, h->Y, h->s));
    TRY(caps::mlp_loss(h->Y, h->T, h->M, B, h->loss, h->s));
    TRY(hipMemcpyAsync(final_loss, h->loss, sizeof(float), hipMemcpyDeviceToHost, h->s));
  }
  TRY(hipStreamSynchronize(h->s));
  return CAPS_OK;
}

}  // extern "C"
