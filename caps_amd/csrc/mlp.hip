// Cost / error model: a 3-layer perceptron (D_IN -> 64 -> 64 -> 2, ReLU,
// FP32) trained on timings and errors of the method itself; its forward,
// backward and Adam kernels are below.  No counterpart in the reference
// (which picks BFS/DFS from a memory rule, multiply.cpp:43-66).
//
// Loss: mean over the batch of sum_k mask_k (y_k - t_k)^2 (mask_k = 0 drops
// a missing target).  Gradients are reduced over the batch in a fixed order
// (per-block partial sums in sample order, then blocks in order), so training
// is deterministic.
#include "mlp.h"

namespace caps {

namespace {

constexpr int T = 256;
constexpr int H = MLP_HIDDEN;
constexpr int DI = MLP_IN;
constexpr int DO = MLP_OUT;

struct Params {
  const float *W1, *b1, *W2, *b2, *W3, *b3;
};

__device__ __forceinline__ Params params_of(const float *p) {
  Params q;
  q.W1 = p + MLP_OFF_W1;
  q.b1 = p + MLP_OFF_B1;
  q.W2 = p + MLP_OFF_W2;
  q.b2 = p + MLP_OFF_B2;
  q.W3 = p + MLP_OFF_W3;
  q.b3 = p + MLP_OFF_B3;
  return q;
}

// one thread per sample: activations to global memory (B x H each)
__global__ __launch_bounds__(T) void mlp_forward_kernel(const float *__restrict__ p,
                                                        const float *__restrict__ X, int B,
                                                        float *__restrict__ A1,
                                                        float *__restrict__ A2,
                                                        float *__restrict__ Y) {
  const int s = blockIdx.x * T + threadIdx.x;
  if (s >= B) return;
  const Params q = params_of(p);
  float x[DI];
#pragma unroll
  for (int i = 0; i < DI; i++) x[i] = X[(long long)s * DI + i];
  float h1[H];
#pragma unroll 4
  for (int j = 0; j < H; j++) {
    float a = q.b1[j];
#pragma unroll
    for (int i = 0; i < DI; i++) a = fmaf(q.W1[j * DI + i], x[i], a);
    h1[j] = a > 0.f ? a : 0.f;
    if (A1) A1[(long long)s * H + j] = h1[j];
  }
  float h2[H];
  for (int j = 0; j < H; j++) {
    float a = q.b2[j];
#pragma unroll 8
    for (int i = 0; i < H; i++) a = fmaf(q.W2[j * H + i], h1[i], a);
    h2[j] = a > 0.f ? a : 0.f;
    if (A2) A2[(long long)s * H + j] = h2[j];
  }
#pragma unroll
  for (int k = 0; k < DO; k++) {
    float a = q.b3[k];
    for (int i = 0; i < H; i++) a = fmaf(q.W3[k * H + i], h2[i], a);
    Y[(long long)s * DO + k] = a;
  }
}

// Per-sample gradient contributions, summed over the block's samples in
// sample order by thread (parameter index) -> partial[block][param].
// Each thread owns a set of parameters; per sample it recomputes the deltas
// it needs from the stored activations (delta3, then delta2, delta1 on the fly).
__global__ __launch_bounds__(T) void mlp_backward_kernel(
    const float *__restrict__ p, const float *__restrict__ X, const float *__restrict__ A1,
    const float *__restrict__ A2, const float *__restrict__ Y, const float *__restrict__ Tg,
    const float *__restrict__ Mk, int B, int per_block, float *__restrict__ partial) {
  __shared__ float d3[DO];
  __shared__ float d2[H];
  __shared__ float d1[H];
  __shared__ float a1[H];
  __shared__ float a2[H];
  __shared__ float xs[DI];
  const Params q = params_of(p);
  const int s0 = blockIdx.x * per_block, s1 = min(B, s0 + per_block);
  float acc[(MLP_NPARAM + T - 1) / T];
#pragma unroll
  for (int r = 0; r < (MLP_NPARAM + T - 1) / T; r++) acc[r] = 0.f;
  const float scale = 2.0f / (float)B;
  for (int s = s0; s < s1; s++) {
    if (threadIdx.x < DO) {
      const int k = threadIdx.x;
      d3[k] = scale * Mk[(long long)s * DO + k] * (Y[(long long)s * DO + k] - Tg[(long long)s * DO + k]);
    }
    if (threadIdx.x < H) {
      a1[threadIdx.x] = A1[(long long)s * H + threadIdx.x];
      a2[threadIdx.x] = A2[(long long)s * H + threadIdx.x];
    }
    if (threadIdx.x < DI) xs[threadIdx.x] = X[(long long)s * DI + threadIdx.x];
    __syncthreads();
    if (threadIdx.x < H) {
      const int j = threadIdx.x;
      float g = 0.f;
#pragma unroll
      for (int k = 0; k < DO; k++) g = fmaf(q.W3[k * H + j], d3[k], g);
      d2[j] = a2[j] > 0.f ? g : 0.f;
    }
    __syncthreads();
    if (threadIdx.x < H) {
      const int j = threadIdx.x;
      float g = 0.f;
      for (int i = 0; i < H; i++) g = fmaf(q.W2[i * H + j], d2[i], g);
      d1[j] = a1[j] > 0.f ? g : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < (MLP_NPARAM + T - 1) / T; r++) {
      const int idx = threadIdx.x + r * T;
      if (idx >= MLP_NPARAM) break;
      float g;
      if (idx < MLP_OFF_B1) {  // W1[j][i]
        const int j = idx / DI, i = idx - j * DI;
        g = d1[j] * xs[i];
      } else if (idx < MLP_OFF_W2) {
        g = d1[idx - MLP_OFF_B1];
      } else if (idx < MLP_OFF_B2) {  // W2[j][i]
        const int t = idx - MLP_OFF_W2, j = t / H, i = t - j * H;
        g = d2[j] * a1[i];
      } else if (idx < MLP_OFF_W3) {
        g = d2[idx - MLP_OFF_B2];
      } else if (idx < MLP_OFF_B3) {  // W3[k][i]
        const int t = idx - MLP_OFF_W3, k = t / H, i = t - k * H;
        g = d3[k] * a2[i];
      } else {
        g = d3[idx - MLP_OFF_B3];
      }
      acc[r] += g;
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < (MLP_NPARAM + T - 1) / T; r++) {
    const int idx = threadIdx.x + r * T;
    if (idx < MLP_NPARAM) partial[(long long)blockIdx.x * MLP_NPARAM + idx] = acc[r];
  }
}

// grad[i] = sum over blocks (in order) of partial[b][i]
__global__ __launch_bounds__(T) void mlp_reduce_kernel(const float *__restrict__ partial,
                                                       int nblocks, float *__restrict__ grad) {
  const int i = blockIdx.x * T + threadIdx.x;
  if (i >= MLP_NPARAM) return;
  float g = 0.f;
  for (int b = 0; b < nblocks; b++) g += partial[(long long)b * MLP_NPARAM + i];
  grad[i] = g;
}

// Adam (Kingma & Ba, 2015) with bias correction
__global__ __launch_bounds__(T) void mlp_adam_kernel(float *__restrict__ p,
                                                     const float *__restrict__ g,
                                                     float *__restrict__ m1,
                                                     float *__restrict__ m2, float lr,
                                                     float beta1, float beta2, float eps,
                                                     float corr1, float corr2) {
  const int i = blockIdx.x * T + threadIdx.x;
  if (i >= MLP_NPARAM) return;
  const float gi = g[i];
  const float a = beta1 * m1[i] + (1.f - beta1) * gi;
  const float b = beta2 * m2[i] + (1.f - beta2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr * (a / corr1) / (sqrtf(b / corr2) + eps);
}

// loss partials: per block, fixed order
__global__ __launch_bounds__(T) void mlp_loss_kernel(const float *__restrict__ Y,
                                                     const float *__restrict__ Tg,
                                                     const float *__restrict__ Mk, int B,
                                                     float *__restrict__ out) {
  __shared__ float red[T];
  float acc = 0.f;
  for (int s = threadIdx.x; s < B; s += T)
    for (int k = 0; k < DO; k++) {
      const float d = Y[(long long)s * DO + k] - Tg[(long long)s * DO + k];
      acc += Mk[(long long)s * DO + k] * d * d;
    }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = T / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0] / (float)B;
}

}  // namespace

hipError_t mlp_forward(const float *p, const float *X, int B, float *A1, float *A2, float *Y,
                       hipStream_t s) {
  if (B <= 0) return hipSuccess;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((B + T - 1) / T), dim3(T), 0, s, p, X, B, A1, A2, Y);
  return hipGetLastError();
}

int mlp_backward_blocks(int B) { return (B + MLP_BACKWARD_PER_BLOCK - 1) / MLP_BACKWARD_PER_BLOCK; }

hipError_t mlp_backward(const float *p, const float *X, const float *A1, const float *A2,
                        const float *Y, const float *Tg, const float *Mk, int B, float *partial,
                        float *grad, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  const int nb = mlp_backward_blocks(B);
  hipLaunchKernelGGL(mlp_backward_kernel, dim3(nb), dim3(T), 0, s, p, X, A1, A2, Y, Tg, Mk, B,
                     MLP_BACKWARD_PER_BLOCK, partial);
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3((MLP_NPARAM + T - 1) / T), dim3(T), 0, s, partial,
                     nb, grad);
  return hipGetLastError();
}

hipError_t mlp_adam(float *p, const float *g, float *m1, float *m2, float lr, float beta1,
                    float beta2, float eps, int step, hipStream_t s) {
  const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((MLP_NPARAM + T - 1) / T), dim3(T), 0, s, p, g, m1,
                     m2, lr, beta1, beta2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t mlp_loss(const float *Y, const float *Tg, const float *Mk, int B, float *out,
                    hipStream_t s) {
  hipLaunchKernelGGL(mlp_loss_kernel, dim3(1), dim3(T), 0, s, Y, Tg, Mk, B, out);
  return hipGetLastError();
}

}  // namespace caps
