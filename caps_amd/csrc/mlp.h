// Cost / error model of the planner: layout of the 3-layer perceptron and the
// launchers of its HIP kernels (mlp.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace caps {

constexpr int MLP_IN = 8;
constexpr int MLP_HIDDEN = 64;
constexpr int MLP_OUT = 2;
// parameter vector: W1 (H x IN, row-major), b1, W2 (H x H), b2, W3 (OUT x H), b3
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_HIDDEN * MLP_IN;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_HIDDEN;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_HIDDEN * MLP_HIDDEN;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_HIDDEN;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_OUT * MLP_HIDDEN;
constexpr int MLP_NPARAM = MLP_OFF_B3 + MLP_OUT;
constexpr int MLP_BACKWARD_PER_BLOCK = 64;

// Y = f(X) for B samples (X: B x IN, Y: B x OUT); A1, A2 (B x H) receive the
// hidden activations when not null.
hipError_t mlp_forward(const float *p, const float *X, int B, float *A1, float *A2, float *Y,
                       hipStream_t s);
int mlp_backward_blocks(int B);
// grad of mean_s sum_k M (Y - T)^2 w.r.t. p; partial has mlp_backward_blocks(B)
// x MLP_NPARAM floats
hipError_t mlp_backward(const float *p, const float *X, const float *A1, const float *A2,
                        const float *Y, const float *T, const float *M, int B, float *partial,
                        float *grad, hipStream_t s);
hipError_t mlp_adam(float *p, const float *g, float *m1, float *m2, float lr, float beta1,
                    float beta2, float eps, int step, hipStream_t s);
hipError_t mlp_loss(const float *Y, const float *T, const float *M, int B, float *out,
                    hipStream_t s);

}  // namespace caps
