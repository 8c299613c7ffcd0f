// Fused Winograd pre-additions and product combinations, batched over the
// sub-problems of a level.  HBM-bound: each kernel streams its quadrant inputs
// once and writes its outputs once, 16 bytes per lane per access.
//
// Per-element arithmetic follows the reference association exactly:
//   BFS  T3..T6: addSubSubSubMatrices call at multiply.cpp:272 (165-178)
//        S5,S3,S4,S7: quadrupleSubMatrices call at multiply.cpp:277 (180-193)
//        C: calcCij (multiply.cpp:195-210)
//   DFS  the non-DAXPY sequence of multiply.cpp:409-451 with its register
//        reuse; consecutive add calls with no product between them are fused.
#include "kernels.h"

namespace caps {

namespace {

template <int OP>
struct Ew;

template <>
struct Ew<EW_BFS_PRE> {
  static constexpr int NIN = 8, NOUT = 8;
  template <typename V>
  __device__ static void apply(const V *x, V *y) {
    const V &A11 = x[0], &A21 = x[1], &A12 = x[2], &A22 = x[3];
    const V &B11 = x[4], &B21 = x[5], &B12 = x[6], &B22 = x[7];
    V T3 = A21 + A22;
    V T4 = T3 - A11;
    V T5 = A11 - A21;
    V T6 = A12 - T4;
    V S5 = B22 - B12;
    V S3 = B12 - B11;
    V S4 = B22 - S3;
    V S7 = S4 - B21;
    y[0] = T3; y[1] = T4; y[2] = T5; y[3] = T6;
    y[4] = S3; y[5] = S4; y[6] = S5; y[7] = S7;
  }
};

template <>
struct Ew<EW_BFS_POST> {
  static constexpr int NIN = 7, NOUT = 4;
  template <typename V>
  __device__ static void apply(const V *q, V *c) {
    V C11 = q[0] + q[1];
    V U1 = q[0] + q[3];
    V U2 = U1 + q[4];
    V C21 = U2 - q[6];
    V C22 = U2 + q[2];
    V U3 = U1 + q[2];
    V C12 = U3 + q[5];
    c[0] = C11; c[1] = C21; c[2] = C12; c[3] = C22;
  }
};

// tripleSubMatrices(S5=B22-B12, S3=B12-B11, S4=B22-S3) then
// addSubMatrices(T3=A21+A22, T5=A11-A21)            (multiply.cpp:421, 424)
template <>
struct Ew<EW_DFS_PRE1> {
  static constexpr int NIN = 6, NOUT = 5;
  template <typename V>
  __device__ static void apply(const V *x, V *y) {
    const V &A11 = x[0], &A21 = x[1], &A22 = x[2];
    const V &B11 = x[3], &B12 = x[4], &B22 = x[5];
    V S5 = B22 - B12;
    V S3 = B12 - B11;
    V S4 = B22 - S3;
    V T3 = A21 + A22;
    V T5 = A11 - A21;
    y[0] = S5; y[1] = S3; y[2] = S4; y[3] = T5; y[4] = T3;
  }
};

// subMatrices(T4 = T3 - A11)                          (multiply.cpp:430)
template <>
struct Ew<EW_DFS_PRE2> {
  static constexpr int NIN = 2, NOUT = 1;
  template <typename V>
  __device__ static void apply(const V *x, V *y) { y[0] = x[0] - x[1]; }
};

// subMatrices(T6 = A12 - T4); subMatrices(S7 = S4 - B21)   (434, 436)
template <>
struct Ew<EW_DFS_PRE3> {
  static constexpr int NIN = 4, NOUT = 2;
  template <typename V>
  __device__ static void apply(const V *x, V *y) {
    y[0] = x[0] - x[1];
    y[1] = x[2] - x[3];
  }
};

// tripleAddMatrices(U1=Q1+Q4, U2=U1+Q5, U3=U1+Q3) then
// addSubMatrices(C22=U2+Q3, C21=U2-Q7)                     (444, 445)
template <>
struct Ew<EW_DFS_POST1> {
  static constexpr int NIN = 5, NOUT = 3;
  template <typename V>
  __device__ static void apply(const V *x, V *y) {
    const V &Q1 = x[0], &Q4 = x[1], &Q5 = x[2], &Q3 = x[3], &Q7 = x[4];
    V U1 = Q1 + Q4;
    V U2 = U1 + Q5;
    V U3 = U1 + Q3;
    y[0] = U3;
    y[1] = U2 + Q3;
    y[2] = U2 - Q7;
  }
};

// addMatrices(C11 = Q1 + Q2)                              (448)
template <>
struct Ew<EW_DFS_POST2> {
  static constexpr int NIN = 2, NOUT = 1;
  template <typename V>
  __device__ static void apply(const V *x, V *y) { y[0] = x[0] + x[1]; }
};

// addMatrices(C12 = U3 + Q6)                              (451)
template <>
struct Ew<EW_DFS_POST3> {
  static constexpr int NIN = 2, NOUT = 1;
  template <typename V>
  __device__ static void apply(const V *x, V *y) { y[0] = x[0] + x[1]; }
};

struct D2 {
  double x, y;
};
__device__ inline D2 operator+(const D2 &a, const D2 &b) { return {a.x + b.x, a.y + b.y}; }
__device__ inline D2 operator-(const D2 &a, const D2 &b) { return {a.x - b.x, a.y - b.y}; }

constexpr int EW_THREADS = 256;

// grid.y = sub-problem, grid.x strides over the n elements of a quadrant.
// Inputs and outputs may alias element-for-element (the DFS register reuse):
// every lane loads all its inputs before it stores.
template <int OP, bool VEC2>
__global__ __launch_bounds__(EW_THREADS) void ew_kernel(const EwPtrs *__restrict__ tab,
                                                        long long n) {
  using Op = Ew<OP>;
  const EwPtrs &P = tab[blockIdx.y];
  const double *in[Op::NIN];
  double *out[Op::NOUT];
#pragma unroll
  for (int k = 0; k < Op::NIN; k++) in[k] = P.in[k];
#pragma unroll
  for (int k = 0; k < Op::NOUT; k++) out[k] = P.out[k];
  const long long stride = (long long)gridDim.x * EW_THREADS;
  if (VEC2) {
    const long long n2 = n >> 1;
    for (long long i = (long long)blockIdx.x * EW_THREADS + threadIdx.x; i < n2; i += stride) {
      D2 x[Op::NIN], y[Op::NOUT];
#pragma unroll
      for (int k = 0; k < Op::NIN; k++) {
        double2 v = reinterpret_cast<const double2 *>(in[k])[i];
        x[k] = {v.x, v.y};
      }
      Op::apply(x, y);
#pragma unroll
      for (int k = 0; k < Op::NOUT; k++)
        reinterpret_cast<double2 *>(out[k])[i] = make_double2(y[k].x, y[k].y);
    }
  } else {
    for (long long i = (long long)blockIdx.x * EW_THREADS + threadIdx.x; i < n; i += stride) {
      double x[Op::NIN], y[Op::NOUT];
#pragma unroll
      for (int k = 0; k < Op::NIN; k++) x[k] = in[k][i];
      Op::apply(x, y);
#pragma unroll
      for (int k = 0; k < Op::NOUT; k++) out[k][i] = y[k];
    }
  }
}

template <int OP>
hipError_t launch_op(const EwPtrs *tab, int count, long long n, bool vec2, hipStream_t s) {
  long long units = vec2 ? (n >> 1) : n;
  long long blocks = (units + EW_THREADS - 1) / EW_THREADS;
  // enough blocks in flight to cover HBM latency: ~8 per CU over the batch
  long long cap = (long long)device_cu_count() * 8 / (count > 0 ? count : 1);
  if (cap < 1) cap = 1;
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  dim3 grid((unsigned)blocks, (unsigned)count);
  if (vec2)
    hipLaunchKernelGGL((ew_kernel<OP, true>), grid, dim3(EW_THREADS), 0, s, tab, n);
  else
    hipLaunchKernelGGL((ew_kernel<OP, false>), grid, dim3(EW_THREADS), 0, s, tab, n);
  return hipGetLastError();
}

}  // namespace

int ew_additions(int op) {
  switch (op) {
    case EW_BFS_PRE: return 8;
    case EW_BFS_POST: return 7;
    case EW_DFS_PRE1: return 5;
    case EW_DFS_PRE2: return 1;
    case EW_DFS_PRE3: return 2;
    case EW_DFS_POST1: return 5;
    case EW_DFS_POST2: return 1;
    case EW_DFS_POST3: return 1;
  }
  return 0;
}

hipError_t launch_ew(int op, const EwPtrs *tab, int count, long long n, bool vec2,
                     hipStream_t s) {
  if (count <= 0 || n <= 0) return hipSuccess;
  // grid.y is limited to 65535 sub-problems per launch
  for (int c0 = 0; c0 < count; c0 += 65535) {
    const EwPtrs *t = tab + c0;
    int c = count - c0 < 65535 ? count - c0 : 65535;
    hipError_t e = hipErrorInvalidValue;
    switch (op) {
      case EW_BFS_PRE: e = launch_op<EW_BFS_PRE>(t, c, n, vec2, s); break;
      case EW_BFS_POST: e = launch_op<EW_BFS_POST>(t, c, n, vec2, s); break;
      case EW_DFS_PRE1: e = launch_op<EW_DFS_PRE1>(t, c, n, vec2, s); break;
      case EW_DFS_PRE2: e = launch_op<EW_DFS_PRE2>(t, c, n, vec2, s); break;
      case EW_DFS_PRE3: e = launch_op<EW_DFS_PRE3>(t, c, n, vec2, s); break;
      case EW_DFS_POST1: e = launch_op<EW_DFS_POST1>(t, c, n, vec2, s); break;
      case EW_DFS_POST2: e = launch_op<EW_DFS_POST2>(t, c, n, vec2, s); break;
      case EW_DFS_POST3: e = launch_op<EW_DFS_POST3>(t, c, n, vec2, s); break;
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace caps
