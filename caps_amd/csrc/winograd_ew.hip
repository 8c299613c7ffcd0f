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
#include "ew_ops.h"
#include "kernels.h"

namespace caps {

namespace {

struct D2 {
  double x, y;
};
__device__ inline D2 operator+(const D2 &a, const D2 &b) { return {a.x + b.x, a.y + b.y}; }
__device__ inline D2 operator-(const D2 &a, const D2 &b) { return {a.x - b.x, a.y - b.y}; }

constexpr int EW_THREADS = 256;

// One entry of a step: block bx of gx strides over the n elements of a
// quadrant.  Inputs and outputs may alias element-for-element (the DFS
// register reuse): every lane loads all its inputs before it stores.
template <int OP, bool VEC2>
__device__ __forceinline__ void ew_body(const EwPtrs &P, long long n, int bx, int gx) {
  using Op = Ew<OP>;
  const double *in[Op::NIN];
  double *out[Op::NOUT];
#pragma unroll
  for (int k = 0; k < Op::NIN; k++) in[k] = P.in[k];
#pragma unroll
  for (int k = 0; k < Op::NOUT; k++) out[k] = P.out[k];
  const long long stride = (long long)gx * EW_THREADS;
  if (VEC2) {
    const long long n2 = n >> 1;
    for (long long i = (long long)bx * EW_THREADS + threadIdx.x; i < n2; i += stride) {
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
    for (long long i = (long long)bx * EW_THREADS + threadIdx.x; i < n; i += stride) {
      double x[Op::NIN], y[Op::NOUT];
#pragma unroll
      for (int k = 0; k < Op::NIN; k++) x[k] = in[k][i];
      Op::apply(x, y);
#pragma unroll
      for (int k = 0; k < Op::NOUT; k++) out[k][i] = y[k];
    }
  }
}

// grid.y = sub-problem, grid.x strides over the n elements of a quadrant
template <int OP, bool VEC2>
__global__ __launch_bounds__(EW_THREADS) void ew_kernel(const EwPtrs *__restrict__ tab,
                                                        long long n) {
  ew_body<OP, VEC2>(tab[blockIdx.y], n, blockIdx.x, gridDim.x);
}

// the same over segments of different sizes (one launch for a phase of an
// uneven batch)
template <int OP, bool VEC2>
__global__ __launch_bounds__(EW_THREADS) void ew_multi_kernel(const char *__restrict__ base,
                                                              const VSeg *__restrict__ segs,
                                                              int nseg) {
  const VSeg &S = segs[vseg_find(segs, nseg, blockIdx.x)];
  const int l = blockIdx.x - S.blk0;
  const EwPtrs *tab = reinterpret_cast<const EwPtrs *>(base + S.arg);
  ew_body<OP, VEC2>(tab[l / S.gx], S.n, l % S.gx, S.gx);
}

template <int OP>
hipError_t launch_op(const EwPtrs *tab, int count, long long n, bool vec2, hipStream_t s) {
  int gx, gy;
  ew_grid(count, n, vec2, gx, gy);
  dim3 grid((unsigned)gx, (unsigned)gy);
  if (vec2)
    hipLaunchKernelGGL((ew_kernel<OP, true>), grid, dim3(EW_THREADS), 0, s, tab, n);
  else
    hipLaunchKernelGGL((ew_kernel<OP, false>), grid, dim3(EW_THREADS), 0, s, tab, n);
  return hipGetLastError();
}

template <int OP>
hipError_t launch_op_multi(const char *base, const VSeg *segs, int nseg, int blocks, bool vec2,
                           hipStream_t s) {
  if (vec2)
    hipLaunchKernelGGL((ew_multi_kernel<OP, true>), dim3(blocks), dim3(EW_THREADS), 0, s, base,
                       segs, nseg);
  else
    hipLaunchKernelGGL((ew_multi_kernel<OP, false>), dim3(blocks), dim3(EW_THREADS), 0, s, base,
                       segs, nseg);
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

void ew_grid(int count, long long n, bool vec2, int &gx, int &gy) {
  long long units = vec2 ? (n >> 1) : n;
  long long blocks = (units + EW_THREADS - 1) / EW_THREADS;
  // enough blocks in flight to cover HBM latency: ~8 per CU over the batch
  long long cap = (long long)device_cu_count() * 8 / (count > 0 ? count : 1);
  if (cap < 1) cap = 1;
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  gx = (int)blocks;
  gy = count;
}

hipError_t launch_ew_multi(int op, const char *base, const VSeg *segs, int nseg, int blocks,
                           bool vec2, hipStream_t s) {
  if (nseg <= 0 || blocks <= 0) return hipSuccess;
  switch (op) {
    case EW_BFS_PRE: return launch_op_multi<EW_BFS_PRE>(base, segs, nseg, blocks, vec2, s);
    case EW_BFS_POST: return launch_op_multi<EW_BFS_POST>(base, segs, nseg, blocks, vec2, s);
    case EW_DFS_PRE1: return launch_op_multi<EW_DFS_PRE1>(base, segs, nseg, blocks, vec2, s);
    case EW_DFS_PRE2: return launch_op_multi<EW_DFS_PRE2>(base, segs, nseg, blocks, vec2, s);
    case EW_DFS_PRE3: return launch_op_multi<EW_DFS_PRE3>(base, segs, nseg, blocks, vec2, s);
    case EW_DFS_POST1: return launch_op_multi<EW_DFS_POST1>(base, segs, nseg, blocks, vec2, s);
    case EW_DFS_POST2: return launch_op_multi<EW_DFS_POST2>(base, segs, nseg, blocks, vec2, s);
    case EW_DFS_POST3: return launch_op_multi<EW_DFS_POST3>(base, segs, nseg, blocks, vec2, s);
  }
  return hipErrorInvalidValue;
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
