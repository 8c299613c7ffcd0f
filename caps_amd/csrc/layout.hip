// CAPS storage layout on the device (one process).
//
// Layout (matrix.h:20-44, communication.cpp:612-708 with nproc = 1): nrec
// levels of quadrants in the order 11, 21, 12, 22, then the m x m leaf split
// into bs x bs blocks, column-major between blocks and within a block.
//
// Both conversions walk the column-major side with consecutive lanes on
// consecutive rows, so that side is fully coalesced; on the CAPS side
// consecutive rows stay consecutive inside a bs-block column (a run of bs
// doubles, the whole leaf column when bs = 1 or bs = m).
#include "kernels.h"

namespace caps {

namespace {

__device__ __forceinline__ long long caps_offset(int n, int nrec, int bs, int i, int j) {
  long long off = 0;
  long long s = n;
  for (int l = 0; l < nrec; l++) {
    long long h = s >> 1;
    int bi = i >= h, bj = j >= h;
    off += (long long)(bi + 2 * bj) * h * h;
    i -= bi * (int)h;
    j -= bj * (int)h;
    s = h;
  }
  long long nb = s / bs;
  return off + ((long long)(i / bs) + (long long)(j / bs) * nb) * bs * bs + (i % bs) +
         (long long)(j % bs) * bs;
}

constexpr int LT = 256;

template <bool TO_CAPS>
__global__ __launch_bounds__(LT) void layout_kernel(const double *__restrict__ src,
                                                    double *__restrict__ dst, int n,
                                                    int nrec, int bs) {
  const long long total = (long long)n * n;
  for (long long e = (long long)blockIdx.x * LT + threadIdx.x; e < total;
       e += (long long)gridDim.x * LT) {
    int i = (int)(e % n), j = (int)(e / n);
    long long o = caps_offset(n, nrec, bs, i, j);
    if (TO_CAPS)
      dst[o] = src[e];
    else
      dst[e] = src[o];
  }
}

// bs-blocked m x m leaf <-> column-major m x m leaf (block_multiply's
// reorder, multiply.cpp:1093-1115 and 1128-1139), batched over grid.y.
template <bool TO_CM>
__global__ __launch_bounds__(LT) void leaf_reorder_kernel(const ReorderPtrs *__restrict__ tab,
                                                          int m, int bs) {
  const double *src = tab[blockIdx.y].src;
  double *dst = tab[blockIdx.y].dst;
  const long long total = (long long)m * m;
  const int nb = m / bs;
  for (long long e = (long long)blockIdx.x * LT + threadIdx.x; e < total;
       e += (long long)gridDim.x * LT) {
    int i = (int)(e % m), j = (int)(e / m);
    long long b = ((long long)(i / bs) + (long long)(j / bs) * nb) * bs * bs + (i % bs) +
                  (long long)(j % bs) * bs;
    if (TO_CM)
      dst[e] = src[b];
    else
      dst[b] = src[e];
  }
}

unsigned grid_for(long long total, int per_problem) {
  long long b = (total + LT - 1) / LT;
  long long cap = (long long)device_cu_count() * 16 / (per_problem > 0 ? per_problem : 1);
  if (cap < 1) cap = 1;
  if (b > cap) b = cap;
  return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

hipError_t launch_to_caps(const double *cm, double *caps, int n, int nrec, int bs,
                          hipStream_t s) {
  hipLaunchKernelGGL((layout_kernel<true>), dim3(grid_for((long long)n * n, 1)), dim3(LT), 0,
                     s, cm, caps, n, nrec, bs);
  return hipGetLastError();
}

hipError_t launch_from_caps(const double *caps, double *cm, int n, int nrec, int bs,
                            hipStream_t s) {
  hipLaunchKernelGGL((layout_kernel<false>), dim3(grid_for((long long)n * n, 1)), dim3(LT), 0,
                     s, caps, cm, n, nrec, bs);
  return hipGetLastError();
}

hipError_t launch_leaf_reorder(const ReorderPtrs *tab, int count, int m, int bs,
                               bool to_colmajor, hipStream_t s) {
  for (int c0 = 0; c0 < count; c0 += 65535) {
    int c = count - c0 < 65535 ? count - c0 : 65535;
    dim3 grid(grid_for((long long)m * m, c), (unsigned)c);
    if (to_colmajor)
      hipLaunchKernelGGL((leaf_reorder_kernel<true>), grid, dim3(LT), 0, s, tab + c0, m, bs);
    else
      hipLaunchKernelGGL((leaf_reorder_kernel<false>), grid, dim3(LT), 0, s, tab + c0, m, bs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace caps
