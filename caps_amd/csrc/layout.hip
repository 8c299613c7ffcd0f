// CAPS storage layout on the device (one process).
//
// Layout (matrix.h:20-44, communication.cpp:612-708 with nproc = 1): nrec
// levels of quadrants in the order 11, 21, 12, 22, then the m x m leaf split
// into bs x bs blocks, column-major between blocks and within a block.
//
// HBM-bound (16 bytes per element moved; stores are non-temporal: the
// result is not read again by this pass).  The tile kernels stage a 64 x 64
// tile of a leaf through LDS: the column-major side moves as 64 columns of 64
// contiguous doubles, and the CAPS side as the tile's runs in storage order:
// with bs dividing 64 the tile covers 64 / bs block columns, each one
// contiguous run of 64 bs doubles (blocks are column-major between blocks);
// with bs a multiple of 64 the tile lies inside one block and its 64 columns
// are runs of 64.  Both sides are fully coalesced.  Other (bs, leaf) shapes
// take the element kernel, which walks the column-major side.
#include "kernels.h"
#include "tilemap.h"

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
constexpr int TS = 64;  // tile side

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

// LDS pitch (doubles) of a tile column.  The CAPS side walks a block column
// as runs of BS rows in BS consecutive tile columns; a 64-bit LDS read serves
// 16 lanes at a time, here 16 / BS columns of BS doubles each, and with the
// pitch 64 + BS those columns start 2 BS banks apart: the 16 lanes read 32
// consecutive banks (pitch 65 put the columns 2 banks apart and overlapped
// them).  Runs of 64 (BS = 0) read one column.
constexpr int tile_pitch(int BS) { return TS + (BS > 0 ? BS : 1); }

// Tile (ti, tj) of one n x n matrix: cm is column-major (leading dimension
// n), caps the CAPS layout of depth nrec with blocks of bs (BS > 0: bs = BS
// divides 64; BS == 0: bs is a multiple of 64).  The leaf size m = n >> nrec
// is a multiple of 64, so the tile lies in one leaf.
template <bool TO_CAPS, int BS>
__device__ __forceinline__ void layout_tile(const double *__restrict__ src,
                                            double *__restrict__ dst, int n, int nrec, int bs,
                                            int ti, int tj,
                                            double (&S)[TS][tile_pitch(BS)]) {
  const int m = n >> nrec;
  const int i0 = ti * TS, j0 = tj * TS;  // tile origin in the matrix
  const int li0 = i0 % m, lj0 = j0 % m;  // and in its leaf
  const long long leaf = quad_leaf(i0 / m, j0 / m, nrec) * (long long)m * m;
  const int lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
  // storage offset (inside the leaf) of the tile's element q in storage order,
  // and its tile coordinates (r, c)
  auto caps_at = [&](int q, int &r, int &c) -> long long {
    if (BS > 0) {
      // q = bcol (64 BS) + brow BS^2 + jj BS + ii
      const int bcol = q / (TS * BS), w = q % (TS * BS);
      const int brow = w / (BS * BS), e = w % (BS * BS);
      r = brow * BS + e % BS;
      c = bcol * BS + e / BS;
      const int nb = m / BS;
      return ((long long)((li0 + r) / BS) + (long long)((lj0 + c) / BS) * nb) * (BS * BS) +
             e;
    } else {
      // inside one bs x bs block: column c of the tile is a run of 64
      r = q & (TS - 1);
      c = q >> 6;
      const int nb = m / bs;
      const long long blk = ((long long)(li0 / bs) + (long long)(lj0 / bs) * nb) * bs * bs;
      return blk + (li0 % bs) + r + (long long)((lj0 % bs) + c) * bs;
    }
  };
  if (TO_CAPS) {
#pragma unroll 4
    for (int c = grp; c < TS; c += LT / 64) S[c][lane] = src[(long long)(j0 + c) * n + i0 + lane];
    __syncthreads();
#pragma unroll 4
    for (int q = threadIdx.x; q < TS * TS; q += LT) {
      int r, c;
      const long long o = caps_at(q, r, c);
      __builtin_nontemporal_store(S[c][r], dst + leaf + o);
    }
  } else {
#pragma unroll 4
    for (int q = threadIdx.x; q < TS * TS; q += LT) {
      int r, c;
      const long long o = caps_at(q, r, c);
      S[c][r] = src[leaf + o];
    }
    __syncthreads();
#pragma unroll 4
    for (int c = grp; c < TS; c += LT / 64)
      __builtin_nontemporal_store(S[c][lane], dst + (long long)(j0 + c) * n + i0 + lane);
  }
  __syncthreads();  // S is reused by the block's next tile
}

// persistent over the (n / 64)^2 tiles of `count` matrices (grid.y = matrix)
template <bool TO_CAPS, int BS>
__global__ __launch_bounds__(LT) void layout_tile_kernel(const ReorderPtrs *__restrict__ tab,
                                                         const double *__restrict__ src1,
                                                         double *__restrict__ dst1, int n,
                                                         int nrec, int bs) {
  __shared__ double S[TS][tile_pitch(BS)];
  const double *src = tab ? tab[blockIdx.y].src : src1;
  double *dst = tab ? tab[blockIdx.y].dst : dst1;
  const int t1 = n / TS, tiles = t1 * t1;
  for (int t = blockIdx.x; t < tiles; t += gridDim.x)
    layout_tile<TO_CAPS, BS>(src, dst, n, nrec, bs, t % t1, t / t1, S);
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

// the tile kernels apply when the leaf is a multiple of 64 and bs divides 64
// or is a multiple of it
bool tiled(int m, int bs) { return m % TS == 0 && (TS % bs == 0 || bs % TS == 0); }

// one tile per workgroup (up to 65,535 per matrix), two when writing small
// blocks (measured at n = 8192: 5.8-6.0 TB/s for bs = 1 and for bs = 8 out of
// the CAPS layout, 5.3 for bs = 8 into it, against 5.1-5.4 with 8 persistent
// workgroups per CU)
template <bool TO_CAPS>
hipError_t launch_tiles(const ReorderPtrs *tab, int count, const double *src, double *dst,
                        int n, int nrec, int bs, hipStream_t s) {
  const long long tiles = (long long)(n / TS) * (n / TS);
  const long long per = (TO_CAPS && bs > 1 && bs < TS) ? 2 : 1;
  long long g = (tiles + per - 1) / per;
  if (g > 65535) g = 65535;
  if (g < 1) g = 1;
  dim3 grid((unsigned)g, (unsigned)count);
#define CAPS_LAYOUT_TILE(B)                                                                    \
  hipLaunchKernelGGL((layout_tile_kernel<TO_CAPS, B>), grid, dim3(LT), 0, s, tab, src, dst, n, \
                     nrec, bs)
  switch (bs % TS == 0 ? 0 : bs) {
    case 0: CAPS_LAYOUT_TILE(0); break;
    case 1: CAPS_LAYOUT_TILE(1); break;
    case 2: CAPS_LAYOUT_TILE(2); break;
    case 4: CAPS_LAYOUT_TILE(4); break;
    case 8: CAPS_LAYOUT_TILE(8); break;
    case 16: CAPS_LAYOUT_TILE(16); break;
    case 32: CAPS_LAYOUT_TILE(32); break;
    default: return hipErrorInvalidValue;
  }
#undef CAPS_LAYOUT_TILE
  return hipGetLastError();
}

}  // namespace

hipError_t launch_to_caps(const double *cm, double *caps, int n, int nrec, int bs,
                          hipStream_t s) {
  if (tiled(n >> nrec, bs)) return launch_tiles<true>(nullptr, 1, cm, caps, n, nrec, bs, s);
  hipLaunchKernelGGL((layout_kernel<true>), dim3(grid_for((long long)n * n, 1)), dim3(LT), 0,
                     s, cm, caps, n, nrec, bs);
  return hipGetLastError();
}

hipError_t launch_from_caps(const double *caps, double *cm, int n, int nrec, int bs,
                            hipStream_t s) {
  if (tiled(n >> nrec, bs)) return launch_tiles<false>(nullptr, 1, caps, cm, n, nrec, bs, s);
  hipLaunchKernelGGL((layout_kernel<false>), dim3(grid_for((long long)n * n, 1)), dim3(LT), 0,
                     s, caps, cm, n, nrec, bs);
  return hipGetLastError();
}

hipError_t launch_leaf_reorder(const ReorderPtrs *tab, int count, int m, int bs,
                               bool to_colmajor, hipStream_t s) {
  for (int c0 = 0; c0 < count; c0 += 65535) {
    int c = count - c0 < 65535 ? count - c0 : 65535;
    hipError_t e;
    if (tiled(m, bs)) {
      // a blocked leaf is the CAPS layout of depth 0
      e = to_colmajor ? launch_tiles<false>(tab + c0, c, nullptr, nullptr, m, 0, bs, s)
                      : launch_tiles<true>(tab + c0, c, nullptr, nullptr, m, 0, bs, s);
    } else {
      dim3 grid(grid_for((long long)m * m, c), (unsigned)c);
      if (to_colmajor)
        hipLaunchKernelGGL((leaf_reorder_kernel<true>), grid, dim3(LT), 0, s, tab + c0, m, bs);
      else
        hipLaunchKernelGGL((leaf_reorder_kernel<false>), grid, dim3(LT), 0, s, tab + c0, m, bs);
      e = hipGetLastError();
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace caps
