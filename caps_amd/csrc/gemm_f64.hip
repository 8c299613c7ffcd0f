// Batched FP64 leaf GEMM on the gfx950 matrix cores (replaces the BLAS call
// square_dgemm_zero, dgemm-blas.cpp:4-13: C = A*B, beta = 0, column-major).
//
// MFMA-bound: v_mfma_f64_16x16x4_f64 (2048 flop per instruction).
// Block tile 128 x 128 (or 64 x 64 for small leaves, the autotuner's tile
// choice), 4 waves in a 2 x 2 grid, each wave 64 x 64 = 4 x 4 MFMA tiles (64
// accumulator doubles per lane; 32 x 32 = 2 x 2 in the small tile).  K is staged through LDS in
// slices of 16 with double buffering: the next slice's global loads are in
// flight while the current slice feeds the MFMAs, one barrier per slice.
//
// LDS: A slice stored [k][row] with row stride 144 doubles and B slice stored
// [col][k] with col stride 18 doubles; with these strides the 64 lanes of an
// MFMA operand read touch every bank exactly twice (two passes of 256 bytes).
//
// One persistent launch covers all tiles of all problems of the batch (the
// 7^b sub-products of the batched levels).  Consecutive tile ids stay on the
// same XCD (blocks are dispatched round-robin over the 8 XCDs) and sweep an
// 8 x 4 patch of tiles, so one XCD's L2 sees 8 A row panels and 4 B column
// panels of one problem at a time.
#include "kernels.h"
#include "tilemap.h"

namespace caps {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

// Block tile BT x BT (BT = 128 or 64), 4 waves in a 2 x 2 grid of
// (BT / 2)^2 wave tiles; K slices of BK = 16.
template <int BT>
struct Tile {
  static constexpr int BM = BT, BN = BT, BK = 16;
  static constexpr int SA = BM + 16;  // As[k][row]
  static constexpr int SB = BK + 2;   // Bs[col][k]
  static constexpr int WT = BT / 2;   // wave tile
  static constexpr int MT = WT / 16;  // MFMA tiles per wave dimension
  static constexpr int IT = BM * BK / 2 / 256;  // v2d loads per thread per operand
};
constexpr int NT = 256;
// global -> registers: the A slice (BM rows x 16 k) and B slice (16 k x BN
// cols) of one K step, two doubles per access.  Ab / Bb point at the slice's
// first element, ld is the column stride (the leaf size).  CHECK: bounds
// against m with the slice origin (row0, k0, col0); only for plain operands.
template <int BT, bool CHECK>
__device__ __forceinline__ void gload_f64(const double *__restrict__ Ab,
                                          const double *__restrict__ Bb, long long ld, int m,
                                          int row0, int col0, int k0, int tid,
                                          v2d (&ra)[Tile<BT>::IT], v2d (&rb)[Tile<BT>::IT]) {
  using T = Tile<BT>;
  constexpr int RP = T::BM / 2;  // v2d per A column
#pragma unroll
  for (int it = 0; it < T::IT; it++) {
    const int idx = tid + NT * it;
    // A: column kc of the slice, rows 2*r2, 2*r2+1
    const int kc = idx / RP, r2 = idx % RP;
    // B: column c of the tile, k rows 2*k2, 2*k2+1
    const int c = idx >> 3, k2 = idx & 7;
    if (!CHECK) {
      ra[it] = *reinterpret_cast<const v2d *>(Ab + kc * ld + 2 * r2);
      rb[it] = *reinterpret_cast<const v2d *>(Bb + c * ld + 2 * k2);
    } else {
      const int ar = row0 + 2 * r2, ak = k0 + kc;
      const int bk = k0 + 2 * k2, bc = col0 + c;
      const bool kok = ak < m;
      ra[it].x = (kok && ar < m) ? Ab[kc * ld + 2 * r2] : 0.0;
      ra[it].y = (kok && ar + 1 < m) ? Ab[kc * ld + 2 * r2 + 1] : 0.0;
      const bool cok = bc < m;
      rb[it].x = (cok && bk < m) ? Bb[c * ld + 2 * k2] : 0.0;
      rb[it].y = (cok && bk + 1 < m) ? Bb[c * ld + 2 * k2 + 1] : 0.0;
    }
  }
}

// registers -> LDS: As[k][row], Bs[col][k]
template <int BT>
__device__ __forceinline__ void lstore_f64(double *As, double *Bs, int tid,
                                           const v2d (&ra)[Tile<BT>::IT],
                                           const v2d (&rb)[Tile<BT>::IT]) {
  using T = Tile<BT>;
  constexpr int RP = T::BM / 2;
#pragma unroll
  for (int it = 0; it < T::IT; it++) {
    const int idx = tid + NT * it;
    const int kc = idx / RP, r2 = idx % RP;
    *reinterpret_cast<v2d *>(&As[kc * T::SA + 2 * r2]) = ra[it];
    const int c = idx >> 3, k2 = idx & 7;
    *reinterpret_cast<v2d *>(&Bs[c * T::SB + 2 * k2]) = rb[it];
  }
}

// One block tile (ti, tj) of C = A * B.  Operands are m x m, stored as a
// quadtree of 2^d x 2^d column-major leaves of size m0 = m >> d in the CAPS
// quadrant order (d = 0: plain column-major); the fast path needs m0 % BT ==
// 0 so every tile and K slice lies in one leaf.
template <int BT, bool CHECK>
__device__ __forceinline__ void dgemm_tile(const GemmPtrs &P, int m, int m0, int d, int ti,
                                           int tj, double (*As)[Tile<BT>::BK * Tile<BT>::SA],
                                           double (*Bs)[Tile<BT>::BN * Tile<BT>::SB]) {
  using T = Tile<BT>;
  constexpr int BM = T::BM, BN = T::BN, BK = T::BK, SA = T::SA, SB = T::SB, WT = T::WT,
                MT = T::MT;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int l16 = lane & 15, lk = lane >> 4;
  const int nk = (m + BK - 1) / BK;
  const double *__restrict__ A = P.A;
  const double *__restrict__ B = P.B;
  double *__restrict__ C = P.C;
  const int row0 = ti * BM, col0 = tj * BN;

  v4d acc[MT][MT];
#pragma unroll
  for (int r = 0; r < MT; r++)
#pragma unroll
    for (int c = 0; c < MT; c++) acc[r][c] = v4d{0.0, 0.0, 0.0, 0.0};

  v2d ra[T::IT], rb[T::IT];
  const long long mm0 = (long long)m0 * m0;
  // first element of the A slice (row0, k0) and B slice (k0, col0)
  auto a_at = [&](int k0) {
    return A + quad_leaf(row0 / m0, k0 / m0, d) * mm0 + (row0 % m0) + (long long)(k0 % m0) * m0;
  };
  auto b_at = [&](int k0) {
    return B + quad_leaf(k0 / m0, col0 / m0, d) * mm0 + (k0 % m0) + (long long)(col0 % m0) * m0;
  };
  gload_f64<BT, CHECK>(a_at(0), b_at(0), m0, m, row0, col0, 0, tid, ra, rb);
  lstore_f64<BT>(As[0], Bs[0], tid, ra, rb);
  __syncthreads();
  for (int kt = 0; kt < nk; kt++) {
    const int buf = kt & 1;
    if (kt + 1 < nk)
      gload_f64<BT, CHECK>(a_at((kt + 1) * BK), b_at((kt + 1) * BK), m0, m, row0, col0,
                           (kt + 1) * BK, tid, ra, rb);
#pragma unroll
    for (int kk = 0; kk < BK / 4; kk++) {
      const int ka = kk * 4 + lk;
      double a[MT], b[MT];
#pragma unroll
      for (int r = 0; r < MT; r++) a[r] = As[buf][ka * SA + wm * WT + r * 16 + l16];
#pragma unroll
      for (int c = 0; c < MT; c++) b[c] = Bs[buf][(wn * WT + c * 16 + l16) * SB + ka];
#pragma unroll
      for (int r = 0; r < MT; r++)
#pragma unroll
        for (int c = 0; c < MT; c++)
          acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], b[c], acc[r][c], 0, 0, 0);
    }
    if (kt + 1 < nk) lstore_f64<BT>(As[buf ^ 1], Bs[buf ^ 1], tid, ra, rb);
    __syncthreads();
  }

  // D layout of v_mfma_f64_16x16x4 on gfx950: lane holds column (lane % 16)
  // and rows (lane / 16) + 4 v in its four result registers v = 0..3
#pragma unroll
  for (int r = 0; r < MT; r++)
#pragma unroll
    for (int c = 0; c < MT; c++) {
      const int row = row0 + wm * WT + r * 16 + lk;
      const int col = col0 + wn * WT + c * 16 + l16;
      double *dst = C + quad_leaf(row0 / m0, col0 / m0, d) * mm0 + (long long)(col % m0) * m0 +
                    (row % m0);
      if (!CHECK) {
#pragma unroll
        for (int v = 0; v < 4; v++) dst[4 * v] = acc[r][c][v];
      } else if (col < m) {
#pragma unroll
        for (int v = 0; v < 4; v++)
          if (row + 4 * v < m) dst[4 * v] = acc[r][c][v];
      }
    }
}

// One persistent launch over all tiles of all problems of one size.
template <int BT, bool CHECK>
__global__ __launch_bounds__(NT) void dgemm_kernel(const GemmPtrs *__restrict__ tab, int m,
                                                   TileMap map, int m0, int d) {
  using T = Tile<BT>;
  __shared__ double As[2][T::BK * T::SA];
  __shared__ double Bs[2][T::BN * T::SB];
  const int G = gridDim.x;
  for (int round = 0; round * G < map.total; round++) {
    const int tile = round * G + xcd_tile(blockIdx.x, G);
    if (tile >= map.total) break;
    int prob, ti, tj;
    tile_coords(map, tile, prob, ti, tj);
    dgemm_tile<BT, CHECK>(tab[prob], m, m0, d, ti, tj, As, Bs);
  }
}

// One persistent launch over the tiles of problems of several sizes (the
// leaf products of an uneven batch): segment s holds gy problems of order n
// with gx x gx tiles each, from global tile blk0 on, arg = its GemmPtrs
// table, aux = quadtree levels | edge checks << 8.
template <int BT>
__global__ __launch_bounds__(NT) void dgemm_multi_kernel(const char *__restrict__ base,
                                                         const VSeg *__restrict__ segs, int nseg,
                                                         int total) {
  using T = Tile<BT>;
  __shared__ double As[2][T::BK * T::SA];
  __shared__ double Bs[2][T::BN * T::SB];
  const int G = gridDim.x;
  for (int round = 0; round * G < total; round++) {
    const int tile = round * G + xcd_tile(blockIdx.x, G);
    if (tile >= total) break;
    const VSeg &S = segs[vseg_find(segs, nseg, tile)];
    const int m = (int)S.n, d = S.aux & 255;
    const int local = tile - S.blk0, per = S.gx * S.gx;
    const int prob = local / per, r = local - prob * per;
    const int ti = r % S.gx, tj = r / S.gx;
    const GemmPtrs &P = reinterpret_cast<const GemmPtrs *>(base + S.arg)[prob];
    if (S.aux >> 8)
      dgemm_tile<BT, true>(P, m, m >> d, d, ti, tj, As, Bs);
    else
      dgemm_tile<BT, false>(P, m, m >> d, d, ti, tj, As, Bs);
  }
}

template <int BT>
hipError_t launch_tile(const GemmPtrs *tab, int count, int m, int levels, hipStream_t s) {
  bool ok;
  TileMap map = make_tilemap(m, BT, count, ok);
  if (!ok) return hipErrorInvalidValue;
  const int m0 = m >> levels;
  const bool fast = (m0 % BT == 0);
  if (!fast && levels > 0) return hipErrorInvalidValue;  // quadtree operands need whole tiles
  // persistent grid: 2 blocks per CU (128 tiles) or 4 (64 tiles), a multiple
  // of the XCD count
  const int G = persistent_grid(BT == 128 ? 2 : 4, device_cu_count(), map.total);
  if (fast)
    hipLaunchKernelGGL((dgemm_kernel<BT, false>), dim3(G), dim3(NT), 0, s, tab, m, map, m0,
                       levels);
  else
    hipLaunchKernelGGL((dgemm_kernel<BT, true>), dim3(G), dim3(NT), 0, s, tab, m, map, m0,
                       levels);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_dgemm_batched(const GemmPtrs *tab, int count, int m, int levels, int tile,
                                hipStream_t s) {
  if (count <= 0 || m <= 0) return hipSuccess;
  const int m0 = m >> levels;
  if (levels < 0 || (m0 << levels) != m) return hipErrorInvalidValue;
  // tile 1: 64 x 64 block tiles (small leaves fill more of the device);
  // quadtree operands (levels > 0) take it only when the leaf is a multiple
  if (dgemm_block_tile(m, levels, tile) == 64) return launch_tile<64>(tab, count, m, levels, s);
  return launch_tile<128>(tab, count, m, levels, s);
}

int dgemm_block_tile(int m, int levels, int tile) {
  const int m0 = m >> levels;
  return (tile == 1 && (levels == 0 || m0 % 64 == 0)) ? 64 : 128;
}

hipError_t launch_dgemm_multi(int bt, const char *base, const VSeg *segs, int nseg, int total,
                              hipStream_t s) {
  if (nseg <= 0 || total <= 0) return hipSuccess;
  const int G = persistent_grid(bt == 128 ? 2 : 4, device_cu_count(), total);
  if (bt == 64)
    hipLaunchKernelGGL((dgemm_multi_kernel<64>), dim3(G), dim3(NT), 0, s, base, segs, nseg,
                       total);
  else
    hipLaunchKernelGGL((dgemm_multi_kernel<128>), dim3(G), dim3(NT), 0, s, base, segs, nseg,
                       total);
  return hipGetLastError();
}

}  // namespace caps
