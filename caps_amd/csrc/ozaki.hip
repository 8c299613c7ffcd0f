// Split-precision leaf kernels (Ozaki scheme II over INT8 moduli); the
// scheme is described in ozaki.h.
//
//   scale  : deterministic 2-norm upper bounds of the rows of A / columns of
//            B as (exponent, scaled sum of squares) pairs reduced in a fixed
//            order, then the power-of-two scale exponents.      HBM-bound
//   split  : A' = trunc(2^sa A), residues mod m_l -> INT8 planes (A planes
//            transposed to row-major through LDS).                HBM-bound
//   igemm  : INT8 planes -> INT32 products on v_mfma_i32_16x16x64_i8, mod m_l
//            in the epilogue, INT8 residue planes out.            MFMA-bound
//   recon  : CRT in FP64 from the residue planes, unscale.        HBM-bound
#include <climits>

#include "kernels.h"
#include "ozaki.h"
#include "ozaki_dev.h"
#include "tilemap.h"

namespace caps {

namespace {

constexpr int T256 = 256;

// ------------------------------------------------------------------ scale
// A: thread per (row, chunk of columns), sequential over the chunk
__global__ __launch_bounds__(T256) void oz_rowsq_kernel(const OzLeaf *__restrict__ tab, int m) {
  const OzLeaf &L = tab[blockIdx.z];
  const int i = blockIdx.x * T256 + threadIdx.x;
  if (i >= m) return;
  const int c = blockIdx.y;
  const int per = (m + OZ_CHUNKS - 1) / OZ_CHUNKS;
  const int k0 = c * per, k1 = min(m, k0 + per);
  int e = NOSCALE;
  double s = 0.0;
  for (int k = k0; k < k1; k++) sq_acc(L.A[(long long)k * m + i], e, s);
  reinterpret_cast<double2 *>(L.pa)[(long long)c * m + i] = make_double2(s, (double)e);
}

// B: one wave per column, lanes strided, fixed butterfly, lane 0 keeps
__global__ __launch_bounds__(T256) void oz_colsq_kernel(const OzLeaf *__restrict__ tab, int m) {
  const OzLeaf &L = tab[blockIdx.y];
  const int j = blockIdx.x * (T256 / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= m) return;
  const double *col = L.B + (long long)j * m;
  int e = NOSCALE;
  double s = 0.0;
  for (int k = lane; k < m; k += 64) sq_acc(col[k], e, s);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    int e2 = __shfl_down(e, off, 64);
    double s2 = __shfl_down(s, off, 64);
    if (lane < off) sq_merge(e, s, e2, s2);
  }
  if (lane == 0) reinterpret_cast<double2 *>(L.pb)[j] = make_double2(s, (double)e);
}

// exponents: ea from the OZ_CHUNKS partials in chunk order, eb from pb
__global__ __launch_bounds__(T256) void oz_finalize_kernel(const OzLeaf *__restrict__ tab, int m,
                                                           const OzakiConsts *__restrict__ kc) {
  const OzLeaf &L = tab[blockIdx.y];
  const int i = blockIdx.x * T256 + threadIdx.x;
  if (i >= m) return;
  const double2 *pa = reinterpret_cast<const double2 *>(L.pa);
  int e = NOSCALE;
  double s = 0.0;
  for (int c = 0; c < OZ_CHUNKS; c++) {
    double2 v = pa[(long long)c * m + i];
    sq_merge(e, s, (int)v.y, v.x);
  }
  L.ea[i] = scale_exponent(e, s, kc->L);
  double2 v = reinterpret_cast<const double2 *>(L.pb)[i];
  L.eb[i] = scale_exponent((int)v.y, v.x, kc->L);
}

// ------------------------------------------------------------------ split
// A (column-major) -> nmod row-major INT8 planes; 64 x 64 tiles through LDS.
template <bool CHECK>
__global__ __launch_bounds__(T256) void oz_split_a_kernel(const OzLeaf *__restrict__ tab, int m,
                                                          int nmod,
                                                          const OzakiConsts *__restrict__ kc) {
  __shared__ double T[64][65];  // [col][row]
  const OzLeaf &L = tab[blockIdx.y];
  const int tiles = (m + 63) / 64;
  const int ti = blockIdx.x % tiles, tk = blockIdx.x / tiles;
  const int r0 = ti * 64, k0 = tk * 64;
  // load: lane -> row, loop over columns (coalesced 512 B per column)
  for (int idx = threadIdx.x; idx < 64 * 64; idx += T256) {
    const int r = idx & 63, c = idx >> 6;
    const int gr = r0 + r, gc = k0 + c;
    T[c][r] = (!CHECK || (gr < m && gc < m)) ? L.A[(long long)gc * m + gr] : 0.0;
  }
  __syncthreads();
  // each thread: one row, 16 consecutive k
  const int r = threadIdx.x >> 2, kq = (threadIdx.x & 3) * 16;
  const int gr = r0 + r;
  if (CHECK && gr >= m) return;
  const int sa = L.ea[gr];
  double x[16];
#pragma unroll
  for (int c = 0; c < 16; c++) x[c] = trunc(ldexp(T[kq + c][r], sa));
  int8_t *base = L.ra + (long long)gr * m + k0 + kq;
  const long long plane = (long long)m * m;
  for (int l = 0; l < nmod; l++) {
    const int mi = kc->m[l];
    const double inv = kc->inv_m[l];
    const int c22 = kc->c22[l];
    if (!CHECK) {
      unsigned w[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        unsigned v = 0;
#pragma unroll
        for (int b = 0; b < 4; b++)
          v |= (unsigned)(sym_residue(x[4 * q + b], mi, inv, c22) & 0xff) << (8 * b);
        w[q] = v;
      }
      *reinterpret_cast<uint4 *>(base + l * plane) = make_uint4(w[0], w[1], w[2], w[3]);
    } else {
#pragma unroll
      for (int c = 0; c < 16; c++)
        if (k0 + kq + c < m) base[l * plane + c] = (int8_t)sym_residue(x[c], mi, inv, c22);
    }
  }
}

// B (column-major) -> nmod column-major INT8 planes; thread per 16 rows of a column
template <bool CHECK>
__global__ __launch_bounds__(T256) void oz_split_b_kernel(const OzLeaf *__restrict__ tab, int m,
                                                          int nmod,
                                                          const OzakiConsts *__restrict__ kc) {
  const OzLeaf &L = tab[blockIdx.y];
  const long long total16 = ((long long)m * m + 15) / 16;
  const long long plane = (long long)m * m;
  for (long long u = (long long)blockIdx.x * T256 + threadIdx.x; u < total16;
       u += (long long)gridDim.x * T256) {
    const long long e0 = u * 16;
    double x[16];
    int sb[16];
    if (!CHECK) {
      // m % 16 == 0: the 16 elements share a column
      const int j = (int)(e0 / m);
      const int s = L.eb[j];
#pragma unroll
      for (int c = 0; c < 16; c += 2) {
        double2 v = *reinterpret_cast<const double2 *>(L.B + e0 + c);
        x[c] = trunc(ldexp(v.x, s));
        x[c + 1] = trunc(ldexp(v.y, s));
      }
      for (int l = 0; l < nmod; l++) {
        const int mi = kc->m[l];
        const double inv = kc->inv_m[l];
        const int c22 = kc->c22[l];
        unsigned w[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          unsigned v = 0;
#pragma unroll
          for (int b = 0; b < 4; b++)
            v |= (unsigned)(sym_residue(x[4 * q + b], mi, inv, c22) & 0xff) << (8 * b);
          w[q] = v;
        }
        *reinterpret_cast<uint4 *>(L.rb + l * plane + e0) = make_uint4(w[0], w[1], w[2], w[3]);
      }
    } else {
#pragma unroll
      for (int c = 0; c < 16; c++) {
        const long long e = e0 + c;
        sb[c] = e < plane ? L.eb[(int)(e / m)] : 0;
        x[c] = e < plane ? trunc(ldexp(L.B[e], sb[c])) : 0.0;
      }
      for (int l = 0; l < nmod; l++) {
        const int mi = kc->m[l];
        const double inv = kc->inv_m[l];
        const int c22 = kc->c22[l];
#pragma unroll
        for (int c = 0; c < 16; c++)
          if (e0 + c < plane) L.rb[l * plane + e0 + c] = (int8_t)sym_residue(x[c], mi, inv, c22);
      }
    }
  }
}

// ------------------------------------------------------------------ INT8 GEMM
typedef int v4i __attribute__((ext_vector_type(4)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));

constexpr int IBM = 128, IBN = 128, IBK = 64;
constexpr int ISTR = IBK + 16;  // LDS row stride (bytes): b128 reads of 16 lanes hit 64 banks

// global -> registers: 128 rows of A (k contiguous) and 128 columns of B
// (k contiguous), 64 k each, 16 bytes per access.  Ab / Bb point at the
// slice's first element, ld is the row (A) / column (B) stride.  CHECK 0:
// none; 1: every byte against m; 2: rows / columns against m (m % 64 == 0,
// so K slices are whole); 3: rows / columns and 16-byte K chunks against m
// (m % 16 == 0, so a chunk is wholly inside or outside).  Checks use the
// slice origin; plain operands only.
typedef __attribute__((address_space(1))) unsigned gu32;  // stores as global_*, not flat_*
typedef __attribute__((address_space(1))) v4i gv4i;
typedef const __attribute__((address_space(1))) v4u gv4u;

template <int CHECK>
__device__ __forceinline__ void gload_i8(const int8_t *__restrict__ Ab,
                                         const int8_t *__restrict__ Bb, long long ld, int m,
                                         int row0, int col0, int k0, int tid, v4u (&ga)[2],
                                         v4u (&gb)[2]) {
#pragma unroll
  for (int it = 0; it < 2; it++) {
    const int idx = tid + T256 * it;
    const int r = idx >> 2, kq = (idx & 3) * 16;
    if (CHECK == 0) {
      ga[it] = *reinterpret_cast<const v4u *>(Ab + r * ld + kq);
      gb[it] = *reinterpret_cast<const v4u *>(Bb + r * ld + kq);
    } else if (CHECK == 2) {
      // the K slice is whole; rows of A / columns of B beyond m read as zero
      const v4u z = {0u, 0u, 0u, 0u};
      ga[it] = row0 + r < m ? *reinterpret_cast<const v4u *>(Ab + r * ld + kq) : z;
      gb[it] = col0 + r < m ? *reinterpret_cast<const v4u *>(Bb + r * ld + kq) : z;
    } else if (CHECK == 3) {
      // 16-byte chunks beyond K, rows of A / columns of B beyond m read as zero
      const v4u z = {0u, 0u, 0u, 0u};
      const bool kin = k0 + kq < m;
      ga[it] = kin && row0 + r < m ? *reinterpret_cast<const v4u *>(Ab + r * ld + kq) : z;
      gb[it] = kin && col0 + r < m ? *reinterpret_cast<const v4u *>(Bb + r * ld + kq) : z;
    } else {
      unsigned wa[4], wb[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        unsigned va = 0, vb = 0;
#pragma unroll
        for (int b = 0; b < 4; b++) {
          const int k = k0 + kq + 4 * q + b;
          if (k < m && row0 + r < m)
            va |= (unsigned)(uint8_t)Ab[r * ld + kq + 4 * q + b] << (8 * b);
          if (k < m && col0 + r < m)
            vb |= (unsigned)(uint8_t)Bb[r * ld + kq + 4 * q + b] << (8 * b);
        }
        wa[q] = va;
        wb[q] = vb;
      }
      ga[it] = v4u{wa[0], wa[1], wa[2], wa[3]};
      gb[it] = v4u{wb[0], wb[1], wb[2], wb[3]};
    }
  }
}

__device__ __forceinline__ void lstore_i8(int8_t *As, int8_t *Bs, int tid, const v4u (&ga)[2],
                                          const v4u (&gb)[2]) {
#pragma unroll
  for (int it = 0; it < 2; it++) {
    const int idx = tid + T256 * it;
    const int r = idx >> 2, kq = (idx & 3) * 16;
    *reinterpret_cast<v4u *>(&As[r * ISTR + kq]) = ga[it];
    *reinterpret_cast<v4u *>(&Bs[r * ISTR + kq]) = gb[it];
  }
}

// One 128 x 128 tile (ti, tj) of problem p = leaf * nmod + l: C_l = A_l *
// B_l (mod m_l).  Operands are m x m quadtrees of d levels over leaves of size
// m0 = m >> d in CAPS quadrant order: A leaves row-major, B and C leaves
// column-major (d = 0: plain matrices).  The fast path needs m0 % 128 == 0.
template <int CHECK>
__device__ __forceinline__ void oz_tile(const RGemm *__restrict__ tab, int m, int nmod,
                                        const OzakiConsts *__restrict__ kc, int m0, int d,
                                        int prob, int ti, int tj, int8_t (*As)[IBM * ISTR],
                                        int8_t (*Bs)[IBN * ISTR]) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int l16 = lane & 15, lk = lane >> 4;
  const int nk = (m + IBK - 1) / IBK;
  const int leaf = prob / nmod, l = prob - leaf * nmod;
  const int8_t *__restrict__ Ap = tab[leaf].a + l * tab[leaf].sa;  // row-major
  const int8_t *__restrict__ Bp = tab[leaf].b + l * tab[leaf].sb;  // column-major
  int8_t *__restrict__ Cp = tab[leaf].c + l * tab[leaf].sc;
  const int mi = kc->m[l];
  const float mf = (float)mi, invf = 1.0f / mf;
  const int row0 = ti * IBM, col0 = tj * IBN;

  // small K: the sums start at 128 for the one-pass reduction res_byte128
  const bool small_k = m <= 16384;
  const int bias = small_k ? 128 : 0;
  v4i acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[r][c] = v4i{bias, bias, bias, bias};

  v4u ga[2], gb[2];
  const long long mm0 = (long long)m0 * m0;
  auto a_at = [&](int k0) {  // A slice origin (row0, k0), row-major leaf
    return Ap + quad_leaf(row0 / m0, k0 / m0, d) * mm0 + (long long)(row0 % m0) * m0 + (k0 % m0);
  };
  auto b_at = [&](int k0) {  // B slice origin (k0, col0), column-major leaf
    return Bp + quad_leaf(k0 / m0, col0 / m0, d) * mm0 + (k0 % m0) + (long long)(col0 % m0) * m0;
  };
  gload_i8<CHECK>(a_at(0), b_at(0), m0, m, row0, col0, 0, tid, ga, gb);
  lstore_i8(As[0], Bs[0], tid, ga, gb);
  __syncthreads();
  for (int kt = 0; kt < nk; kt++) {
    const int buf = kt & 1;
    if (kt + 1 < nk)
      gload_i8<CHECK>(a_at((kt + 1) * IBK), b_at((kt + 1) * IBK), m0, m, row0, col0,
                      (kt + 1) * IBK, tid, ga, gb);
    v4i a[4], b[4];
    // lane group lk supplies k = 16*lk .. 16*lk+15 for both operands
#pragma unroll
    for (int r = 0; r < 4; r++)
      a[r] = *reinterpret_cast<const v4i *>(&As[buf][(wm * 64 + r * 16 + l16) * ISTR + 16 * lk]);
#pragma unroll
    for (int c = 0; c < 4; c++)
      b[c] = *reinterpret_cast<const v4i *>(&Bs[buf][(wn * 64 + c * 16 + l16) * ISTR + 16 * lk]);
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
      for (int c = 0; c < 4; c++)
        acc[r][c] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[r], b[c], acc[r][c], 0, 0, 0);
    if (kt + 1 < nk) lstore_i8(As[buf ^ 1], Bs[buf ^ 1], tid, ga, gb);
    __syncthreads();
  }

  // D layout: lane holds column (lane % 16), rows 4 * (lane / 16) + v
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int row = row0 + wm * 64 + r * 16 + 4 * lk;
      const int col = col0 + wn * 64 + c * 16 + l16;
      int8_t *dst = Cp + quad_leaf(row0 / m0, col0 / m0, d) * mm0 + (long long)(col % m0) * m0 +
                    (row % m0);
      unsigned v = 0;
      if (small_k) {
        const int c14m = kc->c14[l] - 16384;
        const float cl = 0.5f - 128.0f * invf;
#pragma unroll
        for (int q = 0; q < 4; q++) v = res_byte128(acc[r][c][q], c14m, mf, invf, cl, q, v);
        v ^= 0x80808080u;
      } else {
#pragma unroll
        for (int q = 0; q < 4; q++) v |= (unsigned)(mod_i32(acc[r][c][q], mi, invf) & 0xff) << (8 * q);
      }
      if (CHECK == 0) {
        *(gu32 *)dst = v;
      } else if (CHECK >= 2) {
        // m % 16 == 0: the 4 rows of the word are all inside or all outside
        if (col < m && row < m) *(gu32 *)dst = v;
      } else if (col < m) {
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (row + q < m) dst[q] = (int8_t)(v >> (8 * q));
      }
    }
}

template <int CHECK>
__global__ __launch_bounds__(T256) void oz_igemm_kernel(const RGemm *__restrict__ tab, int m,
                                                        int nmod,
                                                        const OzakiConsts *__restrict__ kc,
                                                        TileMap map, int m0, int d) {
  __shared__ __attribute__((aligned(16))) int8_t As[2][IBM * ISTR];
  __shared__ __attribute__((aligned(16))) int8_t Bs[2][IBN * ISTR];
  const int G = gridDim.x;
  for (int round = 0; round * G < map.total; round++) {
    const int tile = round * G + xcd_tile(blockIdx.x, G);
    if (tile >= map.total) break;
    int prob, ti, tj;
    tile_coords(map, tile, prob, ti, tj);
    oz_tile<CHECK>(tab, m, nmod, kc, m0, d, prob, ti, tj, As, Bs);
  }
}

// The residue products of several sizes in one persistent launch: segment
// args are RGemm tables, n = order, gx = tiles per dimension, gy = problems
// (leaves x moduli), aux = quadtree levels | check mode << 8.
__global__ __launch_bounds__(T256) void oz_igemm_multi_kernel(const char *__restrict__ base,
                                                              const VSeg *__restrict__ segs,
                                                              int nseg, int total, int nmod,
                                                              const OzakiConsts *__restrict__ kc) {
  __shared__ __attribute__((aligned(16))) int8_t As[2][IBM * ISTR];
  __shared__ __attribute__((aligned(16))) int8_t Bs[2][IBN * ISTR];
  const int G = gridDim.x;
  for (int round = 0; round * G < total; round++) {
    const int tile = round * G + xcd_tile(blockIdx.x, G);
    if (tile >= total) break;
    const VSeg &S = segs[vseg_find(segs, nseg, tile)];
    const int m = (int)S.n, d = S.aux & 255, check = S.aux >> 8;
    const int local = tile - S.blk0, per = S.gx * S.gx;
    const int prob = local / per, r = local - prob * per;
    const int ti = r % S.gx, tj = r / S.gx;
    const RGemm *tab = reinterpret_cast<const RGemm *>(base + S.arg);
    switch (check) {
      case 0: oz_tile<0>(tab, m, nmod, kc, m >> d, d, prob, ti, tj, As, Bs); break;
      case 2: oz_tile<2>(tab, m, nmod, kc, m >> d, d, prob, ti, tj, As, Bs); break;
      case 3: oz_tile<3>(tab, m, nmod, kc, m >> d, d, prob, ti, tj, As, Bs); break;
      default: oz_tile<1>(tab, m, nmod, kc, m >> d, d, prob, ti, tj, As, Bs); break;
    }
  }
}

// ------------------------------------------------------------------ INT8 GEMM, large tiles
// 256 x 256 block tiles, 8 waves (2 per SIMD) of 128 x 64 (32 accumulator
// tiles of 16 x 16 per wave), K-steps of 128.  Operands go global -> LDS
// directly (MUBUF buffer_load ... lds: each lane's 16 bytes land at the
// wave's LDS base + 16 * lane), double-buffered, so no registers hold the
// next stage.  Per MAC this reads 3/4 of the LDS bytes and half the L2 bytes
// of the 128 x 128 kernel above.  Leaves of m0 % 256 == 0, or (EDGE) plain
// leaves of m % 16 == 0 with partial last tiles.
//
// LDS stage layout per operand: 256 rows x 8 chunks of 16 bytes (k), no
// padding; chunk c of row r sits in slot c ^ swz(r), swz(r) = (r >> 1) & 7.
// The 16 lanes of one ds_read_b128 cycle read one chunk of rows l16 = 0..15;
// two consecutive 128-byte rows span the 256 bytes of the 64 banks, and swz
// gives the 8 rows of each parity 8 distinct slots, so the 16 lanes hit
// distinct banks.  Measured on one launch of 686 GEMMs of 2048^3:
// SQ_LDS_BANK_CONFLICT 2.2e7 cycles, against 5.6e8 with swz(r) = ((r >> 2) &
// 3) << 1 (distinct slots modulo 512 bytes, but not modulo 256); the INT8
// GEMM 0.7 % and the BF16 one 2.5 % faster.
//
// Latency and ordering, each choice measured on 49 x 14 GEMMs of 2048^3:
//  * the LDS-bound loads are MUBUF, not global_load_lds: with the FLAT-class
//    form in the loop the compiler waits for every LDS read with
//    lgkmcnt(0); with MUBUF it waits for exactly the fragment an MFMA needs;
//  * the barrier closing a K-step is a bare s_barrier after this wave's own
//    s_waitcnt vmcnt: __syncthreads adds a release fence that waits for every
//    outstanding load, including the L2 prefetch below;
//  * only the fragments of the first MFMAs are read before they start; the
//    other reads, the second half-step's included, run one per two MFMAs
//    (sched_group_barrier interleave);
//  * an L2 prefetch two K-steps ahead: the 32 tiles one XCD runs at a time
//    form an 8 x 4 patch of one problem (xcd_tile, tile_coords), so each of
//    its blocks touches 1/32 of the patch's next A and B slabs (4 bytes per
//    128-byte line) and the whole XCD finds them in L2;
//  * the next tile's first stage is loaded before the epilogue.
constexpr int JBM = 256, JBK = 128, JT = 512;
constexpr int JSTAGE = JBM * JBK;  // 32 KB per operand per stage
constexpr int JPF = 2;             // prefetch distance, K-steps

__device__ __forceinline__ int jswz(int r) { return (r >> 1) & 7; }

typedef __attribute__((address_space(3))) void *lds_ptr;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const int8_t *p, long long bytes) {
  // raw buffer over one operand's quadtree (bytes < 2^31 for m <= 46340)
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<int8_t *>(p), 0, (int)bytes, 0x00020000);
}

// one stage: wave w loads rows 8 (8 j + w) .. + 7 of A and of B, 8 lanes per
// 128-byte row; for those rows swz(r) = ((w & 1) << 2) | (rl >> 1).
// aoff / boff: byte offsets of the slab origin in the operand.
__device__ __forceinline__ void stage_load(__amdgpu_buffer_rsrc_t ra, __amdgpu_buffer_rsrc_t rb,
                                           int aoff, int boff, int ld, int8_t *As, int8_t *Bs,
                                           int wave, int lane) {
  const int rl = lane >> 3;
  // chunk this lane fetches (same for every j): slot (lane & 7) holds chunk slot ^ swz(r)
  const int c = (lane & 7) ^ (((wave & 1) << 2) | (rl >> 1));
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int r = j * 64 + wave * 8 + rl;
    const int slot0 = (j * 8 + wave) * 64;  // first 16-byte slot of this wave's 1 KB
    __builtin_amdgcn_raw_ptr_buffer_load_lds(ra, (lds_ptr)(As + slot0 * 16), 16, r * ld + c * 16,
                                             aoff, 0, 0);
    __builtin_amdgcn_raw_ptr_buffer_load_lds(rb, (lds_ptr)(Bs + slot0 * 16), 16, r * ld + c * 16,
                                             boff, 0, 0);
  }
}

// stage_load for plain m x m operands (d = 0) of any m % 16 == 0: rows /
// columns at or beyond m and 16-byte K chunks at or beyond m read as zero
// (their buffer offset is the plane size, past num_records: the buffer range
// check returns zeros without a memory access)
__device__ __forceinline__ void stage_load_edge(__amdgpu_buffer_rsrc_t ra,
                                                __amdgpu_buffer_rsrc_t rb, int row0, int col0,
                                                int k, int m, int8_t *As, int8_t *Bs, int wave,
                                                int lane) {
  const int rl = lane >> 3;
  const int c = (lane & 7) ^ (((wave & 1) << 2) | (rl >> 1));
  const int oob = m * m;
  const bool kin = k + c * 16 < m;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int r = j * 64 + wave * 8 + rl;
    const int slot0 = (j * 8 + wave) * 64;
    const int ao = (kin && row0 + r < m) ? (row0 + r) * m + k + c * 16 : oob;
    const int bo = (kin && col0 + r < m) ? (col0 + r) * m + k + c * 16 : oob;
    __builtin_amdgcn_raw_ptr_buffer_load_lds(ra, (lds_ptr)(As + slot0 * 16), 16, ao, 0, 0, 0);
    __builtin_amdgcn_raw_ptr_buffer_load_lds(rb, (lds_ptr)(Bs + slot0 * 16), 16, bo, 0, 0, 0);
  }
}

// block barrier for LDS stages filled by this wave's own loads: the caller
// has waited for them (vmcnt) and consumed every LDS read of the stage it
// leaves; the empty asm statements keep memory operations on their side
__device__ __forceinline__ void stage_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

struct BigTile {
  const int8_t *Ap, *Bp;
  int8_t *Cp;
  __amdgpu_buffer_rsrc_t ra, rb;  // buffer descriptors of the operand planes
  int mi, c14, c22, row0, col0, ti, tj;
  float mf, invf;
  long long cseg;  // C segment stride (RGemm::cseg)
};

// a uniform pointer kept in scalar registers (buffer descriptors built from a
// pointer the compiler thinks may vary per lane are used through per-lane
// loops)
template <typename P>
__device__ __forceinline__ P *uniform_ptr(P *p) {
  const unsigned long long v = reinterpret_cast<unsigned long long>(p);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return reinterpret_cast<P *>(((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ BigTile big_tile(const RGemm *__restrict__ tab, const TileMap &map,
                                            int tile, int nmod,
                                            const OzakiConsts *__restrict__ kc, long long opb) {
  BigTile t;
  int prob, ti, tj;
  tile_coords(map, tile, prob, ti, tj);
  const int leaf = prob / nmod, l = prob - leaf * nmod;
  t.Ap = uniform_ptr(tab[leaf].a + l * tab[leaf].sa);  // row-major leaves
  t.Bp = uniform_ptr(tab[leaf].b + l * tab[leaf].sb);  // column-major leaves
  t.Cp = tab[leaf].c + l * tab[leaf].sc;
  t.cseg = tab[leaf].cseg;
  t.mi = kc->m[l];
  t.mf = (float)t.mi;
  t.invf = 1.0f / t.mf;
  t.c14 = kc->c14[l];
  t.c22 = kc->c22[l];
  t.row0 = ti * JBM;
  t.col0 = tj * JBM;
  t.ti = ti;
  t.tj = tj;
  t.ra = plane_rsrc(t.Ap, opb);
  t.rb = plane_rsrc(t.Bp, opb);
  return t;
}

// SMALL_K: K = m <= 16384, so |sums| <= m 2^14 < 2^28 and the one-pass
// exact reduction res_byte128 applies (sums start at 128); larger K takes the
// general one
// EDGE: plain operands (d = 0) whose size m is a multiple of 16 but not of
// 256: the last tile row / column and K-step are partial, their missing
// elements read as zero and the stores beyond m are skipped.
template <bool SMALL_K, bool EDGE>
__global__ __launch_bounds__(JT) __attribute__((amdgpu_waves_per_eu(2, 2))) void
oz_igemm_big_kernel(const RGemm *__restrict__ tab, int m, int nmod,
                    const OzakiConsts *__restrict__ kc, TileMap map, int m0, int d) {
  __shared__ __attribute__((aligned(1024))) int8_t A0[JSTAGE];
  __shared__ __attribute__((aligned(1024))) int8_t A1[JSTAGE];
  __shared__ __attribute__((aligned(1024))) int8_t B0[JSTAGE];
  __shared__ __attribute__((aligned(1024))) int8_t B1[JSTAGE];
  __shared__ __attribute__((aligned(256))) unsigned PF[2][64];  // prefetch sink
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;  // 2 x 4 waves: rows wm * 128, columns wn * 64
  const int l16 = lane & 15, lk = lane >> 4;
  const int G = gridDim.x;
  const int nk = (m + JBK - 1) / JBK;
  constexpr int bias = SMALL_K ? 128 : 0;
  // read offsets (bytes) of this lane's row l16, chunk lk (kh = 0) and lk + 4 (kh = 1)
  const int s16 = jswz(l16);
  const int rd0 = l16 * JBK + ((lk ^ s16) << 4);
  const int rd1 = l16 * JBK + (((lk + 4) ^ s16) << 4);
  const int aoff = wm * 128 * JBK, boff = wn * 64 * JBK;
  const long long mm0 = (long long)m0 * m0;
  const long long opb = (long long)m * m;  // bytes of one operand quadtree
  // the prefetch needs the 8 x 4 patch order of tile_coords
  const bool pf_wave = map.patch && wave < 2;

  int tile = xcd_tile(blockIdx.x, G);
  if (tile >= map.total) return;
  BigTile t = big_tile(tab, map, tile, nmod, kc, opb);
  // offsets of row `row` / column `col` at k in the quadtree operands; plain
  // operands (d == 0) need no division by the leaf size
  auto a_off = [&](int row, int k) -> long long {
    return d == 0 ? (long long)row * m0 + k
                  : quad_leaf(row / m0, k / m0, d) * mm0 + (long long)(row % m0) * m0 + (k % m0);
  };
  auto b_off = [&](int col, int k) -> long long {
    return d == 0 ? (long long)col * m0 + k
                  : quad_leaf(k / m0, col / m0, d) * mm0 + (long long)(col % m0) * m0 + (k % m0);
  };
  // waves 0 and 1: this block's 1/32 share of the patch's A / B slab at k
  auto prefetch = [&](const BigTile &c, int k) {
    const int w = (c.ti & 7) + 8 * (c.tj & 3);
    const int pr0 = (c.ti & ~7) * JBM, pc0 = (c.tj & ~3) * JBM;
    if (wave == 0)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(c.ra, (lds_ptr)&PF[0][0], 4,
                                               (int)a_off(pr0 + w * 64 + lane, k), 0, 0, 0);
    else
      __builtin_amdgcn_raw_ptr_buffer_load_lds(c.rb, (lds_ptr)&PF[1][0], 4,
                                               (int)b_off(pc0 + w * 32 + (lane & 31), k), 0, 0, 0);
  };
  auto load = [&](const BigTile &c, int k, int8_t *As, int8_t *Bs) {
    if (EDGE)
      stage_load_edge(c.ra, c.rb, c.row0, c.col0, k, m, As, Bs, wave, lane);
    else
      stage_load(c.ra, c.rb, (int)a_off(c.row0, k), (int)b_off(c.col0, k), m0, As, Bs, wave, lane);
  };
  load(t, 0, A0, B0);
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();

  for (int round = 0;; round++) {
    v4i acc[8][4];
#pragma unroll
    for (int r = 0; r < 8; r++)
#pragma unroll
      for (int c = 0; c < 4; c++) acc[r][c] = v4i{bias, bias, bias, bias};
    // scalar loads of the tile context complete here: a pending one would
    // turn every LDS wait of the K loop into lgkmcnt(0)
    __builtin_amdgcn_s_waitcnt(0xc07f);
    for (int kt = 0; kt < nk; kt++) {
      const bool odd = kt & 1;
      if (kt + 1 < nk) load(t, (kt + 1) * JBK, odd ? A0 : A1, odd ? B0 : B1);
      bool pf = false;
      if (pf_wave && kt + JPF < nk) {
        prefetch(t, (kt + JPF) * JBK);
        pf = true;
      }
      const int8_t *Ab = (odd ? A1 : A0) + aoff;
      const int8_t *Bb = (odd ? B1 : B0) + boff;
      v4i a0[8], b0[4], a1[8], b1[4];
#pragma unroll
      for (int c = 0; c < 4; c++) b0[c] = *reinterpret_cast<const v4i *>(Bb + rd0 + c * 16 * JBK);
#pragma unroll
      for (int r = 0; r < 8; r++) a0[r] = *reinterpret_cast<const v4i *>(Ab + rd0 + r * 16 * JBK);
#pragma unroll
      for (int c = 0; c < 4; c++) b1[c] = *reinterpret_cast<const v4i *>(Bb + rd1 + c * 16 * JBK);
#pragma unroll
      for (int r = 0; r < 8; r++) a1[r] = *reinterpret_cast<const v4i *>(Ab + rd1 + r * 16 * JBK);
#pragma unroll
      for (int r = 0; r < 8; r++)
#pragma unroll
        for (int c = 0; c < 4; c++)
          acc[r][c] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a0[r], b0[c], acc[r][c], 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 8; r++)
#pragma unroll
        for (int c = 0; c < 4; c++)
          acc[r][c] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a1[r], b1[c], acc[r][c], 0, 0, 0);
      // order: the stage and prefetch loads, the 6 reads the first MFMAs need
      // (b0, a0[0..1]), then the other 18 reads one per 2 MFMAs, then the rest.
      // After a barrier all 8 waves read at once; issuing 6 reads instead of a
      // whole half-step (12) before the first MFMA shortens that queue (1.5 %
      // faster than reading the first half-step up front)
      __builtin_amdgcn_sched_group_barrier(0x020, 9, 0);
      __builtin_amdgcn_sched_group_barrier(0x100, 6, 0);
#pragma unroll
      for (int i = 0; i < 18; i++) {
        __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      }
      __builtin_amdgcn_sched_group_barrier(0x008, 28, 0);
      // this wave's stage loads are done; its prefetch (issued after them)
      // may still be in flight
      if (pf)
        asm volatile("s_waitcnt vmcnt(1)" ::: "memory");
      else
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      stage_barrier();
    }
    // after the K loop's last barrier both stages are free: the next tile's
    // first stage goes to A0 / B0 before this tile's epilogue, which stages C
    // in A1 / B1
    const int next = (round + 1) * G + xcd_tile(blockIdx.x, G);
    const bool has_next = next < map.total;
    BigTile tn = t;
    if (has_next) tn = big_tile(tab, map, next, nmod, kc, opb);
    if (has_next) load(tn, 0, A0, B0);
    // D layout: lane holds column (lane % 16), rows 4 * (lane / 16) + v of
    // each 16 x 16 block.  The tile's C bytes go through LDS (A1: columns
    // 0..127, B1: 128..255; column c at word 64 c, its word w at w ^ 4 (c &
    // 15), so the 16 columns of one store hit distinct banks) and leave as
    // whole 256-byte columns, 16 bytes per lane.
    unsigned *Ch = reinterpret_cast<unsigned *>(wn < 2 ? A1 : B1);
    auto put = [&](int r, int c, unsigned v) {
      const int col = (wn & 1) * 64 + c * 16 + l16;
      const int w = wm * 32 + r * 4 + lk;
      Ch[col * 64 + (w ^ (4 * l16))] = v;
    };
    if constexpr (SMALL_K) {
      const int c14m = t.c14 - 16384;
      const float cl = 0.5f - 128.0f * t.invf;
#pragma unroll
      for (int r = 0; r < 8; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
          unsigned v = 0;
#pragma unroll
          for (int q = 0; q < 4; q++) v = res_byte128(acc[r][c][q], c14m, t.mf, t.invf, cl, q, v);
          put(r, c, v ^ 0x80808080u);
        }
    } else {
#pragma unroll
      for (int r = 0; r < 8; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
          unsigned v = 0;
#pragma unroll
          for (int q = 0; q < 4; q++)
            v |= (unsigned)(mod_i32(acc[r][c][q], t.mi, t.invf) & 0xff) << (8 * q);
          put(r, c, v);
        }
    }
    // LDS writes done before the barrier; the next tile's first stage (in
    // A0 / B0) may still be loading
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    stage_barrier();
    // lane: 16 bytes (chunk j16) of tile column pass * 32 + (tid >> 4)
    // element e of the C plane at (e >> 7) cseg + (e & 127); a 16-byte chunk
    // (e % 16 == 0) stays in one segment, and a pass moves 32 m0 elements, a
    // multiple of 128 (m0 % 16 == 0)
    const int j16 = tid & 15;
    const long long e0 = quad_leaf(t.row0 / m0, t.col0 / m0, d) * mm0 +
                         (long long)(t.col0 % m0 + (tid >> 4)) * m0 + (t.row0 % m0) + 16 * j16;
    int8_t *Cl = t.Cp + (e0 >> 7) * t.cseg + (e0 & 127);
    const long long pstep = (long long)(m0 >> 2) * t.cseg;
#pragma unroll
    for (int pass = 0; pass < 8; pass++) {
      const int ccl = (pass * 32 + (tid >> 4)) & 127;  // pass < 4: A1, else B1
      const unsigned *Cs = reinterpret_cast<const unsigned *>(pass < 4 ? A1 : B1);
      const v4i v = *reinterpret_cast<const v4i *>(Cs + ccl * 64 + ((4 * j16) ^ (4 * (ccl & 15))));
      // EDGE: m % 16 == 0, so a 16-row chunk is wholly inside or outside
      if (!EDGE || (t.col0 + pass * 32 + (tid >> 4) < m && t.row0 + 16 * j16 < m))
        *(gv4i *)(Cl + pass * pstep) = v;
    }
    if (!has_next) break;
    t = tn;
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    stage_barrier();
  }
}

// ------------------------------------------------------------------ BF16 GEMM on residues
// The BF16 variant of the residue GEMM: INT8 residues (|r| <= 128, exact in
// BF16) are widened when staged into LDS and multiplied on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation.  FP32 holds integers
// exactly below 2^24: after a fold x - m rint(x / m) (|result| <= 1.5 m) at
// most 512 more products of magnitude <= 2^14 are added, so the running sum
// stays below 2^24 and every partial sum is exact.  The result equals the
// INT8 kernel's bit for bit.  128 x 128 tiles, 4 waves of 64 x 64, BK = 64.
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int BSTR = IBK * 2 + 16;  // LDS row stride (bytes) of BF16 rows
constexpr int BFOLD = 512;          // K between folds

// 16 INT8 residues -> 16 BF16 values (8 dwords): exact widening through FP32
__device__ __forceinline__ void widen16(v4u w, unsigned (&o)[8]) {
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int h = 0; h < 2; h++) {
      // (the builtin's result type is unsigned: take it as int before widening)
      const float f0 = (float)(int)__builtin_amdgcn_sbfe(w[q], 16 * h, 8);
      const float f1 = (float)(int)__builtin_amdgcn_sbfe(w[q], 16 * h + 8, 8);
      o[2 * q + h] = (__float_as_uint(f0) >> 16) | (__float_as_uint(f1) & 0xffff0000u);
    }
}

__device__ __forceinline__ void lstore_bf(int8_t *As, int8_t *Bs, int tid, const v4u (&ga)[2],
                                          const v4u (&gb)[2]) {
#pragma unroll
  for (int it = 0; it < 2; it++) {
    const int idx = tid + T256 * it;
    const int r = idx >> 2, kq = (idx & 3) * 16;  // 16 k values = 32 bytes
    unsigned o[8];
    widen16(ga[it], o);
    *reinterpret_cast<v4u *>(&As[r * BSTR + 2 * kq]) = v4u{o[0], o[1], o[2], o[3]};
    *reinterpret_cast<v4u *>(&As[r * BSTR + 2 * kq + 16]) = v4u{o[4], o[5], o[6], o[7]};
    widen16(gb[it], o);
    *reinterpret_cast<v4u *>(&Bs[r * BSTR + 2 * kq]) = v4u{o[0], o[1], o[2], o[3]};
    *reinterpret_cast<v4u *>(&Bs[r * BSTR + 2 * kq + 16]) = v4u{o[4], o[5], o[6], o[7]};
  }
}

__device__ __forceinline__ float fold_f32(float x, float mf, float inv) {
  return fmaf(-rintf(x * inv), mf, x);
}

template <int CHECK>
__global__ __launch_bounds__(T256) void oz_bgemm_kernel(const RGemm *__restrict__ tab, int m,
                                                        int nmod,
                                                        const OzakiConsts *__restrict__ kc,
                                                        TileMap map, int m0, int d) {
  __shared__ __attribute__((aligned(16))) int8_t As[2][IBM * BSTR];
  __shared__ __attribute__((aligned(16))) int8_t Bs[2][IBN * BSTR];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int l16 = lane & 15, lk = lane >> 4;
  const int G = gridDim.x;
  const int nk = (m + IBK - 1) / IBK;
  const int fold_every = BFOLD / IBK;

  for (int round = 0; round * G < map.total; round++) {
    const int tile = round * G + xcd_tile(blockIdx.x, G);
    if (tile >= map.total) break;
    int prob, ti, tj;
    tile_coords(map, tile, prob, ti, tj);
    const int leaf = prob / nmod, l = prob - leaf * nmod;
    const int8_t *__restrict__ Ap = tab[leaf].a + l * tab[leaf].sa;  // row-major
    const int8_t *__restrict__ Bp = tab[leaf].b + l * tab[leaf].sb;  // column-major
    int8_t *__restrict__ Cp = tab[leaf].c + l * tab[leaf].sc;
    const int mi = kc->m[l];
    const float mf = (float)mi, invf = 1.0f / mf;
    const int row0 = ti * IBM, col0 = tj * IBN;

    v4f acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
      for (int c = 0; c < 4; c++) acc[r][c] = v4f{0.f, 0.f, 0.f, 0.f};

    v4u ga[2], gb[2];
    const long long mm0 = (long long)m0 * m0;
    auto a_at = [&](int k0) {
      return Ap + quad_leaf(row0 / m0, k0 / m0, d) * mm0 + (long long)(row0 % m0) * m0 + (k0 % m0);
    };
    auto b_at = [&](int k0) {
      return Bp + quad_leaf(k0 / m0, col0 / m0, d) * mm0 + (k0 % m0) + (long long)(col0 % m0) * m0;
    };
    gload_i8<CHECK>(a_at(0), b_at(0), m0, m, row0, col0, 0, tid, ga, gb);
    lstore_bf(As[0], Bs[0], tid, ga, gb);
    __syncthreads();
    for (int kt = 0; kt < nk; kt++) {
      const int buf = kt & 1;
      if (kt + 1 < nk)
        gload_i8<CHECK>(a_at((kt + 1) * IBK), b_at((kt + 1) * IBK), m0, m, row0, col0,
                        (kt + 1) * IBK, tid, ga, gb);
#pragma unroll
      for (int kh = 0; kh < 2; kh++) {
        // lane group lk supplies k = 32 kh + 8 lk .. + 7 for both operands
        v8bf a[4], b[4];
#pragma unroll
        for (int r = 0; r < 4; r++)
          a[r] = *reinterpret_cast<const v8bf *>(
              &As[buf][(wm * 64 + r * 16 + l16) * BSTR + 2 * (32 * kh + 8 * lk)]);
#pragma unroll
        for (int c = 0; c < 4; c++)
          b[c] = *reinterpret_cast<const v8bf *>(
              &Bs[buf][(wn * 64 + c * 16 + l16) * BSTR + 2 * (32 * kh + 8 * lk)]);
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
          for (int c = 0; c < 4; c++)
            acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[r], b[c], acc[r][c], 0, 0, 0);
      }
      if ((kt + 1) % fold_every == 0) {
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
          for (int c = 0; c < 4; c++)
#pragma unroll
            for (int q = 0; q < 4; q++) acc[r][c][q] = fold_f32(acc[r][c][q], mf, invf);
      }
      if (kt + 1 < nk) lstore_bf(As[buf ^ 1], Bs[buf ^ 1], tid, ga, gb);
      __syncthreads();
    }

    // D layout: lane holds column (lane % 16), rows 4 * (lane / 16) + v.
    // |acc| < 2^24 exact; a fold brings it to |x| <= 1.5 m, then the exact
    // symmetric residue
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const int row = row0 + wm * 64 + r * 16 + 4 * lk;
        const int col = col0 + wn * 64 + c * 16 + l16;
        int8_t *dst = Cp + quad_leaf(row0 / m0, col0 / m0, d) * mm0 +
                      (long long)(col % m0) * m0 + (row % m0);
        int res[4];
#pragma unroll
        for (int q = 0; q < 4; q++)
          res[q] = mod_k((int)fold_f32(acc[r][c][q], mf, invf), kc->c22[l], mi);  // |.| <= 1.5 m
        if (CHECK == 0) {
          unsigned v = 0;
#pragma unroll
          for (int q = 0; q < 4; q++) v |= (unsigned)(res[q] & 0xff) << (8 * q);
          *reinterpret_cast<unsigned *>(dst) = v;
        } else if (CHECK >= 2) {
          // m % 16 == 0: the 4 rows of the word are all inside or all outside
          if (col < m && row < m) {
            unsigned v = 0;
#pragma unroll
            for (int q = 0; q < 4; q++) v |= (unsigned)(res[q] & 0xff) << (8 * q);
            *(gu32 *)dst = v;
          }
        } else if (col < m) {
#pragma unroll
          for (int q = 0; q < 4; q++)
            if (row + q < m) dst[q] = (int8_t)res[q];
        }
      }
  }
}

// ------------------------------------------------------------------ BF16 GEMM, 256 x 256 tiles
// The BF16 variant with the large kernel's tiling: 256 x 256 tiles, 8 waves
// of 128 x 64 (2 per SIMD), K-steps of 64 on v_mfma_f32_16x16x32_bf16 with
// FP32 accumulation folded mod m every BFOLD k as above.  A K-step's INT8
// slabs (256 x 64 bytes of A and of B) come into registers (4 x 16 bytes per
// lane), are widened once to BF16 (cvt_f32 of a sign-extended byte, then
// cvt_pk_bf16: 1.5 VALU per value) and stored to the other LDS stage after
// the step's MFMAs.  A BF16 LDS row is 64 values = 128 bytes, 8 chunks of 16
// with the INT8 kernel's swizzle, so the fragment reads are the INT8
// kernel's.  C leaves through LDS as in the INT8 kernel.  Same results, bit
// for bit, as the INT8 kernels.  Needs m0 % 256 == 0.
constexpr int KBK = 64;                   // K per step (BF16 values)
constexpr int KSTAGE = JBM * KBK * 2;     // 32 KB per operand per stage

__device__ __forceinline__ void widen16_bf16(v4u w, v4u &lo, v4u &hi) {
  typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
  unsigned o[8];
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const float f0 = (float)(int)(signed char)(w[q] >> (16 * h));
      const float f1 = (float)(int)(signed char)(w[q] >> (16 * h + 8));
      o[2 * q + h] = __builtin_bit_cast(unsigned, v2bf{(__bf16)f0, (__bf16)f1});
    }
  lo = v4u{o[0], o[1], o[2], o[3]};
  hi = v4u{o[4], o[5], o[6], o[7]};
}

__global__ __launch_bounds__(JT) __attribute__((amdgpu_waves_per_eu(2, 2))) void
oz_bgemm_big_kernel(const RGemm *__restrict__ tab, int m, int nmod,
                    const OzakiConsts *__restrict__ kc, TileMap map, int m0, int d) {
  __shared__ __attribute__((aligned(1024))) int8_t A0[KSTAGE];
  __shared__ __attribute__((aligned(1024))) int8_t A1[KSTAGE];
  __shared__ __attribute__((aligned(1024))) int8_t B0[KSTAGE];
  __shared__ __attribute__((aligned(1024))) int8_t B1[KSTAGE];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;  // 2 x 4 waves: rows wm * 128, columns wn * 64
  const int l16 = lane & 15, lk = lane >> 4;
  const int G = gridDim.x;
  const int nk = m / KBK;
  const int fold_every = BFOLD / KBK;
  // fragment reads: row l16, chunk lk (k = 8 lk .. + 7) and lk + 4 (k + 32)
  const int s16 = jswz(l16);
  const int rd0 = l16 * 128 + ((lk ^ s16) << 4);
  const int rd1 = l16 * 128 + (((lk + 4) ^ s16) << 4);
  const int aoff = wm * 128 * 128, boff = wn * 64 * 128;
  const long long mm0 = (long long)m0 * m0;
  // staging: lane chunk idx = tid + 512 it: row idx >> 2, 16 INT8 k at 16 (idx & 3)
  int srow[2], skq[2], sdst[2][2];
#pragma unroll
  for (int it = 0; it < 2; it++) {
    const int idx = tid + JT * it;
    srow[it] = idx >> 2;
    skq[it] = idx & 3;
    const int sw = jswz(srow[it]);
    sdst[it][0] = srow[it] * 128 + (((2 * skq[it]) ^ sw) << 4);
    sdst[it][1] = srow[it] * 128 + (((2 * skq[it] + 1) ^ sw) << 4);
  }
  auto a_off = [&](int row, int k) -> long long {
    return d == 0 ? (long long)row * m0 + k
                  : quad_leaf(row / m0, k / m0, d) * mm0 + (long long)(row % m0) * m0 + (k % m0);
  };
  auto b_off = [&](int col, int k) -> long long {
    return d == 0 ? (long long)col * m0 + k
                  : quad_leaf(k / m0, col / m0, d) * mm0 + (long long)(col % m0) * m0 + (k % m0);
  };
  auto gload = [&](const BigTile &c, int k, v4u (&ga)[2], v4u (&gb)[2]) {
#pragma unroll
    for (int it = 0; it < 2; it++) {
      ga[it] = *(const gv4u *)(c.Ap + a_off(c.row0 + srow[it], k + 16 * skq[it]));
      gb[it] = *(const gv4u *)(c.Bp + b_off(c.col0 + srow[it], k + 16 * skq[it]));
    }
  };
  auto lstore = [&](const v4u (&ga)[2], const v4u (&gb)[2], int8_t *As, int8_t *Bs) {
#pragma unroll
    for (int it = 0; it < 2; it++) {
      v4u lo, hi;
      widen16_bf16(ga[it], lo, hi);
      *reinterpret_cast<v4u *>(As + sdst[it][0]) = lo;
      *reinterpret_cast<v4u *>(As + sdst[it][1]) = hi;
      widen16_bf16(gb[it], lo, hi);
      *reinterpret_cast<v4u *>(Bs + sdst[it][0]) = lo;
      *reinterpret_cast<v4u *>(Bs + sdst[it][1]) = hi;
    }
  };

  int tile = xcd_tile(blockIdx.x, G);
  if (tile >= map.total) return;
  BigTile t = big_tile(tab, map, tile, nmod, kc, (long long)m * m);
  v4u ga[2], gb[2];
  gload(t, 0, ga, gb);
  lstore(ga, gb, A0, B0);
  __syncthreads();

  for (int round = 0;; round++) {
    v4f acc[8][4];
#pragma unroll
    for (int r = 0; r < 8; r++)
#pragma unroll
      for (int c = 0; c < 4; c++) acc[r][c] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int kt = 0; kt < nk; kt++) {
      const bool odd = kt & 1;
      if (kt + 1 < nk) gload(t, (kt + 1) * KBK, ga, gb);
      const int8_t *Ab = (odd ? A1 : A0) + aoff;
      const int8_t *Bb = (odd ? B1 : B0) + boff;
      v8bf a0[8], b0[4], a1[8], b1[4];
#pragma unroll
      for (int c = 0; c < 4; c++) b0[c] = *reinterpret_cast<const v8bf *>(Bb + rd0 + c * 16 * 128);
#pragma unroll
      for (int r = 0; r < 8; r++) a0[r] = *reinterpret_cast<const v8bf *>(Ab + rd0 + r * 16 * 128);
#pragma unroll
      for (int c = 0; c < 4; c++) b1[c] = *reinterpret_cast<const v8bf *>(Bb + rd1 + c * 16 * 128);
#pragma unroll
      for (int r = 0; r < 8; r++) a1[r] = *reinterpret_cast<const v8bf *>(Ab + rd1 + r * 16 * 128);
#pragma unroll
      for (int r = 0; r < 8; r++)
#pragma unroll
        for (int c = 0; c < 4; c++)
          acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0[r], b0[c], acc[r][c], 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 8; r++)
#pragma unroll
        for (int c = 0; c < 4; c++)
          acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1[r], b1[c], acc[r][c], 0, 0, 0);
      // as in the INT8 kernel: the first MFMAs' 6 fragment reads, then one
      // read per two MFMAs (1.3 % faster than all 24 reads up front)
      __builtin_amdgcn_sched_group_barrier(0x100, 6, 0);
#pragma unroll
      for (int i = 0; i < 18; i++) {
        __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      }
      __builtin_amdgcn_sched_group_barrier(0x008, 28, 0);
      if ((kt + 1) % fold_every == 0) {
#pragma unroll
        for (int r = 0; r < 8; r++)
#pragma unroll
          for (int c = 0; c < 4; c++)
#pragma unroll
            for (int q = 0; q < 4; q++) acc[r][c][q] = fold_f32(acc[r][c][q], t.mf, t.invf);
      }
      if (kt + 1 < nk) lstore(ga, gb, odd ? A0 : A1, odd ? B0 : B1);
      __syncthreads();
    }
    // after the K loop's last barrier both stages are free: the next tile's
    // first stage goes to registers now and to A0 / B0 after this tile's C is
    // staged in A1 / B1
    const int next = (round + 1) * G + xcd_tile(blockIdx.x, G);
    const bool has_next = next < map.total;
    BigTile tn = t;
    if (has_next) {
      tn = big_tile(tab, map, next, nmod, kc, (long long)m * m);
      gload(tn, 0, ga, gb);
    }
    // C through LDS (A1: columns 0..127, B1: 128..255, as in the INT8 kernel)
    const int c22 = t.c22;
    unsigned *Ch = reinterpret_cast<unsigned *>(wn < 2 ? A1 : B1);
#pragma unroll
    for (int r = 0; r < 8; r++)
#pragma unroll
      for (int c = 0; c < 4; c++) {
        int res[4];
#pragma unroll
        for (int q = 0; q < 4; q++)
          res[q] = mod_k((int)fold_f32(acc[r][c][q], t.mf, t.invf), c22, t.mi);  // |.| <= 1.5 m
        const int col = (wn & 1) * 64 + c * 16 + l16;
        const int w = wm * 32 + r * 4 + lk;
        Ch[col * 64 + (w ^ (4 * l16))] = pack4_lo(res[0], res[1], res[2], res[3]);
      }
    if (has_next) lstore(ga, gb, A0, B0);
    __syncthreads();
    // C segments as in the INT8 kernel
    const int j16 = tid & 15;
    const long long e0 = quad_leaf(t.row0 / m0, t.col0 / m0, d) * mm0 +
                         (long long)(t.col0 % m0 + (tid >> 4)) * m0 + (t.row0 % m0) + 16 * j16;
    int8_t *Cl = t.Cp + (e0 >> 7) * t.cseg + (e0 & 127);
    const long long pstep = (long long)(m0 >> 2) * t.cseg;
#pragma unroll
    for (int pass = 0; pass < 8; pass++) {
      const int ccl = (pass * 32 + (tid >> 4)) & 127;  // pass < 4: A1, else B1
      const unsigned *Cs = reinterpret_cast<const unsigned *>(pass < 4 ? A1 : B1);
      const v4i v = *reinterpret_cast<const v4i *>(Cs + ccl * 64 + ((4 * j16) ^ (4 * (ccl & 15))));
      *(gv4i *)(Cl + pass * pstep) = v;
    }
    if (!has_next) break;
    t = tn;
    __syncthreads();  // the C reads of A1 / B1 end before the next tile's K-step 0 fills them
  }
}

// ------------------------------------------------------------------ reconstruct
template <int NP, bool VEC>
__global__ __launch_bounds__(T256) void oz_recon_kernel(const OzLeaf *__restrict__ tab, int m,
                                                        int nmod,
                                                        const OzakiConsts *__restrict__ kc) {
  const OzLeaf &L = tab[blockIdx.y];
  const long long plane = (long long)m * m;
  const int per = VEC ? 4 : 1;
  const long long units = plane / per;
  for (long long u = (long long)blockIdx.x * T256 + threadIdx.x; u < units;
       u += (long long)gridDim.x * T256) {
    const long long e0 = u * per;
    const int j = (int)(e0 / m), i0 = (int)(e0 - (long long)j * m);
    const int sb = L.eb[j];
    if (VEC) {
      double S[4][NP];
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int p = 0; p < NP; p++) S[q][p] = 0.0;
      for (int l = 0; l < nmod; l++) {
        const unsigned w = *reinterpret_cast<const unsigned *>(L.rc + l * plane + e0);
#pragma unroll
        for (int q = 0; q < 4; q++) crt_acc<NP>(S[q], (int)(int8_t)((w >> (8 * q)) & 0xff), l, kc);
      }
      double out[4];
#pragma unroll
      for (int q = 0; q < 4; q++) out[q] = ldexp(crt_finish<NP>(S[q], kc), -(L.ea[i0 + q] + sb));
      reinterpret_cast<double2 *>(L.C + e0)[0] = make_double2(out[0], out[1]);
      reinterpret_cast<double2 *>(L.C + e0)[1] = make_double2(out[2], out[3]);
    } else {
      double S[NP];
#pragma unroll
      for (int p = 0; p < NP; p++) S[p] = 0.0;
      for (int l = 0; l < nmod; l++) crt_acc<NP>(S, (int)L.rc[l * plane + e0], l, kc);
      L.C[e0] = ldexp(crt_finish<NP>(S, kc), -(L.ea[i0] + sb));
    }
  }
}

template <int NP>
hipError_t launch_recon(const OzLeaf *tab, int count, int m, int nmod, const OzakiConsts *kc,
                        hipStream_t s) {
  const bool vec = (m % 4 == 0);
  const long long units = (long long)m * m / (vec ? 4 : 1);
  long long blocks = (units + T256 - 1) / T256;
  long long cap = (long long)device_cu_count() * 8 / count;
  if (cap < 1) cap = 1;
  if (blocks > cap) blocks = cap;
  dim3 grid((unsigned)blocks, (unsigned)count);
  if (vec)
    hipLaunchKernelGGL((oz_recon_kernel<NP, true>), grid, dim3(T256), 0, s, tab, m, nmod, kc);
  else
    hipLaunchKernelGGL((oz_recon_kernel<NP, false>), grid, dim3(T256), 0, s, tab, m, nmod, kc);
  return hipGetLastError();
}


}  // namespace

hipError_t launch_residue_gemm_bf16(const RGemm *tab, int count, int m, int levels, int nmod,
                                    const OzakiConsts *kc, hipStream_t s) {
  if (count <= 0 || m <= 0) return hipSuccess;
  if (nmod < 2 || nmod > OZ_MAX_MODULI) return hipErrorInvalidValue;
  const int m0 = m >> levels;
  if (levels < 0 || (m0 << levels) != m) return hipErrorInvalidValue;
  bool ok;
  if (m0 % JBM == 0) {
    // large tiles: one block (8 waves, 128 KB of LDS) per CU
    TileMap map = make_tilemap(m, JBM, (long long)count * nmod, ok);
    if (!ok) return hipErrorInvalidValue;
    const int G = persistent_grid(1, device_cu_count(), map.total);
    hipLaunchKernelGGL(oz_bgemm_big_kernel, dim3(G), dim3(JT), 0, s, tab, m, nmod, kc, map, m0,
                       levels);
    return hipGetLastError();
  }
  TileMap map = make_tilemap(m, IBM, (long long)count * nmod, ok);
  if (!ok) return hipErrorInvalidValue;
  const int G = persistent_grid(2, device_cu_count(), map.total);
  const bool fast = (m0 % IBM == 0);
  if (!fast && levels > 0) return hipErrorInvalidValue;
  if (fast)
    hipLaunchKernelGGL((oz_bgemm_kernel<0>), dim3(G), dim3(T256), 0, s, tab, m, nmod, kc, map,
                       m0, levels);
  else if (m % IBK == 0)
    hipLaunchKernelGGL((oz_bgemm_kernel<2>), dim3(G), dim3(T256), 0, s, tab, m, nmod, kc, map,
                       m0, levels);
  else if (m % 16 == 0)
    hipLaunchKernelGGL((oz_bgemm_kernel<3>), dim3(G), dim3(T256), 0, s, tab, m, nmod, kc, map,
                       m0, levels);
  else
    hipLaunchKernelGGL((oz_bgemm_kernel<1>), dim3(G), dim3(T256), 0, s, tab, m, nmod, kc, map,
                       m0, levels);
  return hipGetLastError();
}

// the large-tile kernel with partial edges for plain leaves (levels = 0) of
// m % 16 == 0, m <= 16384, when the padding to 256 x 256 tiles adds at most
// half the area (m >= 512 at worst): at 0.4 of peak it beats the 128 x 128
// kernel there
bool big_edge_gemm(int m, int levels, int tile) {
  if (tile != 0 || levels != 0 || m % 16 != 0 || m % JBM == 0 || m > 16384 || m < 512)
    return false;
  const long long padded = (long long)((m + JBM - 1) / JBM) * JBM;
  return 2 * padded * padded <= 3LL * m * m;
}

// whole 256 x 256 tiles: one block (8 waves, 128 KB of LDS) per CU; buffer
// offsets are 32-bit, so an operand quadtree stays below 2 GiB
static bool big_int8_gemm(int m, int levels, int tile) {
  const int m0 = m >> levels;
  return tile == 0 && m0 % JBM == 0 && (long long)m * m < (1LL << 31);
}

bool residue_gemm_segmented(int m, int levels, int tile, bool bf16) {
  if (m <= 0 || levels < 0 || ((m >> levels) << levels) != m) return false;
  return bf16 ? (m >> levels) % JBM == 0 : big_int8_gemm(m, levels, tile);
}

hipError_t launch_residue_gemm(const RGemm *tab, int count, int m, int levels, int nmod,
                               const OzakiConsts *kc, hipStream_t s, int tile) {
  if (count <= 0 || m <= 0) return hipSuccess;
  if (nmod < 2 || nmod > OZ_MAX_MODULI) return hipErrorInvalidValue;
  const int m0 = m >> levels;
  if (levels < 0 || (m0 << levels) != m) return hipErrorInvalidValue;
  bool ok;
  if (big_int8_gemm(m, levels, tile)) {
    TileMap map = make_tilemap(m, JBM, (long long)count * nmod, ok);
    if (!ok) return hipErrorInvalidValue;
    const int G = persistent_grid(1, device_cu_count(), map.total);
    if (m <= 16384)
      hipLaunchKernelGGL((oz_igemm_big_kernel<true, false>), dim3(G), dim3(JT), 0, s, tab, m,
                         nmod, kc, map, m0, levels);
    else
      hipLaunchKernelGGL((oz_igemm_big_kernel<false, false>), dim3(G), dim3(JT), 0, s, tab, m,
                         nmod, kc, map, m0, levels);
    return hipGetLastError();
  }
  if (big_edge_gemm(m, levels, tile)) {
    // plain operands of any m % 16 == 0 on the large tiles, partial edges
    TileMap map = make_tilemap(m, JBM, (long long)count * nmod, ok);
    if (!ok) return hipErrorInvalidValue;
    const int G = persistent_grid(1, device_cu_count(), map.total);
    hipLaunchKernelGGL((oz_igemm_big_kernel<true, true>), dim3(G), dim3(JT), 0, s, tab, m, nmod,
                       kc, map, m0, levels);
    return hipGetLastError();
  }
  TileMap map = make_tilemap(m, IBM, (long long)count * nmod, ok);
  if (!ok) return hipErrorInvalidValue;
  const int G = persistent_grid(2, device_cu_count(), map.total);
  // the fast path reads 16-byte rows: leaves of a multiple of 128; with m %
  // 64 == 0 the K slices are whole and only edge rows / columns are checked;
  // with m % 16 == 0 whole 16-byte K chunks are checked too; else every byte
  const bool fast = (m0 % IBM == 0);
  if (!fast && levels > 0) return hipErrorInvalidValue;
  if (fast)
    hipLaunchKernelGGL((oz_igemm_kernel<0>), dim3(G), dim3(T256), 0, s, tab, m, nmod, kc, map,
                       m0, levels);
  else if (m % IBK == 0)
    hipLaunchKernelGGL((oz_igemm_kernel<2>), dim3(G), dim3(T256), 0, s, tab, m, nmod, kc, map,
                       m0, levels);
  else if (m % 16 == 0)
    hipLaunchKernelGGL((oz_igemm_kernel<3>), dim3(G), dim3(T256), 0, s, tab, m, nmod, kc, map,
                       m0, levels);
  else
    hipLaunchKernelGGL((oz_igemm_kernel<1>), dim3(G), dim3(T256), 0, s, tab, m, nmod, kc, map,
                       m0, levels);
  return hipGetLastError();
}

int residue_gemm_check(int m, int levels) {
  const int m0 = m >> levels;
  if (m0 % IBM == 0) return 0;
  if (levels > 0) return -1;  // quadtree operands need whole tiles
  return m % IBK == 0 ? 2 : (m % 16 == 0 ? 3 : 1);
}

hipError_t launch_residue_gemm_multi(const char *base, const VSeg *segs, int nseg, int total,
                                     int nmod, const OzakiConsts *kc, hipStream_t s) {
  if (nseg <= 0 || total <= 0) return hipSuccess;
  if (nmod < 2 || nmod > OZ_MAX_MODULI) return hipErrorInvalidValue;
  const int G = persistent_grid(2, device_cu_count(), total);
  hipLaunchKernelGGL(oz_igemm_multi_kernel, dim3(G), dim3(T256), 0, s, base, segs, nseg, total,
                     nmod, kc);
  return hipGetLastError();
}

hipError_t launch_ozaki_step(int step, const OzLeaf *tab, int count, int m, int nmod,
                             const OzakiConsts *kc, hipStream_t s) {
  if (count <= 0 || m <= 0) return hipSuccess;
  if (count > 65535 || nmod < 2 || nmod > OZ_MAX_MODULI) return hipErrorInvalidValue;
  switch (step) {
    case 0: {
      dim3 g1((m + T256 - 1) / T256, OZ_CHUNKS, count);
      hipLaunchKernelGGL(oz_rowsq_kernel, g1, dim3(T256), 0, s, tab, m);
      dim3 g2((m + 3) / 4, count);
      hipLaunchKernelGGL(oz_colsq_kernel, g2, dim3(T256), 0, s, tab, m);
      dim3 g3((m + T256 - 1) / T256, count);
      hipLaunchKernelGGL(oz_finalize_kernel, g3, dim3(T256), 0, s, tab, m, kc);
      return hipGetLastError();
    }
    case 1: {
      const int tiles = (m + 63) / 64;
      dim3 ga((unsigned)(tiles * tiles), count);
      const bool fa = (m % 64 == 0);
      if (fa)
        hipLaunchKernelGGL((oz_split_a_kernel<false>), ga, dim3(T256), 0, s, tab, m, nmod, kc);
      else
        hipLaunchKernelGGL((oz_split_a_kernel<true>), ga, dim3(T256), 0, s, tab, m, nmod, kc);
      const long long u16 = ((long long)m * m + 15) / 16;
      long long blocks = (u16 + T256 - 1) / T256;
      long long cap = (long long)device_cu_count() * 8 / count;
      if (cap < 1) cap = 1;
      if (blocks > cap) blocks = cap;
      dim3 gb((unsigned)blocks, count);
      if (m % 16 == 0)
        hipLaunchKernelGGL((oz_split_b_kernel<false>), gb, dim3(T256), 0, s, tab, m, nmod, kc);
      else
        hipLaunchKernelGGL((oz_split_b_kernel<true>), gb, dim3(T256), 0, s, tab, m, nmod, kc);
      return hipGetLastError();
    }
    case 3: {
      switch (ozaki_npieces(nmod)) {
        case 1:
        case 2: return launch_recon<2>(tab, count, m, nmod, kc, s);
        case 3: return launch_recon<3>(tab, count, m, nmod, kc, s);
        default: return launch_recon<4>(tab, count, m, nmod, kc, s);
      }
    }
  }
  return hipErrorInvalidValue;
}

}  // namespace caps
