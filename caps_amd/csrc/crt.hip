// Kernels of the residue-domain Strassen (crt.h): scale, split, residue
// Winograd steps and the final CRT reconstruction.  All HBM-bound.
#include "crt.h"
#include "ew_ops.h"
#include "kernels.h"
#include "ozaki_dev.h"

namespace caps {

namespace {

constexpr int T = 256;

// ------------------------------------------------------------------ scale
// partial (e, s) of row i of A over columns [lj0, lj1) of leaf column block bj.
// PLAIN: column-major leaves (bs == 1 or bs == m), 8 columns per step.
template <bool PLAIN>
__device__ __forceinline__ void rowpart_body(const CrtMats *__restrict__ M, int bx, int by) {
  const int n = M->n, m = M->m, nrec = M->nrec, bs = M->bs, ch = M->chunks;
  const int i = bx * T + threadIdx.x;
  if (i >= n) return;
  const int c = by;  // bj * chunks + k
  const int bj = c / ch, k = c - bj * ch;
  const int per = (m + ch - 1) / ch;
  const int lj0 = k * per, lj1 = min(m, lj0 + per);
  int e = NOSCALE;
  double s = 0.0;
  if (PLAIN) {
    const double *row = M->A + caps_elem_offset(n, nrec, bs, i, bj * m);  // (i, bj m)
    int lj = lj0;
    for (; lj + 8 <= lj1; lj += 8) {
      double v[8];
#pragma unroll
      for (int t = 0; t < 8; t++) v[t] = row[(long long)(lj + t) * m];
      int e8;
      double s8;
      sq_block8(v, e8, s8);
      sq_merge(e, s, e8, s8);
    }
    for (; lj < lj1; lj++) sq_acc(row[(long long)lj * m], e, s);
  } else {
    for (int lj = lj0; lj < lj1; lj++)
      sq_acc(M->A[caps_elem_offset(n, nrec, bs, i, bj * m + lj)], e, s);
  }
  reinterpret_cast<double2 *>(M->pa)[(long long)c * n + i] = make_double2(s, (double)e);
}

template <bool PLAIN>
__global__ __launch_bounds__(T) void crt_rowpart_kernel(const CrtMats *__restrict__ M) {
  rowpart_body<PLAIN>(M, blockIdx.x, blockIdx.y);
}

// partial of column j of B over leaf row block bi: one wave, fixed butterfly.
// PLAIN: column-major leaves, each lane takes 8 rows (stride 64) per step.
template <bool PLAIN>
__device__ __forceinline__ void colpart_body(const CrtMats *__restrict__ M, int bx) {
  const int n = M->n, m = M->m, nrec = M->nrec, bs = M->bs;
  const long long w = (long long)bx * (T / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int nblk = 1 << nrec;
  if (w >= (long long)n * nblk) return;
  const int j = (int)(w % n), bi = (int)(w / n);
  int e = NOSCALE;
  double s = 0.0;
  if (PLAIN) {
    const double *col = M->B + caps_elem_offset(n, nrec, bs, bi * m, j);  // (bi m, j)
    int li = lane;
    for (; li + 7 * 64 < m; li += 8 * 64) {
      double v[8];
#pragma unroll
      for (int t = 0; t < 8; t++) v[t] = col[li + 64 * t];
      int e8;
      double s8;
      sq_block8(v, e8, s8);
      sq_merge(e, s, e8, s8);
    }
    for (; li < m; li += 64) sq_acc(col[li], e, s);
  } else {
    for (int li = lane; li < m; li += 64)
      sq_acc(M->B[caps_elem_offset(n, nrec, bs, bi * m + li, j)], e, s);
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    int e2 = __shfl_down(e, off, 64);
    double s2 = __shfl_down(s, off, 64);
    if (lane < off) sq_merge(e, s, e2, s2);
  }
  if (lane == 0) reinterpret_cast<double2 *>(M->pb)[(long long)bi * n + j] = make_double2(s, (double)e);
}

template <bool PLAIN>
__global__ __launch_bounds__(T) void crt_colpart_kernel(const CrtMats *__restrict__ M) {
  colpart_body<PLAIN>(M, blockIdx.x);
}

__device__ __forceinline__ void finalize_body(const CrtMats *__restrict__ M,
                                              const OzakiConsts *__restrict__ kc, int bx) {
  const int n = M->n;
  const int i = bx * T + threadIdx.x;
  if (i >= n) return;
  const int nblk = 1 << M->nrec;
  const double2 *pa = reinterpret_cast<const double2 *>(M->pa);
  const double2 *pb = reinterpret_cast<const double2 *>(M->pb);
  int e = NOSCALE;
  double s = 0.0;
  for (int c = 0; c < nblk * M->chunks; c++) {
    double2 v = pa[(long long)c * n + i];
    sq_merge(e, s, (int)v.y, v.x);
  }
  M->sa[i] = scale_exponent(e, s, kc->L);
  e = NOSCALE;
  s = 0.0;
  for (int b = 0; b < nblk; b++) {
    double2 v = pb[(long long)b * n + i];
    sq_merge(e, s, (int)v.y, v.x);
  }
  M->sb[i] = scale_exponent(e, s, kc->L);
}

__global__ __launch_bounds__(T) void crt_finalize_kernel(const CrtMats *__restrict__ M,
                                                         const OzakiConsts *__restrict__ kc) {
  finalize_body(M, kc, blockIdx.x);
}

// ------------------------------------------------------------------ split
// A leaf L -> row-major residue leaves, transposed through LDS.
// FAST (m % 64 == 0 and the leaf column-major, bs == 1 or bs == m): tiles of
// 32 rows x 128 k; the column-major side moves as 256-byte column pieces, the
// row-major side as 128-byte row pieces (8 lanes x 16 residues), K chunks past
// m (m % 128 == 64) skipped.  Column c of the tile in T2[c], row r at slot r ^
// tswz(c); the transposed read has lanes on 8 rows and 8 column groups, the
// swizzle puts 4 of the groups 8 slots apart (the other 4 pairs share a
// slot: a 64-lane 8-byte read spans at least two 256-byte passes anyway).
// Otherwise 64 x 64 tiles with checked, element-wise accesses.
__device__ __forceinline__ int tswz(int c) { return ((c >> 4) & 3) << 3; }

template <bool FAST>
__device__ __forceinline__ void split_a_body(const CrtMats *__restrict__ M,
                                             const OzakiConsts *__restrict__ kc, int bx, int by,
                                             double (&Ts)[64][64]) {
  const int m = M->m, nmod = M->nmod, bs = M->bs;
  const int L = by;
  int bi, bj;
  leaf_coords(L, M->nrec, bi, bj);
  const long long lbase = (long long)L * m * m;
  const long long plane = (long long)M->n * M->n;
  if (FAST) {
    double(*T2)[32] = reinterpret_cast<double(*)[32]>(&Ts[0][0]);  // [128 k][32 rows]
    const int tk = (m + 127) / 128;
    const int r0 = (bx / tk) * 32, k0 = (bx % tk) * 128;
    for (int idx = threadIdx.x; idx < 32 * 128; idx += T) {
      const int r = idx & 31, c = idx >> 5;
      const int lj = k0 + c;
      T2[c][r ^ tswz(c)] = lj < m ? M->A[lbase + r0 + r + (long long)lj * m] : 0.0;
    }
    __syncthreads();
    const int r = threadIdx.x >> 3, kq = (threadIdx.x & 7) * 16;
    if (k0 + kq >= m) return;  // m % 16 == 0: the chunk is wholly past the leaf
    const int li = r0 + r;
    const int sa = M->sa[bi * m + li];
    double x[16];
#pragma unroll
    for (int c = 0; c < 16; c++) x[c] = trunc(ldexp(T2[kq + c][r ^ tswz(kq)], sa));
    int8_t *dst = M->ra + lbase + (long long)li * m + k0 + kq;
    for (int l = 0; l < nmod; l++) {
      const int mi = kc->m[l];
      const double inv = kc->inv_m[l];
      const int c22 = kc->c22[l];
      unsigned w[4];
#pragma unroll
      for (int q = 0; q < 4; q++)
        w[q] = pack4_lo(sym_residue(x[4 * q], mi, inv, c22),
                        sym_residue(x[4 * q + 1], mi, inv, c22),
                        sym_residue(x[4 * q + 2], mi, inv, c22),
                        sym_residue(x[4 * q + 3], mi, inv, c22));
      *reinterpret_cast<uint4 *>(dst + l * plane) = make_uint4(w[0], w[1], w[2], w[3]);
    }
    return;
  }
  const int tiles = (m + 63) / 64;
  const int r0 = (bx % tiles) * 64, k0 = (bx / tiles) * 64;
  for (int idx = threadIdx.x; idx < 64 * 64; idx += T) {
    const int r = idx & 63, c = idx >> 6;
    const int li = r0 + r, lj = k0 + c;
    double v = 0.0;
    if (li < m && lj < m) v = M->A[lbase + leaf_elem_offset(m, bs, li, lj)];
    Ts[c][r ^ tswz(c)] = v;
  }
  __syncthreads();
  const int r = threadIdx.x >> 2, kq = (threadIdx.x & 3) * 16;
  const int li = r0 + r;
  if (li >= m) return;
  const int sa = M->sa[bi * m + li];
  double x[16];
#pragma unroll
  for (int c = 0; c < 16; c++) x[c] = trunc(ldexp(Ts[kq + c][r ^ tswz(kq)], sa));
  int8_t *dst = M->ra + lbase + (long long)li * m + k0 + kq;
  for (int l = 0; l < nmod; l++) {
    const int mi = kc->m[l];
    const double inv = kc->inv_m[l];
    const int c22 = kc->c22[l];
#pragma unroll
    for (int c = 0; c < 16; c++)
      if (k0 + kq + c < m) dst[l * plane + c] = (int8_t)sym_residue(x[c], mi, inv, c22);
  }
}

template <bool FAST>
__global__ __launch_bounds__(T) void crt_split_a_kernel(const CrtMats *__restrict__ M,
                                                        const OzakiConsts *__restrict__ kc) {
  __shared__ double Ts[64][64];  // [col][swizzled row]
  split_a_body<FAST>(M, kc, blockIdx.x, blockIdx.y, Ts);
}

// B leaf L -> column-major residue leaves; thread per 16 consecutive rows.
// FAST: m % 16 == 0 and the leaf is column-major.
template <bool FAST>
__device__ __forceinline__ void split_b_body(const CrtMats *__restrict__ M,
                                             const OzakiConsts *__restrict__ kc, int bx, int by,
                                             int gx) {
  const int m = M->m, nmod = M->nmod, bs = M->bs;
  const int L = by;
  int bi, bj;
  leaf_coords(L, M->nrec, bi, bj);
  const long long lbase = (long long)L * m * m;
  const long long mm = (long long)m * m;
  const long long plane = (long long)M->n * M->n;
  const long long units = FAST ? mm / 16 : mm;
  for (long long u = (long long)bx * T + threadIdx.x; u < units; u += (long long)gx * T) {
    if (FAST) {
      const long long e0 = u * 16;
      const int lj = (int)(e0 / m);
      const int s = M->sb[bj * m + lj];
      double x[16];
#pragma unroll
      for (int c = 0; c < 16; c += 2) {
        double2 v = *reinterpret_cast<const double2 *>(M->B + lbase + e0 + c);
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
        *reinterpret_cast<uint4 *>(M->rb + l * plane + lbase + e0) =
            make_uint4(w[0], w[1], w[2], w[3]);
      }
    } else {
      const int li = (int)(u % m), lj = (int)(u / m);
      const double x = trunc(ldexp(M->B[lbase + leaf_elem_offset(m, bs, li, lj)],
                                   M->sb[bj * m + lj]));
      for (int l = 0; l < nmod; l++)
        M->rb[l * plane + lbase + u] = (int8_t)sym_residue(x, kc->m[l], kc->inv_m[l], kc->c22[l]);
    }
  }
}

template <bool FAST>
__global__ __launch_bounds__(T) void crt_split_b_kernel(const CrtMats *__restrict__ M,
                                                        const OzakiConsts *__restrict__ kc) {
  split_b_body<FAST>(M, kc, blockIdx.x, blockIdx.y, gridDim.x);
}

// ------------------------------------------------------------------ reconstruct
// C' residues (column-major leaves) -> C (CAPS layout, bs blocks)
template <int NP, bool FAST>
__device__ __forceinline__ void recon_body(const CrtMats *__restrict__ M,
                                           const OzakiConsts *__restrict__ kc, int bx, int by,
                                           int gx) {
  const int m = M->m, nmod = M->nmod, bs = M->bs;
  const int L = by;
  int bi, bj;
  leaf_coords(L, M->nrec, bi, bj);
  const long long lbase = (long long)L * m * m;
  const long long mm = (long long)m * m;
  const long long plane = (long long)M->n * M->n;
  const int per = FAST ? 4 : 1;
  const long long units = mm / per;
  for (long long u = (long long)bx * T + threadIdx.x; u < units; u += (long long)gx * T) {
    const long long e0 = u * per;
    const int li = (int)(e0 % m), lj = (int)(e0 / m);
    const int sb = M->sb[bj * m + lj];
    if (FAST) {
      double S[4][NP];
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int p = 0; p < NP; p++) S[q][p] = 0.0;
      for (int l = 0; l < nmod; l++) {
        const unsigned w = *reinterpret_cast<const unsigned *>(M->rc + l * plane + lbase + e0);
#pragma unroll
        for (int q = 0; q < 4; q++) crt_acc<NP>(S[q], (int)(int8_t)((w >> (8 * q)) & 0xff), l, kc);
      }
      double out[4];
#pragma unroll
      for (int q = 0; q < 4; q++)
        out[q] = ldexp(crt_finish<NP>(S[q], kc), -(M->sa[bi * m + li + q] + sb));
      double2 *dst = reinterpret_cast<double2 *>(M->C + lbase + e0);
      dst[0] = make_double2(out[0], out[1]);
      dst[1] = make_double2(out[2], out[3]);
    } else {
      double S[NP];
#pragma unroll
      for (int p = 0; p < NP; p++) S[p] = 0.0;
      for (int l = 0; l < nmod; l++) crt_acc<NP>(S, (int)M->rc[l * plane + lbase + e0], l, kc);
      M->C[lbase + leaf_elem_offset(m, bs, li, lj)] =
          ldexp(crt_finish<NP>(S, kc), -(M->sa[bi * m + li] + sb));
    }
  }
}

template <int NP, bool FAST>
__global__ __launch_bounds__(T) void crt_recon_kernel(const CrtMats *__restrict__ M,
                                                      const OzakiConsts *__restrict__ kc) {
  recon_body<NP, FAST>(M, kc, blockIdx.x, blockIdx.y, gridDim.x);
}

// ------------------------------------------------------------------ segmented CRT steps
// One launch per kernel over the items of an uneven batch: segment args are
// CrtMats descriptions (crt_grid gives each item the grid it has alone).
template <int K, bool F, int NP>
__global__ __launch_bounds__(T) void crt_multi_kernel(const char *__restrict__ base,
                                                      const VSeg *__restrict__ segs, int nseg,
                                                      const OzakiConsts *__restrict__ kc) {
  __shared__ double Ts[K == CRT_SPLIT_A ? 64 : 1][K == CRT_SPLIT_A ? 64 : 1];
  const VSeg &S = segs[vseg_find(segs, nseg, blockIdx.x)];
  const int l = blockIdx.x - S.blk0;
  const int bx = l % S.gx, by = l / S.gx;
  const CrtMats *M = reinterpret_cast<const CrtMats *>(base + S.arg);
  if constexpr (K == CRT_ROWPART) rowpart_body<F>(M, bx, by);
  if constexpr (K == CRT_COLPART) colpart_body<F>(M, bx);
  if constexpr (K == CRT_FINALIZE) finalize_body(M, kc, bx);
  if constexpr (K == CRT_SPLIT_A) split_a_body<F>(M, kc, bx, by, Ts);
  if constexpr (K == CRT_SPLIT_B) split_b_body<F>(M, kc, bx, by, S.gx);
  if constexpr (K == CRT_RECON) recon_body<NP, F>(M, kc, bx, by, S.gx);
}

// ------------------------------------------------------------------ residue Winograd steps
template <int OP, bool VEC4>
__device__ __forceinline__ void rew_body(const RPtrs &P, long long n, int l, int bx, int gx,
                                         const OzakiConsts *__restrict__ kc) {
  using Op = Ew<OP>;
  const int mi = kc->m[l], c22 = kc->c22[l];  // sums of at most 4 residues: |y| <= 512
  const int8_t *in[Op::NIN];
  int8_t *out[Op::NOUT];
#pragma unroll
  for (int k = 0; k < Op::NIN; k++) in[k] = P.in[k] + l * P.sin[k];
#pragma unroll
  for (int k = 0; k < Op::NOUT; k++) out[k] = P.out[k] + l * P.sout[k];
  const long long stride = (long long)gx * T;
  if (VEC4) {
    const long long n4 = n >> 2;
    for (long long i = (long long)bx * T + threadIdx.x; i < n4; i += stride) {
      unsigned w[Op::NIN];
#pragma unroll
      for (int k = 0; k < Op::NIN; k++) w[k] = reinterpret_cast<const unsigned *>(in[k])[i];
      unsigned o[Op::NOUT];
#pragma unroll
      for (int k = 0; k < Op::NOUT; k++) o[k] = 0;
#pragma unroll
      for (int b = 0; b < 4; b++) {
        int x[Op::NIN], y[Op::NOUT];
#pragma unroll
        for (int k = 0; k < Op::NIN; k++) x[k] = (int)(int8_t)((w[k] >> (8 * b)) & 0xff);
        Op::apply(x, y);
#pragma unroll
        for (int k = 0; k < Op::NOUT; k++)
          o[k] |= (unsigned)(mod_k(y[k], c22, mi) & 0xff) << (8 * b);
      }
#pragma unroll
      for (int k = 0; k < Op::NOUT; k++) reinterpret_cast<unsigned *>(out[k])[i] = o[k];
    }
  } else {
    for (long long i = (long long)bx * T + threadIdx.x; i < n; i += stride) {
      int x[Op::NIN], y[Op::NOUT];
#pragma unroll
      for (int k = 0; k < Op::NIN; k++) x[k] = (int)in[k][i];
      Op::apply(x, y);
#pragma unroll
      for (int k = 0; k < Op::NOUT; k++) out[k][i] = (int8_t)mod_k(y[k], c22, mi);
    }
  }
}

// grid (blocks, entries, moduli)
template <int OP, bool VEC4>
__global__ __launch_bounds__(T) void rew_kernel(const RPtrs *__restrict__ tab, long long n,
                                                const OzakiConsts *__restrict__ kc) {
  rew_body<OP, VEC4>(tab[blockIdx.y], n, blockIdx.z, blockIdx.x, gridDim.x, kc);
}

// segments: gx blocks x (entries * moduli), aux = moduli
template <int OP, bool VEC4>
__global__ __launch_bounds__(T) void rew_multi_kernel(const char *__restrict__ base,
                                                      const VSeg *__restrict__ segs, int nseg,
                                                      const OzakiConsts *__restrict__ kc) {
  const VSeg &S = segs[vseg_find(segs, nseg, blockIdx.x)];
  const int l = blockIdx.x - S.blk0;
  const int by = l / S.gx;
  const RPtrs *tab = reinterpret_cast<const RPtrs *>(base + S.arg);
  rew_body<OP, VEC4>(tab[by / S.aux], S.n, by % S.aux, l % S.gx, S.gx, kc);
}

template <int OP>
hipError_t launch_rew_op(const RPtrs *tab, int count, long long n, int nmod, bool vec4,
                         const OzakiConsts *kc, hipStream_t s) {
  int gx, gy;
  rew_grid(count, n, nmod, vec4, gx, gy);
  dim3 grid((unsigned)gx, (unsigned)count, (unsigned)nmod);
  if (vec4)
    hipLaunchKernelGGL((rew_kernel<OP, true>), grid, dim3(T), 0, s, tab, n, kc);
  else
    hipLaunchKernelGGL((rew_kernel<OP, false>), grid, dim3(T), 0, s, tab, n, kc);
  return hipGetLastError();
}

template <int OP>
hipError_t launch_rew_op_multi(const char *base, const VSeg *segs, int nseg, int blocks,
                               bool vec4, const OzakiConsts *kc, hipStream_t s) {
  if (vec4)
    hipLaunchKernelGGL((rew_multi_kernel<OP, true>), dim3(blocks), dim3(T), 0, s, base, segs,
                       nseg, kc);
  else
    hipLaunchKernelGGL((rew_multi_kernel<OP, false>), dim3(blocks), dim3(T), 0, s, base, segs,
                       nseg, kc);
  return hipGetLastError();
}

template <int K, bool F>
hipError_t launch_crt_multi_k(int np, const char *base, const VSeg *segs, int nseg, int blocks,
                              const OzakiConsts *kc, hipStream_t s) {
  if (K != CRT_RECON || np <= 2)
    hipLaunchKernelGGL((crt_multi_kernel<K, F, 2>), dim3(blocks), dim3(T), 0, s, base, segs, nseg,
                       kc);
  else if (np == 3)
    hipLaunchKernelGGL((crt_multi_kernel<K, F, 3>), dim3(blocks), dim3(T), 0, s, base, segs, nseg,
                       kc);
  else
    hipLaunchKernelGGL((crt_multi_kernel<K, F, 4>), dim3(blocks), dim3(T), 0, s, base, segs, nseg,
                       kc);
  return hipGetLastError();
}

template <int K>
hipError_t launch_crt_multi_f(bool flag, int np, const char *base, const VSeg *segs, int nseg,
                              int blocks, const OzakiConsts *kc, hipStream_t s) {
  return flag ? launch_crt_multi_k<K, true>(np, base, segs, nseg, blocks, kc, s)
              : launch_crt_multi_k<K, false>(np, base, segs, nseg, blocks, kc, s);
}

unsigned capped_blocks(long long units, long long per_grid) {
  long long b = (units + T - 1) / T;
  long long cap = (long long)device_cu_count() * 8 / (per_grid > 0 ? per_grid : 1);
  if (cap < 1) cap = 1;
  if (b > cap) b = cap;
  return (unsigned)(b < 1 ? 1 : b);
}

template <int NP>
hipError_t launch_recon(const CrtMats *M, const CrtMats &h, const OzakiConsts *kc, hipStream_t s) {
  int gx, gy;
  const bool fast = crt_grid(CRT_RECON, h, gx, gy);
  dim3 grid((unsigned)gx, (unsigned)gy);
  if (fast)
    hipLaunchKernelGGL((crt_recon_kernel<NP, true>), grid, dim3(T), 0, s, M, kc);
  else
    hipLaunchKernelGGL((crt_recon_kernel<NP, false>), grid, dim3(T), 0, s, M, kc);
  return hipGetLastError();
}

}  // namespace

void rew_grid(int count, long long n, int nmod, bool vec4, int &gx, int &gy) {
  const long long units = vec4 ? (n >> 2) : n;
  long long blocks = (units + T - 1) / T;
  long long cap = (long long)device_cu_count() * 8 / ((long long)count * nmod);
  if (cap < 1) cap = 1;
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  gx = (int)blocks;
  gy = count * nmod;
}

bool crt_grid(int k, const CrtMats &h, int &gx, int &gy) {
  const int leaves = 1 << (2 * h.nrec);
  const int nblk = 1 << h.nrec;
  const bool plain = (h.bs == 1 || h.bs == h.m);
  const long long mm = (long long)h.m * h.m;
  gy = 1;
  switch (k) {
    case CRT_ROWPART:
      gx = (h.n + T - 1) / T;
      gy = nblk * h.chunks;
      return plain;
    case CRT_COLPART:
      gx = (int)(((long long)h.n * nblk + 3) / 4);
      return plain;
    case CRT_FINALIZE:
      gx = (h.n + T - 1) / T;
      return false;
    case CRT_SPLIT_A: {
      const bool fast = plain && h.m % 64 == 0;
      const int tiles = (h.m + 63) / 64;
      // fast: 32-row x 128-k tiles; else 64 x 64
      gx = fast ? (h.m / 32) * ((h.m + 127) / 128) : tiles * tiles;
      gy = leaves;
      return fast;
    }
    case CRT_SPLIT_B: {
      const bool fb = plain && h.m % 16 == 0;
      gx = (int)capped_blocks(fb ? mm / 16 : mm, leaves);
      gy = leaves;
      return fb;
    }
    case CRT_RECON: {
      const bool fast = (h.m % 4 == 0) && plain;
      gx = (int)capped_blocks(fast ? mm / 4 : mm, leaves);
      gy = leaves;
      return fast;
    }
  }
  gx = 0;
  return false;
}

hipError_t launch_crt_multi(int k, bool flag, int np, const char *base, const VSeg *segs,
                            int nseg, int blocks, const OzakiConsts *kc, hipStream_t s) {
  if (nseg <= 0 || blocks <= 0) return hipSuccess;
  switch (k) {
    case CRT_ROWPART: return launch_crt_multi_f<CRT_ROWPART>(flag, np, base, segs, nseg, blocks, kc, s);
    case CRT_COLPART: return launch_crt_multi_f<CRT_COLPART>(flag, np, base, segs, nseg, blocks, kc, s);
    case CRT_FINALIZE: return launch_crt_multi_f<CRT_FINALIZE>(false, np, base, segs, nseg, blocks, kc, s);
    case CRT_SPLIT_A: return launch_crt_multi_f<CRT_SPLIT_A>(flag, np, base, segs, nseg, blocks, kc, s);
    case CRT_SPLIT_B: return launch_crt_multi_f<CRT_SPLIT_B>(flag, np, base, segs, nseg, blocks, kc, s);
    case CRT_RECON: return launch_crt_multi_f<CRT_RECON>(flag, np, base, segs, nseg, blocks, kc, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_residue_ew_multi(int op, const char *base, const VSeg *segs, int nseg,
                                   int blocks, bool vec4, const OzakiConsts *kc, hipStream_t s) {
  if (nseg <= 0 || blocks <= 0) return hipSuccess;
  switch (op) {
    case EW_BFS_PRE: return launch_rew_op_multi<EW_BFS_PRE>(base, segs, nseg, blocks, vec4, kc, s);
    case EW_BFS_POST: return launch_rew_op_multi<EW_BFS_POST>(base, segs, nseg, blocks, vec4, kc, s);
    case EW_DFS_PRE1: return launch_rew_op_multi<EW_DFS_PRE1>(base, segs, nseg, blocks, vec4, kc, s);
    case EW_DFS_PRE2: return launch_rew_op_multi<EW_DFS_PRE2>(base, segs, nseg, blocks, vec4, kc, s);
    case EW_DFS_PRE3: return launch_rew_op_multi<EW_DFS_PRE3>(base, segs, nseg, blocks, vec4, kc, s);
    case EW_DFS_POST1: return launch_rew_op_multi<EW_DFS_POST1>(base, segs, nseg, blocks, vec4, kc, s);
    case EW_DFS_POST2: return launch_rew_op_multi<EW_DFS_POST2>(base, segs, nseg, blocks, vec4, kc, s);
    case EW_DFS_POST3: return launch_rew_op_multi<EW_DFS_POST3>(base, segs, nseg, blocks, vec4, kc, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_crt_step(int step, const CrtMats *M, const CrtMats &h, const OzakiConsts *kc,
                           hipStream_t s) {
  const int leaves = 1 << (2 * h.nrec);
  if (leaves > 65535) return hipErrorInvalidValue;
  int gx, gy;
  bool f;
  switch (step) {
    case 0: {
      f = crt_grid(CRT_ROWPART, h, gx, gy);
      if (f)
        hipLaunchKernelGGL((crt_rowpart_kernel<true>), dim3(gx, gy), dim3(T), 0, s, M);
      else
        hipLaunchKernelGGL((crt_rowpart_kernel<false>), dim3(gx, gy), dim3(T), 0, s, M);
      f = crt_grid(CRT_COLPART, h, gx, gy);
      if (f)
        hipLaunchKernelGGL((crt_colpart_kernel<true>), dim3(gx), dim3(T), 0, s, M);
      else
        hipLaunchKernelGGL((crt_colpart_kernel<false>), dim3(gx), dim3(T), 0, s, M);
      crt_grid(CRT_FINALIZE, h, gx, gy);
      hipLaunchKernelGGL(crt_finalize_kernel, dim3(gx), dim3(T), 0, s, M, kc);
      return hipGetLastError();
    }
    case 1: {
      f = crt_grid(CRT_SPLIT_A, h, gx, gy);
      if (f)
        hipLaunchKernelGGL((crt_split_a_kernel<true>), dim3(gx, gy), dim3(T), 0, s, M, kc);
      else
        hipLaunchKernelGGL((crt_split_a_kernel<false>), dim3(gx, gy), dim3(T), 0, s, M, kc);
      if (h.fused_b) return hipGetLastError();
      f = crt_grid(CRT_SPLIT_B, h, gx, gy);
      if (f)
        hipLaunchKernelGGL((crt_split_b_kernel<true>), dim3(gx, gy), dim3(T), 0, s, M, kc);
      else
        hipLaunchKernelGGL((crt_split_b_kernel<false>), dim3(gx, gy), dim3(T), 0, s, M, kc);
      return hipGetLastError();
    }
    case 2: {
      switch (ozaki_npieces(h.nmod)) {
        case 1:
        case 2: return launch_recon<2>(M, h, kc, s);
        case 3: return launch_recon<3>(M, h, kc, s);
        default: return launch_recon<4>(M, h, kc, s);
      }
    }
  }
  return hipErrorInvalidValue;
}

hipError_t launch_residue_ew(int op, const RPtrs *tab, int count, long long n, int nmod,
                             bool vec4, const OzakiConsts *kc, hipStream_t s) {
  if (count <= 0 || n <= 0) return hipSuccess;
  if (nmod < 2 || nmod > OZ_MAX_MODULI) return hipErrorInvalidValue;
  for (int c0 = 0; c0 < count; c0 += 65535) {
    const RPtrs *t = tab + c0;
    const int c = count - c0 < 65535 ? count - c0 : 65535;
    hipError_t e = hipErrorInvalidValue;
    switch (op) {
      case EW_BFS_PRE: e = launch_rew_op<EW_BFS_PRE>(t, c, n, nmod, vec4, kc, s); break;
      case EW_BFS_POST: e = launch_rew_op<EW_BFS_POST>(t, c, n, nmod, vec4, kc, s); break;
      case EW_DFS_PRE1: e = launch_rew_op<EW_DFS_PRE1>(t, c, n, nmod, vec4, kc, s); break;
      case EW_DFS_PRE2: e = launch_rew_op<EW_DFS_PRE2>(t, c, n, nmod, vec4, kc, s); break;
      case EW_DFS_PRE3: e = launch_rew_op<EW_DFS_PRE3>(t, c, n, nmod, vec4, kc, s); break;
      case EW_DFS_POST1: e = launch_rew_op<EW_DFS_POST1>(t, c, n, nmod, vec4, kc, s); break;
      case EW_DFS_POST2: e = launch_rew_op<EW_DFS_POST2>(t, c, n, nmod, vec4, kc, s); break;
      case EW_DFS_POST3: e = launch_rew_op<EW_DFS_POST3>(t, c, n, nmod, vec4, kc, s); break;
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace caps
