// Two consecutive batched ('b') Winograd levels fused into single passes.
//
// Pre-additions: the 16 depth-2 sub-blocks of A (of B) at one position give
// all 49 depth-2 operands T_j(T_i) (S_j(S_i)) at that position; one launch
// reads each input element once and writes the 40 operands that are not
// plain sub-blocks (the 9 with i, j in {T1, T2, T7} / {S1, S2, S6} are passed
// by pointer).  Post-combinations: the 49 depth-2 products at one position
// give C's 16 depth-2 sub-blocks.  The expression trees are exactly those of
// two strassenBFS levels (multiply.cpp:272, 277, 195-210) applied one after
// the other, so FP64 results are bit-identical to the unfused levels; on
// residues every step is exact modulo m_l.
//
// HBM-bound: per position, FP64 pre reads 16 x 8 B and writes 40 x 8 B per
// side; FP64 post reads 49 x 8 B and writes 16 x 8 B.  The residue post is
// fused with the CRT reconstruction: 49 x nmod bytes in, 16 FP64 results out.
#include "ew_ops.h"
#include "fused2.h"
#include "ozaki_dev.h"
#include "tilemap.h"

namespace caps {

namespace {

constexpr int T = 256;

// level-1 operands at one sub-position from the four quadrant values (order
// 11, 21, 12, 22) -> T1..T7 (A side) or S1..S7 (B side)
template <typename V>
__device__ __forceinline__ void level_ops(int side, const V (&q)[4], V (&o)[7]) {
  V x[8], y[8];
  if (side == 0) {
    x[0] = q[0]; x[1] = q[1]; x[2] = q[2]; x[3] = q[3];
    x[4] = q[0]; x[5] = q[0]; x[6] = q[0]; x[7] = q[0];
  } else {
    x[0] = q[0]; x[1] = q[0]; x[2] = q[0]; x[3] = q[0];
    x[4] = q[0]; x[5] = q[1]; x[6] = q[2]; x[7] = q[3];
  }
  Ew<EW_BFS_PRE>::apply(x, y);  // y: T3 T4 T5 T6 S3 S4 S5 S7
  if (side == 0) {
    o[0] = q[0]; o[1] = q[2]; o[2] = y[0]; o[3] = y[1]; o[4] = y[2]; o[5] = y[3]; o[6] = q[3];
  } else {
    o[0] = q[0]; o[1] = q[1]; o[2] = y[4]; o[3] = y[5]; o[4] = y[6]; o[5] = q[3]; o[6] = y[7];
  }
}

// ---------------------------------------------------------------- FP64 pre
// thread: 2 consecutive positions (16-byte accesses); grid.y: side
__global__ __launch_bounds__(T) void bb_pre_f64_kernel(const BBPre *__restrict__ P, long long q) {
  const int side = blockIdx.y;
  const double *in = reinterpret_cast<const double *>(P->in[side]);
  const long long n2 = q >> 1;
  for (long long u = (long long)blockIdx.x * T + threadIdx.x; u < n2;
       u += (long long)gridDim.x * T) {
    double x0[4][4], x1[4][4];  // [q1][q2], two positions
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) {
        const double2 v = reinterpret_cast<const double2 *>(in + (4 * a + b) * q)[u];
        x0[a][b] = v.x;
        x1[a][b] = v.y;
      }
#pragma unroll
    for (int h = 0; h < 2; h++) {
      double (&x)[4][4] = h ? x1 : x0;
      double l1[7][4];  // level-1 operand i at sub-position q2
#pragma unroll
      for (int b = 0; b < 4; b++) {
        double qv[4] = {x[0][b], x[1][b], x[2][b], x[3][b]}, o[7];
        level_ops(side, qv, o);
#pragma unroll
        for (int i = 0; i < 7; i++) l1[i][b] = o[i];
      }
#pragma unroll
      for (int i = 0; i < 7; i++) {
        double qv[4] = {l1[i][0], l1[i][1], l1[i][2], l1[i][3]}, o[7];
        level_ops(side, qv, o);
#pragma unroll
        for (int j = 0; j < 7; j++) {
          double *dst = reinterpret_cast<double *>(P->out[side][7 * i + j]);
          if (dst) dst[2 * u + h] = o[j];
        }
      }
    }
  }
}

// ---------------------------------------------------------------- FP64 post
// thread: 2 consecutive positions
__global__ __launch_bounds__(T) void bb_post_f64_kernel(const BBPost *__restrict__ P, long long q) {
  const long long n2 = q >> 1;
  for (long long u = (long long)blockIdx.x * T + threadIdx.x; u < n2;
       u += (long long)gridDim.x * T) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
      double l1[7][4];  // product i of level 1 at sub-position q2
#pragma unroll
      for (int i = 0; i < 7; i++) {
        double qv[7], c[4];
#pragma unroll
        for (int j = 0; j < 7; j++)
          qv[j] = reinterpret_cast<const double *>(P->q[7 * i + j])[2 * u + h];
        Ew<EW_BFS_POST>::apply(qv, c);  // C11 C21 C12 C22 of product i
#pragma unroll
        for (int b = 0; b < 4; b++) l1[i][b] = c[b];
      }
#pragma unroll
      for (int b = 0; b < 4; b++) {
        double qv[7], c[4];
#pragma unroll
        for (int i = 0; i < 7; i++) qv[i] = l1[i][b];
        Ew<EW_BFS_POST>::apply(qv, c);
#pragma unroll
        for (int a = 0; a < 4; a++)
          reinterpret_cast<double *>(P->c)[(4 * a + b) * q + 2 * u + h] = c[a];
      }
    }
  }
}

// ---------------------------------------------------------------- residue pre
// thread: 4 consecutive positions (4-byte accesses); grid.y: side; grid.z: modulus
__global__ __launch_bounds__(T) void bb_pre_res_kernel(const BBPre *__restrict__ P, long long q,
                                                       const OzakiConsts *__restrict__ kc) {
  const int side = blockIdx.y, l = blockIdx.z;
  const int mi = kc->m[l];
  const float inv = 1.0f / (float)mi;
  const int8_t *in = reinterpret_cast<const int8_t *>(P->in[side]) + l * P->sin[side];
  const long long n4 = q >> 2;
  for (long long u = (long long)blockIdx.x * T + threadIdx.x; u < n4;
       u += (long long)gridDim.x * T) {
    unsigned w[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) w[a][b] = reinterpret_cast<const unsigned *>(in + (4 * a + b) * q)[u];
    unsigned out[49];
#pragma unroll
    for (int k = 0; k < 49; k++) out[k] = 0;
#pragma unroll
    for (int byte = 0; byte < 4; byte++) {
      int l1[7][4];
#pragma unroll
      for (int b = 0; b < 4; b++) {
        int qv[4], o[7];
#pragma unroll
        for (int a = 0; a < 4; a++) qv[a] = (int)(int8_t)((w[a][b] >> (8 * byte)) & 0xff);
        level_ops(side, qv, o);
#pragma unroll
        for (int i = 0; i < 7; i++) l1[i][b] = o[i];
      }
#pragma unroll
      for (int i = 0; i < 7; i++) {
        int qv[4] = {l1[i][0], l1[i][1], l1[i][2], l1[i][3]}, o[7];
        level_ops(side, qv, o);
#pragma unroll
        for (int j = 0; j < 7; j++)
          out[7 * i + j] |= (unsigned)(mod_i32(o[j], mi, inv) & 0xff) << (8 * byte);
      }
    }
#pragma unroll
    for (int k = 0; k < 49; k++) {
      int8_t *dst = reinterpret_cast<int8_t *>(P->out[side][k]);
      if (dst) reinterpret_cast<unsigned *>(dst + l * P->sout[side][k])[u] = out[k];
    }
  }
}

// ---------------------------------------------------------------- residue post + CRT
// thread: one position; loops over the moduli, accumulating the CRT sums of
// C's 16 sub-blocks, then reconstructs, unscales and writes FP64 C
template <int NP>
__global__ __launch_bounds__(T) void bb_post_crt_kernel(const BBPost *__restrict__ P, long long q,
                                                        const OzakiConsts *__restrict__ kc) {
  const int m0 = P->m0, sub = P->sub, nrec2 = P->nrec2, bs = P->bs;
  const long long mm0 = (long long)m0 * m0;
  for (long long u = (long long)blockIdx.x * T + threadIdx.x; u < q;
       u += (long long)gridDim.x * T) {
    double S[16][NP];
#pragma unroll
    for (int c = 0; c < 16; c++)
#pragma unroll
      for (int p = 0; p < NP; p++) S[c][p] = 0.0;
    for (int l = 0; l < kc->nmod; l++) {
      const int mi = kc->m[l];
      const float inv = 1.0f / (float)mi;
      int l1[7][4];
#pragma unroll
      for (int i = 0; i < 7; i++) {
        int qv[7], c[4];
#pragma unroll
        for (int j = 0; j < 7; j++)
          qv[j] = (int)reinterpret_cast<const int8_t *>(P->q[7 * i + j])[l * P->sq[7 * i + j] + u];
        Ew<EW_BFS_POST>::apply(qv, c);
#pragma unroll
        for (int b = 0; b < 4; b++) l1[i][b] = c[b];  // |.| <= 512
      }
#pragma unroll
      for (int b = 0; b < 4; b++) {
        int qv[7], c[4];
#pragma unroll
        for (int i = 0; i < 7; i++) qv[i] = l1[i][b];
        Ew<EW_BFS_POST>::apply(qv, c);  // |.| <= 2048
#pragma unroll
        for (int a = 0; a < 4; a++) crt_acc<NP>(S[4 * a + b], mod_i32(c[a], mi, inv), l, kc);
      }
    }
    // position u of a depth-2 sub-block: leaf L2 (quadtree of nrec2 levels), column-major leaf
    const long long L2 = u / mm0, e = u - L2 * mm0;
    const int li = (int)(e % m0), lj = (int)(e / m0);
    int bi, bj;
    leaf_coords((int)L2, nrec2, bi, bj);
    const long long dst_in_sub = L2 * mm0 + leaf_elem_offset(m0, bs, li, lj);
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) {
        const int gi = (a & 1) * 2 * sub + (b & 1) * sub + bi * m0 + li;
        const int gj = (a >> 1) * 2 * sub + (b >> 1) * sub + bj * m0 + lj;
        const double X = crt_finish<NP>(S[4 * a + b], kc);
        P->out_c[(4 * a + b) * q + dst_in_sub] = ldexp(X, -(P->sa[gi] + P->sb[gj]));
      }
  }
}

unsigned blocks_for(long long units, int per_grid) {
  long long b = (units + T - 1) / T;
  long long cap = (long long)device_cu_count() * 8 / (per_grid > 0 ? per_grid : 1);
  if (cap < 1) cap = 1;
  if (b > cap) b = cap;
  return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

hipError_t launch_bb_pre(const BBPre *P, long long q, bool residue, int nmod,
                         const OzakiConsts *kc, hipStream_t s) {
  if (!residue) {
    if (q % 2) return hipErrorInvalidValue;
    dim3 g(blocks_for(q / 2, 2), 2);
    hipLaunchKernelGGL(bb_pre_f64_kernel, g, dim3(T), 0, s, P, q);
  } else {
    if (q % 4) return hipErrorInvalidValue;
    dim3 g(blocks_for(q / 4, 2 * nmod), 2, nmod);
    hipLaunchKernelGGL(bb_pre_res_kernel, g, dim3(T), 0, s, P, q, kc);
  }
  return hipGetLastError();
}

hipError_t launch_bb_post(const BBPost *P, long long q, bool residue, int npieces,
                          const OzakiConsts *kc, hipStream_t s) {
  if (!residue) {
    if (q % 2) return hipErrorInvalidValue;
    hipLaunchKernelGGL(bb_post_f64_kernel, dim3(blocks_for(q / 2, 1)), dim3(T), 0, s, P, q);
  } else {
    dim3 g(blocks_for(q, 1));
    switch (npieces) {
      case 1:
      case 2: hipLaunchKernelGGL((bb_post_crt_kernel<2>), g, dim3(T), 0, s, P, q, kc); break;
      case 3: hipLaunchKernelGGL((bb_post_crt_kernel<3>), g, dim3(T), 0, s, P, q, kc); break;
      default: hipLaunchKernelGGL((bb_post_crt_kernel<4>), g, dim3(T), 0, s, P, q, kc); break;
    }
  }
  return hipGetLastError();
}

}  // namespace caps
