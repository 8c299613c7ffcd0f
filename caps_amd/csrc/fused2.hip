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
// side; FP64 post reads 49 x 8 B and writes 16 x 8 B.  At the top of a
// residue-domain multiply the splits of B and A are fused with their
// pre-additions (16 FP64 values in, the residues of all 49 operands out; A's
// transposed to row-major leaves on the way), and the residue post with the
// CRT reconstruction: 49 x nmod bytes in, 16 FP64 results out (its VALU work,
// about 4,900 lane instructions per position for 14 moduli, is close to its
// memory time).
#include "ew_ops.h"
#include "fused2.h"
#include "ozaki_dev.h"
#include "tilemap.h"

namespace caps {

namespace {

constexpr int T = 256;

// global-memory views (accesses through them are global_*, not flat_*)
typedef __attribute__((address_space(1))) unsigned short gushort;
typedef const __attribute__((address_space(1))) unsigned gcuint;
typedef __attribute__((address_space(1))) unsigned guint;
typedef __attribute__((address_space(1))) double gdouble;
typedef const __attribute__((address_space(1))) int gcint;
typedef unsigned v4u __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) v4u gv4u;

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
// thread: 2 consecutive positions, 16-byte loads and stores; grid.y: side
__global__ __launch_bounds__(T) void bb_pre_f64_kernel(const BBPre *__restrict__ P, long long q) {
  const int side = blockIdx.y;
  const double *in = reinterpret_cast<const double *>(P->in[side]);
  const long long n2 = q >> 1;
  for (long long u = (long long)blockIdx.x * T + threadIdx.x; u < n2;
       u += (long long)gridDim.x * T) {
    double l1[2][7][4];  // [position][level-1 operand i][sub-position q2]
#pragma unroll
    for (int b = 0; b < 4; b++) {
      double q0[4], q1[4];
#pragma unroll
      for (int a = 0; a < 4; a++) {
        const double2 v = reinterpret_cast<const double2 *>(in + (4 * a + b) * q)[u];
        q0[a] = v.x;
        q1[a] = v.y;
      }
      double o0[7], o1[7];
      level_ops(side, q0, o0);
      level_ops(side, q1, o1);
#pragma unroll
      for (int i = 0; i < 7; i++) {
        l1[0][i][b] = o0[i];
        l1[1][i][b] = o1[i];
      }
    }
    char *out = reinterpret_cast<char *>(P->out[side]);
    int k = 0;  // index among the written operands (compile-time after unrolling)
#pragma unroll
    for (int i = 0; i < 7; i++) {
      double q0[4] = {l1[0][i][0], l1[0][i][1], l1[0][i][2], l1[0][i][3]};
      double q1[4] = {l1[1][i][0], l1[1][i][1], l1[1][i][2], l1[1][i][3]};
      double o0[7], o1[7];
      level_ops(side, q0, o0);
      level_ops(side, q1, o1);
#pragma unroll
      for (int j = 0; j < 7; j++) {
        if (bb_plain(side, i) && bb_plain(side, j)) continue;
        reinterpret_cast<double2 *>(out + k * P->step)[u] = make_double2(o0[j], o1[j]);
        k++;
      }
    }
  }
}

// ---------------------------------------------------------------- FP64 post
// thread: 2 consecutive positions, 16-byte loads and stores
__global__ __launch_bounds__(T) void bb_post_f64_kernel(const BBPost *__restrict__ P, long long q) {
  const long long n2 = q >> 1;
  for (long long u = (long long)blockIdx.x * T + threadIdx.x; u < n2;
       u += (long long)gridDim.x * T) {
    double l1[2][7][4];  // [position][product i of level 1][sub-position q2]
#pragma unroll
    for (int i = 0; i < 7; i++) {
      double q0[7], q1[7], c0[4], c1[4];
#pragma unroll
      for (int j = 0; j < 7; j++) {
        const double2 v =
            reinterpret_cast<const double2 *>(reinterpret_cast<const char *>(P->q) +
                                              (7 * i + j) * P->step)[u];
        q0[j] = v.x;
        q1[j] = v.y;
      }
      Ew<EW_BFS_POST>::apply(q0, c0);  // C11 C21 C12 C22 of product i
      Ew<EW_BFS_POST>::apply(q1, c1);
#pragma unroll
      for (int b = 0; b < 4; b++) {
        l1[0][i][b] = c0[b];
        l1[1][i][b] = c1[b];
      }
    }
#pragma unroll
    for (int b = 0; b < 4; b++) {
      double q0[7], q1[7], c0[4], c1[4];
#pragma unroll
      for (int i = 0; i < 7; i++) {
        q0[i] = l1[0][i][b];
        q1[i] = l1[1][i][b];
      }
      Ew<EW_BFS_POST>::apply(q0, c0);
      Ew<EW_BFS_POST>::apply(q1, c1);
#pragma unroll
      for (int a = 0; a < 4; a++)
        reinterpret_cast<double2 *>(reinterpret_cast<double *>(P->c) + (4 * a + b) * q)[u] =
            make_double2(c0[a], c1[a]);
    }
  }
}

// ---------------------------------------------------------------- residue pre
// thread: 4 consecutive positions (one 4-byte word per operand); grid.y:
// modulus; one instantiation per side.  Depth-2 values are sums of at most 16
// residues: |x| <= 2048, reduced by mod_k.  Offsets within a plane are 32-bit
// (q < 2^31).
template <int SIDE>
__global__ __launch_bounds__(T) void bb_pre_res_kernel(const BBPre *__restrict__ P, long long q,
                                                       const OzakiConsts *__restrict__ kc) {
  const int l = blockIdx.y;
  const int mi = kc->m[l], c22 = kc->c22[l];
  const char *in = reinterpret_cast<const char *>(P->in[SIDE]) + l * P->sin[SIDE];
  char *obase = reinterpret_cast<char *>(P->out[SIDE]) + l * P->sout[SIDE];
  const long long step = P->step;
  const unsigned qq = (unsigned)q, n4 = (unsigned)(q >> 2);
  for (unsigned u = blockIdx.x * T + threadIdx.x; u < n4; u += gridDim.x * T) {
    unsigned w[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) w[a][b] = *(gcuint *)(in + (4 * a + b) * qq + 4 * u);
    int l1[4][7][4];  // [position][level-1 operand i][sub-position b]
#pragma unroll
    for (int h = 0; h < 4; h++)
#pragma unroll
      for (int b = 0; b < 4; b++) {
        int qv[4], o[7];
#pragma unroll
        for (int a = 0; a < 4; a++) qv[a] = (int)__builtin_amdgcn_sbfe(w[a][b], 8 * h, 8);
        level_ops(SIDE, qv, o);
#pragma unroll
        for (int i = 0; i < 7; i++) l1[h][i][b] = o[i];
      }
    int k = 0;  // index among the written operands (compile-time after unrolling)
#pragma unroll
    for (int i = 0; i < 7; i++) {
      int o[4][7];
#pragma unroll
      for (int h = 0; h < 4; h++) level_ops(SIDE, l1[h][i], o[h]);
#pragma unroll
      for (int j = 0; j < 7; j++) {
        if (bb_plain(SIDE, i) && bb_plain(SIDE, j)) continue;
        const unsigned v = pack4_lo(mod_k(o[0][j], c22, mi), mod_k(o[1][j], c22, mi),
                                    mod_k(o[2][j], c22, mi), mod_k(o[3][j], c22, mi));
        __builtin_nontemporal_store(v, (guint *)(obase + k * step + 4 * u));
        k++;
      }
    }
  }
}

// ------------------------------------------------ residue pre of B fused with the split
// thread: 2 consecutive positions of one leaf column (m0 even).  The 16 FP64
// inputs are scaled by their global columns' exponents and truncated once,
// exactly as the separate split does; then for every modulus they are reduced
// to symmetric residues and combined into all 49 depth-2 operands.  Per
// position: 16 x 8 bytes in, 49 x nmod bytes out.
__global__ __launch_bounds__(T) void bb_split_b_kernel(const BBSplitB *__restrict__ P,
                                                       long long q,
                                                       const OzakiConsts *__restrict__ kc) {
  const int nmod = kc->nmod;
  const unsigned n2 = (unsigned)(q >> 1);
  const unsigned m0 = (unsigned)P->m0, mm0 = m0 * m0;
  const int sub = P->sub;
  for (unsigned u = blockIdx.x * T + threadIdx.x; u < n2; u += gridDim.x * T) {
    const unsigned p = 2 * u;
    const unsigned L2 = p / mm0, e = p - L2 * mm0;
    int bi, bj;
    leaf_coords((int)L2, P->nrec2, bi, bj);
    const int cj = bj * (int)m0 + (int)(e / m0);
    int sc[4];  // exponent of column half (a >> 1, b >> 1) as index 2 (a >> 1) + (b >> 1)
#pragma unroll
    for (int h = 0; h < 4; h++) sc[h] = P->sb[(h >> 1) * 2 * sub + (h & 1) * sub + cj];
    double x[16][2];
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const double2 v = *reinterpret_cast<const double2 *>(P->B + (long long)k * q + p);
      const int s = sc[2 * (k >> 3) + ((k & 3) >> 1)];
      x[k][0] = trunc(ldexp(v.x, s));
      x[k][1] = trunc(ldexp(v.y, s));
    }
    for (int l = 0; l < nmod; l++) {
      const int mi = kc->m[l];
      const double inv = kc->inv_m[l];
      const int c22 = kc->c22[l];
      int8_t *ob = P->out + l * P->sout;
      int l1[2][7][4];  // [position][level-1 operand i][sub-position b]
#pragma unroll
      for (int h = 0; h < 2; h++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
          int qv[4], o[7];
#pragma unroll
          for (int a = 0; a < 4; a++) qv[a] = sym_residue(x[4 * a + b][h], mi, inv, c22);
          level_ops(1, qv, o);
#pragma unroll
          for (int i = 0; i < 7; i++) l1[h][i][b] = o[i];
        }
#pragma unroll
      for (int i = 0; i < 7; i++) {
        int o0[7], o1[7];
        level_ops(1, l1[0][i], o0);
        level_ops(1, l1[1][i], o1);
#pragma unroll
        for (int j = 0; j < 7; j++) {
          // plain operands are input residues, already reduced
          const bool plain = bb_plain(1, i) && bb_plain(1, j);
          const int r0 = plain ? o0[j] : mod_k(o0[j], c22, mi);
          const int r1 = plain ? o1[j] : mod_k(o1[j], c22, mi);
          const unsigned v = (unsigned)(r0 & 0xff) | ((unsigned)(r1 & 0xff) << 8);
          __builtin_nontemporal_store((unsigned short)v, (gushort *)(ob + (7 * i + j) * q) + u);
        }
      }
    }
  }
}

// ------------------------------------------------ residue pre of A fused with the split
// The leaf GEMM reads A's operands as row-major leaves, and A's leaves are
// column-major, so the pass transposes on the way.  Block: a tile of 16 rows x
// 128 columns of one leaf in all 16 depth-2 sub-blocks; thread: 2 consecutive
// columns of one row, lanes along the rows (each 8-byte load of a wave reads
// four whole 128-byte column runs).  The FP64 inputs are scaled by their
// global rows' exponents and truncated once, as the separate split does; per
// modulus the 49 operand bytes of the tile's 2,048 positions go through LDS
// ([operand][row][column], 128-byte rows whose 16-byte chunk j sits in slot
// j ^ ((row >> 1) & 7), so that the 16 rows a 16-lane cycle touches spread
// over the 64 banks: conflict cycles 4.5e7 per launch against 5.6e7 with rows
// padded to 144 bytes, 0.81 against 0.825 ms) and leave as whole 128-byte
// rows of the row-major leaves (64-byte pieces, 16 x 64 tiles, took 1.15
// against 0.82 ms; 8-byte pieces instead of 16, 0.83 against 0.81 ms).
// Per position: 16 x 8 bytes in, 49 x nmod bytes out.
constexpr int SA_R = 16, SA_C = 128, SA_T = 1024, SA_PITCH = 128;

__device__ __forceinline__ int sa_slot(int row, int chunk) { return chunk ^ ((row >> 1) & 7); }
static_assert(SA_T == SA_R * SA_C / 2, "two positions per thread");

__global__ __launch_bounds__(SA_T) void bb_split_a_kernel(const BBSplitA *__restrict__ P,
                                                          long long q,
                                                          const OzakiConsts *__restrict__ kc) {
  __shared__ __align__(16) unsigned char S[49 * SA_R * SA_PITCH];
  const int nmod = kc->nmod;
  const int m0 = P->m0, sub = P->sub;
  const long long mm0 = (long long)m0 * m0;
  const int tr = m0 / SA_R, tc = m0 / SA_C;  // row and column tiles per leaf
  const unsigned L2 = blockIdx.x / (unsigned)(tr * tc), w = blockIdx.x % (unsigned)(tr * tc);
  const int r0 = (int)(w % tr) * SA_R, c0 = (int)(w / tr) * SA_C;
  int bi, bj;
  leaf_coords((int)L2, P->nrec2, bi, bj);
  const int t = threadIdx.x, r = t & (SA_R - 1), c = 2 * (t / SA_R);
  const int ri = bi * m0 + r0 + r;  // row within a depth-2 sub-block
  int sr[4];  // exponent of row half (a & 1, b & 1) as index 2 (a & 1) + (b & 1)
#pragma unroll
  for (int h = 0; h < 4; h++) sr[h] = P->sa[(h >> 1) * 2 * sub + (h & 1) * sub + ri];
  const double *src = P->A + L2 * mm0 + (long long)(c0 + c) * m0 + r0 + r;
  double x[16][2];
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const double *pk = src + (long long)k * q;
    const int s = sr[2 * ((k >> 2) & 1) + (k & 1)];
    x[k][0] = trunc(ldexp(pk[0], s));
    x[k][1] = trunc(ldexp(pk[m0], s));
  }
  unsigned char *sw = S + r * SA_PITCH + (sa_slot(r, c >> 4) << 4) + (c & 15);
  for (int l = 0; l < nmod; l++) {
    const int mi = kc->m[l];
    const double inv = kc->inv_m[l];
    const int c22 = kc->c22[l];
    int l1[2][7][4];  // [column][level-1 operand i][sub-position b]
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int b = 0; b < 4; b++) {
        int qv[4], o[7];
#pragma unroll
        for (int a = 0; a < 4; a++) qv[a] = sym_residue(x[4 * a + b][h], mi, inv, c22);
        level_ops(0, qv, o);
#pragma unroll
        for (int i = 0; i < 7; i++) l1[h][i][b] = o[i];
      }
#pragma unroll
    for (int i = 0; i < 7; i++) {
      int o0[7], o1[7];
      level_ops(0, l1[0][i], o0);
      level_ops(0, l1[1][i], o1);
#pragma unroll
      for (int j = 0; j < 7; j++) {
        const bool plain = bb_plain(0, i) && bb_plain(0, j);
        const int v0 = plain ? o0[j] : mod_k(o0[j], c22, mi);
        const int v1 = plain ? o1[j] : mod_k(o1[j], c22, mi);
        *reinterpret_cast<unsigned short *>(sw + (7 * i + j) * SA_R * SA_PITCH) =
            (unsigned short)((v0 & 0xff) | ((v1 & 0xff) << 8));
      }
    }
    __syncthreads();
    // 49 operands x 16 rows x 8 pieces of 16 bytes
    int8_t *ob = P->out + l * P->sout + L2 * mm0 + (long long)r0 * m0 + c0;
    for (int idx = t; idx < 49 * SA_R * 8; idx += SA_T) {
      const int ch = idx & 7, rr = (idx >> 3) & (SA_R - 1), k = idx >> 7;
      const v4u v = *reinterpret_cast<const v4u *>(S + (k * SA_R + rr) * SA_PITCH + 16 * sa_slot(rr, ch));
      __builtin_nontemporal_store(v, (gv4u *)(ob + k * q + (long long)rr * m0 + 16 * ch));
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- residue post + CRT
// One block per PPOS consecutive positions of the depth-2 sub-blocks, two
// phases through LDS (each product is read in PPOS-byte segments: 128 keeps
// them whole cache lines; with cseg = 49 x 128 the 49 segments of a modulus
// are one contiguous run):
//   1. lane per (4 positions, modulus): 49 product words, both levels of the
//      post-combination (on 16-bit pairs), the 16 residues of C's sub-blocks
//      into LDS;
//   2. lane per (position, sub-block of C), consecutive lanes on consecutive
//      positions: CRT over the moduli from LDS, unscale, FP64 store.
typedef short v2s __attribute__((ext_vector_type(2)));

// the signed bytes 0, 1 and 2, 3 of w as pairs of 16-bit integers: v_perm_b32
// selectors 8 and 9 replicate the sign bit of bytes 1 and 3 of its second
// operand, 10 and 11 those of bytes 1 and 3 of its first (here w << 8, whose
// bytes 1 and 3 are bytes 0 and 2 of w)
__device__ __forceinline__ void sext_pairs(unsigned w, v2s &lo, v2s &hi) {
  const unsigned s = w << 8;
  lo = __builtin_bit_cast(v2s, __builtin_amdgcn_perm(s, w, 0x08010a00u));
  hi = __builtin_bit_cast(v2s, __builtin_amdgcn_perm(s, w, 0x09030b02u));
}

constexpr int PPOS = 128;
constexpr int PG = PPOS / 4;  // 4-position words per block
static_assert(T % PPOS == 0, "phase 2 maps a lane to one position");

template <int NP>
__global__ __launch_bounds__(T) void bb_post_crt_kernel(const BBPost *__restrict__ P, long long q,
                                                        const OzakiConsts *__restrict__ kc) {
  __shared__ unsigned R[OZ_MAX_MODULI][16][PG];
  const int nmod = kc->nmod;
  const long long pos0 = (long long)blockIdx.x * PPOS;
  for (int unit = threadIdx.x; unit < nmod * PG; unit += T) {
    const int g = unit % PG, l = unit / PG;
    const long long u4 = pos0 / 4 + g;
    if (4 * u4 >= q) continue;  // q % 4 == 0
    const int mi = kc->m[l], c22 = kc->c22[l];
    // this lane's word of product 0; product k lies k * step bytes further
    // (a wave-uniform offset)
    const long long u = 4 * u4;
    const char *qb =
        reinterpret_cast<const char *>(P->q) + l * P->sq + (u >> 7) * P->cseg + (u & 127);
    const long long step = P->step;
    unsigned w[49];
#pragma unroll
    for (int k = 0; k < 49; k++) w[k] = *(gcuint *)(qb + k * step);
    // both levels on pairs of 16-bit integers, positions (0, 1) and (2, 3) of
    // the words (|level 1| <= 512, |level 2| <= 2048)
    v2s l1[7][4][2];  // [level-1 operand i][sub-position bb][pair]
#pragma unroll
    for (int i = 0; i < 7; i++) {
      v2s qlo[7], qhi[7], clo[4], chi[4];
#pragma unroll
      for (int j = 0; j < 7; j++) sext_pairs(w[7 * i + j], qlo[j], qhi[j]);
      Ew<EW_BFS_POST>::apply(qlo, clo);
      Ew<EW_BFS_POST>::apply(qhi, chi);
#pragma unroll
      for (int bb = 0; bb < 4; bb++) {
        l1[i][bb][0] = clo[bb];
        l1[i][bb][1] = chi[bb];
      }
    }
#pragma unroll
    for (int bb = 0; bb < 4; bb++) {
      v2s qlo[7], qhi[7], clo[4], chi[4];
#pragma unroll
      for (int i = 0; i < 7; i++) {
        qlo[i] = l1[i][bb][0];
        qhi[i] = l1[i][bb][1];
      }
      Ew<EW_BFS_POST>::apply(qlo, clo);
      Ew<EW_BFS_POST>::apply(qhi, chi);
#pragma unroll
      for (int aa = 0; aa < 4; aa++)
        R[l][4 * aa + bb][g] = pack4_lo(mod_k(clo[aa].x, c22, mi), mod_k(clo[aa].y, c22, mi),
                                        mod_k(chi[aa].x, c22, mi), mod_k(chi[aa].y, c22, mi));
    }
  }
  __syncthreads();

  const int m0 = P->m0, sub = P->sub, nrec2 = P->nrec2, bs = P->bs;
  const long long mm0 = (long long)m0 * m0;
  // lane -> one position pl (T is a multiple of PPOS) and outputs o = o0, o0 + T / PPOS, ...
  const int pl = threadIdx.x % PPOS, o0 = threadIdx.x / PPOS;
  const long long u = pos0 + pl;
  if (u >= q) return;
  // position u of a depth-2 sub-block: leaf L2 (quadtree of nrec2 levels), column-major leaf
  const long long L2 = u / mm0, e = u - L2 * mm0;
  const int li = (int)(e % m0), lj = (int)(e / m0);
  int bi, bj;
  leaf_coords((int)L2, nrec2, bi, bj);
  const long long dst_in_sub = L2 * mm0 + leaf_elem_offset(m0, bs, li, lj);
  const int ri = bi * m0 + li, cj = bj * m0 + lj;
  const unsigned lds_shift = 8 * (pl & 3);
  // outputs o = o0 + 2 i (T / PPOS = 2): sub-block (aa, bb) = (i >> 1, o0 + 2 (i & 1)),
  // so the lane's 8 outputs use 2 row and 4 column exponents, loaded once
  static_assert(T / PPOS == 2, "outputs o0 + 2 i");
  const gcint *ea = (const gcint *)P->sa + o0 * sub + ri;
  const gcint *eb = (const gcint *)P->sb + cj;
  const int xa[2] = {ea[0], ea[2 * sub]};
  const int xb[4] = {eb[0], eb[sub], eb[2 * sub], eb[3 * sub]};
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int o = o0 + 2 * i;
    double S[NP];
#pragma unroll
    for (int p = 0; p < NP; p++) S[p] = 0.0;
    for (int l = 0; l < nmod; l++)
      crt_acc<NP>(S, (int)__builtin_amdgcn_sbfe(R[l][o][pl >> 2], lds_shift, 8), l, kc);
    const double X = crt_finish<NP>(S, kc);
    // row exponent of sub-block row (aa & 1, bb & 1), column of (aa >> 1, bb >> 1)
    const int e = xa[(i >> 1) & 1] + xb[2 * (i >> 2) + (i & 1)];
    __builtin_nontemporal_store(ldexp(X, -e), (gdouble *)P->out_c + o * q + dst_in_sub);
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
                         const OzakiConsts *kc, hipStream_t s, bool a_only) {
  if (!residue) {
    if (q % 2) return hipErrorInvalidValue;
    dim3 g(blocks_for(q / 2, 2), 2);
    hipLaunchKernelGGL(bb_pre_f64_kernel, g, dim3(T), 0, s, P, q);
  } else {
    if (q % 4 || q >= (1LL << 31)) return hipErrorInvalidValue;
    dim3 g(blocks_for(q / 4, 2 * nmod), nmod);
    hipLaunchKernelGGL(bb_pre_res_kernel<0>, g, dim3(T), 0, s, P, q, kc);
    if (!a_only) hipLaunchKernelGGL(bb_pre_res_kernel<1>, g, dim3(T), 0, s, P, q, kc);
  }
  return hipGetLastError();
}

hipError_t launch_bb_split_b(const BBSplitB *P, long long q, const OzakiConsts *kc,
                             hipStream_t s) {
  if (q % 4 || q >= (1LL << 31)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(bb_split_b_kernel, dim3(blocks_for(q / 2, 1)), dim3(T), 0, s, P, q, kc);
  return hipGetLastError();
}

hipError_t launch_bb_split_a(const BBSplitA *P, long long q, int m0, const OzakiConsts *kc,
                             hipStream_t s) {
  if (m0 <= 0 || m0 % SA_C || q % ((long long)m0 * m0) || q >= (1LL << 31))
    return hipErrorInvalidValue;
  const long long blocks = q / (SA_R * SA_C);  // one per 16 x 128 tile
  hipLaunchKernelGGL(bb_split_a_kernel, dim3((unsigned)blocks), dim3(SA_T), 0, s, P, q, kc);
  return hipGetLastError();
}

hipError_t launch_bb_post(const BBPost *P, long long q, bool residue, int npieces,
                          const OzakiConsts *kc, hipStream_t s) {
  if (!residue) {
    if (q % 2) return hipErrorInvalidValue;
    hipLaunchKernelGGL(bb_post_f64_kernel, dim3(blocks_for(q / 2, 1)), dim3(T), 0, s, P, q);
  } else {
    if (q % 4) return hipErrorInvalidValue;
    dim3 g((unsigned)((q + PPOS - 1) / PPOS));
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
