// Split-precision leaf product (Ozaki scheme II: integer modular emulation of
// an FP64 GEMM; K. Ozaki, Y. Uchino, T. Imamura, "Ozaki Scheme II: A
// GEMM-oriented emulation of floating-point matrix multiplication using an
// integer modular technique", 2025).  No counterpart in the reference, whose
// leaf is a BLAS dgemm (dgemm-blas.cpp:4-13).
//
// For a leaf product C = A * B (m x m, column-major FP64):
//  1. scale: row i of A by 2^sa_i, column j of B by 2^sb_j, powers of two
//     chosen from deterministic 2-norm upper bounds so that
//        sum_k |A'_ik| |B'_kj| <= ||A'_i.|| ||B'_.j|| <= L^2 < M / 2,
//     with A' = trunc(2^sa A), B' = trunc(2^sb B) (exact FP64 integers) and
//     M the product of the moduli;
//  2. split: A' mod m_l and B' mod m_l as symmetric residues (INT8 planes),
//     A-side planes row-major (k contiguous), B-side column-major;
//  3. product: C'_l = A'_l B'_l on the INT8 matrix cores with exact INT32
//     accumulation, reduced mod m_l in the epilogue (INT8 output plane);
//  4. reconstruct: X = sum_l C'_l W_l  (mod M) by the Chinese remainder
//     theorem in FP64 with the weights W_l split into 37-bit pieces, so every
//     partial sum is exact; X = A'B' exactly (assembled in 128 bits), returned
//     correctly rounded to FP64, then C = 2^-(sa_i + sb_j) X.
// Integer-valued leaves whose products fit 53 bits come out exact.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

namespace caps {

constexpr int OZ_MAX_MODULI = 16;  // M <= 2^128: X fits 128 bits
constexpr int OZ_MAX_PIECES = 4;
constexpr int OZ_PIECE_BITS = 37;
constexpr int OZ_DEFAULT_MODULI = 14;
// row / column chunks of the deterministic sum-of-squares reduction
constexpr int OZ_CHUNKS = 8;

struct OzakiConsts {
  int nmod;
  int npieces;
  int m[OZ_MAX_MODULI];
  int c14[OZ_MAX_MODULI];                  // 2^14 mod m_l, symmetric (|c14| <= m_l / 2)
  int c22[OZ_MAX_MODULI];                  // round(2^22 / m_l): quotients of small sums (mod_k)
  double inv_m[OZ_MAX_MODULI];
  double w[OZ_MAX_MODULI][OZ_MAX_PIECES];  // CRT weight W_l in base 2^37 digits
  double Mp[OZ_MAX_PIECES];                // M in base 2^37 digits
  double invM;                             // 1 / M (rounded)
  double L;                                // per-side norm bound, L^2 < M/2
  double log2M;
};

struct OzLeaf {
  const double *A;  // m x m column-major
  const double *B;
  double *C;
  int *ea;          // m row exponents of A
  int *eb;          // m column exponents of B
  int8_t *ra;       // nmod planes, m x m row-major
  int8_t *rb;       // nmod planes, m x m column-major
  int8_t *rc;       // nmod planes, m x m column-major (product residues)
  double *pa;       // OZ_CHUNKS x m partial row sums of squares (A)
  double *pb;       // m column sums of squares (B)
};

// One leaf product of the INT8 GEMM step: nmod residue planes per operand,
// plane l of an operand at base + l * stride (bytes).  a: row-major m x m
// (k contiguous), b and c: column-major m x m.  Element e of C's plane l sits
// at c + l * sc + (e >> 7) * cseg + (e & 127): cseg = 128 is the plain plane;
// a larger cseg interleaves the 128-byte segments of several products (the
// 256-tile kernels only: residue_gemm_segmented says which launches take them).
struct RGemm {
  const int8_t *a;
  const int8_t *b;
  int8_t *c;
  long long sa, sb, sc;
  long long cseg;
};

// Host: moduli and CRT constants for nmod moduli (2..OZ_MAX_MODULI).
// Returns false if nmod is out of range.
bool ozaki_make_consts(int nmod, OzakiConsts &k);

// The moduli used (greedy pairwise-coprime, descending from 256).
const int *ozaki_moduli_table();

// Number of 37-bit pieces of M for nmod moduli.
int ozaki_npieces(int nmod);

// step: 0 = scale, 1 = split, 3 = reconstruct (OzLeaf tables)
hipError_t launch_ozaki_step(int step, const OzLeaf *d_tab, int count, int m, int nmod,
                             const OzakiConsts *d_consts, hipStream_t s);

// C_l = A_l B_l (mod m_l) for every leaf of the table and every modulus, one
// persistent launch.  Operands are m x m quadtrees of `levels` quadrant levels
// over leaves of size m >> levels (A leaves row-major, B and C column-major);
// levels > 0 needs leaves that are multiples of 128.
// the same products on BF16 MFMA with FP32 accumulation (bit-identical results)
hipError_t launch_residue_gemm_bf16(const RGemm *d_tab, int count, int m, int levels, int nmod,
                                    const OzakiConsts *d_consts, hipStream_t s);
// The 128 x 128 residue products over segments of different sizes in one
// persistent launch (kernels.h VSeg: args are RGemm tables, n = order, gx =
// tiles per dimension, gy = leaves x moduli, aux = quadtree levels | check
// mode << 8, the mode residue_gemm_check gives; -1: not supported).
int residue_gemm_check(int m, int levels);
// launch_residue_gemm takes the large tiles with partial edges for this
// (plain leaves, m % 16 == 0, little padding)
bool big_edge_gemm(int m, int levels, int tile);
// launch_residue_gemm (INT8) / launch_residue_gemm_bf16 (bf16) runs these
// leaves on the 256-tile kernel for whole tiles, which honours RGemm::cseg
bool residue_gemm_segmented(int m, int levels, int tile, bool bf16);
hipError_t launch_residue_gemm_multi(const char *d_tables, const VSeg *d_segs, int nseg,
                                     int total_tiles, int nmod, const OzakiConsts *d_consts,
                                     hipStream_t s);
// tile: 0 = 256 x 256 block tiles where the leaves allow, 1 = 128 x 128
hipError_t launch_residue_gemm(const RGemm *d_tab, int count, int m, int levels, int nmod,
                               const OzakiConsts *d_consts, hipStream_t s, int tile = 0);

}  // namespace caps
