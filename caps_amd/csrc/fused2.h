// Two consecutive batched Winograd levels in single passes (fused2.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"
#include "ozaki.h"

namespace caps {

// Level-1 operands that are plain quadrants (passed by pointer, not
// written): T1 = A11, T2 = A12, T7 = A22 and S1 = B11, S2 = B21, S6 = B22.
// Depth-2 operand (i, j) is plain iff i and j both are.
__host__ __device__ constexpr bool bb_plain(int side, int t) {
  return side == 0 ? (t == 0 || t == 1 || t == 6) : (t == 0 || t == 1 || t == 5);
}

// pre-additions of two 'b' levels for both operands of one sub-problem
struct BBPre {
  const void *in[2];      // A (side 0) and B (side 1): 16 depth-2 sub-blocks of q elements,
                          // sub-block (a, b) at element (4a + b) * q
  void *out[2];           // the 40 written depth-2 operands (i, j) = operand j of level-1
                          // operand i, not plain, in (i, j) order, one every `step` bytes
  long long step;         // bytes between consecutive operands (q elements)
  long long sin[2];       // residue mode: plane strides (bytes)
  long long sout[2];
};

// post-combinations of two 'b' levels for one sub-problem
struct BBPost {
  const void *q;          // the 49 depth-2 products (i, j), product 7 i + j at (7 i + j) * step
  long long step;         // bytes
  long long sq;           // residue mode: plane stride
  // residue mode: position u of a product at (u >> 7) * cseg + (u & 127)
  // (128: plain planes; 49 * 128 with step = 128: the products interleaved
  // in 128-byte segments, as the 256-tile leaf GEMM writes them)
  long long cseg;
  double *c;              // FP64 mode: C's 16 depth-2 sub-blocks at (4a + b) * q
  // residue mode: reconstruct into FP64 C (CAPS layout) directly
  double *out_c;
  const int *sa, *sb;     // scale exponents of the global rows / columns
  int sub;                // depth-2 sub-block size (n / 4)
  int m0;                 // leaf size of the layout
  int nrec2;              // quadrant levels inside a depth-2 sub-block (nrec - 2)
  int bs;
};

// residue mode at the top of the recursion: the B side of the pre-additions
// fused with B's split.  FP64 B (column-major leaves) in, the residues of
// all 49 depth-2 operands out (the plain ones too).
struct BBSplitB {
  const double *B;        // 16 depth-2 sub-blocks of q elements, (a, b) at element (4a + b) * q
  int8_t *out;            // operand (i, j) of modulus l at out + l * sout + (7 i + j) * q
  long long sout;         // plane stride (bytes)
  const int *sb;          // scale exponents of B's global columns
  int sub;                // depth-2 sub-block size (n / 4)
  int m0;                 // leaf size (even)
  int nrec2;              // quadrant levels inside a depth-2 sub-block
};

// residue mode at the top of the recursion: the A side of the pre-additions
// fused with A's split.  FP64 A (column-major leaves) in, the residues of all
// 49 depth-2 operands out as row-major leaves (the layout the leaf GEMM reads
// A in; the plain operands too).
struct BBSplitA {
  const double *A;        // 16 depth-2 sub-blocks of q elements, (a, b) at element (4a + b) * q
  int8_t *out;            // operand (i, j) of modulus l at out + l * sout + (7 i + j) * q
  long long sout;         // plane stride (bytes)
  const int *sa;          // scale exponents of A's global rows
  int sub;                // depth-2 sub-block size (n / 4)
  int m0;                 // leaf size (a multiple of 128)
  int nrec2;              // quadrant levels inside a depth-2 sub-block
};

// FP64: q % 2 == 0; residue: q % 4 == 0.  a_only (residue): side B is
// produced by launch_bb_split_b instead.
hipError_t launch_bb_pre(const BBPre *d_P, long long q, bool residue, int nmod,
                         const OzakiConsts *d_consts, hipStream_t s, bool a_only = false);
// q % 4 == 0, q < 2^31, m0 even
hipError_t launch_bb_split_b(const BBSplitB *d_P, long long q, const OzakiConsts *d_consts,
                             hipStream_t s);
// m0 % 128 == 0 (host copy of the table's m0), q < 2^31
hipError_t launch_bb_split_a(const BBSplitA *d_P, long long q, int m0,
                             const OzakiConsts *d_consts, hipStream_t s);
hipError_t launch_bb_post(const BBPost *d_P, long long q, bool residue, int npieces,
                          const OzakiConsts *d_consts, hipStream_t s);

}  // namespace caps
