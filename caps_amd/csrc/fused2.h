// Two consecutive batched Winograd levels in single passes (fused2.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"
#include "ozaki.h"

namespace caps {

// pre-additions of two 'b' levels for both operands of one sub-problem
struct BBPre {
  const void *in[2];      // A (side 0) and B (side 1): 16 depth-2 sub-blocks of q elements,
                          // sub-block (a, b) at element (4a + b) * q
  void *out[2][49];       // depth-2 operand (i, j) = operand j of level-1 operand i;
                          // null: a plain sub-block passed by pointer
  long long sin[2];       // residue mode: plane strides (bytes)
  long long sout[2][49];
};

// post-combinations of two 'b' levels for one sub-problem
struct BBPost {
  const void *q[49];      // depth-2 products (i, j)
  long long sq[49];       // residue mode: plane strides
  double *c;              // FP64 mode: C's 16 depth-2 sub-blocks at (4a + b) * q
  // residue mode: reconstruct into FP64 C (CAPS layout) directly
  double *out_c;
  const int *sa, *sb;     // scale exponents of the global rows / columns
  int sub;                // depth-2 sub-block size (n / 4)
  int m0;                 // leaf size of the layout
  int nrec2;              // quadrant levels inside a depth-2 sub-block (nrec - 2)
  int bs;
};

// FP64: q % 2 == 0; residue: q % 4 == 0
hipError_t launch_bb_pre(const BBPre *d_P, long long q, bool residue, int nmod,
                         const OzakiConsts *d_consts, hipStream_t s);
hipError_t launch_bb_post(const BBPost *d_P, long long q, bool residue, int npieces,
                          const OzakiConsts *d_consts, hipStream_t s);

}  // namespace caps
