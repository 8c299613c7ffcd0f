// Residue-domain Strassen ("CRT mode", CAPS_MODE_CRT_INT8).
//
// The whole multiply runs over the integers modulo M = prod m_l:
//   1. scale once: row i of A by 2^sa_i, column j of B by 2^sb_j, from
//      deterministic 2-norm bounds over the full rows / columns, so that
//      |(A'B')_ij| <= ||A'_i.|| ||B'_.j|| <= L^2 < M/2 with A' = trunc(2^sa A),
//      B' = trunc(2^sb B) exact FP64 integers;
//   2. split once: A', B' -> symmetric residues modulo each m_l (INT8 planes);
//      A-side leaves are stored row-major (k contiguous) for the MFMA operand;
//   3. run the Winograd recursion of the chosen pattern on the residue planes:
//      every pre-addition and combination is exact modulo m_l, and the leaf
//      products are the INT8 MFMA GEMMs with INT32 accumulation (mod m_l);
//   4. reconstruct once: C' = A'B' from its residues by the CRT in FP64, then
//      C = 2^-(sa_i + sb_j) C'.
// Strassen-Winograd is an identity over any commutative ring, so C' is the
// exact integer product: the result carries only the truncation of step 1
// and the final rounding, not the FP64 rounding of the Winograd additions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"
#include "ozaki.h"

namespace caps {

struct CrtMats {
  const double *A;  // n x n CAPS layout (nrec, bs)
  const double *B;
  double *C;
  int8_t *ra;  // residues of A': nmod planes of n*n bytes, leaves row-major
  int8_t *rb;  // residues of B': leaves column-major
  int8_t *rc;  // residues of C': leaves column-major
  int *sa;     // n row exponents of A
  int *sb;     // n column exponents of B
  double *pa;  // row partials (e, s) of A: [2^nrec * chunks][n] pairs
  double *pb;  // column partials of B: [2^nrec][n] pairs
  int n, nrec, bs, m, nmod, chunks;
  int fused_b;  // B's split is fused with the top pre-additions (no rb planes)
  int fused_a;  // A's split is fused with the top pre-additions (no ra planes)
};

// residue Winograd step: like EwPtrs with a plane stride per pointer
struct RPtrs {
  const int8_t *in[8];
  int8_t *out[8];
  long long sin[8];
  long long sout[8];
};

// step 0 = scale, 1 = split, 2 = reconstruct
hipError_t launch_crt_step(int step, const CrtMats *d_mats, const CrtMats &host_mats,
                           const OzakiConsts *d_consts, hipStream_t s);

// Winograd add step `op` (EwOp) on residue planes, all nmod moduli
hipError_t launch_residue_ew(int op, const RPtrs *d_tab, int count, long long n, int nmod,
                             bool vec4, const OzakiConsts *d_consts, hipStream_t s);

// ---- segmented launches over the items of an uneven batch (kernels.h VSeg)
// the kernels of the scale, split and reconstruct steps
enum CrtKernel : int {
  CRT_ROWPART = 0,
  CRT_COLPART = 1,
  CRT_FINALIZE = 2,
  CRT_SPLIT_A = 3,
  CRT_SPLIT_B = 4,
  CRT_RECON = 5
};
// the grid kernel k has for one item alone (gx x gy) and its template flag
// (column-major leaves / aligned sizes); items merged in one launch share it
bool crt_grid(int k, const CrtMats &h, int &gx, int &gy);
hipError_t launch_crt_multi(int k, bool flag, int np, const char *d_tables, const VSeg *d_segs,
                            int nseg, int blocks, const OzakiConsts *d_consts, hipStream_t s);
// residue Winograd step over segments (gx blocks x entries * moduli, aux = moduli)
void rew_grid(int count, long long n, int nmod, bool vec4, int &gx, int &gy);
hipError_t launch_residue_ew_multi(int op, const char *d_tables, const VSeg *d_segs, int nseg,
                                   int blocks, bool vec4, const OzakiConsts *d_consts,
                                   hipStream_t s);

// chunks of a leaf row used by the row-partials reduction
inline int crt_row_chunks(int m) {
  int c = m / 256;
  return c < 1 ? 1 : (c > 8 ? 8 : c);
}

}  // namespace caps
