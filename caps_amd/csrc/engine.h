// Host engine: turns (descriptor, pattern, mode) into a list of batched
// device launches, owns the device workspace and replays the list as a
// captured hipGraph.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/caps_amd.h"
#include "crt.h"
#include "kernels.h"

namespace caps {

// the whole recursion on residues (INT8 or BF16 leaf products)
inline bool residue_mode(int mode) {
  return mode == CAPS_MODE_CRT_INT8 || mode == CAPS_MODE_CRT_BF16;
}

// A matrix operand of a batched level: FP64 (plane stride unused) or nmod
// INT8 residue planes, plane l at p + l * ps bytes.
struct Mat {
  char *p;
  long long ps;
};

// one sub-problem of a batched level: C = A * B, all s x s in CAPS layout.
// cseg: residue leaves only, C's segment stride (RGemm::cseg; 128 = plain)
struct Prob {
  Mat A, B, C;
  long long cseg = 128;
};

enum StepKind : int {
  STEP_EW = 0,       // fused Winograd adds (EwOp in `op`)
  STEP_GEMM_F64 = 1, // batched FP64 leaf GEMM
  STEP_REORDER = 2,  // bs-block <-> column-major leaf (op = 1: to column-major)
  STEP_OZ_SCALE = 3, // split mode: per-row / per-column scaling exponents
  STEP_OZ_SPLIT = 4, // split mode: scaled operands -> INT8 residue planes
  STEP_OZ_GEMM = 5,  // split mode: INT8 GEMM per modulus, residue outputs
  STEP_OZ_RECON = 6, // split mode: CRT reconstruction + unscaling
  STEP_REW = 7,      // residue mode: Winograd adds on residue planes (EwOp in `op`)
  STEP_CRT_SCALE = 8,  // residue mode: scale exponents of the whole A and B
  STEP_CRT_SPLIT = 9,  // residue mode: A, B -> residue planes
  STEP_CRT_RECON = 10, // residue mode: C residues -> C
  STEP_BB_PRE = 11,    // two 'b' levels of pre-additions in one pass (op = residue)
  STEP_BB_POST = 12,   // two 'b' levels of combinations (residue: + reconstruction)
  STEP_BB_SPLIT_B = 13,  // residue mode: B's split fused with its two top 'b' levels
  // segmented launches (kernels.h VSeg) merging one step of several
  // problems of an uneven batch: table = segment array, count = segments,
  // n = flat blocks (or tiles) of the launch
  STEP_MULTI_EW = 14,    // FP64 Winograd adds, op = EwOp, vec2 = 16-byte accesses
  STEP_MULTI_REW = 15,   // residue Winograd adds, op = EwOp, vec2 = 4-byte accesses
  STEP_MULTI_GEMM = 16,  // FP64 leaf GEMMs, bs = block tile
  STEP_MULTI_RGEMM = 17, // INT8 residue leaf GEMMs (128 x 128 tiles)
  STEP_MULTI_CRT = 18,   // residue scale / split / reconstruct kernel op (CrtKernel), vec2 = flag
  STEP_BB_SPLIT_A = 19,  // residue mode: A's split fused with its two top 'b' levels, bs = m0
};

struct Step {
  int kind;
  int op;
  int count;      // sub-problems in the batch
  long long n;    // elements per quadrant (EW) or leaf size m (GEMM/REORDER)
  int bs;
  int vec2;
  int levels;     // leaf GEMMs: quadrant levels inside the operands (cutoff)
  int tile;       // leaf GEMMs: 0 default block tile, 1 the smaller one
  size_t table;   // byte offset of this step's pointer table
  size_t aux;     // byte offset of an auxiliary workspace region
};

// A plan owns one device workspace, so it belongs to one stream: calls with
// the same buffers on different streams get separate plans and never share
// a workspace.
struct PlanKey {
  const void *A, *B, *C;
  caps_desc desc;
  std::string pattern;
  int mode, moduli, cutoff;
  bool fuse;
  std::string batch;  // batched plans: every item's pointers, descriptor and pattern
  const void *stream;
  int tile;           // leaf-GEMM block tile: 0 default, 1 the smaller one
  bool operator==(const PlanKey &o) const;
};

struct Plan {
  PlanKey key;
  std::vector<Step> steps;
  std::vector<unsigned char> host_tables;
  void *d_tables = nullptr;
  void *d_work = nullptr;
  size_t work_bytes = 0;
  hipGraphExec_t graph = nullptr;
  hipStream_t capture_stream = nullptr;
  long long additions = 0, multiplications = 0, leaves = 0;
  std::string exec_pattern;
  unsigned long long last_use = 0;
  // split-mode constants (moduli, CRT weights), device copy in d_tables
  size_t oz_const = 0;
  int oz_moduli = 0;
  // residue mode: host copies of the residue descriptions (Step.aux indexes)
  std::vector<CrtMats> crt_host;
  // batched plans: the key of every item
  std::string batch_key;
};

// one independent multiply of a batched call
struct BatchItem {
  const double *A, *B;
  double *C;
  caps_desc desc;
  std::string pattern;  // resolved, one char per recursed level
  int cutoff = 0;       // layout levels multiplied classically
};

// Build a plan: pattern must already be resolved to one char per level.
// work_base = nullptr computes sizes only.
int build_plan(Plan &plan, const double *A, const double *B, double *C,
               const caps_desc &d, const std::string &pattern, int mode, int moduli,
               int cutoff, void *work_base, bool fuse = true, int tile = 0);

// Build one plan for independent multiplies of possibly different shapes:
// items with the same (descriptor, pattern, cutoff) form a group that shares
// batched steps.  If the groups' workspaces together fit merge_budget bytes,
// each group gets its own region and the same step of different groups runs
// as one segmented launch (one leaf-GEMM launch for all shapes, one launch
// per Winograd phase, one per scale / split / reconstruct kernel); otherwise
// the groups run one after another in one shared region.
int build_batch_plan(Plan &plan, const std::vector<BatchItem> &items, int mode, int moduli,
                     void *work_base, int tile = 0, size_t merge_budget = 0);

// Enqueue every step of the plan on `s` (no graph).
hipError_t run_steps(const Plan &plan, hipStream_t s);

// Same, with a HIP event pair around every step; adds each step's elapsed
// milliseconds to ms[timer of its kind] and counts launches in n[].  Waits
// for the stream.
hipError_t run_steps_timed(const Plan &plan, hipStream_t s, double *ms, long long *n);

// Workspace (bytes) a resolved pattern needs.
size_t plan_workspace_bytes(const caps_desc &d, const std::string &pattern, int mode,
                            int moduli, int cutoff = 0);

}  // namespace caps
