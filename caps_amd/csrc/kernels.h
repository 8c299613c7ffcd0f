// Host-side declarations of the HIP kernel launchers (gfx950 only).
//
// Every launcher enqueues on the given stream and returns the hipError_t of
// the launch.  Pointer tables (one entry per sub-problem of a batched level)
// live in device memory; the host engine builds them once per plan.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace caps {

// ---------------------------------------------------------------- Winograd adds
// Fused elementwise steps of one Winograd level.  Each evaluates, per element,
// exactly the expression tree of the reference's add routines so the FP64
// results are bit-identical to them (multiply.cpp:110-210, 272, 277, 409-451).
enum EwOp : int {
  EW_BFS_PRE = 0,   // in A11 A21 A12 A22 B11 B21 B12 B22 -> T3 T4 T5 T6 S3 S4 S5 S7
  EW_BFS_POST = 1,  // in Q1..Q7 -> C11 C21 C12 C22                (calcCij)
  EW_DFS_PRE1 = 2,  // in A11 A21 A22 B11 B12 B22 -> S5 S3 S4 T5 T3
  EW_DFS_PRE2 = 3,  // in T3 A11 -> T4
  EW_DFS_PRE3 = 4,  // in A12 T4 S4 B21 -> T6 S7
  EW_DFS_POST1 = 5, // in Q1 Q4 Q5 Q3 Q7 -> U3 C22 C21
  EW_DFS_POST2 = 6, // in Q1 Q2 -> C11
  EW_DFS_POST3 = 7, // in U3 Q6 -> C12
  EW_NUM_OPS = 8
};

struct EwPtrs {
  const double *in[8];
  double *out[8];
};

// Number of additions the reference counts for one element of each op.
int ew_additions(int op);

hipError_t launch_ew(int op, const EwPtrs *d_tab, int count, long long n,
                     bool vec2, hipStream_t s);

// ---------------------------------------------------------------- layout
// column-major n x n <-> CAPS layout (nproc = 1)
hipError_t launch_to_caps(const double *cm, double *caps, int n, int nrec,
                          int bs, hipStream_t s);
hipError_t launch_from_caps(const double *caps, double *cm, int n, int nrec,
                            int bs, hipStream_t s);

// bs-blocked leaf <-> column-major leaf, batched (multiply.cpp:1093-1139)
struct ReorderPtrs {
  const double *src;
  double *dst;
};
hipError_t launch_leaf_reorder(const ReorderPtrs *d_tab, int count, int m,
                               int bs, bool to_colmajor, hipStream_t s);

// ---------------------------------------------------------------- FP64 GEMM
struct GemmPtrs {
  const double *A;
  const double *B;
  double *C;
};

// C_p = A_p * B_p for every table entry, beta = 0.  Operands are m x m
// quadtrees of `levels` quadrant levels over column-major leaves of size
// m >> levels (levels = 0: plain column-major; levels > 0 needs leaves that
// are multiples of 128).  One persistent launch over all tiles of all problems.
// tile: 0 = 128 x 128 block tiles, 1 = 64 x 64.
hipError_t launch_dgemm_batched(const GemmPtrs *d_tab, int count, int m, int levels, int tile,
                                hipStream_t s);

int device_cu_count();

// ---------------------------------------------------------------- segmented launches
// One launch over the same step of several independent problems (the uneven
// batch): segment s owns the flat blocks [blk0, blk0 + gx * gy) and runs them
// as a gx x gy grid over its own argument, which sits `arg` bytes into the
// plan's device tables (segment tables are built before those tables have a
// device address).  n is the segment's element count where the step has one.
struct VSeg {
  long long arg;
  long long n;
  int blk0, gx, gy, aux;
};

// segment of flat block b (segments sorted by blk0, the first at 0)
__device__ __forceinline__ int vseg_find(const VSeg *__restrict__ segs, int nseg, int b) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (segs[mid].blk0 <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// FP64 leaf GEMMs over segments of different sizes, one persistent launch:
// segment args are GemmPtrs tables, n = order, gx = tiles per dimension, gy =
// problems, blk0 = first global tile, aux = quadtree levels | edge checks <<
// 8.  bt: block tile (dgemm_block_tile).
int dgemm_block_tile(int m, int levels, int tile);
hipError_t launch_dgemm_multi(int bt, const char *d_tables, const VSeg *d_segs, int nseg,
                              int total_tiles, hipStream_t s);

// Winograd add step `op` over segments whose args are EwPtrs tables (gy =
// entries, n = elements per quadrant); vec2: every entry 16-byte aligned
hipError_t launch_ew_multi(int op, const char *d_tables, const VSeg *d_segs, int nseg,
                           int blocks, bool vec2, hipStream_t s);
// the local grid launch_ew gives one step of `count` entries of n elements
void ew_grid(int count, long long n, bool vec2, int &gx, int &gy);

}  // namespace caps
