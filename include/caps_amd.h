/*
 * caps_amd.h - C ABI of the MI355X (gfx950) CAPS Strassen-Winograd multiply.
 *
 * Plain C: pointers, sizes and status codes only.  Every entry point returns
 * a caps_status (0 = success); caps_last_error() gives a message for the
 * calling thread's last failure.  Nothing here aborts the process.
 *
 * Reference interfaces replaced (paths relative to the CAPS source tree):
 *   caps_multiply / caps_multiply_device
 *       -> void multiply(double*, double*, double*, MatDescriptor, char*)
 *          (multiply.h:13, multiply.cpp:9-25); the C++ overload with that
 *          exact signature is in include/caps/multiply.h.
 *   caps_desc        -> struct MatDescriptor (matrix.h:8-16), same 7 ints in
 *                       the same order (DESC_SIZE = 7, matrix.h:18).
 *   caps_verify_desc -> verifyDescriptor (matrix.cpp:3-7) for one device.
 *   caps_num_entries -> numEntriesPerProc (matrix.cpp:9-12).
 *   caps_to_layout / caps_from_layout (+ _device)
 *       -> distributeFrom1Proc / collectTo1Proc with one process
 *          (communication.cpp:612-708, communication.h).
 *   caps_get_counters / caps_reset_counters
 *       -> the counters printed by printCounters (counters.cpp:33-74).
 *   caps_set_available_memory -> setAvailableMemory (memory.cpp:31-33), the
 *       budget that drives the automatic BFS/DFS choice (memory.cpp:22-29).
 */
#ifndef CAPS_AMD_H
#define CAPS_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same fields, order and size as MatDescriptor (matrix.h:8-16).  On one
 * device nproc = nprocr = nprocc = nproc_summa = 1. */
typedef struct caps_desc {
  int lda;
  int nrec;
  int nproc;
  int nprocr;
  int nprocc;
  int nproc_summa;
  int bs;
} caps_desc;

typedef enum caps_status {
  CAPS_OK = 0,
  CAPS_EINVAL = 1,      /* bad descriptor, size or argument               */
  CAPS_EPATTERN = 2,    /* pattern char not in [bBdDhH] or too short      */
  CAPS_ENOMEM = 3,      /* device or host allocation failed / over budget */
  CAPS_EHIP = 4,        /* a HIP runtime call failed                      */
  CAPS_ENODEV = 5,      /* no gfx950 device                               */
  CAPS_EUNSUPPORTED = 6 /* valid request this build does not implement   */
} caps_status;

/* Leaf (base-case) product arithmetic. */
typedef enum caps_mode {
  CAPS_MODE_FP64 = 0,       /* FP64 MFMA leaf GEMM                           */
  CAPS_MODE_OZAKI_INT8 = 1, /* leaf-only split: FP64 Winograd additions; each
                               leaf product from FP64 operands scaled to
                               integers, split into residues modulo
                               pairwise-coprime moduli <= 256 (INT8), exact
                               INT8 MFMA products with INT32 accumulation and
                               CRT reconstruction in FP64                    */
  CAPS_MODE_CRT_INT8 = 2,   /* residue-domain Strassen: A and B scaled and
                               split into INT8 residues once, every Winograd
                               addition exact modulo each m_l, INT8 MFMA leaf
                               products, one CRT reconstruction of C; exact
                               for integer products below 2^53, otherwise
                               only the initial scaling truncates (nrec <= 7) */
  CAPS_MODE_CRT_BF16 = 3,   /* as CAPS_MODE_CRT_INT8 with the leaf products on
                               BF16 MFMA (residues widened to BF16, FP32
                               accumulation folded mod m every 512 k, so
                               exact): the same bits, at the BF16 rate      */
  CAPS_MODE_AUTO = 99       /* the planner (autotuner / learned model) picks */
} caps_mode;

typedef struct caps_opts {
  int mode;      /* caps_mode                                              */
  int moduli;    /* split modes: number of moduli (0 = default 14, 2..16)  */
  int flags;     /* CAPS_FLAG_* below                                      */
  int cutoff;    /* recursion cutoff: the last `cutoff` of the layout's nrec
                    levels are not Strassen steps; their sub-blocks are
                    multiplied classically by the leaf GEMM reading the
                    quadtree layout directly (0 = recurse to the leaves, as
                    the reference does; > 0 needs leaves that are multiples
                    of 128, bs = 1 or bs = leaf size, and mode FP64 or
                    CRT_INT8)                                              */
} caps_opts;

/* Re-use a captured hipGraph for repeated calls with the same pointers,
 * descriptor, pattern, options and stream (default on).  A cached plan owns
 * one device workspace and belongs to one stream: calls on different streams
 * never share a workspace, and calls on one stream run in stream order. */
#define CAPS_FLAG_NO_GRAPH 1
/* Force the reference's automatic rule when pattern == NULL (DFS at every
 * level on one process, multiply.cpp:43-66) instead of the planner's choice. */
#define CAPS_FLAG_REFERENCE_AUTO 2
/* Run the steps eagerly with a HIP event pair around each launch and add the
 * elapsed times to the phase timers of caps_get_counters (the reference's
 * startTimer / stopTimer around each add, dgemm and reorder,
 * counters.cpp:19-31).  Synchronises the stream before returning. */
#define CAPS_FLAG_TIMERS 4
/* Run two leading 'b' levels as separate passes instead of the fused
 * two-level kernels (same results; for testing and timing comparisons). */
#define CAPS_FLAG_NO_FUSION 8
/* Leaf GEMMs on the smaller block tile: INT8 residue products 128 x 128
 * instead of 256 x 256, FP64 64 x 64 instead of 128 x 128 (same results; the
 * autotuner's tile dimension, caps_candidate.tile). */
#define CAPS_FLAG_SMALL_TILE 16

/* Select the device (HIP ordinal) for the calling thread and check that it
 * is gfx950.  Optional: the first call of any entry point does it for the
 * current device. */
int caps_init(int device);

/* Message for the calling thread's last failed call ("" if none). */
const char *caps_last_error(void);

/* verifyDescriptor (matrix.cpp:3-7) for one device, plus bs | leaf size. */
int caps_verify_desc(const caps_desc *desc);

/* numEntriesPerProc (matrix.cpp:9-12): lda*lda/nproc/nproc_summa. */
long long caps_num_entries(const caps_desc *desc);

/* C = A * B for n x n FP64 matrices in the CAPS layout (host pointers).
 * A and B are only read; C is overwritten.  pattern: one char per recursion
 * level from [bBdDhH] ('b' = batched level, 'd'/'h' = recursive level), or
 * NULL for the planner's choice.  opts may be NULL (FP64 mode).
 * multiply.h:13 / multiply.cpp:9-25. */
int caps_multiply(const double *A, const double *B, double *C,
                  const caps_desc *desc, const char *pattern,
                  const caps_opts *opts);

/* Same on device-resident buffers, enqueued on `stream` (a hipStream_t, NULL
 * = the default stream).  Returns after enqueueing; the caller synchronises.
 * Device workspace is owned by the library and cached per plan. */
int caps_multiply_device(const double *dA, const double *dB, double *dC,
                         const caps_desc *desc, const char *pattern,
                         const caps_opts *opts, void *stream);

/* Same as caps_multiply_device, for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops): it calls no
 * hipMalloc / hipFree.  No counterpart in the reference.  The pointer tables
 * and workspace live in a scratch array of CAPS_NOALLOC_SCRATCH_MIB MiB that
 * every device the library runs on holds from module load until the process
 * ends; a call whose plan does not fit returns CAPS_ENOMEM (e.g. FP64 n = 2048
 * 'bd' fits, residue mode n = 1024 'bb' fits).  No graph: the steps are
 * enqueued one by one on `stream`.  CAPS_MODE_AUTO is not accepted.  All calls
 * on one device share its scratch (it keeps the last call's contents; no step
 * reads scratch it has not written in the same call): calls that overlap on
 * one device, on different streams, race. */
#define CAPS_NOALLOC_SCRATCH_MIB 256
int caps_multiply_device_noalloc(const double *dA, const double *dB, double *dC,
                                 const caps_desc *desc, const char *pattern,
                                 const caps_opts *opts, void *stream);

/* `count` independent multiplies C_i = A_i * B_i of possibly different
 * sizes and depths (device pointers, CAPS layout each), patterns[i] or NULL
 * per item.  opts is NULL (defaults) or an array of `count` caps_opts: one
 * leaf mode (CAPS_MODE_AUTO is not accepted) and one modulus count for all
 * items, a cutoff per item, flags from opts[0].  Items with the same
 * descriptor, pattern and cutoff share batched launches; all launches of the
 * call go into one captured hipGraph replayed as one launch. */
int caps_multiply_batch_device(int count, const double *const *dA, const double *const *dB,
                               double *const *dC, const caps_desc *descs,
                               const char *const *patterns, const caps_opts *opts,
                               void *stream);

/* Column-major n x n <-> CAPS layout, one process
 * (distributeFrom1Proc / collectTo1Proc, communication.cpp:612-708). */
int caps_to_layout(const double *colmajor, double *caps, const caps_desc *desc);
int caps_from_layout(const double *caps, double *colmajor, const caps_desc *desc);
int caps_to_layout_device(const double *dcolmajor, double *dcaps,
                          const caps_desc *desc, void *stream);
int caps_from_layout_device(const double *dcaps, double *dcolmajor,
                            const caps_desc *desc, void *stream);

/* The reference's counters (counters.cpp:6-12, 33-47): additions counts one
 * per element of every add kernel plus m^3 per leaf; multiplications m^3
 * per leaf (multiply.cpp:1085-1086). */
/* Phase timers, in the reference's order (counters.h:13-18) plus the split
 * mode's own phases. */
enum {
  CAPS_TIMER_COMM = 0,      /* always 0 on one device                     */
  CAPS_TIMER_ADD = 1,       /* Winograd pre-additions and combinations    */
  CAPS_TIMER_MUL = 2,       /* leaf GEMM launches                         */
  CAPS_TIMER_REARRANGE = 3, /* bs-block <-> column-major leaf reorders    */
  CAPS_TIMER_COMM_SUMMA = 4,/* always 0 on one device                     */
  CAPS_TIMER_SPLIT = 5,     /* split mode: scaling and residue splitting  */
  CAPS_TIMER_RECON = 6,     /* split mode: CRT reconstruction             */
  CAPS_NUM_TIMERS = 8
};

typedef struct caps_counters {
  long long additions;
  long long multiplications;
  long long leaf_products;   /* number of base-case products               */
  double peak_workspace_mb;  /* device workspace of the last plan, MiB     */
  char pattern[256];         /* "BFS,DFS,..." like setExecutionType        */
  double timers_ms[CAPS_NUM_TIMERS]; /* accumulated under CAPS_FLAG_TIMERS */
  long long timed_launches[CAPS_NUM_TIMERS];
} caps_counters;

int caps_get_counters(caps_counters *out);
int caps_reset_counters(void);

/* Memory budget in doubles for the automatic BFS/DFS choice
 * (setAvailableMemory, memory.cpp:31-33).  With a NULL pattern a level is
 * batched ('b') only while the plan's device workspace fits the budget,
 * greedily from the top level, else recursive ('d'), as enoughMemory
 * (memory.cpp:22-29) gates the reference's BFS steps.  doubles > 0 sets the
 * budget (capped at the free device memory); 0 restores the default, half of
 * the free device memory (2^28 doubles, memory.cpp:4, when no device
 * answers). */
int caps_set_available_memory(long long doubles);

/* Release cached plans, graphs and workspace. */
int caps_release(void);

/* ------------------------------------------------------------------------
 * Autotuner (no counterpart in the reference): times candidate configurations
 * on the device and measures their normwise error max|C - C*| / max|C*|
 * against a high-precision product C* (CRT mode with 16 moduli); the best
 * configuration meeting an error target goes into a tuning table consulted by
 * calls with CAPS_MODE_AUTO.
 * ------------------------------------------------------------------------ */
typedef struct caps_candidate {
  int mode;          /* caps_mode                                           */
  int moduli;        /* split modes                                         */
  int cutoff;        /* caps_opts.cutoff                                    */
  int tile;          /* 0: default leaf-GEMM block tile, 1: the smaller one
                        (CAPS_FLAG_SMALL_TILE)                              */
  char pattern[32];  /* one char per level, NUL-terminated                  */
  double ms;         /* median milliseconds per multiply                    */
  double err;        /* normwise error vs the reference product (-1: none)  */
} caps_candidate;

/* Mode masks for the autotuner and planner */
#define CAPS_MASK_FP64 (1 << CAPS_MODE_FP64)
#define CAPS_MASK_OZAKI_INT8 (1 << CAPS_MODE_OZAKI_INT8)
#define CAPS_MASK_CRT_INT8 (1 << CAPS_MODE_CRT_INT8)
#define CAPS_MASK_CRT_BF16 (1 << CAPS_MODE_CRT_BF16)

/* High-precision product used as the error reference (device pointers). */
int caps_reference_product(const double *dA, const double *dB, double *dRef,
                           const caps_desc *desc, void *stream);
/* max|C - Ref| / max|Ref| over the n*n entries (device pointers, synchronises). */
int caps_normwise_error(const double *dC, const double *dRef, const caps_desc *desc,
                        double *err, void *stream);
/* Time one candidate (median of `reps` runs after one warm-up, HIP events on
 * `stream`) and, if dRef is not NULL, its error; fills cand->ms and ->err. */
int caps_measure(const double *dA, const double *dB, double *dC, const caps_desc *desc,
                 const double *dRef, int reps, caps_candidate *cand, void *stream);
/* Exhaustive search over the candidates allowed by mode_mask (every b/d
 * pattern over the recursed levels up to 4 levels, prefix patterns b^k d^l
 * beyond, cutoff, moduli 12..16 for split modes, and the leaf-GEMM block tile)
 * on the given operands; all measured candidates go to out[0..*n_out) (up to max_out),
 * the fastest one with err <= max_err to *best and to the tuning table. */
int caps_autotune(const double *dA, const double *dB, double *dC, const caps_desc *desc,
                  int mode_mask, double max_err, int reps, caps_candidate *out, int max_out,
                  int *n_out, caps_candidate *best, void *stream);
/* The leaf mode, moduli, cutoff and pattern that a call with pattern NULL
 * and these options would run (opts NULL: FP64 with the memory-budget
 * pattern rule above; CAPS_MODE_AUTO: the tuning table, then the learned
 * model, then the default), without multiplying; out->ms and out->err are the
 * learned model's predicted time and error when it made the choice, else -1.
 * Needs no device unless the learned model is consulted.  multiplyInternal's
 * automatic choice, multiply.cpp:43-66. */
int caps_plan_choice(const caps_desc *desc, const caps_opts *opts, caps_candidate *out);
/* The error target and allowed modes that CAPS_MODE_AUTO calls look up. */
int caps_planner_config(double max_err, int mode_mask, double exponent_spread);
int caps_tune_save(const char *path);
int caps_tune_load(const char *path);
/* The candidates caps_autotune would measure, and the 8 model features of a
 * candidate: log2 n, exponent spread, #b levels, #d levels, cutoff, moduli,
 * mode (0 FP64, 1 residue INT8, 2 leaf split, 3 residue BF16), and the
 * leaf's power-of-two alignment log2(gcd(n / 2^nrec, 256)) / 8. */
int caps_enumerate_candidates(const caps_desc *desc, int mode_mask, caps_candidate *out,
                              int max_out, int *n_out);
int caps_candidate_features(const caps_desc *desc, const caps_candidate *c, double spread,
                            float *features8);

/* ------------------------------------------------------------------------
 * Planner cost / error model (no counterpart in the reference, which picks
 * BFS or DFS from a memory rule, multiply.cpp:43-66): a 3-layer perceptron
 * 8 -> 64 -> 64 -> 2 (ReLU, FP32) whose forward, backward and Adam steps run
 * as HIP kernels.  Inputs and targets are the caller's normalised features
 * (row-major B x 8) and targets (B x 2) with a 0/1 mask per target.  The
 * parameter vector is W1 (64x8), b1, W2 (64x64), b2, W3 (2x64), b3.
 * ------------------------------------------------------------------------ */
typedef struct caps_mlp caps_mlp;

/* number of parameters (not a status) */
int caps_mlp_nparams(void);
/* He-uniform weights from a splitmix64 stream seeded with `seed`, zero biases */
int caps_mlp_create(caps_mlp **out, unsigned long long seed);
int caps_mlp_destroy(caps_mlp *m);
int caps_mlp_get_params(caps_mlp *m, float *out, int n);
/* also resets the optimiser state */
int caps_mlp_set_params(caps_mlp *m, const float *in, int n);
/* Y (B x 2) = model(X (B x 8)), host arrays */
int caps_mlp_forward(caps_mlp *m, const float *X, int B, float *Y);
/* gradient of mean_s sum_k M (Y - T)^2 with respect to the parameters, and
 * the loss itself (loss may be NULL) */
int caps_mlp_gradient(caps_mlp *m, const float *X, const float *T, const float *M, int B,
                      float *grad, float *loss);
/* Adam (lr, betas 0.9 / 0.999, eps 1e-8) over `epochs` seeded shuffles in
 * minibatches of `batch`; final_loss (may be NULL) over all B samples */
int caps_mlp_train(caps_mlp *m, const float *X, const float *T, const float *M, int B,
                   int epochs, int batch, float lr, unsigned long long seed,
                   float *final_loss);
/* Install m (NULL: remove) as the planner of CAPS_MODE_AUTO calls that have
 * no tuning-table entry.  norm: feature means[8], feature scales[8], target
 * means[2], target scales[2]; targets are log2(ms) and log10(error).  The
 * caller keeps ownership of m; caps_mlp_destroy(m) uninstalls it. */
int caps_planner_use_model(caps_mlp *m, const float *norm, int nnorm);

#ifdef __cplusplus
}
#endif

#endif /* CAPS_AMD_H */
