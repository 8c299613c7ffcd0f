/*
 * CPU oracle for the CAPS Strassen-Winograd multiply path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library is
 * compared against.  It is loaded by tests/, by __graft_entry__.smoke() and by
 * the cpu_baseline leg of bench.py; nothing in caps_amd/ links, loads or calls
 * it, and the product never falls back to it.
 *
 * It restates, in plain C, the arithmetic of the one-process CAPS path:
 *   - the CAPS storage layout for nproc = nproc_summa = 1
 *       (matrix.h:20-44, communication.cpp:612-708)
 *   - multiply / multiplyInternal pattern dispatch (multiply.cpp:9-108)
 *   - strassenBFS arithmetic without the MPI redistribution
 *       (multiply.cpp:212-310, calcCij 195-210)
 *   - strassenDFS with its register reuse (multiply.cpp:313-455, non-DAXPY)
 *   - block_multiply with the bs-block <-> column-major reorder
 *       (multiply.cpp:1082-1141) and square_dgemm_zero (dgemm-blas.cpp:4-13)
 *   - the flop counters (counters.cpp:33-47, multiply.cpp:111 etc.)
 *   - POSIX drand48 (randombenchmark.cpp:19-22 fills with 2*drand48()-1)
 * plus two checkers with no reference counterpart:
 *   - a compensated classical GEMM (Ogita, Rump, Oishi, "Accurate sum and dot
 *     product", SIAM J. Sci. Comput. 26(6), 2005: Dot2 = TwoProduct via FMA
 *     followed by TwoSum accumulation)
 *   - an exact int64 classical GEMM for integer-valued inputs.
 *
 * How each function is pinned (see tests/test_oracle.py):
 *   layout      : closed-form index vs. a recursive placement, bijectivity,
 *                 round trip, ramp matrix
 *   multiply    : the reference's 110 one-process fixture rows
 *                 (tests/golden/test*.params), bit-exact; BFS and DFS orders
 *                 bit-identical to each other on real inputs; integer inputs
 *                 equal the exact int64 product
 *   leaf dgemm  : brute-force triple loop on tiny shapes, int64 exactness
 *   dot2 gemm   : exact on integers; against python fractions on small n
 *   drand48     : against the C library's drand48/srand48
 *
 * Leaf GEMM summation order: every C(i,j) is accumulated over k = 0..n-1 in
 * increasing order, starting from 0, with separate multiply and add roundings
 * (built with -ffp-contract=off), whatever the thread count or blocking.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_OK 0
#define ORACLE_EBAD_PATTERN 1
#define ORACLE_EBAD_DESC 2
#define ORACLE_ENOMEM 3

/* ------------------------------------------------------------------ */
/* counters (counters.cpp:6-47)                                        */
/* ------------------------------------------------------------------ */
static long long g_additions = 0;
static long long g_multiplications = 0;

void oracle_reset_counters(void) { g_additions = 0; g_multiplications = 0; }
long long oracle_additions(void) { return g_additions; }
long long oracle_multiplications(void) { return g_multiplications; }

static int g_threads = 1;

/* ------------------------------------------------------------------ */
/* layout (matrix.h:20-44; communication.cpp:612-708, nproc = 1)       */
/* ------------------------------------------------------------------ */

/* Offset of element (i,j) of an n x n matrix stored in the CAPS layout with
 * nrec recursion levels and block size bs, for one process.  Quadrants are in
 * the order 11, 21, 12, 22 (communication.cpp:646-653); inside a leaf of size
 * m the bs x bs blocks are column-major between blocks
 * (communication.cpp:630-635) and column-major within a block
 * (communication.cpp:563-566). */
long long oracle_caps_offset(int n, int nrec, int bs, int i, int j) {
  long long off = 0;
  long long s = n;
  for (int l = 0; l < nrec; l++) {
    long long h = s / 2;
    int bi = i >= h, bj = j >= h;
    off += (long long)(bi + 2 * bj) * h * h;
    i -= bi * (int)h;
    j -= bj * (int)h;
    s = h;
  }
  long long nb = s / bs;
  off += ((long long)(i / bs) + (long long)(j / bs) * nb) * bs * bs;
  off += (i % bs) + (long long)(j % bs) * bs;
  return off;
}

/* Recursive placement, restating distributeFrom1ProcRec for one process
 * (communication.cpp:612-655): quadrant recursion, then block copies. */
static void to_caps_rec(int lda, int nrec, int bs, double *O, const double *I,
                        int ldi) {
  if (nrec == 0) {
    int nb = lda / bs;
    for (int bj = 0; bj < nb; bj++)
      for (int bi = 0; bi < nb; bi++) {
        double *o = O + ((long long)bi + (long long)bj * nb) * bs * bs;
        const double *in = I + (long long)bi * bs + (long long)bj * bs * ldi;
        for (int c = 0; c < bs; c++)
          memcpy(o + (long long)c * bs, in + (long long)c * ldi,
                 (size_t)bs * sizeof(double));
      }
    return;
  }
  int h = lda / 2;
  long long q = (long long)h * h;
  to_caps_rec(h, nrec - 1, bs, O, I, ldi);
  to_caps_rec(h, nrec - 1, bs, O + q, I + h, ldi);
  to_caps_rec(h, nrec - 1, bs, O + 2 * q, I + (long long)h * ldi, ldi);
  to_caps_rec(h, nrec - 1, bs, O + 3 * q, I + (long long)h * ldi + h, ldi);
}

static void from_caps_rec(int lda, int nrec, int bs, double *O, const double *I,
                          int ldo) {
  if (nrec == 0) {
    int nb = lda / bs;
    for (int bj = 0; bj < nb; bj++)
      for (int bi = 0; bi < nb; bi++) {
        const double *in = I + ((long long)bi + (long long)bj * nb) * bs * bs;
        double *o = O + (long long)bi * bs + (long long)bj * bs * ldo;
        for (int c = 0; c < bs; c++)
          memcpy(o + (long long)c * ldo, in + (long long)c * bs,
                 (size_t)bs * sizeof(double));
      }
    return;
  }
  int h = lda / 2;
  long long q = (long long)h * h;
  from_caps_rec(h, nrec - 1, bs, O, I, ldo);
  from_caps_rec(h, nrec - 1, bs, O + h, I + q, ldo);
  from_caps_rec(h, nrec - 1, bs, O + (long long)h * ldo, I + 2 * q, ldo);
  from_caps_rec(h, nrec - 1, bs, O + (long long)h * ldo + h, I + 3 * q, ldo);
}

/* column-major n x n -> CAPS layout (distributeFrom1Proc, nproc = 1) */
void oracle_to_caps(int n, int nrec, int bs, const double *colmajor,
                    double *caps) {
  to_caps_rec(n, nrec, bs, caps, colmajor, n);
}

/* CAPS layout -> column-major n x n (collectTo1Proc, nproc = 1) */
void oracle_from_caps(int n, int nrec, int bs, const double *caps,
                      double *colmajor) {
  from_caps_rec(n, nrec, bs, colmajor, caps, n);
}

/* verifyDescriptor for one process (matrix.cpp:3-7) plus bs | leaf */
int oracle_valid_desc(int lda, int nrec, int bs) {
  if (lda <= 0 || nrec < 0 || bs <= 0 || nrec > 30) return 0;
  long long d = (1LL << nrec) * bs;
  return lda % d == 0;
}

/* ------------------------------------------------------------------ */
/* elementwise add kernels (multiply.cpp:110-210), same association    */
/* ------------------------------------------------------------------ */
typedef long long ll;

static void add2(ll n, double *C, const double *A, const double *B) {
  g_additions += n;
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (ll i = 0; i < n; i++) C[i] = A[i] + B[i];
}
static void sub2(ll n, double *C, const double *A, const double *B) {
  g_additions += n;
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (ll i = 0; i < n; i++) C[i] = A[i] - B[i];
}
/* tripleSubMatrices (127-138): outputs written in order per element */
static void sub3(ll n, double *T1, const double *S11, const double *S12,
                 double *T2, const double *S21, const double *S22, double *T3,
                 const double *S31, const double *S32) {
  g_additions += 3 * n;
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (ll i = 0; i < n; i++) {
    T1[i] = S11[i] - S12[i];
    T2[i] = S21[i] - S22[i];
    T3[i] = S31[i] - S32[i];
  }
}
/* tripleAddMatrices (140-151) */
static void add3(ll n, double *T1, const double *S11, const double *S12,
                 double *T2, const double *S21, const double *S22, double *T3,
                 const double *S31, const double *S32) {
  g_additions += 3 * n;
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (ll i = 0; i < n; i++) {
    T1[i] = S11[i] + S12[i];
    T2[i] = S21[i] + S22[i];
    T3[i] = S31[i] + S32[i];
  }
}
/* addSubMatrices (153-163) */
static void addsub(ll n, double *T1, const double *S11, const double *S12,
                   double *T2, const double *S21, const double *S22) {
  g_additions += 2 * n;
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (ll i = 0; i < n; i++) {
    T1[i] = S11[i] + S12[i];
    T2[i] = S21[i] - S22[i];
  }
}

/* ------------------------------------------------------------------ */
/* leaf GEMM (dgemm-blas.cpp:4-13): C = A*B, column-major, beta = 0     */
/* ------------------------------------------------------------------ */

/* Reference triple loop, k ascending.  Used for tiny shapes and to pin the
 * blocked version below. */
void oracle_dgemm_naive(int n, const double *A, const double *B, double *C) {
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      double c = 0.0;
      for (int k = 0; k < n; k++) c += A[i + (ll)k * n] * B[k + (ll)j * n];
      C[i + (ll)j * n] = c;
    }
}

/* Register-blocked version with the same per-element order (k ascending,
 * separate roundings), parallel over column blocks. */
#define MB 16
#define NBJ 4
void oracle_dgemm(int n, const double *A, const double *B, double *C) {
  int jb_count = (n + NBJ - 1) / NBJ;
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads)
  for (int jb = 0; jb < jb_count; jb++) {
    int j0 = jb * NBJ;
    int nj = n - j0 < NBJ ? n - j0 : NBJ;
    for (int i0 = 0; i0 < n; i0 += MB) {
      int ni = n - i0 < MB ? n - i0 : MB;
      if (ni == MB && nj == NBJ) {
        double acc[NBJ][MB];
        for (int jj = 0; jj < NBJ; jj++)
          for (int ii = 0; ii < MB; ii++) acc[jj][ii] = 0.0;
        for (int k = 0; k < n; k++) {
          const double *a = A + i0 + (ll)k * n;
          for (int jj = 0; jj < NBJ; jj++) {
            double b = B[k + (ll)(j0 + jj) * n];
            for (int ii = 0; ii < MB; ii++) acc[jj][ii] += a[ii] * b;
          }
        }
        for (int jj = 0; jj < NBJ; jj++)
          for (int ii = 0; ii < MB; ii++)
            C[i0 + ii + (ll)(j0 + jj) * n] = acc[jj][ii];
      } else {
        for (int jj = 0; jj < nj; jj++)
          for (int ii = 0; ii < ni; ii++) {
            double c = 0.0;
            for (int k = 0; k < n; k++)
              c += A[i0 + ii + (ll)k * n] * B[k + (ll)(j0 + jj) * n];
            C[i0 + ii + (ll)(j0 + jj) * n] = c;
          }
      }
    }
  }
}

/* Compensated classical GEMM (Dot2), column-major, n x n.  Every C(i,j) is
 * the Dot2 result of row i of A and column j of B, accurate as if computed
 * in twice the working precision then rounded (Ogita-Rump-Oishi Thm 5.3). */
void oracle_dgemm_dot2(int n, const double *A, const double *B, double *C) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads)
  for (int j = 0; j < n; j++) {
    const double *b = B + (ll)j * n;
    for (int i0 = 0; i0 < n; i0 += MB) {
      int ni = n - i0 < MB ? n - i0 : MB;
      double s[MB], c[MB];
      for (int ii = 0; ii < MB; ii++) { s[ii] = 0.0; c[ii] = 0.0; }
      for (int k = 0; k < n; k++) {
        const double *a = A + i0 + (ll)k * n;
        double bk = b[k];
        for (int ii = 0; ii < ni; ii++) {
          double p = a[ii] * bk;
          double ep = fma(a[ii], bk, -p);       /* TwoProduct */
          double x = s[ii] + p;                 /* TwoSum     */
          double z = x - s[ii];
          double es = (s[ii] - (x - z)) + (p - z);
          s[ii] = x;
          c[ii] += ep + es;
        }
      }
      for (int ii = 0; ii < ni; ii++) C[i0 + ii + (ll)j * n] = s[ii] + c[ii];
    }
  }
}

/* Dot2 for a subset of entries: rows[r], cols[r] -> out[r].  Used to check
 * large products on a sample of entries. */
void oracle_dot2_entries(int n, const double *A, const double *B, int count,
                         const int *rows, const int *cols, double *out) {
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int r = 0; r < count; r++) {
    int i = rows[r], j = cols[r];
    double s = 0.0, c = 0.0;
    for (int k = 0; k < n; k++) {
      double a = A[i + (ll)k * n], bk = B[k + (ll)j * n];
      double p = a * bk;
      double ep = fma(a, bk, -p);
      double x = s + p;
      double z = x - s;
      double es = (s - (x - z)) + (p - z);
      s = x;
      c += ep + es;
    }
    out[r] = s + c;
  }
}

/* Exact classical product of integer-valued matrices (|entries| < 2^31,
 * partial sums < 2^63), column-major. */
void oracle_igemm_exact(int n, const double *A, const double *B, double *C) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads)
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      long long acc = 0;
      for (int k = 0; k < n; k++)
        acc += (long long)A[i + (ll)k * n] * (long long)B[k + (ll)j * n];
      C[i + (ll)j * n] = (double)acc;
    }
}

/* ------------------------------------------------------------------ */
/* recursion (multiply.cpp:9-108, 212-455, 1082-1141)                  */
/* ------------------------------------------------------------------ */
typedef struct {
  int lda, nrec, bs;
} odesc;

static const char *g_pattern; /* one char per level, like multiply.cpp:7 */
static int g_depth;
static int g_err;
static double *g_work; /* 3*m*m reorder space (multiply.cpp:14-16) */

static void mult_internal(const double *A, const double *B, double *C,
                          odesc d);

/* block_multiply (multiply.cpp:1082-1141) */
static void block_multiply(const double *A, const double *B, double *C,
                           odesc d) {
  ll m = d.lda;
  g_additions += m * m * m;
  g_multiplications += m * m * m;
  int nb = d.lda / d.bs;
  if (d.bs != 1 && d.bs != d.lda) {
    double *AA = g_work, *BB = g_work + m * m, *CC = g_work + 2 * m * m;
    for (int col = 0; col < d.lda; col++) {
      int cb = col / d.bs, cr = col - cb * d.bs;
      for (int rb = 0; rb < nb; rb++) {
        memcpy(AA + col * m + (ll)rb * d.bs,
               A + ((ll)rb + (ll)cb * nb) * d.bs * d.bs + (ll)cr * d.bs,
               (size_t)d.bs * sizeof(double));
        memcpy(BB + col * m + (ll)rb * d.bs,
               B + ((ll)rb + (ll)cb * nb) * d.bs * d.bs + (ll)cr * d.bs,
               (size_t)d.bs * sizeof(double));
      }
    }
    oracle_dgemm(d.lda, AA, BB, CC);
    for (int col = 0; col < d.lda; col++) {
      int cb = col / d.bs, cr = col - cb * d.bs;
      for (int rb = 0; rb < nb; rb++)
        memcpy(C + ((ll)rb + (ll)cb * nb) * d.bs * d.bs + (ll)cr * d.bs,
               CC + col * m + (ll)rb * d.bs, (size_t)d.bs * sizeof(double));
    }
  } else {
    oracle_dgemm(d.lda, A, B, C);
  }
}

/* strassenBFS arithmetic for one process (multiply.cpp:212-310): all seven
 * T_i / S_i formed first (272, 277), seven recursive products into one
 * contiguous Q buffer (289-299), then calcCij (195-210, 307). */
static void strassen_bfs(const double *A, const double *B, double *C,
                         odesc d) {
  odesc h = d;
  h.lda /= 2;
  h.nrec -= 1;
  ll q = (ll)h.lda * h.lda;
  const double *A11 = A, *A21 = A + q, *A12 = A + 2 * q, *A22 = A + 3 * q;
  const double *B11 = B, *B21 = B + q, *B12 = B + 2 * q, *B22 = B + 3 * q;
  double *C11 = C, *C21 = C + q, *C12 = C + 2 * q, *C22 = C + 3 * q;
  double *LF = malloc(sizeof(double) * 4 * q);
  double *RF = malloc(sizeof(double) * 4 * q);
  double *Q = malloc(sizeof(double) * 7 * q);
  if (!LF || !RF || !Q) {
    free(LF); free(RF); free(Q);
    g_err = ORACLE_ENOMEM;
    return;
  }
  double *T3 = LF, *T4 = LF + q, *T5 = LF + 2 * q, *T6 = LF + 3 * q;
  double *S3 = RF, *S4 = RF + q, *S5 = RF + 2 * q, *S7 = RF + 3 * q;
  /* addSubSubSubMatrices(T3 = A21+A22, T4 = T3-A11, T5 = A11-A21,
   *                      T6 = A12-T4)                       (multiply.cpp:272) */
  g_additions += 4 * q;
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (ll i = 0; i < q; i++) {
    T3[i] = A21[i] + A22[i];
    T4[i] = T3[i] - A11[i];
    T5[i] = A11[i] - A21[i];
    T6[i] = A12[i] - T4[i];
  }
  /* quadrupleSubMatrices(S5 = B22-B12, S3 = B12-B11, S4 = B22-S3,
   *                      S7 = S4-B21)                       (multiply.cpp:277) */
  g_additions += 4 * q;
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (ll i = 0; i < q; i++) {
    S5[i] = B22[i] - B12[i];
    S3[i] = B12[i] - B11[i];
    S4[i] = B22[i] - S3[i];
    S7[i] = S4[i] - B21[i];
  }
  const double *T[7] = {A11, A12, T3, T4, T5, T6, A22};
  const double *S[7] = {B11, B21, S3, S4, S5, B22, S7};
  for (int p = 0; p < 7 && !g_err; p++) mult_internal(T[p], S[p], Q + p * q, h);
  if (!g_err) {
    const double *Q1 = Q, *Q2 = Q + q, *Q3 = Q + 2 * q, *Q4 = Q + 3 * q,
                 *Q5 = Q + 4 * q, *Q6 = Q + 5 * q, *Q7 = Q + 6 * q;
    g_additions += 7 * q;
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (ll j = 0; j < q; j++) {
      C11[j] = Q1[j] + Q2[j];
      double U1 = Q1[j] + Q4[j];
      double U2 = U1 + Q5[j];
      C21[j] = U2 - Q7[j];
      C22[j] = U2 + Q3[j];
      double U3 = U1 + Q3[j];
      C12[j] = U3 + Q6[j];
    }
  }
  free(LF);
  free(RF);
  free(Q);
}

/* strassenDFS, non-DAXPY branch (multiply.cpp:409-454): C's quadrants are
 * used as scratch registers R1 = C21, R3 = C11, R4 = C22, R6 = C12 plus two
 * extra half-size buffers R2, R5. */
static void strassen_dfs(const double *A, const double *B, double *C,
                         odesc d) {
  odesc h = d;
  h.lda /= 2;
  h.nrec -= 1;
  ll q = (ll)h.lda * h.lda;
  const double *A11 = A, *A21 = A + q, *A12 = A + 2 * q, *A22 = A + 3 * q;
  const double *B11 = B, *B21 = B + q, *B12 = B + 2 * q, *B22 = B + 3 * q;
  double *C11 = C, *C21 = C + q, *C12 = C + 2 * q, *C22 = C + 3 * q;
  double *R1 = C21, *R3 = C11, *R4 = C22, *R6 = C12;
  double *R2 = malloc(sizeof(double) * q);
  double *R5 = malloc(sizeof(double) * q);
  if (!R2 || !R5) {
    free(R2); free(R5);
    g_err = ORACLE_ENOMEM;
    return;
  }
  double *S5 = R1, *S3 = R2, *S4 = R3;
  sub3(q, S5, B22, B12, S3, B12, B11, S4, B22, S3);
  double *T5 = R4, *T3 = R6;
  addsub(q, T3, A21, A22, T5, A11, A21);
  double *Q5 = R5;
  mult_internal(T5, S5, Q5, h);
  double *Q3 = R4;
  if (!g_err) mult_internal(T3, S3, Q3, h);
  double *T4 = R6;
  sub2(q, T4, T3, A11);
  double *Q4 = R2;
  if (!g_err) mult_internal(T4, S4, Q4, h);
  double *T6 = R6;
  sub2(q, T6, A12, T4);
  double *S7 = R3;
  sub2(q, S7, S4, B21);
  double *Q7 = R1;
  if (!g_err) mult_internal(A22, S7, Q7, h);
  double *Q1 = R3;
  if (!g_err) mult_internal(A11, B11, Q1, h);
  double *U1 = R2, *U2 = R5, *U3 = R2;
  add3(q, U1, Q1, Q4, U2, U1, Q5, U3, U1, Q3);
  addsub(q, C22, U2, Q3, C21, U2, Q7);
  double *Q2 = R5;
  if (!g_err) mult_internal(A12, B21, Q2, h);
  add2(q, C11, Q1, Q2);
  double *Q6 = R5;
  if (!g_err) mult_internal(T6, B22, Q6, h);
  add2(q, C12, U3, Q6);
  free(R5);
  free(R2);
}

/* multiplyInternal for nproc = nproc_summa = 1 (multiply.cpp:28-108).  With
 * no pattern, the one-process rule is DFS at every level (43-66: BFS needs
 * nproc > 1 or 7^nrec == nproc).  'h' with nproc_summa == 1 is DFS (91-92).
 * 'b' is the BFS arithmetic (the one-GPU meaning of a BFS step). */
static void mult_internal(const double *A, const double *B, double *C,
                          odesc d) {
  if (g_err) return;
  if (d.nrec == 0) {
    block_multiply(A, B, C, d);
    return;
  }
  char c = g_pattern ? g_pattern[g_depth] : 'd';
  g_depth++;
  if (c == 'b' || c == 'B')
    strassen_bfs(A, B, C, d);
  else if (c == 'd' || c == 'D' || c == 'h' || c == 'H')
    strassen_dfs(A, B, C, d);
  else
    g_err = ORACLE_EBAD_PATTERN; /* multiply.cpp:100-104 */
  g_depth--;
}

/* multiply (multiply.cpp:9-25) for one process, CAPS layout in and out. */
int oracle_multiply(const double *A, const double *B, double *C, int lda,
                    int nrec, int bs, const char *pattern, int threads) {
  if (!oracle_valid_desc(lda, nrec, bs)) return ORACLE_EBAD_DESC;
  g_threads = threads > 0 ? threads : 1;
  odesc d = {lda, nrec, bs};
  ll m = lda >> nrec;
  g_work = NULL;
  if (bs != 1) {
    g_work = malloc(sizeof(double) * 3 * m * m);
    if (!g_work) return ORACLE_ENOMEM;
  }
  g_pattern = pattern;
  g_depth = 0;
  g_err = 0;
  mult_internal(A, B, C, d);
  free(g_work);
  g_work = NULL;
  return g_err;
}

void oracle_set_threads(int threads) { g_threads = threads > 0 ? threads : 1; }

/* ------------------------------------------------------------------ */
/* POSIX drand48: X' = (0x5DEECE66D X + 0xB) mod 2^48, result X'/2^48; */
/* default state 0x1234ABCD330E, srand48(s) = (s << 16) | 0x330E.       */
/* ------------------------------------------------------------------ */
uint64_t oracle_drand48_default_state(void) { return 0x1234ABCD330EULL; }
uint64_t oracle_srand48_state(long seed) {
  return (((uint64_t)(uint32_t)seed) << 16) | 0x330EULL;
}

/* fill p[0..count) with 2*drand48()-1 (randombenchmark.cpp:19-22); returns
 * the advanced state through *state. */
void oracle_fill_uniform(double *p, long long count, uint64_t *state) {
  uint64_t x = *state;
  const uint64_t mask = (1ULL << 48) - 1;
  for (long long i = 0; i < count; i++) {
    x = (0x5DEECE66DULL * x + 0xBULL) & mask;
    p[i] = 2.0 * ldexp((double)x, -48) - 1.0;
  }
  *state = x;
}

/* raw drand48 values (no 2x-1), for checking against the C library */
void oracle_drand48_seq(double *p, long long count, uint64_t *state) {
  uint64_t x = *state;
  const uint64_t mask = (1ULL << 48) - 1;
  for (long long i = 0; i < count; i++) {
    x = (0x5DEECE66DULL * x + 0xBULL) & mask;
    p[i] = ldexp((double)x, -48);
  }
  *state = x;
}
