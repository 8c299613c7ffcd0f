// Plan builder and step runner.
//
// The recursion of multiplyInternal (multiply.cpp:28-108) is unrolled on the
// host into batched device launches:
//   'b' (strassenBFS, multiply.cpp:212-310): the seven T_i / S_i of every
//        sub-problem are formed in one launch and the level's batch grows
//        seven-fold; T1 = A11, T2 = A12, T7 = A22, S1 = B11, S2 = B21,
//        S6 = B22 are passed by pointer, as the reference passes them to
//        gatherMatrices1 (266-270).  After the recursion, calcCij combines
//        the seven products (307).
//   'd' / 'h' (strassenDFS, multiply.cpp:313-455, non-DAXPY): the seven
//        products run one after another over the whole batch, with the
//        reference's register reuse (C's quadrants plus two half-size
//        buffers R2, R5), so its workspace stays 2 h^2 per sub-problem.
//   leaf (block_multiply, multiply.cpp:1082-1141): one batched GEMM launch
//        over every sub-problem of the batch; for 1 < bs < m the leaves are
//        reordered to column-major around it (1093-1139).
// Workspace is a stack: each level allocates on entry and releases on exit,
// which is safe because every launch is ordered on one stream.
#include "engine.h"

#include <algorithm>
#include <cstring>

#include "ozaki.h"

namespace caps {

bool PlanKey::operator==(const PlanKey &o) const {
  return A == o.A && B == o.B && C == o.C && memcmp(&desc, &o.desc, sizeof(desc)) == 0 &&
         pattern == o.pattern && mode == o.mode && moduli == o.moduli;
}

namespace {

constexpr size_t ALIGN = 256;

struct Builder {
  Plan &plan;
  char *base;  // workspace base (nullptr while sizing)
  size_t top = 0, peak = 0;
  int bs;
  int mode;
  int moduli;
  std::string pattern;
  int top_nrec;

  size_t alloc(size_t bytes) {
    size_t off = top;
    top += (bytes + ALIGN - 1) / ALIGN * ALIGN;
    peak = std::max(peak, top);
    return off;
  }
  double *dptr(size_t off) { return reinterpret_cast<double *>(base + off); }

  size_t add_table(const void *data, size_t bytes) {
    auto &t = plan.host_tables;
    size_t off = (t.size() + 63) / 64 * 64;
    t.resize(off + bytes);
    memcpy(t.data() + off, data, bytes);
    return off;
  }

  void ew(int op, const std::vector<EwPtrs> &tab, long long q) {
    Step st{};
    st.kind = STEP_EW;
    st.op = op;
    st.count = (int)tab.size();
    st.n = q;
    // 16-byte accesses need every quadrant offset 16-byte aligned
    st.vec2 = (q % 2 == 0);
    for (auto &e : tab)
      for (int k = 0; k < 8; k++)
        if ((reinterpret_cast<uintptr_t>(e.in[k]) | reinterpret_cast<uintptr_t>(e.out[k])) & 15)
          st.vec2 = 0;
    st.table = add_table(tab.data(), tab.size() * sizeof(EwPtrs));
    plan.steps.push_back(st);
    plan.additions += (long long)ew_additions(op) * q * (long long)tab.size();
  }

  void level_tag(int depth, char c) {
    const char *t = (c == 'b' || c == 'B') ? "BFS" : (c == 'h' || c == 'H') ? "HYB" : "DFS";
    // setExecutionType (counters.cpp:49-59): one 4-char field per level
    size_t want = (size_t)(depth + 1) * 4;
    if (plan.exec_pattern.size() < want) plan.exec_pattern.resize(want, ' ');
    memcpy(&plan.exec_pattern[(size_t)depth * 4], t, 3);
    plan.exec_pattern[(size_t)depth * 4 + 3] = ',';
  }

  void leaf(const std::vector<Prob> &batch, int m) {
    const long long m3 = (long long)m * m * m;
    plan.additions += m3 * (long long)batch.size();
    plan.multiplications += m3 * (long long)batch.size();
    plan.leaves += (long long)batch.size();
    const size_t mark = top;
    std::vector<Prob> work = batch;
    const bool reorder = (bs != 1 && bs != m);
    const size_t msz = (size_t)m * m * sizeof(double);
    if (reorder) {
      size_t AA = alloc(msz * batch.size()), BB = alloc(msz * batch.size()),
             CC = alloc(msz * batch.size());
      std::vector<ReorderPtrs> ra, rb;
      for (size_t p = 0; p < batch.size(); p++) {
        work[p].A = dptr(AA + p * msz);
        work[p].B = dptr(BB + p * msz);
        work[p].C = dptr(CC + p * msz);
        ra.push_back({batch[p].A, const_cast<double *>(work[p].A)});
        rb.push_back({batch[p].B, const_cast<double *>(work[p].B)});
      }
      ra.insert(ra.end(), rb.begin(), rb.end());
      Step st{};
      st.kind = STEP_REORDER;
      st.op = 1;
      st.count = (int)ra.size();
      st.n = m;
      st.bs = bs;
      st.table = add_table(ra.data(), ra.size() * sizeof(ReorderPtrs));
      plan.steps.push_back(st);
    }
    if (mode == CAPS_MODE_FP64) {
      std::vector<GemmPtrs> g;
      for (auto &p : work) g.push_back({p.A, p.B, p.C});
      Step st{};
      st.kind = STEP_GEMM_F64;
      st.count = (int)g.size();
      st.n = m;
      st.table = add_table(g.data(), g.size() * sizeof(GemmPtrs));
      plan.steps.push_back(st);
    } else {
      ozaki_leaf_steps(work, m);
    }
    if (reorder) {
      std::vector<ReorderPtrs> rc;
      for (size_t p = 0; p < batch.size(); p++) rc.push_back({work[p].C, batch[p].C});
      Step st{};
      st.kind = STEP_REORDER;
      st.op = 0;
      st.count = (int)rc.size();
      st.n = m;
      st.bs = bs;
      st.table = add_table(rc.data(), rc.size() * sizeof(ReorderPtrs));
      plan.steps.push_back(st);
    }
    top = mark;
  }

  // ---------------------------------------------------------------- split mode
  // Leaf product of a batch through the Ozaki-II CRT scheme (ozaki.h).
  void ozaki_leaf_steps(const std::vector<Prob> &work, int m);

  void bfs(const std::vector<Prob> &batch, int s, int nrec, int depth) {
    const int h = s / 2;
    const long long q = (long long)h * h;
    const size_t qb = (size_t)q * sizeof(double);
    const size_t mark = top;
    const size_t nb = batch.size();
    size_t LF = alloc(4 * qb * nb), RF = alloc(4 * qb * nb), P = alloc(7 * qb * nb);
    std::vector<EwPtrs> pre(nb), post(nb);
    std::vector<Prob> next;
    next.reserve(7 * nb);
    for (size_t p = 0; p < nb; p++) {
      const double *A = batch[p].A, *B = batch[p].B;
      double *C = batch[p].C;
      double *lf = dptr(LF + 4 * qb * p), *rf = dptr(RF + 4 * qb * p), *pp = dptr(P + 7 * qb * p);
      EwPtrs &e = pre[p];
      for (int k = 0; k < 4; k++) {
        e.in[k] = A + k * q;          // A11 A21 A12 A22
        e.in[4 + k] = B + k * q;      // B11 B21 B12 B22
        e.out[k] = lf + k * q;        // T3 T4 T5 T6
        e.out[4 + k] = rf + k * q;    // S3 S4 S5 S7
      }
      const double *T[7] = {A, A + 2 * q, lf, lf + q, lf + 2 * q, lf + 3 * q, A + 3 * q};
      const double *S[7] = {B, B + q, rf, rf + q, rf + 2 * q, B + 3 * q, rf + 3 * q};
      for (int i = 0; i < 7; i++) next.push_back({T[i], S[i], pp + i * q});
      EwPtrs &o = post[p];
      for (int i = 0; i < 7; i++) o.in[i] = pp + i * q;
      o.in[7] = nullptr;
      for (int k = 0; k < 4; k++) o.out[k] = C + k * q;  // C11 C21 C12 C22
      for (int k = 4; k < 8; k++) o.out[k] = nullptr;
    }
    ew(EW_BFS_PRE, pre, q);
    rec(next, h, nrec - 1, depth + 1);
    ew(EW_BFS_POST, post, q);
    top = mark;
  }

  void dfs(const std::vector<Prob> &batch, int s, int nrec, int depth) {
    const int h = s / 2;
    const long long q = (long long)h * h;
    const size_t qb = (size_t)q * sizeof(double);
    const size_t mark = top;
    const size_t nb = batch.size();
    size_t R2o = alloc(qb * nb), R5o = alloc(qb * nb);
    struct Regs {
      const double *A11, *A21, *A12, *A22, *B11, *B21, *B12, *B22;
      double *R1, *R2, *R3, *R4, *R5, *R6;
    };
    std::vector<Regs> r(nb);
    for (size_t p = 0; p < nb; p++) {
      const double *A = batch[p].A, *B = batch[p].B;
      double *C = batch[p].C;
      r[p] = {A, A + q, A + 2 * q, A + 3 * q, B, B + q, B + 2 * q, B + 3 * q,
              C + q, dptr(R2o + qb * p), C, C + 3 * q, dptr(R5o + qb * p), C + 2 * q};
    }
    auto table = [&](auto fill) {
      std::vector<EwPtrs> t(nb);
      for (size_t p = 0; p < nb; p++) {
        EwPtrs e{};
        fill(r[p], e);
        t[p] = e;
      }
      return t;
    };
    auto sub = [&](auto pick) {
      std::vector<Prob> b(nb);
      for (size_t p = 0; p < nb; p++) b[p] = pick(r[p]);
      rec(b, h, nrec - 1, depth + 1);
    };
    // S5=B22-B12 (R1), S3=B12-B11 (R2), S4=B22-S3 (R3); T3=A21+A22 (R6), T5=A11-A21 (R4)
    ew(EW_DFS_PRE1, table([](const Regs &x, EwPtrs &e) {
         e.in[0] = x.A11; e.in[1] = x.A21; e.in[2] = x.A22;
         e.in[3] = x.B11; e.in[4] = x.B12; e.in[5] = x.B22;
         e.out[0] = x.R1; e.out[1] = x.R2; e.out[2] = x.R3; e.out[3] = x.R4; e.out[4] = x.R6;
       }), q);
    sub([](const Regs &x) { return Prob{x.R4, x.R1, x.R5}; });   // Q5 = T5 S5
    sub([](const Regs &x) { return Prob{x.R6, x.R2, x.R4}; });   // Q3 = T3 S3
    ew(EW_DFS_PRE2, table([](const Regs &x, EwPtrs &e) {       // T4 = T3 - A11
         e.in[0] = x.R6; e.in[1] = x.A11; e.out[0] = x.R6;
       }), q);
    sub([](const Regs &x) { return Prob{x.R6, x.R3, x.R2}; });   // Q4 = T4 S4
    ew(EW_DFS_PRE3, table([](const Regs &x, EwPtrs &e) {       // T6 = A12-T4, S7 = S4-B21
         e.in[0] = x.A12; e.in[1] = x.R6; e.in[2] = x.R3; e.in[3] = x.B21;
         e.out[0] = x.R6; e.out[1] = x.R3;
       }), q);
    sub([](const Regs &x) { return Prob{x.A22, x.R3, x.R1}; });  // Q7 = A22 S7
    sub([](const Regs &x) { return Prob{x.A11, x.B11, x.R3}; }); // Q1 = A11 B11
    ew(EW_DFS_POST1, table([](const Regs &x, EwPtrs &e) {      // U's, C22, C21
         e.in[0] = x.R3; e.in[1] = x.R2; e.in[2] = x.R5; e.in[3] = x.R4; e.in[4] = x.R1;
         e.out[0] = x.R2; e.out[1] = x.R4; e.out[2] = x.R1;
       }), q);
    sub([](const Regs &x) { return Prob{x.A12, x.B21, x.R5}; }); // Q2 = A12 B21
    ew(EW_DFS_POST2, table([](const Regs &x, EwPtrs &e) {      // C11 = Q1 + Q2
         e.in[0] = x.R3; e.in[1] = x.R5; e.out[0] = x.R3;
       }), q);
    sub([](const Regs &x) { return Prob{x.R6, x.B22, x.R5}; });  // Q6 = T6 B22
    ew(EW_DFS_POST3, table([](const Regs &x, EwPtrs &e) {      // C12 = U3 + Q6
         e.in[0] = x.R2; e.in[1] = x.R5; e.out[0] = x.R6;
       }), q);
    top = mark;
  }

  void rec(const std::vector<Prob> &batch, int s, int nrec, int depth) {
    if (nrec == 0) {
      leaf(batch, s);
      return;
    }
    char c = pattern[depth];
    level_tag(depth, c);
    if (c == 'b' || c == 'B')
      bfs(batch, s, nrec, depth);
    else
      dfs(batch, s, nrec, depth);
  }
};

}  // namespace

// Split-mode leaf: workspace and tables for the Ozaki-II steps (ozaki.h).
void Builder::ozaki_leaf_steps(const std::vector<Prob> &work, int m) {
  const int nmod = moduli;
  const size_t nb = work.size();
  const size_t plane = (size_t)m * m;  // bytes per INT8 plane
  if (plan.oz_moduli == 0) {
    OzakiConsts k;
    ozaki_make_consts(nmod, k);
    plan.oz_const = add_table(&k, sizeof(k));
    plan.oz_moduli = nmod;
  }
  const size_t ex_bytes = (size_t)m * sizeof(int);
  const size_t pa_bytes = (size_t)OZ_CHUNKS * m * 2 * sizeof(double);
  const size_t pb_bytes = (size_t)m * 2 * sizeof(double);
  const size_t mark = top;
  size_t EA = alloc(ex_bytes * nb), EB = alloc(ex_bytes * nb);
  size_t PA = alloc(pa_bytes * nb), PB = alloc(pb_bytes * nb);
  size_t RA = alloc(plane * nmod * nb), RB = alloc(plane * nmod * nb);
  size_t RC = alloc(plane * nmod * nb);
  std::vector<OzLeaf> t(nb);
  for (size_t p = 0; p < nb; p++) {
    OzLeaf &L = t[p];
    L.A = work[p].A;
    L.B = work[p].B;
    L.C = work[p].C;
    L.ea = reinterpret_cast<int *>(base + EA + ex_bytes * p);
    L.eb = reinterpret_cast<int *>(base + EB + ex_bytes * p);
    L.pa = reinterpret_cast<double *>(base + PA + pa_bytes * p);
    L.pb = reinterpret_cast<double *>(base + PB + pb_bytes * p);
    L.ra = reinterpret_cast<int8_t *>(base + RA + plane * nmod * p);
    L.rb = reinterpret_cast<int8_t *>(base + RB + plane * nmod * p);
    L.rc = reinterpret_cast<int8_t *>(base + RC + plane * nmod * p);
  }
  const size_t tab = add_table(t.data(), t.size() * sizeof(OzLeaf));
  for (int kind : {STEP_OZ_SCALE, STEP_OZ_SPLIT, STEP_OZ_GEMM, STEP_OZ_RECON}) {
    Step st{};
    st.kind = kind;
    st.count = (int)nb;
    st.n = m;
    st.op = nmod;
    st.table = tab;
    plan.steps.push_back(st);
  }
  top = mark;
}

int build_plan(Plan &plan, const double *A, const double *B, double *C, const caps_desc &d,
               const std::string &pattern, int mode, int moduli, void *work_base) {
  plan.steps.clear();
  plan.host_tables.clear();
  plan.additions = plan.multiplications = plan.leaves = 0;
  plan.exec_pattern.clear();
  plan.oz_moduli = 0;
  Builder b{plan, reinterpret_cast<char *>(work_base)};
  b.bs = d.bs;
  b.mode = mode;
  b.moduli = moduli;
  b.pattern = pattern;
  b.top_nrec = d.nrec;
  std::vector<Prob> top{{A, B, C}};
  b.rec(top, d.lda, d.nrec, 0);
  plan.work_bytes = b.peak;
  return CAPS_OK;
}

size_t plan_workspace_bytes(const caps_desc &d, const std::string &pattern, int mode,
                            int moduli) {
  Plan p;
  build_plan(p, nullptr, nullptr, nullptr, d, pattern, mode, moduli, nullptr);
  return p.work_bytes;
}

namespace {

hipError_t run_one(const Plan &plan, const Step &st, hipStream_t s) {
  const char *tb = reinterpret_cast<const char *>(plan.d_tables);
  {
    hipError_t e = hipSuccess;
    switch (st.kind) {
      case STEP_EW:
        e = launch_ew(st.op, reinterpret_cast<const EwPtrs *>(tb + st.table), st.count, st.n,
                      st.vec2 != 0, s);
        break;
      case STEP_GEMM_F64:
        e = launch_dgemm_batched(reinterpret_cast<const GemmPtrs *>(tb + st.table), st.count,
                                 (int)st.n, s);
        break;
      case STEP_REORDER:
        e = launch_leaf_reorder(reinterpret_cast<const ReorderPtrs *>(tb + st.table), st.count,
                                (int)st.n, st.bs, st.op == 1, s);
        break;
      case STEP_OZ_SCALE:
      case STEP_OZ_SPLIT:
      case STEP_OZ_GEMM:
      case STEP_OZ_RECON:
        e = launch_ozaki_step(st.kind - STEP_OZ_SCALE,
                              reinterpret_cast<const OzLeaf *>(tb + st.table), st.count,
                              (int)st.n, st.op,
                              reinterpret_cast<const OzakiConsts *>(tb + plan.oz_const), s);
        break;
      default:
        e = hipErrorInvalidValue;
    }
    return e;
  }
}

int timer_of(int kind) {
  switch (kind) {
    case STEP_EW: return CAPS_TIMER_ADD;
    case STEP_GEMM_F64:
    case STEP_OZ_GEMM: return CAPS_TIMER_MUL;
    case STEP_REORDER: return CAPS_TIMER_REARRANGE;
    case STEP_OZ_SCALE:
    case STEP_OZ_SPLIT: return CAPS_TIMER_SPLIT;
    case STEP_OZ_RECON: return CAPS_TIMER_RECON;
  }
  return CAPS_TIMER_COMM;
}

}  // namespace

hipError_t run_steps(const Plan &plan, hipStream_t s) {
  for (const Step &st : plan.steps) {
    hipError_t e = run_one(plan, st, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t run_steps_timed(const Plan &plan, hipStream_t s, double *ms, long long *n) {
  const size_t ns = plan.steps.size();
  std::vector<hipEvent_t> ev(ns + 1);
  hipError_t e = hipSuccess;
  size_t made = 0;
  for (; made < ev.size(); made++) {
    e = hipEventCreate(&ev[made]);
    if (e != hipSuccess) break;
  }
  if (e == hipSuccess) e = hipEventRecord(ev[0], s);
  for (size_t i = 0; i < ns && e == hipSuccess; i++) {
    e = run_one(plan, plan.steps[i], s);
    if (e == hipSuccess) e = hipEventRecord(ev[i + 1], s);
  }
  if (e == hipSuccess) e = hipEventSynchronize(ev[ns]);
  for (size_t i = 0; i < ns && e == hipSuccess; i++) {
    float t = 0.f;
    e = hipEventElapsedTime(&t, ev[i], ev[i + 1]);
    const int k = timer_of(plan.steps[i].kind);
    ms[k] += t;
    n[k] += 1;
  }
  for (size_t i = 0; i < made; i++) hipEventDestroy(ev[i]);
  return e;
}

}  // namespace caps
