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
// The same recursion runs on FP64 matrices (CAPS_MODE_FP64, and
// CAPS_MODE_OZAKI_INT8 whose leaves are split products) or on INT8 residue
// planes (CAPS_MODE_CRT_INT8, crt.h), where a matrix is nmod planes.
// Workspace is a stack: each level allocates on entry and releases on exit,
// which is safe because every launch is ordered on one stream.
#include "engine.h"

#include <algorithm>
#include <cstring>

#include "crt.h"
#include "fused2.h"
#include "ozaki.h"

namespace caps {

bool PlanKey::operator==(const PlanKey &o) const {
  return A == o.A && B == o.B && C == o.C && memcmp(&desc, &o.desc, sizeof(desc)) == 0 &&
         pattern == o.pattern && mode == o.mode && moduli == o.moduli && cutoff == o.cutoff &&
         fuse == o.fuse && batch == o.batch && stream == o.stream && tile == o.tile;
}

namespace {

constexpr size_t ALIGN = 256;

struct Builder {
  Plan &plan;
  char *base;  // workspace base (nullptr while sizing)
  size_t top = 0, peak = 0;
  int bs = 1;
  int mode = CAPS_MODE_FP64;
  int moduli = 0;
  int cutoff = 0;  // layout levels below the Strassen recursion
  std::string pattern;
  bool fuse = true;      // fuse two leading 'b' levels when the sizes allow
  int tile = 0;          // leaf-GEMM block tile (0 default, 1 smaller)
  bool residue = false;  // matrices are nmod INT8 planes
  int es = 8;            // bytes per element: 8 (FP64) or 1 (residue)
  int planes = 1;        // 1 or nmod
  bool broken = false;   // an internal layout invariant failed (plan unusable)
  std::vector<Step> *out = nullptr;  // where steps go (default: plan.steps)

  std::vector<Step> &steps() { return out ? *out : plan.steps; }

  size_t alloc(size_t bytes) {
    size_t off = top;
    top += (bytes + ALIGN - 1) / ALIGN * ALIGN;
    peak = std::max(peak, top);
    return off;
  }

  // `count` matrices of `elems` elements each, one block per plane:
  // matrix p at base + p * elems * es, plane stride count * elems * es
  Mat alloc_mats(size_t count, long long elems) {
    const size_t per_plane = count * (size_t)elems * es;
    size_t off = alloc(per_plane * planes);
    return {base + off, (long long)per_plane};
  }
  Mat at(Mat M, long long elems) const { return {M.p + elems * es, M.ps}; }

  size_t add_table(const void *data, size_t bytes) {
    auto &t = plan.host_tables;
    size_t off = (t.size() + 63) / 64 * 64;
    t.resize(off + bytes);
    memcpy(t.data() + off, data, bytes);
    return off;
  }

  void ensure_consts() {
    if (plan.oz_moduli == 0) {
      OzakiConsts k;
      if (!ozaki_make_consts(moduli, k)) {  // modulus count outside 2..16
        broken = true;
        return;
      }
      plan.oz_const = add_table(&k, sizeof(k));
      plan.oz_moduli = moduli;
    }
  }

  void push(int kind, int op, int count, long long n, int bs_, size_t table, int vec = 0,
            int levels = 0) {
    Step st{};
    st.levels = levels;
    st.tile = tile;
    st.kind = kind;
    st.op = op;
    st.count = count;
    st.n = n;
    st.bs = bs_;
    st.vec2 = vec;
    st.table = table;
    steps().push_back(st);
  }

  // one Winograd add step over the batch: ins[p][k], outs[p][k]
  void ew(int op, const std::vector<std::vector<Mat>> &ins,
          const std::vector<std::vector<Mat>> &outs, long long q) {
    const int count = (int)ins.size();
    if (!residue) {
      std::vector<EwPtrs> t(ins.size());
      bool aligned = (q % 2 == 0);  // 16-byte accesses
      for (size_t p = 0; p < ins.size(); p++) {
        EwPtrs e{};
        for (size_t k = 0; k < ins[p].size(); k++)
          e.in[k] = reinterpret_cast<const double *>(ins[p][k].p);
        for (size_t k = 0; k < outs[p].size(); k++)
          e.out[k] = reinterpret_cast<double *>(outs[p][k].p);
        for (int k = 0; k < 8; k++)
          if ((reinterpret_cast<uintptr_t>(e.in[k]) | reinterpret_cast<uintptr_t>(e.out[k])) & 15)
            aligned = false;
        t[p] = e;
      }
      push(STEP_EW, op, count, q, 0, add_table(t.data(), t.size() * sizeof(EwPtrs)), aligned);
    } else {
      ensure_consts();
      std::vector<RPtrs> t(ins.size());
      bool aligned = (q % 4 == 0);  // 4-byte accesses
      for (size_t p = 0; p < ins.size(); p++) {
        RPtrs e{};
        for (size_t k = 0; k < ins[p].size(); k++) {
          e.in[k] = reinterpret_cast<const int8_t *>(ins[p][k].p);
          e.sin[k] = ins[p][k].ps;
        }
        for (size_t k = 0; k < outs[p].size(); k++) {
          e.out[k] = reinterpret_cast<int8_t *>(outs[p][k].p);
          e.sout[k] = outs[p][k].ps;
        }
        for (int k = 0; k < 8; k++)
          if (((reinterpret_cast<uintptr_t>(e.in[k]) | reinterpret_cast<uintptr_t>(e.out[k])) & 3) ||
              ((e.sin[k] | e.sout[k]) & 3))
            aligned = false;
        t[p] = e;
      }
      push(STEP_REW, op, count, q, 0, add_table(t.data(), t.size() * sizeof(RPtrs)), aligned);
    }
    plan.additions += (long long)ew_additions(op) * q * (long long)ins.size();
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
    if (residue) {
      // residue leaves are already row-major (A) / column-major (B, C)
      ensure_consts();
      std::vector<RGemm> g(batch.size());
      const bool segmented = residue_gemm_segmented(m, cutoff, tile, mode == CAPS_MODE_CRT_BF16);
      for (size_t p = 0; p < batch.size(); p++) {
        g[p] = {reinterpret_cast<const int8_t *>(batch[p].A.p),
                reinterpret_cast<const int8_t *>(batch[p].B.p),
                reinterpret_cast<int8_t *>(batch[p].C.p), batch[p].A.ps, batch[p].B.ps,
                batch[p].C.ps, batch[p].cseg};
        if (batch[p].cseg != 128 && !segmented) broken = true;
      }
      // bs field: 1 = BF16 MFMA leaf products
      push(STEP_OZ_GEMM, moduli, (int)g.size(), m, mode == CAPS_MODE_CRT_BF16 ? 1 : 0,
           add_table(g.data(), g.size() * sizeof(RGemm)), 0, cutoff);
      return;
    }
    const size_t mark = top;
    std::vector<Prob> work = batch;
    // bs-blocked leaves of the layout (size m >> cutoff) need a reorder
    const bool reorder = (bs != 1 && bs != (m >> cutoff));
    const long long mm = (long long)m * m;
    if (reorder) {
      Mat AA = alloc_mats(batch.size(), mm), BB = alloc_mats(batch.size(), mm),
          CC = alloc_mats(batch.size(), mm);
      std::vector<ReorderPtrs> rr;
      for (size_t p = 0; p < batch.size(); p++) {
        work[p].A = at(AA, (long long)p * mm);
        work[p].B = at(BB, (long long)p * mm);
        work[p].C = at(CC, (long long)p * mm);
        rr.push_back({reinterpret_cast<const double *>(batch[p].A.p),
                      reinterpret_cast<double *>(work[p].A.p)});
      }
      for (size_t p = 0; p < batch.size(); p++)
        rr.push_back({reinterpret_cast<const double *>(batch[p].B.p),
                      reinterpret_cast<double *>(work[p].B.p)});
      push(STEP_REORDER, 1, (int)rr.size(), m, bs,
           add_table(rr.data(), rr.size() * sizeof(ReorderPtrs)));
    }
    if (mode == CAPS_MODE_FP64) {
      std::vector<GemmPtrs> g;
      for (auto &p : work)
        g.push_back({reinterpret_cast<const double *>(p.A.p),
                     reinterpret_cast<const double *>(p.B.p), reinterpret_cast<double *>(p.C.p)});
      push(STEP_GEMM_F64, 0, (int)g.size(), m, 0, add_table(g.data(), g.size() * sizeof(GemmPtrs)),
           0, cutoff);
    } else {
      ozaki_leaf_steps(work, m);
    }
    if (reorder) {
      std::vector<ReorderPtrs> rc;
      for (size_t p = 0; p < batch.size(); p++)
        rc.push_back({reinterpret_cast<const double *>(work[p].C.p),
                      reinterpret_cast<double *>(batch[p].C.p)});
      push(STEP_REORDER, 0, (int)rc.size(), m, bs,
           add_table(rc.data(), rc.size() * sizeof(ReorderPtrs)));
    }
    top = mark;
  }

  // Leaf-only split (CAPS_MODE_OZAKI_INT8): FP64 leaf product through the
  // Ozaki-II CRT scheme (ozaki.h).
  void ozaki_leaf_steps(const std::vector<Prob> &work, int m) {
    const int nmod = moduli;
    const size_t nb = work.size();
    const size_t plane = (size_t)m * m;  // bytes per INT8 plane
    ensure_consts();
    const size_t ex_bytes = (size_t)m * sizeof(int);
    const size_t pa_bytes = (size_t)OZ_CHUNKS * m * 2 * sizeof(double);
    const size_t pb_bytes = (size_t)m * 2 * sizeof(double);
    const size_t mark = top;
    size_t EA = alloc(ex_bytes * nb), EB = alloc(ex_bytes * nb);
    size_t PA = alloc(pa_bytes * nb), PB = alloc(pb_bytes * nb);
    size_t RA = alloc(plane * nmod * nb), RB = alloc(plane * nmod * nb);
    size_t RC = alloc(plane * nmod * nb);
    std::vector<OzLeaf> t(nb);
    std::vector<RGemm> g(nb);
    for (size_t p = 0; p < nb; p++) {
      OzLeaf &L = t[p];
      L.A = reinterpret_cast<const double *>(work[p].A.p);
      L.B = reinterpret_cast<const double *>(work[p].B.p);
      L.C = reinterpret_cast<double *>(work[p].C.p);
      L.ea = reinterpret_cast<int *>(base + EA + ex_bytes * p);
      L.eb = reinterpret_cast<int *>(base + EB + ex_bytes * p);
      L.pa = reinterpret_cast<double *>(base + PA + pa_bytes * p);
      L.pb = reinterpret_cast<double *>(base + PB + pb_bytes * p);
      L.ra = reinterpret_cast<int8_t *>(base + RA + plane * nmod * p);
      L.rb = reinterpret_cast<int8_t *>(base + RB + plane * nmod * p);
      L.rc = reinterpret_cast<int8_t *>(base + RC + plane * nmod * p);
      g[p] = {L.ra, L.rb, L.rc, (long long)plane, (long long)plane, (long long)plane, 128};
    }
    const size_t tab = add_table(t.data(), t.size() * sizeof(OzLeaf));
    const size_t gtab = add_table(g.data(), g.size() * sizeof(RGemm));
    push(STEP_OZ_SCALE, nmod, (int)nb, m, 0, tab);
    push(STEP_OZ_SPLIT, nmod, (int)nb, m, 0, tab);
    push(STEP_OZ_GEMM, nmod, (int)nb, m, 0, gtab);
    push(STEP_OZ_RECON, nmod, (int)nb, m, 0, tab);
    top = mark;
  }

  void bfs(const std::vector<Prob> &batch, int s, int nrec, int depth) {
    const int h = s / 2;
    const long long q = (long long)h * h;
    const size_t mark = top;
    const size_t nb = batch.size();
    Mat LF = alloc_mats(4 * nb, q), RF = alloc_mats(4 * nb, q), P = alloc_mats(7 * nb, q);
    std::vector<std::vector<Mat>> pre_in(nb), pre_out(nb), post_in(nb), post_out(nb);
    std::vector<Prob> next;
    next.reserve(7 * nb);
    for (size_t p = 0; p < nb; p++) {
      const Mat A = batch[p].A, B = batch[p].B, C = batch[p].C;
      const Mat lf = at(LF, 4 * q * (long long)p), rf = at(RF, 4 * q * (long long)p),
                pp = at(P, 7 * q * (long long)p);
      for (int k = 0; k < 4; k++) pre_in[p].push_back(at(A, k * q));    // A11 A21 A12 A22
      for (int k = 0; k < 4; k++) pre_in[p].push_back(at(B, k * q));    // B11 B21 B12 B22
      for (int k = 0; k < 4; k++) pre_out[p].push_back(at(lf, k * q));  // T3 T4 T5 T6
      for (int k = 0; k < 4; k++) pre_out[p].push_back(at(rf, k * q));  // S3 S4 S5 S7
      const Mat T[7] = {A, at(A, 2 * q), lf, at(lf, q), at(lf, 2 * q), at(lf, 3 * q), at(A, 3 * q)};
      const Mat S[7] = {B, at(B, q), rf, at(rf, q), at(rf, 2 * q), at(B, 3 * q), at(rf, 3 * q)};
      for (int i = 0; i < 7; i++) next.push_back({T[i], S[i], at(pp, i * q)});
      for (int i = 0; i < 7; i++) post_in[p].push_back(at(pp, i * q));
      for (int k = 0; k < 4; k++) post_out[p].push_back(at(C, k * q));  // C11 C21 C12 C22
    }
    ew(EW_BFS_PRE, pre_in, pre_out, q);
    rec(next, h, nrec - 1, depth + 1);
    ew(EW_BFS_POST, post_in, post_out, q);
    top = mark;
  }

  void dfs(const std::vector<Prob> &batch, int s, int nrec, int depth) {
    const int h = s / 2;
    const long long q = (long long)h * h;
    const size_t mark = top;
    const size_t nb = batch.size();
    Mat R2o = alloc_mats(nb, q), R5o = alloc_mats(nb, q);
    struct Regs {
      Mat A11, A21, A12, A22, B11, B21, B12, B22;
      Mat R1, R2, R3, R4, R5, R6;
    };
    std::vector<Regs> r(nb);
    for (size_t p = 0; p < nb; p++) {
      const Mat A = batch[p].A, B = batch[p].B, C = batch[p].C;
      r[p] = {A,        at(A, q),  at(A, 2 * q), at(A, 3 * q),
              B,        at(B, q),  at(B, 2 * q), at(B, 3 * q),
              at(C, q), at(R2o, q * (long long)p), C, at(C, 3 * q),
              at(R5o, q * (long long)p), at(C, 2 * q)};
    }
    using V = std::vector<Mat>;
    auto step = [&](int op, V (*pick_in)(const Regs &), V (*pick_out)(const Regs &)) {
      std::vector<V> ins(nb), outs(nb);
      for (size_t p = 0; p < nb; p++) {
        ins[p] = pick_in(r[p]);
        outs[p] = pick_out(r[p]);
      }
      ew(op, ins, outs, q);
    };
    auto sub = [&](Prob (*pick)(const Regs &)) {
      std::vector<Prob> b(nb);
      for (size_t p = 0; p < nb; p++) b[p] = pick(r[p]);
      rec(b, h, nrec - 1, depth + 1);
    };
    // S5=B22-B12 (R1), S3=B12-B11 (R2), S4=B22-S3 (R3); T5=A11-A21 (R4), T3=A21+A22 (R6)
    step(EW_DFS_PRE1, [](const Regs &x) { return V{x.A11, x.A21, x.A22, x.B11, x.B12, x.B22}; },
         [](const Regs &x) { return V{x.R1, x.R2, x.R3, x.R4, x.R6}; });
    sub([](const Regs &x) { return Prob{x.R4, x.R1, x.R5}; });   // Q5 = T5 S5
    sub([](const Regs &x) { return Prob{x.R6, x.R2, x.R4}; });   // Q3 = T3 S3
    step(EW_DFS_PRE2, [](const Regs &x) { return V{x.R6, x.A11}; },  // T4 = T3 - A11
         [](const Regs &x) { return V{x.R6}; });
    sub([](const Regs &x) { return Prob{x.R6, x.R3, x.R2}; });   // Q4 = T4 S4
    step(EW_DFS_PRE3, [](const Regs &x) { return V{x.A12, x.R6, x.R3, x.B21}; },  // T6, S7
         [](const Regs &x) { return V{x.R6, x.R3}; });
    sub([](const Regs &x) { return Prob{x.A22, x.R3, x.R1}; });  // Q7 = A22 S7
    sub([](const Regs &x) { return Prob{x.A11, x.B11, x.R3}; }); // Q1 = A11 B11
    step(EW_DFS_POST1, [](const Regs &x) { return V{x.R3, x.R2, x.R5, x.R4, x.R1}; },  // U3 C22 C21
         [](const Regs &x) { return V{x.R2, x.R4, x.R1}; });
    sub([](const Regs &x) { return Prob{x.A12, x.B21, x.R5}; }); // Q2 = A12 B21
    step(EW_DFS_POST2, [](const Regs &x) { return V{x.R3, x.R5}; },  // C11 = Q1 + Q2
         [](const Regs &x) { return V{x.R3}; });
    sub([](const Regs &x) { return Prob{x.R6, x.B22, x.R5}; });  // Q6 = T6 B22
    step(EW_DFS_POST3, [](const Regs &x) { return V{x.R2, x.R5}; },  // C12 = U3 + Q6
         [](const Regs &x) { return V{x.R6}; });
    top = mark;
  }

  // the fused FP64 passes move 16-byte pairs
  static bool aligned16(const Prob &p) {
    return ((reinterpret_cast<uintptr_t>(p.A.p) | reinterpret_cast<uintptr_t>(p.B.p) |
             reinterpret_cast<uintptr_t>(p.C.p)) & 15) == 0;
  }

  // Can two 'b' levels starting at `depth` run as fused passes?
  bool fusable(size_t batch, int s, int nrec, int depth) const {
    if (!fuse || batch != 1 || nrec - cutoff < 2) return false;
    if (!(pattern[depth] == 'b' || pattern[depth] == 'B')) return false;
    if (!(pattern[depth + 1] == 'b' || pattern[depth + 1] == 'B')) return false;
    const long long q = (long long)(s / 4) * (s / 4);
    return residue ? (q % 4 == 0) : (q % 2 == 0);
  }

  // Two fused 'b' levels for one sub-problem (fused2.h).  In residue mode at
  // the top, `crt` carries the reconstruction target and the combinations are
  // fused with it.
  void bfs2(const Prob &pb, int s, int nrec, int depth, const CrtMats *crt) {
    const int h = s / 2, sub = s / 4;
    const long long q = (long long)sub * sub;
    const size_t mark = top;
    level_tag(depth, pattern[depth]);
    level_tag(depth + 1, pattern[depth + 1]);
    plan.additions += 15LL * h * h + 7LL * 15 * q;
    // plain sub-blocks passed by pointer: level-1 / level-2 operand -> quadrant
    const int quadA[7] = {0, 2, -1, -1, -1, -1, 3};  // T1 = A11, T2 = A12, T7 = A22
    const int quadB[7] = {0, 1, -1, -1, -1, 3, -1};  // S1 = B11, S2 = B21, S6 = B22
    const bool fused_b = crt && crt->fused_b, fused_a = crt && crt->fused_a;
    Mat LA = alloc_mats(fused_a ? 49 : 40, q), LB = alloc_mats(fused_b ? 49 : 40, q),
        P2 = alloc_mats(49, q);
    // residue products written by the leaf GEMM itself (no level between) on
    // its 256-tile kernel: the 49 products of a modulus interleave in 128-byte
    // segments, so that the post-combination reads one contiguous 49 x 128
    // bytes per modulus for 128 positions
    const bool seg = residue && nrec - 2 == cutoff && q % 128 == 0 &&
                     residue_gemm_segmented(sub, cutoff, tile, mode == CAPS_MODE_CRT_BF16);
    const long long pstep = seg ? 128 : q * es, cseg = seg ? 49 * 128 : 128;
    // the kernels address the 40 written operands of a side as out + k * step
    // (k counts the non-plain (i, j) in order) and product (i, j) as
    // q + (7 i + j) * step
    BBPre pre{};
    pre.in[0] = pb.A.p;
    pre.in[1] = pb.B.p;
    pre.sin[0] = pb.A.ps;
    pre.sin[1] = pb.B.ps;
    pre.out[0] = LA.p;
    pre.out[1] = LB.p;
    pre.sout[0] = LA.ps;
    pre.sout[1] = LB.ps;
    pre.step = q * es;
    std::vector<Prob> next(49);
    int ka = 0, kb = 0;
    for (int i = 0; i < 7; i++)
      for (int j = 0; j < 7; j++) {
        const int k = 7 * i + j;
        Mat a, b;
        if (fused_a) {
          a = at(LA, (long long)k * q);
        } else if (bb_plain(0, i) && bb_plain(0, j)) {
          a = at(pb.A, (4 * quadA[i] + quadA[j]) * q);
        } else {
          a = at(LA, (long long)ka++ * q);
        }
        if (fused_b) {
          b = at(LB, (long long)k * q);
        } else if (bb_plain(1, i) && bb_plain(1, j)) {
          b = at(pb.B, (4 * quadB[i] + quadB[j]) * q);
        } else {
          b = at(LB, (long long)kb++ * q);
        }
        next[k] = {a, b, {P2.p + k * pstep, P2.ps}, cseg};
      }
    if (ka != (fused_a ? 0 : 40) || kb != (fused_b ? 0 : 40)) broken = true;
    if (residue) ensure_consts();
    if (fused_b) {
      BBSplitB sp{};
      sp.B = crt->B;
      sp.out = reinterpret_cast<int8_t *>(LB.p);
      sp.sout = LB.ps;
      sp.sb = crt->sb;
      sp.sub = sub;
      sp.m0 = crt->m;
      sp.nrec2 = crt->nrec - 2;
      push(STEP_BB_SPLIT_B, 1, 1, q, 0, add_table(&sp, sizeof(sp)));
    }
    if (fused_a) {
      BBSplitA sp{};
      sp.A = crt->A;
      sp.out = reinterpret_cast<int8_t *>(LA.p);
      sp.sout = LA.ps;
      sp.sa = crt->sa;
      sp.sub = sub;
      sp.m0 = crt->m;
      sp.nrec2 = crt->nrec - 2;
      push(STEP_BB_SPLIT_A, 1, 1, q, crt->m, add_table(&sp, sizeof(sp)));
    } else {
      push(STEP_BB_PRE, residue, 1, q, 0, add_table(&pre, sizeof(pre)), fused_b ? 1 : 0);
    }
    rec(next, sub, nrec - 2, depth + 2);
    BBPost post{};
    post.q = P2.p;
    post.step = pstep;
    post.sq = P2.ps;
    post.cseg = cseg;
    for (int k = 0; k < 49; k++)
      if (next[k].C.p != P2.p + k * pstep || next[k].C.ps != P2.ps) broken = true;
    if (!residue) {
      post.c = reinterpret_cast<double *>(pb.C.p);
    } else {
      post.out_c = crt->C;
      post.sa = crt->sa;
      post.sb = crt->sb;
      post.sub = sub;
      post.m0 = crt->m;
      post.nrec2 = crt->nrec - 2;
      post.bs = crt->bs;
    }
    push(STEP_BB_POST, residue, 1, q, 0, add_table(&post, sizeof(post)));
    top = mark;
  }

  void rec(const std::vector<Prob> &batch, int s, int nrec, int depth) {
    if (nrec == cutoff) {
      leaf(batch, s);
      return;
    }
    if (!residue && fusable(batch.size(), s, nrec, depth) && aligned16(batch[0])) {
      bfs2(batch[0], s, nrec, depth, nullptr);
      return;
    }
    char c = pattern[depth];
    level_tag(depth, c);
    if (c == 'b' || c == 'B')
      bfs(batch, s, nrec, depth);
    else
      dfs(batch, s, nrec, depth);
  }

  // residue mode: scale and split every A, B of the group once, recurse on
  // the residues of the whole group, reconstruct each C
  void crt_group(const std::vector<const BatchItem *> &items) {
    const caps_desc &d = items[0]->desc;
    const int n = d.lda, m = n >> d.nrec;
    const long long n2 = (long long)n * n;
    ensure_consts();
    std::vector<Prob> top_batch;
    std::vector<size_t> mats;
    // two fused 'b' levels at the top: the reconstruction is fused with the
    // combinations; with an even leaf size B's split is fused with its
    // pre-additions too, and with leaves a multiple of 128 A's split as well
    // (no residue planes of C, B and A).  Otherwise A's split is a pass of
    // its own, which transposes its leaves to row-major on the way.
    const bool top2 = items.size() == 1 && fusable(1, n, d.nrec, 0) && (d.bs == 1 || d.bs == m);
    const bool fused_b = top2 && m % 2 == 0;
    const bool fused_a = fused_b && m % 128 == 0;
    for (const BatchItem *it : items) {
      CrtMats M{};
      M.A = it->A;
      M.B = it->B;
      M.C = it->C;
      M.n = n;
      M.nrec = d.nrec;
      M.bs = d.bs;
      M.m = m;
      M.nmod = moduli;
      M.chunks = crt_row_chunks(m);
      const int nblk = 1 << d.nrec;
      M.sa = reinterpret_cast<int *>(base + alloc(sizeof(int) * n));
      M.sb = reinterpret_cast<int *>(base + alloc(sizeof(int) * n));
      M.pa = reinterpret_cast<double *>(base + alloc(sizeof(double) * 2 * n * nblk * M.chunks));
      M.pb = reinterpret_cast<double *>(base + alloc(sizeof(double) * 2 * n * nblk));
      M.fused_b = fused_b;
      M.fused_a = fused_a;
      Mat RA = fused_a ? Mat{} : alloc_mats(1, n2);
      Mat RB = fused_b ? Mat{} : alloc_mats(1, n2);
      Mat RC = top2 ? Mat{} : alloc_mats(1, n2);
      M.ra = reinterpret_cast<int8_t *>(RA.p);
      M.rb = reinterpret_cast<int8_t *>(RB.p);
      M.rc = reinterpret_cast<int8_t *>(RC.p);
      const size_t mt = add_table(&M, sizeof(M));
      plan.crt_host.push_back(M);
      top_batch.push_back({RA, RB, RC});
      mats.push_back(mt);
    }
    // every item's scale, then every item's split (the items are independent;
    // consecutive steps of one kind can share a segmented launch)
    const size_t h0 = plan.crt_host.size() - items.size();
    for (size_t i = 0; i < items.size(); i++) {
      push(STEP_CRT_SCALE, moduli, 1, n, d.bs, mats[i]);
      steps().back().aux = h0 + i;
    }
    for (size_t i = 0; i < items.size() && !fused_a; i++) {
      push(STEP_CRT_SPLIT, moduli, 1, n, d.bs, mats[i]);
      steps().back().aux = h0 + i;
    }
    if (top2) {
      bfs2(top_batch[0], n, d.nrec, 0, &plan.crt_host.back());
      return;
    }
    rec(top_batch, n, d.nrec, 0);
    for (size_t i = 0; i < items.size(); i++) {
      push(STEP_CRT_RECON, moduli, 1, n, d.bs, mats[i]);
      steps().back().aux = plan.crt_host.size() - items.size() + i;
    }
  }
};

hipError_t run_one(const Plan &plan, const Step &st, hipStream_t s) {
  const char *tb = reinterpret_cast<const char *>(plan.d_tables);
  const OzakiConsts *kc = reinterpret_cast<const OzakiConsts *>(tb + plan.oz_const);
  switch (st.kind) {
    case STEP_EW:
      return launch_ew(st.op, reinterpret_cast<const EwPtrs *>(tb + st.table), st.count, st.n,
                       st.vec2 != 0, s);
    case STEP_GEMM_F64:
      return launch_dgemm_batched(reinterpret_cast<const GemmPtrs *>(tb + st.table), st.count,
                                  (int)st.n, st.levels, st.tile, s);
    case STEP_REORDER:
      return launch_leaf_reorder(reinterpret_cast<const ReorderPtrs *>(tb + st.table), st.count,
                                 (int)st.n, st.bs, st.op == 1, s);
    case STEP_OZ_GEMM:
      if (st.bs == 1)
        return launch_residue_gemm_bf16(reinterpret_cast<const RGemm *>(tb + st.table), st.count,
                                        (int)st.n, st.levels, st.op, kc, s);
      return launch_residue_gemm(reinterpret_cast<const RGemm *>(tb + st.table), st.count,
                                 (int)st.n, st.levels, st.op, kc, s, st.tile);
    case STEP_OZ_SCALE:
    case STEP_OZ_SPLIT:
    case STEP_OZ_RECON:
      return launch_ozaki_step(st.kind - STEP_OZ_SCALE,
                               reinterpret_cast<const OzLeaf *>(tb + st.table), st.count,
                               (int)st.n, st.op, kc, s);
    case STEP_REW:
      return launch_residue_ew(st.op, reinterpret_cast<const RPtrs *>(tb + st.table), st.count,
                               st.n, plan.oz_moduli, st.vec2 != 0, kc, s);
    case STEP_BB_PRE:
      return launch_bb_pre(reinterpret_cast<const BBPre *>(tb + st.table), st.n, st.op != 0,
                           plan.oz_moduli, kc, s, st.vec2 != 0);
    case STEP_BB_SPLIT_B:
      return launch_bb_split_b(reinterpret_cast<const BBSplitB *>(tb + st.table), st.n, kc, s);
    case STEP_BB_SPLIT_A:
      return launch_bb_split_a(reinterpret_cast<const BBSplitA *>(tb + st.table), st.n, st.bs,
                               kc, s);

    case STEP_BB_POST:
      return launch_bb_post(reinterpret_cast<const BBPost *>(tb + st.table), st.n, st.op != 0,
                            st.op ? ozaki_npieces(plan.oz_moduli) : 0, kc, s);
    case STEP_CRT_SCALE:
    case STEP_CRT_SPLIT:
    case STEP_CRT_RECON:
      return launch_crt_step(st.kind - STEP_CRT_SCALE,
                             reinterpret_cast<const CrtMats *>(tb + st.table),
                             plan.crt_host[st.aux], kc, s);
    case STEP_MULTI_EW:
      return launch_ew_multi(st.op, tb, reinterpret_cast<const VSeg *>(tb + st.table), st.count,
                             (int)st.n, st.vec2 != 0, s);
    case STEP_MULTI_REW:
      return launch_residue_ew_multi(st.op, tb, reinterpret_cast<const VSeg *>(tb + st.table),
                                     st.count, (int)st.n, st.vec2 != 0, kc, s);
    case STEP_MULTI_GEMM:
      return launch_dgemm_multi(st.bs, tb, reinterpret_cast<const VSeg *>(tb + st.table),
                                st.count, (int)st.n, s);
    case STEP_MULTI_RGEMM:
      return launch_residue_gemm_multi(tb, reinterpret_cast<const VSeg *>(tb + st.table),
                                       st.count, (int)st.n, plan.oz_moduli, kc, s);
    case STEP_MULTI_CRT:
      return launch_crt_multi(st.op, st.vec2 != 0, ozaki_npieces(plan.oz_moduli), tb,
                              reinterpret_cast<const VSeg *>(tb + st.table), st.count, (int)st.n,
                              kc, s);
  }
  return hipErrorInvalidValue;
}

int timer_of(const Step &st) {
  if (st.kind == STEP_BB_POST && st.op == 0) return CAPS_TIMER_ADD;
  switch (st.kind) {
    case STEP_EW:
    case STEP_REW:
    case STEP_BB_PRE: return CAPS_TIMER_ADD;
    case STEP_BB_POST: return CAPS_TIMER_RECON;
    case STEP_GEMM_F64:
    case STEP_OZ_GEMM: return CAPS_TIMER_MUL;
    case STEP_REORDER: return CAPS_TIMER_REARRANGE;
    case STEP_OZ_SCALE:
    case STEP_OZ_SPLIT:
    case STEP_CRT_SCALE:
    case STEP_CRT_SPLIT:
    case STEP_BB_SPLIT_A:
    case STEP_BB_SPLIT_B: return CAPS_TIMER_SPLIT;
    case STEP_OZ_RECON:
    case STEP_CRT_RECON: return CAPS_TIMER_RECON;
    case STEP_MULTI_EW:
    case STEP_MULTI_REW: return CAPS_TIMER_ADD;
    case STEP_MULTI_GEMM:
    case STEP_MULTI_RGEMM: return CAPS_TIMER_MUL;
    case STEP_MULTI_CRT: return st.op == CRT_RECON ? CAPS_TIMER_RECON : CAPS_TIMER_SPLIT;
  }
  return CAPS_TIMER_COMM;
}

}  // namespace

int build_plan(Plan &plan, const double *A, const double *B, double *C, const caps_desc &d,
               const std::string &pattern, int mode, int moduli, int cutoff, void *work_base,
               bool fuse, int tile) {
  plan.steps.clear();
  plan.host_tables.clear();
  plan.additions = plan.multiplications = plan.leaves = 0;
  plan.exec_pattern.clear();
  plan.oz_moduli = 0;
  plan.crt_host.clear();
  Builder b{plan, reinterpret_cast<char *>(work_base)};
  b.bs = d.bs;
  b.mode = mode;
  b.moduli = moduli;
  b.cutoff = cutoff;
  b.fuse = fuse;
  b.tile = tile;
  b.pattern = pattern;
  b.residue = residue_mode(mode);
  b.es = b.residue ? 1 : 8;
  b.planes = b.residue ? moduli : 1;
  if (b.residue) {
    BatchItem it{A, B, C, d, pattern, cutoff};
    b.crt_group({&it});
  } else {
    std::vector<Prob> top{{{reinterpret_cast<char *>(const_cast<double *>(A)), 0},
                           {reinterpret_cast<char *>(const_cast<double *>(B)), 0},
                           {reinterpret_cast<char *>(C), 0}}};
    b.rec(top, d.lda, d.nrec, 0);
  }
  plan.work_bytes = b.peak;
  return b.broken ? CAPS_EINVAL : CAPS_OK;
}

namespace {

// Which steps of different problems can share one segmented launch: equal
// keys merge; kind < 0 never merges.
struct MergeKey {
  int kind = -1, a = 0, b = 0;
  bool operator==(const MergeKey &o) const { return kind == o.kind && a == o.a && b == o.b; }
};

constexpr int MERGE_CRT_SCALE = 100, MERGE_CRT_SPLIT = 101, MERGE_CRT_RECON = 102;

MergeKey merge_key(const Plan &plan, const Step &st) {
  MergeKey k;
  int gx, gy;
  switch (st.kind) {
    case STEP_EW: k = {STEP_MULTI_EW, st.op, st.vec2}; break;
    case STEP_REW: k = {STEP_MULTI_REW, st.op, st.vec2}; break;
    case STEP_GEMM_F64:
      k = {STEP_MULTI_GEMM, dgemm_block_tile((int)st.n, st.levels, st.tile), 0};
      break;
    case STEP_OZ_GEMM: {
      // INT8 products that would take the 128 x 128 kernel alone; those on
      // 256-tile leaves keep their own launch, and so do those that take the
      // 256 tiles with partial edges when they fill the device (256 tiles)
      const int m = (int)st.n, m0 = m >> st.levels;
      const long long t256 =
          (long long)((m + 255) / 256) * ((m + 255) / 256) * st.count * st.op;  // op: moduli
      const bool big = (st.tile == 0 && m0 % 256 == 0 && (long long)m * m < (1LL << 31)) ||
                       (big_edge_gemm(m, st.levels, st.tile) && t256 >= 256);
      if (st.bs == 0 && !big && residue_gemm_check(m, st.levels) >= 0)
        k = {STEP_MULTI_RGEMM, 0, 0};
      break;
    }
    case STEP_CRT_SCALE:
      k = {STEP_MULTI_CRT, MERGE_CRT_SCALE, crt_grid(CRT_ROWPART, plan.crt_host[st.aux], gx, gy)};
      break;
    case STEP_CRT_SPLIT: {
      const CrtMats &h = plan.crt_host[st.aux];
      const int fa = crt_grid(CRT_SPLIT_A, h, gx, gy), fb = crt_grid(CRT_SPLIT_B, h, gx, gy);
      k = {STEP_MULTI_CRT, MERGE_CRT_SPLIT, fa | (fb << 1) | (h.fused_b << 2)};
      break;
    }
    case STEP_CRT_RECON:
      k = {STEP_MULTI_CRT, MERGE_CRT_RECON, crt_grid(CRT_RECON, plan.crt_host[st.aux], gx, gy)};
      break;
  }
  return k;
}

size_t push_segs(Plan &plan, const std::vector<VSeg> &segs) {
  auto &t = plan.host_tables;
  const size_t off = (t.size() + 63) / 64 * 64;
  t.resize(off + segs.size() * sizeof(VSeg));
  memcpy(t.data() + off, segs.data(), segs.size() * sizeof(VSeg));
  return off;
}

Step multi_step(int kind, int op, int flag, const std::vector<VSeg> &segs, long long total,
                size_t table) {
  Step st{};
  st.kind = kind;
  st.op = op;
  st.vec2 = flag;
  st.bs = flag;
  st.count = (int)segs.size();
  st.n = total;
  st.table = table;
  return st;
}

// one CRT kernel over the items of `heads`
void emit_crt(Plan &plan, const std::vector<const Step *> &heads, int k, bool flag) {
  std::vector<VSeg> segs;
  int blk = 0;
  for (const Step *st : heads) {
    VSeg v{};
    v.arg = (long long)st->table;
    v.blk0 = blk;
    crt_grid(k, plan.crt_host[st->aux], v.gx, v.gy);
    blk += v.gx * v.gy;
    segs.push_back(v);
  }
  const size_t off = push_segs(plan, segs);
  plan.steps.push_back(multi_step(STEP_MULTI_CRT, k, flag, segs, blk, off));
}

// the merged launch(es) for steps with key k
void emit_merged(Plan &plan, const MergeKey &k, const std::vector<const Step *> &heads,
                 int moduli) {
  if (k.kind == STEP_MULTI_CRT) {
    if (k.a == MERGE_CRT_SCALE) {
      emit_crt(plan, heads, CRT_ROWPART, k.b);
      emit_crt(plan, heads, CRT_COLPART, k.b);
      emit_crt(plan, heads, CRT_FINALIZE, false);
    } else if (k.a == MERGE_CRT_SPLIT) {
      emit_crt(plan, heads, CRT_SPLIT_A, k.b & 1);
      if (!(k.b & 4)) emit_crt(plan, heads, CRT_SPLIT_B, (k.b >> 1) & 1);
    } else {
      emit_crt(plan, heads, CRT_RECON, k.b);
    }
    return;
  }
  std::vector<VSeg> segs;
  long long total = 0;
  for (const Step *st : heads) {
    VSeg v{};
    v.arg = (long long)st->table;
    v.n = st->n;
    v.blk0 = (int)total;
    if (k.kind == STEP_MULTI_EW) {
      ew_grid(st->count, st->n, st->vec2 != 0, v.gx, v.gy);
    } else if (k.kind == STEP_MULTI_REW) {
      rew_grid(st->count, st->n, moduli, st->vec2 != 0, v.gx, v.gy);
      v.aux = moduli;
    } else if (k.kind == STEP_MULTI_GEMM) {
      const int m = (int)st->n, bt = k.a;
      v.gx = (m + bt - 1) / bt;
      v.gy = st->count;
      v.aux = st->levels | (((m >> st->levels) % bt != 0) << 8);
    } else {  // STEP_MULTI_RGEMM
      const int m = (int)st->n;
      v.gx = (m + 127) / 128;
      v.gy = st->count * moduli;
      v.aux = st->levels | (residue_gemm_check(m, st->levels) << 8);
    }
    // elementwise segments count blocks (gx x gy), GEMM segments tiles
    // (gx x gx per problem, gy problems)
    const bool gemm = k.kind == STEP_MULTI_GEMM || k.kind == STEP_MULTI_RGEMM;
    total += gemm ? (long long)v.gx * v.gx * v.gy : (long long)v.gx * v.gy;
    segs.push_back(v);
  }
  const size_t off = push_segs(plan, segs);
  // op: the Winograd op of an add step; flag: 16/4-byte accesses (adds) or
  // the FP64 block tile
  plan.steps.push_back(multi_step(k.kind, k.a, k.kind == STEP_MULTI_GEMM ? k.a : k.b, segs,
                                  total, off));
}

// Interleave the groups' step lists: repeatedly take the key shared by the
// most groups at their current step (ties: the earliest group), launch those
// steps together, advance those groups.  Each group's own order is kept; the
// groups are independent (own operands and workspace), so any interleaving
// is valid.
void merge_groups(Plan &plan, const std::vector<std::vector<Step>> &gsteps, int moduli) {
  const size_t G = gsteps.size();
  std::vector<size_t> pos(G, 0);
  for (;;) {
    std::vector<MergeKey> keys(G);
    int best = -1, best_n = 0;
    for (size_t g = 0; g < G; g++) {
      if (pos[g] >= gsteps[g].size()) continue;
      keys[g] = merge_key(plan, gsteps[g][pos[g]]);
      int n = 0;
      if (keys[g].kind >= 0)
        for (size_t h = 0; h < G; h++)
          if (pos[h] < gsteps[h].size() && merge_key(plan, gsteps[h][pos[h]]) == keys[g]) n++;
      if (best < 0 || n > best_n) {
        best = (int)g;
        best_n = n;
      }
    }
    if (best < 0) return;
    // a run of per-item CRT steps in one group merges on its own too
    const bool run = keys[best].kind == STEP_MULTI_CRT &&
                     pos[best] + 1 < gsteps[best].size() &&
                     merge_key(plan, gsteps[best][pos[best] + 1]) == keys[best];
    if (keys[best].kind < 0 || (best_n < 2 && !run)) {
      plan.steps.push_back(gsteps[best][pos[best]++]);
      continue;
    }
    std::vector<const Step *> heads;
    for (size_t g = 0; g < G; g++)
      if (pos[g] < gsteps[g].size() && keys[g] == keys[best]) {
        heads.push_back(&gsteps[g][pos[g]++]);
        // per-item scale / split / reconstruct steps of one group are
        // independent of each other: a run of them joins the same launch
        if (keys[best].kind == STEP_MULTI_CRT)
          while (pos[g] < gsteps[g].size() && merge_key(plan, gsteps[g][pos[g]]) == keys[best])
            heads.push_back(&gsteps[g][pos[g]++]);
      }
    emit_merged(plan, keys[best], heads, moduli);
  }
}

}  // namespace

int build_batch_plan(Plan &plan, const std::vector<BatchItem> &items, int mode, int moduli,
                     void *work_base, int tile, size_t merge_budget) {
  plan.steps.clear();
  plan.host_tables.clear();
  plan.additions = plan.multiplications = plan.leaves = 0;
  plan.exec_pattern.clear();
  plan.oz_moduli = 0;
  plan.crt_host.clear();
  // group by (descriptor, pattern, cutoff), in first-appearance order
  std::vector<std::vector<const BatchItem *>> groups;
  for (const BatchItem &it : items) {
    bool placed = false;
    for (auto &g : groups)
      if (memcmp(&g[0]->desc, &it.desc, sizeof(caps_desc)) == 0 && g[0]->pattern == it.pattern &&
          g[0]->cutoff == it.cutoff) {
        g.push_back(&it);
        placed = true;
        break;
      }
    if (!placed) groups.push_back({&it});
  }
  // Merged schedule: every group gets a workspace region of its own and the
  // same step of different groups runs as one segmented launch.  When the
  // regions together exceed merge_budget, the groups run one after another
  // in one shared region instead.
  auto build_group = [&](Plan &into, const std::vector<const BatchItem *> &g, char *base,
                         std::vector<Step> *out, size_t &peak) {
    Builder b{into, base};
    b.out = out;
    b.bs = g[0]->desc.bs;
    b.mode = mode;
    b.moduli = moduli;
    b.cutoff = g[0]->cutoff;
    b.pattern = g[0]->pattern;
    b.residue = residue_mode(mode);
    b.es = b.residue ? 1 : 8;
    b.planes = b.residue ? moduli : 1;
    b.fuse = true;
    b.tile = tile;
    if (b.residue) {
      b.crt_group(g);
    } else {
      std::vector<Prob> batch;
      for (const BatchItem *it : g)
        batch.push_back({{reinterpret_cast<char *>(const_cast<double *>(it->A)), 0},
                         {reinterpret_cast<char *>(const_cast<double *>(it->B)), 0},
                         {reinterpret_cast<char *>(it->C), 0}});
      b.rec(batch, g[0]->desc.lda, g[0]->desc.nrec, 0);
    }
    peak = (b.peak + ALIGN - 1) / ALIGN * ALIGN;
    return !b.broken;
  };
  // each group's workspace, from a throwaway build with the same items
  std::vector<size_t> peaks(groups.size());
  size_t sum = 0, mx = 0;
  for (size_t g = 0; g < groups.size(); g++) {
    Plan scratch;
    std::vector<Step> tmp;
    if (!build_group(scratch, groups[g], nullptr, &tmp, peaks[g])) return CAPS_EINVAL;
    sum += peaks[g];
    mx = std::max(mx, peaks[g]);
  }
  const bool merge = (groups.size() > 1 || groups[0].size() > 1) && sum <= merge_budget;
  char *base = reinterpret_cast<char *>(work_base);
  if (merge) {
    std::vector<std::vector<Step>> gsteps(groups.size());
    size_t off = 0;
    for (size_t g = 0; g < groups.size(); g++) {
      size_t pk = 0;
      if (!build_group(plan, groups[g], base ? base + off : nullptr, &gsteps[g], pk) ||
          pk != peaks[g])
        return CAPS_EINVAL;
      off += peaks[g];
    }
    merge_groups(plan, gsteps, moduli);
    plan.work_bytes = sum;
  } else {
    for (auto &g : groups) {
      size_t pk = 0;
      if (!build_group(plan, g, base, nullptr, pk)) return CAPS_EINVAL;
    }
    plan.work_bytes = mx;
  }
  plan.exec_pattern = "BATCH,";
  return CAPS_OK;
}

size_t plan_workspace_bytes(const caps_desc &d, const std::string &pattern, int mode,
                            int moduli, int cutoff) {
  Plan p;
  build_plan(p, nullptr, nullptr, nullptr, d, pattern, mode, moduli, cutoff, nullptr, true);
  return p.work_bytes;
}

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
    const int k = timer_of(plan.steps[i]);
    ms[k] += t;
    n[k] += 1;
  }
  for (size_t i = 0; i < made; i++) hipEventDestroy(ev[i]);
  return e;
}

}  // namespace caps
