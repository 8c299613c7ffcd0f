// Candidate configurations, the exhaustive autotuner and the learned planner.
//
// A candidate is (leaf mode, moduli, cutoff, b/d pattern over the recursed
// levels).  The autotuner measures every candidate on the caller's operands
// (time: HIP events; error: against the 16-modulus residue-domain product)
// and keeps the fastest one meeting the error target.  The learned planner
// scores the same candidates with the cost / error model instead (one forward
// pass of the 3-layer perceptron over all of them) and picks the predicted
// fastest whose predicted error meets the target.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/caps_amd.h"
#include "engine.h"
#include "mlp.h"
#include "ozaki.h"
#include "planner.h"
#include "tune.h"

namespace caps {

int set_error(int code, const char *msg);

// b/d patterns over L recursed levels: the prefix patterns b^k d^(L-k) first
// (k = L .. 0; the learned planner's features count b and d levels, so among
// patterns it cannot tell apart it keeps the first), then, for L <= 4, every
// other pattern
static std::vector<std::string> level_patterns(int L) {
  std::vector<std::string> out;
  for (int k = L; k >= 0; k--)
    out.push_back(std::string((size_t)k, 'b') + std::string((size_t)(L - k), 'd'));
  if (L <= 4)
    for (int bits = 0; bits < (1 << L); bits++) {
      std::string p((size_t)L, 'd');
      for (int l = 0; l < L; l++)
        if (bits & (1 << (L - 1 - l))) p[(size_t)l] = 'b';
      if (std::find(out.begin(), out.end(), p) == out.end()) out.push_back(p);
    }
  return out;
}

std::vector<caps_candidate> enumerate_candidates(const caps_desc &d, int mask) {
  std::vector<caps_candidate> out;
  const int m0 = d.lda >> d.nrec;
  const bool plain_leaf = (d.bs == 1 || d.bs == m0);
  struct ModeSet {
    int mode;
    std::vector<int> moduli;
  };
  std::vector<ModeSet> modes;
  if (mask & CAPS_MASK_FP64) modes.push_back({CAPS_MODE_FP64, {0}});
  if ((mask & CAPS_MASK_CRT_INT8) && d.nrec <= 7)
    modes.push_back({CAPS_MODE_CRT_INT8, {12, 13, 14, 15, 16}});
  if (mask & CAPS_MASK_OZAKI_INT8) modes.push_back({CAPS_MODE_OZAKI_INT8, {14, 16}});
  if ((mask & CAPS_MASK_CRT_BF16) && d.nrec <= 7)
    modes.push_back({CAPS_MODE_CRT_BF16, {14, 16}});
  for (const ModeSet &ms : modes)
    for (int cutoff = 0; cutoff <= d.nrec; cutoff++) {
      if (cutoff > 0 && (ms.mode == CAPS_MODE_OZAKI_INT8 || m0 % 128 != 0 || !plain_leaf)) break;
      const int L = d.nrec - cutoff;
      // leaf-GEMM block tiles: FP64 128 or 64; INT8 residue products 256 or
      // 128 where the layout leaf allows the 256 tiles; BF16 one tile
      int tiles = 1;
      if (ms.mode == CAPS_MODE_FP64) tiles = 2;
      if ((ms.mode == CAPS_MODE_CRT_INT8 || ms.mode == CAPS_MODE_OZAKI_INT8) && m0 % 256 == 0)
        tiles = 2;
      for (const std::string &p : level_patterns(L))
        for (int mod : ms.moduli)
          for (int tile = 0; tile < tiles; tile++) {
            caps_candidate c{};
            c.mode = ms.mode;
            c.moduli = mod;
            c.cutoff = cutoff;
            c.tile = tile;
            snprintf(c.pattern, sizeof(c.pattern), "%s",
                     (p + std::string((size_t)cutoff, 'b')).c_str());
            c.ms = -1;
            c.err = -1;
            out.push_back(c);
          }
    }
  return out;
}

void candidate_features(const caps_desc &d, const caps_candidate &c, double spread, float *f) {
  int nb = 0, nd = 0;
  for (int l = 0; l < d.nrec - c.cutoff; l++) (c.pattern[l] == 'b' || c.pattern[l] == 'B') ? nb++ : nd++;
  f[0] = (float)std::log2((double)d.lda);
  f[1] = (float)spread;
  f[2] = (float)nb;
  f[3] = (float)nd;
  f[4] = (float)c.cutoff;
  f[5] = (float)c.moduli;
  // mode as one ordinal: 0 FP64, 1 residue (INT8), 2 leaf split, 3 residue (BF16)
  f[6] = c.mode == CAPS_MODE_CRT_INT8     ? 1.f
         : c.mode == CAPS_MODE_OZAKI_INT8 ? 2.f
         : c.mode == CAPS_MODE_CRT_BF16   ? 3.f
                                          : 0.f;
  // power-of-two alignment of the layout's leaf, log2(gcd(leaf, 256)) / 8: it
  // picks the leaf GEMM's path (256 tiles, 128 tiles, edge-checked variants)
  const int leaf = d.lda >> d.nrec;
  int a = 0;
  while (a < 8 && leaf > 0 && (leaf >> a) % 2 == 0) a++;
  f[7] = (float)a / 8.f;
}

namespace {

struct ModelSlot {
  std::mutex mu;
  caps_mlp *m = nullptr;
  float norm[2 * MLP_IN + 2 * MLP_OUT];
};

ModelSlot &model() {
  static ModelSlot s;
  return s;
}

}  // namespace

// caps_mlp_destroy of the installed planner model uninstalls it first (the
// library does not own the model; a destroyed one must not stay installed)
void planner_forget_model(const caps_mlp *m) {
  ModelSlot &ms = model();
  std::lock_guard<std::mutex> g(ms.mu);
  if (ms.m == m) ms.m = nullptr;
}

bool mlp_planner_choose(const caps_desc &d, int mask, double max_err, double spread, int &mode,
                        int &moduli, int &cutoff, std::string &pat, double *pred) {
  ModelSlot &ms = model();
  std::lock_guard<std::mutex> g(ms.mu);
  if (!ms.m) return false;
  // the model has no tile feature: it scores the default tiles, the
  // autotuner searches the tile dimension
  std::vector<caps_candidate> cands;
  for (const caps_candidate &c : enumerate_candidates(d, mask))
    if (c.tile == 0) cands.push_back(c);
  if (cands.empty()) return false;
  const int B = (int)cands.size();
  std::vector<float> X((size_t)B * MLP_IN), Y((size_t)B * MLP_OUT);
  for (int i = 0; i < B; i++) {
    float f[MLP_IN];
    candidate_features(d, cands[i], spread, f);
    for (int k = 0; k < MLP_IN; k++)
      X[(size_t)i * MLP_IN + k] = (f[k] - ms.norm[k]) / ms.norm[MLP_IN + k];
  }
  if (caps_mlp_forward(ms.m, X.data(), B, Y.data()) != CAPS_OK) return false;
  const float *tm = ms.norm + 2 * MLP_IN, *ts = tm + MLP_OUT;
  const double target = std::log10(max_err);
  int best = -1;
  double best_t = INFINITY, best_e = 0;
  for (int i = 0; i < B; i++) {
    const double lt = Y[(size_t)i * MLP_OUT] * ts[0] + tm[0];      // log2(ms)
    const double le = Y[(size_t)i * MLP_OUT + 1] * ts[1] + tm[1];  // log10(error)
    if (le <= target - 0.25 && lt < best_t) {  // a quarter-decade of safety on the error
      best_t = lt;
      best_e = le;
      best = i;
    }
  }
  if (best < 0) return false;
  if (pred) {
    pred[0] = std::exp2(best_t);
    pred[1] = std::pow(10.0, best_e);
  }
  mode = cands[best].mode;
  moduli = cands[best].moduli ? cands[best].moduli : OZ_DEFAULT_MODULI;
  cutoff = cands[best].cutoff;
  pat.assign(cands[best].pattern, (size_t)d.nrec);
  return true;
}

}  // namespace caps

using namespace caps;

extern "C" {

int caps_planner_use_model(caps_mlp *m, const float *norm, int nnorm) {
  ModelSlot &ms = model();
  std::lock_guard<std::mutex> g(ms.mu);
  if (m && (!norm || nnorm != 2 * MLP_IN + 2 * MLP_OUT))
    return set_error(CAPS_EINVAL, "caps_planner_use_model: norm needs 20 floats");
  ms.m = m;
  if (m) memcpy(ms.norm, norm, sizeof(ms.norm));
  return CAPS_OK;
}

int caps_enumerate_candidates(const caps_desc *d, int mode_mask, caps_candidate *out,
                              int max_out, int *n_out) {
  if (!d || !n_out) return set_error(CAPS_EINVAL, "caps_enumerate_candidates: NULL");
  std::vector<caps_candidate> c = enumerate_candidates(*d, mode_mask);
  *n_out = (int)c.size();
  for (int i = 0; i < (int)c.size() && i < max_out && out; i++) out[i] = c[i];
  return CAPS_OK;
}

int caps_candidate_features(const caps_desc *d, const caps_candidate *c, double spread,
                            float *features8) {
  if (!d || !c || !features8) return set_error(CAPS_EINVAL, "caps_candidate_features: NULL");
  candidate_features(*d, *c, spread, features8);
  return CAPS_OK;
}

int caps_autotune(const double *dA, const double *dB, double *dC, const caps_desc *desc,
                  int mode_mask, double max_err, int reps, caps_candidate *out, int max_out,
                  int *n_out, caps_candidate *best, void *stream) {
  int rc = caps_verify_desc(desc);
  if (rc) return rc;
  if (!dA || !dB || !dC || reps < 1) return set_error(CAPS_EINVAL, "caps_autotune: bad arguments");
  std::vector<caps_candidate> cands = enumerate_candidates(*desc, mode_mask);
  if (cands.empty()) return set_error(CAPS_EUNSUPPORTED, "caps_autotune: no candidate");
  const size_t n2 = (size_t)desc->lda * desc->lda;
  double *ref = nullptr;
  if (hipMalloc(&ref, n2 * sizeof(double)) != hipSuccess)
    return set_error(CAPS_ENOMEM, "caps_autotune: reference buffer");
  rc = caps_reference_product(dA, dB, ref, desc, stream);
  int found = -1;
  int measured = 0;
  for (size_t i = 0; i < cands.size() && !rc; i++) {
    size_t freeb = 0, totalb = 0;
    hipMemGetInfo(&freeb, &totalb);
    if (plan_workspace_bytes(*desc, cands[i].pattern, cands[i].mode, cands[i].moduli,
                             cands[i].cutoff) > freeb / 10 * 8)
      continue;  // does not fit
    int r = caps_measure(dA, dB, dC, desc, ref, reps, &cands[i], stream);
    if (r) {
      rc = r;
      break;
    }
    if (out && measured < max_out) out[measured] = cands[i];
    measured++;
    if (cands[i].err <= max_err && (found < 0 || cands[i].ms < cands[found].ms)) found = (int)i;
  }
  hipStreamSynchronize((hipStream_t)stream);
  hipFree(ref);
  if (rc) return rc;
  if (n_out) *n_out = measured;
  if (found < 0) return set_error(CAPS_EUNSUPPORTED, "caps_autotune: no candidate met the error target");
  if (best) *best = cands[found];
  TuneEntry e{desc->lda, desc->nrec, desc->bs, mode_mask, max_err, cands[found]};
  tune_store(e);
  return CAPS_OK;
}

}  // extern "C"
