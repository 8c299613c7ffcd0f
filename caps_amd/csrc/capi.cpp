// C ABI of the library (include/caps_amd.h): argument checking, pattern
// resolution, the plan cache, hipGraph replay and host-pointer staging.
//
// Errors never abort: every entry point returns a caps_status and records a
// message for caps_last_error() (the reference prints and exits instead,
// multiply.cpp:94-104).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/caps_amd.h"
#include "engine.h"
#include "ozaki.h"
#include "planner.h"
#include "tune.h"

// Device scratch of caps_multiply_device_noalloc: every GPU the library runs
// on holds its own copy from module load until the process ends.  Host code
// never uses this name as an address: the per-device address comes from
// hipGetSymbolAddress (State::scratch).
static __device__ __align__(256) char scratch[(size_t)CAPS_NOALLOC_SCRATCH_MIB << 20];

namespace caps {

namespace {

thread_local std::string g_error;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
  return code;
}

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(CAPS_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_));             \
  } while (0)

struct State {
  std::mutex mu;
  int device = -1;
  int cus = 256;
  std::vector<std::unique_ptr<Plan>> plans;
  unsigned long long clock = 0;
  long long additions = 0, multiplications = 0, leaves = 0;
  double peak_mb = 0;
  std::string exec_pattern;
  double timers_ms[CAPS_NUM_TIMERS] = {};
  long long timed[CAPS_NUM_TIMERS] = {};
  long long available_doubles = 1LL << 28;  // memory.cpp:4
  // set by caps_set_available_memory; unset, the planner takes half of the
  // free device memory
  bool budget_set = false;
  // CAPS_MODE_AUTO: error target, allowed modes, expected exponent spread
  double auto_max_err = 1e-14;
  int auto_mask = CAPS_MASK_FP64 | CAPS_MASK_CRT_INT8;
  double auto_spread = 10.0;
  // host staging buffers (device side) for the host-pointer entry points
  double *stage[3] = {nullptr, nullptr, nullptr};
  size_t stage_elems = 0;
  hipStream_t stage_stream = nullptr;
  // caps_multiply_device_noalloc: the address of each device's scratch array,
  // and host-only plans (their tables are copied into the scratch per call)
  std::map<int, char *> scratch;
  std::vector<std::unique_ptr<Plan>> host_plans;
};

State &st() {
  static State s;
  return s;
}

constexpr size_t MAX_PLANS = 24;
// plans with more launches than this (deep 'd' recursions: 7^d leaf
// products one after another) are launched directly instead of captured
constexpr size_t MAX_GRAPH_STEPS = 16384;

int ensure_device_locked(State &s) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (s.device == dev) return CAPS_OK;
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, dev));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(CAPS_ENODEV, "device %d is %s, this library is built for gfx950", dev,
                prop.gcnArchName);
  if (!s.scratch.count(dev)) {  // once per device, with that device current
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(CAPS_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                  hipGetErrorString(e));
    s.scratch[dev] = static_cast<char *>(p);
  }
  s.device = dev;
  s.cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  return CAPS_OK;
}

void free_plan(Plan &p) {
  if (p.graph) hipGraphExecDestroy(p.graph);
  if (p.capture_stream) hipStreamDestroy(p.capture_stream);
  if (p.d_tables) hipFree(p.d_tables);
  if (p.d_work) hipFree(p.d_work);
  p.graph = nullptr;
  p.capture_stream = nullptr;
  p.d_tables = p.d_work = nullptr;
}

void evict_locked(State &s, size_t keep) {
  if (s.plans.size() <= keep) return;
  hipDeviceSynchronize();
  std::sort(s.plans.begin(), s.plans.end(),
            [](const std::unique_ptr<Plan> &a, const std::unique_ptr<Plan> &b) {
              return a->last_use > b->last_use;
            });
  while (s.plans.size() > keep) {
    free_plan(*s.plans.back());
    s.plans.pop_back();
  }
}

int check_desc(const caps_desc *d) {
  if (!d) return fail(CAPS_EINVAL, "desc is NULL");
  if (d->nproc != 1 || d->nprocr != 1 || d->nprocc != 1 || d->nproc_summa != 1)
    return fail(CAPS_EINVAL,
                "one device needs nproc = nprocr = nprocc = nproc_summa = 1 (got %d %d %d %d)",
                d->nproc, d->nprocr, d->nprocc, d->nproc_summa);
  if (d->lda <= 0 || d->bs <= 0 || d->nrec < 0 || d->nrec > 24)
    return fail(CAPS_EINVAL, "bad descriptor lda=%d nrec=%d bs=%d", d->lda, d->nrec, d->bs);
  long long div = (1LL << d->nrec) * d->bs;
  if (d->lda % div != 0)  // matrix.cpp:4-5
    return fail(CAPS_EINVAL, "lda=%d is not a multiple of 2^nrec * bs = %lld", d->lda, div);
  if ((long long)d->lda * d->lda > (1LL << 34))
    return fail(CAPS_EINVAL, "lda=%d too large", d->lda);
  return CAPS_OK;
}

int resolve_pattern(const caps_desc &d, const char *pattern, const caps_opts *o,
                    std::string &out, int mode, int moduli, long long budget_doubles,
                    bool budget_set) {
  if (pattern) {
    out.clear();
    for (int l = 0; l < d.nrec; l++) {
      char c = pattern[l];
      if (c == '\0')
        return fail(CAPS_EPATTERN, "pattern \"%s\" shorter than nrec=%d", pattern, d.nrec);
      if (!strchr("bBdDhH", c))
        return fail(CAPS_EPATTERN, "invalid pattern char '%c' in \"%s\"", c, pattern);
      out.push_back(c);
    }
    return CAPS_OK;
  }
  if (o && (o->flags & CAPS_FLAG_REFERENCE_AUTO)) {
    out.assign((size_t)d.nrec, 'd');  // one process: DFS at every level (multiply.cpp:43-66)
    return CAPS_OK;
  }
  out = choose_pattern(d, mode, moduli, budget_doubles, budget_set);
  return CAPS_OK;
}

int resolve_mode(const caps_desc &d, const caps_opts *o, int &mode, int &moduli) {
  mode = o ? o->mode : CAPS_MODE_FP64;
  moduli = (o && o->moduli) ? o->moduli : OZ_DEFAULT_MODULI;
  if (mode == CAPS_MODE_AUTO) mode = choose_mode(d);
  if (mode != CAPS_MODE_FP64 && mode != CAPS_MODE_OZAKI_INT8 && mode != CAPS_MODE_CRT_INT8 &&
      mode != CAPS_MODE_CRT_BF16)
    return fail(CAPS_EINVAL, "unknown mode %d", mode);
  if (mode != CAPS_MODE_FP64 && (moduli < 2 || moduli > OZ_MAX_MODULI))
    return fail(CAPS_EINVAL, "moduli=%d outside 2..%d", moduli, OZ_MAX_MODULI);
  if (residue_mode(mode) && d.nrec > 7)
    return fail(CAPS_EUNSUPPORTED, "residue-domain mode supports nrec <= 7 (got %d)", d.nrec);
  if (mode == CAPS_MODE_FP64) moduli = 0;
  return CAPS_OK;
}

int check_cutoff(const caps_desc &d, int mode, int cutoff) {
  if (cutoff < 0 || cutoff > d.nrec)
    return fail(CAPS_EINVAL, "cutoff=%d outside 0..nrec=%d", cutoff, d.nrec);
  if (cutoff == 0) return CAPS_OK;
  const int m0 = d.lda >> d.nrec;
  if (mode == CAPS_MODE_OZAKI_INT8)
    return fail(CAPS_EUNSUPPORTED, "cutoff > 0 needs mode FP64 or CRT_INT8");
  if (m0 % 128 != 0)
    return fail(CAPS_EUNSUPPORTED, "cutoff > 0 needs leaves that are multiples of 128 (leaf %d)",
                m0);
  if (d.bs != 1 && d.bs != m0)
    return fail(CAPS_EUNSUPPORTED, "cutoff > 0 needs bs = 1 or bs = leaf size");
  return CAPS_OK;
}

// allocate, build and upload a plan whose builder is `build(plan, work)`
template <typename F>
int make_plan_locked(State &s, const PlanKey &key, F build, Plan *&out) {
  evict_locked(s, MAX_PLANS - 1);
  auto plan = std::make_unique<Plan>();
  plan->key = key;
  if (build(*plan, nullptr) != CAPS_OK) return fail(CAPS_EINVAL, "internal: plan layout check failed");
  const size_t need = plan->work_bytes;
  if (need) {
    size_t freeb = 0, totalb = 0;
    HIP_TRY(hipMemGetInfo(&freeb, &totalb));
    if (need > freeb) {
      evict_locked(s, 0);
      HIP_TRY(hipMemGetInfo(&freeb, &totalb));
      if (need > freeb)
        return fail(CAPS_ENOMEM, "plan needs %.1f MiB of device workspace, %.1f MiB free",
                    need / 1048576.0, freeb / 1048576.0);
    }
    hipError_t e = hipMalloc(&plan->d_work, need);
    if (e != hipSuccess)
      return fail(CAPS_ENOMEM, "hipMalloc(%zu) for workspace: %s", need, hipGetErrorString(e));
  }
  if (build(*plan, plan->d_work) != CAPS_OK) {
    free_plan(*plan);
    return fail(CAPS_EINVAL, "internal: plan layout check failed");
  }
  if (!plan->host_tables.empty()) {
    hipError_t e = hipMalloc(&plan->d_tables, plan->host_tables.size());
    if (e == hipSuccess)
      e = hipMemcpy(plan->d_tables, plan->host_tables.data(), plan->host_tables.size(),
                    hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      free_plan(*plan);
      return fail(CAPS_EHIP, "pointer tables: %s", hipGetErrorString(e));
    }
  }
  plan->last_use = ++s.clock;
  out = plan.get();
  s.plans.push_back(std::move(plan));
  return CAPS_OK;
}

int get_plan_locked(State &s, const double *A, const double *B, double *C, const caps_desc &d,
                    const std::string &pat, int mode, int moduli, int cutoff, bool fuse,
                    hipStream_t stream, int tile, Plan *&out) {
  PlanKey key{A, B, C, d, pat, mode, moduli, cutoff, fuse, std::string(), stream, tile};
  for (auto &p : s.plans)
    if (p->key == key) {
      out = p.get();
      p->last_use = ++s.clock;
      return CAPS_OK;
    }
  return make_plan_locked(
      s, key,
      [&](Plan &pl, void *work) {
        return build_plan(pl, A, B, C, d, pat, mode, moduli, cutoff, work, fuse, tile);
      },
      out);
}

int capture_graph(Plan &p) {
  HIP_TRY(hipStreamCreateWithFlags(&p.capture_stream, hipStreamNonBlocking));
  HIP_TRY(hipStreamBeginCapture(p.capture_stream, hipStreamCaptureModeThreadLocal));
  hipError_t le = run_steps(p, p.capture_stream);
  hipGraph_t g = nullptr;
  hipError_t ee = hipStreamEndCapture(p.capture_stream, &g);
  if (le != hipSuccess) {
    if (g) hipGraphDestroy(g);
    return fail(CAPS_EHIP, "kernel launch during capture: %s", hipGetErrorString(le));
  }
  if (ee != hipSuccess) return fail(CAPS_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(ee));
  hipError_t ie = hipGraphInstantiate(&p.graph, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (ie != hipSuccess) {
    p.graph = nullptr;
    return fail(CAPS_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(ie));
  }
  return CAPS_OK;
}

// CAPS_MODE_AUTO: the tuning table, then the learned model, then a default
// (residue-domain split when allowed: it meets a 1e-14 bar; else FP64)
int auto_choose(State &s, const caps_desc &d, const char *pattern, int &mode, int &moduli,
                int &cutoff, std::string &pat, double *pred = nullptr, int *tile = nullptr) {
  TuneEntry e;
  if (tile) *tile = 0;
  if (tune_lookup(d, s.auto_mask, s.auto_max_err, e)) {
    // a table entry goes through the same checks as a caller's options
    caps_opts te{e.c.mode, e.c.moduli, 0, e.c.cutoff};
    int rc = resolve_mode(d, &te, mode, moduli);
    if (rc) return rc;
    cutoff = e.c.cutoff;
    if (tile) *tile = e.c.tile ? 1 : 0;
    if (!pattern && (int)strlen(e.c.pattern) >= d.nrec)
      return resolve_pattern(d, e.c.pattern, &te, pat, mode, moduli, s.available_doubles,
                             s.budget_set);
  } else if (mlp_planner_choose(d, s.auto_mask, s.auto_max_err, s.auto_spread, mode, moduli,
                                cutoff, pat, pred)) {
    if (!pattern) return CAPS_OK;
  } else {
    mode = (s.auto_mask & CAPS_MASK_CRT_INT8) ? CAPS_MODE_CRT_INT8 : CAPS_MODE_FP64;
    moduli = mode == CAPS_MODE_FP64 ? 0 : OZ_DEFAULT_MODULI;
    cutoff = 0;
  }
  if (mode == CAPS_MODE_FP64) moduli = 0;
  caps_opts o{mode, moduli, 0, cutoff};
  return resolve_pattern(d, pattern, &o, pat, mode, moduli, s.available_doubles, s.budget_set);
}

int multiply_device_locked(State &s, const double *dA, const double *dB, double *dC,
                           const caps_desc *desc, const char *pattern, const caps_opts *opts,
                           hipStream_t stream) {
  int rc = check_desc(desc);
  if (rc) return rc;
  if (!dA || !dB || !dC) return fail(CAPS_EINVAL, "NULL matrix pointer");
  int mode, moduli, cutoff;
  int tile = (opts && (opts->flags & CAPS_FLAG_SMALL_TILE)) ? 1 : 0;
  std::string pat;
  if (opts && opts->mode == CAPS_MODE_AUTO) {
    int tuned_tile = 0;
    rc = auto_choose(s, *desc, pattern, mode, moduli, cutoff, pat, nullptr, &tuned_tile);
    if (rc) return rc;
    tile |= tuned_tile;
    if (pattern) {  // the caller's pattern with the chosen leaf mode
      caps_opts o{mode, moduli, 0, cutoff};
      rc = resolve_pattern(*desc, pattern, &o, pat, mode, moduli, s.available_doubles,
                           s.budget_set);
      if (rc) return rc;
    }
  } else {
    rc = resolve_mode(*desc, opts, mode, moduli);
    if (rc) return rc;
    cutoff = opts ? opts->cutoff : 0;
    rc = resolve_pattern(*desc, pattern, opts, pat, mode, moduli, s.available_doubles,
                         s.budget_set);
    if (rc) return rc;
  }
  rc = check_cutoff(*desc, mode, cutoff);
  if (rc) return rc;
  if (residue_mode(mode) && desc->nrec > 7)
    return fail(CAPS_EUNSUPPORTED, "residue-domain mode supports nrec <= 7 (got %d)", desc->nrec);
  Plan *plan = nullptr;
  const bool fuse = !(opts && (opts->flags & CAPS_FLAG_NO_FUSION));
  rc = get_plan_locked(s, dA, dB, dC, *desc, pat, mode, moduli, cutoff, fuse, stream, tile,
                       plan);
  if (rc) return rc;
  const bool timed = opts && (opts->flags & CAPS_FLAG_TIMERS);
  const bool use_graph = !timed && !(opts && (opts->flags & CAPS_FLAG_NO_GRAPH)) &&
                         plan->steps.size() <= MAX_GRAPH_STEPS;
  if (timed) {
    hipError_t e = run_steps_timed(*plan, stream, s.timers_ms, s.timed);
    if (e != hipSuccess) return fail(CAPS_EHIP, "timed launch: %s", hipGetErrorString(e));
  } else if (use_graph && !plan->graph) {
    rc = capture_graph(*plan);
    if (rc) return rc;
  }
  if (timed) {
  } else if (use_graph) {
    HIP_TRY(hipGraphLaunch(plan->graph, stream));
  } else {
    hipError_t e = run_steps(*plan, stream);
    if (e != hipSuccess) return fail(CAPS_EHIP, "kernel launch: %s", hipGetErrorString(e));
  }
  s.additions += plan->additions;
  s.multiplications += plan->multiplications;
  s.leaves += plan->leaves;
  s.peak_mb = plan->work_bytes / 1048576.0;
  s.exec_pattern = plan->exec_pattern;
  return CAPS_OK;
}

int stage_locked(State &s, size_t elems) {
  if (s.stage_elems >= elems && s.stage[0]) return CAPS_OK;
  for (auto &p : s.stage)
    if (p) {
      hipFree(p);
      p = nullptr;
    }
  s.stage_elems = 0;
  for (auto &p : s.stage) {
    hipError_t e = hipMalloc(&p, elems * sizeof(double));
    if (e != hipSuccess) return fail(CAPS_ENOMEM, "staging hipMalloc: %s", hipGetErrorString(e));
  }
  if (!s.stage_stream) HIP_TRY(hipStreamCreateWithFlags(&s.stage_stream, hipStreamNonBlocking));
  s.stage_elems = elems;
  return CAPS_OK;
}

}  // namespace

int device_cu_count() { return st().cus; }

int set_error(int code, const char *msg) { return fail(code, "%s", msg); }

}  // namespace caps

using namespace caps;

extern "C" {

int caps_init(int device) {
  State &s = st();
  std::lock_guard<std::mutex> g(s.mu);
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(CAPS_ENODEV, "device %d of %d", device, n);
  HIP_TRY(hipSetDevice(device));
  return ensure_device_locked(s);
}

const char *caps_last_error(void) { return g_error.c_str(); }

int caps_verify_desc(const caps_desc *desc) { return check_desc(desc); }

long long caps_num_entries(const caps_desc *d) {
  if (!d || d->nproc <= 0 || d->nproc_summa <= 0) return -1;
  long long lda = d->lda;
  return lda * lda / d->nproc / d->nproc_summa;
}

int caps_multiply_device(const double *dA, const double *dB, double *dC, const caps_desc *desc,
                         const char *pattern, const caps_opts *opts, void *stream) {
  State &s = st();
  std::lock_guard<std::mutex> g(s.mu);
  int rc = ensure_device_locked(s);
  if (rc) return rc;
  return multiply_device_locked(s, dA, dB, dC, desc, pattern, opts, (hipStream_t)stream);
}

int caps_multiply_device_noalloc(const double *dA, const double *dB, double *dC,
                                 const caps_desc *desc, const char *pattern,
                                 const caps_opts *opts, void *stream) {
  State &s = st();
  std::lock_guard<std::mutex> g(s.mu);
  int rc = ensure_device_locked(s);
  if (rc) return rc;
  rc = check_desc(desc);
  if (rc) return rc;
  if (!dA || !dB || !dC) return fail(CAPS_EINVAL, "NULL matrix pointer");
  if (opts && opts->mode == CAPS_MODE_AUTO)
    return fail(CAPS_EUNSUPPORTED, "caps_multiply_device_noalloc takes an explicit leaf mode");
  int mode, moduli;
  rc = resolve_mode(*desc, opts, mode, moduli);
  if (rc) return rc;
  const int cutoff = opts ? opts->cutoff : 0;
  const size_t cap = (size_t)CAPS_NOALLOC_SCRATCH_MIB << 20;
  // an automatic pattern must fit the scratch as well as the budget
  const size_t budget =
      std::min(planner_budget_bytes(s.available_doubles, s.budget_set), cap);
  std::string pat;
  rc = resolve_pattern(*desc, pattern, opts, pat, mode, moduli,
                       (long long)(budget / sizeof(double)), true);
  if (rc) return rc;
  rc = check_cutoff(*desc, mode, cutoff);
  if (rc) return rc;
  const bool fuse = !(opts && (opts->flags & CAPS_FLAG_NO_FUSION));
  const int tile = (opts && (opts->flags & CAPS_FLAG_SMALL_TILE)) ? 1 : 0;
  hipStream_t hs = (hipStream_t)stream;
  char *base = s.scratch[s.device];
  // host-only plans, one per (buffers, shape, options, device): their tables
  // hold this device's scratch addresses
  PlanKey key{dA, dB, dC, *desc, pat, mode, moduli, cutoff, fuse,
              "scratch of device " + std::to_string(s.device), nullptr, tile};
  Plan *p = nullptr;
  for (auto &hp : s.host_plans)
    if (hp->key == key) {
      p = hp.get();
      p->last_use = ++s.clock;
    }
  if (!p) {
    auto plan = std::make_unique<Plan>();
    plan->key = key;
    if (build_plan(*plan, dA, dB, dC, *desc, pat, mode, moduli, cutoff, nullptr, fuse, tile) !=
        CAPS_OK)
      return fail(CAPS_EINVAL, "internal: plan layout check failed");
    const size_t toff = (plan->host_tables.size() + 255) / 256 * 256;
    if (toff + plan->work_bytes > cap)
      return fail(CAPS_ENOMEM,
                  "plan needs %.1f MiB of tables and workspace; caps_multiply_device_noalloc "
                  "has %d MiB of scratch",
                  (toff + plan->work_bytes) / 1048576.0, CAPS_NOALLOC_SCRATCH_MIB);
    if (build_plan(*plan, dA, dB, dC, *desc, pat, mode, moduli, cutoff, base + toff, fuse,
                   tile) != CAPS_OK ||
        (plan->host_tables.size() + 255) / 256 * 256 != toff)
      return fail(CAPS_EINVAL, "internal: plan layout check failed");
    plan->d_tables = base;
    plan->d_work = base + toff;
    plan->last_use = ++s.clock;
    // host memory only; keep the 16 most recent
    if (s.host_plans.size() >= 16) {
      auto oldest = std::min_element(
          s.host_plans.begin(), s.host_plans.end(),
          [](const std::unique_ptr<Plan> &a, const std::unique_ptr<Plan> &b) {
            return a->last_use < b->last_use;
          });
      s.host_plans.erase(oldest);
    }
    p = plan.get();
    s.host_plans.push_back(std::move(plan));
  }
  // the tables go to the front of the scratch in stream order (a pageable
  // source is read before the call returns), then the steps run eagerly
  HIP_TRY(hipMemcpyAsync(base, p->host_tables.data(), p->host_tables.size(),
                         hipMemcpyHostToDevice, hs));
  hipError_t e = run_steps(*p, hs);
  if (e != hipSuccess) return fail(CAPS_EHIP, "kernel launch: %s", hipGetErrorString(e));
  s.additions += p->additions;
  s.multiplications += p->multiplications;
  s.leaves += p->leaves;
  s.peak_mb = p->work_bytes / 1048576.0;
  s.exec_pattern = p->exec_pattern;
  return CAPS_OK;
}

int caps_multiply(const double *A, const double *B, double *C, const caps_desc *desc,
                  const char *pattern, const caps_opts *opts) {
  State &s = st();
  std::lock_guard<std::mutex> g(s.mu);
  int rc = ensure_device_locked(s);
  if (rc) return rc;
  rc = check_desc(desc);
  if (rc) return rc;
  if (!A || !B || !C) return fail(CAPS_EINVAL, "NULL matrix pointer");
  const size_t n2 = (size_t)desc->lda * desc->lda;
  rc = stage_locked(s, n2);
  if (rc) return rc;
  hipStream_t ss = s.stage_stream;
  HIP_TRY(hipMemcpyAsync(s.stage[0], A, n2 * sizeof(double), hipMemcpyHostToDevice, ss));
  HIP_TRY(hipMemcpyAsync(s.stage[1], B, n2 * sizeof(double), hipMemcpyHostToDevice, ss));
  rc = multiply_device_locked(s, s.stage[0], s.stage[1], s.stage[2], desc, pattern, opts, ss);
  if (rc) {
    hipStreamSynchronize(ss);
    return rc;
  }
  HIP_TRY(hipMemcpyAsync(C, s.stage[2], n2 * sizeof(double), hipMemcpyDeviceToHost, ss));
  HIP_TRY(hipStreamSynchronize(ss));
  return CAPS_OK;
}

static int layout_device(const double *src, double *dst, const caps_desc *desc, void *stream,
                         bool to_caps) {
  State &s = st();
  std::lock_guard<std::mutex> g(s.mu);
  int rc = ensure_device_locked(s);
  if (rc) return rc;
  rc = check_desc(desc);
  if (rc) return rc;
  if (!src || !dst) return fail(CAPS_EINVAL, "NULL matrix pointer");
  if (src == dst) return fail(CAPS_EINVAL, "layout conversion cannot run in place");
  hipError_t e = to_caps ? launch_to_caps(src, dst, desc->lda, desc->nrec, desc->bs,
                                          (hipStream_t)stream)
                         : launch_from_caps(src, dst, desc->lda, desc->nrec, desc->bs,
                                            (hipStream_t)stream);
  if (e != hipSuccess) return fail(CAPS_EHIP, "layout kernel: %s", hipGetErrorString(e));
  return CAPS_OK;
}

int caps_to_layout_device(const double *dcm, double *dcaps, const caps_desc *desc, void *stream) {
  return layout_device(dcm, dcaps, desc, stream, true);
}

int caps_from_layout_device(const double *dcaps, double *dcm, const caps_desc *desc,
                            void *stream) {
  return layout_device(dcaps, dcm, desc, stream, false);
}

static int layout_host(const double *src, double *dst, const caps_desc *desc, bool to_caps) {
  State &s = st();
  std::lock_guard<std::mutex> g(s.mu);
  int rc = ensure_device_locked(s);
  if (rc) return rc;
  rc = check_desc(desc);
  if (rc) return rc;
  if (!src || !dst) return fail(CAPS_EINVAL, "NULL matrix pointer");
  const size_t n2 = (size_t)desc->lda * desc->lda;
  rc = stage_locked(s, n2);
  if (rc) return rc;
  hipStream_t ss = s.stage_stream;
  HIP_TRY(hipMemcpyAsync(s.stage[0], src, n2 * sizeof(double), hipMemcpyHostToDevice, ss));
  hipError_t e = to_caps ? launch_to_caps(s.stage[0], s.stage[1], desc->lda, desc->nrec,
                                          desc->bs, ss)
                         : launch_from_caps(s.stage[0], s.stage[1], desc->lda, desc->nrec,
                                            desc->bs, ss);
  if (e != hipSuccess) return fail(CAPS_EHIP, "layout kernel: %s", hipGetErrorString(e));
  HIP_TRY(hipMemcpyAsync(dst, s.stage[1], n2 * sizeof(double), hipMemcpyDeviceToHost, ss));
  HIP_TRY(hipStreamSynchronize(ss));
  return CAPS_OK;
}

int caps_to_layout(const double *cm, double *caps, const caps_desc *desc) {
  return layout_host(cm, caps, desc, true);
}

int caps_from_layout(const double *caps, double *cm, const caps_desc *desc) {
  return layout_host(caps, cm, desc, false);
}

int caps_get_counters(caps_counters *out) {
  if (!out) return fail(CAPS_EINVAL, "NULL counters");
  State &s = st();
  std::lock_guard<std::mutex> g(s.mu);
  memset(out, 0, sizeof(*out));
  out->additions = s.additions;
  out->multiplications = s.multiplications;
  out->leaf_products = s.leaves;
  out->peak_workspace_mb = s.peak_mb;
  snprintf(out->pattern, sizeof(out->pattern), "%s", s.exec_pattern.c_str());
  for (int k = 0; k < CAPS_NUM_TIMERS; k++) {
    out->timers_ms[k] = s.timers_ms[k];
    out->timed_launches[k] = s.timed[k];
  }
  return CAPS_OK;
}

int caps_reset_counters(void) {
  State &s = st();
  std::lock_guard<std::mutex> g(s.mu);
  s.additions = s.multiplications = s.leaves = 0;
  s.peak_mb = 0;
  s.exec_pattern.clear();
  for (int k = 0; k < CAPS_NUM_TIMERS; k++) {
    s.timers_ms[k] = 0;
    s.timed[k] = 0;
  }
  return CAPS_OK;
}

int caps_set_available_memory(long long doubles) {
  if (doubles < 0) return fail(CAPS_EINVAL, "available memory must not be negative");
  State &s = st();
  std::lock_guard<std::mutex> g(s.mu);
  s.budget_set = doubles > 0;
  s.available_doubles = doubles > 0 ? doubles : 1LL << 28;
  return CAPS_OK;
}

int caps_multiply_batch_device(int count, const double *const *dA, const double *const *dB,
                               double *const *dC, const caps_desc *descs,
                               const char *const *patterns, const caps_opts *opts,
                               void *stream) {
  State &s = st();
  std::lock_guard<std::mutex> g(s.mu);
  int rc = ensure_device_locked(s);
  if (rc) return rc;
  if (count <= 0 || !dA || !dB || !dC || !descs)
    return fail(CAPS_EINVAL, "caps_multiply_batch_device: bad arguments");
  // opts: NULL (defaults for every item) or one caps_opts per item; the
  // items share one leaf mode and one modulus count
  int mode = CAPS_MODE_FP64, moduli = 0;
  for (int i = 0; i < count; i++) {
    const caps_opts *oi = opts ? &opts[i] : nullptr;
    if (oi && oi->mode == CAPS_MODE_AUTO)
      return fail(CAPS_EUNSUPPORTED, "batched calls take an explicit leaf mode");
    int mi_, ni_;
    rc = resolve_mode(descs[i], oi, mi_, ni_);
    if (rc) return rc;
    if (i == 0) {
      mode = mi_;
      moduli = ni_;
    } else if (mi_ != mode || ni_ != moduli) {
      return fail(CAPS_EINVAL, "batch item %d: every item must use the same mode and moduli", i);
    }
  }
  std::vector<BatchItem> items((size_t)count);
  // items of one group are batched together: each NULL-pattern item gets an
  // equal share of the budget
  const long long share =
      (long long)(planner_budget_bytes(s.available_doubles, s.budget_set) / sizeof(double)) /
      count;
  std::string key;
  key.reserve((size_t)count * 64);
  for (int i = 0; i < count; i++) {
    rc = check_desc(&descs[i]);
    if (rc) return rc;
    if (!dA[i] || !dB[i] || !dC[i]) return fail(CAPS_EINVAL, "NULL matrix pointer in item %d", i);
    const int cutoff = opts ? opts[i].cutoff : 0;
    rc = check_cutoff(descs[i], mode, cutoff);
    if (rc) return rc;
    if (residue_mode(mode) && descs[i].nrec > 7)
      return fail(CAPS_EUNSUPPORTED, "residue-domain mode supports nrec <= 7");
    items[i].A = dA[i];
    items[i].B = dB[i];
    items[i].C = dC[i];
    items[i].desc = descs[i];
    items[i].cutoff = cutoff;
    caps_opts o{mode, moduli, 0, cutoff};
    rc = resolve_pattern(descs[i], patterns ? patterns[i] : nullptr, &o, items[i].pattern, mode,
                         moduli, share > 0 ? share : 1, true);
    if (rc) return rc;
    char buf[176];
    snprintf(buf, sizeof(buf), "%p %p %p %d %d %d %d %s;", (const void *)dA[i],
             (const void *)dB[i], (void *)dC[i], descs[i].lda, descs[i].nrec, descs[i].bs, cutoff,
             items[i].pattern.c_str());
    key += buf;
  }
  const int tile = (opts && (opts[0].flags & CAPS_FLAG_SMALL_TILE)) ? 1 : 0;
  // merged groups (segmented launches) unless CAPS_FLAG_NO_FUSION, within the
  // memory budget
  const bool merge = !(opts && (opts[0].flags & CAPS_FLAG_NO_FUSION));
  const size_t merge_budget = merge ? planner_budget_bytes(s.available_doubles, s.budget_set) : 0;
  PlanKey pk{nullptr, nullptr, nullptr, descs[0], std::string(), mode, moduli, 0, merge, key,
             stream, tile};
  Plan *plan = nullptr;
  for (auto &p : s.plans)
    if (p->key == pk) {
      plan = p.get();
      p->last_use = ++s.clock;
    }
  if (!plan) {
    rc = make_plan_locked(
        s, pk,
        [&](Plan &pl, void *work) {
          return build_batch_plan(pl, items, mode, moduli, work, tile, merge_budget);
        },
        plan);
    if (rc) return rc;
  }
  const bool use_graph = !(opts && (opts->flags & (CAPS_FLAG_NO_GRAPH | CAPS_FLAG_TIMERS))) &&
                         plan->steps.size() <= MAX_GRAPH_STEPS;
  if (opts && (opts->flags & CAPS_FLAG_TIMERS)) {
    hipError_t e = run_steps_timed(*plan, (hipStream_t)stream, s.timers_ms, s.timed);
    if (e != hipSuccess) return fail(CAPS_EHIP, "timed launch: %s", hipGetErrorString(e));
  } else if (use_graph) {
    if (!plan->graph) {
      rc = capture_graph(*plan);
      if (rc) return rc;
    }
    HIP_TRY(hipGraphLaunch(plan->graph, (hipStream_t)stream));
  } else {
    hipError_t e = run_steps(*plan, (hipStream_t)stream);
    if (e != hipSuccess) return fail(CAPS_EHIP, "kernel launch: %s", hipGetErrorString(e));
  }
  s.additions += plan->additions;
  s.multiplications += plan->multiplications;
  s.leaves += plan->leaves;
  s.peak_mb = plan->work_bytes / 1048576.0;
  s.exec_pattern = plan->exec_pattern;
  return CAPS_OK;
}

int caps_reference_product(const double *dA, const double *dB, double *dRef,
                           const caps_desc *desc, void *stream) {
  State &s = st();
  std::lock_guard<std::mutex> g(s.mu);
  int rc = ensure_device_locked(s);
  if (rc) return rc;
  rc = check_desc(desc);
  if (rc) return rc;
  // CRT mode with 16 moduli: exact product of operands truncated ~63 bits
  // below their row / column norms
  const int nrec = desc->nrec < 7 ? desc->nrec : 7;
  caps_desc d = *desc;
  if (nrec != desc->nrec) return fail(CAPS_EUNSUPPORTED, "reference product needs nrec <= 7");
  caps_opts o{CAPS_MODE_CRT_INT8, 16, 0, 0};
  std::string pat((size_t)d.nrec, 'd');
  return multiply_device_locked(s, dA, dB, dRef, &d, pat.c_str(), &o, (hipStream_t)stream);
}

int caps_normwise_error(const double *dC, const double *dRef, const caps_desc *desc, double *err,
                        void *stream) {
  if (!dC || !dRef || !desc || !err) return fail(CAPS_EINVAL, "caps_normwise_error: NULL");
  State &s = st();
  std::lock_guard<std::mutex> g(s.mu);
  int rc = ensure_device_locked(s);
  if (rc) return rc;
  HIP_TRY(normwise_error(dC, dRef, (long long)desc->lda * desc->lda, err, (hipStream_t)stream));
  return CAPS_OK;
}

int caps_measure(const double *dA, const double *dB, double *dC, const caps_desc *desc,
                 const double *dRef, int reps, caps_candidate *cand, void *stream) {
  if (!cand || reps < 1) return fail(CAPS_EINVAL, "caps_measure: bad arguments");
  State &s = st();
  std::lock_guard<std::mutex> g(s.mu);
  int rc = ensure_device_locked(s);
  if (rc) return rc;
  rc = check_desc(desc);
  if (rc) return rc;
  hipStream_t hs = (hipStream_t)stream;
  caps_opts o{cand->mode, cand->moduli, cand->tile ? CAPS_FLAG_SMALL_TILE : 0, cand->cutoff};
  std::string pat;
  if (cand->pattern[0]) {
    pat = cand->pattern;
  } else {
    int mode = cand->mode, moduli = cand->moduli ? cand->moduli : OZ_DEFAULT_MODULI;
    rc = resolve_pattern(*desc, nullptr, &o, pat, mode, moduli, s.available_doubles,
                         s.budget_set);
    if (rc) return rc;
    snprintf(cand->pattern, sizeof(cand->pattern), "%s", pat.c_str());
  }
  rc = multiply_device_locked(s, dA, dB, dC, desc, pat.c_str(), &o, hs);  // plan + graph
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(hs));
  std::vector<hipEvent_t> ev(2 * reps);
  for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
  for (int r = 0; r < reps && !rc; r++) {
    HIP_TRY(hipEventRecord(ev[2 * r], hs));
    rc = multiply_device_locked(s, dA, dB, dC, desc, pat.c_str(), &o, hs);
    HIP_TRY(hipEventRecord(ev[2 * r + 1], hs));
  }
  HIP_TRY(hipEventSynchronize(ev[2 * reps - 1]));
  std::vector<double> t(reps);
  for (int r = 0; r < reps; r++) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ev[2 * r], ev[2 * r + 1]));
    t[r] = ms;
  }
  for (auto &e : ev) hipEventDestroy(e);
  if (rc) return rc;
  std::sort(t.begin(), t.end());
  cand->ms = t[reps / 2];
  cand->err = -1.0;
  if (dRef) {
    double err = 0;
    HIP_TRY(normwise_error(dC, dRef, (long long)desc->lda * desc->lda, &err, hs));
    cand->err = err;
  }
  return CAPS_OK;
}

int caps_plan_choice(const caps_desc *desc, const caps_opts *opts, caps_candidate *out) {
  if (!out) return fail(CAPS_EINVAL, "caps_plan_choice: NULL output");
  State &s = st();
  std::lock_guard<std::mutex> g(s.mu);
  int rc = check_desc(desc);
  if (rc) return rc;
  int mode, moduli, cutoff;
  std::string pat;
  double pred[2] = {-1.0, -1.0};
  int tile = (opts && (opts->flags & CAPS_FLAG_SMALL_TILE)) ? 1 : 0;
  if (opts && opts->mode == CAPS_MODE_AUTO) {
    int tuned_tile = 0;
    rc = auto_choose(s, *desc, nullptr, mode, moduli, cutoff, pat, pred, &tuned_tile);
    tile |= tuned_tile;
  } else {
    rc = resolve_mode(*desc, opts, mode, moduli);
    cutoff = opts ? opts->cutoff : 0;
    if (!rc) rc = resolve_pattern(*desc, nullptr, opts, pat, mode, moduli, s.available_doubles,
                                  s.budget_set);
  }
  if (!rc) rc = check_cutoff(*desc, mode, cutoff);
  if (rc) return rc;
  memset(out, 0, sizeof(*out));
  out->mode = mode;
  out->moduli = moduli;
  out->cutoff = cutoff;
  out->tile = tile;
  snprintf(out->pattern, sizeof(out->pattern), "%s", pat.c_str());
  out->ms = pred[0];
  out->err = pred[1];
  return CAPS_OK;
}

int caps_planner_config(double max_err, int mode_mask, double exponent_spread) {
  if (!(max_err > 0) || (mode_mask & ~(CAPS_MASK_FP64 | CAPS_MASK_OZAKI_INT8 | CAPS_MASK_CRT_INT8 |
                                        CAPS_MASK_CRT_BF16)) ||
      !mode_mask)
    return fail(CAPS_EINVAL, "caps_planner_config: bad error target or mode mask");
  State &s = st();
  std::lock_guard<std::mutex> g(s.mu);
  s.auto_max_err = max_err;
  s.auto_mask = mode_mask;
  s.auto_spread = exponent_spread;
  return CAPS_OK;
}

int caps_tune_save(const char *path) {
  if (!path) return fail(CAPS_EINVAL, "NULL path");
  return tune_save(path);
}

int caps_tune_load(const char *path) {
  if (!path) return fail(CAPS_EINVAL, "NULL path");
  return tune_load(path);
}

int caps_release(void) {
  State &s = st();
  std::lock_guard<std::mutex> g(s.mu);
  evict_locked(s, 0);
  s.host_plans.clear();
  for (auto &p : s.stage)
    if (p) {
      hipFree(p);
      p = nullptr;
    }
  s.stage_elems = 0;
  return CAPS_OK;
}

}  // extern "C"
