// Autotuner: measurement of one configuration (time and error on the device)
// and the exhaustive search over recursion pattern x cutoff x leaf mode x
// moduli, with a tuning table the planner consults.  No counterpart in the
// reference (whose BFS/DFS choice is a memory rule, multiply.cpp:43-66).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/caps_amd.h"
#include "kernels.h"
#include "tune.h"

namespace caps {

int set_error(int code, const char *msg);

namespace {

constexpr int T = 256;

// max |x - y| and max |y| as order-independent integer maxima of the FP64
// bit patterns (non-negative doubles order like their bits)
__global__ __launch_bounds__(T) void normwise_kernel(const double *__restrict__ x,
                                                     const double *__restrict__ y, long long n,
                                                     unsigned long long *out) {
  unsigned long long md = 0, my = 0;
  for (long long i = (long long)blockIdx.x * T + threadIdx.x; i < n;
       i += (long long)gridDim.x * T) {
    const double d = fabs(x[i] - y[i]);
    const double a = fabs(y[i]);
    const unsigned long long bd = (d != d) ? 0x7ff8000000000000ull : __double_as_longlong(d);
    md = max(md, bd);
    my = max(my, (unsigned long long)__double_as_longlong(a));
  }
  for (int off = 32; off >= 1; off >>= 1) {
    md = max(md, (unsigned long long)__shfl_down((long long)md, off, 64));
    my = max(my, (unsigned long long)__shfl_down((long long)my, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(out, md);
    atomicMax(out + 1, my);
  }
}

struct Table {
  std::mutex mu;
  std::vector<TuneEntry> entries;
};

Table &table() {
  static Table t;
  return t;
}

}  // namespace

hipError_t normwise_error(const double *x, const double *y, long long n, double *err,
                          hipStream_t s) {
  unsigned long long *d = nullptr;
  hipError_t e = hipMallocAsync(reinterpret_cast<void **>(&d), 2 * sizeof(unsigned long long), s);
  if (e != hipSuccess) return e;
  // cleared by a kernel-free copy from a zeroed host pair (no memset node)
  static const unsigned long long zero[2] = {0, 0};
  e = hipMemcpyAsync(d, zero, sizeof(zero), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    long long blocks = std::min<long long>((n + T - 1) / T, (long long)device_cu_count() * 8);
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(normwise_kernel, dim3((unsigned)blocks), dim3(T), 0, s, x, y, n, d);
    e = hipGetLastError();
  }
  unsigned long long h[2] = {0, 0};
  if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  hipFreeAsync(d, s);
  if (e != hipSuccess) return e;
  double md, my;
  memcpy(&md, &h[0], 8);
  memcpy(&my, &h[1], 8);
  *err = my > 0 ? md / my : (md > 0 ? INFINITY : 0.0);
  return hipSuccess;
}

bool tune_lookup(const caps_desc &d, int mode_mask, double max_err, TuneEntry &out) {
  Table &t = table();
  std::lock_guard<std::mutex> g(t.mu);
  for (const TuneEntry &e : t.entries)
    if (e.lda == d.lda && e.nrec == d.nrec && e.bs == d.bs && e.mode_mask == mode_mask &&
        e.max_err == max_err) {
      out = e;
      return true;
    }
  return false;
}

void tune_store(const TuneEntry &e) {
  Table &t = table();
  std::lock_guard<std::mutex> g(t.mu);
  for (TuneEntry &x : t.entries)
    if (x.lda == e.lda && x.nrec == e.nrec && x.bs == e.bs && x.mode_mask == e.mode_mask &&
        x.max_err == e.max_err) {
      x = e;
      return;
    }
  t.entries.push_back(e);
}

int tune_save(const char *path) {
  Table &t = table();
  std::lock_guard<std::mutex> g(t.mu);
  FILE *f = fopen(path, "w");
  if (!f) return set_error(CAPS_EINVAL, "caps_tune_save: cannot open file");
  fprintf(f, "# lda nrec bs mode_mask max_err -> mode moduli cutoff pattern ms err tile\n");
  for (const TuneEntry &e : t.entries)
    fprintf(f, "%d %d %d %d %.17g %d %d %d %s %.9g %.9g %d\n", e.lda, e.nrec, e.bs, e.mode_mask,
            e.max_err, e.c.mode, e.c.moduli, e.c.cutoff, e.c.pattern[0] ? e.c.pattern : "-",
            e.c.ms, e.c.err, e.c.tile);
  fclose(f);
  return CAPS_OK;
}

// one line of a tuning file: every field in range, or false
bool valid_entry(const TuneEntry &e, const char *pat) {
  if (e.lda <= 0 || e.nrec < 0 || e.nrec > 24 || e.bs <= 0 || !(e.max_err > 0)) return false;
  if (e.c.mode != CAPS_MODE_FP64 && e.c.mode != CAPS_MODE_OZAKI_INT8 &&
      e.c.mode != CAPS_MODE_CRT_INT8 && e.c.mode != CAPS_MODE_CRT_BF16)
    return false;
  if (e.c.mode == CAPS_MODE_FP64 ? e.c.moduli != 0
                                 : (e.c.moduli != 0 && (e.c.moduli < 2 || e.c.moduli > 16)))
    return false;
  if (e.c.cutoff < 0 || e.c.cutoff > e.nrec) return false;
  if (e.c.tile != 0 && e.c.tile != 1) return false;
  if (strcmp(pat, "-") == 0) return true;
  if ((int)strlen(pat) != e.nrec) return false;
  for (const char *c = pat; *c; c++)
    if (!strchr("bBdDhH", *c)) return false;
  return true;
}

// All-or-nothing: a file with any malformed entry loads nothing.  The tile
// column is optional (0 when absent).
int tune_load(const char *path) {
  FILE *f = fopen(path, "r");
  if (!f) return set_error(CAPS_EINVAL, "caps_tune_load: cannot open file");
  char line[512];
  std::vector<TuneEntry> got;
  int lineno = 0;
  while (fgets(line, sizeof(line), f)) {
    lineno++;
    const char *p = line;
    while (*p == ' ' || *p == '\t') p++;
    if (*p == '#' || *p == '\n' || *p == '\0') continue;
    TuneEntry e{};
    char pat[64] = {0};
    const int got_fields =
        sscanf(p, "%d %d %d %d %lg %d %d %d %63s %lg %lg %d", &e.lda, &e.nrec, &e.bs,
               &e.mode_mask, &e.max_err, &e.c.mode, &e.c.moduli, &e.c.cutoff, pat, &e.c.ms,
               &e.c.err, &e.c.tile);
    if (got_fields < 11 || !valid_entry(e, pat)) {
      fclose(f);
      char msg[96];
      snprintf(msg, sizeof(msg), "caps_tune_load: malformed entry on line %d", lineno);
      return set_error(CAPS_EINVAL, msg);
    }
    snprintf(e.c.pattern, sizeof(e.c.pattern), "%s", strcmp(pat, "-") ? pat : "");
    got.push_back(e);
  }
  fclose(f);
  for (const TuneEntry &e : got) tune_store(e);
  return CAPS_OK;
}

}  // namespace caps
