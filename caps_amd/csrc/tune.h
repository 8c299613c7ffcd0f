// Autotuner support (tune.hip): device error measurement and the tuning table.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/caps_amd.h"

namespace caps {

struct TuneEntry {
  int lda, nrec, bs, mode_mask;
  double max_err;
  caps_candidate c;
};

// max|x - y| / max|y| over n doubles, on stream s (synchronises)
hipError_t normwise_error(const double *x, const double *y, long long n, double *err,
                          hipStream_t s);

bool tune_lookup(const caps_desc &d, int mode_mask, double max_err, TuneEntry &out);
void tune_store(const TuneEntry &e);
int tune_save(const char *path);
int tune_load(const char *path);

}  // namespace caps
