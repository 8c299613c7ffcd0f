// Greedy BFS/DFS choice under a memory budget.  The reference's automatic
// rule (multiply.cpp:43-66 with enoughMemory, memory.cpp:22-29) only takes a
// BFS step when there are several processes; on one device the analogue is
// batching, limited by workspace alone.
#include "planner.h"

#include <hip/hip_runtime.h>

#include "engine.h"

namespace caps {

std::string choose_pattern(const caps_desc &d, int mode, int moduli, long long budget_doubles) {
  std::string pat((size_t)d.nrec, 'd');
  size_t freeb = 0, totalb = 0;
  size_t budget = (size_t)budget_doubles * sizeof(double);
  if (hipMemGetInfo(&freeb, &totalb) == hipSuccess && freeb / 2 > budget) budget = freeb / 2;
  for (int l = 0; l < d.nrec; l++) {
    std::string trial = pat;
    trial[(size_t)l] = 'b';
    if (plan_workspace_bytes(d, trial, mode, moduli, 0) <= budget) pat = trial;
  }
  return pat;
}

int choose_mode(const caps_desc &) { return CAPS_MODE_FP64; }

}  // namespace caps
