// Greedy BFS/DFS choice under a memory budget.  The reference's automatic
// rule (multiply.cpp:43-66 with enoughMemory, memory.cpp:22-29) takes a BFS
// step only while the workspace it needs fits the budget set by
// setAvailableMemory (memory.cpp:31-33; drivers pass -m / -k,
// randombenchmark.cpp:45-48), and only with several processes.  On one device
// a 'b' level is batching, limited by workspace alone, so the same budget
// decides it: a level becomes 'b' when the whole plan's device workspace still
// fits.
#include "planner.h"

#include <hip/hip_runtime.h>

#include "engine.h"

namespace caps {

size_t planner_budget_bytes(long long budget_doubles, bool budget_set) {
  size_t freeb = 0, totalb = 0;
  const bool have_free = hipMemGetInfo(&freeb, &totalb) == hipSuccess;
  if (budget_set) {
    // the caller's budget, capped at what the device has free
    size_t b = (size_t)budget_doubles * sizeof(double);
    return have_free && b > freeb ? freeb : b;
  }
  // no budget set: half of the free device memory (the reference's default
  // of 2^28 doubles, memory.cpp:4, when no device answers)
  return have_free ? freeb / 2 : (size_t)(1LL << 28) * sizeof(double);
}

std::string choose_pattern(const caps_desc &d, int mode, int moduli, long long budget_doubles,
                           bool budget_set) {
  std::string pat((size_t)d.nrec, 'd');
  const size_t budget = planner_budget_bytes(budget_doubles, budget_set);
  for (int l = 0; l < d.nrec; l++) {
    std::string trial = pat;
    trial[(size_t)l] = 'b';
    if (plan_workspace_bytes(d, trial, mode, moduli, 0) <= budget) pat = trial;
  }
  return pat;
}

int choose_mode(const caps_desc &) { return CAPS_MODE_FP64; }

}  // namespace caps
