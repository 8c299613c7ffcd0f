// Plan choice when the caller passes no pattern (or mode CAPS_MODE_AUTO).
#pragma once
#include <string>

#include "../../include/caps_amd.h"

namespace caps {

// Workspace budget in bytes: the caller's (budget_set, capped at the free
// device memory) or, by default, half of the free device memory.
size_t planner_budget_bytes(long long budget_doubles, bool budget_set);

// One char per level: 'b' (batched) at every level whose workspace still fits
// the budget, 'd' below.  On one device a 'b' level only costs memory
// ((7/4)^levels growth), never communication, so batching is preferred.
std::string choose_pattern(const caps_desc &d, int mode, int moduli, long long budget_doubles,
                           bool budget_set);

// Leaf arithmetic when the caller asks for CAPS_MODE_AUTO.
int choose_mode(const caps_desc &d);

// The learned planner (tune_plan.cpp): false when no model is installed or no
// candidate is predicted to meet max_err.  pred (may be NULL) receives the
// model's predicted milliseconds and normwise error of the choice.
bool mlp_planner_choose(const caps_desc &d, int mask, double max_err, double spread, int &mode,
                        int &moduli, int &cutoff, std::string &pat, double *pred = nullptr);

// Uninstall m if it is the installed planner model (caps_mlp_destroy calls it).
void planner_forget_model(const caps_mlp *m);

}  // namespace caps
