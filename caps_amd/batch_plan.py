"""Per-problem plans for uneven batches (caps_multiply_batch_device).

The reference multiplies one matrix per call and picks BFS or DFS per level
from its memory rule (multiply.cpp:43-66).  A batch of independent problems
of different sizes needs a layout depth, a b/d pattern and a cutoff for every
size; here the learned cost / error model proposes them (one forward pass per
candidate descriptor through caps_plan_choice) and, optionally, the device
confirms the proposal by timing the model's best choice at each depth
(caps_measure).  Without an installed model the rule is the deepest layout
whose leaves are at least 256 (or the deepest with leaves >= 64), every level
batched.

    sizes = uneven_sizes()                      # 512 sizes, seed 2024
    plans = choose_plans(sizes, MODE_CRT_INT8)  # {n: Plan}
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

import caps_amd


@dataclass
class Plan:
    nrec: int
    pattern: str
    cutoff: int
    moduli: int
    predicted_ms: float = -1.0
    measured_ms: float = -1.0


def uneven_sizes(count: int = 512, seed: int = 2024, lo: int = 128, hi: int = 4096,
                 mult: int = 64) -> list:
    """n = mult * round(exp(U[ln lo, ln hi]) / mult), clipped to [lo, hi]."""
    rng = np.random.default_rng(seed)
    n = mult * np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), count)) / mult)
    return [int(x) for x in np.clip(n, lo, hi)]


def depths(n: int, min_leaf: int = 64) -> list:
    """Layout depths r with n % 2^r == 0 and leaves of at least min_leaf."""
    return [r for r in range(0, 12) if n % (1 << r) == 0 and (n >> r) >= min_leaf]


def default_plan(n: int, moduli: int) -> Plan:
    ds = depths(n, 256) or depths(n, 64) or [0]
    r = max(ds)
    return Plan(r, "b" * r, 0, moduli)


_MASKS = {caps_amd.MODE_FP64: caps_amd.MASK_FP64,
          caps_amd.MODE_CRT_INT8: caps_amd.MASK_CRT_INT8,
          caps_amd.MODE_CRT_BF16: caps_amd.MASK_CRT_BF16}


def choose_plans(sizes, mode: int, max_err: float = 1e-14, spread: float = 0.0,
                 measure_with=None, reps: int = 3, moduli: int | None = None) -> dict:
    """One Plan per distinct size.  The model (if installed, e.g. by
    planner_train.load_model) scores every candidate of every depth; the
    predicted-fastest depth wins.  With measure_with=torch the model's choice
    at each depth is also timed on the device and the measured-fastest wins.
    Split modes share one modulus count in a batch: `moduli` if given, else
    the largest chosen one."""
    caps_amd.planner_config(max_err, _MASKS[mode], spread)
    out = {}
    for n in sorted(set(int(x) for x in sizes)):
        best = None
        for r in depths(n):
            desc = caps_amd.MatDescriptor.one_device(n, r, 1)
            try:
                c = caps_amd.plan_choice(desc, caps_amd.MODE_AUTO)
            except caps_amd.CapsError:
                continue
            if c["mode"] != mode or c["predicted_ms"] < 0:
                continue  # no model, or it found nothing for this mode
            p = Plan(r, c["pattern"], c["cutoff"], moduli or c["moduli"], c["predicted_ms"])
            if measure_with is not None:
                p.measured_ms = _time_plan(measure_with, n, p, mode, reps)
            key = p.measured_ms if measure_with is not None else p.predicted_ms
            if best is None or key < (best.measured_ms if measure_with is not None
                                      else best.predicted_ms):
                best = p
        out[n] = best if best is not None else default_plan(n, caps_amd.DEFAULT_MODULI)
    if mode == caps_amd.MODE_FP64:
        for p in out.values():
            p.moduli = 0
    else:
        k = moduli if moduli else max(p.moduli for p in out.values())
        for p in out.values():
            p.moduli = k
    return out


def _time_plan(torch, n: int, p: Plan, mode: int, reps: int) -> float:
    desc = caps_amd.MatDescriptor.one_device(n, p.nrec, 1)
    A = torch.rand(n * n, device="cuda", dtype=torch.float64) * 2 - 1
    B = torch.rand(n * n, device="cuda", dtype=torch.float64) * 2 - 1
    C = torch.empty_like(A)
    cand = caps_amd.Candidate(mode, p.moduli if mode != caps_amd.MODE_FP64 else 0, p.cutoff, 0,
                              p.pattern.encode(), -1.0, -1.0)
    caps_amd.measure(A, B, C, desc, cand, None, reps)
    return cand.ms
