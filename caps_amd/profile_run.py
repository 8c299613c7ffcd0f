"""A short, profiler-friendly run of the multiply: a few calls of one mode on
device-resident inputs, then a plain device-to-device copy of known size
(the calibration for hardware byte counters such as FETCH_SIZE and
WRITE_SIZE: the copy moves exactly --copy-mib MiB in and out).

    python -m caps_amd.profile_run --n 8192 --nrec 2 --pattern bb --mode crt --calls 3
    python -m caps_amd.profile_run --n 8192 --nrec 2 --bs 8 --mode layout --calls 20
    python -m caps_amd.profile_run --mode batch_crt --calls 5
    rocprofv3 --kernel-trace --stats --output-format csv -d OUT -o run -- \\
        python3 -m caps_amd.profile_run --mode crt --calls 3

Inputs are u * 2^e, u uniform [-1, 1), e uniform integer in [-10, 10].
"""
from __future__ import annotations

import argparse
import sys

import caps_amd

MODES = {"fp64": caps_amd.MODE_FP64, "crt": caps_amd.MODE_CRT_INT8,
         "crt_bf16": caps_amd.MODE_CRT_BF16, "leaf": caps_amd.MODE_OZAKI_INT8, "layout": None,
         "batch_fp64": caps_amd.MODE_FP64, "batch_crt": caps_amd.MODE_CRT_INT8}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--nrec", type=int, default=2)
    ap.add_argument("--pattern", default="bb")
    ap.add_argument("--bs", type=int, default=1)
    ap.add_argument("--mode", choices=sorted(MODES), default="crt")
    ap.add_argument("--moduli", type=int, default=14)
    ap.add_argument("--calls", type=int, default=3)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the plan's kernels one by one instead of as a graph")
    a = ap.parse_args(argv)

    import torch
    if not torch.cuda.is_available():
        print("profile_run: no GPU", file=sys.stderr)
        return 2
    if a.mode.startswith("batch"):
        run_batch(torch, a)
    else:
        run_one(torch, a)
    if a.copy_mib > 0:
        words = a.copy_mib * (1 << 20) // 8
        src = torch.ones(words, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[profile_run] copy {a.copy_mib} MiB", file=sys.stderr, flush=True)
    caps_amd.release()
    return 0


def run_one(torch, a) -> None:
    n = a.n
    g = torch.Generator(device="cuda")
    g.manual_seed(1)

    def rand():
        u = torch.rand(n * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
        e = torch.randint(-10, 11, (n * n,), generator=g, device="cuda", dtype=torch.int32)
        return torch.ldexp(u, e.to(torch.float64))

    dA, dB = rand(), rand()
    dC = torch.empty_like(dA)
    desc = caps_amd.MatDescriptor.one_device(n, a.nrec, a.bs)
    mode = MODES[a.mode]
    moduli = 0 if mode == caps_amd.MODE_FP64 else a.moduli
    flags = caps_amd.FLAG_NO_GRAPH if a.no_graph else 0
    for i in range(a.calls):
        if mode is None:
            # column-major -> CAPS layout and back (16 bytes per element each way)
            caps_amd.to_layout_device(dA, dC, desc)
            caps_amd.from_layout_device(dC, dB, desc)
        else:
            caps_amd.multiply_device(dA, dB, dC, desc, a.pattern, mode=mode, moduli=moduli,
                                     flags=flags)
        torch.cuda.synchronize()
        print(f"[profile_run] {a.mode} n={n} nrec={a.nrec} pattern={a.pattern} call {i + 1}",
              file=sys.stderr, flush=True)


def run_batch(torch, a) -> None:
    """The 512-problem uneven batch (n = 64 round(exp(U[ln 128, ln 4096]) / 64),
    seed 2024, uniform [-1, 1)) with the shipped learned planner's plans, as
    one caps_multiply_batch_device call per replay."""
    import os
    from caps_amd import batch_plan, planner_train
    mode = MODES[a.mode]
    planner_train.load_model(os.path.join(os.path.dirname(__file__), "planner_model.npz"))
    sizes = batch_plan.uneven_sizes()
    err = 1e-12 if mode == caps_amd.MODE_FP64 else 1e-14
    plans = batch_plan.choose_plans(sizes, mode, max_err=err, spread=0.0,
                                    moduli=None if mode == caps_amd.MODE_FP64 else a.moduli)
    g = torch.Generator(device="cuda")
    g.manual_seed(77)
    As, Bs, Cs, descs, pats, cuts = [], [], [], [], [], []
    for n in sizes:
        p = plans[n]
        As.append(torch.rand(n * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1)
        Bs.append(torch.rand(n * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1)
        Cs.append(torch.empty_like(As[-1]))
        descs.append(caps_amd.MatDescriptor.one_device(n, p.nrec, 1))
        pats.append(p.pattern)
        cuts.append(p.cutoff)
    K = next(iter(plans.values())).moduli
    for i in range(a.calls):
        caps_amd.multiply_batch_device(As, Bs, Cs, descs, pats, mode, K, cutoffs=cuts)
        torch.cuda.synchronize()
        print(f"[profile_run] {a.mode} 512 problems call {i + 1}", file=sys.stderr, flush=True)
    caps_amd.MLP.clear_planner()


if __name__ == "__main__":
    sys.exit(main())
