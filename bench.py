"""Benchmark: effective GFLOP/s (2 n^3 / t) of the Strassen-Winograd multiply
at n = 8192 on one MI355X, split precision (INT8 residues, CRT) against plain
FP64, with the accuracy gates of the metric.

Workload (one "step" = one multiply of device-resident CAPS-layout inputs):
  n = 8192, nrec = 2, pattern "bb" (two batched levels: 49 leaf products of
  2048^2 in one launch), bs = 1; A, B entries u * 2^e, u uniform [-1, 1),
  e uniform integer in [-10, 10] (fixed seeds).  The split run (residue-
  domain Strassen: INT8 residues of the scaled operands, exact Winograd
  additions modulo each modulus, INT8 MFMA leaf products, one CRT
  reconstruction) is the reported value; the plain FP64 run of the same
  multiply and the leaf-only split (FP64 additions) are reported beside it.
  After timing:
    * normwise error max|C - C*| / max|C*| over all of C against the device
      residue product with 16 moduli, and on 16 sampled rows against the
      compensated (Dot2) classical product from the CPU oracle,
    * an integer pair in [-64, 64] through the split path checked exactly
      (Freivalds' test over all of C plus sampled rows, exact in FP64).
Beside it (config): plain FP64 at n = 7168, r = 3, 'bdd'; the uneven batch
of 512 problems (n in [128, 4096], plans from the learned planner) in FP64
and split precision.  The roofline entry is for the leaf GEMM launch (INT8
MFMA), timed live with HIP events on the launch stream through the
library's phase timers (20 eager launches by default).  cpu_baseline is the
reference's own randombenchmark built from its sources (oracle/ref.mk) when
present, else the oracle's restatement.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batch]
                    [--workload split|batch]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_INT8_TOPS = 5000.0   # MI355X dense INT8 (TOP/s)
PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 matrix
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_inputs(torch, n, seed, integer=False):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    if integer:
        return torch.randint(-64, 65, (n * n,), generator=g, device="cuda",
                             dtype=torch.int32).to(torch.float64)
    u = torch.rand(n * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    e = torch.randint(-10, 11, (n * n,), generator=g, device="cuda", dtype=torch.int32)
    return torch.ldexp(u, e.to(torch.float64))


def time_steps(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def profiled_traffic(kernel, path=os.path.join(ROOT, "profiles", "crt_int8_n8192_bb.md")):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary
    (FETCH_SIZE and WRITE_SIZE runs, calibrated by a device copy), or None."""
    try:
        in_traffic = False
        for line in open(path):
            if line.startswith("## "):
                in_traffic = line.startswith("## HBM traffic")
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if in_traffic and len(cells) >= 4 and cells[0] == f"`{kernel}`":
                return round((float(cells[2]) + float(cells[3])) * 1048576)
    except (OSError, ValueError):
        return None
    return None


def leaf_kernel_ms(caps_amd, torch, dA, dB, dC, desc, pattern, mode, moduli, reps=3):
    """Average duration of the leaf GEMM launch, HIP events on its stream."""
    caps_amd.reset_counters()
    for _ in range(reps):
        caps_amd.multiply_device(dA, dB, dC, desc, pattern, mode=mode, moduli=moduli,
                                 flags=caps_amd.FLAG_TIMERS)
    torch.cuda.synchronize()
    c = caps_amd.counters()
    t, k = c["timers_ms"], c["timed_launches"]
    per = {name: (t[name] / k[name] if k[name] else 0.0) for name in t}
    per_call = {name: t[name] / reps for name in t}
    return per, per_call


def sampled_error(oracle, A_cm, B_cm, C_cm, rows):
    n = A_cm.shape[0]
    rr = np.repeat(rows, n).astype(np.int32)
    cc = np.tile(np.arange(n), len(rows)).astype(np.int32)
    ref = oracle.dot2_entries(A_cm, B_cm, rr, cc).reshape(len(rows), n)
    got = C_cm[rows, :]
    return float(np.abs(got - ref).max() / np.abs(ref).max())


def exact_integer_check(A_cm, B_cm, C_cm, rows, seed=11):
    """Freivalds over all of C plus exact sampled rows.  Entries <= 64, K <=
    8192: every partial sum stays below 2^53, so FP64 BLAS is exact here."""
    rng = np.random.default_rng(seed)
    x = rng.integers(-8, 9, (A_cm.shape[0], 2)).astype(np.float64)
    lhs = C_cm @ x
    rhs = A_cm @ (B_cm @ x)
    ok = bool(np.array_equal(lhs, rhs))
    ok &= bool(np.array_equal(C_cm[rows, :], A_cm[rows, :] @ B_cm))
    return ok


def cpu_threads():
    """Threads for the CPU leg: the host cores this process may use, capped by
    OMP_NUM_THREADS when set (a shared GPU host gives each GPU a share)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(avail, int(omp))) if omp and omp.isdigit() else avail


def cpu_baseline(oracle, n=8192, r=2):
    """The reference's own randombenchmark (built from its sources into
    oracle/_ref by oracle/ref.mk, MKL dgemm_) at the metric's size, one MPI
    process, all DFS levels (the reference's one-process path,
    multiply.cpp:43-66); else the oracle's restatement at n=4096."""
    threads = cpu_threads()
    host = {"host_cpu_count": os.cpu_count(), "threads_used": threads,
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}
    ref = oracle.run_reference_benchmark(n, r, "d" * r, threads)
    if ref is not None:
        return {"value": round(ref["effective_gflops"], 3), "unit": "GFLOP/s", "cores": threads,
                "kind": "reference",
                "sample": f"reference randombenchmark -s {n} -r {r} -p {'d' * r} (one process, "
                          f"NUM_THREADS=16 OpenMP additions, MKL dgemm_ with {threads} threads), "
                          f"{ref['time_s']:.2f} s; effective Gflop/s = 2n^3/t (counters.cpp:69)",
                **host}
    n = 4096
    g = oracle.Drand48()
    a = g.uniform(n * n)
    b = g.uniform(n * n)
    t0 = time.perf_counter()
    oracle.multiply(a, b, n, r, 1, "d" * r, threads)
    t = time.perf_counter() - t0
    return {"value": round(2.0 * n ** 3 / t / 1e9, 3), "unit": "GFLOP/s", "cores": threads,
            "kind": "port",
            "sample": f"one oracle multiply n={n} nrec={r} pattern 'dd' bs=1, uniform [-1,1) "
                      f"from drand48, {threads} OpenMP threads, {t:.2f} s", **host}


def run_uneven_batch(torch, caps_amd, sizes, plans, mode, steps, warmup, check=True,
                     compare_single=False):
    """512 independent problems (entries uniform [-1, 1)) in one batched plan,
    replayed as one graph: ms per batch, effective GFLOP/s (sum of 2n^3 / t)
    and the worst normwise error against the 16-modulus reference product."""
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

    def run_batch():
        caps_amd.multiply_batch_device(As, Bs, Cs, descs, pats, mode, K, cutoffs=cuts)

    eff = sum(2.0 * n ** 3 for n in sizes)
    t = time_steps(torch, run_batch, steps, warmup)
    out = {"ms": round(t * 1e3, 3), "gflops": round(eff / t / 1e9, 2), "moduli": K,
           "depths": sorted(set(p.nrec for p in plans.values())),
           "distinct_sizes": len(plans)}
    if compare_single:
        def run_each():
            for a, b, c, d, pt, k in zip(As, Bs, Cs, descs, pats, cuts):
                caps_amd.multiply_device(a, b, c, d, pt, mode=mode, moduli=K, cutoff=k)
        t1 = time_steps(torch, run_each, max(2, steps // 2), 1)
        out["ms_one_call_per_problem"] = round(t1 * 1e3, 3)
        out["gflops_one_call_per_problem"] = round(eff / t1 / 1e9, 2)
        run_batch()
    if check:
        torch.cuda.synchronize()
        worst = 0.0
        for i, (a, b, c, d) in enumerate(zip(As, Bs, Cs, descs)):
            ref = torch.empty_like(c)
            caps_amd.reference_product(a, b, ref, d)
            worst = max(worst, caps_amd.normwise_error(c, ref, d))
            del ref
            if i % 64 == 63:
                caps_amd.release()
        out["max_error"] = worst
        out["error_metric"] = ("max over problems of max|C-C*|/max|C*|, C* = 16-modulus "
                               "residue product (caps_reference_product)")
    del As, Bs, Cs
    caps_amd.release()
    torch.cuda.empty_cache()
    return out


def uneven_batch_section(args, torch, caps_amd, measure=False, compare_single=False):
    """Cfg4: 512 sizes n = 64 round(exp(U[ln 128, ln 4096]) / 64) (seed 2024);
    per-size depth, pattern and cutoff proposed by the shipped learned planner
    (measure=True: the proposal at every depth timed on the device, fastest
    kept); FP64 and residue-domain INT8 both run."""
    from caps_amd import batch_plan, planner_train
    model, _ = planner_train.load_model(os.path.join(ROOT, "caps_amd", "planner_model.npz"))
    sizes = batch_plan.uneven_sizes()
    res = {"workload": "512 independent multiplies, n = 64 round(exp(U[ln128, ln4096])/64), "
                       "seed 2024, uniform [-1, 1); one batched plan replayed as one graph",
           "sum_2n3": sum(2.0 * n ** 3 for n in sizes),
           "plans": "learned planner" + (" + device timing per depth" if measure else "")}
    for mode, name, err in ((caps_amd.MODE_FP64, "fp64", 1e-12),
                            (caps_amd.MODE_CRT_INT8, "split", 1e-14)):
        t0 = time.perf_counter()
        plans = batch_plan.choose_plans(sizes, mode, max_err=err, spread=0.0,
                                        measure_with=torch if measure else None,
                                        moduli=None if mode == caps_amd.MODE_FP64 else args.moduli)
        caps_amd.release()
        log(f"[bench] batch {name}: plans for {len(plans)} sizes in "
            f"{time.perf_counter() - t0:.1f} s")
        r = run_uneven_batch(torch, caps_amd, sizes, plans, mode, args.steps, args.warmup,
                             check=not args.no_check, compare_single=compare_single)
        r["planning_s"] = round(time.perf_counter() - t0, 2)
        log(f"[bench] batch {name}: {r['ms']:.2f} ms, {r['gflops']:.0f} GFLOP/s, "
            f"max error {r.get('max_error', float('nan')):.3e}")
        res[name] = r
    caps_amd.MLP.clear_planner()
    caps_amd.planner_config(1e-14, caps_amd.MASK_FP64 | caps_amd.MASK_CRT_INT8, 10.0)
    del model
    return res


def batch_workload(args, torch, caps_amd):
    """The uneven batch as its own bench line (--workload batch): plans
    confirmed on the device and compared with one call per problem."""
    res = uneven_batch_section(args, torch, caps_amd, measure=True, compare_single=True)
    result = {
        "metric": "effective GFLOP/s (sum of 2n^3 / t), 512 uneven problems n in [128, 4096]",
        "value": res["split"]["gflops"],
        "unit": "GFLOP/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res["split"]["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 in/out, int8 MFMA leaf products",
        "data": "synthetic: 512 sizes n = 64 round(exp(U[ln128, ln4096])/64), seed 2024; "
                "entries uniform [-1, 1)",
        "config": res,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(result), flush=True)


def fp64_bdd_workload(args, torch, caps_amd, oracle, n=7168, r=3, pat="bdd"):
    """Plain FP64 Strassen at n=7168, r=3, 'bdd' (the BFS level as 7 batched
    sub-products, then two recursive levels, FP64 MFMA leaves of 896^2), entries
    uniform [-1, 1): time per multiply, effective GFLOP/s and the normwise error
    of 16 sampled full rows against the Dot2 product."""
    log(f"[bench] plain FP64 n={n} nrec={r} pattern '{pat}'")
    desc = caps_amd.MatDescriptor.one_device(n, r, 1)
    g = torch.Generator(device="cuda")
    g.manual_seed(2024)
    A = torch.rand(n * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    B = torch.rand(n * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    C = torch.empty_like(A)

    def run():
        caps_amd.multiply_device(A, B, C, desc, pat, mode=caps_amd.MODE_FP64)

    t = time_steps(torch, run, args.steps, args.warmup)
    out = {"ms_per_step": round(t * 1e3, 4), "gflops": round(2.0 * n ** 3 / t / 1e9, 2)}
    log(f"[bench] plain FP64 {pat}: {t * 1e3:.3f} ms/step")
    if not args.no_check:
        rows = np.sort(np.random.default_rng(7).choice(n, 16, replace=False))
        out["error"] = sampled_error(oracle, oracle.from_caps(A.cpu().numpy(), n, r, 1),
                                     oracle.from_caps(B.cpu().numpy(), n, r, 1),
                                     oracle.from_caps(C.cpu().numpy(), n, r, 1), rows)
        out["error_metric"] = "max|C-C*|/max|C*| over 16 sampled full rows, C* = Dot2"
        log(f"[bench] plain FP64 {pat} error {out['error']:.3e}")
    del A, B, C
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=("split", "batch"), default="split",
                    help="split: the n=8192 headline; batch: 512 uneven problems")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--nrec", type=int, default=2)
    ap.add_argument("--pattern", default="bb")
    ap.add_argument("--moduli", type=int, default=14)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-fp64-bdd", action="store_true",
                    help="skip the plain FP64 n=7168 r=3 'bdd' measurement")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 512-problem uneven batch (config.batch)")
    ap.add_argument("--gemm-reps", type=int, default=20,
                    help="eager launches timed with HIP events for the roofline GEMM")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU BFS exchange is a design, not part of this build)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import caps_amd
    import oracle

    if not os.path.exists(caps_amd.LIB_PATH):
        from caps_amd import build
        build.build()
    torch.cuda.set_device(0)
    caps_amd.load()
    if args.workload == "batch":
        batch_workload(args, torch, caps_amd)
        return
    n, r, pat, K = args.n, args.nrec, args.pattern, args.moduli
    desc = caps_amd.MatDescriptor.one_device(n, r, 1)
    eff = 2.0 * n ** 3

    log(f"[bench] inputs n={n}")
    dA = make_inputs(torch, n, 1234)
    dB = make_inputs(torch, n, 5678)
    dC_split = torch.empty_like(dA)
    dC_plain = torch.empty_like(dA)
    dC_leaf = torch.empty_like(dA)
    SPLIT = caps_amd.MODE_CRT_INT8

    def run_split():
        caps_amd.multiply_device(dA, dB, dC_split, desc, pat, mode=SPLIT, moduli=K)

    def run_plain():
        caps_amd.multiply_device(dA, dB, dC_plain, desc, pat, mode=caps_amd.MODE_FP64)

    def run_leaf():
        caps_amd.multiply_device(dA, dB, dC_leaf, desc, pat, mode=caps_amd.MODE_OZAKI_INT8,
                                 moduli=K)

    dC_bf16 = torch.empty_like(dA)

    def run_bf16():
        caps_amd.multiply_device(dA, dB, dC_bf16, desc, pat, mode=caps_amd.MODE_CRT_BF16,
                                 moduli=K)

    log("[bench] split precision (residue-domain Strassen, INT8 MFMA, CRT)")
    t_split = time_steps(torch, run_split, args.steps, args.warmup)
    log(f"[bench] split {t_split * 1e3:.3f} ms/step")
    log("[bench] plain FP64")
    t_plain = time_steps(torch, run_plain, args.steps, args.warmup)
    log(f"[bench] plain {t_plain * 1e3:.3f} ms/step")
    log("[bench] leaf-only split (FP64 Winograd additions)")
    t_leaf = time_steps(torch, run_leaf, max(2, args.steps // 2), 1)
    log(f"[bench] leaf split {t_leaf * 1e3:.3f} ms/step")
    log("[bench] residue-domain Strassen with BF16 MFMA leaf products")
    t_bf16 = time_steps(torch, run_bf16, max(2, args.steps // 2), 1)
    log(f"[bench] bf16 variant {t_bf16 * 1e3:.3f} ms/step")

    # roofline of the dominant kernel: the leaf GEMM launch (all 49 x K INT8 GEMMs)
    per_launch, per_call = leaf_kernel_ms(caps_amd, torch, dA, dB, dC_split, desc, pat,
                                          SPLIT, K, reps=args.gemm_reps)
    m = n >> r
    leaves = 7 ** r
    int_ops = K * 2.0 * leaves * m ** 3
    gemm_ms = per_launch["mul"]
    achieved = int_ops / (gemm_ms * 1e-3) / 1e12
    plain_launch, plain_call = leaf_kernel_ms(caps_amd, torch, dA, dB, dC_plain, desc, pat,
                                              caps_amd.MODE_FP64, 0, reps=args.gemm_reps)
    f64_ach = 2.0 * leaves * m ** 3 / (plain_launch["mul"] * 1e-3) / 1e12
    log(f"[bench] split phases per call (ms): {per_call}")
    log(f"[bench] plain phases per call (ms): {plain_call}")

    result = {
        "metric": "effective GFLOP/s (2n^3/t) at n=8192, err<=1e-14; split-precision vs plain FP64",
        "value": round(eff / t_split / 1e9, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_split * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 in/out, int8 MFMA leaf products",
        "data": "synthetic: entries u*2^e, u~U[-1,1), e~U{-10..10}, fixed seeds",
        "config": {"workload": f"CAPS multiply n={n} nrec={r} pattern '{pat}' bs=1, "
                               f"{leaves} leaf products of {m}^2 per launch",
                   "n": n, "nrec": r, "pattern": pat, "moduli": K,
                   "split_mode": "residue-domain Strassen (CAPS_MODE_CRT_INT8)",
                   "plain_fp64_gflops": round(eff / t_plain / 1e9, 2),
                   "leaf_split_gflops": round(eff / t_leaf / 1e9, 2),
                   "bf16_variant_gflops": round(eff / t_bf16 / 1e9, 2),
                   "bf16_variant_ms_per_step": round(t_bf16 * 1e3, 4),
                   "plain_fp64_ms_per_step": round(t_plain * 1e3, 4),
                   "split_over_plain": round(t_plain / t_split, 3),
                   "phases_ms_split": {k: round(v, 4) for k, v in per_call.items() if v},
                   "phases_ms_plain": {k: round(v, 4) for k, v in plain_call.items() if v},
                   "plain_leaf_gemm_tflops": round(f64_ach, 2),
                   "plain_leaf_gemm_frac_fp64_peak": round(f64_ach / PEAK_FP64_TFLOPS, 4)},
        "roofline": {"bound": "mfma",
                     "kernel": "oz_igemm_big_kernel<true, false> (INT8 MFMA, all leaves x "
                               "moduli)",
                     "achieved": round(achieved, 2), "peak": PEAK_INT8_TOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_INT8_TOPS, 4),
                     "traffic": profiled_traffic("oz_igemm_big_kernel<true, false>"),
                     "note": "INT8 multiply-adds counted as 2 ops; per launch "
                             f"{int_ops:.4e} ops = moduli x 2 x 49 x 2048^3, "
                             f"{gemm_ms:.3f} ms"},
    }

    if not args.no_check:
        log("[bench] accuracy: all of C against the 16-modulus reference product")
        dRef = torch.empty_like(dA)
        caps_amd.reference_product(dA, dB, dRef, desc)
        full = {name: caps_amd.normwise_error(c, dRef, desc)
                for name, c in (("split", dC_split), ("plain", dC_plain), ("leaf_split", dC_leaf),
                                ("bf16_variant", dC_bf16))}
        del dRef
        caps_amd.release()
        log(f"[bench] full-C errors {full}")
        log("[bench] accuracy: copying operands to host")
        A_cm = oracle.from_caps(dA.cpu().numpy(), n, r, 1)
        B_cm = oracle.from_caps(dB.cpu().numpy(), n, r, 1)
        rng = np.random.default_rng(99)
        rows = np.sort(rng.choice(n, 16, replace=False))
        err_split = sampled_error(oracle, A_cm, B_cm,
                                  oracle.from_caps(dC_split.cpu().numpy(), n, r, 1), rows)
        err_plain = sampled_error(oracle, A_cm, B_cm,
                                  oracle.from_caps(dC_plain.cpu().numpy(), n, r, 1), rows)
        err_leaf = sampled_error(oracle, A_cm, B_cm,
                                 oracle.from_caps(dC_leaf.cpu().numpy(), n, r, 1), rows)
        bf16_same = bool(torch.equal(dC_bf16, dC_split))
        log(f"[bench] error split {err_split:.3e} plain {err_plain:.3e} leaf {err_leaf:.3e}")
        del A_cm, B_cm
        log("[bench] integer pair through the split path")
        iA = make_inputs(torch, n, 4321, integer=True)
        iB = make_inputs(torch, n, 8765, integer=True)
        iC = torch.empty_like(iA)
        caps_amd.multiply_device(iA, iB, iC, desc, pat, mode=SPLIT, moduli=K)
        torch.cuda.synchronize()
        exact = exact_integer_check(oracle.from_caps(iA.cpu().numpy(), n, r, 1),
                                    oracle.from_caps(iB.cpu().numpy(), n, r, 1),
                                    oracle.from_caps(iC.cpu().numpy(), n, r, 1), rows)
        log(f"[bench] integer pair bit-exact: {exact}")
        result["config"]["error_split"] = err_split
        result["config"]["error_plain"] = err_plain
        result["config"]["error_leaf_split"] = err_leaf
        result["config"]["bf16_variant_bitwise_equal_to_int8"] = bf16_same
        result["config"]["error_metric"] = ("max|C-C*|/max|C*| over 16 sampled full rows, "
                                            "C* = Dot2 compensated classical product")
        result["config"]["error_full_c"] = full
        result["config"]["error_full_c_metric"] = (
            "max|C-C*|/max|C*| over all n^2 entries, C* = residue-domain product with 16 "
            "moduli (operands truncated 63 bits below their row / column norms)")
        result["config"]["integer_pair_bit_exact"] = exact
        result["config"]["meets_error_bar"] = err_split <= 1e-14 and full["split"] <= 1e-14

    del dA, dB, dC_split, dC_plain, dC_leaf, dC_bf16
    caps_amd.release()
    torch.cuda.empty_cache()
    if not args.no_fp64_bdd:
        result["config"]["plain_fp64_n7168_bdd"] = fp64_bdd_workload(args, torch, caps_amd, oracle)

    if not args.no_batch:
        log("[bench] uneven batch: 512 problems, plans from the learned planner")
        result["config"]["batch"] = uneven_batch_section(args, torch, caps_amd)

    if not args.no_cpu:
        log("[bench] cpu baseline (reference randombenchmark, else the oracle restatement)")
        result["cpu_baseline"] = cpu_baseline(oracle)
    else:
        result["cpu_baseline"] = None

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
