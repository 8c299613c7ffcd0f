"""Measurements for the planner's cost / error model: for seeded problem
shapes (size n, CAPS layout depth, exponent spread of the entries) every
candidate configuration the autotuner enumerates is run on the device and
its time (HIP events) and normwise error (against the 16-modulus residue
product) are recorded.

    python -m caps_amd.planner_data --split train --shapes 170 --part 0 --parts 4 \\
        --out OUT/planner_train_0.npz

Shapes (splits train / test): n = 256 round(exp(U[ln 512, ln 16384]) / 256);
layout depth r = the largest with leaves n / 2^r >= 512 and n divisible by 2^r
(at least 0).  Splits batch_train / batch_test cover the sizes of uneven
batches: n = 64 round(exp(U[ln 128, ln 4096]) / 64), r uniform over the depths
with n divisible by 2^r and leaves of at least 64.  Entries u 2^e with u
uniform [-1, 1), e uniform integer in [-spread, spread], spread uniform in
{0, ..., 20}.  Every split has its own seed.
Candidates: plain FP64 and the residue-domain mode (moduli 12..16, or the
subset given by --crt-moduli), every pattern b^k d^(L-k) and every
admissible cutoff; a candidate whose plan does not fit the device's free
memory (deep all-'b' plans at the largest sizes) is skipped, as the
exhaustive autotuner skips it.  Each shape's rows are
written as soon as the shape is done, so a part can be resumed.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

import caps_amd

SEEDS = {"train": 1000, "test": 2000, "batch_train": 3000, "batch_test": 4000}


def shape(split: str, i: int):
    rng = np.random.default_rng(SEEDS[split] * 100003 + i)
    if split.startswith("batch"):
        n = int(64 * np.round(np.exp(rng.uniform(np.log(128), np.log(4096))) / 64))
        n = min(max(n, 128), 4096)
        ds = [r for r in range(12) if n % (1 << r) == 0 and (n >> r) >= 64]
        r = int(ds[int(rng.integers(0, len(ds)))])
    else:
        n = int(256 * np.round(np.exp(rng.uniform(np.log(512), np.log(16384))) / 256))
        n = min(max(n, 512), 16384)
        r = 0
        while n % (2 << r) == 0 and (n >> (r + 1)) >= 512:
            r += 1
    spread = int(rng.integers(0, 21))
    return n, r, spread


def inputs(torch, n, spread, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    out = []
    for _ in range(2):
        u = torch.rand(n * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
        if spread:
            e = torch.randint(-spread, spread + 1, (n * n,), generator=g, device="cuda",
                              dtype=torch.int32)
            u = torch.ldexp(u, e.to(torch.float64))
        out.append(u)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--split", choices=sorted(SEEDS), required=True)
    ap.add_argument("--shapes", type=int, required=True, help="shapes in the whole split")
    ap.add_argument("--part", type=int, default=0)
    ap.add_argument("--parts", type=int, default=1)
    ap.add_argument("--reps", type=int, default=1)
    ap.add_argument("--budget-s", type=float, default=900.0,
                    help="stop after this many seconds (rows so far are kept)")
    ap.add_argument("--crt-moduli", default="",
                    help="comma-separated moduli counts of the residue-domain candidates to keep "
                         "(default: all the library enumerates)")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)

    import torch
    if not torch.cuda.is_available():
        print("planner_data: no GPU", file=sys.stderr)
        return 2
    caps_amd.load()
    mask = caps_amd.MASK_FP64 | caps_amd.MASK_CRT_INT8
    rows = {k: [] for k in ("shape", "n", "nrec", "spread", "features", "ms", "err", "mode",
                            "moduli", "cutoff", "pattern")}
    if os.path.exists(a.out):  # resume
        with np.load(a.out) as z:
            for k in rows:
                rows[k] = list(z[k])
    done = set(int(s) for s in rows["shape"])
    t0 = time.time()
    for i in range(a.part, a.shapes, a.parts):
        if i in done:
            continue
        if time.time() - t0 > a.budget_s:
            print(f"[planner_data] budget reached before shape {i}", flush=True)
            break
        n, r, spread = shape(a.split, i)
        desc = caps_amd.MatDescriptor.one_device(n, r, 1)
        A, B = inputs(torch, n, spread, 7919 * i + SEEDS[a.split])
        C = torch.empty_like(A)
        Ref = torch.empty_like(A)
        caps_amd.reference_product(A, B, Ref, desc)
        # the model's features count 'b' and 'd' levels and carry no tile:
        # the data keeps the prefix patterns b^k d^l and the default tiles
        cands = [c for c in caps_amd.candidates(desc, mask)
                 if c.tile == 0 and "db" not in c.pattern.decode()[:r - c.cutoff]]
        if a.crt_moduli:
            keep = {int(x) for x in a.crt_moduli.split(",")}
            cands = [c for c in cands if c.mode != caps_amd.MODE_CRT_INT8 or c.moduli in keep]
        ts = time.time()
        skipped = 0
        for c in cands:
            try:
                caps_amd.measure(A, B, C, desc, c, Ref, a.reps)
            except caps_amd.CapsError as e:
                if e.code != 3:  # only a plan that does not fit the device is skipped
                    raise
                skipped += 1
                caps_amd.release()
                torch.cuda.empty_cache()
                continue
            rows["shape"].append(i)
            rows["n"].append(n)
            rows["nrec"].append(r)
            rows["spread"].append(spread)
            rows["features"].append(caps_amd.candidate_features(desc, c, float(spread)))
            rows["ms"].append(c.ms)
            rows["err"].append(c.err)
            rows["mode"].append(c.mode)
            rows["moduli"].append(c.moduli)
            rows["cutoff"].append(c.cutoff)
            rows["pattern"].append(c.pattern.decode())
        del A, B, C, Ref
        caps_amd.release()
        torch.cuda.empty_cache()
        np.savez(a.out, **{k: np.array(v) for k, v in rows.items()})
        print(f"[planner_data] {a.split} shape {i}: n={n} nrec={r} spread={spread} "
              f"{len(cands) - skipped} candidates in {time.time() - ts:.1f} s, {skipped} too large "
              f"for the device (rows {len(rows['ms'])})",
              flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
