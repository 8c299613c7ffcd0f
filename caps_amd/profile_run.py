"""A short, profiler-friendly run of the multiply: a few calls of one mode on
device-resident inputs, then a plain device-to-device copy of known size
(the calibration for hardware byte counters such as FETCH_SIZE and
WRITE_SIZE: the copy moves exactly --copy-mib MiB in and out).

    python -m caps_amd.profile_run --n 8192 --nrec 2 --pattern bb --mode crt --calls 3
    python -m caps_amd.profile_run --n 8192 --nrec 2 --bs 8 --mode layout --calls 20
    rocprofv3 --kernel-trace --stats --output-format csv -d OUT -o run -- \\
        python3 -m caps_amd.profile_run --mode crt --calls 3

Inputs are u * 2^e, u uniform [-1, 1), e uniform integer in [-10, 10].
"""
from __future__ import annotations

import argparse
import sys

import caps_amd

MODES = {"fp64": caps_amd.MODE_FP64, "crt": caps_amd.MODE_CRT_INT8,
         "leaf": caps_amd.MODE_OZAKI_INT8, "layout": None}


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
    if a.copy_mib > 0:
        words = a.copy_mib * (1 << 20) // 8
        src = torch.ones(words, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[profile_run] copy {a.copy_mib} MiB", file=sys.stderr, flush=True)
    caps_amd.release()
    return 0


if __name__ == "__main__":
    sys.exit(main())
