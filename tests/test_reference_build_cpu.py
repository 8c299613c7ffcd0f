"""The reference's own randombenchmark, compiled from its sources by
oracle/ref.mk into oracle/_ref/ (where the reference tree was present), checks
the C restatement's counters: for one process and DFS patterns the two must
report the same additions, multiplications and execution pattern
(printCounters, counters.cpp:49-74).  Skipped where the binary was not built."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.skipif(not os.path.exists(oracle.REF_BENCH),
                                reason="oracle/_ref/randombenchmark not built here")


def reference_counters(n, nrec, pattern, bs):
    """(additions, multiplications, pattern) of the reference's report, which
    prints additions - multiplications and 2 x multiplications
    (counters.cpp:67-68)."""
    env = dict(os.environ, MKL_THREADING_LAYER="GNU", MKL_NUM_THREADS="2", OMP_NUM_THREADS="2")
    args = [oracle.REF_BENCH, "-s", str(n), "-r", str(nrec), "-b", str(bs), "-m", "256"]
    if pattern:
        args += ["-p", pattern]
    r = subprocess.run(args, capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    flops = re.search(r"flops=(\d+) \((\d+)\+(\d+)\)", r.stdout)
    pat = re.search(r"Execution pattern (\S*)", r.stdout)
    shown_adds, shown_mults = int(flops.group(2)), int(flops.group(3))
    assert int(flops.group(1)) == shown_adds + shown_mults
    return shown_adds + shown_mults // 2, shown_mults // 2, pat.group(1)


@pytest.mark.parametrize("n,nrec,pattern,bs", [
    (256, 2, "dd", 1),
    (448, 3, "ddd", 1),
    (512, 1, "d", 4),
    (384, 2, None, 2),     # the reference's automatic choice (all DFS with one process)
])
def test_reference_counters_equal_oracle(n, nrec, pattern, bs):
    adds, mults, pat = reference_counters(n, nrec, pattern, bs)
    a = np.zeros(n * n)
    oracle.reset_counters()
    oracle.multiply(a, a, n, nrec, bs, pattern or "d" * nrec, 2)
    assert oracle.counters() == (adds, mults)
    assert pat == "DFS," * nrec
