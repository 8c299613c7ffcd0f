"""The C++ drivers built on the drop-in headers (include/caps/*.h): the
reference's correctness harness semantics (fromfile1.cpp:124-145, "Maximum
deviation" must be 0 on the integer fixtures, README:35) and the benchmark
report (counters.cpp:61-74)."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "caps_amd", "bin")
GOLD = os.path.join(ROOT, "tests", "golden")


def _run(args, timeout=120):
    print("[gpu] driver", " ".join(args), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.mark.parametrize("size,bs,r,pattern,mode", [
    (392, 7, 3, "bdd", "fp64"), (392, 1, 3, None, "crt"), (196, 7, 2, "db", "leaf"),
    (28, 2, 1, "d", "fp64")])
def test_fromfile1_fixture_exact(size, bs, r, pattern, mode):
    exe = os.path.join(BIN, "fromfile1")
    if not os.path.exists(exe):
        pytest.skip("drivers not built")
    args = [exe, "-i", f"{GOLD}/test{size}.in", "-b", str(bs), "-r", str(r),
            "-c", f"{GOLD}/test{size}.correct", "--mode", mode]
    if pattern:
        args += ["-p", pattern]
    out = _run(args)
    assert "Maximum deviation 0.000000e+00" in out, out


def test_randombenchmark_report():
    exe = os.path.join(BIN, "randombenchmark")
    if not os.path.exists(exe):
        pytest.skip("drivers not built")
    out = _run([exe, "-s", "1024", "-r", "3", "-p", "bdd"])
    assert "Execution pattern BFS,DFS,DFS," in out
    m = re.search(r"effective Gflop/s ([0-9.]+)", out)
    assert m and float(m.group(1)) > 0
    # flops = additions + multiplications with the reference's counting
    adds = 15 * (512 ** 2) + 7 * 15 * (256 ** 2) + 49 * 15 * (128 ** 2) + 343 * 128 ** 3
    assert f"flops={adds + 343 * 128 ** 3} " in out


def test_invalid_pattern_exits_like_reference():
    exe = os.path.join(BIN, "randombenchmark")
    if not os.path.exists(exe):
        pytest.skip("drivers not built")
    r = subprocess.run([exe, "-s", "64", "-r", "2", "-p", "dx"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode != 0 and "invalid pattern" in r.stdout
