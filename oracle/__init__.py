"""CPU oracle for the CAPS Strassen-Winograd path (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker.  The HIP library in caps_amd/ never
loads it and has no CPU fallback.

The arithmetic lives in oracle.c (a plain-C restatement of the reference's
one-process path, each routine citing the reference file:line it follows);
this module loads liboracle.so with ctypes and adds numpy conveniences plus
the fixture reader for the reference's tests/*.in / *.correct / *.params
files (fromfile1.cpp:45-57 reads row-major text into column-major storage).
"""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"],
                       check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


REF_TREE = "/root/reference"
REF_BENCH = os.path.join(_HERE, "_ref", "randombenchmark")


def build_reference(ref: str = REF_TREE) -> str | None:
    """Compile the reference's own randombenchmark from its sources with the
    committed recipe (ref.mk: MPICH wrapper, MKL dgemm_) into oracle/_ref/.
    Only where the reference tree exists; elsewhere the prebuilt binary (if
    any) is used as is.  Returns the binary's path or None."""
    if os.path.isdir(ref):
        r = subprocess.run(["make", "-s", "-B", "-f", os.path.join(_HERE, "ref.mk"), f"REF={ref}"],
                           capture_output=True, text=True)
        if r.returncode != 0:
            print("oracle: reference build failed:\n" + r.stderr[-2000:])
            return None
    return REF_BENCH if os.path.exists(REF_BENCH) else None


def run_reference_benchmark(n: int, nrec: int, pattern: str, threads: int,
                            timeout: float = 300.0) -> dict | None:
    """Run oracle/_ref/randombenchmark (one MPI process, MKL with `threads`
    GNU OpenMP threads) and parse its printCounters report
    (counters.cpp:61-74).  None if the binary is absent or fails."""
    if not os.path.exists(REF_BENCH):
        return None
    env = dict(os.environ, MKL_THREADING_LAYER="GNU", MKL_NUM_THREADS=str(threads),
               OMP_NUM_THREADS=str(threads))
    try:
        r = subprocess.run([REF_BENCH, "-s", str(n), "-r", str(nrec), "-p", pattern,
                            "-m", str(max(1536, 3 * n * n * 8 // (1 << 20) * 4))],
                           capture_output=True, text=True, env=env, timeout=timeout)
    except (OSError, subprocess.TimeoutExpired):
        return None
    m = re.search(r"time ([0-9.eE+-]+) actual Gflop/s ([0-9.eE+-]+) per process effective "
                  r"Gflop/s ([0-9.eE+-]+)", r.stdout)
    if r.returncode != 0 or not m:
        return None
    return {"time_s": float(m.group(1)), "actual_gflops": float(m.group(2)),
            "effective_gflops": float(m.group(3)), "report": r.stdout.strip().splitlines()[-4:]}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_caps_offset.restype = ctypes.c_longlong
        L.oracle_caps_offset.argtypes = [ctypes.c_int] * 5
        for name in ("oracle_to_caps", "oracle_from_caps"):
            getattr(L, name).argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.oracle_valid_desc.argtypes = [ctypes.c_int] * 3
        L.oracle_multiply.argtypes = [_dp, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        for name in ("oracle_dgemm", "oracle_dgemm_naive", "oracle_dgemm_dot2",
                     "oracle_igemm_exact"):
            getattr(L, name).argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.oracle_dot2_entries.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int), _dp]
        L.oracle_additions.restype = ctypes.c_longlong
        L.oracle_multiplications.restype = ctypes.c_longlong
        L.oracle_drand48_default_state.restype = ctypes.c_uint64
        L.oracle_srand48_state.restype = ctypes.c_uint64
        L.oracle_srand48_state.argtypes = [ctypes.c_long]
        for name in ("oracle_fill_uniform", "oracle_drand48_seq"):
            getattr(L, name).argtypes = [_dp, ctypes.c_longlong, ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def threads_default() -> int:
    return max(1, min(16, os.cpu_count() or 1))


# ---------------------------------------------------------------- layout
def caps_offset(n: int, nrec: int, bs: int, i: int, j: int) -> int:
    return lib().oracle_caps_offset(n, nrec, bs, i, j)


def to_caps(M: np.ndarray, nrec: int, bs: int) -> np.ndarray:
    """Column-major n x n (a numpy array indexed M[i, j]) -> CAPS vector."""
    n = M.shape[0]
    cm = np.asfortranarray(M, dtype=np.float64)
    out = np.empty(n * n, dtype=np.float64)
    lib().oracle_to_caps(n, nrec, bs, cm.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return out


def from_caps(v: np.ndarray, n: int, nrec: int, bs: int) -> np.ndarray:
    """CAPS vector -> n x n numpy matrix (Fortran-ordered)."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty((n, n), dtype=np.float64, order="F")
    lib().oracle_from_caps(n, nrec, bs, v.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return out


# ---------------------------------------------------------------- multiply
class OracleError(RuntimeError):
    pass


def multiply(A: np.ndarray, B: np.ndarray, lda: int, nrec: int, bs: int,
             pattern: str | None = None, threads: int | None = None) -> np.ndarray:
    """CAPS-layout multiply (multiply.cpp:9-25 semantics, one process)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    C = np.empty(lda * lda, dtype=np.float64)
    pat = pattern.encode() if pattern is not None else None
    rc = lib().oracle_multiply(A.ctypes.data_as(_dp), B.ctypes.data_as(_dp),
                               C.ctypes.data_as(_dp), lda, nrec, bs, pat,
                               threads or threads_default())
    if rc != 0:
        raise OracleError(f"oracle_multiply failed with code {rc}")
    return C


def counters() -> tuple[int, int]:
    L = lib()
    return L.oracle_additions(), L.oracle_multiplications()


def reset_counters() -> None:
    lib().oracle_reset_counters()


# ---------------------------------------------------------------- GEMMs
def _gemm(name: str, A: np.ndarray, B: np.ndarray, threads: int | None) -> np.ndarray:
    n = A.shape[0]
    a = np.asfortranarray(A, dtype=np.float64)
    b = np.asfortranarray(B, dtype=np.float64)
    c = np.empty((n, n), dtype=np.float64, order="F")
    L = lib()
    L.oracle_set_threads(threads or threads_default())
    getattr(L, name)(n, a.ctypes.data_as(_dp), b.ctypes.data_as(_dp), c.ctypes.data_as(_dp))
    return c


def dgemm(A, B, threads=None):
    """Leaf GEMM of the oracle (k ascending per entry)."""
    return _gemm("oracle_dgemm", A, B, threads)


def dgemm_naive(A, B):
    return _gemm("oracle_dgemm_naive", A, B, 1)


def dgemm_dot2(A, B, threads=None):
    """Compensated classical GEMM (Dot2)."""
    return _gemm("oracle_dgemm_dot2", A, B, threads)


def igemm_exact(A, B, threads=None):
    return _gemm("oracle_igemm_exact", A, B, threads)


def dot2_entries(A: np.ndarray, B: np.ndarray, rows, cols, threads=None) -> np.ndarray:
    n = A.shape[0]
    a = np.asfortranarray(A, dtype=np.float64)
    b = np.asfortranarray(B, dtype=np.float64)
    r = np.ascontiguousarray(rows, dtype=np.int32)
    c = np.ascontiguousarray(cols, dtype=np.int32)
    out = np.empty(len(r), dtype=np.float64)
    L = lib()
    L.oracle_set_threads(threads or threads_default())
    ip = ctypes.POINTER(ctypes.c_int)
    L.oracle_dot2_entries(n, a.ctypes.data_as(_dp), b.ctypes.data_as(_dp), len(r),
                          r.ctypes.data_as(ip), c.ctypes.data_as(ip), out.ctypes.data_as(_dp))
    return out


# ---------------------------------------------------------------- drand48
class Drand48:
    """POSIX drand48 stream; default state is the unseeded POSIX state, which
    is what randombenchmark uses (no srand48 call without --random)."""

    def __init__(self, seed: int | None = None):
        L = lib()
        self.state = ctypes.c_uint64(L.oracle_drand48_default_state() if seed is None
                                     else L.oracle_srand48_state(seed))

    def uniform(self, count: int) -> np.ndarray:
        """count values of 2*drand48()-1 (randombenchmark.cpp:19-22)."""
        out = np.empty(count, dtype=np.float64)
        lib().oracle_fill_uniform(out.ctypes.data_as(_dp), count, ctypes.byref(self.state))
        return out

    def raw(self, count: int) -> np.ndarray:
        out = np.empty(count, dtype=np.float64)
        lib().oracle_drand48_seq(out.ctypes.data_as(_dp), count, ctypes.byref(self.state))
        return out


# ---------------------------------------------------------------- fixtures
GOLDEN = os.path.join(os.path.dirname(_HERE), "tests", "golden")


def read_fixture(size: int):
    """Return (A, B, C) as n x n float64 matrices from tests/golden/test<size>.*.
    The text is row-major; fromfile1.cpp:45-57 stores IA[i + j*lda]."""
    with open(os.path.join(GOLDEN, f"test{size}.in")) as f:
        toks = f.read().split()
    n = int(toks[0])
    vals = np.array(toks[1:], dtype=np.float64)
    assert vals.size == 2 * n * n, "fixture size mismatch"
    A = vals[: n * n].reshape(n, n)
    B = vals[n * n:].reshape(n, n)
    with open(os.path.join(GOLDEN, f"test{size}.correct")) as f:
        C = np.array(f.read().split(), dtype=np.float64).reshape(n, n)
    return A, B, C


_PARAM_RE = re.compile(r"mpirun -n (\d+) \./fromfile1 -i tests/test(\d+)\.in -b (\d+) -r (\d+)")


def read_params(size: int):
    """All (nproc, bs, nrec) rows of tests/golden/test<size>.params."""
    rows = []
    with open(os.path.join(GOLDEN, f"test{size}.params")) as f:
        for line in f:
            m = _PARAM_RE.search(line)
            if m:
                rows.append((int(m.group(1)), int(m.group(3)), int(m.group(4))))
    return rows


FIXTURE_SIZES = (28, 42, 56, 196, 294, 392)
