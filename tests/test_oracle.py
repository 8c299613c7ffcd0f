"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong sign, a swapped quadrant, a transposed operand) fails it:
  * the reference's own integer fixtures (tests/golden, copied data files of
    the reference's tests/ directory) for every one-process .params row,
  * closed-form layout index vs. recursive placement, bijectivity, round trip,
  * exact int64 products, brute-force triple loops, exact rational Dot2,
  * the C library's drand48 for the input generator,
  * the reference's flop-counter formulas.
"""
import ctypes
import ctypes.util
from fractions import Fraction

import numpy as np
import pytest

import oracle


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


# ------------------------------------------------------------------ layout
@pytest.mark.parametrize("n,r,bs", [(28, 2, 7), (56, 3, 1), (392, 3, 7), (16, 1, 4),
                                    (64, 2, 16), (48, 0, 12), (8, 3, 1)])
def test_layout_index_matches_recursive_placement(n, r, bs):
    # ramp matrix (idea of comm-test.cpp:20-21): value = column-major index
    M = np.arange(n * n, dtype=np.float64).reshape(n, n, order="F")
    v = oracle.to_caps(M, r, bs)
    # every element lands where the closed-form index says
    for i in range(0, n, max(1, n // 13)):
        for j in range(0, n, max(1, n // 11)):
            assert v[oracle.caps_offset(n, r, bs, i, j)] == M[i, j]
    # bijective: the ramp's values are a permutation
    assert np.array_equal(np.sort(v), np.arange(n * n, dtype=np.float64))
    # round trip
    assert np.array_equal(oracle.from_caps(v, n, r, bs), M)


def test_layout_quadrant_order():
    # quadrant order 11, 21, 12, 22 (communication.cpp:646-653)
    n = 4
    M = np.arange(16, dtype=np.float64).reshape(4, 4, order="F")
    v = oracle.to_caps(M, 1, 1)
    assert list(v[0:4]) == [M[0, 0], M[1, 0], M[0, 1], M[1, 1]]      # A11
    assert list(v[4:8]) == [M[2, 0], M[3, 0], M[2, 1], M[3, 1]]      # A21
    assert list(v[8:12]) == [M[0, 2], M[1, 2], M[0, 3], M[1, 3]]     # A12
    assert list(v[12:16]) == [M[2, 2], M[3, 2], M[2, 3], M[3, 3]]    # A22


# ------------------------------------------------------------------ fixtures
def _fixture_rows():
    rows = []
    for size in oracle.FIXTURE_SIZES:
        for nproc, bs, r in oracle.read_params(size):
            if nproc == 1:
                rows.append((size, bs, r))
    return rows


def test_fixture_row_count():
    # 110 single-process rows across the six .params files
    assert len(_fixture_rows()) == 110


_FIX = {}


def _fixture(size):
    if size not in _FIX:
        _FIX[size] = oracle.read_fixture(size)
    return _FIX[size]


@pytest.mark.parametrize("size,bs,r", _fixture_rows())
def test_oracle_fixture_rows_exact(size, bs, r):
    A, B, C = _fixture(size)
    a = oracle.to_caps(A, r, bs)
    b = oracle.to_caps(B, r, bs)
    c = oracle.multiply(a, b, size, r, bs, None, threads=2)
    assert np.array_equal(oracle.from_caps(c, size, r, bs), C)


@pytest.mark.parametrize("size", oracle.FIXTURE_SIZES)
def test_oracle_fixture_bfs_patterns(size):
    # multi-rank rows (-n 7^k) map to k leading 'b' steps on one device
    A, B, C = _fixture(size)
    for nproc, bs, r in oracle.read_params(size):
        if nproc == 1:
            continue
        k = {7: 1, 49: 2, 343: 3}[nproc]
        pat = "b" * k + "d" * (r - k)
        c = oracle.multiply(oracle.to_caps(A, r, bs), oracle.to_caps(B, r, bs), size, r, bs, pat, 2)
        assert np.array_equal(oracle.from_caps(c, size, r, bs), C), (nproc, bs, r, pat)


def test_fixture_answer_is_exact_product():
    # the .correct files are the exact integer products of the .in matrices
    for size in (28, 42, 56):
        A, B, C = _fixture(size)
        assert np.array_equal(oracle.igemm_exact(A, B), C)


# ------------------------------------------------------------------ BFS == DFS
@pytest.mark.parametrize("pattern", ["bbb", "ddd", "bdb", "dbd", "hhd", "BDb"])
def test_bfs_dfs_bit_identical_on_reals(pattern):
    n, r = 256, 3
    rng = np.random.default_rng(7)
    a = rng.uniform(-1, 1, n * n)
    b = rng.uniform(-1, 1, n * n)
    ref = oracle.multiply(a, b, n, r, 1, "ddd", 2)
    got = oracle.multiply(a, b, n, r, 1, pattern, 2)
    assert np.array_equal(ref, got)
    assert np.array_equal(ref, oracle.multiply(a, b, n, r, 1, None, 2))


def test_strassen_close_to_classical_on_reals():
    n, r = 256, 3
    rng = np.random.default_rng(3)
    A = rng.uniform(-1, 1, (n, n))
    B = rng.uniform(-1, 1, (n, n))
    c = oracle.from_caps(oracle.multiply(oracle.to_caps(A, r, 1), oracle.to_caps(B, r, 1),
                                         n, r, 1, "bdd", 2), n, r, 1)
    ref = oracle.dgemm_dot2(A, B)
    assert np.abs(c - ref).max() / np.abs(ref).max() < 1e-14
    # and a transposed operand would not be
    assert np.abs(c - oracle.dgemm_dot2(A.T, B)).max() / np.abs(ref).max() > 1e-2


def test_bad_pattern_and_descriptor():
    a = np.zeros(64)
    with pytest.raises(oracle.OracleError):
        oracle.multiply(a, a, 8, 2, 1, "dx")      # invalid char (multiply.cpp:100-104)
    with pytest.raises(oracle.OracleError):
        oracle.multiply(a, a, 8, 2, 1, "d")       # pattern shorter than nrec
    with pytest.raises(oracle.OracleError):
        oracle.multiply(a, a, 8, 2, 3, None)      # lda % (2^nrec * bs) != 0


# ------------------------------------------------------------------ counters
def test_counters_match_reference_formula():
    # DFS level: 15 half-size additions per sub-problem (multiply.cpp:421-451);
    # leaf: m^3 additions and m^3 multiplications (multiply.cpp:1085-1086).
    n, r = 1024, 3
    m = n >> r
    a = np.zeros(n * n)
    oracle.reset_counters()
    oracle.multiply(a, a, n, r, 1, "ddd", 4)
    adds, mults = oracle.counters()
    level_adds = sum(7 ** l * 15 * (n >> (l + 1)) ** 2 for l in range(r))
    assert level_adds == 22_855_680
    assert mults == 7 ** r * m ** 3
    assert adds == level_adds + 7 ** r * m ** 3
    oracle.reset_counters()
    oracle.multiply(a, a, n, r, 1, "bbd", 4)
    assert oracle.counters() == (adds, mults)   # BFS: 4 + 4 + 7 additions per level


# ------------------------------------------------------------------ GEMMs
def test_blocked_dgemm_equals_triple_loop():
    rng = np.random.default_rng(11)
    for n in (1, 5, 16, 37, 64):
        A = rng.uniform(-1, 1, (n, n))
        B = rng.uniform(-1, 1, (n, n))
        assert np.array_equal(oracle.dgemm(A, B, threads=3), oracle.dgemm_naive(A, B))


def test_dgemm_k_order():
    # entry = ((0 + a0 b0) + a1 b1) + a2 b2 ... in k order
    A = np.array([[1e16, 1.0, -1e16]])
    B = np.array([[1.0], [1.0], [1.0]])
    A3 = np.zeros((3, 3)); A3[0] = A[0]
    B3 = np.zeros((3, 3)); B3[:, 0] = B[:, 0]
    assert oracle.dgemm_naive(A3, B3)[0, 0] == 0.0          # (1e16 + 1) - 1e16 = 0
    assert oracle.dgemm(A3, B3)[0, 0] == 0.0
    assert oracle.dgemm_dot2(A3, B3)[0, 0] == 1.0           # compensated: exact


def test_integer_products_exact():
    rng = np.random.default_rng(5)
    n = 96
    A = rng.integers(-64, 65, (n, n)).astype(np.float64)
    B = rng.integers(-64, 65, (n, n)).astype(np.float64)
    exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
    assert np.array_equal(oracle.igemm_exact(A, B), exact)
    assert np.array_equal(oracle.dgemm(A, B), exact)
    assert np.array_equal(oracle.dgemm_dot2(A, B), exact)


def test_dot2_against_exact_rationals():
    rng = np.random.default_rng(9)
    n = 10
    # ill-conditioned sums: large cancelling terms plus small ones
    A = rng.uniform(-1, 1, (n, n)) * np.ldexp(1.0, rng.integers(-30, 30, (n, n)))
    B = rng.uniform(-1, 1, (n, n)) * np.ldexp(1.0, rng.integers(-30, 30, (n, n)))
    C = oracle.dgemm_dot2(A, B)
    for i in range(n):
        for j in range(n):
            ex = sum(Fraction(A[i, k]) * Fraction(B[k, j]) for k in range(n))
            exf = float(ex)
            # Dot2 is as accurate as twice-working precision then rounded
            assert abs(C[i, j] - exf) <= 2 * np.spacing(abs(exf)) + 1e-300
    # sampled-entry variant agrees with the full one
    rows = [0, 3, 9]
    cols = [2, 3, 0]
    assert np.array_equal(oracle.dot2_entries(A, B, rows, cols), C[rows, cols])


# ------------------------------------------------------------------ drand48
def test_drand48_matches_libc():
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.drand48.restype = ctypes.c_double
    libc.srand48.argtypes = [ctypes.c_long]
    libc.srand48(12345)
    want = np.array([libc.drand48() for _ in range(1000)])
    got = oracle.Drand48(12345).raw(1000)
    assert np.array_equal(want, got)
    u = oracle.Drand48(12345).uniform(1000)
    assert np.array_equal(u, 2 * want - 1)


def test_drand48_default_state():
    # X0 = 0x1234ABCD330E; first output of the unseeded stream
    x1 = (0x5DEECE66D * 0x1234ABCD330E + 0xB) % (1 << 48)
    assert oracle.Drand48().raw(1)[0] == x1 / 2.0 ** 48
