"""Pins the exact-integer model of the split modes (oracle/split.py) before it
checks the GPU: number theory of the moduli, CRT round trips, exactness on
integers, the Cauchy-Schwarz bound that makes the reconstruction unique, the
numpy Winograd restatement against the C oracle, and the error it reaches."""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle
from oracle import split


def test_moduli_pairwise_coprime_and_int8():
    ms = split.moduli(16)
    assert ms[:5] == [256, 255, 253, 251, 247]
    assert len(ms) == 16 and all(m <= 256 for m in ms)
    for a in range(16):
        for b in range(a + 1, 16):
            assert math.gcd(ms[a], ms[b]) == 1
    # 14 moduli give a 110.7-bit M
    assert 110 < math.log2(split.modulus_product(14)) < 111


@pytest.mark.parametrize("count", [2, 7, 14, 16])
def test_crt_round_trip(count):
    M = split.modulus_product(count)
    rng = np.random.default_rng(count)
    for _ in range(200):
        x = int(rng.integers(-2**62, 2**62)) * int(rng.integers(1, 2**40)) % M - M // 2
        r = split.residues(x, count)
        assert all(-128 <= v <= 127 for v in r)
        assert split.crt(r, count) == x


def test_norm_limit_bounds_products():
    for count in (8, 14, 16):
        L = split.norm_limit(count)
        M = split.modulus_product(count)
        assert Fraction(L) ** 2 < Fraction(M, 2)
        assert Fraction(L) ** 2 > Fraction(M, 2) * (1 - Fraction(1, 2**29))


def test_scale_exponent_is_largest_admissible():
    L = split.norm_limit(14)
    rng = np.random.default_rng(3)
    for _ in range(50):
        v = rng.uniform(-1, 1, 33) * np.ldexp(1.0, rng.integers(-30, 30, 33))
        t = split.scale_exponent(v, L)
        nrm = math.sqrt(float(sum(Fraction(x) ** 2 for x in v)))
        assert math.ldexp(nrm, t) <= L
        assert math.ldexp(nrm, t + 1) > L * (1 - 1e-6)
    assert split.scale_exponent(np.zeros(5), L) == 0


def test_cauchy_schwarz_bound_holds():
    """sum_k |A'_ik||B'_kj| < M/2 for every entry: the CRT value is unique."""
    rng = np.random.default_rng(4)
    n = 24
    A = rng.uniform(-1, 1, (n, n)) * np.ldexp(1.0, rng.integers(-10, 11, (n, n)))
    B = rng.uniform(-1, 1, (n, n)) * np.ldexp(1.0, rng.integers(-10, 11, (n, n)))
    _, sa, sb, Ai, Bi = split.scaled_product(A, B, 14)
    bound = split.exact_int_matmul(np.abs(Ai), np.abs(Bi))
    M = split.modulus_product(14)
    assert max(bound.ravel()) < M // 2
    # and the scaling used most of the range (within 2 bits per side)
    assert max(bound.ravel()) > M // 2 // 64


def test_integer_inputs_exact():
    rng = np.random.default_rng(5)
    n = 32
    A = rng.integers(-64, 65, (n, n)).astype(np.float64)
    B = rng.integers(-64, 65, (n, n)).astype(np.float64)
    exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
    assert np.array_equal(split.crt_mode(A, B, 14), exact)
    assert np.array_equal(split.leaf_split_mode(A, B, 2, 14), exact)
    # even with few moduli, integers of this size stay exact
    assert np.array_equal(split.crt_mode(A, B, 6), exact)


def test_numpy_winograd_matches_c_oracle():
    """The numpy restatement used by the leaf-split model is bit-identical to
    the C oracle when both use the oracle's leaf dgemm."""
    rng = np.random.default_rng(6)
    n, r = 64, 3
    A = rng.uniform(-1, 1, (n, n))
    B = rng.uniform(-1, 1, (n, n))
    got = split.winograd(A, B, r, lambda a, b: oracle.dgemm_naive(a, b))
    ref = oracle.from_caps(oracle.multiply(oracle.to_caps(A, r, 1), oracle.to_caps(B, r, 1),
                                           n, r, 1, "bbb", 2), n, r, 1)
    assert np.array_equal(got, ref)
    # a swapped product would not be
    bad = split.winograd(A, B, r, lambda a, b: oracle.dgemm_naive(b, a))
    assert not np.array_equal(bad, ref)


def test_exact_int_matmul_against_python_ints():
    rng = np.random.default_rng(8)
    A = np.trunc(rng.uniform(-1, 1, (9, 13)) * 2.0 ** 56)
    B = np.trunc(rng.uniform(-1, 1, (13, 7)) * 2.0 ** 56)
    want = np.array([[sum(int(A[i, k]) * int(B[k, j]) for k in range(13)) for j in range(7)]
                     for i in range(9)], dtype=object)
    assert (split.exact_int_matmul(A, B) == want).all()


def test_exact_int_matmul_wide_entries():
    """Entries up to 2^63 (16 moduli) take four pieces."""
    rng = np.random.default_rng(9)
    A = np.trunc(rng.uniform(-1, 1, (5, 11)) * 2.0 ** 63)
    B = np.trunc(rng.uniform(-1, 1, (11, 6)) * 2.0 ** 62)
    want = np.array([[sum(int(A[i, k]) * int(B[k, j]) for k in range(11)) for j in range(6)]
                     for i in range(5)], dtype=object)
    assert (split.exact_int_matmul(A, B) == want).all()


def test_crt_mode_error_level():
    """On u * 2^e data the exact-product model is within 1e-15 normwise of
    the compensated classical product, far below Winograd's FP64 floor."""
    rng = np.random.default_rng(7)
    n = 64
    A = rng.uniform(-1, 1, (n, n)) * np.ldexp(1.0, rng.integers(-10, 11, (n, n)))
    B = rng.uniform(-1, 1, (n, n)) * np.ldexp(1.0, rng.integers(-10, 11, (n, n)))
    ref = oracle.dgemm_dot2(A, B)
    err = np.abs(split.crt_mode(A, B, 14) - ref).max() / np.abs(ref).max()
    assert err < 1e-15
    # with too few moduli the truncation shows
    err8 = np.abs(split.crt_mode(A, B, 8) - ref).max() / np.abs(ref).max()
    assert err8 > 1e-12
