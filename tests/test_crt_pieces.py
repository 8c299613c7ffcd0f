"""The device CRT (crt_acc + crt_finish in caps_amd/csrc/ozaki_dev.h),
restated with Python integers and exactly emulated FP64 roundings.  The
kernels hand it symmetric residues (|r| <= 128); the checks go further, to
any inputs congruent to the residues with |v| <= 2048, and confirm the
bounds its comments state (partial sums below 2^52, |k| < 2^15, |D_p| <
2^53, all exact in FP64) and that the assembled X is the CRT value, on
worst cases and random cases."""
from fractions import Fraction
import math

import numpy as np

from oracle import split

BITS = 37


def consts(count):
    ms = split.moduli(count)
    M = split.modulus_product(count)
    npieces = (M.bit_length() + BITS - 1) // BITS
    piece = lambda x, p: (x >> (BITS * p)) & ((1 << BITS) - 1)
    W = [(M // m) * pow(M // m, -1, m) for m in ms]
    w = [[piece(Wl, p) for p in range(npieces)] for Wl in W]
    Mp = [piece(M, p) for p in range(npieces)]
    Md = float(M)  # the host's M.to_double() is <= M and within an ulp of it
    invM = 1.0 / Md
    return ms, M, npieces, w, Mp, invM


def rn(x):
    """Round an exact rational to the nearest double (ties to even)."""
    return float(Fraction(x))


def device_crt(v, count, C):
    ms, M, npieces, w, Mp, invM = C
    S = [0] * npieces
    for l, r in enumerate(v):
        for p in range(npieces):
            S[p] += r * w[l][p]           # fma(rd, w, S): exact while |S| < 2^53
            assert abs(S[p]) < 2 ** 52
    approx = rn(S[npieces - 1])
    for p in range(npieces - 2, -1, -1):
        approx = rn(Fraction(approx) * 2 ** BITS + S[p])   # one fma
    k = round(rn(Fraction(approx) * Fraction(invM)))      # rint(approx * invM)
    assert abs(k) < 2 ** 15
    D = [S[p] - k * Mp[p] for p in range(npieces)]        # fma(-k, Mp, S): exact
    assert all(abs(d) < 2 ** 53 for d in D)
    return sum(D[p] << (BITS * p) for p in range(npieces))


def unreduced(X, ms, rng, worst=False):
    v = []
    for m in ms:
        r = X % m
        r = r - m if r > (m - 1) // 2 else r
        lo, hi = -((2048 + r) // m), (2048 - r) // m      # r + t m within [-2048, 2048]
        t = (hi if r >= 0 else lo) if worst else int(rng.integers(lo, hi + 1))
        v.append(r + t * m)
    return v


def test_device_crt_on_residues():
    rng = np.random.default_rng(12)
    C = consts(14)
    ms, M = C[0], C[1]
    lim = (M // 2) - (M >> 31)
    for X in [0, -1, lim, -lim] + [int(rng.integers(-2 ** 62, 2 ** 62)) * (lim // 2 ** 62)
                                   for _ in range(200)]:
        assert device_crt(split.residues(X, 14), 14, C) == X


def test_device_crt_on_unreduced_sums():
    rng = np.random.default_rng(11)
    for count in (14, 16):
        C = consts(count)
        ms, M = C[0], C[1]
        lim = (M // 2) - (M >> 31)                        # |X| <= (1/2 - 2^-31) M
        cases = [0, 1, -1, lim, -lim, lim - 12345, -(lim - 777)]
        cases += [int(rng.integers(-2 ** 62, 2 ** 62)) * (lim // 2 ** 62) for _ in range(300)]
        for X in cases:
            for worst in (False, True):
                assert device_crt(unreduced(X, ms, rng, worst), count, C) == X


def test_device_crt_catches_a_dropped_piece():
    """Not vacuous: losing the top weight piece gives a different X."""
    rng = np.random.default_rng(3)
    C = consts(14)
    ms, M, npieces, w, Mp, invM = C
    w2 = [row[:-1] + [0] for row in w]
    bad = (ms, M, npieces, w2, Mp, invM)
    X = (M // 3)
    assert device_crt(unreduced(X, ms, rng), 14, bad) != X
