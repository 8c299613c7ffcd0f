"""The device CRT (crt_acc + crt_finish in caps_amd/csrc/ozaki_dev.h),
restated with Python integers and exactly emulated FP64 roundings.  The
kernels hand it symmetric residues (|r| <= 128); the checks go further, to
any inputs congruent to the residues with |v| <= 2048, and confirm the
bounds its comments state (partial sums below 2^52, |k| < 2^15, |D_p| <
2^53, all exact in FP64) and that the assembled X is the CRT value, on
worst cases and random cases.  The final rounding of X = D_2 2^74 + D_1 2^37
+ D_0 to FP64 (TwoSum twice and one addition rounded to odd, in Python
floats, which round to nearest even like the device) is checked against
Python's correctly rounded int -> float conversion on sums built to sit at,
next to and away from rounding ties."""
from fractions import Fraction
import math
import random
import struct

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


# ------------------------------------------------------------ FP64 rounding of X

def two_sum(a, b):
    s = a + b
    bv = s - a
    return s, (a - (s - bv)) + (b - bv)


def add_round_odd(a, b, odd=True):
    s, e = two_sum(a, b)
    if not odd:
        return s
    bits = struct.unpack("<q", struct.pack("<d", s))[0]
    if e != 0.0 and not (bits & 1):
        bits += 1 if (e > 0.0) == (s > 0.0) else -1
    return struct.unpack("<d", struct.pack("<q", bits))[0]


def finish3(D, odd=True):
    """crt_finish's three-piece rounding, operation by operation."""
    uh, ul = two_sum(float(D[1]) * 2.0 ** 37, float(D[0]))
    th, tl = two_sum(float(D[2]) * 2.0 ** 74, uh)
    return th + add_round_odd(tl, ul, odd)


def pieces(X, rng):
    """D_0 + D_1 2^37 + D_2 2^74 == X with every |D_p| < 2^53, the digits
    moved off the canonical ones by random carries of either sign."""
    D0 = X % (1 << 37)
    D1 = (X >> 37) % (1 << 37)
    D2 = X >> 74
    t0 = rng.randrange(-2 ** 14, 2 ** 14)
    t1 = rng.randrange(-2 ** 14, 2 ** 14)
    D = [D0 + t0 * (1 << 37), D1 - t0 + t1 * (1 << 37), D2 - t1]
    assert sum(d << (37 * p) for p, d in enumerate(D)) == X
    return D if all(abs(d) < 2 ** 53 for d in D) else [D0, D1, D2]


def near_ties(rng, count):
    """Integers at, one off and far from the midpoint of two doubles, with
    magnitudes 2^60 .. 2^126 and both signs."""
    out = []
    for _ in range(count):
        e = rng.randrange(60, 126)
        y = (1 << e) + rng.randrange(0, 1 << 52) * (1 << (e - 52))   # a double
        half = 1 << (e - 53)                                        # half its ulp
        sign = rng.choice((1, -1))
        for d in (-1, 0, 1):
            out.append(sign * (y + half + d))
        out.append(sign * (y + rng.randrange(0, 2 * half)))
    return out


def test_three_piece_rounding_is_correct_rounding():
    rng = random.Random(2024)
    cases = near_ties(rng, 4000) + [0, 1, -1, (1 << 53) + 1, -(1 << 90) - 1]
    for X in cases:
        D = pieces(X, rng)
        assert finish3(D) == float(X), (X, D)


def test_three_piece_rounding_needs_round_to_odd():
    """Not vacuous: with the middle addition rounded to nearest instead of to
    odd, some of the tie cases above round the wrong way."""
    rng = random.Random(2024)
    wrong = sum(finish3(pieces(X, rng), odd=False) != float(X) for X in near_ties(rng, 4000))
    assert wrong > 0
