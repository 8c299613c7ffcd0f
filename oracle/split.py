"""Exact-integer model of the two split-precision modes (TEST INFRASTRUCTURE ONLY).

Only tests/ import this module, as the checker of the HIP library's split
modes; the library never calls it.  Pure Python / numpy with exact integers
(Python ints, Fractions), so it is for small matrices only.

Both modes scale FP64 operands by powers of two to integers, multiply the
integers exactly and round once.  With identical scale exponents the library's
output equals this model bit for bit:

  CRT mode (residue-domain Strassen):
      sa_i from the 2-norm of row i of the whole A, sb_j from column j of B,
      A' = trunc(2^sa A), B' = trunc(2^sb B), C = 2^-(sa+sb) RN(A'B').
      The Winograd recursion runs on exact residues, so the result does not
      depend on the pattern or the recursion depth.
  leaf-split mode (FP64 Winograd additions, Ozaki-II leaves):
      the reference's FP64 recursion (multiply.cpp:212-310 arithmetic) with
      each leaf product replaced by the scaled exact product of that leaf.

The scale exponent follows the library's rule: with N = ||row||_2 computed
here exactly (a correctly rounded sum of squares), nm = sqrt(s) (1 + 2^-20),
sa = the largest t with 2^t 2^e nm <= L.  The library's sum of squares is
floating point, so the two exponents can differ only when 2^t 2^e nm lies
within ~1e-13 relative of L; random test data never hits that tie.

Parameters (moduli, M, L) are computed here independently of the library:
the moduli are the pairwise-coprime integers <= 256 taken greedily downward,
M their product, L = sqrt(M/2 (1 - 2^-30)) (1 - 2^-40) with M first rounded
toward zero to 53 bits, so that ||A'_i|| ||B'_j|| <= L^2 < M/2 bounds the
CRT-reconstructed integers (Ozaki, Uchino, Imamura, "Ozaki Scheme II", 2025).
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np


def moduli(count: int) -> list[int]:
    out: list[int] = []
    c = 256
    while len(out) < count and c >= 2:
        if all(math.gcd(c, o) == 1 for o in out):
            out.append(c)
        c -= 1
    return out


def modulus_product(count: int) -> int:
    M = 1
    for m in moduli(count):
        M *= m
    return M


def _trunc53(x: int) -> float:
    """x > 0 rounded toward zero to 53 bits, as a float."""
    b = x.bit_length()
    if b <= 53:
        return float(x)
    return math.ldexp(float(x >> (b - 53)), b - 53)


def norm_limit(count: int) -> float:
    Md = _trunc53(modulus_product(count))
    D = Md * 0.5 * (1.0 - math.ldexp(1.0, -30))
    return math.sqrt(D) * (1.0 - math.ldexp(1.0, -40))


def _ilogb(x: float) -> int:
    return math.frexp(x)[1] - 1


def scale_exponent(v: np.ndarray, L: float) -> int:
    """Largest t with 2^t * ||v||_2 bound <= L (0 for a zero vector)."""
    nz = [float(x) for x in v if x != 0.0]
    if not nz:
        return 0
    e = max(_ilogb(abs(x)) for x in nz)
    s = float(sum(Fraction(math.ldexp(x, -e)) ** 2 for x in nz))
    nm = math.sqrt(s) * (1.0 + math.ldexp(1.0, -20))
    t = _ilogb(L) - e - _ilogb(nm)
    if math.ldexp(nm, e + t) > L:
        t -= 1
    return t


def _scaled(M: np.ndarray, exps, axis: int) -> np.ndarray:
    """trunc(2^exp * M) as FP64 integers (exact), exps per row (axis 0) or
    column (axis 1)."""
    e = np.asarray(exps, dtype=np.int64)
    return np.trunc(np.ldexp(M, e[:, None] if axis == 0 else e[None, :]))


def _pieces(X: np.ndarray, count: int):
    """X = sum_p X_p 2^(19 p) exactly, |X_p| < 2^19, for FP64 integers
    |X| < 2^(19 count) (every step is an exact FP64 operation)."""
    out = []
    R = X
    for p in range(count - 1, 0, -1):
        Xp = np.trunc(np.ldexp(R, -19 * p))
        R = R - np.ldexp(Xp, 19 * p)
        out.append(Xp)
    out.append(R)
    return out[::-1]


def exact_int_matmul(Ai: np.ndarray, Bi: np.ndarray) -> np.ndarray:
    """Exact product of FP64 integer matrices (|entries| < 2^76, K < 2^15) as
    a numpy object array of Python ints: FP64 GEMMs of 19-bit pieces whose
    partial sums stay below 2^53 (exact in any order), recombined with exact
    integers."""
    assert Ai.shape[1] < 2 ** 15
    big = max(np.abs(Ai).max(initial=0), np.abs(Bi).max(initial=0))
    assert big < 2.0 ** 76
    count = 3 if big < 2.0 ** 57 else 4
    pa, pb = _pieces(Ai, count), _pieces(Bi, count)
    to_int = np.vectorize(int, otypes=[object])
    C = np.zeros((Ai.shape[0], Bi.shape[1]), dtype=object)
    for p in range(count):
        for q in range(count):
            G = pa[p] @ pb[q]
            C = C + to_int(G) * (1 << (19 * (p + q)))
    return C


def scaled_product(A: np.ndarray, B: np.ndarray, count: int):
    """2^-(sa+sb) RN(A'B') for column-major numpy matrices A (rows scaled) and
    B (columns scaled); also returns (sa, sb, A', B') with A', B' as FP64
    integers."""
    L = norm_limit(count)
    sa = [scale_exponent(A[i, :], L) for i in range(A.shape[0])]
    sb = [scale_exponent(B[:, j], L) for j in range(B.shape[1])]
    Ai = _scaled(A, sa, 0)
    Bi = _scaled(B, sb, 1)
    Ci = exact_int_matmul(Ai, Bi)
    C = np.empty(Ci.shape, dtype=np.float64)
    for (i, j), x in np.ndenumerate(Ci):
        C[i, j] = math.ldexp(float(x), -(sa[i] + sb[j]))  # float(int) rounds to nearest even
    return C, sa, sb, Ai, Bi


def crt_mode(A: np.ndarray, B: np.ndarray, count: int = 14) -> np.ndarray:
    """The CRT (residue-domain) mode's result for n x n A, B (any pattern)."""
    return scaled_product(A, B, count)[0]


# ---------------------------------------------------------------- leaf-split mode
def _quads(X: np.ndarray):
    h = X.shape[0] // 2
    return X[:h, :h], X[h:, :h], X[:h, h:], X[h:, h:]   # 11, 21, 12, 22


def winograd(A: np.ndarray, B: np.ndarray, r: int, leaf) -> np.ndarray:
    """Winograd recursion with the reference's association (BFS arithmetic,
    multiply.cpp:272, 277, 195-210; identical to the DFS order) on n x n
    numpy matrices; `leaf(A, B)` multiplies at depth r.  Elementwise FP64
    adds in numpy are single IEEE operations, as in the reference loops."""
    if r == 0:
        return leaf(A, B)
    A11, A21, A12, A22 = _quads(A)
    B11, B21, B12, B22 = _quads(B)
    T3 = A21 + A22
    T4 = T3 - A11
    T5 = A11 - A21
    T6 = A12 - T4
    S5 = B22 - B12
    S3 = B12 - B11
    S4 = B22 - S3
    S7 = S4 - B21
    T = [A11, A12, T3, T4, T5, T6, A22]
    S = [B11, B21, S3, S4, S5, B22, S7]
    Q = [winograd(np.ascontiguousarray(t), np.ascontiguousarray(s), r - 1, leaf)
         for t, s in zip(T, S)]
    C11 = Q[0] + Q[1]
    U1 = Q[0] + Q[3]
    U2 = U1 + Q[4]
    C21 = U2 - Q[6]
    C22 = U2 + Q[2]
    U3 = U1 + Q[2]
    C12 = U3 + Q[5]
    return np.block([[C11, C12], [C21, C22]])


def leaf_split_mode(A: np.ndarray, B: np.ndarray, r: int, count: int = 14) -> np.ndarray:
    """The leaf-split mode's result: FP64 Winograd, scaled exact leaves."""
    return winograd(A, B, r, lambda a, b: scaled_product(a, b, count)[0])


def residues(X: int, count: int) -> list[int]:
    """Symmetric residues of X modulo each modulus (as INT8 values)."""
    out = []
    for m in moduli(count):
        r = X % m
        hi = (m - 1) // 2
        out.append(r - m if r > hi else r)
    return out


def crt(res: list[int], count: int) -> int:
    """Inverse of residues(): the X in (-M/2, M/2) with those residues."""
    ms = moduli(count)
    M = modulus_product(count)
    x = 0
    for r, m in zip(res, ms):
        Mi = M // m
        x += r * Mi * pow(Mi, -1, m)
    x %= M
    return x - M if x > M // 2 else x
