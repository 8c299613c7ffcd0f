"""The one-pass reduction of the residue GEMM's epilogue (res_byte128 in
caps_amd/csrc/ozaki_dev.h), restated in numpy with the same FP32 roundings,
checked exhaustively: for every modulus and every reduced sum y the kernel
can form, the byte it stores is the INT8 value of the symmetric residue of
the product sum.  An FP32 fma is emulated exactly: the product of two FP32
values and the addend fit a float64 without rounding, so one rounding to
float32 follows."""
import numpy as np

from oracle import split


def fma32(a, b, c):
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(np.float32)


def integer_step(xb, m):
    """y = xb + (xb >> 14)(c14 - 2^14): one mad24, exact in INT32."""
    c14 = ((1 << 14) + m // 2) % m - m // 2  # symmetric 2^14 mod m
    xb = np.asarray(xb, dtype=np.int64)
    return (xb >> 14) * (c14 - 16384) + xb


def float_step(y, m):
    """The FP32 part of res_byte128 on a reduced sum y, then the XOR 0x80."""
    mf = np.float32(m)
    inv = np.float32(1.0) / mf                                # correctly rounded
    cl = np.float32(np.float32(0.5) - np.float32(128.0) * inv)
    yf = np.asarray(y).astype(np.float32)
    assert np.array_equal(yf.astype(np.int64), y)             # exact conversion
    k = np.floor(fma32(yf, np.full_like(yf, inv), np.full_like(yf, cl)))
    rb = fma32(-k, np.full_like(k, mf), yf)
    assert np.all(rb == np.round(rb)) and rb.min() >= 0 and rb.max() <= 255
    return (rb.astype(np.int64) ^ 0x80).astype(np.uint8)


def kernel_byte(xb, m):
    return float_step(integer_step(xb, m), m)


def want_byte(x, m):
    r = np.mod(np.asarray(x, dtype=np.int64) + m // 2, m) - m // 2   # [-m/2, m/2)
    return (r & 0xFF).astype(np.uint8)


def test_reduction_exhaustive_over_reduced_range():
    # |x| <= 2^28 gives |xb >> 14| <= 2^14 and |y| <= 2^14 * 128 + 2^14: every
    # such y (y = xb + 128 mod m) goes through the FP32 part exactly
    lim = (1 << 14) * 128 + (1 << 14) + 1
    y = np.arange(-lim, lim + 1, dtype=np.int64)
    for m in split.moduli(16):
        assert np.array_equal(float_step(y, m), want_byte(y - 128, m)), m


def test_integer_step_range_and_congruence():
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.integers(-(1 << 28), (1 << 28) + 1, 1_000_000),
                        np.array([-(1 << 28), (1 << 28)])])
    for m in split.moduli(16):
        y = integer_step(x + 128, m)
        assert np.abs(y).max() <= (1 << 14) * 128 + (1 << 14)
        assert np.array_equal(np.mod(y - (x + 128), m), np.zeros_like(y))


def test_reduction_extreme_sums():
    # K = 2^14 products of magnitude <= 2^14: |x| <= 2^28, plus the bias
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.integers(-(1 << 28), (1 << 28) + 1, 2_000_000),
                        np.arange(-(1 << 28), -(1 << 28) + 5000),
                        np.arange((1 << 28) - 5000, (1 << 28) + 1)])
    for m in split.moduli(16):
        assert np.array_equal(kernel_byte(x + 128, m), want_byte(x, m)), m


def test_reduction_catches_a_wrong_offset():
    """The check is not vacuous: dropping the 128/m term of the offset breaks it."""
    xb = np.arange(-50000, 50000, dtype=np.int64)
    m = 255
    y = xb  # xb >> 14 is 0 or -1..-4 here; the integer step is exact either way
    c14 = ((1 << 14) + m // 2) % m - m // 2
    y = (xb >> 14) * (c14 - 16384) + xb
    yf = y.astype(np.float32)
    inv = np.float32(1.0) / np.float32(m)
    k = np.floor(fma32(yf, np.full_like(yf, inv), np.full_like(yf, np.float32(0.5))))
    rb = fma32(-k, np.full_like(k, np.float32(m)), yf)
    ok = (rb >= 0) & (rb <= 255)
    got = (rb[ok].astype(np.int64) ^ 0x80).astype(np.uint8)
    assert not (ok.all() and np.array_equal(got, want_byte(xb - 128, m)))


def mod_k(x, m):
    """mod_k in caps_amd/csrc/ozaki_dev.h: q = (x c22 + 2^21) >> 22, r = x - q m."""
    c22 = ((1 << 23) // m + 1) >> 1
    x = np.asarray(x, dtype=np.int64)
    q = (x * c22 + (1 << 21)) >> 22
    return x - q * m


def sym(x, m):
    r = np.mod(np.asarray(x, dtype=np.int64) + m // 2, m) - m // 2
    hi = (m - 1) // 2
    return np.where(r > hi, r - m, r)


def test_mod_k_exhaustive():
    x = np.arange(-8192, 8193, dtype=np.int64)
    for m in split.moduli(16):
        r = mod_k(x, m)
        assert np.array_equal(np.mod(r - x, m), np.zeros_like(x)), m
        assert r.min() >= -(m // 2) and r.max() <= m // 2, m
        # for odd m the result is the symmetric residue; for m = 256 a tie gives -128
        assert np.array_equal(r, sym(x, m)) or m % 2 == 0, m
        assert np.array_equal(r & 0xFF, sym(x, m) & 0xFF), m


def test_mod_k_fails_outside_its_range():
    """Not vacuous: far outside |x| <= 8192 the one-step quotient goes wrong."""
    x = np.arange(1 << 20, (1 << 20) + 200000, dtype=np.int64)
    m = 255
    assert not np.array_equal(mod_k(x, m), sym(x, m))


def test_sym_residue_on_large_fp64_integers():
    """sym_residue: one FP64 quotient step (x * inv rounded, rint, x - q m
    exact in one fma) leaves |r| < 12 m, then mod_k; the INT8 byte equals
    the symmetric residue of x for FP64 integers up to 2^63."""
    rng = np.random.default_rng(21)
    e = rng.integers(0, 63, 20000)
    xs = np.trunc(np.ldexp(rng.uniform(-1, 1, 20000), e))
    xs = np.concatenate([xs, [2.0 ** 63, -2.0 ** 63, 2.0 ** 62 + 2.0 ** 10, 0.0, 1.0, -1.0]])
    for m in split.moduli(16):
        inv = 1.0 / m
        q = np.rint(xs * inv)
        r = np.array([int(x) - int(qq) * m for x, qq in zip(xs, q)], dtype=np.int64)
        assert np.abs(r).max() < 12 * m
        want = np.array([((int(x) + m // 2) % m) - m // 2 for x in xs], dtype=np.int64)
        assert np.array_equal(mod_k(r, m) & 0xFF, want & 0xFF), m
