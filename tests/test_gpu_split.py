"""GPU parity of the two split-precision modes through the C ABI.

  * CAPS_MODE_CRT_INT8 (residue-domain Strassen): bit-exact against the
    exact-integer model oracle/split.py on real inputs, for every pattern;
    bit-exact on integers (fixtures, [-64, 64]); normwise error <= 1e-14
    against the compensated (Dot2) product on u * 2^e inputs.
  * CAPS_MODE_OZAKI_INT8 (FP64 Winograd additions, split leaves): bit-exact
    against the model's FP64 Winograd with scaled exact leaves.
"""
import numpy as np
import pytest

import caps_amd
import oracle
from oracle import split

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

CRT = caps_amd.MODE_CRT_INT8
LEAF = caps_amd.MODE_OZAKI_INT8


def run(a, b, n, r, bs=1, pattern=None, mode=CRT, moduli=0, flags=0):
    print(f"[gpu] multiply n={n} r={r} bs={bs} pattern={pattern} mode={mode} "
          f"moduli={moduli} flags={flags}", flush=True)
    d = caps_amd.MatDescriptor.one_device(n, r, bs)
    dA = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    dB = torch.from_numpy(np.ascontiguousarray(b)).cuda()
    dC = torch.empty_like(dA)
    caps_amd.multiply_device(dA, dB, dC, d, pattern, mode, moduli, flags)
    torch.cuda.synchronize()
    return dC.cpu().numpy()


def realdata(n, seed, spread=10):
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1, 1, (n, n)) * np.ldexp(1.0, rng.integers(-spread, spread + 1, (n, n)))
    B = rng.uniform(-1, 1, (n, n)) * np.ldexp(1.0, rng.integers(-spread, spread + 1, (n, n)))
    return A, B


@pytest.fixture(scope="module", autouse=True)
def _setup():
    oracle.build()
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    caps_amd.load()
    yield
    caps_amd.release()


@pytest.mark.parametrize("n,r,bs,patterns,moduli", [
    (64, 2, 1, ["bb", "dd", "bd", "db"], 14),
    (96, 1, 1, ["b", "d"], 14),
    (128, 2, 8, ["bb", "dh"], 16),     # bs-blocked leaves
    (56, 3, 1, ["bbb", "ddd"], 12),    # odd leaves (m = 7)
    (256, 1, 1, ["b"], 14),            # 128 x 128 leaves: the fast GEMM path
    (512, 1, 1, ["b", "d"], 14),       # 256 x 256 leaves: the large-tile GEMM
    (384, 1, 1, ["b", "d"], 14),       # 192 x 192 leaves: whole K slices, edge rows checked
    (256, 2, 64, ["bb", "bd"], 14),    # bs = leaf size; 'bb': B's split fused with the adds
    (512, 2, 128, ["bb", "bd"], 14),   # bs = leaf size = 128: A's split fused with the adds too
    (232, 2, 1, ["bb"], 14),           # even leaves (m = 58), not a multiple of 4
    (160, 1, 1, ["b", "d"], 14),       # m = 80: 16-byte K chunks checked
    (2 * 544, 1, 1, ["b"], 14),        # m = 544 (17 x 32), as at n = 4352, r = 3
])
def test_crt_mode_bit_exact_vs_model(n, r, bs, patterns, moduli):
    A, B = realdata(n, 100 + n)
    want = split.crt_mode(A, B, moduli)
    a, b = oracle.to_caps(A, r, bs), oracle.to_caps(B, r, bs)
    for pat in patterns:
        got = oracle.from_caps(run(a, b, n, r, bs, pat, CRT, moduli), n, r, bs)
        assert np.array_equal(got, want), (pat, np.abs(got - want).max())


def test_leaf_split_mode_bit_exact_vs_model():
    n, r = 64, 2
    A, B = realdata(n, 7)
    want = split.leaf_split_mode(A, B, r, 14)
    for pat in ("bb", "dd"):
        got = oracle.from_caps(run(oracle.to_caps(A, r, 1), oracle.to_caps(B, r, 1), n, r, 1,
                                   pat, LEAF, 14), n, r, 1)
        assert np.array_equal(got, want), pat


@pytest.mark.parametrize("size", oracle.FIXTURE_SIZES)
def test_crt_mode_fixtures_exact(size):
    A, B, C = oracle.read_fixture(size)
    for nproc, bs, r in oracle.read_params(size):
        k = {1: 0, 7: 1, 49: 2, 343: 3}[nproc]
        pat = "b" * k + "d" * (r - k)
        c = run(oracle.to_caps(A, r, bs), oracle.to_caps(B, r, bs), size, r, bs, pat, CRT)
        assert np.array_equal(oracle.from_caps(c, size, r, bs), C), (bs, r, pat)


def test_crt_mode_integers_exact_bb():
    n, r = 1024, 2
    rng = np.random.default_rng(64)
    A = rng.integers(-64, 65, (n, n)).astype(np.float64)
    B = rng.integers(-64, 65, (n, n)).astype(np.float64)
    c = run(oracle.to_caps(A, r, 1), oracle.to_caps(B, r, 1), n, r, 1, "bb", CRT)
    assert np.array_equal(oracle.from_caps(c, n, r, 1), oracle.igemm_exact(A, B))


@pytest.mark.parametrize("n,r,pattern", [(1024, 2, "bb"), (1024, 3, "bdd"), (768, 1, "d")])
def test_crt_mode_error_below_bar(n, r, pattern):
    """u * 2^e inputs, e in [-10, 10]: normwise error <= 1e-14 (the metric's
    bar) -- and below the FP64 Winograd floor of the same pattern."""
    A, B = realdata(n, n + r)
    a, b = oracle.to_caps(A, r, 1), oracle.to_caps(B, r, 1)
    ref = oracle.dgemm_dot2(A, B)
    got = oracle.from_caps(run(a, b, n, r, 1, pattern, CRT, 14), n, r, 1)
    err = float(np.abs(got - ref).max() / np.abs(ref).max())
    fp64 = oracle.from_caps(oracle.multiply(a, b, n, r, 1, pattern, 8), n, r, 1)
    err_w = float(np.abs(fp64 - ref).max() / np.abs(ref).max())
    print(f"crt {n} {pattern}: error {err:.3e}, FP64 Winograd {err_w:.3e}")
    assert err <= 1e-14
    assert err < err_w


def test_crt_mode_pattern_independent():
    n, r = 512, 2
    A, B = realdata(n, 77)
    a, b = oracle.to_caps(A, r, 1), oracle.to_caps(B, r, 1)
    ref = run(a, b, n, r, 1, "bb", CRT)
    for pat in ("dd", "bd", "db", "hb"):
        assert np.array_equal(run(a, b, n, r, 1, pat, CRT), ref), pat


def test_crt_mode_cutoff_bit_exact_vs_model():
    n, nrec = 512, 2
    A, B = realdata(n, 31)
    want = split.crt_mode(A, B, 14)
    d = caps_amd.MatDescriptor.one_device(n, nrec, 1)
    dA = torch.from_numpy(oracle.to_caps(A, nrec, 1)).cuda()
    dB = torch.from_numpy(oracle.to_caps(B, nrec, 1)).cuda()
    for cutoff, pat in ((1, "bb"), (2, "bb"), (1, "db")):
        dC = torch.empty_like(dA)
        print(f"[gpu] crt cutoff={cutoff} pattern={pat}", flush=True)
        caps_amd.multiply_device(dA, dB, dC, d, pat, mode=CRT, moduli=14, cutoff=cutoff)
        torch.cuda.synchronize()
        assert np.array_equal(oracle.from_caps(dC.cpu().numpy(), n, nrec, 1), want), cutoff


def test_crt_large_tile_gemm_bit_exact_vs_model():
    """Leaves of 256 and quadtree operands of 512 (cutoff) take the 256 x 256
    tile residue GEMM (under 'bb' with no cutoff it writes the 49 products of a
    modulus interleaved in 128-byte segments): same bits as the exact-integer
    model."""
    n, nrec = 1024, 2
    A, B = realdata(n, 41)
    want = split.crt_mode(A, B, 14)
    d = caps_amd.MatDescriptor.one_device(n, nrec, 1)
    dA = torch.from_numpy(oracle.to_caps(A, nrec, 1)).cuda()
    dB = torch.from_numpy(oracle.to_caps(B, nrec, 1)).cuda()
    for cutoff, pat in ((0, "bb"), (0, "db"), (1, "bb"), (1, "db")):
        dC = torch.empty_like(dA)
        print(f"[gpu] crt n=1024 cutoff={cutoff} pattern={pat}", flush=True)
        caps_amd.multiply_device(dA, dB, dC, d, pat, mode=CRT, moduli=14, cutoff=cutoff)
        torch.cuda.synchronize()
        got = oracle.from_caps(dC.cpu().numpy(), n, nrec, 1)
        assert np.array_equal(got, want), (cutoff, pat, np.abs(got - want).max())


@pytest.mark.parametrize("n,nrec,pattern", [(512, 2, "bb"), (1024, 3, "bbd"), (256, 3, "bbb"),
                                             (56, 3, "bbb"), (512, 2, "bd"),
                                             # leaves of 128, 4 x 4 of them per depth-2 sub-block
                                             (2048, 4, "bbdd")])
def test_crt_fused_levels_bit_exact(n, nrec, pattern):
    """Fused two-level residue passes with the reconstruction folded into the
    combinations: same bits as the unfused residue recursion and the model."""
    A, B = realdata(n, n + 5)
    a, b = oracle.to_caps(A, nrec, 1), oracle.to_caps(B, nrec, 1)
    fused = run(a, b, n, nrec, 1, pattern, CRT, 14)
    plain = run(a, b, n, nrec, 1, pattern, CRT, 14, flags=caps_amd.FLAG_NO_FUSION)
    assert np.array_equal(fused, plain)
    if n <= 512:
        assert np.array_equal(oracle.from_caps(fused, n, nrec, 1), split.crt_mode(A, B, 14))


BF16 = caps_amd.MODE_CRT_BF16


@pytest.mark.parametrize("n,r,patterns", [(256, 1, ["b"]), (384, 1, ["b", "d"]), (56, 3, ["bbb"]),
                                          (224, 1, ["b"])])
def test_crt_bf16_bit_exact_vs_model(n, r, patterns):
    """BF16 leaf products (FP32 accumulation folded mod m): the exact model's
    bits, on the fast, edge-checked and byte-checked GEMM paths."""
    A, B = realdata(n, 300 + n)
    want = split.crt_mode(A, B, 14)
    a, b = oracle.to_caps(A, r, 1), oracle.to_caps(B, r, 1)
    for pat in patterns:
        got = oracle.from_caps(run(a, b, n, r, 1, pat, BF16, 14), n, r, 1)
        assert np.array_equal(got, want), pat


@pytest.mark.parametrize("n,r,pattern,cutoff", [(4096, 1, "b", 0), (2048, 2, "bb", 1),
                                                (1536, 2, "bd", 0),
                                                # leaves of 256 under 'bb': both 256-tile
                                                # kernels write the products interleaved
                                                (1024, 2, "bb", 0),
                                                # leaves of 640 and 1344: the INT8 side takes
                                                # the 256-tile kernel with partial edges
                                                (2560, 2, "bb", 0), (2688, 1, "b", 0)])
def test_crt_bf16_equals_int8_mode(n, r, pattern, cutoff):
    """Large leaves (K up to 2048: several FP32 folds) and the quadtree
    operands of a cutoff: the BF16 and INT8 leaf GEMMs give the same bits."""
    g = torch.Generator(device="cuda")
    g.manual_seed(n + cutoff)
    u = torch.rand(2 * n * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    e = torch.randint(-10, 11, (2 * n * n,), generator=g, device="cuda", dtype=torch.int32)
    x = torch.ldexp(u, e.to(torch.float64))
    dA, dB = x[: n * n].contiguous(), x[n * n:].contiguous()
    d = caps_amd.MatDescriptor.one_device(n, r, 1)
    outs = []
    for mode in (CRT, BF16):
        print(f"[gpu] crt n={n} r={r} pattern={pattern} cutoff={cutoff} mode={mode}", flush=True)
        dC = torch.empty_like(dA)
        caps_amd.multiply_device(dA, dB, dC, d, pattern, mode=mode, moduli=14, cutoff=cutoff)
        torch.cuda.synchronize()
        outs.append(dC.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("n,nrec,pattern,cutoff", [(2048, 3, "bbb", 1), (4096, 2, "bb", 0)])
def test_crt_interleaved_products_equal_plain(n, nrec, pattern, cutoff):
    """Under two leading 'b' levels the 256-tile leaf GEMMs write the 49
    depth-2 products of a modulus interleaved in 128-byte segments (here on
    2 x 2 quadtrees of 256-leaves below a cutoff, and on leaves of 1024); the
    unfused path keeps plain product planes.  INT8 and BF16 leaf products, fused
    and unfused, give the same bits, within the error bar of the 16-modulus
    reference product."""
    g = torch.Generator(device="cuda")
    g.manual_seed(n + 7 * nrec)
    u = torch.rand(2 * n * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    e = torch.randint(-10, 11, (2 * n * n,), generator=g, device="cuda", dtype=torch.int32)
    x = torch.ldexp(u, e.to(torch.float64))
    dA, dB = x[: n * n].contiguous(), x[n * n:].contiguous()
    d = caps_amd.MatDescriptor.one_device(n, nrec, 1)
    outs = []
    for mode, flags in ((CRT, caps_amd.FLAG_NO_FUSION), (CRT, 0), (BF16, 0)):
        print(f"[gpu] crt n={n} nrec={nrec} pattern={pattern} cutoff={cutoff} mode={mode} "
              f"flags={flags}", flush=True)
        dC = torch.empty_like(dA)
        caps_amd.multiply_device(dA, dB, dC, d, pattern, mode=mode, moduli=14, flags=flags,
                                 cutoff=cutoff)
        torch.cuda.synchronize()
        outs.append(dC)
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[2])
    ref = torch.empty_like(dA)
    caps_amd.reference_product(dA, dB, ref, d)
    torch.cuda.synchronize()
    assert caps_amd.normwise_error(outs[1], ref, d) <= 1e-14

