"""GPU parity of the HIP path against the CPU oracle (oracle/), through the C ABI.

Bars (written in each test):
  * integer-valued inputs (the reference's fixtures, Cfg1/Cfg3-style integer
    pairs): bit-exact in every mode;
  * real inputs with 1 x 1 leaves: bit-exact against the oracle (every leaf
    product is one rounded multiply, so all remaining arithmetic is the
    Winograd additions, whose association must match the reference);
  * real inputs otherwise: normwise error max|C - C*| / max|C*| against the
    compensated (Dot2) classical product, with the tolerance in the test.
"""
import numpy as np
import pytest

import caps_amd
import oracle

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def run(a, b, n, r, bs=1, pattern=None, mode=caps_amd.MODE_FP64, moduli=0, flags=0):
    print(f"[gpu] multiply n={n} r={r} bs={bs} pattern={pattern} mode={mode} "
          f"moduli={moduli} flags={flags}", flush=True)
    d = caps_amd.MatDescriptor.one_device(n, r, bs)
    dA, dB = _dev(a), _dev(b)
    dC = torch.empty_like(dA)
    caps_amd.multiply_device(dA, dB, dC, d, pattern, mode, moduli, flags)
    torch.cuda.synchronize()
    return dC.cpu().numpy()


def normwise(C, ref):
    return float(np.abs(C - ref).max() / np.abs(ref).max())


@pytest.fixture(scope="module", autouse=True)
def _setup():
    oracle.build()
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    caps_amd.load()
    yield
    caps_amd.release()


# ------------------------------------------------------------------ layout
@pytest.mark.parametrize("n,r,bs", [(28, 2, 7), (392, 3, 7), (64, 2, 16), (1024, 3, 1),
                                    (48, 0, 12), (56, 3, 1),
                                    # LDS-staged 64 x 64 tiles: bs dividing 64, bs a
                                    # multiple of 64, bs = leaf
                                    (512, 2, 8), (512, 1, 32), (256, 1, 2), (1024, 1, 64),
                                    (512, 0, 128), (1024, 2, 256), (768, 1, 4), (192, 0, 1)])
def test_layout_device_matches_oracle(n, r, bs):
    rng = np.random.default_rng(n + r)
    M = rng.standard_normal((n, n))
    d = caps_amd.MatDescriptor.one_device(n, r, bs)
    cm = _dev(M.ravel(order="F"))
    caps = torch.empty_like(cm)
    print(f"[gpu] layout n={n} r={r} bs={bs}", flush=True)
    caps_amd.to_layout_device(cm, caps, d)
    back = torch.empty_like(cm)
    caps_amd.from_layout_device(caps, back, d)
    torch.cuda.synchronize()
    assert np.array_equal(caps.cpu().numpy(), oracle.to_caps(M, r, bs))
    assert np.array_equal(back.cpu().numpy(), M.ravel(order="F"))
    # host-pointer entry points
    assert np.array_equal(caps_amd.distribute_from_1proc(d, M), oracle.to_caps(M, r, bs))
    assert np.array_equal(caps_amd.collect_to_1proc(d, oracle.to_caps(M, r, bs)), M)


# ------------------------------------------------------------------ fixtures
def _rows(size):
    out = []
    for nproc, bs, r in oracle.read_params(size):
        k = {1: 0, 7: 1, 49: 2, 343: 3}[nproc]
        out.append((bs, r, "b" * k + "d" * (r - k)))
    return out


_FIX = {}


def _fixture(size):
    if size not in _FIX:
        _FIX[size] = oracle.read_fixture(size)
    return _FIX[size]


@pytest.mark.parametrize("size", oracle.FIXTURE_SIZES)
def test_fixture_rows_bit_exact_fp64(size):
    """Every row of tests/golden/test<size>.params; -n 7^k maps to k leading
    'b' levels.  Exact (maximum deviation 0, README:35)."""
    A, B, C = _fixture(size)
    for bs, r, pat in _rows(size):
        a, b = oracle.to_caps(A, r, bs), oracle.to_caps(B, r, bs)
        c = run(a, b, size, r, bs, pat)
        assert np.array_equal(oracle.from_caps(c, size, r, bs), C), (bs, r, pat)
        # the planner's own pattern and all-BFS give the same bits
        c2 = run(a, b, size, r, bs, None)
        assert np.array_equal(c2, c), (bs, r, "auto")


@pytest.mark.parametrize("size", [28, 56, 196])
def test_fixture_rows_bit_exact_split(size):
    """Split precision (INT8 residues, CRT) on the integer fixtures: exact."""
    A, B, C = _fixture(size)
    for bs, r, pat in _rows(size):
        a, b = oracle.to_caps(A, r, bs), oracle.to_caps(B, r, bs)
        c = run(a, b, size, r, bs, pat, mode=caps_amd.MODE_OZAKI_INT8)
        assert np.array_equal(oracle.from_caps(c, size, r, bs), C), (bs, r, pat)


# ------------------------------------------------------------------ Winograd adds, bitwise
@pytest.mark.parametrize("n,r,pattern", [(32, 5, "bbbbb"), (16, 4, "dddd"), (32, 5, "bdbdb"),
                                         (16, 4, "dbbd")])
def test_unit_leaves_bit_exact_on_reals(n, r, pattern):
    """Leaves of size 1: each leaf product is a single rounded multiply on
    both sides, so the result pins every Winograd addition's association
    (multiply.cpp:272, 277, 195-210, 409-451) bit-for-bit."""
    rng = np.random.default_rng(1000 + n + r)
    a = rng.uniform(-1, 1, n * n) * np.ldexp(1.0, rng.integers(-20, 20, n * n))
    b = rng.uniform(-1, 1, n * n)
    ref = oracle.multiply(a, b, n, r, 1, pattern, 4)
    got = run(a, b, n, r, 1, pattern, flags=caps_amd.FLAG_NO_GRAPH)
    assert np.array_equal(got, ref)


# ------------------------------------------------------------------ configurations
@pytest.mark.parametrize("n,r,bs", [(512, 2, 8), (1024, 2, 64), (768, 1, 4), (448, 3, 7)])
def test_blocked_leaves_give_the_column_major_bits(n, r, bs):
    """1 < bs < leaf: block_multiply reorders the leaves to column-major around
    the dgemm (multiply.cpp:1093-1139); a pure permutation, so the product of
    the same matrices stored with bs blocks has the bits of the bs = 1 run."""
    rng = np.random.default_rng(n + bs)
    A = rng.uniform(-1, 1, (n, n))
    B = rng.uniform(-1, 1, (n, n))
    pat = "b" + "d" * (r - 1)
    c1 = oracle.from_caps(run(oracle.to_caps(A, r, 1), oracle.to_caps(B, r, 1), n, r, 1, pat),
                          n, r, 1)
    cb = oracle.from_caps(run(oracle.to_caps(A, r, bs), oracle.to_caps(B, r, bs), n, r, bs, pat),
                          n, r, bs)
    assert np.array_equal(c1, cb)


def test_cfg1_ddd_integers_exact_and_reals_close():
    """n=1024, r=3, 'ddd', bs=1: [-8,8] integers exact; uniform [-1,1) from
    drand48 (randombenchmark.cpp:19-22) within 1e-14 normwise of Dot2."""
    n, r = 1024, 3
    rng = np.random.default_rng(42)
    A = rng.integers(-8, 9, (n, n)).astype(np.float64)
    B = rng.integers(-8, 9, (n, n)).astype(np.float64)
    exact = oracle.igemm_exact(A, B)
    c = run(oracle.to_caps(A, r, 1), oracle.to_caps(B, r, 1), n, r, 1, "ddd")
    assert np.array_equal(oracle.from_caps(c, n, r, 1), exact)
    g = oracle.Drand48()
    a = g.uniform(n * n)
    b = g.uniform(n * n)
    c = run(a, b, n, r, 1, "ddd")
    ref = oracle.dgemm_dot2(oracle.from_caps(a, n, r, 1), oracle.from_caps(b, n, r, 1))
    err = normwise(oracle.from_caps(c, n, r, 1), ref)
    # the reference association's own error (Winograd's FP64 additions)
    o = oracle.multiply(a, b, n, r, 1, "ddd", 8)
    err_o = normwise(oracle.from_caps(o, n, r, 1), ref)
    print(f"cfg1 real normwise error {err:.3e} (oracle Winograd {err_o:.3e})")
    # same algorithm, different leaf summation order: within 25% of the oracle
    assert err <= 1.25 * err_o + 1e-16
    assert err < 5e-14


@pytest.mark.parametrize("m", [128, 256, 200, 97, 1])
def test_leaf_gemm_accuracy(m):
    """nrec = 0: the FP64 MFMA GEMM alone against Dot2, componentwise
    |C - C*| <= 2 m u (|A||B|) (standard bound gamma_m with slack)."""
    rng = np.random.default_rng(m)
    A = rng.uniform(-1, 1, (m, m))
    B = rng.uniform(-1, 1, (m, m))
    c = run(oracle.to_caps(A, 0, 1), oracle.to_caps(B, 0, 1), m, 0, 1, "")
    C = oracle.from_caps(c, m, 0, 1)
    ref = oracle.dgemm_dot2(A, B)
    bound = 2 * m * np.finfo(float).eps / 2 * (np.abs(A) @ np.abs(B))
    assert np.all(np.abs(C - ref) <= bound + 1e-300)


@pytest.mark.parametrize("n,r,pattern", [(512, 2, "bb"), (768, 2, "bd"), (1024, 3, "bdd")])
def test_fp64_reals_close_to_dot2(n, r, pattern):
    rng = np.random.default_rng(n)
    A = rng.uniform(-1, 1, (n, n))
    B = rng.uniform(-1, 1, (n, n))
    a, b = oracle.to_caps(A, r, 1), oracle.to_caps(B, r, 1)
    c = run(a, b, n, r, 1, pattern)
    ref = oracle.dgemm_dot2(A, B)
    err = normwise(oracle.from_caps(c, n, r, 1), ref)
    err_o = normwise(oracle.from_caps(oracle.multiply(a, b, n, r, 1, pattern, 8), n, r, 1), ref)
    print(f"fp64 {n} {pattern} normwise error {err:.3e} (oracle Winograd {err_o:.3e})")
    assert err <= 1.25 * err_o + 1e-16


@pytest.mark.parametrize("n,r,pattern,moduli", [(512, 2, "bb", 14), (1024, 2, "bb", 14),
                                                (640, 1, "b", 14), (256, 0, "", 16)])
def test_leaf_split_reals_close_to_dot2(n, r, pattern, moduli):
    """Leaf-only split (FP64 Winograd additions, Ozaki-II leaves) on u * 2^e
    inputs, e uniform in [-10, 10]: the leaf products add at most 1e-15 over
    the FP64 Winograd floor; with no recursion the error is <= 1e-14."""
    rng = np.random.default_rng(7 * n + r)
    A = rng.uniform(-1, 1, (n, n)) * np.ldexp(1.0, rng.integers(-10, 11, (n, n)))
    B = rng.uniform(-1, 1, (n, n)) * np.ldexp(1.0, rng.integers(-10, 11, (n, n)))
    a, b = oracle.to_caps(A, r, 1), oracle.to_caps(B, r, 1)
    c = run(a, b, n, r, 1, pattern, mode=caps_amd.MODE_OZAKI_INT8, moduli=moduli)
    ref = oracle.dgemm_dot2(A, B)
    err = normwise(oracle.from_caps(c, n, r, 1), ref)
    err_o = normwise(oracle.from_caps(oracle.multiply(a, b, n, r, 1, pattern or None, 8),
                                      n, r, 1), ref) if r else 0.0
    print(f"leaf split {n} {pattern} moduli={moduli} normwise error {err:.3e} "
          f"(oracle Winograd {err_o:.3e})")
    assert err <= err_o + 1e-15 if r else err < 1e-14


def test_split_integers_exact_bb():
    """[-64, 64] integers through two batched levels: bit-exact."""
    n, r = 1024, 2
    rng = np.random.default_rng(64)
    A = rng.integers(-64, 65, (n, n)).astype(np.float64)
    B = rng.integers(-64, 65, (n, n)).astype(np.float64)
    c = run(oracle.to_caps(A, r, 1), oracle.to_caps(B, r, 1), n, r, 1, "bb",
            mode=caps_amd.MODE_OZAKI_INT8)
    assert np.array_equal(oracle.from_caps(c, n, r, 1), oracle.igemm_exact(A, B))


# ------------------------------------------------------------------ graphs, counters, errors
def test_graph_replays_match_eager():
    n, r = 512, 2
    rng = np.random.default_rng(5)
    a = rng.uniform(-1, 1, n * n)
    b = rng.uniform(-1, 1, n * n)
    d = caps_amd.MatDescriptor.one_device(n, r, 1)
    dA, dB = _dev(a), _dev(b)
    eager = torch.empty_like(dA)
    caps_amd.multiply_device(dA, dB, eager, d, "bd", flags=caps_amd.FLAG_NO_GRAPH)
    torch.cuda.synchronize()
    ref = eager.cpu().numpy()
    dC = torch.empty_like(dA)
    for rep in range(4):
        dC.fill_(np.nan)
        caps_amd.multiply_device(dA, dB, dC, d, "bd")   # captured once, replayed
        x = torch.ones(1024, device="cuda").sum()       # unrelated eager work between replays
        torch.cuda.synchronize()
        assert np.array_equal(dC.cpu().numpy(), ref), rep
        assert float(x) == 1024.0


def test_counters_match_oracle():
    n, r = 256, 3
    a = np.zeros(n * n)
    for pat in ("ddd", "bdb"):
        caps_amd.reset_counters()
        oracle.reset_counters()
        run(a, a, n, r, 1, pat)
        oracle.multiply(a, a, n, r, 1, pat, 2)
        c = caps_amd.counters()
        adds, mults = oracle.counters()
        assert (c["additions"], c["multiplications"]) == (adds, mults)
        assert c["leaf_products"] == 7 ** r
        assert c["pattern"] == ",".join("BFS" if ch == "b" else "DFS" for ch in pat) + ","


def test_errors_are_reported_not_fatal():
    a = np.zeros(64)
    d = caps_amd.MatDescriptor.one_device(8, 2, 1)
    with pytest.raises(caps_amd.CapsError) as e:
        caps_amd.multiply(a, a, np.zeros(64), d, "dx")
    assert e.value.code == 2
    with pytest.raises(caps_amd.CapsError) as e:
        caps_amd.multiply(a, a, np.zeros(64), d, "d")
    assert e.value.code == 2
    with pytest.raises(caps_amd.CapsError) as e:
        caps_amd.multiply(a, a, np.zeros(64), d, "dd", mode=caps_amd.MODE_OZAKI_INT8, moduli=40)
    assert e.value.code == 1
    # the library still works afterwards
    out = np.zeros(64)
    caps_amd.multiply(np.ones(64), np.ones(64), out, d, "dd")
    assert np.all(out == 8.0)


# ------------------------------------------------------------------ recursion cutoff
@pytest.mark.parametrize("n,nrec,cutoff,pattern", [(512, 2, 2, ""), (512, 2, 1, "b"),
                                                   (512, 2, 1, "d"), (1024, 3, 2, "b")])
def test_cutoff_equals_shallower_layout(n, nrec, cutoff, pattern):
    """Stopping the recursion `cutoff` levels early (classical GEMM over the
    quadtree of leaves) gives the same bits as the layout with nrec - cutoff
    levels and full recursion: same pre-additions, same leaf summation order."""
    rng = np.random.default_rng(n + cutoff)
    A = rng.uniform(-1, 1, (n, n))
    B = rng.uniform(-1, 1, (n, n))
    d = caps_amd.MatDescriptor.one_device(n, nrec, 1)
    dA, dB = _dev(oracle.to_caps(A, nrec, 1)), _dev(oracle.to_caps(B, nrec, 1))
    dC = torch.empty_like(dA)
    print(f"[gpu] cutoff n={n} nrec={nrec} cutoff={cutoff} pattern={pattern}", flush=True)
    caps_amd.multiply_device(dA, dB, dC, d, pattern + "b" * cutoff, cutoff=cutoff)
    torch.cuda.synchronize()
    got = oracle.from_caps(dC.cpu().numpy(), n, nrec, 1)
    r2 = nrec - cutoff
    ref = oracle.from_caps(run(oracle.to_caps(A, r2, 1), oracle.to_caps(B, r2, 1), n, r2, 1,
                               pattern), n, r2, 1)
    assert np.array_equal(got, ref)


def test_cutoff_rejected_where_unsupported():
    d = caps_amd.MatDescriptor.one_device(256, 2, 1)   # leaves of 64: not a multiple of 128
    x = torch.zeros(256 * 256, dtype=torch.float64, device="cuda")
    with pytest.raises(caps_amd.CapsError) as e:
        caps_amd.multiply_device(x, x, x.clone(), d, "bb", cutoff=1)
    assert e.value.code == 6
    with pytest.raises(caps_amd.CapsError) as e:
        caps_amd.multiply_device(x, x, x.clone(), d, "bb", cutoff=3)
    assert e.value.code == 1


# ------------------------------------------------------------------ fused two-level passes
@pytest.mark.parametrize("n,nrec,pattern", [(512, 2, "bb"), (1024, 3, "bbd"), (256, 4, "bbbb"),
                                            (32, 5, "bbbbb")])
def test_fused_levels_bit_identical_to_unfused(n, nrec, pattern):
    """The fused two-level pre/post passes evaluate the same expression trees
    as two separate 'b' levels: identical bits on real inputs."""
    rng = np.random.default_rng(n + nrec)
    a = rng.uniform(-1, 1, n * n) * np.ldexp(1.0, rng.integers(-8, 8, n * n))
    b = rng.uniform(-1, 1, n * n)
    fused = run(a, b, n, nrec, 1, pattern)
    plain = run(a, b, n, nrec, 1, pattern, flags=caps_amd.FLAG_NO_FUSION)
    assert np.array_equal(fused, plain)
    if n <= 64:
        assert np.array_equal(fused, oracle.multiply(a, b, n, nrec, 1, pattern, 4))


def test_deep_dfs_plan_runs_without_graph():
    """A pattern with 7^5 sequential leaf products (about 33k launches) is
    above the graph-capture limit and runs as direct launches: same bits as
    the oracle on integers."""
    n, r, pat = 1024, 5, "ddddd"
    rng = np.random.default_rng(55)
    A = rng.integers(-8, 9, (n, n)).astype(np.float64)
    B = rng.integers(-8, 9, (n, n)).astype(np.float64)
    a, b = oracle.to_caps(A, r, 1), oracle.to_caps(B, r, 1)
    got = run(a, b, n, r, 1, pat)
    want = oracle.multiply(a, b, n, r, 1, pat, 8)
    assert np.array_equal(got, want)


def test_fp64_bdd_full_size():
    """Plain FP64 Strassen at its full benchmark size: n=7168, r=3, 'bdd'
    (the BFS level as 7 batched sub-products, FP64 MFMA leaves of 896^2),
    entries uniform [-1, 1).  Size-independent properties: the batched level
    gives the same bits as the recursive one ('ddd'), since the FP64 adds use
    the same association; the normwise error against the 16-modulus residue
    product (correctly rounded for these inputs) stays within 1e-13, a bound
    that a dropped Winograd term or a swapped operand would exceed by far."""
    n, r = 7168, 3
    print(f"[gpu] full-size FP64 n={n} r={r} bdd / ddd", flush=True)
    d = caps_amd.MatDescriptor.one_device(n, r, 1)
    g = torch.Generator(device="cuda")
    g.manual_seed(2024)
    A = torch.rand(n * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    B = torch.rand(n * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    C1, C2, ref = torch.empty_like(A), torch.empty_like(A), torch.empty_like(A)
    caps_amd.multiply_device(A, B, C1, d, "bdd", caps_amd.MODE_FP64)
    caps_amd.multiply_device(A, B, C2, d, "ddd", caps_amd.MODE_FP64)
    caps_amd.reference_product(A, B, ref, d)
    torch.cuda.synchronize()
    assert torch.equal(C1, C2)
    err = caps_amd.normwise_error(C1, ref, d)
    print(f"[gpu] n={n} bdd normwise error {err:.3e}", flush=True)
    assert 0 < err <= 1e-13
    del A, B, C1, C2, ref
    caps_amd.release()
    torch.cuda.empty_cache()
