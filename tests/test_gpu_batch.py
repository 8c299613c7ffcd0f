"""GPU parity of the uneven batch entry point (caps_multiply_batch_device):
a set of independent problems of different sizes, depths and patterns runs as
one captured graph.  FP64: on integer inputs each result is the oracle's, bit
for bit; on real inputs it is bit-identical to the same problem run alone
(the leaf GEMM accumulates in MFMA order, so real-data FP64 results match the
oracle only to rounding).  Residue-domain mode: the exact-integer model's
bits on real inputs."""
import ctypes

import numpy as np
import pytest

import caps_amd
import oracle
from oracle import split

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _setup():
    oracle.build()
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    caps_amd.load()
    yield
    caps_amd.release()


# (n, nrec, bs, pattern): repeated shapes share batched launches
ITEMS = [(64, 2, 1, "bb"), (128, 2, 1, "bd"), (64, 2, 1, "bb"), (56, 3, 1, "bbb"),
         (192, 1, 1, "b"), (128, 2, 8, "dd"), (64, 2, 1, "bb")]


def _problems(seed, integers=False):
    rng = np.random.default_rng(seed)
    out = []
    for n, r, bs, pat in ITEMS:
        if integers:
            A = rng.integers(-64, 65, (n, n)).astype(np.float64)
            B = rng.integers(-64, 65, (n, n)).astype(np.float64)
        else:
            A = rng.uniform(-1, 1, (n, n)) * np.ldexp(1.0, rng.integers(-8, 9, (n, n)))
            B = rng.uniform(-1, 1, (n, n)) * np.ldexp(1.0, rng.integers(-8, 9, (n, n)))
        out.append((A, B, n, r, bs, pat))
    return out


def _run_batch(probs, mode, moduli=0, flags=0):
    print(f"[gpu] batch of {len(probs)} mode={mode} flags={flags}", flush=True)
    As, Bs, Cs, descs, pats = [], [], [], [], []
    for A, B, n, r, bs, pat in probs:
        As.append(torch.from_numpy(oracle.to_caps(A, r, bs)).cuda())
        Bs.append(torch.from_numpy(oracle.to_caps(B, r, bs)).cuda())
        Cs.append(torch.full_like(As[-1], float("nan")))
        descs.append(caps_amd.MatDescriptor.one_device(n, r, bs))
        pats.append(pat)
    caps_amd.multiply_batch_device(As, Bs, Cs, descs, pats, mode, moduli, flags)
    torch.cuda.synchronize()
    return [oracle.from_caps(c.cpu().numpy(), p[2], p[3], p[4]) for c, p in zip(Cs, probs)]


def _run_single(A, B, n, r, bs, pat, mode=caps_amd.MODE_FP64):
    d = caps_amd.MatDescriptor.one_device(n, r, bs)
    dA = torch.from_numpy(oracle.to_caps(A, r, bs)).cuda()
    dB = torch.from_numpy(oracle.to_caps(B, r, bs)).cuda()
    dC = torch.empty_like(dA)
    caps_amd.multiply_device(dA, dB, dC, d, pat, mode)
    torch.cuda.synchronize()
    return oracle.from_caps(dC.cpu().numpy(), n, r, bs)


@pytest.mark.parametrize("flags", [0, caps_amd.FLAG_NO_GRAPH])
def test_batch_fp64_integers_match_oracle(flags):
    probs = _problems(1, integers=True)
    got = _run_batch(probs, caps_amd.MODE_FP64, 0, flags)
    for (A, B, n, r, bs, pat), g in zip(probs, got):
        want = oracle.from_caps(oracle.multiply(oracle.to_caps(A, r, bs), oracle.to_caps(B, r, bs),
                                                n, r, bs, pat, 2), n, r, bs)
        assert np.array_equal(g, want), (n, r, bs, pat)


def test_batch_fp64_reals_match_single_calls():
    probs = _problems(4)
    got = _run_batch(probs, caps_amd.MODE_FP64)
    for (A, B, n, r, bs, pat), g in zip(probs, got):
        one = _run_single(A, B, n, r, bs, pat)
        assert np.array_equal(g, one), (n, r, bs, pat, np.abs(g - one).max())
        want = oracle.from_caps(oracle.multiply(oracle.to_caps(A, r, bs), oracle.to_caps(B, r, bs),
                                                n, r, bs, pat, 2), n, r, bs)
        assert np.abs(g - want).max() <= 1e-13 * np.abs(want).max()


def test_batch_crt_matches_model():
    probs = _problems(2)
    got = _run_batch(probs, caps_amd.MODE_CRT_INT8, 14)
    for (A, B, n, r, bs, pat), g in zip(probs, got):
        assert np.array_equal(g, split.crt_mode(A, B, 14)), (n, r, bs, pat)


def test_batch_replay_and_single_item():
    """Replaying the cached batch graph on new data, and a batch of one."""
    probs = _problems(3)
    first = _run_batch(probs, caps_amd.MODE_FP64)
    again = _run_batch(probs, caps_amd.MODE_FP64)
    for a, b in zip(first, again):
        assert np.array_equal(a, b)
    A, B, n, r, bs, pat = probs[1]
    one = _run_batch([probs[1]], caps_amd.MODE_FP64)[0]
    assert np.array_equal(one, first[1])


def test_batch_rejects_auto_and_bad_items():
    d = caps_amd.MatDescriptor.one_device(64, 2, 1)
    x = torch.zeros(64 * 64, dtype=torch.float64, device="cuda")
    with pytest.raises(caps_amd.CapsError):
        caps_amd.multiply_batch_device([x], [x], [x], [d], None, caps_amd.MODE_AUTO)
    bad = caps_amd.MatDescriptor.one_device(64, 2, 3)   # 16 % 3 != 0
    with pytest.raises(caps_amd.CapsError):
        caps_amd.multiply_batch_device([x], [x], [x], [bad], None, caps_amd.MODE_FP64)


def test_batch_per_item_cutoff_and_mode_agreement():
    """Per-item cutoffs: FP64 results equal single calls with the same
    cutoff; residue-domain results equal the exact model.  Items whose modes
    differ are rejected."""
    rng = np.random.default_rng(12)
    spec = [(512, 2, "bb", 0), (512, 2, "bb", 1), (256, 1, "b", 0), (512, 2, "db", 1)]
    probs = []
    for n, r, pat, cut in spec:
        A = rng.uniform(-1, 1, (n, n)) * np.ldexp(1.0, rng.integers(-8, 9, (n, n)))
        B = rng.uniform(-1, 1, (n, n)) * np.ldexp(1.0, rng.integers(-8, 9, (n, n)))
        probs.append((A, B, n, r, pat, cut))

    def run_batch(mode, moduli):
        print(f"[gpu] batch per-item cutoff mode={mode}", flush=True)
        As = [torch.from_numpy(oracle.to_caps(A, r, 1)).cuda() for A, B, n, r, p, c in probs]
        Bs = [torch.from_numpy(oracle.to_caps(B, r, 1)).cuda() for A, B, n, r, p, c in probs]
        Cs = [torch.empty_like(a) for a in As]
        descs = [caps_amd.MatDescriptor.one_device(n, r, 1) for A, B, n, r, p, c in probs]
        caps_amd.multiply_batch_device(As, Bs, Cs, descs, [p[4] for p in probs], mode, moduli,
                                       cutoffs=[p[5] for p in probs])
        torch.cuda.synchronize()
        return [oracle.from_caps(c.cpu().numpy(), p[2], p[3], 1) for c, p in zip(Cs, probs)]

    for (A, B, n, r, pat, cut), g in zip(probs, run_batch(caps_amd.MODE_FP64, 0)):
        d = caps_amd.MatDescriptor.one_device(n, r, 1)
        dA = torch.from_numpy(oracle.to_caps(A, r, 1)).cuda()
        dB = torch.from_numpy(oracle.to_caps(B, r, 1)).cuda()
        dC = torch.empty_like(dA)
        caps_amd.multiply_device(dA, dB, dC, d, pat, caps_amd.MODE_FP64, cutoff=cut)
        torch.cuda.synchronize()
        assert np.array_equal(g, oracle.from_caps(dC.cpu().numpy(), n, r, 1)), (n, pat, cut)
    for (A, B, n, r, pat, cut), g in zip(probs, run_batch(caps_amd.MODE_CRT_INT8, 14)):
        assert np.array_equal(g, split.crt_mode(A, B, 14)), (n, pat, cut)

    x = torch.zeros(64 * 64, dtype=torch.float64, device="cuda")
    d = caps_amd.MatDescriptor.one_device(64, 2, 1)
    O = (caps_amd.Options * 2)(caps_amd._opts(caps_amd.MODE_FP64, 0, 0),
                               caps_amd._opts(caps_amd.MODE_CRT_INT8, 14, 0))
    P = ctypes.c_void_p * 2
    p = P(x.data_ptr(), x.data_ptr())
    D = (caps_amd.MatDescriptor * 2)(d, d)
    assert caps_amd.load().caps_multiply_batch_device(2, p, p, p, D, None, O, None) != 0
    assert b"same mode" in caps_amd.load().caps_last_error()
