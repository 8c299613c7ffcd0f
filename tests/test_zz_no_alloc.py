"""The no-allocation entry point (caps_multiply_device_noalloc) against
caps_multiply_device on the same operands: bit-identical C, on two calls in a
row (the second smaller, so the scratch holds the first call's contents), in
FP64 and in the residue mode, plus the size limit (a plan that does not fit
is refused with CAPS_ENOMEM and the next call still works).  Each call uses
the library's fixed per-device scratch; nothing else runs while it does."""
import pytest

import caps_amd

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _setup():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    caps_amd.load()
    yield
    caps_amd.release()


def _operands(n, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    u = torch.rand(n * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    e = torch.randint(-10, 11, (n * n,), generator=g, device="cuda", dtype=torch.int32)
    A = torch.ldexp(u, e.to(torch.float64))
    B = torch.rand(n * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    return A, B


def _pair(n, r, pattern, mode, seed, flags=0):
    d = caps_amd.MatDescriptor.one_device(n, r, 1)
    A, B = _operands(n, seed)
    want = torch.full_like(A, float("nan"))
    got = torch.full_like(A, float("nan"))
    print(f"[gpu] noalloc n={n} r={r} {pattern} mode={mode} flags={flags}", flush=True)
    caps_amd.multiply_device(A, B, want, d, pattern, mode=mode, flags=flags)
    caps_amd.multiply_device_noalloc(A, B, got, d, pattern, mode=mode, flags=flags)
    torch.cuda.synchronize()
    return want, got


@pytest.mark.parametrize("mode,first,second", [
    (caps_amd.MODE_FP64, (1024, 2, "bb"), (512, 2, "bd")),
    (caps_amd.MODE_CRT_INT8, (1024, 2, "bb"), (768, 1, "b")),
    (caps_amd.MODE_FP64, (2048, 2, "bd"), (448, 3, "dbd")),
])
def test_noalloc_matches_multiply_device(mode, first, second):
    for i, (n, r, p) in enumerate((first, second)):
        want, got = _pair(n, r, p, mode, 10 + i)
        assert torch.equal(want, got), (mode, n, r, p)


def test_noalloc_rejects_what_does_not_fit():
    n, r = 4096, 2
    d = caps_amd.MatDescriptor.one_device(n, r, 1)
    A, B = _operands(n, 3)
    C = torch.empty_like(A)
    with pytest.raises(caps_amd.CapsError) as e:
        caps_amd.multiply_device_noalloc(A, B, C, d, "bb")
    assert e.value.code == 3 and "scratch" in str(e.value)
    with pytest.raises(caps_amd.CapsError):
        caps_amd.multiply_device_noalloc(A, B, C, d, "bb", mode=caps_amd.MODE_AUTO)
    # a call that fits still works afterwards
    want, got = _pair(512, 2, "bb", caps_amd.MODE_CRT_INT8, 4)
    assert torch.equal(want, got)
