"""Autotuner and AUTO-mode planner on the GPU: every measured candidate is a
correct configuration (checked against the high-precision reference), the
choice meets the error target, and AUTO calls reuse the tuned choice."""
import numpy as np
import pytest

import caps_amd
import oracle
from oracle import split

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _inputs(n, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    u = torch.rand(n * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    e = torch.randint(-10, 11, (n * n,), generator=g, device="cuda", dtype=torch.int32)
    return torch.ldexp(u, e.to(torch.float64))


def test_reference_product_accuracy():
    n, r = 512, 2
    A, B = _inputs(n, 1), _inputs(n, 2)
    d = caps_amd.MatDescriptor.one_device(n, r, 1)
    ref = torch.empty_like(A)
    print("[gpu] reference product n=512", flush=True)
    caps_amd.reference_product(A, B, ref, d)
    torch.cuda.synchronize()
    Am = oracle.from_caps(A.cpu().numpy(), n, r, 1)
    Bm = oracle.from_caps(B.cpu().numpy(), n, r, 1)
    dot2 = oracle.dgemm_dot2(Am, Bm)
    got = oracle.from_caps(ref.cpu().numpy(), n, r, 1)
    # the reference product is the 16-moduli residue product: exact integer
    # product of the scaled operands, rounded once
    assert np.array_equal(got, split.crt_mode(Am, Bm, 16))
    # Dot2 is accurate to about one rounding, so the two agree to 1 ulp of max|C|
    assert np.abs(got - dot2).max() / np.abs(dot2).max() <= 2.0 ** -52
    # the device error measure agrees with numpy on a perturbed copy
    C = ref.clone()
    C[123] += 1e-9 * float(ref.abs().max())
    err = caps_amd.normwise_error(C, ref, d)
    assert abs(err - 1e-9) < 1e-12


def test_autotune_picks_a_valid_fast_candidate():
    n, r = 1024, 3
    A, B = _inputs(n, 3), _inputs(n, 4)
    C = torch.empty_like(A)
    d = caps_amd.MatDescriptor.one_device(n, r, 1)
    print("[gpu] autotune n=1024 nrec=3", flush=True)
    best, all_ = caps_amd.autotune(A, B, C, d, caps_amd.MASK_FP64 | caps_amd.MASK_CRT_INT8,
                                   1e-14, reps=2)
    assert len(all_) >= 10
    for c in all_:
        assert c.ms > 0 and c.err >= 0
        if c.mode == caps_amd.MODE_CRT_INT8 and c.moduli >= 14:
            assert c.err < 1e-14          # residue mode meets the bar
    ok = [c for c in all_ if c.err <= 1e-14]
    assert best.ms == min(c.ms for c in ok)
    # AUTO calls now use the tuned configuration
    caps_amd.planner_config(1e-14, caps_amd.MASK_FP64 | caps_amd.MASK_CRT_INT8)
    caps_amd.reset_counters()
    caps_amd.multiply_device(A, B, C, d, None, mode=caps_amd.MODE_AUTO)
    torch.cuda.synchronize()
    ref = torch.empty_like(A)
    caps_amd.reference_product(A, B, ref, d)
    assert caps_amd.normwise_error(C, ref, d) <= 1e-14
