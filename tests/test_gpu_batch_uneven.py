"""The uneven batch at its full workload: 512 independent square problems,
n = 64 round(exp(U[ln 128, ln 4096]) / 64) (seed 2024), entries uniform
[-1, 1), all in one caps_multiply_batch_device call (one captured graph), with
each size's layout depth, b/d pattern and cutoff from the shipped learned
planner (caps_amd/planner_model.npz).  Bars:
  * FP64: every result bit-identical to the same problem multiplied alone
    (same plan, one call per problem) - batching changes no arithmetic;
  * residue-domain INT8: normwise error <= 1e-14 against the 16-modulus
    reference product for every problem, and against the compensated (Dot2)
    classical product of the CPU oracle on sampled rows of 8 problems.
The batch reuses the per-problem arithmetic of the reference's BFS level
(multiply.cpp:212-310) for every problem."""
import os

import numpy as np
import pytest

import caps_amd
import oracle
from caps_amd import batch_plan, planner_train

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

MODEL = os.path.join(os.path.dirname(caps_amd.__file__), "planner_model.npz")


@pytest.fixture(scope="module", autouse=True)
def _setup():
    oracle.build()
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    caps_amd.load()
    model, _ = planner_train.load_model(MODEL)
    yield
    caps_amd.MLP.clear_planner()
    caps_amd.planner_config(1e-14, caps_amd.MASK_FP64 | caps_amd.MASK_CRT_INT8, 10.0)
    del model
    caps_amd.release()
    torch.cuda.empty_cache()


def _inputs(sizes, plans, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    As, Bs, Cs, descs, pats, cuts = [], [], [], [], [], []
    for n in sizes:
        p = plans[n]
        As.append(torch.rand(n * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1)
        Bs.append(torch.rand(n * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1)
        Cs.append(torch.full_like(As[-1], float("nan")))
        descs.append(caps_amd.MatDescriptor.one_device(n, p.nrec, 1))
        pats.append(p.pattern)
        cuts.append(p.cutoff)
    return As, Bs, Cs, descs, pats, cuts


def test_uneven_batch_fp64_bit_identical_to_single_calls():
    sizes = batch_plan.uneven_sizes()
    assert len(sizes) == 512 and min(sizes) >= 128 and max(sizes) <= 4096
    assert all(n % 64 == 0 for n in sizes)
    plans = batch_plan.choose_plans(sizes, caps_amd.MODE_FP64, max_err=1e-12)
    print(f"[gpu] fp64 batch: {len(set(sizes))} sizes, depths "
          f"{sorted(set(p.nrec for p in plans.values()))}", flush=True)
    As, Bs, Cs, descs, pats, cuts = _inputs(sizes, plans, 77)
    caps_amd.multiply_batch_device(As, Bs, Cs, descs, pats, caps_amd.MODE_FP64, 0, cutoffs=cuts)
    torch.cuda.synchronize()
    print("[gpu] fp64 batch done; single calls", flush=True)
    bad = []
    for i, (a, b, c, d, p, k) in enumerate(zip(As, Bs, Cs, descs, pats, cuts)):
        alone = torch.empty_like(c)
        caps_amd.multiply_device(a, b, alone, d, p, mode=caps_amd.MODE_FP64, cutoff=k,
                                 flags=caps_amd.FLAG_NO_GRAPH)
        torch.cuda.synchronize()
        if not torch.equal(alone, c):
            bad.append((i, d.lda))
        del alone
        if i % 64 == 63:
            caps_amd.release()
    assert not bad, f"{len(bad)} problems differ from their single call, e.g. {bad[:5]}"


def test_uneven_batch_residue_mode_error_bar():
    sizes = batch_plan.uneven_sizes()
    plans = batch_plan.choose_plans(sizes, caps_amd.MODE_CRT_INT8, max_err=1e-14, moduli=14)
    K = next(iter(plans.values())).moduli
    print(f"[gpu] residue batch: {K} moduli, depths "
          f"{sorted(set(p.nrec for p in plans.values()))}", flush=True)
    As, Bs, Cs, descs, pats, cuts = _inputs(sizes, plans, 78)
    caps_amd.multiply_batch_device(As, Bs, Cs, descs, pats, caps_amd.MODE_CRT_INT8, K,
                                   cutoffs=cuts)
    torch.cuda.synchronize()
    print("[gpu] residue batch done; reference products", flush=True)
    worst, errs = 0.0, []
    for i, (a, b, c, d) in enumerate(zip(As, Bs, Cs, descs)):
        ref = torch.empty_like(c)
        caps_amd.reference_product(a, b, ref, d)
        e = caps_amd.normwise_error(c, ref, d)
        errs.append(e)
        worst = max(worst, e)
        del ref
        if i % 64 == 63:
            caps_amd.release()
    print(f"[gpu] residue batch worst error {worst:.3e}", flush=True)
    assert worst <= 1e-14
    # Dot2 on 4 sampled rows of the 8 largest problems
    order = sorted(range(len(sizes)), key=lambda i: -sizes[i])[:8]
    rng = np.random.default_rng(5)
    for i in order:
        n, d = sizes[i], descs[i]
        A = oracle.from_caps(As[i].cpu().numpy(), n, d.nrec, 1)
        B = oracle.from_caps(Bs[i].cpu().numpy(), n, d.nrec, 1)
        C = oracle.from_caps(Cs[i].cpu().numpy(), n, d.nrec, 1)
        rows = np.sort(rng.choice(n, 4, replace=False))
        rr = np.repeat(rows, n).astype(np.int32)
        cc = np.tile(np.arange(n), len(rows)).astype(np.int32)
        ref = oracle.dot2_entries(A, B, rr, cc).reshape(len(rows), n)
        e = float(np.abs(C[rows] - ref).max() / np.abs(ref).max())
        assert e <= 1e-14, (i, n, e)
