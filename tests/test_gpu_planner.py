"""What the learned planner gains over the method without it.

The measurements in tests/golden/planner_runs_{train,test}.npz (sizes 512 to
16384, multiples of 256) and planner_runs_batch_{train,test}.npz (the sizes of
uneven batches: 128 to 4096, multiples of 64, every depth with leaves of at
least 64) were made on an MI355X by caps_amd.planner_data: every candidate
configuration of each seeded shape, timed on the device, with its normwise
error against the 16-modulus residue product.  Here the library's perceptron
is trained on the training shapes of both domains by its own HIP kernels and
its choices on the held-out shapes are scored with the recorded times (no new
timing):
  * the model's choice is faster than the no-model default (residue mode,
    14 moduli, all levels batched) on average,
  * its mean regret against the exhaustive search stays small, also when
    restricted to plain FP64 under 1e-12 (the FP64 mode of uneven batches),
  * it never picks a configuration that misses the error target.
"""
import os

import numpy as np
import pytest

import caps_amd
from caps_amd import planner_train as pt

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
TRAIN = os.path.join(GOLD, "planner_runs_train.npz")
TEST = os.path.join(GOLD, "planner_runs_test.npz")
BATCH_TRAIN = os.path.join(GOLD, "planner_runs_batch_train.npz")
BATCH_TEST = os.path.join(GOLD, "planner_runs_batch_test.npz")
SHIPPED = os.path.join(os.path.dirname(caps_amd.__file__), "planner_model.npz")


@pytest.fixture(scope="module")
def runs():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    caps_amd.load()
    # training rows of both domains (shape ids matter only for scoring); the
    # held-out files are scored one by one (ids are unique within a split)
    return (pt.load_runs([TRAIN, BATCH_TRAIN]), pt.load_runs([TEST]),
            pt.load_runs([BATCH_TEST]))


def test_model_beats_no_model_default(runs):
    train, test, btest = runs
    print(f"[gpu] planner: {len(train['ms'])} training rows, "
          f"{len(np.unique(test['shape']))} + {len(np.unique(btest['shape']))} held-out shapes",
          flush=True)
    model, norm, losses = pt.fit(train, epochs=150, seed=0)
    for held, bar in ((test, 0.25), (btest, 0.35)):
        rep = pt.evaluate(model, norm, held)
        print(rep, flush=True)
        assert rep["model"]["violations"] == 0
        assert rep["model"]["mean_regret"] < rep["default"]["mean_regret"]
        assert rep["model"]["mean_regret"] < bar
        # the leaf-alignment feature lets the fit see which GEMM path a size
        # takes (recorded times differ up to 5x between neighbouring sizes)
        assert rep["rmse_log2_ms"] < 0.5
        assert rep["rmse_log10_err"] < 0.5
    rep = pt.evaluate(model, norm, btest, 1e-12, modes=(caps_amd.MODE_FP64,))
    print(rep, flush=True)
    assert rep["model"]["violations"] == 0 and rep["model"]["mean_regret"] < 0.15


def test_shipped_model_regret(runs):
    """The model shipped as caps_amd/planner_model.npz (trained on the same
    measurements by caps_amd.planner_train) scores the held-out shapes without
    retraining: no error-target violation, mean regret below the default's."""
    _, test, btest = runs
    model, norm = pt.load_model(SHIPPED, install=False)
    for held, bar in ((test, 0.1), (btest, 0.25)):
        rep = pt.evaluate(model, norm, held)
        print(rep, flush=True)
        assert rep["model"]["violations"] == 0
        assert rep["model"]["mean_regret"] < rep["default"]["mean_regret"]
        assert rep["model"]["mean_regret"] < bar
    rep = pt.evaluate(model, norm, btest, 1e-12, modes=(caps_amd.MODE_FP64,))
    print(rep, flush=True)
    assert rep["model"]["violations"] == 0 and rep["model"]["mean_regret"] < 0.05


def test_installed_model_drives_auto_mode(runs):
    """With the model installed, CAPS_MODE_AUTO picks a configuration whose
    result meets the error target on a fresh problem."""
    train, _, _ = runs
    model, norm, _ = pt.fit(train, epochs=60, seed=1)
    model.use_as_planner(norm)
    try:
        n, r = 1024, 2
        g = torch.Generator(device="cuda")
        g.manual_seed(5)
        A = (torch.rand(n * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1)
        B = (torch.rand(n * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1)
        C = torch.empty_like(A)
        Ref = torch.empty_like(A)
        d = caps_amd.MatDescriptor.one_device(n, r, 1)
        caps_amd.planner_config(1e-14, caps_amd.MASK_FP64 | caps_amd.MASK_CRT_INT8, 0.0)
        caps_amd.multiply_device(A, B, C, d, None, mode=caps_amd.MODE_AUTO)
        caps_amd.reference_product(A, B, Ref, d)
        torch.cuda.synchronize()
        assert caps_amd.normwise_error(C, Ref, d) <= 1e-14
    finally:
        caps_amd.MLP.clear_planner()
