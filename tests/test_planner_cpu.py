"""The planner evaluation (caps_amd.planner_train.evaluate) on synthetic
runs, with stand-in predictors: a perfect one has zero regret and no
violations; one that ignores the error target violates it; the no-model
default and the reference rule are scored from their own rows."""
import numpy as np

import caps_amd
from caps_amd import planner_train as pt


def _runs():
    # 3 shapes x 4 candidates: (mode, moduli, cutoff, pattern, ms, err)
    cands = [
        (caps_amd.MODE_FP64, 0, 0, "dd", 10.0, 2e-14),          # reference rule, misses the bar
        (caps_amd.MODE_FP64, 0, 0, "bb", 8.0, 3e-14),           # fast FP64, misses the bar
        (caps_amd.MODE_CRT_INT8, 14, 0, "bb", 6.0, 5e-15),     # default
        (caps_amd.MODE_CRT_INT8, 14, 1, "bb", 4.0, 4e-15),     # best feasible
    ]
    rows = {k: [] for k in ("shape", "n", "nrec", "spread", "features", "ms", "err", "mode",
                            "moduli", "cutoff", "pattern")}
    for s in range(3):
        for i, (mode, mod, cut, pat, ms, err) in enumerate(cands):
            rows["shape"].append(s)
            rows["n"].append(1024)
            rows["nrec"].append(2)
            rows["spread"].append(0)
            f = np.zeros(8, np.float32)
            f[0] = i  # the stand-in predictors key on this
            rows["features"].append(f)
            rows["ms"].append(ms * (1 + s))
            rows["err"].append(err)
            rows["mode"].append(mode)
            rows["moduli"].append(mod)
            rows["cutoff"].append(cut)
            rows["pattern"].append(pat)
    return {k: np.array(v) for k, v in rows.items()}


class _Stand:
    """forward() returns normalised (log2 ms, log10 err) from a table keyed
    by feature 0; with the identity normaliser below these are the values."""

    def __init__(self, table):
        self.table = table

    def forward(self, X):
        return np.array([self.table[int(round(x[0]))] for x in X], dtype=np.float32)


NORM = np.concatenate([np.zeros(8), np.ones(8), np.zeros(2), np.ones(2)]).astype(np.float32)


def test_perfect_predictor_has_zero_regret():
    runs = _runs()
    truth = {i: (np.log2(ms), np.log10(err)) for i, (_, _, _, _, ms, err) in enumerate([
        (0, 0, 0, 0, 10.0, 2e-14), (0, 0, 0, 0, 8.0, 3e-14), (0, 0, 0, 0, 6.0, 5e-15),
        (0, 0, 0, 0, 4.0, 4e-15)])}
    rep = pt.evaluate(_Stand(truth), NORM, runs)
    assert rep["model"]["mean_regret"] == 0.0 and rep["model"]["violations"] == 0
    assert abs(rep["default"]["mean_regret"] - 0.5) < 1e-12 and rep["default"]["violations"] == 0
    assert rep["reference"]["violations"] == 3
    assert rep["shapes"] == 3 and rep["model_fallbacks_to_exhaustive"] == 0


def test_predictor_blind_to_error_violates():
    runs = _runs()
    blind = {0: (3.0, -16.0), 1: (1.0, -16.0), 2: (2.6, -16.0), 3: (2.0, -16.0)}
    rep = pt.evaluate(_Stand(blind), NORM, runs)
    assert rep["model"]["violations"] == 3  # it picks the fast FP64 candidate


def test_no_candidate_predicted_feasible_falls_back_to_exhaustive():
    runs = _runs()
    pessimist = {i: (1.0, -10.0) for i in range(4)}
    rep = pt.evaluate(_Stand(pessimist), NORM, runs)
    assert rep["model_fallbacks_to_exhaustive"] == 3
    assert rep["model"]["mean_regret"] == 0.0


def test_feature_restatement_matches_library():
    """planner_train.features_from_columns equals the library's
    caps_candidate_features on every candidate of a range of layouts."""
    rows = {k: [] for k in ("n", "nrec", "spread", "cutoff", "pattern", "moduli", "mode")}
    want = []
    mask = caps_amd.MASK_FP64 | caps_amd.MASK_CRT_INT8 | caps_amd.MASK_OZAKI_INT8
    for n, r in [(512, 0), (1024, 1), (2304, 2), (4352, 3), (6912, 4), (8192, 2), (12800, 4),
                 (96, 1), (56, 3)]:
        d = caps_amd.MatDescriptor.one_device(n, r, 1)
        for c in caps_amd.candidates(d, mask):
            for spread in (0.0, 7.0):
                want.append(caps_amd.candidate_features(d, c, spread))
                for k, v in (("n", n), ("nrec", r), ("spread", spread), ("cutoff", c.cutoff),
                             ("pattern", c.pattern.decode()), ("moduli", c.moduli),
                             ("mode", c.mode)):
                    rows[k].append(v)
    runs = {k: np.array(v) for k, v in rows.items()}
    got = pt.features_from_columns(runs)
    assert np.array_equal(got, np.array(want, dtype=np.float32))
    # leaves 512 (256-aligned), 576, 544, 432 (64-, 32-, 16-aligned), 7 (odd)
    assert set(np.unique(got[:, 7]).tolist()) == {1.0, 0.75, 0.625, 0.5, 0.0}


def test_planner_shapes():
    """The seeded shapes of caps_amd.planner_data: n a multiple of 256 in
    [512, 16384], the deepest layout with leaves >= 512 that divides n, spread
    in 0..20; the training and held-out seeds give different shape lists."""
    from caps_amd import planner_data as pd
    for split in ("train", "test"):
        for i in range(300):
            n, r, spread = pd.shape(split, i)
            assert n % 256 == 0 and 512 <= n <= 16384
            assert n % (1 << r) == 0 and (n >> r) >= 512
            assert not (n % (2 << r) == 0 and (n >> (r + 1)) >= 512)
            assert 0 <= spread <= 20
    train = [pd.shape("train", i) for i in range(50)]
    test = [pd.shape("test", i) for i in range(50)]
    assert train != test
    assert pd.shape("train", 7) == pd.shape("train", 7)


def test_batch_planner_shapes():
    """The batch-domain shapes: n a multiple of 64 in [128, 4096], a depth
    drawn from those with leaves >= 64 dividing n, spread in 0..20; the two
    batch splits differ from each other."""
    from caps_amd import planner_data as pd
    seen = set()
    for split in ("batch_train", "batch_test"):
        for i in range(300):
            n, r, spread = pd.shape(split, i)
            assert n % 64 == 0 and 128 <= n <= 4096
            assert n % (1 << r) == 0 and (n >> r) >= 64
            assert 0 <= spread <= 20
            seen.add((n, r))
    # depth 6 needs n = 4096, the top of the range; 0..5 all turn up
    assert {0, 1, 2, 3, 4, 5} <= {r for _, r in seen}
    assert [pd.shape("batch_train", i) for i in range(50)] != \
        [pd.shape("batch_test", i) for i in range(50)]


def test_mode_restricted_evaluation():
    """evaluate(modes=FP64, target) chooses among FP64 rows only: with a looser
    target the fast FP64 candidate is feasible and the perfect predictor
    takes it; the no-model default (a residue-mode row) is not scored."""
    runs = _runs()
    truth = {0: (np.log2(10.0), np.log10(2e-14)), 1: (np.log2(8.0), np.log10(3e-14)),
             2: (np.log2(6.0), np.log10(5e-15)), 3: (np.log2(4.0), np.log10(4e-15))}
    rep = pt.evaluate(_Stand(truth), NORM, runs, target=1e-12, modes=(caps_amd.MODE_FP64,))
    assert rep["model"]["mean_regret"] == 0.0 and rep["model"]["violations"] == 0
    assert rep["default"]["shapes"] == 0
    assert abs(rep["reference"]["mean_regret"] - 0.25) < 1e-12
    # under 1e-14 no FP64 row is feasible: every shape is skipped
    rep = pt.evaluate(_Stand(truth), NORM, runs, modes=(caps_amd.MODE_FP64,))
    assert rep["shapes_without_feasible_candidate"] == 3

