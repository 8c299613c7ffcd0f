"""Train the planner's cost / error model on measured runs and evaluate it on
held-out shapes against the exhaustive autotuner.

    python -m caps_amd.planner_train --train RUNS_TRAIN.npz ... --test RUNS_TEST.npz ... \\
        --epochs 300 --model caps_amd/planner_model.npz

Each --test file is scored on its own, also with the choice restricted to
plain FP64 candidates under --fp64-target (the FP64 mode of uneven batches).

Runs come from caps_amd.planner_data.  The model is the library's 3-layer
perceptron (8 -> 64 -> 64 -> 2, ReLU, FP32), trained by the library's own HIP
forward / backward / Adam kernels (caps_mlp_train) on normalised features
(log2 n, exponent spread, #b, #d, cutoff, moduli, mode, leaf alignment; see
features_from_columns) and targets (log2 ms, log10 normwise error).

Evaluation per held-out shape, with the error target 1e-14:
  exhaustive  the fastest measured candidate whose measured error meets the
              target (what caps_autotune returns);
  model       the candidate with the smallest predicted time among those
              whose predicted log10 error is at least a quarter-decade below
              the target (the library's rule in CAPS_MODE_AUTO); shapes
              where no candidate qualifies fall back to the exhaustive search;
  default     the choice without a model or tuning table: residue-domain
              mode, 14 moduli, all levels batched, no cutoff;
  reference   the reference's one-process rule: plain FP64, all levels 'd'.
Regret = t(choice) / t(exhaustive) - 1; a choice whose measured error
misses the target is a violation.
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

import caps_amd

TARGET = 1e-14
MARGIN = 0.25  # decades, as in the library's planner


_MODE_CODE = {caps_amd.MODE_FP64: 0, caps_amd.MODE_CRT_INT8: 1, caps_amd.MODE_OZAKI_INT8: 2,
              caps_amd.MODE_CRT_BF16: 3}


def features_from_columns(runs) -> np.ndarray:
    """The library's 8 candidate features (caps_candidate_features), restated
    from the recorded columns: log2 n, exponent spread, #b and #d levels above
    the cutoff, cutoff, moduli, mode ordinal, log2(gcd(n / 2^nrec, 256)) / 8."""
    N = len(runs["n"])
    f = np.zeros((N, 8), np.float32)
    for i in range(N):
        n, nrec, cut = int(runs["n"][i]), int(runs["nrec"][i]), int(runs["cutoff"][i])
        pat = str(runs["pattern"][i])[:nrec - cut]
        nb = sum(ch in "bB" for ch in pat)
        leaf = n >> nrec
        a = 0
        while a < 8 and leaf > 0 and (leaf >> a) % 2 == 0:
            a += 1
        f[i] = (np.float32(np.log2(float(n))), np.float32(float(runs["spread"][i])), nb,
                len(pat) - nb, cut, int(runs["moduli"][i]), _MODE_CODE[int(runs["mode"][i])],
                np.float32(a) / np.float32(8))
    return f


def load_runs(paths) -> dict:
    parts = [dict(np.load(p)) for p in paths]
    out = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    # shape ids are unique within a split; parts of one split never overlap.
    # Features are recomputed with the library's current definition.
    out["features"] = features_from_columns(out)
    return out


def targets(runs):
    T = np.stack([np.log2(np.maximum(runs["ms"], 1e-6)),
                  np.log10(np.maximum(runs["err"], 1e-17))], axis=1).astype(np.float32)
    return T


def normaliser(X, T):
    xm, xs = X.mean(0), X.std(0)
    xs = np.where(xs > 1e-6, xs, 1.0)
    tm, ts = T.mean(0), T.std(0)
    ts = np.where(ts > 1e-6, ts, 1.0)
    return np.concatenate([xm, xs, tm, ts]).astype(np.float32)


def fit(train, epochs=300, seed=0, lr=2e-3, batch=256):
    X = np.asarray(train["features"], dtype=np.float32)
    T = targets(train)
    norm = normaliser(X, T)
    Xn = (X - norm[:8]) / norm[8:16]
    Tn = (T - norm[16:18]) / norm[18:20]
    M = np.ones_like(Tn)
    model = caps_amd.MLP(seed=seed)
    # three stages of decreasing step size
    losses = []
    for frac, scale in ((0.5, 1.0), (0.3, 0.3), (0.2, 0.1)):
        losses.append(model.train(Xn, Tn, M, epochs=max(1, int(epochs * frac)), batch=batch,
                                  lr=lr * scale, seed=seed + len(losses)))
    return model, norm, losses


def predict(model, norm, X):
    Y = model.forward((np.asarray(X, np.float32) - norm[:8]) / norm[8:16])
    return Y * norm[18:20] + norm[16:18]  # (log2 ms, log10 err)


def _default_index(rows_idx, runs):
    for j in rows_idx:
        p = str(runs["pattern"][j])
        if (runs["mode"][j] == caps_amd.MODE_CRT_INT8 and runs["moduli"][j] == 14 and
                runs["cutoff"][j] == 0 and set(p) <= {"b"}):
            return j
    return None


def _reference_index(rows_idx, runs):
    for j in rows_idx:
        p = str(runs["pattern"][j])
        if runs["mode"][j] == caps_amd.MODE_FP64 and runs["cutoff"][j] == 0 and set(p) <= {"d"}:
            return j
    return None


def evaluate(model, norm, test, target: float = TARGET, modes=None) -> dict:
    """Scores per held-out shape (see the module docstring); `modes`
    restricts every choice to candidates of those modes (e.g. FP64 only,
    with its own error target)."""
    P = predict(model, norm, test["features"])
    T = targets(test)
    res = {"rmse_log2_ms": float(np.sqrt(np.mean((P[:, 0] - T[:, 0]) ** 2))),
           "rmse_log10_err": float(np.sqrt(np.mean((P[:, 1] - T[:, 1]) ** 2)))}
    shapes = np.unique(test["shape"])
    reg = {"model": [], "default": [], "reference": []}
    viol = {"model": 0, "default": 0, "reference": 0}
    fallbacks = 0
    no_feasible = 0
    for s in shapes:
        idx = np.nonzero(test["shape"] == s)[0]
        if modes is not None:
            idx = idx[np.isin(test["mode"][idx], list(modes))]
            if len(idx) == 0:
                continue
        ms, err = test["ms"][idx], test["err"][idx]
        feas = err <= target
        if not feas.any():
            no_feasible += 1
            continue
        best = ms[feas].min()
        ok = P[idx, 1] <= np.log10(target) - MARGIN
        if ok.any():
            k = idx[ok][np.argmin(P[idx[ok], 0])]
        else:
            fallbacks += 1
            k = idx[feas][np.argmin(ms[feas])]
        for name, j in (("model", k), ("default", _default_index(idx, test)),
                        ("reference", _reference_index(idx, test))):
            if j is None:
                continue
            reg[name].append(test["ms"][j] / best - 1.0)
            viol[name] += int(test["err"][j] > target)
    for name, r in reg.items():
        if not r:  # no row of this rule among the scored modes
            res[name] = {"shapes": 0, "mean_regret": float("nan"), "median_regret": float("nan"),
                         "p90_regret": float("nan"), "violations": 0}
            continue
        r = np.array(r)
        res[name] = {"shapes": int(len(r)), "mean_regret": float(np.mean(r)),
                     "median_regret": float(np.median(r)),
                     "p90_regret": float(np.percentile(r, 90)), "violations": viol[name]}
    res["shapes"] = int(len(shapes))
    res["model_fallbacks_to_exhaustive"] = fallbacks
    res["shapes_without_feasible_candidate"] = no_feasible
    return res


def save_model(path, model, norm):
    np.savez(path, params=model.params(), norm=norm)


def load_model(path, install=True):
    """Load a saved model; with install=True it becomes the planner of
    CAPS_MODE_AUTO calls (after any tuning-table entry)."""
    with np.load(path) as z:
        params, norm = z["params"], z["norm"]
    model = caps_amd.MLP()
    model.set_params(params)
    if install:
        model.use_as_planner(norm)
    return model, norm


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", nargs="+", required=True)
    ap.add_argument("--test", nargs="+", required=True)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--model", default=None, help="save the trained model here (.npz)")
    ap.add_argument("--fp64-target", type=float, default=1e-12,
                    help="also score FP64-only choices against this error target")
    a = ap.parse_args(argv)
    train = load_runs(a.train)
    model, norm, losses = fit(train, a.epochs, a.seed)
    # each held-out file on its own (shape ids are unique within a split only)
    rep = {"train_rows": int(len(train["ms"])), "final_losses": losses}
    for path in a.test:
        test = load_runs([path])
        name = os.path.basename(path)
        rep[name] = evaluate(model, norm, test)
        rep[name]["test_rows"] = int(len(test["ms"]))
        rep[name]["fp64_only"] = evaluate(model, norm, test, a.fp64_target,
                                          modes=(caps_amd.MODE_FP64,))
    if a.model:
        save_model(a.model, model, norm)
    print(json.dumps(rep, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
