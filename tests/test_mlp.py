"""The planner's cost / error model: the float64 oracle (oracle/mlp.py) is
pinned by finite differences on the CPU; the GPU forward, backward and Adam
kernels are checked against it (FP32 tolerances written in each test)."""
import numpy as np
import pytest

from oracle import mlp as om


def _data(B=300, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((B, om.N_IN))
    T = np.stack([np.sin(X[:, 0]) + 0.3 * X[:, 1], X[:, 2] * X[:, 3] - 0.5], axis=1)
    M = np.ones_like(T)
    M[::7, 1] = 0.0    # some samples without an error target
    return X, T, M


def _params(seed=1):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(om.NPARAM) * 0.2


def test_oracle_gradient_matches_finite_differences():
    X, T, M = _data(64)
    p = _params()
    g = om.gradient(p, X, T, M)
    rng = np.random.default_rng(2)
    idx = np.concatenate([rng.choice(om.NPARAM, 40, replace=False),
                          [om.OFF_B1, om.OFF_B2 + 3, om.OFF_W3 + 5, om.OFF_B3, om.OFF_B3 + 1]])
    fd = om.finite_difference(p, X, T, M, idx)
    assert np.allclose(g[idx], fd, rtol=1e-5, atol=1e-8)


def test_oracle_mask_drops_targets():
    X, T, M = _data(50)
    p = _params()
    T2 = T.copy()
    T2[M == 0] += 100.0          # masked targets do not matter
    assert om.loss(p, X, T, M) == om.loss(p, X, T2, M)
    assert np.array_equal(om.gradient(p, X, T, M), om.gradient(p, X, T2, M))


@pytest.mark.gpu
def test_gpu_forward_and_gradient_match_oracle():
    import caps_amd
    m = caps_amd.MLP(seed=3)
    X, T, M = _data(300)
    p = m.params().astype(np.float64)
    print("[gpu] mlp forward/gradient B=300", flush=True)
    Y = m.forward(X)
    Yo, _ = om.forward(p, X.astype(np.float32))
    assert np.allclose(Y, Yo, rtol=1e-4, atol=1e-4)
    g, loss = m.gradient(X, T, M)
    go = om.gradient(p, X.astype(np.float32), T.astype(np.float32), M)
    assert abs(loss - om.loss(p, X.astype(np.float32), T.astype(np.float32), M)) < 1e-4 * max(1, loss)
    scale = np.abs(go).max()
    assert np.abs(g - go).max() <= 1e-4 * scale
    m.close()


@pytest.mark.gpu
def test_gpu_training_deterministic_and_learns():
    import caps_amd
    X, T, M = _data(2000, seed=5)
    losses, params = [], []
    for _ in range(2):
        m = caps_amd.MLP(seed=11)
        print("[gpu] mlp train", flush=True)
        start = m.gradient(X, T, M)[1]
        losses.append(m.train(X, T, M, epochs=60, batch=128, lr=3e-3, seed=7))
        params.append(m.params())
        m.close()
    assert losses[0] < 0.25 * start
    assert np.array_equal(params[0], params[1])      # fixed-order reductions: bitwise repeatable
    assert losses[0] == losses[1]
