"""Rank-prediction MLP (device forward / backward / Adam) and the
learned-rank rounding.  Gradients are checked against the numpy
restatement (oracle/mlp_ref.py); the trained network's gain over the
randomized rounding without it is measured on decaying-spectrum TTs."""
import numpy as np
import pytest

import oracle.mlp_ref as R
import tensor_train_amd as T

pytestmark = pytest.mark.gpu


# 300 rows fit one slice of the batch reduction behind dW and db; 5000 rows
# split into several slices whose partial sums are added afterwards.
@pytest.mark.parametrize("n", [300, 5000])
def test_mlp_gradients_match_cpu(n):
    rng = np.random.default_rng(0)
    m = T.RankMLP(13, 64, 3, seed=5)
    dims = [13, 64, 64, 64, 1]
    p = m.get_params()
    assert p.size == sum(a * b + b for a, b in zip(dims[:-1], dims[1:]))
    X = rng.standard_normal((n, 13)).astype(np.float32)
    y = rng.uniform(0, 1, n).astype(np.float32)
    loss, g = m.grad(X, y)
    loss_ref, g_ref = R.loss_and_grad(p, dims, X, y)
    assert abs(loss - loss_ref) <= 1e-4 * max(1.0, loss_ref)
    assert np.abs(g - g_ref).max() <= 2e-4 * max(1.0, np.abs(g_ref).max())
    # Finite-difference check of a few parameters through the CPU model.
    for idx in rng.integers(0, p.size, 5):
        h = 1e-6  # float64 model; small enough not to cross ReLU kinks
        pp, pm = p.astype(np.float64).copy(), p.astype(np.float64).copy()
        pp[idx] += h
        pm[idx] -= h
        fd = (R.loss_and_grad(pp, dims, X, y)[0] - R.loss_and_grad(pm, dims, X, y)[0]) / (2 * h)
        assert abs(fd - g_ref[idx]) <= 1e-7 + 1e-4 * abs(fd)


def test_mlp_training_fits_a_smooth_target():
    rng = np.random.default_rng(1)
    X = rng.uniform(-1, 1, (2000, 13)).astype(np.float32)
    y = (0.5 + 0.3 * np.sin(2 * X[:, 0]) * X[:, 1]).astype(np.float32)
    m = T.RankMLP(13, 64, 3, seed=2)
    l0 = m.grad(X, y)[0]
    l1 = m.train(X, y, steps=600, lr=2e-3)
    assert l1 < 0.05 * l0
    pred = m.predict(X)
    assert np.abs(pred - y).mean() < 0.03


def _labels(tts, plain):
    lab = []
    for t, p in zip(tts, plain):
        for k in range(1, t.d):
            lab.append(p.ranks[k] / t.ranks[k])
    return np.array(lab, np.float32)


def test_learned_rounding_beats_untrained_policy():
    eps = 1e-6
    rng = np.random.default_rng(3)
    train = T.build_decay_batch(64, 10, 4, 32, rng.uniform(0.05, 0.3, 64), 100 + np.arange(64))
    plain_tr = T.round_batched(train, eps=eps)
    X = T.bond_features(train, eps)
    y = _labels(train, plain_tr)
    m = T.RankMLP(X.shape[1], 64, 3, seed=4)
    m.train(X, y, steps=1500, lr=2e-3)
    test = T.build_decay_batch(24, 10, 4, 32, rng.uniform(0.05, 0.3, 24), 900 + np.arange(24))
    plain = T.round_batched(test, eps=eps)
    outs, err, fb, ranks = T.round_learned_batched(m, test, eps, oversample=4, seed=7)
    assert fb.sum() <= 2
    for y_, e, f in zip(outs, err, fb):
        if not f:
            assert e <= eps
    # Gain over the method without the network: the untrained network's
    # targets either fall back or carry larger ranks.
    m0 = T.RankMLP(X.shape[1], 64, 3, seed=99)
    outs0, err0, fb0, ranks0 = T.round_learned_batched(m0, test, eps, oversample=4, seed=7)
    learned_work = sum(sum(r) for r in ranks)
    plain_ranks = sum(sum(p.ranks) for p in plain)
    assert learned_work <= 1.6 * plain_ranks
    assert fb0.sum() > fb.sum() or sum(sum(r) for r in ranks0) > learned_work
