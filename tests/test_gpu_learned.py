"""Rank-prediction MLP (device forward / backward / Adam, FP32 and BF16
inference) and the learned-rank rounding.  Gradients and the BF16 forward are
checked against the numpy restatement (oracle/mlp_ref.py); the trained
network's gain is measured at config-4 shapes (d = 32, n = 4, ranks up to
128, bond spectra 10^(-j s)) against the same randomized rounding driven by
non-learned rank policies: the stored ranks, and fixed fractions of them."""
import time

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
    l1 = m.train(X, y, steps=1500, lr=2e-3)
    assert l1 < 0.05 * l0
    pred = m.predict(X)
    assert np.abs(pred - y).mean() < 0.03


def test_mlp_training_keeps_lowest_loss():
    """The loss returned is that of the parameters left in the network (the
    lowest seen), recomputed here by a separate gradient call."""
    rng = np.random.default_rng(2)
    X = rng.uniform(-1, 1, (3000, 13)).astype(np.float32)
    y = (0.4 + 0.2 * np.tanh(3 * X[:, 2]) + 0.1 * X[:, 5] * X[:, 7]).astype(np.float32)
    m = T.RankMLP(13, 64, 3, seed=3)
    l_a = m.train(X, y, steps=400, lr=5e-3)
    l_b = m.grad(X, y)[0]
    assert abs(l_a - l_b) <= 1e-6 * l_b + 1e-12
    l_c = m.train(X, y, steps=400, lr=5e-3)
    assert l_c <= l_a * (1 + 1e-6)  # the second call starts from the kept parameters


def test_bf16_inference_matches_restatement():
    rng = np.random.default_rng(6)
    m = T.RankMLP(13, 64, 3, seed=8)
    X = rng.uniform(-1, 1, (4000, 13)).astype(np.float32)
    y = (0.5 + 0.3 * X[:, 0] * X[:, 3]).astype(np.float32)
    m.train(X, y, steps=300, lr=2e-3)
    p = m.get_params()
    dims = [13, 64, 64, 64, 1]
    f32 = m.predict(X)
    m.set_bf16_inference(True)
    b16 = m.predict(X)
    m.set_bf16_inference(False)
    assert np.array_equal(m.predict(X), f32)
    ref16 = R.forward_bf16(p, dims, X)
    ref32 = R.forward(p, dims, X)[-1][:, 0]
    # FP32 accumulation of BF16 products against exact accumulation: a few
    # float32 ulps of the 64-term sums, except where such a difference moves
    # a hidden activation across a BF16 rounding boundary (its next-layer
    # input then changes by one BF16 ulp, 2^-8 relative): rare, and bounded.
    d16 = np.abs(b16 - ref16)
    assert np.mean(d16 > 1e-6) <= 0.01
    assert d16.max() <= 1e-3
    assert np.abs(f32 - ref32).max() <= 1e-5 * max(1.0, np.abs(ref32).max())
    # BF16 keeps 8 significant bits: the outputs move by about 2^-8 of the
    # activations' scale, not more.
    assert 0 < np.abs(b16 - f32).max() <= 0.05


def _config4_batch(count, seed, rng):
    s = rng.uniform(0.05, 0.3, count)
    return T.build_decay_batch(count, 32, 4, 128, s, seed + np.arange(count))


def _best_time(ctx, fn, reps=3):
    best = np.inf
    for _ in range(reps):
        ctx.synchronize()
        t0 = time.perf_counter()
        out = fn()
        ctx.synchronize()
        best = min(best, time.perf_counter() - t0)
    return out, best


def test_learned_rounding_beats_non_learned_policies():
    """Config-4 shapes: the trained network's target ranks against the same
    randomized rounding (same seed, same plain fallback) with the stored
    ranks and with fixed fractions of them as targets.  Measured on
    rounds/s (best of three passes), fallback rate and mean target rank."""
    eps = 1e-6
    ctx = T.Context.default(0)
    rng = np.random.default_rng(3)
    train = _config4_batch(200, 7_000_000, rng)
    plain_tr = T.round_batched(train, eps=eps)
    X = T.bond_features(train, eps)
    y = np.array([p.ranks[k] / t.ranks[k] for t, p in zip(train, plain_tr) for k in range(1, t.d)], np.float32)
    m = T.RankMLP(X.shape[1], 64, 3, seed=4)
    m.train(X, y, steps=1500, lr=2e-3)
    del train, plain_tr
    test = _config4_batch(512, 900, rng)
    plain = T.round_batched(test, eps=eps)
    n = len(test)
    # Unbiased predictor: median rank error within one rank.
    Xt = T.bond_features(test, eps)
    stored = np.array([t.ranks[k] for t in test for k in range(1, t.d)], np.float64)
    lab = np.array([p.ranks[k] for p in plain for k in range(1, p.d)], np.float64)
    assert abs(np.median(m.predict(Xt) * stored - lab)) <= 1.0
    mean_plain = np.mean([np.mean(p.ranks[1:-1]) for p in plain])

    policies = {"stored": [t.ranks for t in test]}
    for f in (0.6, 0.8):
        policies[f"fraction_{f}"] = [[1] + [min(t.ranks[k], int(np.ceil(f * t.ranks[k])) + 6)
                                            for k in range(1, t.d)] + [1] for t in test]
    res = {}
    for name, rk in policies.items():
        T.round_randomized_batched(test, rk, eps, seed=7)
        (_, _, fb), t = _best_time(ctx, lambda: T.round_randomized_batched(test, rk, eps, seed=7))
        res[name] = (n / t, float(fb.mean()))
    m.set_bf16_inference(True)
    T.round_learned_batched(m, test, eps, seed=7)
    (outs, err, fb, ranks), t = _best_time(ctx, lambda: T.round_learned_batched(m, test, eps, seed=7))
    learned_rps, learned_fb = n / t, float(fb.mean())
    mean_target = np.mean([np.mean(r[1:-1]) for r in ranks])
    print(f"rounds/s, fallback: learned {learned_rps:.0f} {learned_fb:.3f} (mean target {mean_target:.1f}, "
          f"plain {mean_plain:.1f}); " + ", ".join(f"{k} {v[0]:.0f} {v[1]:.3f}" for k, v in res.items()), flush=True)
    assert learned_fb <= 0.02
    for e, f in zip(err, fb):
        if not f:
            assert e <= eps
    assert mean_target <= 1.4 * mean_plain
    # Faster than every non-learned policy, at no higher fallback rate.
    for name, (rps, f) in res.items():
        assert learned_rps >= 1.3 * rps, name
        assert learned_fb <= f + 1e-12, name
    # The outputs are within eps of the inputs (sampled, device-side norms).
    for i in range(0, n, 97):
        dif = T.sub(outs[i], test[i])
        assert T.norm(dif) <= eps * T.norm(test[i]) * (1 + 1e-6)
