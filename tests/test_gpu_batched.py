"""Batched plain rounding (tta_round_batched): every TT of a batch of uneven
d, mode sizes and ranks matches the oracle's round of that TT (identical
ranks, relative error <= 1e-12)."""
import numpy as np
import pytest

import oracle as O
import tensor_train_amd as T
from tests.ttutil import rel_err

pytestmark = pytest.mark.gpu


def cfg3_like(t, rng):
    d = int(rng.integers(3, 12))
    n = int(rng.choice([2, 4, 8]))
    ru, rv = int(rng.integers(2, 20)), int(rng.integers(2, 20))
    u = O.OTT.random([n] * d, ru, 1000 + t)
    v = O.OTT.random([n] * d, rv, 500000 + t)
    return O.add(u, O.add(O.scale(u, 0.5), v))


@pytest.mark.parametrize("eps,cap", [(1e-6, 0), (1e-10, 0), (1e-3, 5)])
def test_round_batched_matches_oracle(eps, cap):
    rng = np.random.default_rng(7)
    inputs = [cfg3_like(t, rng) for t in range(24)]
    inputs.append(O.OTT.random([3, 4, 5], 6, 9))
    inputs.append(O.OTT.from_cores([np.arange(1.0, 6.0).reshape(1, 5, 1)]))        # d = 1
    inputs.append(O.OTT.from_cores([np.zeros((1, 3, 1)), np.zeros((1, 2, 1))]))  # zero TT
    dev = [T.TensorTrain.from_cores(o.cores()) for o in inputs]
    outs = T.round_batched(dev, eps=eps, max_rank=cap)
    for o, y in zip(inputs, outs):
        ref = O.round(o, eps, cap)
        assert y.ranks == ref.ranks
        if O.norm(ref) > 0:
            assert rel_err([c for c in y.cores], ref) <= 1e-12
        else:
            assert np.all(y.to_dense() == 0)


def test_round_batched_equals_single_path():
    rng = np.random.default_rng(11)
    inputs = [cfg3_like(t, rng) for t in range(6)]
    dev = [T.TensorTrain.from_cores(o.cores()) for o in inputs]
    outs = T.round_batched(dev, eps=1e-8)
    for d_in, y in zip(dev, outs):
        single = T.round(d_in, eps=1e-8)
        assert y.ranks == single.ranks
        a, b = y.to_dense(), single.to_dense()
        assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b)


def test_build_sum_batch_structure():
    d, n, ru, rv = [4, 1, 5], [2, 3, 4], [3, 2, 4], [2, 3, 2]
    cs = T.build_sum_batch(d, n, ru, rv, [11, 12, 13], [21, 22, 23])
    for t, c in enumerate(cs):
        cores = c.cores
        if d[t] == 1:
            assert cores[0].shape == (1, n[t], 1)
            continue
        Ru = O.OTT.random([n[t]] * d[t], ru[t], 0).ranks
        Rv = O.OTT.random([n[t]] * d[t], rv[t], 0).ranks
        assert c.ranks == [1] + [2 * a + b for a, b in zip(Ru[1:-1], Rv[1:-1])] + [1]
        u, v = [], []
        for k, ck in enumerate(cores):
            if k == 0:
                u.append(ck[:, :, :Ru[1]])
                assert np.array_equal(ck[:, :, Ru[1]:2 * Ru[1]], 0.5 * u[-1])
                v.append(ck[:, :, 2 * Ru[1]:])
            elif k == d[t] - 1:
                u.append(ck[:Ru[k]])
                assert np.array_equal(ck[Ru[k]:2 * Ru[k]], u[-1])
                v.append(ck[2 * Ru[k]:])
            else:
                u.append(ck[:Ru[k], :, :Ru[k + 1]])
                assert np.array_equal(ck[Ru[k]:2 * Ru[k], :, Ru[k + 1]:2 * Ru[k + 1]], u[-1])
                v.append(ck[2 * Ru[k]:, :, 2 * Ru[k + 1]:])
                assert np.all(ck[:Ru[k], :, Ru[k + 1]:] == 0) and np.all(ck[2 * Ru[k]:, :, :2 * Ru[k + 1]] == 0)
        U, V = O.OTT.from_cores(u), O.OTT.from_cores(v)
        dense = c.to_dense()
        np.testing.assert_allclose(dense, 1.5 * U.to_dense() + V.to_dense(), rtol=0, atol=1e-12)
        allv = np.concatenate([x.ravel() for x in u + v])
        assert abs(allv.mean()) < 0.3 and 0.5 < allv.var() < 1.6
        r = T.round(c, eps=1e-12)
        assert all(x <= a + b for x, a, b in zip(r.ranks, Ru, Rv))


def test_build_decay_batch_spectrum():
    s = [0.05, 0.2, 0.3]
    ts = T.build_decay_batch(3, 6, 4, 16, s, [5, 6, 7])
    for si, t in zip(s, ts):
        cores = t.cores
        assert t.ranks == [1, 4, 16, 16, 16, 4, 1]
        for k in range(5):
            M = cores[k].reshape(-1, cores[k].shape[2])
            want = 10.0 ** (-2.0 * si * np.arange(M.shape[1]))
            np.testing.assert_allclose(M.T @ M, np.diag(want), rtol=0, atol=1e-13)
        assert abs(np.linalg.norm(cores[-1]) - 1.0) < 1e-14
