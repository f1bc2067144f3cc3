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
