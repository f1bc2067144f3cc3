"""Mixed batched rounding (tta_round_batched, TTA_PREC_MIXED): SVD-sweep
steps certified by the rank-revealing CholQR2 (FP64 Gram, pivoted Cholesky,
CholQR2, FP64 residual correction), the rest run as in the plain mode.

Bar: relative error <= 1e-6 against the FP64 oracle's round of the same
input, the error bound ||a - round(a)|| <= eps ||a||, and ranks never below
the plain ranks.  On sums of TTs (the config-3 construction) the certified
steps only drop exact null directions, so the ranks equal the oracle's and
the error sits at rounding level."""
import numpy as np
import pytest

import oracle as O
import tensor_train_amd as T
from tests.ttutil import rel_err

pytestmark = pytest.mark.gpu


def dev_rel(a, ref):
    """||a - ref|| / ||ref|| by QR on the device (no squaring)."""
    dif = T.sub(a, ref)
    T.right_orthogonalize(dif)
    r = ref.copy()
    T.right_orthogonalize(r)
    return float(np.linalg.norm(dif.cores[0]) / np.linalg.norm(r.cores[0]))


def _sum_batch(count, seed, dmax=40):
    rng = np.random.default_rng(seed)
    d = rng.integers(8, dmax + 1, count)
    n = rng.choice([2, 4, 8], count)
    ru = rng.integers(8, 65, count)
    rv = rng.integers(8, 65, count)
    d[0], n[0], ru[0], rv[0] = 12, 8, 64, 64  # stored ranks 192
    d[1], n[1], ru[1], rv[1] = 40, 2, 60, 50
    return T.build_sum_batch(d, n, ru, rv, 2000 + np.arange(count), 800000 + np.arange(count)), d


def test_mixed_config3_shapes_against_oracle():
    eps = 1e-6
    batch, d = _sum_batch(40, 31)
    ctx = T.Context.default(0)
    ctx.prof_enable(True)
    mixed = T.round_batched(batch, eps=eps, precision="mixed")
    tried, held = ctx.rank_revealing_counts()
    ctx.prof_enable(False)
    assert tried > 0 and held >= 0.5 * tried
    plain = T.round_batched(batch, eps=eps)
    for i, (m, p) in enumerate(zip(mixed, plain)):
        assert all(a >= b for a, b in zip(m.ranks, p.ranks)), i
        assert dev_rel(m, batch[i]) <= eps
    # Samples against the oracle (short TTs keep the CPU time small).
    for i in [0] + [j for j in range(2, 40) if d[j] <= 14][:3]:
        x = O.OTT.from_cores(batch[i].cores)
        ref = O.round(x, eps, 0, 0)
        assert mixed[i].ranks == ref.ranks, i
        err = rel_err(mixed[i].cores, ref)
        assert err <= 1e-6, (i, err)
        assert err <= 1e-12  # sums: only exact null directions are dropped


def test_mixed_long_sum_against_plain():
    """d = 40, n = 2: the mixed result against the plain one (which matches
    the oracle in test_gpu_config_shapes.py): equal ranks, and a difference
    far inside the 1e-6 bar (CholQR2's rounding, about kappa u per step,
    summed over 39 steps: measured 1.1e-12)."""
    batch, _ = _sum_batch(4, 77)
    mixed = T.round_batched(batch[:2], eps=1e-6, precision="mixed")
    plain = T.round_batched(batch[:2], eps=1e-6)
    for m, p in zip(mixed, plain):
        assert m.ranks == p.ranks
        assert dev_rel(m, p) <= 1e-10


def test_mixed_decaying_spectra():
    """Config-4 construction: singular values decay smoothly across the cut,
    so the certificate may cut less than the rank rule; ranks >= plain, the
    error bound holds and the result is within 1e-6 of the oracle's."""
    eps = 1e-6
    rng = np.random.default_rng(8)
    s = rng.uniform(0.05, 0.3, 12)
    batch = T.build_decay_batch(12, 16, 4, 64, s, 300 + np.arange(12))
    mixed = T.round_batched(batch, eps=eps, precision="mixed")
    plain = T.round_batched(batch, eps=eps)
    for i in (0, 5, 11):
        assert all(a >= b for a, b in zip(mixed[i].ranks, plain[i].ranks))
        assert dev_rel(mixed[i], batch[i]) <= eps
        x = O.OTT.from_cores(batch[i].cores)
        ref = O.round(x, eps, 0, 0)
        assert dev_rel(mixed[i], T.TensorTrain.from_cores(ref.cores())) <= 1e-6


@pytest.mark.parametrize("eps,cap", [(1e-10, 0), (1e-3, 5), (0.0, 0)])
def test_mixed_edge_cases(eps, cap):
    rng = np.random.default_rng(5)
    inputs = []
    for t in range(10):
        dd = int(rng.integers(3, 9))
        n = int(rng.choice([2, 3, 4]))
        u = O.OTT.random([n] * dd, int(rng.integers(2, 12)), 40 + t)
        v = O.OTT.random([n] * dd, int(rng.integers(2, 12)), 90 + t)
        inputs.append(O.add(u, O.add(O.scale(u, 0.5), v)))
    inputs.append(O.OTT.from_cores([np.arange(1.0, 6.0).reshape(1, 5, 1)]))        # d = 1
    inputs.append(O.OTT.from_cores([np.zeros((1, 3, 1)), np.zeros((1, 2, 1))]))  # zero TT
    dev = [T.TensorTrain.from_cores(o.cores()) for o in inputs]
    outs = T.round_batched(dev, eps=eps, max_rank=cap, precision="mixed")
    for o, y in zip(inputs, outs):
        ref = O.round(o, eps, cap, 0)
        if O.norm(ref) == 0:
            assert np.all(y.to_dense() == 0)
            continue
        if cap:
            assert max(y.ranks) <= cap
        assert rel_err(y.cores, ref) <= (1e-6 if eps < 1e-6 else 2 * eps + 1e-12)
        if eps <= 1e-10:
            assert y.ranks == ref.ranks
