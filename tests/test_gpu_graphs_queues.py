"""HIP-graph replay of the block-Jacobi sweeps (tta_set_graphs) and the
persistent work-queue launches of the batched whole-problem kernels: results
are bit-identical to the directly enqueued launches, and each graph replay is
checked against the eager result on several calls in a row with eager work
and synchronizations between them."""
import numpy as np
import pytest

import oracle as O
import tensor_train_amd as T
from tests.ttutil import cfg2_input

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert a.ranks == b.ranks
    for x, y in zip(a.cores, b.cores):
        assert np.array_equal(x, y)


def test_graph_replay_single_tt_jacobi():
    """matvec_round with every SVD step running its block-Jacobi SVD
    (certificate off; q = 256): graphs on and off give identical cores, on
    three graph calls with eager calls in between."""
    A, x = cfg2_input(d=10)
    ctx = T.Context.default(0)
    Ad = T.TensorTrainMatrix.from_cores([c.reshape(c.shape[0], 4, 4, c.shape[2]) for c in A.cores()])
    xd = T.TensorTrain.from_cores(x.cores())
    ctx.set_rank_certificate(False)
    try:
        ref = T.matvec_round(Ad, xd, eps=1e-8)
        for _ in range(3):
            ctx.set_graphs(True)
            got = T.matvec_round(Ad, xd, eps=1e-8)
            ctx.set_graphs(False)
            ctx.synchronize()
            _same(got, ref)
            eager = T.matvec_round(Ad, xd, eps=1e-8)
            _same(eager, ref)
    finally:
        ctx.set_graphs(False)
        ctx.set_rank_certificate(True)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replay_batched(precision):
    rng = np.random.default_rng(21)
    s = rng.uniform(0.05, 0.3, 40)
    batch = T.build_decay_batch(40, 12, 4, 96, s, 60 + np.arange(40))
    ctx = T.Context.default(0)
    ref = T.round_batched(batch, eps=1e-6, precision=precision)
    for _ in range(2):
        ctx.set_graphs(True)
        try:
            got = T.round_batched(batch, eps=1e-6, precision=precision)
        finally:
            ctx.set_graphs(False)
        for a, b in zip(got, ref):
            _same(a, b)


def test_persistent_queue_small_problems_against_oracle():
    """Many uneven small problems (q <= 64): the persistent whole-problem
    Jacobi and pivoted Cholesky take them largest first from a work queue;
    every result still matches the oracle."""
    rng = np.random.default_rng(9)
    inputs = []
    for t in range(70):
        d = int(rng.integers(3, 7))
        n = int(rng.choice([2, 3, 4]))
        u = O.OTT.random([n] * d, int(rng.integers(2, 16)), 300 + t)
        v = O.OTT.random([n] * d, int(rng.integers(2, 16)), 700 + t)
        inputs.append(O.add(u, O.add(O.scale(u, 0.5), v)))
    dev = [T.TensorTrain.from_cores(o.cores()) for o in inputs]
    for precision in ("fp64", "mixed"):
        outs = T.round_batched(dev, eps=1e-10, precision=precision)
        for o, y in zip(inputs, outs):
            ref = O.round(o, 1e-10, 0, 0)
            assert y.ranks == ref.ranks
            d = y.to_dense() - ref.to_dense()
            assert np.linalg.norm(d) <= 1e-12 * np.linalg.norm(ref.to_dense())


def test_autotuned_tiles_identical():
    """The grouped GEMM's 64 and 32 tiles accumulate every output element in
    the same order: batched rounding gives bit-identical cores with the tuner
    exploring both tiles (its first launches of each shape class), with the
    tuner's choices, and with the tuner off (64 tile only)."""
    rng = np.random.default_rng(31)
    d = rng.integers(4, 12, 30)
    n = rng.choice([2, 4], 30)
    ru = rng.integers(2, 24, 30)
    rv = rng.integers(2, 24, 30)
    batch = T.build_sum_batch(d, n, ru, rv, 10 + np.arange(30), 90 + np.arange(30))
    ctx = T.Context.default(0)
    ctx.set_autotune(False)
    try:
        ref = T.round_batched(batch, eps=1e-9)
        ref_m = T.round_batched(batch, eps=1e-9, precision="mixed")
    finally:
        ctx.set_autotune(True)
    for _ in range(3):
        got = T.round_batched(batch, eps=1e-9)
        got_m = T.round_batched(batch, eps=1e-9, precision="mixed")
        ctx.synchronize()
        for a, b in zip(got, ref):
            _same(a, b)
        for a, b in zip(got_m, ref_m):
            _same(a, b)
    keys, decided, small = ctx.autotune_stats()
    assert keys > 0 and decided > 0
