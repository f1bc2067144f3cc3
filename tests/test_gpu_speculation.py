"""Speculative sweeps (tta_set_speculation).

The SVD sweep of round decides CholQR2 acceptance and the rank certificates
on the device, into one flag read once per sweep; a sweep with a failed
decision is undone from a copy of the TT and run again with host decisions
(the QR sweep decides on the host step by step).  Checked here: speculation on and off give bit-identical cores
(and the oracle's ranks and tensor) where every decision holds (the MPO
product, both gauges, and the naive matvec_round built on round) and where
some fail (config 1, whose bonds the rank rule cuts; a zero-padded TT, on
which CholQR2 breaks down); the counters show each fallback, and after one
the next sweep of that kind runs checked.  Chained blocks are switched off
here (their own tests are in test_gpu_chain.py)."""
import numpy as np
import pytest

import oracle as O
import tensor_train_amd as T
from tests.test_gpu_als import to_dev
from tests.ttutil import cfg1_input, cfg2_input, rel_err

pytestmark = pytest.mark.gpu


def cores3(t):
    return [c.reshape(c.shape[0], -1, c.shape[-1]) for c in t.cores]


def _run(ctx, fn, on):
    # Chained blocks (which run only in speculative sweeps) switched off, so
    # that both runs factor step by step and their cores must be identical.
    ctx.set_chain(False)
    ctx.set_speculation(on)
    ctx.prof_enable(True)
    try:
        out = fn()
        return out, ctx.speculation_counts()
    finally:
        ctx.prof_enable(False)
        ctx.set_speculation(True)
        ctx.set_chain(True)


def _same(a, b):
    assert a.ranks == b.ranks
    for x, y in zip(a.cores, b.cores):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("gauge", ["none", "right_canonical"])
def test_speculation_identical_on_mpo_product(gauge):
    A, x = cfg2_input(d=10)
    ref = O.round(O.matvec(A, x), 1e-8, 0, 1 if gauge == "right_canonical" else 0)
    ctx = T.Context.default()
    y = T.matvec(to_dev(A), to_dev(x))
    on, (sw, failed) = _run(ctx, lambda: T.round(y, eps=1e-8, gauge=gauge), True)
    off, (sw0, _) = _run(ctx, lambda: T.round(y, eps=1e-8, gauge=gauge), False)
    assert (sw, failed, sw0) == (1, 0, 0)
    _same(on, off)
    assert on.ranks == ref.ranks and rel_err(cores3(on), ref) <= 1e-13


def test_speculation_identical_on_matvec_round():
    A, x = cfg2_input(d=12)
    ref = O.round(O.matvec(A, x), 1e-8)
    ctx = T.Context.default()
    Ad, xd = to_dev(A), to_dev(x)
    on, (sw, failed) = _run(ctx, lambda: T.matvec_round(Ad, xd, eps=1e-8), True)
    off, _ = _run(ctx, lambda: T.matvec_round(Ad, xd, eps=1e-8), False)
    assert sw == 1 and failed == 0
    _same(on, off)
    assert on.ranks == ref.ranks and rel_err(cores3(on), ref) <= 1e-13


def test_speculation_falls_back_where_the_rule_cuts():
    # Every interior bond of config 1 truncates 24 -> 16: the speculative SVD
    # sweep fails its certificates and is run again; the next call's SVD
    # sweep runs checked.
    c, _, _ = cfg1_input()
    ref = O.round(c, 1e-10)
    ctx = T.Context.default()
    cd = to_dev(c)
    (r1, r2), (sw, failed) = _run(ctx, lambda: (T.round(cd, eps=1e-10), T.round(cd, eps=1e-10)), True)
    off, _ = _run(ctx, lambda: T.round(cd, eps=1e-10), False)
    assert (sw, failed) == (1, 1)
    _same(r1, off)
    _same(r2, off)
    assert r1.ranks == ref.ranks and rel_err(cores3(r1), ref) <= 1e-12


def test_speculation_falls_back_on_cholqr_breakdown():
    # Exact zero padding of a rank-4 TT to bond 40: the Grams are singular,
    # so the QR sweep's CholQR2 declines (Householder takes those steps), and
    # the padding is cut, so the SVD sweep fails its certificates.
    a = O.OTT.random([4, 4, 4, 4, 4, 4], 4, 9)
    pad = []
    cores = a.cores()
    for k, cc in enumerate(cores):
        rl = cc.shape[0] + (36 if k > 0 else 0)
        rr = cc.shape[2] + (36 if k < len(cores) - 1 else 0)
        p = np.zeros((rl, cc.shape[1], rr))
        p[: cc.shape[0], :, : cc.shape[2]] = cc
        pad.append(p)
    b = O.OTT.from_cores(pad)
    ref = O.round(b, 1e-10)
    ctx = T.Context.default()
    bd = to_dev(b)
    on, (sw, failed) = _run(ctx, lambda: T.round(bd, eps=1e-10), True)
    off, _ = _run(ctx, lambda: T.round(bd, eps=1e-10), False)
    assert sw == 1 and failed == 1
    _same(on, off)
    assert on.ranks == ref.ranks and rel_err(cores3(on), ref) <= 1e-12


def test_speculation_rank_deficient_sum_with_cap():
    """round(x + x, max_rank < 2r) with bonds >= 32: the sum's Grams are
    singular, so a speculative CholQR2 breaks down and its (discarded) factors
    hold non-finite values while the sweep continues; the capped SVD steps
    run the Jacobi SVD on them before the sweep's flag is read.  The sorted
    singular values must still index every column exactly once (no
    out-of-range gather), and the undone sweep must give the checked result."""
    x = O.OTT.random([4] * 8, 40, 91)
    s = O.add(x, x)
    cap = 30
    ref = O.round(s, 1e-12, cap, 0)
    ctx = T.Context.default()
    on, (sw, failed) = _run(ctx, lambda: T.round(to_dev(s), eps=1e-12, max_rank=cap), True)
    off, _ = _run(ctx, lambda: T.round(to_dev(s), eps=1e-12, max_rank=cap), False)
    _same(on, off)
    assert on.ranks == ref.ranks
    assert max(on.ranks) == cap
    assert rel_err(cores3(on), ref) <= 1e-12
