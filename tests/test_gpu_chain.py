"""Chained sweep blocks (tta_set_chain).

Inside a speculative sweep, runs of CholQR2-shaped steps are factored as one
block: the Grams chained by products, the Choleskys and the orthonormalizing
products batched, a second pass by the same chain.  In exact arithmetic the
result is the sequential sweep's, so the checks are: the oracle's ranks and
tensor (to 1e-13 relative, as the sequential path), agreement with the
chain switched off to rounding, orthonormal unfoldings, the block counters,
and the fallbacks (a truncating SVD sweep, singular Grams) giving the
oracle's result."""
import numpy as np
import pytest

import oracle as O
import tensor_train_amd as T
from tests.test_gpu_als import to_dev
from tests.ttutil import cfg2_input, rel_err

pytestmark = pytest.mark.gpu


def cores3(t):
    return [c.reshape(c.shape[0], -1, c.shape[-1]) for c in t.cores]


def _run(ctx, fn, chain):
    # set_speculation also clears the record of an earlier fallback (after
    # one, the next sweep of that kind would run checked).
    ctx.set_speculation(True)
    ctx.set_chain(chain)
    ctx.prof_enable(True)
    try:
        out = fn()
        return out, ctx.chain_stats(), ctx.speculation_counts()
    finally:
        ctx.prof_enable(False)
        ctx.set_chain(True)


def _orth_err(t, gauge):
    """Largest |U^T U - I| over the orthonormal unfoldings of the result:
    left unfoldings of cores 0..d-2 (gauge none), right unfoldings of cores
    1..d-1 (right canonical)."""
    cs = cores3(t)
    worst = 0.0
    idx = range(len(cs) - 1) if gauge == "none" else range(1, len(cs))
    for k in idx:
        c = cs[k]
        U = c.reshape(-1, c.shape[2]) if gauge == "none" else c.reshape(c.shape[0], -1).T
        worst = max(worst, np.abs(U.T @ U - np.eye(U.shape[1])).max())
    return worst


@pytest.mark.parametrize("gauge", ["none", "right_canonical"])
def test_chain_on_mpo_product(gauge):
    A, x = cfg2_input(d=12)
    ref = O.round(O.matvec(A, x), 1e-8, 0, 1 if gauge == "right_canonical" else 0)
    ctx = T.Context.default()
    y = T.matvec(to_dev(A), to_dev(x))
    on, (blocks, steps, one, rej), (sw, failed) = _run(ctx, lambda: T.round(y, eps=1e-8, gauge=gauge), True)
    off, (b0, _, _, _), _ = _run(ctx, lambda: T.round(y, eps=1e-8, gauge=gauge), False)
    # One block per sweep over the rank-256 interior (product ranks
    # 1, 32, 128, 256, .., 256, 128, 32, 1; the ends take single steps);
    # the QR sweep's block is orthonormal after its first pass (condition
    # numbers about 10), the SVD sweep's always takes both.
    assert (blocks, one, rej, sw, failed, b0) == (2, 1, 0, 1, 0, 0)
    assert steps >= 10
    assert on.ranks == ref.ranks == off.ranks
    assert rel_err(cores3(on), ref) <= 1e-13
    assert rel_err(cores3(on), O.OTT.from_cores(cores3(off))) <= 1e-13
    assert _orth_err(on, gauge) <= 1e-13


def test_chain_matvec_round_full_size():
    """The d = 20 benchmark workload (product ranks 256) against the oracle."""
    A, x = cfg2_input(d=20)
    ref = O.round(O.matvec(A, x), 1e-8)
    ctx = T.Context.default()
    Ad, xd = to_dev(A), to_dev(x)
    y, (blocks, steps, one, rej), (sw, failed) = _run(ctx, lambda: T.matvec_round(Ad, xd, eps=1e-8), True)
    assert (blocks, one, rej, sw, failed) == (2, 1, 0, 1, 0)
    assert steps >= 25
    assert y.ranks == ref.ranks
    assert rel_err(cores3(y), ref) <= 1e-12
    assert _orth_err(y, "none") <= 1e-13


def _varied_tt(seed):
    """Gaussian cores with unequal ranks inside one chainable run (every
    step tall: r_k <= 4 r_{k+1} and 4 r_k >= r_{k+1})."""
    rng = np.random.default_rng(seed)
    ranks = [1, 4, 16, 40, 64, 100, 150, 96, 200, 70, 33, 16, 4, 1]
    cores = [rng.standard_normal((ranks[k], 4, ranks[k + 1])) / np.sqrt(4.0 * ranks[k]) for k in range(len(ranks) - 1)]
    return O.OTT.from_cores(cores)


@pytest.mark.parametrize("gauge", ["none", "right_canonical"])
def test_chain_unequal_ranks(gauge):
    a = _varied_tt(5)
    ref = O.round(a, 1e-12, 0, 1 if gauge == "right_canonical" else 0)
    ctx = T.Context.default()
    ad = to_dev(a)
    on, (blocks, _, _, _), (sw, failed) = _run(ctx, lambda: T.round(ad, eps=1e-12, gauge=gauge), True)
    off, _, _ = _run(ctx, lambda: T.round(ad, eps=1e-12, gauge=gauge), False)
    assert blocks >= 2 and failed == 0
    assert on.ranks == ref.ranks == off.ranks
    assert rel_err(cores3(on), ref) <= 1e-13
    assert rel_err(cores3(on), O.OTT.from_cores(cores3(off))) <= 1e-13
    assert _orth_err(on, gauge) <= 1e-13


def test_chain_falls_back_where_the_rule_cuts():
    """A decaying-spectrum TT (d = 10, bonds 64, singular values
    10^(-0.05 j)) rounded at eps = 1e-2: the QR sweep's block holds; the SVD
    sweep's certificates fail where the rule cuts (whether its block is
    accepted or run step by step), and the sweep is run again checked,
    giving the oracle's result."""
    t = T.build_decay_batch(1, 10, 4, 64, [0.05], [7])[0]
    a = O.OTT.from_cores(cores3(t))
    ref = O.round(a, 1e-2)
    assert max(ref.ranks) < 64
    ctx = T.Context.default()
    ctx.set_speculation(True)
    r, (blocks, _, _, rej), (sw, failed) = _run(ctx, lambda: T.round(t, eps=1e-2), True)
    ctx.set_speculation(True)
    # One block per sweep, accepted or (where the chained second pass leaves
    # the polar form's range) rejected and run step by step.
    assert blocks + rej == 2 and blocks >= 1 and failed == 1
    assert r.ranks == ref.ranks
    assert rel_err(cores3(r), ref) <= 1e-12


def test_chain_rejects_singular_grams():
    """x + x with bonds 80: every Gram of the sum is singular, so a block's
    Choleskys break down; the block is rejected on the host with the TT
    untouched and its steps run one by one (CholQR2 declines in turn and
    Householder takes them); the sum's bonds are cut, so the speculative SVD
    sweep is undone; the result is the oracle's."""
    x = O.OTT.random([4] * 8, 40, 91)
    s = O.add(x, x)
    ref = O.round(s, 1e-12)
    ctx = T.Context.default()
    on, (blocks, _, _, rej), (sw, failed) = _run(ctx, lambda: T.round(to_dev(s), eps=1e-12), True)
    ctx.set_speculation(True)
    assert rej >= 1 and failed >= 1
    assert on.ranks == ref.ranks
    assert rel_err(cores3(on), ref) <= 1e-12


def _graded_tt(d, r, decay, seed):
    """Left-orthogonal Gaussian cores with every bond spectrum rescaled to
    decay^j (the full-size QTT workload's construction, at bond r)."""
    x = O.OTT.random([4] * d, r, seed)
    O.left_orthogonalize(x)
    cs = x.cores()
    for k in range(len(cs) - 1):
        cs[k] = cs[k] * (decay ** np.arange(cs[k].shape[2]))[None, None, :]
    return O.OTT.from_cores(cs)


def test_chain_rejected_block_keeps_the_sweep():
    """Graded bonds 0.85^j (condition about 1e9 over 128 values): a chained
    block's second pass can fall outside the polar form's range where the
    gradings of neighbouring cores meet, although each single CholQR2 step
    would pass.  Whether or not a block is rejected, the QR sweep is not
    undone for it, and the result is the oracle's."""
    g = _graded_tt(10, 128, 0.85, 23)
    ref = O.round(g, 1e-10)
    ctx = T.Context.default()
    gd = to_dev(g)
    r, (blocks, _, _, rej), (sw, failed) = _run(ctx, lambda: T.round(gd, eps=1e-10), True)
    off, _, (sw0, failed0) = _run(ctx, lambda: T.round(gd, eps=1e-10), False)
    # The chain changes no decision of the sweeps themselves.
    assert failed <= failed0
    assert blocks + rej >= 1
    assert r.ranks == ref.ranks == off.ranks
    assert rel_err(cores3(r), ref) <= 1e-12
