"""The full-rank certificate of the device SVD steps (tta_set_rank_certificate).

A step whose triangular factor R satisfies 1/||R^-1||_F > delta (with margin)
keeps every singular value under the reference's rank rule
(tt_eigen_bridge.hpp:172-197), so the step may keep Q R instead of U S V^T:
the same ranks and the same tensor in another orthogonal gauge.  Checked here:
certificate on and off give equal ranks, both match the oracle's full-SVD
rounding within 1e-13 on an MPO product (1e-12 elsewhere), the certificate
holds on the non-truncating bonds of that product and fails wherever the rule
cuts, and it is not tried when max_rank binds."""
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
    ctx.set_rank_certificate(on)
    ctx.prof_enable(True)
    try:
        out = fn()
        return out, ctx.rank_certificates()
    finally:
        ctx.prof_enable(False)
        ctx.set_rank_certificate(True)


@pytest.mark.parametrize("gauge", [0, 1])
def test_certificate_on_mpo_product(gauge):
    A, x = cfg2_input(d=12)
    ref = O.round(O.matvec(A, x), 1e-8, 0, gauge)
    Ad, xd = to_dev(A), to_dev(x)
    ctx = T.Context.default()
    y = T.matvec(Ad, xd)
    on, (tried, held) = _run(ctx, lambda: T.round(y, eps=1e-8, gauge="right_canonical" if gauge else "none"), True)
    off, (tried0, held0) = _run(ctx, lambda: T.round(y, eps=1e-8, gauge="right_canonical" if gauge else "none"),
                                False)
    assert tried0 == 0 and held0 == 0
    assert tried >= 8 and held == tried
    assert on.ranks == off.ranks == ref.ranks
    # Both device paths sit at the FP64 noise floor of the oracle, far inside
    # the 1e-12 bar: the certified steps keep Q R, and the SVD steps form the
    # absorbed factor from the polar-corrected rotations by one GEMM.
    assert rel_err(cores3(on), ref) <= 1e-13
    assert rel_err(cores3(off), ref) <= 1e-13
    cs = on.cores
    if gauge == 0:
        for c in cs[:-1]:
            M = c.reshape(-1, c.shape[-1])
            assert np.abs(M.T @ M - np.eye(M.shape[1])).max() <= 1e-13
    else:
        for c in cs[1:]:
            M = c.reshape(c.shape[0], -1)
            assert np.abs(M @ M.T - np.eye(M.shape[0])).max() <= 1e-13


def test_certificate_on_matvec_round_path():
    # The default matvec_round (naive: apply, then round) uses it as well.
    A, x = cfg2_input(d=10)
    ref = O.round(O.matvec(A, x), 1e-8)
    ctx = T.Context.default()
    y, (tried, held) = _run(ctx, lambda: T.matvec_round(to_dev(A), to_dev(x), eps=1e-8), True)
    assert held > 0
    assert y.ranks == ref.ranks and rel_err(cores3(y), ref) <= 1e-12


def test_certificate_fails_where_the_rule_cuts():
    # Config 1: every interior bond of add(add(a, b), a) truncates 24 -> 16.
    c, a, b = cfg1_input()
    ref = O.round(c, 1e-10)
    ctx = T.Context.default()
    r, (tried, held) = _run(ctx, lambda: T.round(to_dev(c), eps=1e-10), True)
    assert r.ranks == ref.ranks == [1, 4, 16, 16, 16, 16, 16, 16, 16, 4, 1]
    assert rel_err(cores3(r), ref) <= 1e-12
    assert held < tried
    # Graded spectra (0.9^j QTT bonds, eps 1e-10) cut at most bonds.
    y = O.OTT.random([2] * 16, 24, 5)
    for eps in (1e-2, 1e-6):
        ref = O.round(y, eps)
        r, _ = _run(ctx, lambda: T.round(to_dev(y), eps=eps), True)
        assert r.ranks == ref.ranks and rel_err(cores3(r), ref) <= 1e-12


def test_certificate_not_tried_under_binding_max_rank():
    A, x = cfg2_input(d=8)
    Y = O.matvec(A, x)
    ref = O.round(Y, 1e-8, 20)
    ctx = T.Context.default()
    r, (tried, held) = _run(ctx, lambda: T.round(to_dev(Y), eps=1e-8, max_rank=20), True)
    assert r.ranks == ref.ranks and max(r.ranks) <= 20
    assert rel_err(cores3(r), ref) <= 1e-12
    # Only bonds whose size is already within the cap are tried.
    assert tried <= 4


def test_certificate_rank_deficient_input_falls_back():
    # Exact zero padding: zero pivots make the certificate fail, the SVD drops
    # the padding as the oracle does.
    a = O.OTT.random([3, 4, 5, 4, 3], 4, 9)
    cores = a.cores()
    pad = []
    for k, cc in enumerate(cores):
        rl = cc.shape[0] + (3 if k > 0 else 0)
        rr = cc.shape[2] + (3 if k < len(cores) - 1 else 0)
        p = np.zeros((rl, cc.shape[1], rr))
        p[: cc.shape[0], :, : cc.shape[2]] = cc
        pad.append(p)
    b = O.OTT.from_cores(pad)
    ref = O.round(b, 1e-10)
    ctx = T.Context.default()
    r, (tried, held) = _run(ctx, lambda: T.round(to_dev(b), eps=1e-10), True)
    assert r.ranks == ref.ranks and rel_err(cores3(r), ref) <= 1e-12
