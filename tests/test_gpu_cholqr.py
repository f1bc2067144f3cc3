"""CholQR2 for the tall QR factorizations of the device sweeps
(tta_set_cholqr).

CholQR2 gives a backward-stable Q R for condition numbers up to about 1e7 and
declines (falling back to Householder) on breakdown or when its second pass is
far from the identity.  Checked here: it is used on the MPO products of the
benchmark workload, CholQR2 on and off give equal ranks and tensors matching
the oracle within 1e-13 with orthonormal cores, and inputs it must decline
(exact zero padding, steeply graded spectra) still match the oracle."""
import numpy as np
import pytest

import oracle as O
import tensor_train_amd as T
from tests.test_gpu_als import to_dev
from tests.ttutil import cfg2_input, rel_err

pytestmark = pytest.mark.gpu


def cores3(t):
    return [c.reshape(c.shape[0], -1, c.shape[-1]) for c in t.cores]


def _with(ctx, fn, cholqr=True, cert=True):
    ctx.set_cholqr(cholqr)
    ctx.set_rank_certificate(cert)
    ctx.prof_enable(True)
    try:
        out = fn()
        return out, ctx.cholqr_counts()
    finally:
        ctx.prof_enable(False)
        ctx.set_cholqr(True)
        ctx.set_rank_certificate(True)


@pytest.mark.parametrize("cert", [True, False])
def test_cholqr_on_mpo_product(cert):
    A, x = cfg2_input(d=10)
    Y = O.matvec(A, x)
    ref = O.round(Y, 1e-8)
    ctx = T.Context.default()
    Yd = T.matvec(to_dev(A), to_dev(x))
    on, (tried, used) = _with(ctx, lambda: T.round(Yd, eps=1e-8), True, cert)
    off, (tried0, used0) = _with(ctx, lambda: T.round(Yd, eps=1e-8), False, cert)
    assert used > 0 and used == tried and tried0 == 0
    assert on.ranks == off.ranks == ref.ranks
    assert rel_err(cores3(on), ref) <= 1e-13
    assert rel_err(cores3(off), ref) <= 1e-13
    for c in on.cores[:-1]:
        M = c.reshape(-1, c.shape[-1])
        assert np.abs(M.T @ M - np.eye(M.shape[1])).max() <= 1e-13


def test_cholqr_right_orthogonalize():
    A, x = cfg2_input(d=8)
    Y = O.matvec(A, x)
    ctx = T.Context.default()
    Yd = T.matvec(to_dev(A), to_dev(x))
    R, (tried, used) = _with(ctx, lambda: (lambda t: (T.right_orthogonalize(t), t)[1])(Yd.copy()))
    assert used > 0
    assert rel_err(cores3(R), Y) <= 1e-14
    for c in R.cores[1:]:
        M = c.reshape(c.shape[0], -1)
        assert np.abs(M @ M.T - np.eye(M.shape[0])).max() <= 1e-13


def test_cholqr_declines_on_rank_deficient_and_graded():
    ctx = T.Context.default()
    # Exact zero padding of a rank-4 TT to bond 40: singular Grams.
    a = O.OTT.random([4, 4, 4, 4, 4, 4], 4, 9)
    cores = a.cores()
    pad = []
    for k, cc in enumerate(cores):
        rl = cc.shape[0] + (36 if k > 0 else 0)
        rr = cc.shape[2] + (36 if k < len(cores) - 1 else 0)
        p = np.zeros((rl, cc.shape[1], rr))
        p[: cc.shape[0], :, : cc.shape[2]] = cc
        pad.append(p)
    b = O.OTT.from_cores(pad)
    ref = O.round(b, 1e-10)
    r, (tried, used) = _with(ctx, lambda: T.round(to_dev(b), eps=1e-10))
    assert r.ranks == ref.ranks and rel_err(cores3(r), ref) <= 1e-12
    assert used < tried or tried == 0
    # Graded bond spectra 10^-(j/4) over 64 values: condition 1e16.
    x = O.OTT.random([4] * 8, 64, 13)
    O.left_orthogonalize(x)
    cs = x.cores()
    for k in range(len(cs) - 1):
        cs[k] = cs[k] * (10.0 ** (-np.arange(cs[k].shape[2]) / 4.0))[None, None, :]
    g = O.OTT.from_cores(cs)
    for eps in (0.0, 1e-10):
        ref = O.round(g, eps)
        r, _ = _with(ctx, lambda: T.round(to_dev(g), eps=eps))
        assert r.ranks == ref.ranks and rel_err(cores3(r), ref) <= 1e-12


def test_cholqr_forms_across_condition_numbers():
    # Bond spectra 10^-(s j) over 64 values (d = 8, bonds 64, so that the
    # QR sweep factors two tall 256 x 64 unfoldings and the SVD sweep three):
    # whichever form each factorization takes (the polar second pass for
    # small E, the triangular one for n max|E| between 1e-5 and 1e-4,
    # Householder beyond, the last covered by the graded test above), the
    # result reproduces the oracle.  Here the triangular form is taken for
    # decay exponents in a narrow window near 8.5 (measured), so the sweep
    # steps by 1/32 through 8.25 .. 8.75 to reach both CholQR2 forms whatever
    # the last bits of the factorizations before them.
    ctx = T.Context.default()
    forms = {"triangular": 0, "declined": 0, "polar": 0}
    for logk in [2.0, 4.0, 6.0] + list(8.25 + np.arange(17) / 32.0):
        x = O.OTT.random([4] * 8, 64, 17)
        O.left_orthogonalize(x)
        cs = x.cores()
        for k in range(len(cs) - 1):
            cs[k] = cs[k] * (10.0 ** (-np.arange(cs[k].shape[2]) * logk / 63.0))[None, None, :]
        g = O.OTT.from_cores(cs)
        ref = O.round(g, 0.0)
        ctx.set_cholqr(True)
        ctx.set_speculation(False)
        ctx.prof_enable(True)
        try:
            r = T.round(to_dev(g), eps=0.0)
            tried, used = ctx.cholqr_counts()
            tri = ctx.cholqr_triangular_count()
        finally:
            ctx.prof_enable(False)
            ctx.set_speculation(True)
        assert r.ranks == ref.ranks and rel_err(cores3(r), ref) <= 1e-12
        forms["triangular"] += tri
        forms["declined"] += tried - used
        forms["polar"] += used - tri
    assert forms["polar"] > 0 and forms["triangular"] > 0, forms
