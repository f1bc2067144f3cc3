"""Factorizations past the staged kernels' sizes: Householder panels taller
than 2048 rows and Jacobi SVDs of more than 640 columns (the Gram-based block
rounds), against LAPACK (numpy) and against the oracle's round."""
import numpy as np
import pytest

import oracle as O
import tensor_train_amd as T
from tests.ttutil import rel_err

pytestmark = pytest.mark.gpu


def _graded(m, q, decades, seed):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((m, q)))
    V, _ = np.linalg.qr(rng.standard_normal((q, q)))
    s = 10.0 ** (-decades * np.arange(q) / q)
    return (U * s) @ V.T, s


@pytest.mark.parametrize("m,q,decades", [(900, 700, 12), (1100, 1100, 8), (2600, 800, 10)])
def test_svd_thin_large_q(m, q, decades):
    M, s = _graded(m, q, decades, m + q)
    U, sv, Vt = T.svd_thin(M)
    ref = np.linalg.svd(M, compute_uv=False)
    # Two backward-stable SVDs agree to a few q u sigma_max (LAPACK's own
    # error bound is of that size).
    assert np.abs(sv - ref).max() <= 10 * q * 2.2e-16 * ref[0]
    # Orthogonality of accumulated rotations: about q u per column pair.
    assert np.abs(U.T @ U - np.eye(q)).max() <= 10 * q * 2.2e-16
    assert np.abs(Vt @ Vt.T - np.eye(q)).max() <= 10 * q * 2.2e-16
    assert np.linalg.norm((U * sv) @ Vt - M) <= 10 * q * 2.2e-16 * np.linalg.norm(M)


def test_qr_tall_panels():
    rng = np.random.default_rng(4)
    M = rng.standard_normal((5000, 300))
    Q, R = T.qr_thin(M)
    Qo, Ro = O.qr_thin(M)
    assert np.abs(Q.T @ Q - np.eye(300)).max() <= 1e-13
    assert np.abs(Q @ R - M).max() <= 1e-12 * np.abs(M).max()
    assert np.abs(Q - Qo).max() <= 1e-11


@pytest.mark.parametrize("eps,cert", [(1e-3, True), (1e-8, False)])
def test_round_bond_700_vs_oracle(eps, cert):
    """random([30]^4, 700): the SVD step at the middle bond has 700 columns.
    With eps = 1e-8 nothing is cut, so the full-rank certificate is switched
    off to make the step run its SVD."""
    x = O.OTT.random([30] * 4, 700, 5)
    ref = O.round(x, eps, 0, 0)
    ctx = T.Context.default(0)
    ctx.set_rank_certificate(cert)
    try:
        ctx.prof_enable(True)
        y = T.round(T.TensorTrain.from_cores(x.cores()), eps=eps)
        calls, _ = ctx.jacobi_sweeps()
        unconv = ctx.jacobi_unconverged()
        ctx.prof_enable(False)
    finally:
        ctx.set_rank_certificate(True)
    assert calls >= 1 and unconv == 0
    assert y.ranks == ref.ranks
    assert max(ref.ranks) >= 640
    assert rel_err(y.cores, ref) <= 1e-12


def test_round_batched_bond_700_equals_single():
    xs = [O.OTT.random([30] * 4, 700, 6), O.OTT.random([30] * 4, 660, 7), O.OTT.random([4] * 6, 12, 8)]
    dev = [T.TensorTrain.from_cores(o.cores()) for o in xs]
    outs = T.round_batched(dev, eps=1e-3)
    for d_in, o in zip(dev, outs):
        single = T.round(d_in, eps=1e-3)
        assert o.ranks == single.ranks
        a, b = T.sub(o, single), single.copy()
        T.right_orthogonalize(a)
        T.right_orthogonalize(b)
        assert np.linalg.norm(a.cores[0]) <= 1e-12 * np.linalg.norm(b.cores[0])
