"""The Gram-driven block-Jacobi rounds (one problem of 41..512 columns or a
few of 65..512, tta_set_jacobi_gram) against LAPACK and against the staged
rounds (q = 24 runs the whole-problem kernel either way): the
singular values to backward-stable accuracy (absolute errors ~ u s_max, as
LAPACK and Eigen's BDCSVD give), orthonormal U, the product reproduced, the
same sweep counts on graded spectra, and matvec_round with every SVD run
equal to the oracle in ranks and within 1e-12."""
import numpy as np
import pytest

import oracle as O
import tensor_train_amd as T

pytestmark = pytest.mark.gpu


def _graded(q, decades, zero_from, seed):
    rng = np.random.default_rng(seed)
    U0, _ = np.linalg.qr(rng.standard_normal((q, q)))
    V0, _ = np.linalg.qr(rng.standard_normal((q, q)))
    sv = 10.0 ** (-decades * np.arange(q) / (q - 1))
    sv[zero_from:] = 0.0
    return (U0 * sv) @ V0.T


@pytest.mark.parametrize("q,decades", [(24, 5), (48, 8), (64, 6), (72, 6), (130, 4), (256, 12), (300, 8), (512, 4)])
def test_gram_rounds_vs_lapack_and_staged(q, decades):
    ctx = T.Context.default()
    M = _graded(q, decades, q - q // 5, q)
    s_ref = np.linalg.svd(M, compute_uv=False)
    out = {}
    try:
        for gram in (False, True):
            ctx.set_jacobi_gram(gram)
            ctx.prof_enable(True)
            U, s, Vt = T.svd_thin(M)
            out[gram] = (U, s, Vt, ctx.jacobi_sweeps()[1], ctx.jacobi_unconverged())
            ctx.prof_enable(False)
    finally:
        ctx.set_jacobi_gram(True)
    for gram, (U, s, Vt, sweeps, unconverged) in out.items():
        assert unconverged == 0
        assert np.abs(s - s_ref).max() <= 1e-13, gram
        assert np.abs(U.T @ U - np.eye(q)).max() <= 1e-12, gram
        keep = s > 1e-12
        assert np.abs((U[:, keep] * s[keep]) @ Vt[keep] - M).max() <= 1e-13, gram
    # The rounds decide the same rotations up to rounding: sweep counts agree
    # within one (for q <= 64 the staged path is the whole-problem kernel,
    # whose sweeps are not counted).
    if q > 64:
        assert abs(out[True][3] - out[False][3]) <= 1


def test_gram_rounds_matvec_round_every_svd_vs_oracle():
    # d = 10 with ranks 256 in the middle: every SVD step of the sweep runs
    # (certificate off) through the Gram rounds.
    A = O.OTT.random_matrix([4] * 10, [4] * 10, 8, 3)
    x = O.OTT.random([4] * 10, 32, 4)
    yref = O.matvec_round(A, x, 1e-8, 0, 0)
    ctx = T.Context.default()
    ctx.set_rank_certificate(False)
    try:
        Ad = T.TensorTrainMatrix.from_cores([c.reshape(c.shape[0], 4, 4, c.shape[2]) for c in A.cores()])
        yd = T.matvec_round(Ad, T.TensorTrain.from_cores(x.cores()), eps=1e-8)
    finally:
        ctx.set_rank_certificate(True)
    assert yd.ranks == yref.ranks
    d = yref.to_dense()
    assert np.linalg.norm(yd.to_dense() - d) / np.linalg.norm(d) <= 1e-12
