"""ALS rounding on the GPU (include/core/round/tt_als.hpp,
tt_matvec_als_round.hpp, tt_matmat_als_round.hpp, tt_matrix_als_round.hpp)
against the oracle on the same inputs and the same starts.

Bar: equal ranks, equal iteration counts and convergence flags, and a
relative Frobenius difference <= 1e-11.  ALS runs a fixed sequence of
least-squares solves whose result is independent of the QR sign conventions
(every iterate depends only on the interface subspaces), so the two
implementations agree to rounding, amplified by the conditioning of the
local problems over a few sweeps; 1e-11 leaves a factor of ~100 over the
differences seen on these inputs.  The random cold start of the apply paths
is the reference generator on both sides, bit for bit."""
import numpy as np
import pytest

import oracle as O
import tensor_train_amd as T
from tests.ttutil import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-11


def to_dev(o):
    if o.m is not None:
        cores = [c.reshape(c.shape[0], m, n, c.shape[2]) for c, m, n in zip(o.cores(), o.m, o.n)]
        return T.TensorTrainMatrix.from_cores(cores)
    return T.TensorTrain.from_cores(o.cores())


def cores3(t):
    return [c.reshape(c.shape[0], -1, c.shape[-1]) for c in t.cores]


def _np_tt_svd(X, shape, eps=1e-14):
    cores, r, M = [], 1, X.reshape(1, -1)
    for k in range(len(shape) - 1):
        M = M.reshape(r * shape[k], -1)
        U, s, Vt = np.linalg.svd(M, full_matrices=False)
        keep = max(1, int(np.sum(s > eps * s[0])))
        cores.append(U[:, :keep].reshape(r, shape[k], keep))
        M = s[:keep, None] * Vt[:keep]
        r = keep
    cores.append(M.reshape(r, shape[-1], 1))
    return cores


def _padded_case():
    rng = np.random.default_rng(13)
    shape = [4, 5, 6, 4]
    X = rng.standard_normal(int(np.prod(shape)))
    cores = _np_tt_svd(X, shape)
    pad = []
    for k, c in enumerate(cores):
        rl = c.shape[0] + (5 if k > 0 else 0)
        rr = c.shape[2] + (5 if k < len(cores) - 1 else 0)
        p = np.zeros((rl, c.shape[1], rr))
        p[: c.shape[0], :, : c.shape[2]] = c
        pad.append(p)
    return shape, X, O.OTT.from_cores(cores), O.OTT.from_cores(pad)


def _same(dev, info, ref, it, ch, cv, tol=TOL):
    assert dev.ranks == ref.ranks
    assert (info.iters_run, info.converged) == (it, cv)
    if ch > 1e-8:
        assert abs(info.final_resid - ch) <= 1e-6 * ch
    assert rel_err(cores3(dev), ref) <= tol


def test_als_reference_cases_match_oracle():
    # tests/core/test_tt_round.cpp:117-226 on both sides.
    shape, X, A, B = _padded_case()
    cap = max(A.ranks)
    Bd, Ad = to_dev(B), to_dev(A)
    for warm, wd in ((None, None), (A, Ad)):
        ref, it, ch, cv = O.round_als(B, 0.0, cap, 3, 1e-12, warm=warm)
        dev, info = T.round_ex(Bd, method="als", max_rank=cap, max_iters=3, tol=1e-12, warm_start=wd)
        assert np.linalg.norm(dev.to_dense() - X) / np.linalg.norm(X) <= 1e-10
        assert dev.ranks == ref.ranks
        assert np.linalg.norm(dev.to_dense() - ref.to_dense()) <= 1e-10 * np.linalg.norm(X)
    Bp = O.axpby(1.0, A, 1e-3, O.OTT.random(shape, 4, 99))
    ref, it, ch, cv = O.round_als(Bp, 0.0, cap, 3, 1e-12, warm=A)
    dev, info = T.round_ex(to_dev(Bp), method="als", max_rank=cap, max_iters=3, tol=1e-12, warm_start=Ad)
    assert info.iters_run <= 5
    _same(dev, info, ref, it, ch, cv)


@pytest.mark.parametrize("shape,ra,rb,cap,iters", [
    ([3, 4, 3, 4, 3], 4, 6, 3, 4),
    ([4] * 8, 8, 8, 6, 3),
    ([2, 5, 3, 4, 2, 3, 4], 6, 5, 5, 5),
    ([4] * 12, 16, 16, 12, 3),
])
def test_als_cold_matches_oracle(shape, ra, rb, cap, iters):
    Y = O.axpby(1.0, O.OTT.random(shape, ra, 61), 0.1, O.OTT.random(shape, rb, 62))
    ref, it, ch, cv = O.round_als(Y, 0.0, cap, iters, 0.0)
    dev, info = T.round_ex(to_dev(Y), method="als", max_rank=cap, max_iters=iters, tol=0.0)
    _same(dev, info, ref, it, ch, cv)


def test_als_converges_like_oracle():
    # Convergence at tol 1e-10: the stopping sweep is the same on both sides.
    shape = [3, 4, 3, 4, 3]
    Y = O.axpby(1.0, O.OTT.random(shape, 3, 71), 1e-2, O.OTT.random(shape, 5, 72))
    ref, it, ch, cv = O.round_als(Y, 0.0, 3, 100, 1e-10)
    dev, info = T.round_ex(to_dev(Y), method="als", max_rank=3, max_iters=100, tol=1e-10)
    assert cv and info.converged and info.iters_run == it
    assert rel_err(cores3(dev), ref) <= TOL


@pytest.mark.parametrize("eps", [1e-1, 1e-3])
def test_als_eps_path_matches_oracle(eps):
    shape = [4, 4, 4, 4, 4, 4]
    Y = O.axpby(1.0, O.OTT.random(shape, 6, 81), 0.05, O.OTT.random(shape, 8, 82))
    ref, it, ch, cv = O.round_als(Y, eps, 0, 4, 0.0)
    dev, info = T.round_ex(to_dev(Y), method="als", eps=eps, max_iters=4, tol=0.0)
    _same(dev, info, ref, it, ch, cv)


def test_als_matrix_round_matches_oracle():
    # round(tt_matrix, als) packs the matrix (tt_matrix_als_round.hpp:40-72);
    # a TT-matrix warm start is used when its (m, n) match.
    A = O.OTT.random_matrix([2, 3, 2, 3], [3, 2, 2, 2], 4, 91)
    B = O.OTT.random_matrix([2, 3, 2, 3], [3, 2, 2, 2], 5, 92)
    Y = O.axpby(1.0, A, 0.1, B)
    ref, it, ch, cv = O.round_als(Y, 0.0, 3, 3, 0.0)
    dev, info = T.round_ex(to_dev(Y), method="als", max_rank=3, max_iters=3, tol=0.0)
    assert isinstance(dev, T.TensorTrainMatrix)
    _same(dev, info, ref, it, ch, cv)
    W = O.round(Y, 0.0, 3)
    ref, it, ch, cv = O.round_als(Y, 0.0, 3, 3, 0.0, warm=W)
    dev, info = T.round_ex(to_dev(Y), method="als", max_rank=3, max_iters=3, tol=0.0, warm_start=to_dev(W))
    _same(dev, info, ref, it, ch, cv)


@pytest.mark.parametrize("seed", [0, 0x3030])
def test_als_matvec_matches_oracle(seed):
    A = O.OTT.random_matrix([2, 3, 2, 3, 2], [3, 2, 2, 2, 3], 3, 7)
    x = O.OTT.random([3, 2, 2, 2, 3], 4, 8)
    ref, it, ch, cv = O.matvec_round_als(A, x, 0.0, 5, 4, 0.0, seed=seed)
    dev, info = T.matvec_round_ex(to_dev(A), to_dev(x), method="als", max_rank=5, max_iters=4, tol=0.0, seed=seed)
    _same(dev, info, ref, it, ch, cv)
    # eps path: the SVD round of the product is the warm start.
    ref, it, ch, cv = O.matvec_round_als(A, x, 1e-2, 0, 3, 0.0)
    dev, info = T.matvec_round_ex(to_dev(A), to_dev(x), method="als", eps=1e-2, max_iters=3, tol=0.0)
    _same(dev, info, ref, it, ch, cv)


def test_als_matmat_matches_oracle():
    A = O.OTT.random_matrix([2, 3, 2, 3], [3, 2, 2, 2], 2, 7)
    B = O.OTT.random_matrix([3, 2, 2, 2], [2, 2, 3, 2], 3, 9)
    ref, it, ch, cv = O.matmat_round_als(A, B, 0.0, 4, 4, 0.0, seed=0x4040)
    dev, info = T.matmat_round_ex(to_dev(A), to_dev(B), method="als", max_rank=4, max_iters=4, tol=0.0,
                                  seed=0x4040)
    assert isinstance(dev, T.TensorTrainMatrix)
    _same(dev, info, ref, it, ch, cv)


def test_als_info_and_errors():
    shape = [3, 4, 3, 4]
    Y = O.axpby(1.0, O.OTT.random(shape, 3, 41), 0.5, O.OTT.random(shape, 4, 42))
    Yd = to_dev(Y)
    ref, it, ch, cv = O.round_als(Y, 0.0, 3, 0, 1e-12)
    dev, info = T.round_ex(Yd, method="als", max_rank=3, max_iters=0)
    assert (info.iters_run, info.final_resid, info.converged) == (0, 0.0, False)
    assert rel_err(cores3(dev), ref) <= TOL
    v = T.TensorTrain.from_cores([np.arange(1.0, 6.0).reshape(1, 5, 1)])
    dev, info = T.round_ex(v, method="als", max_rank=1, max_iters=5)
    assert (info.iters_run, info.final_resid, info.converged) == (0, 0.0, True)
    assert np.allclose(dev.to_dense(), np.arange(1.0, 6.0))
    with pytest.raises(T.TTAError, match="max_rank <= 0 and eps <= 0"):
        T.round(Yd, method="als")
    with pytest.raises(T.TTAError, match="eps is NaN"):
        T.round(Yd, method="als", eps=float("nan"), max_rank=2)
    with pytest.raises(T.TTAError, match="warm_start max rank"):
        T.round(Yd, method="als", max_rank=2, warm_start=Yd)
    # The svd method reports {1, 0, true}.
    _, info = T.round_ex(Yd, eps=1e-8)
    assert (info.iters_run, info.final_resid, info.converged) == (1, 0.0, True)
