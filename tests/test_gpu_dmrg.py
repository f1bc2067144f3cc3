"""Two-site DMRG rounding on the GPU (include/core/round/tt_dmrg.hpp,
tt_matvec_dmrg_round.hpp, tt_matmat_dmrg_round.hpp, tt_matrix_dmrg_round.hpp)
against the oracle on the same inputs and starts.

Bar: equal ranks, iteration counts and convergence flags, relative Frobenius
difference <= 1e-11.  Every bond solve is a truncated SVD of a target that
depends only on the interface subspaces, so the QR and SVD sign conventions
of the two sides do not enter; the inputs have clear gaps at each cut."""
import numpy as np
import pytest

import oracle as O
import tensor_train_amd as T
from tests.test_gpu_als import _padded_case, _same, cores3, to_dev
from tests.ttutil import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-11


def test_dmrg_reference_cases_match_oracle():
    # tests/core/test_tt_round.cpp:227-384 on both sides.
    shape, X, A, B = _padded_case()
    cap = max(A.ranks)
    Bd, Ad = to_dev(B), to_dev(A)
    for warm, wd in ((None, None), (A, Ad)):
        ref, it, ch, cv = O.round_dmrg(B, 1e-12, cap, 3, 1e-12, warm=warm)
        dev, info = T.round_ex(Bd, method="dmrg", eps=1e-12, max_rank=cap, max_iters=3, tol=1e-12, warm_start=wd)
        assert np.linalg.norm(dev.to_dense() - X) / np.linalg.norm(X) <= 1e-10
        assert dev.ranks == ref.ranks and max(dev.ranks) <= cap
        assert np.linalg.norm(dev.to_dense() - ref.to_dense()) <= 1e-10 * np.linalg.norm(X)
    Bp = O.axpby(1.0, A, 1e-3, O.OTT.random(shape, 4, 99))
    ref, it, ch, cv = O.round_dmrg(Bp, 1e-12, cap, 3, 1e-12, warm=A)
    dev, info = T.round_ex(to_dev(Bp), method="dmrg", eps=1e-12, max_rank=cap, max_iters=3, tol=1e-12,
                           warm_start=Ad)
    assert info.iters_run <= 5
    _same(dev, info, ref, it, ch, cv, TOL)
    ref, it, ch, cv = O.round_dmrg(B, 1e-10, 0, 3, 1e-12)
    dev, info = T.round_ex(Bd, method="dmrg", eps=1e-10, max_iters=3, tol=1e-12)
    assert np.linalg.norm(dev.to_dense() - X) / np.linalg.norm(X) <= 1e-8
    assert dev.ranks == ref.ranks and max(dev.ranks) <= max(B.ranks)


@pytest.mark.parametrize("shape,ra,rb,eps,cap,iters", [
    ([3, 4, 3, 4, 3], 4, 6, 1e-3, 0, 3),
    ([4] * 8, 8, 8, 1e-2, 0, 3),
    ([4] * 8, 8, 8, 1e-12, 6, 3),
    ([2, 5, 3, 4, 2, 3, 4], 6, 5, 1e-1, 0, 4),
    ([4] * 12, 16, 16, 1e-2, 12, 2),
])
def test_dmrg_cold_matches_oracle(shape, ra, rb, eps, cap, iters):
    Y = O.axpby(1.0, O.OTT.random(shape, ra, 61), 0.1, O.OTT.random(shape, rb, 62))
    ref, it, ch, cv = O.round_dmrg(Y, eps, cap, iters, 0.0)
    dev, info = T.round_ex(to_dev(Y), method="dmrg", eps=eps, max_rank=cap, max_iters=iters, tol=0.0)
    _same(dev, info, ref, it, ch, cv, TOL)


def test_dmrg_converges_like_oracle():
    shape = [3, 4, 3, 4, 3]
    Y = O.axpby(1.0, O.OTT.random(shape, 3, 71), 1e-2, O.OTT.random(shape, 5, 72))
    ref, it, ch, cv = O.round_dmrg(Y, 1e-12, 3, 100, 1e-10)
    dev, info = T.round_ex(to_dev(Y), method="dmrg", eps=1e-12, max_rank=3, max_iters=100, tol=1e-10)
    assert cv and info.converged and info.iters_run == it
    assert rel_err(cores3(dev), ref) <= TOL


def test_dmrg_matrix_round_matches_oracle():
    A = O.OTT.random_matrix([2, 3, 2, 3], [3, 2, 2, 2], 4, 91)
    B = O.OTT.random_matrix([2, 3, 2, 3], [3, 2, 2, 2], 5, 92)
    Y = O.axpby(1.0, A, 0.1, B)
    ref, it, ch, cv = O.round_dmrg(Y, 1e-2, 0, 3, 0.0)
    dev, info = T.round_ex(to_dev(Y), method="dmrg", eps=1e-2, max_iters=3, tol=0.0)
    assert isinstance(dev, T.TensorTrainMatrix)
    _same(dev, info, ref, it, ch, cv, TOL)


@pytest.mark.parametrize("eps,cap,seed", [(1e-12, 5, 0x3030), (1e-2, 0, 0)])
def test_dmrg_matvec_matches_oracle(eps, cap, seed):
    A = O.OTT.random_matrix([2, 3, 2, 3, 2], [3, 2, 2, 2, 3], 3, 7)
    x = O.OTT.random([3, 2, 2, 2, 3], 4, 8)
    ref, it, ch, cv = O.matvec_round_dmrg(A, x, eps, cap, 3, 0.0, seed=seed)
    dev, info = T.matvec_round_ex(to_dev(A), to_dev(x), method="dmrg", eps=eps, max_rank=cap, max_iters=3, tol=0.0,
                                  seed=seed)
    _same(dev, info, ref, it, ch, cv, TOL)


def test_dmrg_matmat_matches_oracle():
    A = O.OTT.random_matrix([2, 3, 2, 3], [3, 2, 2, 2], 2, 7)
    B = O.OTT.random_matrix([3, 2, 2, 2], [2, 2, 3, 2], 3, 9)
    ref, it, ch, cv = O.matmat_round_dmrg(A, B, 1e-12, 4, 3, 0.0, seed=0x4040)
    dev, info = T.matmat_round_ex(to_dev(A), to_dev(B), method="dmrg", eps=1e-12, max_rank=4, max_iters=3,
                                  tol=0.0, seed=0x4040)
    assert isinstance(dev, T.TensorTrainMatrix)
    _same(dev, info, ref, it, ch, cv, TOL)


def test_dmrg_edge_cases_and_errors():
    shape = [3, 4, 3, 4]
    Y = O.axpby(1.0, O.OTT.random(shape, 3, 41), 0.5, O.OTT.random(shape, 4, 42))
    Yd = to_dev(Y)
    ref, it, ch, cv = O.round_dmrg(Y, 1e-2, 3, 0, 1e-12)
    dev, info = T.round_ex(Yd, method="dmrg", eps=1e-2, max_rank=3, max_iters=0)
    assert (info.iters_run, info.final_resid, info.converged) == (0, 0.0, False)
    assert rel_err(cores3(dev), ref) <= TOL
    z = O.OTT.from_cores([np.zeros((1, 3, 2)), np.zeros((2, 4, 2)), np.zeros((2, 3, 1))])
    ref, it, ch, cv = O.round_dmrg(z, 1e-6, 0, 5, 1e-12)
    dev, info = T.round_ex(to_dev(z), method="dmrg", eps=1e-6, max_iters=5)
    assert dev.ranks == ref.ranks == [1, 1, 1, 1] and not np.any(dev.to_dense())
    assert (info.iters_run, info.converged) == (it, cv)
    v = T.TensorTrain.from_cores([np.arange(1.0, 6.0).reshape(1, 5, 1)])
    dev, info = T.round_ex(v, method="dmrg", eps=1e-3, max_iters=5)
    assert (info.iters_run, info.final_resid, info.converged) == (0, 0.0, True)
    for bad in (0.0, -1.0, float("nan")):
        with pytest.raises(T.TTAError, match="eps must be > 0"):
            T.round(Yd, method="dmrg", eps=bad, max_rank=3)
    with pytest.raises(T.TTAError, match="warm_start max rank"):
        T.round(Yd, method="dmrg", eps=1e-3, max_rank=2, warm_start=Yd)
