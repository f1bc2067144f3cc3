"""CPU tests of the oracle's two-site DMRG rounding
(include/core/round/tt_dmrg.hpp, tt_matvec_dmrg_round.hpp,
tt_matmat_dmrg_round.hpp, include/core/detail/dmrg_local_solve.hpp of
bdaszuta/tensor_train).

Pins, none of which reuses the oracle's environment or merge code:
  * the assertions of the reference's own DMRG cases
    (tests/core/test_tt_round.cpp:227-384);
  * d = 2 closed form: L = R = [1], so every bond solve is the truncated SVD
    of Y as an n0 x n1 matrix (numpy's SVD and the truncation rule);
  * the final R->L split at bond 0 is the truncated SVD of the dense
    projection of Y onto the right interface of X_2..X_{d-1};
  * at a converged fixed point every one-site partial gradient of
    ||Y - X||^2 vanishes (brute-force dense contraction);
  * the zero tensor, max_iters = 0, d = 1 and the apply-path starts."""
import numpy as np
import pytest

import oracle as O
from tests.test_oracle_als import _np_tt_svd, _pad, _partial_gradients, _rel


def _reference_case():
    rng = np.random.default_rng(13)
    shape = [4, 5, 6, 4]
    T = rng.standard_normal(int(np.prod(shape)))
    cores = _np_tt_svd(T, shape)
    return shape, T, O.OTT.from_cores(cores), O.OTT.from_cores(_pad(cores, 5))


def test_dmrg_reference_cases():
    # tests/core/test_tt_round.cpp:227-384.
    shape, T, A, B = _reference_case()
    cap = max(A.ranks)
    X, it, ch, cv = O.round_dmrg(B, 1e-12, cap, 3, 1e-12)
    assert _rel(X.to_dense(), T) <= 1e-10 and max(X.ranks) <= cap
    X, it, ch, cv = O.round_dmrg(B, 1e-12, cap, 3, 1e-12, warm=A)
    assert _rel(X.to_dense(), T) <= 1e-10
    e_round = _rel(O.round(B, 1e-10, cap).to_dense(), T)
    e_dmrg = _rel(O.round_dmrg(B, 1e-12, cap, 3, 1e-13)[0].to_dense(), T)
    assert e_dmrg <= 1.1 * e_round + 1e-12
    Bp = O.axpby(1.0, A, 1e-3, O.OTT.random(shape, 4, 99))
    X, it, ch, cv = O.round_dmrg(Bp, 1e-12, cap, 3, 1e-12, warm=A)
    assert it <= 5 and _rel(X.to_dense(), Bp.to_dense()) <= 1e-2
    X, it, ch, cv = O.round_dmrg(B, 1e-10, 0, 3, 1e-12)
    assert _rel(X.to_dense(), T) <= 1e-8 and max(X.ranks) <= max(B.ranks)


def _np_truncate(s, delta, max_rank=0):
    tail, r = 0.0, len(s)
    for i in range(len(s) - 1, -1, -1):
        if tail + s[i] ** 2 > delta ** 2:
            break
        tail += s[i] ** 2
        r = i
    r = max(r, 1)
    return min(r, max_rank) if max_rank > 0 else r


@pytest.mark.parametrize("eps,cap", [(1e-1, 0), (1e-3, 0), (1e-12, 3)])
def test_dmrg_d2_is_truncated_matrix_svd(eps, cap):
    rng = np.random.default_rng(7)
    U = np.linalg.qr(rng.standard_normal((6, 6)))[0]
    V = np.linalg.qr(rng.standard_normal((7, 6)))[0]
    s = 10.0 ** -np.arange(6.0)
    M = U @ np.diag(s) @ V.T
    Y = O.OTT.from_cores([np.eye(6).reshape(1, 6, 6), M.reshape(6, 7, 1)])
    X, it, ch, cv = O.round_dmrg(Y, eps, cap, 5, 1e-12)
    sv = np.linalg.svd(M, compute_uv=False)
    r = _np_truncate(sv, eps * np.linalg.norm(M), cap)
    Uf, sf, Vf = np.linalg.svd(M)
    want = (Uf[:, :r] * sf[:r]) @ Vf[:r]
    assert X.ranks == [1, r, 1]
    assert np.abs(X.to_dense() - want.ravel()).max() <= 1e-14 * sv[0]
    assert cv and it == 2


def test_dmrg_last_split_is_dense_two_site_projection():
    shape = [3, 4, 5, 4]
    Yt = O.axpby(1.0, O.OTT.random(shape, 4, 21), 0.3, O.OTT.random(shape, 5, 22))
    eps = 0.05
    for iters in (1, 2):
        X, it, ch, cv = O.round_dmrg(Yt, eps, 0, iters, 0.0)
        cs = X.cores()
        for c in cs[1:]:
            Ru = c.reshape(c.shape[0], -1)
            assert np.abs(Ru @ Ru.T - np.eye(c.shape[0])).max() <= 1e-13
        Phi = np.ones((1, 1))
        for c in reversed(cs[2:]):
            Phi = (c.reshape(-1, c.shape[2]) @ Phi).reshape(c.shape[0], -1)
        Th = (Yt.to_dense().reshape(shape[0] * shape[1], -1) @ Phi.T).reshape(shape[0], -1)
        Uf, sf, Vf = np.linalg.svd(Th, full_matrices=False)
        delta = eps * np.linalg.norm(Yt.to_dense()) / np.sqrt(len(shape) - 1)
        r = _np_truncate(sf, delta)
        assert cs[0].shape[2] == r
        want = (Uf[:, :r] * sf[:r]) @ Vf[:r]
        got = cs[0].reshape(shape[0], r) @ cs[1].reshape(r, -1)
        assert np.abs(got - want).max() <= 1e-13 * sf[0]


def test_dmrg_fixed_point_is_stationary():
    shape = [3, 4, 3, 4, 3]
    Yt = O.axpby(1.0, O.OTT.random(shape, 3, 5), 1e-2, O.OTT.random(shape, 6, 6))
    Y = Yt.to_dense()
    X, it, ch, cv = O.round_dmrg(Yt, 1e-14, 3, 200, 1e-14)
    assert cv and it < 200
    scale = np.linalg.norm(Y) * max(np.linalg.norm(c) for c in X.cores())
    for g in _partial_gradients(X.cores(), Y):
        assert np.linalg.norm(g) <= 1e-9 * scale


def test_dmrg_edge_cases():
    shape = [3, 4, 3, 4]
    Yt = O.axpby(1.0, O.OTT.random(shape, 3, 41), 0.5, O.OTT.random(shape, 4, 42))
    # max_iters = 0: the cold start round_right_canonical(Y, eps, max_rank).
    X, it, ch, cv = O.round_dmrg(Yt, 1e-2, 3, 0, 1e-12)
    assert (it, ch, cv) == (0, 0.0, False)
    assert _rel(X.to_dense(), O.round(Yt, 1e-2, 3, gauge=1).to_dense()) <= 1e-13
    # d = 1: the SVD round with {0, 0, true}.
    v = O.OTT.from_cores([np.arange(1.0, 6.0).reshape(1, 5, 1)])
    X, it, ch, cv = O.round_dmrg(v, 1e-3, 0, 5, 1e-12)
    assert (it, ch, cv) == (0, 0.0, True) and np.allclose(X.to_dense(), np.arange(1.0, 6.0))
    # Zero input: rank-1 zero result, converged at the second sweep.
    z = O.OTT.from_cores([np.zeros((1, 3, 2)), np.zeros((2, 4, 2)), np.zeros((2, 3, 1))])
    X, it, ch, cv = O.round_dmrg(z, 1e-6, 0, 5, 1e-12)
    assert X.ranks == [1, 1, 1, 1] and not np.any(X.to_dense())
    assert cv and it == 2


def test_dmrg_apply_paths():
    # matvec: random start at max_rank > 0, else round_right_canonical(Y, eps, 0).
    A = O.OTT.random_matrix([2, 3, 2, 3], [3, 2, 2, 2], 2, 7)
    x = O.OTT.random([3, 2, 2, 2], 2, 8)
    Y = O.matvec(A, x)
    Yd = Y.to_dense()
    X, it, ch, cv = O.matvec_round_dmrg(A, x, 1e-12, 4, 20, 1e-12, seed=0x3030)
    assert _rel(X.to_dense(), Yd) <= 1e-10 and cv
    X, it, ch, cv = O.matvec_round_dmrg(A, x, 1e-2, 0, 20, 1e-12)
    e_svd = _rel(O.round(Y, 1e-2).to_dense(), Yd)
    assert _rel(X.to_dense(), Yd) <= 1.1 * e_svd + 1e-12
    B = O.OTT.random_matrix([3, 2, 2, 2], [2, 2, 3, 2], 2, 9)
    C = O.matmat(A, B)
    Xc, it, ch, cv = O.matmat_round_dmrg(A, B, 1e-12, 4, 20, 1e-12, seed=0x4040)
    assert Xc.m == C.m and _rel(Xc.to_dense(), C.to_dense()) <= 1e-10
