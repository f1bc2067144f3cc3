"""CPU tests of the oracle's ALS rounding (include/core/round/tt_als.hpp,
tt_matvec_als_round.hpp, tt_matmat_als_round.hpp of bdaszuta/tensor_train).

Pins, none of which reuses the oracle's environment contractions:
  * the assertions of the reference's own ALS cases
    (tests/core/test_tt_round.cpp:117-226, test_tt_matrix_round.cpp:138-215,
    test_tt_matvec_round.cpp);
  * stationarity: at a converged ALS fixed point every partial gradient
    d||Y - X||^2 / dX_k, formed from brute-force dense contractions, is zero;
  * the last local update is the dense projection of Y onto the right
    interface of X_1..X_{d-1};
  * monotonicity: from the SVD warm start each local update is a least-squares
    minimiser, so the error never exceeds the SVD truncation error;
  * exact recovery of a tensor whose ranks are within the cap, the info
    semantics (first change 1.0, iteration cap, max_iters = 0, d <= 1)."""
import numpy as np
import pytest

import oracle as O
from tests.ttutil import dense_from_cores


def _np_tt_svd(T, shape, eps=1e-14):
    cores, r, M = [], 1, T.reshape(1, -1)
    d = len(shape)
    for k in range(d - 1):
        M = M.reshape(r * shape[k], -1)
        U, s, Vt = np.linalg.svd(M, full_matrices=False)
        keep = max(1, int(np.sum(s > eps * s[0])))
        cores.append(U[:, :keep].reshape(r, shape[k], keep))
        M = s[:keep, None] * Vt[:keep]
        r = keep
    cores.append(M.reshape(r, shape[-1], 1))
    return cores


def _pad(cores, pad):
    out = []
    for k, c in enumerate(cores):
        rl = c.shape[0] + (pad if k > 0 else 0)
        rr = c.shape[2] + (pad if k < len(cores) - 1 else 0)
        p = np.zeros((rl, c.shape[1], rr))
        p[: c.shape[0], :, : c.shape[2]] = c
        out.append(p)
    return out


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _reference_case():
    """The reference's fixture: a dense 4x5x6x4 normal tensor, its exact TT A
    and B = A with every bond zero-padded by 5."""
    rng = np.random.default_rng(13)
    shape = [4, 5, 6, 4]
    T = rng.standard_normal(int(np.prod(shape)))
    cores = _np_tt_svd(T, shape)
    A = O.OTT.from_cores(cores)
    B = O.OTT.from_cores(_pad(cores, 5))
    return shape, T, A, B, max(A.ranks)


def test_als_reference_cold_warm_vs_svd_warm_pert():
    # tests/core/test_tt_round.cpp:117-226.
    shape, T, A, B, cap = _reference_case()
    X, it, ch, cv = O.round_als(B, 0.0, cap, 3, 1e-12)
    assert _rel(X.to_dense(), T) <= 1e-10
    assert max(X.ranks[1:-1]) <= cap
    assert it <= 3
    X, it, ch, cv = O.round_als(B, 0.0, cap, 3, 1e-12, warm=A)
    assert _rel(X.to_dense(), T) <= 1e-10
    e_round = _rel(O.round(B, 1e-10, cap).to_dense(), T)
    e_als = _rel(O.round_als(B, 0.0, cap, 3, 1e-13)[0].to_dense(), T)
    assert e_als <= 10.0 * max(e_round, 1e-13)
    noise = O.OTT.random(shape, 4, 99)
    Bp = O.axpby(1.0, A, 1e-3, noise)
    X, it, ch, cv = O.round_als(Bp, 0.0, cap, 3, 1e-12, warm=A)
    assert it <= 5
    assert _rel(X.to_dense(), Bp.to_dense()) <= 1e-2


def _partial_gradients(cores, Y):
    """d||Y - X||^2 / dX_k for every core, by dense contraction: the residual
    contracted with everything left and right of site k."""
    d = len(cores)
    shape = [c.shape[1] for c in cores]
    Rz = (Y - dense_from_cores(cores)).reshape(shape)
    grads = []
    for k in range(d):
        left = np.ones((1, 1))  # (N_left, r_k)
        for j in range(k):
            left = (left @ cores[j].reshape(cores[j].shape[0], -1)).reshape(-1, cores[j].shape[2])
        right = np.ones((1, 1))  # (r_{k+1}, N_right)
        for j in range(d - 1, k, -1):
            right = (cores[j].reshape(-1, cores[j].shape[2]) @ right).reshape(cores[j].shape[0], -1)
        Rk = Rz.reshape(left.shape[0], shape[k], right.shape[1])
        g = -2.0 * np.einsum("la,lib,cb->aic", left, Rk, right)
        grads.append(g)
    return grads


@pytest.mark.parametrize("shape,rY,cap", [([3, 4, 3, 4, 3], 6, 3), ([2, 3, 4, 3, 2, 3], 5, 4)])
def test_als_fixed_point_is_stationary(shape, rY, cap):
    # A converged ALS iterate minimises ||Y - X|| in every core: all partial
    # gradients vanish.  A dropped term or transposed operand in either
    # environment moves the fixed point off this set.
    a = O.OTT.random(shape, cap, 5)
    b = O.OTT.random(shape, rY, 6)
    Yt = O.axpby(1.0, a, 1e-2, b)
    Y = Yt.to_dense()
    X, it, ch, cv = O.round_als(Yt, 0.0, cap, 200, 1e-14)
    assert cv and it < 200
    g = _partial_gradients(X.cores(), Y)
    scale = np.linalg.norm(Y) * max(np.linalg.norm(c) for c in X.cores())
    for gk in g:
        assert np.linalg.norm(gk) <= 1e-9 * scale


def test_als_last_update_is_dense_projection():
    # After the R->L sweep X_1..X_{d-1} are right-orthogonal and X_0 is the
    # projection of Y onto their interface: X_0 = Y_mat Phi^T.
    shape = [3, 4, 5, 4]
    Yt = O.axpby(1.0, O.OTT.random(shape, 4, 21), 0.3, O.OTT.random(shape, 5, 22))
    for iters in (1, 2):
        X, it, ch, cv = O.round_als(Yt, 0.0, 3, iters, 0.0)
        cs = X.cores()
        for c in cs[1:]:
            Ru = c.reshape(c.shape[0], -1)
            assert np.abs(Ru @ Ru.T - np.eye(c.shape[0])).max() <= 1e-13
        Phi = np.ones((1, 1))
        for c in reversed(cs[1:]):
            Phi = (c.reshape(-1, c.shape[2]) @ Phi).reshape(c.shape[0], -1)
        want = Yt.to_dense().reshape(shape[0], -1) @ Phi.T
        assert np.abs(cs[0].reshape(shape[0], -1) - want).max() <= 1e-13 * np.abs(want).max()


@pytest.mark.parametrize("eps", [1e-2, 1e-4])
def test_als_eps_path_never_worse_than_svd(eps):
    # max_rank <= 0: target = largest SVD rank at eps, warm start = that SVD
    # round, and each local update is a least-squares step, so the error
    # cannot exceed the SVD truncation error.
    shape = [4, 4, 4, 4, 4]
    Yt = O.axpby(1.0, O.OTT.random(shape, 6, 31), 0.05, O.OTT.random(shape, 8, 32))
    Y = Yt.to_dense()
    S = O.round(Yt, eps)
    X, it, ch, cv = O.round_als(Yt, eps, 0, 20, 1e-12)
    assert max(X.ranks) <= max(S.ranks)
    assert _rel(X.to_dense(), Y) <= _rel(S.to_dense(), Y) * (1 + 1e-10) + 1e-14
    # One iteration is already a monotone step from the SVD start.
    X1, it1, ch1, cv1 = O.round_als(Yt, eps, 0, 1, 1e-12)
    assert it1 == 1 and ch1 == 1.0 and not cv1
    assert _rel(X1.to_dense(), Y) <= _rel(S.to_dense(), Y) * (1 + 1e-10) + 1e-14
    assert _rel(X.to_dense(), Y) <= _rel(X1.to_dense(), Y) * (1 + 1e-10) + 1e-14


def test_als_info_semantics_and_edge_cases():
    shape = [3, 4, 3, 4]
    Yt = O.axpby(1.0, O.OTT.random(shape, 3, 41), 0.5, O.OTT.random(shape, 4, 42))
    # max_iters = 0: the right-orthogonalized start, nothing run.
    X0, it, ch, cv = O.round_als(Yt, 0.0, 3, 0, 1e-12)
    S = O.round(Yt, 0.0, 3, gauge=1)
    assert it == 0 and ch == 0.0 and not cv
    assert _rel(X0.to_dense(), S.to_dense()) <= 1e-12
    # The first sweep has no baseline: change 1.0.
    X1, it, ch, cv = O.round_als(Yt, 0.0, 3, 1, 1e-12)
    assert it == 1 and ch == 1.0 and not cv
    # Recovery at a cap above the true ranks converges at the second sweep.
    X, it, ch, cv = O.round_als(Yt, 0.0, 7, 20, 1e-12)
    assert cv and it == 2 and ch <= 1e-12
    assert _rel(X.to_dense(), Yt.to_dense()) <= 1e-13
    # d == 1 reduces to the SVD round with {0, 0, true}.
    v = O.OTT.from_cores([np.arange(1.0, 6.0).reshape(1, 5, 1)])
    X, it, ch, cv = O.round_als(v, 0.0, 1, 5, 1e-12)
    assert (it, ch, cv) == (0, 0.0, True)
    assert np.allclose(X.to_dense(), np.arange(1.0, 6.0))


def test_als_matvec_and_matmat_known_answers():
    # tests/core/test_tt_matvec_round.cpp / test_tt_matrix_round.cpp:325-471:
    # at a cap of the product's exact rank ALS recovers A x (A B) from a random
    # cold start; a tight cap stays within 5x of the SVD error; a warm start
    # from the SVD answer converges in a few sweeps.
    A = O.OTT.random_matrix([2, 3, 2, 3], [3, 2, 2, 2], 2, 7)
    x = O.OTT.random([3, 2, 2, 2], 2, 8)
    Y = O.matvec(A, x)
    Yd = Y.to_dense()
    cap = max(O.round(Y, 1e-13).ranks)
    X, it, ch, cv = O.matvec_round_als(A, x, 0.0, cap, 20, 1e-12, seed=0x3030)
    assert _rel(X.to_dense(), Yd) <= 1e-10
    e_svd = _rel(O.round(Y, 0.0, 2).to_dense(), Yd)
    X2 = O.matvec_round_als(A, x, 0.0, 2, 20, 1e-12, seed=5)[0]
    assert _rel(X2.to_dense(), Yd) <= 5.0 * e_svd + 1e-12
    # From the SVD answer each sweep is a least-squares step: never worse.
    W = O.round(Y, 0.0, 2)
    Xw = O.matvec_round_als(A, x, 0.0, 2, 3, 0.0, warm=W)[0]
    assert _rel(Xw.to_dense(), Yd) <= e_svd * (1 + 1e-10)
    # From the converged answer itself the change is at rounding level at once.
    F = O.matvec_round_als(A, x, 0.0, 2, 400, 1e-14, warm=W)[0]
    Xf, itf, chf, cvf = O.matvec_round_als(A, x, 0.0, 2, 20, 1e-10, warm=F)
    assert cvf and itf <= 3
    B = O.OTT.random_matrix([3, 2, 2, 2], [2, 2, 3, 2], 2, 9)
    C = O.matmat(A, B)
    Cd = C.to_dense()
    capc = max(O.round(C, 1e-13).ranks)
    Xc, it, ch, cv = O.matmat_round_als(A, B, 0.0, capc, 20, 1e-12, seed=0x4040)
    assert Xc.m == C.m and Xc.n == C.n
    assert _rel(Xc.to_dense(), Cd) <= 1e-10
    # eps path: the SVD answer is the warm start, so never worse than it.
    Ce = O.round(C, 1e-2)
    Xe = O.matmat_round_als(A, B, 1e-2, 0, 20, 1e-12)[0]
    assert _rel(Xe.to_dense(), Cd) <= _rel(Ce.to_dense(), Cd) * (1 + 1e-10)


def test_als_gauge_invariant_warm_start():
    # The iterates depend on the interface subspaces only: a warm start under
    # an arbitrary invertible gauge gives the same tensor.
    shape = [3, 4, 3, 4]
    Yt = O.axpby(1.0, O.OTT.random(shape, 3, 51), 0.2, O.OTT.random(shape, 4, 52))
    W = O.round(Yt, 0.0, 3)
    cs = [c.copy() for c in W.cores()]
    rng = np.random.default_rng(3)
    for k in range(len(cs) - 1):
        r = cs[k].shape[2]
        G = rng.standard_normal((r, r)) + 3 * np.eye(r)
        cs[k] = np.einsum("aib,bc->aic", cs[k], G)
        cs[k + 1] = np.einsum("cb,bid->cid", np.linalg.inv(G), cs[k + 1])
    W2 = O.OTT.from_cores(cs)
    Xa = O.round_als(Yt, 0.0, 3, 4, 0.0, warm=W)[0]
    Xb = O.round_als(Yt, 0.0, 3, 4, 0.0, warm=W2)[0]
    assert _rel(Xa.to_dense(), Xb.to_dense()) <= 1e-11


def test_frob_norm_apply_vs_dense():
    # include/core/algebra/tt_matrix_frob_norm_apply.hpp: ||A x|| and ||A B||
    # against the dense products.
    A = O.OTT.random_matrix([2, 3, 2], [3, 2, 2], 3, 17)
    x = O.OTT.random([3, 2, 2], 4, 18)
    B = O.OTT.random_matrix([3, 2, 2], [2, 2, 3], 2, 19)

    def dense_mat(M):
        rows, cols = int(np.prod(M.m)), int(np.prod(M.n))
        T = M.to_dense().reshape([v for pair in zip(M.m, M.n) for v in pair])
        d = len(M.m)
        return T.transpose(list(range(0, 2 * d, 2)) + list(range(1, 2 * d, 2))).reshape(rows, cols)

    Ad = dense_mat(A)
    want = np.linalg.norm(Ad @ x.to_dense())
    assert abs(O.frob_norm_apply(A, x) - want) <= 1e-13 * want
    want = np.linalg.norm(Ad @ dense_mat(B))
    assert abs(O.frob_norm_apply(A, B) - want) <= 1e-13 * want
