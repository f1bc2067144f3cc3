"""CPU tests of the FP64 oracle: pinned against numpy/LAPACK golden vectors,
hand-written known answers, closed forms, brute-force dense products and the
assertions of the reference's own hot-path tests (tests/core/test_tt_round.cpp,
test_tt_invariants.cpp, test_tt_ops.cpp, test_tt_matrix_apply.cpp,
test_tt_matrix_apply_round.cpp, test_tt_orthogonalize.cpp,
test_tt_validation.cpp of bdaszuta/tensor_train)."""
import os

import numpy as np
import pytest

import oracle as O
from tests.ttutil import cfg1_input, cfg2_input, dense_from_cores, rel_err

GOLD = os.path.join(os.path.dirname(__file__), "golden", "linalg_golden.npz")


# ---------------------------------------------------------------- linalg pins
def test_qr_matches_lapack_golden():
    g = np.load(GOLD)
    for i in range(int(g["nshapes"])):
        M = g[f"M{i}"]
        Q, R = O.qr_thin(M)
        scale = max(1.0, np.abs(M).max())
        # Householder QR with the dlarfg sign convention is unique.
        assert np.abs(Q - g[f"Q{i}"]).max() <= 1e-13, i
        assert np.abs(R - g[f"R{i}"]).max() <= 1e-13 * scale, i
        k = min(M.shape)
        assert np.abs(Q.T @ Q - np.eye(k)).max() <= 1e-14 * max(k, 1) * 4


def test_svd_matches_lapack_golden():
    g = np.load(GOLD)
    mats = [(g[f"M{i}"], g[f"s{i}"]) for i in range(int(g["nshapes"]))]
    mats += [(g["Mdef"], g["sdef"]), (g["Mgrad"], g["sgrad"])]
    for M, s_ref in mats:
        U, s, Vt = O.svd_thin(M)
        assert np.abs(s - s_ref).max() <= 1e-13 * s_ref[0]
        assert np.all(np.diff(s) <= 0)
        assert np.abs(U @ np.diag(s) @ Vt - M).max() <= 1e-13 * s_ref[0]
        k = len(s)
        # U is exactly orthonormal on the side the Jacobi rotations build.
        if M.shape[0] <= M.shape[1]:
            assert np.abs(U.T @ U - np.eye(k)).max() <= 1e-13
        else:
            assert np.abs(Vt @ Vt.T - np.eye(k)).max() <= 1e-13


def test_svd_graded_spectrum_relative_accuracy():
    # The stored matrix carries ~1e-16 absolute rounding from its formation, so
    # sigma_j = 10^(-j/2) is fixed to ~1e-16/sigma_j relative; check where that
    # is below 1e-9, and everywhere to 1e-14 absolute against LAPACK.
    g = np.load(GOLD)
    U, s, Vt = O.svd_thin(g["Mgrad"])
    rel = np.abs(s - g["sgrad"]) / g["sgrad"]
    assert rel[:12].max() <= 1e-9
    assert np.abs(s - np.linalg.svd(g["Mgrad"], compute_uv=False)).max() <= 1e-14


@pytest.mark.parametrize("s,delta,cap,want", [
    ([3.0, 2.0, 1.0], 0.0, 0, 3),
    ([3.0, 2.0, 1.0], 1.0, 0, 2),         # tail 1 <= 1 (equality kept), 1+4 > 1 stops
    ([3.0, 2.0, 1.0], np.sqrt(5.0), 0, 1),
    ([3.0, 2.0, 1.0], 100.0, 0, 1),       # never below 1
    ([0.0, 0.0], 0.0, 0, 1),              # exact zeros dropped, floor 1
    ([3.0, 2.0, 1.0], 0.0, 2, 2),         # max_rank cap
    ([3.0, 0.0, 0.0], 0.0, 0, 1),
    ([], 0.5, 0, 1),
    ([5.0, 4.0, 3.0, 2.0], 1.9999999, 0, 4),
])
def test_truncate_eps_rank_known_answers(s, delta, cap, want):
    assert O.truncate_eps_rank(np.array(s, dtype=float), delta, cap) == want


# ---------------------------------------------------------------- generators
def test_random_rank_rule():
    assert O.OTT.random([4] * 10, 8, 1).ranks == [1, 4, 8, 8, 8, 8, 8, 8, 8, 8, 1]
    # Unlimited rank: the forward cap 2^k, and the last bond keeps 32 (r_d = 1 set last).
    assert O.OTT.random([2] * 6, 0, 1).ranks == [1, 2, 4, 8, 16, 32, 1]
    assert O.OTT.random([3, 4, 5], 100, 1).ranks == [1, 3, 12, 1]
    A = O.OTT.random_matrix([4] * 20, [4] * 20, 8, 3)
    assert A.ranks == [1] + [8] * 19 + [1]
    x = O.OTT.random([4] * 20, 32, 4)
    assert x.ranks == [1, 4, 16] + [32] * 17 + [1]


def test_random_is_seeded_normal():
    a = np.concatenate([c.ravel() for c in O.OTT.random([8] * 6, 32, 7).cores()])
    b = np.concatenate([c.ravel() for c in O.OTT.random([8] * 6, 32, 7).cores()])
    c = np.concatenate([c.ravel() for c in O.OTT.random([8] * 6, 32, 8).cores()])
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert abs(a.mean()) < 0.02 and abs(a.var() - 1.0) < 0.03


def test_product_generator_bit_identical():
    # The library's host generator (compiled by hipcc) draws the same values.
    import tensor_train_amd as T

    for shape, mr, seed in [([4] * 10, 8, 1), ([2, 3, 5, 7], 5, 99), ([8, 2, 4], 64, 12345)]:
        got = T.random_host(shape, mr, seed)
        ref = O.OTT.random(shape, mr, seed).cores()
        assert len(got) == len(ref)
        assert all(np.array_equal(g, r) for g, r in zip(got, ref))


# ---------------------------------------------------------------- algebra
def test_add_axpby_scale_dense():
    a = O.OTT.random([2, 3, 2], 4, 0)
    b = O.OTT.random([2, 3, 2], 4, 1)
    ab = O.axpby(2.5, a, -0.7, b)
    assert ab.ranks == [1] + [x + y for x, y in zip(a.ranks[1:-1], b.ranks[1:-1])] + [1] == [1, 4, 8, 1]
    np.testing.assert_allclose(ab.to_dense(), 2.5 * a.to_dense() - 0.7 * b.to_dense(), rtol=0, atol=1e-14)
    np.testing.assert_allclose(O.scale(a, 3.0).to_dense(), 3.0 * a.to_dense(), rtol=0, atol=1e-14)
    one = O.axpby(2.0, O.OTT.from_cores([np.arange(4.0).reshape(1, 4, 1)]), 1.0,
                  O.OTT.from_cores([np.ones((1, 4, 1))]))
    assert np.array_equal(one.to_dense(), 2 * np.arange(4.0) + 1)


def test_to_dense_known_answer():
    # tests/core/test_tt_core.cpp:133 -- outer product [1,2] x [10,20,30].
    t = O.OTT.from_cores([np.array([1.0, 2.0]).reshape(1, 2, 1), np.array([10.0, 20.0, 30.0]).reshape(1, 3, 1)])
    assert np.array_equal(t.to_dense(), [10, 20, 30, 20, 40, 60])


def test_inner_norm_vs_dense():
    a = O.OTT.random([3, 4, 5, 2], 6, 31)
    b = O.OTT.random([3, 4, 5, 2], 5, 32)
    assert abs(O.inner(a, b) - a.to_dense() @ b.to_dense()) <= 1e-12 * np.abs(a.to_dense()) @ np.abs(b.to_dense())
    assert abs(O.norm(a) - np.linalg.norm(a.to_dense())) <= 1e-13 * np.linalg.norm(a.to_dense())


def _dense_matrix(A):
    """Dense (prod m) x (prod n) matrix of a TT-matrix."""
    cores = [c.reshape(c.shape[0], m, n, c.shape[2]) for c, m, n in zip(A.cores(), A.m, A.n)]
    T = np.ones((1, 1, 1))
    for c in cores:
        T = np.einsum("ijr,rmns->imjns", T, c).reshape(T.shape[0] * c.shape[1], T.shape[1] * c.shape[2], c.shape[3])
    return T[:, :, 0]


def test_matvec_matmat_vs_brute_force():
    # tests/core/test_tt_matrix_apply.cpp:76-105 -- A x and A B vs dense.
    A = O.OTT.random_matrix([2, 3, 2], [3, 2, 4], 3, 5)
    x = O.OTT.random([3, 2, 4], 2, 6)
    B = O.OTT.random_matrix([3, 2, 4], [2, 2, 3], 2, 7)
    y = O.matvec(A, x)
    assert y.ranks == [a * b for a, b in zip(A.ranks, x.ranks)]
    np.testing.assert_allclose(y.to_dense(), _dense_matrix(A) @ x.to_dense(), rtol=0, atol=1e-12)
    C = O.matmat(A, B)
    np.testing.assert_allclose(_dense_matrix(C), _dense_matrix(A) @ _dense_matrix(B), rtol=0, atol=1e-12)


def test_identity_matvec():
    # I x = x (tests/core/test_tt_matrix_apply.cpp:59-74).
    I = O.OTT.from_cores([np.eye(n).reshape(1, n * n, 1) for n in (2, 3, 4)], [2, 3, 4], [2, 3, 4])
    x = O.OTT.random([2, 3, 4], 3, 1)
    np.testing.assert_allclose(O.matvec(I, x).to_dense(), x.to_dense(), rtol=0, atol=1e-15)


# ---------------------------------------------------------------- gauge
def test_right_orthogonalize_properties():
    # tests/core/test_tt_orthogonalize.cpp:23-48, test_tt_invariants.cpp:156-175.
    a = O.OTT.random([3, 4, 5, 3], 6, 99)
    dense = a.to_dense()
    O.right_orthogonalize(a)
    for c in a.cores()[1:]:
        M = c.reshape(c.shape[0], -1)
        assert np.linalg.norm(M @ M.T - np.eye(M.shape[0])) <= 1e-14 * M.shape[0]
    np.testing.assert_allclose(a.to_dense(), dense, rtol=0, atol=1e-12 * np.abs(dense).max())
    assert abs(np.linalg.norm(a.cores()[0]) - np.linalg.norm(dense)) <= 1e-14 * (np.linalg.norm(dense) + 1)
    # r_eff = min(n * r_right, r_left): the last core (6, 3, 1) shrinks to 3.
    assert a.ranks == [1, 3, 6, 3, 1]


def test_left_orthogonalize_properties():
    a = O.OTT.random([3, 4, 5, 3], 6, 98)
    dense = a.to_dense()
    O.left_orthogonalize(a)
    for c in a.cores()[:-1]:
        M = c.reshape(-1, c.shape[2])
        assert np.linalg.norm(M.T @ M - np.eye(M.shape[1])) <= 1e-14 * M.shape[1]
    np.testing.assert_allclose(a.to_dense(), dense, rtol=0, atol=1e-12 * np.abs(dense).max())


# ---------------------------------------------------------------- round
@pytest.mark.parametrize("gauge", [0, 1])
def test_config1_closed_form(gauge):
    c, a, b = cfg1_input()
    assert c.ranks == [1, 12] + [24] * 8 + [1]
    r = O.round(c, 1e-10, 0, gauge)
    assert r.ranks == [1, 4, 16, 16, 16, 16, 16, 16, 16, 4, 1]
    dense = 2 * a.to_dense() + b.to_dense()
    assert np.linalg.norm(r.to_dense() - dense) / np.linalg.norm(dense) <= 1e-10


def test_round_right_canonical_properties():
    # round_right_canonical (tt_round.hpp:151-191): cores 1..d-1 are
    # right-orthogonal, the error bound holds, and the dense result equals
    # the gauge-none rounding (the truncated projection is unique when the
    # singular-value gaps at the cuts are clear).
    a = O.OTT.random([3, 4, 2, 5, 3], 6, 11)
    b = O.OTT.random([3, 4, 2, 5, 3], 6, 12)
    c = O.add(O.add(a, b), a)
    r1 = O.round(c, 1e-9, 0, 1)
    r0 = O.round(c, 1e-9, 0, 0)
    assert r1.ranks == r0.ranks
    for core in r1.cores()[1:]:
        M = core.reshape(core.shape[0], -1)
        assert np.abs(M @ M.T - np.eye(M.shape[0])).max() <= 1e-14 * 10
    dense = c.to_dense()
    assert np.linalg.norm(r1.to_dense() - dense) <= 1e-9 * np.linalg.norm(dense)
    assert np.linalg.norm(r1.to_dense() - r0.to_dense()) <= 1e-12 * np.linalg.norm(dense)
    # A real cut: the first core absorbs the norm, the rest stay orthogonal.
    x = O.OTT.random([2, 3, 4, 3, 2], 5, 21)
    rx = O.round(x, 0.05, 0, 1)
    assert np.linalg.norm(rx.to_dense() - x.to_dense()) <= 0.05 * np.linalg.norm(x.to_dense())
    assert abs(np.linalg.norm(rx.cores()[0]) - np.linalg.norm(rx.to_dense())) <= 1e-12 * np.linalg.norm(x.to_dense())


def _tt_svd(dense, shape, eps=1e-14):
    """Plain TT-SVD (numpy) of a dense tensor: the input builder of
    tests/core/test_tt_round.cpp:20-50."""
    d = len(shape)
    cores, r = [], 1
    M = dense.reshape(r * shape[0], -1)
    for k in range(d - 1):
        U, s, Vt = np.linalg.svd(M, full_matrices=False)
        rk = max(1, int(np.sum(s > eps * s[0])))
        cores.append(U[:, :rk].reshape(r, shape[k], rk))
        M = (s[:rk, None] * Vt[:rk]).reshape(rk * shape[k + 1], -1)
        r = rk
    cores.append(M.reshape(r, shape[-1], 1))
    return cores


def test_round_recovers_zero_padded_tt():
    # tests/core/test_tt_round.cpp:80-115: pad every bond by +5 zeros, round(1e-10)
    # recovers the tensor with ranks <= the original; max_rank=2 is honoured.
    rng = np.random.default_rng(13)
    shape = [4, 5, 6, 4]
    dense = rng.standard_normal(int(np.prod(shape)))
    cores = _tt_svd(dense, shape)
    orig = [1] + [c.shape[2] for c in cores]
    padded = []
    for k, c in enumerate(cores):
        rl = c.shape[0] + (5 if k > 0 else 0)
        rr = c.shape[2] + (5 if k < len(cores) - 1 else 0)
        p = np.zeros((rl, c.shape[1], rr))
        p[: c.shape[0], :, : c.shape[2]] = c
        padded.append(p)
    t = O.OTT.from_cores(padded)
    r = O.round(t, 1e-10)
    assert all(a <= b for a, b in zip(r.ranks, orig))
    assert np.linalg.norm(r.to_dense() - dense) / np.linalg.norm(dense) <= 1e-8
    r2 = O.round(t, 1e-10, 2)
    assert max(r2.ranks) <= 2


def test_round_invariants():
    # tests/core/test_tt_invariants.cpp
    z = O.OTT.from_cores([np.zeros((1, n, 1)) for n in (3, 4, 2)])
    rz = O.round(z, 1e-10)
    assert rz.ranks == [1, 1, 1, 1] and np.all(rz.to_dense() == 0)
    ones7 = O.OTT.from_cores([np.ones((1, 7, 1))])
    assert np.array_equal(O.round(ones7, 1e-10).to_dense(), np.ones(7))
    r = O.round(O.OTT.random([3, 4, 5, 3], 6, 42), 1e-10)
    for c in r.cores()[:-1]:  # left-orthogonal after the SVD sweep (:145-154)
        M = c.reshape(-1, c.shape[2])
        assert np.linalg.norm(M.T @ M - np.eye(M.shape[1])) <= 1e-14 * M.shape[1] * 10
    a = O.OTT.random([2, 3, 4, 2], 8, 7)
    na = O.norm(a)
    assert O.norm(O.round(a, 1e-2)) <= na * (1 + 1e-14)
    assert O.norm(O.round(a, 0.0, 3)) <= na * (1 + 1e-14)
    a = O.OTT.random([2, 3, 4], 4, 0)
    b = O.OTT.random([2, 3, 4], 4, 1)
    assert np.abs(O.round(O.add(a, b), 1e-10).to_dense() - a.to_dense() - b.to_dense()).max() <= 1e-13
    a = O.OTT.random([2, 3, 4, 2, 3], 6, 123)
    r1 = O.round(a, 1e-6)
    r2 = O.round(r1, 1e-6)
    assert np.abs(r1.to_dense() - r2.to_dense()).max() <= 1e-14 * (np.linalg.norm(r1.to_dense()) + 1)
    a = O.OTT.random([2, 3, 4, 2], 8, 99)
    ar = O.round(a, 0.01)
    assert np.linalg.norm(a.to_dense() - ar.to_dense()) <= 0.01 * np.linalg.norm(a.to_dense()) * 1.01


def test_round_add_self():
    # tests/core/test_tt_ops.cpp:160-191: round(add(A, A), 1e-12) = 2A, ranks <= ranks(A).
    a = O.OTT.random([3, 4, 3, 2], 5, 77)
    r = O.round(O.add(a, a), 1e-12)
    assert all(x <= y for x, y in zip(r.ranks, a.ranks))
    np.testing.assert_allclose(r.to_dense(), 2 * a.to_dense(), rtol=0, atol=1e-9)


def test_round_mode_size_one():
    shape = [1, 5, 1, 3, 1]
    buf = np.arange(1.0, 16.0)
    t = O.OTT.from_cores(_tt_svd(buf, shape))
    assert np.abs(O.round(t, 1e-10).to_dense() - buf).max() <= 2e-14 * 15


def test_config2_reduced_structure():
    A, x = cfg2_input(d=6)
    y = O.matvec(A, x)
    assert y.ranks == [1, 32, 128, 256, 256, 256, 1]
    yr = O.matvec_round(A, x, 1e-8, 0, 0)
    assert yr.ranks == [1, 4, 16, 64, 16, 4, 1]
    # svd (naive) == round(matvec) by construction; error bound vs the exact product.
    assert rel_err(yr.cores(), y) <= 1e-8


def test_streaming_vs_naive():
    # tests/core/test_tt_matrix_apply_round.cpp:100-129: within 10 eps ||y||.
    A = O.OTT.random_matrix([2, 3, 2, 3], [3, 2, 2, 2], 4, 11)
    x = O.OTT.random([3, 2, 2, 2], 5, 12)
    eps = 1e-6
    naive = O.matvec_round(A, x, eps, 0, 0)
    stream = O.matvec_round(A, x, eps, 0, 1)
    yd = O.matvec(A, x).to_dense()
    assert np.linalg.norm(stream.to_dense() - naive.to_dense()) <= 10 * eps * np.linalg.norm(yd)
    assert np.linalg.norm(stream.to_dense() - yd) <= 10 * eps * np.linalg.norm(yd)
    B = O.OTT.random_matrix([3, 2, 2, 2], [2, 2, 3, 2], 3, 13)
    mm = O.matmat_round(A, B, eps, 0, 0)
    mn = O.matmat_round(A, B, eps, 0, 2)
    exact = _dense_matrix(O.matmat(A, B))
    assert np.linalg.norm(_dense_matrix(mm) - exact) <= 10 * eps * np.linalg.norm(exact)
    assert np.linalg.norm(_dense_matrix(mn) - exact) <= 10 * eps * np.linalg.norm(exact)


def test_fused_core_identity_environment():
    # tests/core/test_tt_fused_apply_core.cpp:57-164 via the streaming path with
    # eps = 0: the first fused core with F = [[1]] equals the unfused product.
    A = O.OTT.random_matrix([2, 2], [3, 2], 3, 21)
    x = O.OTT.random([3, 2], 2, 22)
    s = O.matvec_round(A, x, 0.0, 0, 1)
    np.testing.assert_allclose(s.to_dense(), O.matvec(A, x).to_dense(), rtol=0, atol=1e-13)


def test_rel_err_resolves_tiny_differences():
    # The error measure (QR-based, no squaring) sees a 1e-14 perturbation.
    a = O.OTT.random([4] * 6, 8, 5)
    cores = a.cores()
    cores[2] = cores[2] * (1 + 1e-14)
    e = rel_err(cores, a)
    assert 5e-15 < e < 5e-14
    # An exact copy leaves only the rounding of the block QR of [a, -a].
    assert rel_err(a.cores(), a) <= 2e-15
    assert np.allclose(dense_from_cores(a.cores()), a.to_dense())


def _np_tt_svd(T, shape, eps, max_rank=0):
    """Independent TT-SVD with numpy's LAPACK SVD and the reference's rank
    rule (for pinning the oracle's from_dense)."""
    d = len(shape)
    flat = np.asarray(T, np.float64).ravel()
    nrm = np.linalg.norm(flat)
    delta = eps * nrm / np.sqrt(max(d - 1, 1))
    cores, r, C = [], 1, flat
    for k in range(d - 1):
        M = C.reshape(r * shape[k], -1)
        U, s, Vt = np.linalg.svd(M, full_matrices=False)
        rk = O.truncate_eps_rank(s, delta, max_rank)
        cores.append(U[:, :rk].reshape(r, shape[k], rk))
        C = (s[:rk, None] * Vt[:rk]).ravel()
        r = rk
    cores.append(C.reshape(r, shape[-1], 1))
    return cores


def test_from_dense_known_answers():
    # tests/core/test_tt_svd.cpp:24-104: rank-1 outer product recovered with
    # ranks 1; random tensor within 10 eps; zero tensor -> rank-1 zeros.
    u, v, w = np.array([1.0, 2, 3]), np.array([1.0, -1, 0.5, 0.25]), np.array([2.0, 0.5, -1, 1, 3])
    T = np.einsum("i,j,k->ijk", u, v, w)
    t = O.from_dense(T, [3, 4, 5], 1e-12)
    assert t.ranks == [1, 1, 1, 1]
    assert np.linalg.norm(t.to_dense().ravel() - T.ravel()) <= 1e-10 * np.linalg.norm(T)
    R = np.random.default_rng(42).standard_normal((4, 5, 6))
    t = O.from_dense(R, [4, 5, 6], 1e-8)
    assert np.linalg.norm(t.to_dense().ravel() - R.ravel()) <= 10 * 1e-8 * np.linalg.norm(R)
    z = O.from_dense(np.zeros(24), [2, 3, 4], 1e-10)
    assert z.ranks == [1, 1, 1, 1] and not np.any(z.to_dense())


@pytest.mark.parametrize("shape,rank,eps", [([3, 4, 2, 5], 4, 1e-12), ([4, 4, 4, 4, 4], 6, 1e-6), ([7, 2, 9], 3, 0.0)])
def test_from_dense_matches_numpy_tt_svd(shape, rank, eps):
    a = O.OTT.random(shape, rank, 17)
    D = a.to_dense().ravel()
    D = D + 1e-4 * np.random.default_rng(3).standard_normal(D.size)  # full rank, decaying
    t = O.from_dense(D, shape, eps)
    ref = _np_tt_svd(D, shape, eps)
    assert t.ranks == [1] + [c.shape[2] for c in ref]
    Dt = t.to_dense().ravel()
    Dr = dense_from_cores(ref).ravel()
    assert np.linalg.norm(Dt - Dr) <= 1e-12 * np.linalg.norm(D)
    assert np.linalg.norm(Dt - D) <= max(eps, 1e-13) * np.linalg.norm(D) * 1.0000001  # eps = 0: rounding only


def test_from_dense_qtt_and_matrix():
    # QTT of a linear function of the index has ranks 2 (from_dense.hpp:109-199).
    x = np.arange(64.0).reshape(8, 8)
    q = O.from_dense(x, [8, 8], 1e-12, method="qtt")
    assert q.ranks == [1, 2, 2, 2, 2, 2, 1]
    # The QTT index order is the binary digits of (i, j), most significant first.
    assert np.allclose(q.to_dense().ravel(), x.ravel(), rtol=0, atol=1e-12)
    M = np.random.default_rng(5).standard_normal((12, 20))
    tm = O.matrix_from_dense(M, [3, 4], [4, 5], 1e-13)
    # to_dense of a TT-matrix is (m0, n0, m1, n1); back to (m0 m1) x (n0 n1).
    Dm = tm.to_dense().reshape(3, 4, 4, 5).transpose(0, 2, 1, 3).reshape(12, 20)
    assert np.linalg.norm(Dm - M) <= 1e-12 * np.linalg.norm(M)


def test_soft_threshold_closed_form():
    # d = 2: the TT is one matrix, and soft thresholding is U max(s - tau, 0) V^T
    # of its LAPACK SVD; tau <= 0 and d = 1 return copies
    # (include/core/round/tt_soft_threshold.hpp:42-52).
    a = O.OTT.random([6, 7], 5, 3)
    D = a.to_dense().reshape(6, 7)
    U, s, Vt = np.linalg.svd(D, full_matrices=False)
    for tau in (0.4 * s[2], 0.9 * s[0], 1.1 * s[0]):
        r = O.soft_threshold(a, tau)
        ref = (U * np.maximum(s - tau, 0.0)) @ Vt
        assert r.ranks[1] == max(1, int(np.sum(s > tau)))
        assert np.abs(r.to_dense().reshape(6, 7) - ref).max() <= 1e-13 * max(1.0, np.abs(D).max())
    assert np.array_equal(O.soft_threshold(a, 0.0).to_dense(), a.to_dense())
    one = O.OTT.random([5], 1, 2)
    assert np.array_equal(O.soft_threshold(one, 3.0).to_dense(), one.to_dense())
    # Every step shrinks: the norm never grows.
    b = O.OTT.random([3, 4, 3, 2], 6, 9)
    assert O.norm(O.soft_threshold(b, 0.1)) <= O.norm(b) * (1 + 1e-14)


def test_round_without_truncation_reproduces_input():
    # eps = 0 on an MPO product (no singular value is cut): the rounded TT is
    # the input tensor in another gauge, to rounding.  The Jacobi SVD
    # accumulates its rotations in extended precision, so thousands of
    # rotations per bond leave no FP64 drift (with FP64 accumulation this
    # product drifts by ~1e-13 per bond at rank 256).
    A, x = cfg2_input(d=8)
    Y = O.matvec(A, x)
    for gauge in (0, 1):
        r = O.round(Y, 0.0, 0, gauge)
        assert r.ranks[1:-1] == [4, 16, 64, 256, 64, 16, 4]
        assert rel_err(r.cores(), Y) <= 1e-13


def test_round_batch_threads_equal_round():
    """The threaded batch entry (the CPU baseline of the batched workloads)
    returns exactly what round returns for each TT."""
    tts = []
    for t in range(6):
        u = O.OTT.random([3] * 6, 5 + t, 70 + t)
        tts.append(O.add(u, O.OTT.random([3] * 6, 4, 90 + t)))
    outs = O.round_batch(tts, 1e-8, 0, threads=3)
    for x, y in zip(tts, outs):
        ref = O.round(x, 1e-8, 0, 0)
        assert y.ranks == ref.ranks
        for a, b in zip(y.cores(), ref.cores()):
            assert np.array_equal(a, b)


def test_round_bf16_matches_torch():
    """The MLP restatement's bfloat16 rounding against torch's (nearest, ties
    to even), on random values, exact ties and specials."""
    import torch

    import oracle.mlp_ref as R

    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32) * np.float32(10.0) ** rng.integers(-30, 30, 10000),
                        np.array([1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8), 0.0, -0.0, np.inf, -np.inf,
                                  3.4e38, 1e-40], np.float32)]).astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = R.round_bf16(x)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
