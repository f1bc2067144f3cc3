"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle on the
same seeded inputs.  Bar: relative Frobenius error <= 1e-12 and identical
truncated ranks (the inputs here have well separated singular values at
every cut)."""
import numpy as np
import pytest

import oracle as O
import tensor_train_amd as T
from tests.ttutil import cfg1_input, cfg2_input, dense_from_cores, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


def to_dev(o: O.OTT):
    if o.m is not None:
        cores = [c.reshape(c.shape[0], m, n, c.shape[2]) for c, m, n in zip(o.cores(), o.m, o.n)]
        return T.TensorTrainMatrix.from_cores(cores)
    return T.TensorTrain.from_cores(o.cores())


def dev_cores3(t):
    return [c.reshape(c.shape[0], -1, c.shape[-1]) for c in t.cores]


def left_orth_err(c):
    M = c.reshape(-1, c.shape[-1])
    return np.linalg.norm(M.T @ M - np.eye(M.shape[1]))


def right_orth_err(c):
    M = c.reshape(c.shape[0], -1)
    return np.linalg.norm(M @ M.T - np.eye(M.shape[0]))


@pytest.mark.parametrize("gauge", [0, 1])
def test_round_config1(gauge):
    c, a, b = cfg1_input()
    ref = O.round(c, 1e-10, 0, gauge)
    y = T.round(to_dev(c), eps=1e-10, gauge=gauge)
    assert y.ranks == ref.ranks == [1, 4, 16, 16, 16, 16, 16, 16, 16, 4, 1]
    assert rel_err(dev_cores3(y), ref) <= TOL
    dense = 2 * a.to_dense() + b.to_dense()
    assert np.linalg.norm(y.to_dense() - dense) / np.linalg.norm(dense) <= 1e-10
    cores = y.cores
    if gauge == 1:
        assert all(right_orth_err(cores[k]) <= 1e-12 for k in range(1, len(cores)))
    else:
        assert all(left_orth_err(cores[k]) <= 1e-12 for k in range(len(cores) - 1))


CASES = [
    # shape, max_rank, seed, eps, cap
    ([3, 4, 5, 3], 6, 42, 1e-10, 0),
    ([2, 3, 4, 2, 3], 6, 123, 1e-6, 0),
    ([2, 3, 4, 2], 8, 7, 1e-2, 0),
    ([2, 3, 4, 2], 8, 7, 0.0, 3),
    ([1, 5, 1, 3, 1], 4, 5, 1e-12, 0),
    ([7], 1, 3, 1e-10, 0),
    ([6, 5], 5, 9, 1e-10, 0),
    ([4] * 6, 40, 11, 1e-8, 0),
    ([8, 2, 8, 2, 8], 30, 17, 1e-9, 0),
    ([2] * 12, 16, 19, 1e-10, 0),
]


@pytest.mark.parametrize("gauge", [0, 1])
@pytest.mark.parametrize("shape,mr,seed,eps,cap", CASES)
def test_round_random(shape, mr, seed, eps, cap, gauge):
    a = O.OTT.random(shape, mr, seed)
    b = O.OTT.random(shape, max(1, mr // 2), seed + 1000)
    c = O.add(a, O.add(O.scale(a, 0.5), b))
    ref = O.round(c, eps, cap, gauge)
    y = T.round(to_dev(c), eps=eps, max_rank=cap, gauge=gauge)
    assert y.ranks == ref.ranks
    assert rel_err(dev_cores3(y), ref) <= TOL


def test_round_zero_and_d1():
    z = O.OTT.from_cores([np.zeros((1, 3, 1)), np.zeros((1, 4, 1)), np.zeros((1, 2, 1))])
    y = T.round(to_dev(z), eps=1e-10)
    assert y.ranks == [1, 1, 1, 1]
    assert np.all(y.to_dense() == 0)
    one = O.OTT.from_cores([np.arange(1.0, 8.0).reshape(1, 7, 1)])
    y1 = T.round(to_dev(one), eps=1e-10)
    assert np.array_equal(y1.to_dense(), np.arange(1.0, 8.0))


def test_right_orthogonalize():
    a = O.OTT.random([3, 4, 5, 3, 2], 7, 99)
    dense = a.to_dense()
    ya = to_dev(a)
    T.right_orthogonalize(ya)
    ref = a.copy()
    O.right_orthogonalize(ref)
    assert ya.ranks == ref.ranks
    cores = ya.cores
    for k in range(1, len(cores)):
        assert right_orth_err(cores[k]) <= 1e-12
    assert np.linalg.norm(ya.to_dense() - dense) / np.linalg.norm(dense) <= 1e-12
    assert abs(np.linalg.norm(cores[0]) - np.linalg.norm(dense)) <= 1e-13 * np.linalg.norm(dense)


def test_left_orthogonalize():
    a = O.OTT.random([3, 4, 5, 3, 2], 7, 98)
    dense = a.to_dense()
    ya = to_dev(a)
    T.left_orthogonalize(ya)
    ref = a.copy()
    O.left_orthogonalize(ref)
    assert ya.ranks == ref.ranks
    cores = ya.cores
    for k in range(len(cores) - 1):
        assert left_orth_err(cores[k]) <= 1e-12
    assert np.linalg.norm(ya.to_dense() - dense) / np.linalg.norm(dense) <= 1e-12


def test_matvec_and_matmat_raw():
    A = O.OTT.random_matrix([2, 3, 2], [3, 2, 4], 3, 5)
    x = O.OTT.random([3, 2, 4], 2, 6)
    B = O.OTT.random_matrix([3, 2, 4], [2, 2, 3], 2, 7)
    yref = O.matvec(A, x)
    y = T.matvec(to_dev(A), to_dev(x))
    assert y.ranks == yref.ranks
    assert np.max(np.abs(y.to_dense() - yref.to_dense())) <= 1e-13 * np.max(np.abs(yref.to_dense()))
    Cref = O.matmat(A, B)
    C = T.matmat(to_dev(A), to_dev(B))
    assert C.ranks == Cref.ranks
    got = dense_from_cores(dev_cores3(C))
    assert np.max(np.abs(got - Cref.to_dense())) <= 1e-13 * np.max(np.abs(Cref.to_dense()))


@pytest.mark.parametrize("method", [0, 1])
def test_matvec_round_small(method):
    A, x = cfg2_input(d=6, rA=8, rx=32)
    ref = O.matvec_round(A, x, 1e-8, 0, method)
    y = T.matvec_round(to_dev(A), to_dev(x), eps=1e-8, method=["svd", "svd_streaming"][method])
    assert y.ranks == ref.ranks
    assert rel_err(dev_cores3(y), ref) <= TOL


@pytest.mark.parametrize("method", [0, 2])
def test_matmat_round_small(method):
    A = O.OTT.random_matrix([2, 3, 2, 2], [3, 2, 2, 3], 4, 21)
    B = O.OTT.random_matrix([3, 2, 2, 3], [2, 2, 3, 2], 3, 22)
    ref = O.matmat_round(A, B, 1e-9, 0, method)
    C = T.matmat_round(to_dev(A), to_dev(B), eps=1e-9, method=["svd", "x", "svd_naive"][method])
    assert C.ranks == ref.ranks
    assert rel_err(dev_cores3(C), ref) <= TOL


def test_inner_norm():
    a = O.OTT.random([3, 4, 5, 2], 6, 31)
    b = O.OTT.random([3, 4, 5, 2], 5, 32)
    da, db = to_dev(a), to_dev(b)
    ip = T.inner(da, db)
    ref = O.inner(a, b)
    assert abs(ip - ref) <= 1e-12 * abs(a.to_dense()) @ abs(b.to_dense())
    assert abs(T.norm(da) - O.norm(a)) <= 1e-13 * O.norm(a)


def test_axpby_scale():
    a = O.OTT.random([2, 3, 2], 4, 0)
    b = O.OTT.random([2, 3, 2], 4, 1)
    ref = O.axpby(2.5, a, -0.7, b)
    got = T.axpby(2.5, to_dev(a), -0.7, to_dev(b))
    assert got.ranks == ref.ranks
    assert np.max(np.abs(got.to_dense() - ref.to_dense())) <= 1e-14 * np.max(np.abs(ref.to_dense()))
    s = T.scale(to_dev(a), -3.0)
    assert np.allclose(s.to_dense(), -3.0 * a.to_dense(), rtol=0, atol=1e-14 * np.max(np.abs(a.to_dense())))
    one_a = O.OTT.from_cores([np.arange(5.0).reshape(1, 5, 1)])
    one_b = O.OTT.from_cores([np.ones((1, 5, 1))])
    got1 = T.axpby(2.0, to_dev(one_a), 3.0, to_dev(one_b))
    assert np.array_equal(got1.to_dense(), 2.0 * np.arange(5.0) + 3.0)


def test_method_errors():
    a = to_dev(O.OTT.random([2, 3, 2], 2, 0))
    with pytest.raises(T.TTAError) as e:
        T.round(a, method="svd_streaming")
    assert e.value.code == 3
    A = to_dev(O.OTT.random_matrix([2, 2], [2, 2], 2, 1))
    x = to_dev(O.OTT.random([2, 2], 2, 2))
    with pytest.raises(T.TTAError):
        T.matvec_round(A, x, method="svd_naive")
    with pytest.raises(T.TTAError):
        T.matmat_round(A, A, method="svd_streaming")
    bad = to_dev(O.OTT.random([3, 2], 2, 3))
    with pytest.raises(T.TTAError) as e2:
        T.matvec(A, bad)
    assert e2.value.code == 2
