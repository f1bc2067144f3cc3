"""Dense to TT on the GPU (from_dense, include/core/factory/from_dense.hpp and
tt_svd.hpp) against the oracle on the same data: identical ranks and a
relative Frobenius difference <= 1e-12; the reference's own known answers
(tests/core/test_tt_svd.cpp); QTT and TT-matrix forms; error handling."""
import numpy as np
import pytest

import oracle as O
import tensor_train_amd as T
from tests.ttutil import rel_err

pytestmark = pytest.mark.gpu


def test_from_dense_known_answers():
    u, v, w = np.array([1.0, 2, 3]), np.array([1.0, -1, 0.5, 0.25]), np.array([2.0, 0.5, -1, 1, 3])
    X = np.einsum("i,j,k->ijk", u, v, w)
    t = T.from_dense(X, eps=1e-12)
    assert t.ranks == [1, 1, 1, 1]
    assert np.linalg.norm(t.to_dense().ravel() - X.ravel()) <= 1e-12 * np.linalg.norm(X)
    z = T.from_dense(np.zeros((2, 3, 4)), eps=1e-10)
    assert z.ranks == [1, 1, 1, 1] and not np.any(z.to_dense())
    bad = np.ones((2, 3))
    bad[1, 1] = np.nan
    with pytest.raises(T.TTAError):
        T.from_dense(bad)


@pytest.mark.parametrize("shape,rank,eps", [([3, 4, 2, 5], 4, 1e-12), ([4] * 8, 8, 1e-6), ([2] * 14, 12, 1e-9),
                                            ([16, 3, 40], 5, 0.0), ([64, 64], 20, 1e-3)])
def test_from_dense_matches_oracle(shape, rank, eps):
    a = O.OTT.random(shape, rank, 23)
    D = a.to_dense().ravel()
    D = D + 1e-5 * np.random.default_rng(len(shape)).standard_normal(D.size)
    ref = O.from_dense(D, shape, eps)
    t = T.from_dense(D, shape=shape, eps=eps)
    assert t.ranks == ref.ranks
    assert rel_err(t.cores, ref) <= 1e-12


def test_from_dense_qtt_and_matrix():
    x = np.sin(np.linspace(0, 3, 256)).reshape(16, 16)
    ref = O.from_dense(x, [16, 16], 1e-10, method="qtt")
    t = T.from_dense(x, eps=1e-10, method="qtt")
    assert t.ranks == ref.ranks and t.d == 8
    assert rel_err(t.cores, ref) <= 1e-12
    M = np.random.default_rng(8).standard_normal((24, 30))
    refm = O.matrix_from_dense(M, [4, 6], [5, 6], 1e-12)
    tm = T.matrix_from_dense(M, [4, 6], [5, 6], eps=1e-12)
    assert tm.ranks == refm.ranks
    assert rel_err([c.reshape(c.shape[0], -1, c.shape[-1]) for c in tm.cores], refm) <= 1e-12
    with pytest.raises(T.TTAError):
        T.from_dense(np.ones((3, 4)), eps=1e-10, method="qtt")  # 3 is not a power of 2
