"""Large single-instance QTT path (the d = 50, n = 2 workload at reduced
ranks): x with N(0,1) cores, left-orthogonalized and rescaled to a
geometric bond spectrum 0.9^j; an MPO with N(0, 1/(r*4)) cores; the device
matvec_round and round against the FP64 oracle on identical inputs (ranks
equal, relative Frobenius error <= 1e-12)."""
import numpy as np
import pytest

import oracle as O
import tensor_train_amd as T
from tests.ttutil import rel_err

pytestmark = pytest.mark.gpu

D = 50


def qtt_x(rank, seed):
    x = O.OTT.random([2] * D, rank, seed)
    O.left_orthogonalize(x)
    cores = x.cores()
    for k in range(D - 1):
        cores[k] = cores[k] * (0.9 ** np.arange(cores[k].shape[2]))[None, None, :]
    return O.OTT.from_cores(cores)


def qtt_mpo(rank, seed):
    A = O.OTT.random_matrix([2] * D, [2] * D, rank, seed)
    cores = [c / np.sqrt(c.shape[0] * 4.0) for c in A.cores()]
    return O.OTT.from_cores(cores, [2] * D, [2] * D)


def to_dev(o):
    if o.m is not None:
        cores = [c.reshape(c.shape[0], m, n, c.shape[2]) for c, m, n in zip(o.cores(), o.m, o.n)]
        return T.TensorTrainMatrix.from_cores(cores)
    return T.TensorTrain.from_cores(o.cores())


@pytest.mark.parametrize("xr,ar", [(32, 4), (48, 4)])
def test_qtt_matvec_round(xr, ar):
    x = qtt_x(xr, 11)
    A = qtt_mpo(ar, 12)
    ref = O.matvec_round(A, x, 1e-10, 0, 0)
    y = T.matvec_round(to_dev(A), to_dev(x), eps=1e-10)
    assert y.ranks == ref.ranks
    assert rel_err(y.cores, ref) <= 1e-12


def test_qtt_round_graded():
    x = qtt_x(96, 13)
    ref = O.round(x, 1e-10, 0, 0)
    y = T.round(to_dev(x), eps=1e-10)
    assert y.ranks == ref.ranks
    assert max(ref.ranks) < 96  # the 0.9^j spectrum is cut
    assert rel_err(y.cores, ref) <= 1e-12
