"""soft_threshold on the GPU (include/core/round/tt_soft_threshold.hpp)
against the oracle on the same inputs: equal ranks, relative Frobenius
difference <= 1e-12 (the shrunk spectra have clear gaps at tau)."""
import numpy as np
import pytest

import oracle as O
import tensor_train_amd as T
from tests.ttutil import rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,rank,tau", [([6, 7], 5, 0.5), ([3, 4, 3, 2, 5], 6, 0.2), ([4] * 8, 16, 0.05),
                                            ([4] * 6, 12, 1e6)])
def test_soft_threshold_matches_oracle(shape, rank, tau):
    a = O.OTT.random(shape, rank, 31)
    ref = O.soft_threshold(a, tau)
    t = T.soft_threshold(T.TensorTrain.from_cores(a.cores()), tau)
    assert t.ranks == ref.ranks
    if np.linalg.norm(ref.to_dense()) == 0.0:
        assert not np.any(t.to_dense())
    else:
        assert rel_err(t.cores, ref) <= 1e-12


def test_soft_threshold_noop_copies():
    a = T.random([3, 4, 5], 4, 7)
    for tau in (0.0, -1.0):
        b = T.soft_threshold(a, tau)
        assert b.ranks == a.ranks and np.array_equal(b.to_dense(), a.to_dense())
