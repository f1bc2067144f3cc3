"""frob_norm_apply on the GPU (include/core/algebra/tt_matrix_frob_norm_apply.hpp)
against the oracle: relative difference <= 1e-12 (a norm, not a difference of
nearly equal tensors, so the squaring in the environment sweep costs nothing)."""
import numpy as np
import pytest

import oracle as O
import tensor_train_amd as T
from tests.test_gpu_als import to_dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rs,cs,ra,rx", [([2, 3, 2], [3, 2, 2], 3, 4), ([4] * 8, [4] * 8, 4, 16),
                                         ([4] * 20, [4] * 20, 8, 32)])
def test_frob_norm_apply_matvec(rs, cs, ra, rx):
    A = O.OTT.random_matrix(rs, cs, ra, 5)
    x = O.OTT.random(cs, rx, 6)
    ref = O.frob_norm_apply(A, x)
    got = T.frob_norm_apply(to_dev(A), to_dev(x))
    assert abs(got - ref) <= 1e-12 * ref
    assert abs(got - T.norm(T.matvec(to_dev(A), to_dev(x)))) <= 1e-12 * ref


def test_frob_norm_apply_matmat():
    A = O.OTT.random_matrix([2, 3, 2, 3], [3, 2, 2, 2], 3, 7)
    B = O.OTT.random_matrix([3, 2, 2, 2], [2, 2, 3, 2], 4, 8)
    ref = O.frob_norm_apply(A, B)
    got = T.frob_norm_apply(to_dev(A), to_dev(B))
    assert abs(got - ref) <= 1e-12 * ref


def test_frob_norm_apply_errors():
    A = to_dev(O.OTT.random_matrix([2, 3], [3, 2], 2, 1))
    x = to_dev(O.OTT.random([2, 2], 2, 2))
    with pytest.raises(T.TTAError, match="mode mismatch"):
        T.frob_norm_apply(A, x)
