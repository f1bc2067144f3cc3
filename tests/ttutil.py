"""Shared helpers of the test suite: seeded inputs of the benchmark
configurations (built by the oracle's restatement of the reference
generators) and error measures."""
import numpy as np

import oracle as O


def dense_from_cores(cores):
    if not cores:
        return np.ones(1)
    T = cores[0].reshape(-1, cores[0].shape[-1])
    for c in cores[1:]:
        T = (T @ c.reshape(c.shape[0], -1)).reshape(-1, c.shape[-1])
    return T.ravel()


def cfg1_input():
    """add(add(a, b), a) with a = random([4]^10, 8, 1), b = random([4]^10, 8, 2)."""
    a = O.OTT.random([4] * 10, 8, 1)
    b = O.OTT.random([4] * 10, 8, 2)
    return O.add(O.add(a, b), a), a, b


def cfg2_input(d=20, rA=8, rx=32):
    """A = random([4]^d, [4]^d, rA, 3) with core k scaled by 1/sqrt(r_k m_k n_k)
    (r_k the left rank), x = random([4]^d, rx, 4)."""
    A = O.OTT.random_matrix([4] * d, [4] * d, rA, 3)
    cores = A.cores()
    cores = [c / np.sqrt(c.shape[0] * 16.0) for c in cores]
    A = O.OTT.from_cores(cores, [4] * d, [4] * d)
    x = O.OTT.random([4] * d, rx, 4)
    return A, x


def rel_err(cores_a, ref: "O.OTT"):
    """||a - ref|| / ||ref|| with both norms taken as ||core0|| after the
    oracle right-orthogonalizes the exact difference TT (no squaring, so
    it resolves ~1e-15)."""
    a = O.OTT.from_cores(cores_a, ref.m, ref.n)
    diff = O.axpby(1.0, a, -1.0, ref)
    O.right_orthogonalize(diff)
    num = np.linalg.norm(diff.cores()[0])
    r = ref.copy()
    O.right_orthogonalize(r)
    den = np.linalg.norm(r.cores()[0])
    return num / den if den > 0 else num
