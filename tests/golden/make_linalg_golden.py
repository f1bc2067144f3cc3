"""Writes tests/golden/linalg_golden.npz: thin QR and SVD of small seeded
matrices computed by numpy's LAPACK (dgeqrf + dorgqr, dgesdd) in this
container.  These are library-routine vectors, not reference outputs: they
pin the oracle's Householder QR (LAPACK sign convention, shared by Eigen's
HouseholderQR) and the singular values of its Jacobi SVD.

Run:  python tests/golden/make_linalg_golden.py
"""
import os

import numpy as np

SHAPES = [(7, 3), (3, 7), (5, 5), (40, 12), (12, 40), (96, 24), (1, 5), (5, 1), (64, 64)]


def main():
    rng = np.random.default_rng(20240601)
    out = {}
    for idx, (m, n) in enumerate(SHAPES):
        M = rng.standard_normal((m, n))
        Q, R = np.linalg.qr(M)
        s = np.linalg.svd(M, compute_uv=False)
        out[f"M{idx}"] = M
        out[f"Q{idx}"] = Q
        out[f"R{idx}"] = R
        out[f"s{idx}"] = s
    # Rank-deficient 24 x 24 of rank 16 (as the padded cores of add(add(a,b),a)).
    A = rng.standard_normal((24, 16)) @ rng.standard_normal((16, 24))
    out["Mdef"] = A
    out["sdef"] = np.linalg.svd(A, compute_uv=False)
    # Graded spectrum 10^-j over 30 singular values.
    U, _ = np.linalg.qr(rng.standard_normal((40, 30)))
    V, _ = np.linalg.qr(rng.standard_normal((30, 30)))
    sg = 10.0 ** (-0.5 * np.arange(30))
    out["Mgrad"] = (U * sg) @ V.T
    out["sgrad"] = sg
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "linalg_golden.npz")
    np.savez_compressed(path, nshapes=len(SHAPES), **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
