"""GPU kernel-level pins: the device thin QR and thin SVD (qr_thin /
svd_thin of the reference) against LAPACK golden vectors and against the
oracle at the shapes the TT sweeps produce."""
import os

import numpy as np
import pytest

import oracle as O
import tensor_train_amd as T

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "linalg_golden.npz")


def test_qr_golden():
    g = np.load(GOLD)
    for i in range(int(g["nshapes"])):
        M = g[f"M{i}"]
        Q, R = T.qr_thin(M)
        assert np.abs(Q - g[f"Q{i}"]).max() <= 1e-13, (i, M.shape)
        assert np.abs(R - g[f"R{i}"]).max() <= 1e-13 * max(1.0, np.abs(M).max()), (i, M.shape)


# Includes shapes factored in one 16-column panel (no trailing update, or
# one for the columns of a wide matrix) and shapes with a short last panel.
@pytest.mark.parametrize("shape", [(1024, 256), (2048, 100), (300, 300), (96, 24), (640, 33), (17, 70), (4, 256),
                                   (16, 16), (64, 4), (256, 16), (200, 13), (10, 40), (64, 64), (100, 70)])
def test_qr_large_vs_oracle(shape):
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    M = rng.standard_normal(shape)
    Q, R = T.qr_thin(M)
    Qo, Ro = O.qr_thin(M)
    k = min(shape)
    assert np.abs(Q.T @ Q - np.eye(k)).max() <= 1e-13
    assert np.abs(Q @ R - M).max() <= 1e-13 * np.abs(M).max() * np.sqrt(shape[0])
    assert np.abs(np.tril(R, -1)).max() == 0.0
    # Same sign convention => same factors as the oracle's Householder.
    assert np.abs(Q - Qo).max() <= 1e-11
    assert np.abs(R - Ro).max() <= 1e-11 * np.abs(Ro).max()


def test_qr_rank_deficient():
    rng = np.random.default_rng(5)
    M = rng.standard_normal((400, 40)) @ rng.standard_normal((40, 120))
    Q, R = T.qr_thin(M)
    assert np.abs(Q.T @ Q - np.eye(120)).max() <= 1e-13
    assert np.abs(Q @ R - M).max() <= 1e-12 * np.abs(M).max()


def test_svd_golden():
    g = np.load(GOLD)
    mats = [(g[f"M{i}"], g[f"s{i}"]) for i in range(int(g["nshapes"]))]
    mats += [(g["Mdef"], g["sdef"]), (g["Mgrad"], g["sgrad"])]
    for M, s_ref in mats:
        U, s, Vt = T.svd_thin(M)
        assert np.abs(s - s_ref).max() <= 1e-13 * s_ref[0], M.shape
        assert np.abs(U.T @ U - np.eye(len(s))).max() <= 1e-13
        big = s > 1e-8 * s[0]
        assert np.abs((U[:, big] * s[big]) @ Vt[big] - M).max() <= 1e-12 * s_ref[0] + np.sum(s[~big])


@pytest.mark.parametrize("shape", [(1024, 256), (256, 256), (128, 300), (500, 64)])
def test_svd_large_vs_numpy(shape):
    rng = np.random.default_rng(shape[0] + 3 * shape[1])
    M = rng.standard_normal(shape)
    U, s, Vt = T.svd_thin(M)
    s_ref = np.linalg.svd(M, compute_uv=False)
    assert np.abs(s - s_ref).max() <= 1e-12 * s_ref[0]
    assert np.abs(U.T @ U - np.eye(len(s))).max() <= 1e-12
    assert np.abs((U * s) @ Vt - M).max() <= 1e-12 * s_ref[0]


def test_wave_sum_dpp_matches_bpermute():
    import ctypes

    ctx = T.Context.default()
    rng = np.random.default_rng(3)
    x = np.ascontiguousarray(rng.standard_normal(64 * 97) * 10.0 ** rng.integers(-5, 5, 64 * 97))
    bad = ctypes.c_int(-1)
    ctx.check(ctx.lib.tta_selfcheck_wave_sum(ctx.h, x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(x),
                                             ctypes.byref(bad)))
    assert bad.value == 0


@pytest.mark.parametrize("decades", [4, 12, 20])
def test_svd_graded_large_dynamic_range(decades):
    # Block Jacobi on spectra spanning many decades, plus an exactly
    # rank-deficient block: backward-stable accuracy
    # (absolute errors ~ u * s_max, as LAPACK and Eigen's BDCSVD give).
    rng = np.random.default_rng(decades)
    q = 256
    U0, _ = np.linalg.qr(rng.standard_normal((q, q)))
    V0, _ = np.linalg.qr(rng.standard_normal((q, q)))
    sv = 10.0 ** (-decades * np.arange(q) / (q - 1))
    sv[200:] = 0.0
    M = (U0 * sv) @ V0.T
    U, s, Vt = T.svd_thin(M)
    s_ref = np.linalg.svd(M, compute_uv=False)
    assert np.abs(s - s_ref).max() <= 1e-13
    assert np.abs(U.T @ U - np.eye(q)).max() <= 1e-12
    keep = s > 1e-12
    assert np.abs((U[:, keep] * s[keep]) @ Vt[keep] - M).max() <= 1e-13


@pytest.mark.parametrize("q", [48, 100])
def test_svd_negligible_columns_converge(q):
    # Columns whose squared norms underflow to zero (entries ~1e-170) are
    # negligible: the Jacobi sweep must stop rotating them instead of
    # applying angle-zero rotations forever, and the factorization stays
    # backward stable.
    rng = np.random.default_rng(q)
    M = rng.standard_normal((2 * q, q))
    M[:, 7] *= 1e-170
    M[:, q // 2] *= 1e-200
    ctx = T.Context.default()
    ctx.prof_enable(True)
    U, s, Vt = T.svd_thin(M)
    calls, sweeps = ctx.jacobi_sweeps()
    ctx.prof_enable(False)
    if q > 64:
        assert calls == 1 and sweeps <= 24  # the sweep cap is 30
    s_ref = np.linalg.svd(M, compute_uv=False)
    assert np.abs(s - s_ref).max() <= 1e-13 * s_ref[0]
    assert np.abs(U[:, :q - 2].T @ U[:, :q - 2] - np.eye(q - 2)).max() <= 1e-12
    assert np.abs((U * s) @ Vt - M).max() <= 1e-13 * s_ref[0]
