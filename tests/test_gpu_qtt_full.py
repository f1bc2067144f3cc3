"""The large single-instance QTT workload at its full size: d = 50, n = 2,
x of rank 256 (Gaussian cores, left-orthogonalized, bond spectrum rescaled to
0.9^j), MPO of rank 16 (N(0, 1/(4 r))), matvec_round at eps = 1e-10 with
product ranks up to 4096 (8192 x 4096 QR factorizations and 4096-column
SVDs).  The oracle cannot round this on the CPU in test time, so the result
is checked by properties that hold for the rounding's output y_hat = P y
(P the orthogonal projector onto the kept left frame):
  * ranks within the structural bounds min(2^k, 2^(50-k), 16 * 256);
  * <y, y_hat> = ||y_hat||^2 (y_hat is a projection of y);
  * ||y||^2 - ||y_hat||^2 = ||y - y_hat||^2 >= 0 and tiny;
  * sampled entries: the root mean square of y_hat(i) - y(i) over random
    multi-indices is within 10 eps of the root mean square of y(i);
  * interior cores have orthogonal columns.
The same construction at reduced ranks is compared with the oracle directly
in test_gpu_qtt.py."""
import time

import numpy as np
import pytest

import tensor_train_amd as T

pytestmark = pytest.mark.gpu

D = 50


def cfg5_inputs(rx=256, ra=16, seed_x=11, seed_a=12):
    x = T.random([2] * D, rx, seed_x)
    T.left_orthogonalize(x)
    cores = x.cores
    for k in range(D - 1):
        cores[k] = cores[k] * (0.9 ** np.arange(cores[k].shape[2]))[None, None, :]
    x = T.TensorTrain.from_cores(cores)
    A = T.random_matrix([2] * D, [2] * D, ra, seed_a)
    rk = A.ranks
    for k in range(D):
        T.scale_core(A, k, 1.0 / np.sqrt(rk[k] * 4.0))
    return A, x


def _entry(t, idx):
    """t(idx) as inner(t, e_idx) with the one-hot rank-1 TT e_idx."""
    cores = []
    for i in idx:
        c = np.zeros((1, 2, 1))
        c[0, i, 0] = 1.0
        cores.append(c)
    return T.inner(t, T.TensorTrain.from_cores(cores))


def test_qtt_full_size_properties():
    eps = 1e-10
    A, x = cfg5_inputs()
    ctx = T.Context.default(0)
    ctx.synchronize()
    t0 = time.perf_counter()
    yh = T.matvec_round(A, x, eps=eps)
    ctx.synchronize()
    print(f"matvec_round d=50 x rank 256, MPO rank 16: {time.perf_counter() - t0:.1f} s, "
          f"max rank {max(yh.ranks)}", flush=True)
    ry = [a * b for a, b in zip(A.ranks, x.ranks)]
    for k, r in enumerate(yh.ranks):
        assert 1 <= r <= min(2 ** k, 2 ** (D - k), ry[k])
    assert max(ry) == 4096 and max(yh.ranks) > 640
    y = T.matvec(A, x)
    ny2 = T.inner(y, y)
    nh2 = T.inner(yh, yh)
    ip = T.inner(y, yh)
    print(f"||y||^2 {ny2:.6e}, ||yh||^2 {nh2:.6e}, <y,yh> {ip:.6e}", flush=True)
    assert abs(ip - nh2) <= 1e-12 * ny2
    assert -1e-12 * ny2 <= ny2 - nh2 <= 1e-12 * ny2
    rng = np.random.default_rng(5)
    ys, dif = [], []
    for _ in range(200):
        idx = rng.integers(0, 2, D)
        a, b = _entry(y, idx), _entry(yh, idx)
        ys.append(a)
        dif.append(b - a)
    rms_y = np.sqrt(np.mean(np.square(ys)))
    rms_d = np.sqrt(np.mean(np.square(dif)))
    print(f"sampled entries: rms y {rms_y:.3e}, rms (yh - y) {rms_d:.3e}", flush=True)
    assert rms_y > 0 and rms_d <= 10 * eps * rms_y
    for k in (12, 25):
        M = yh.cores[k].reshape(-1, yh.ranks[k + 1])
        G = M.T @ M
        dg = np.sqrt(np.diag(G))
        off = G / np.outer(dg, dg) - np.eye(G.shape[0])
        assert np.abs(off).max() <= 1e-10
