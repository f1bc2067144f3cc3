"""GPU parity at the benchmark configurations' own shapes.

* config 4: a batch of 48 decaying-spectrum TTs (d = 32, n = 4, ranks
  min(4^k, 4^(32-k), 128)) -- batch >= 32 and 64 < q <= 128, so the SVD
  steps run the grouped block-Jacobi kernel;
* config 3: sum TTs add(u, add(scale(u, 0.5), v)) with stored ranks up to
  192, and two with 16 modes per core, whose QR factorizations have 2560 rows
  (the panel kernel for panels taller than 2048 rows);
* config 2 at its full size (d = 20, product ranks 256);
* a near-tie at the cut, where equal ranks are not required (the exemption
  for a relative singular-value gap <= 1e-3), and
* round / right_orthogonalize of a TT-matrix called directly.

Sampled TTs are compared with the oracle's round of the same device-built
input: identical ranks and relative Frobenius error <= 1e-12.  For the large
cases the error is the QR-based difference norm taken on the device
(||core 0|| of the right-orthogonalized difference TT; the device
right_orthogonalize is itself checked against the oracle in
test_gpu_parity.py)."""
import numpy as np
import pytest

import oracle as O
import tensor_train_amd as T
from tests.ttutil import cfg2_input, rel_err

pytestmark = pytest.mark.gpu


def dev_rel(a, ref_cores):
    """||a - ref|| / ||ref|| by QR on the device (no squaring)."""
    ref = T.TensorTrain.from_cores(ref_cores)
    dif = T.sub(a, ref)
    T.right_orthogonalize(dif)
    r = ref.copy()
    T.right_orthogonalize(r)
    return float(np.linalg.norm(dif.cores[0]) / np.linalg.norm(r.cores[0]))


def test_config4_shapes_batched_grouped_jacobi():
    count, eps = 48, 1e-6
    rng = np.random.default_rng(404)
    s = rng.uniform(0.05, 0.3, count)
    batch = T.build_decay_batch(count, 32, 4, 128, s, 7000 + np.arange(count))
    ctx = T.Context.default(0)
    ctx.prof_enable(True)
    outs = T.round_batched(batch, eps=eps)
    calls, sweeps = ctx.jacobi_sweeps()
    assert calls > 0 and sweeps > 0  # block Jacobi (q > 64) ran on the batch
    assert ctx.jacobi_unconverged() == 0
    ctx.prof_enable(False)
    lo = int(np.argmin(s))
    for i in (lo, int(np.argmax(s)), 17):
        assert max(batch[i].ranks) == 128  # SVD steps of 128 columns
        x = O.OTT.from_cores(batch[i].cores)
        ref = O.round(x, eps, 0, 0)
        assert outs[i].ranks == ref.ranks, (i, s[i])
        if i == lo:
            assert max(ref.ranks) > 64  # slow decay: most of the 128 are kept
        assert dev_rel(outs[i], ref.cores()) <= 1e-12


def test_config3_shapes_batched_ranks_to_192():
    eps = 1e-6
    rng = np.random.default_rng(303)
    count = 36
    d = rng.integers(8, 13, count)
    n = rng.choice([2, 4, 8], count)
    ru = rng.integers(8, 65, count)
    rv = rng.integers(8, 65, count)
    # Stored ranks 192 (n = 8 reaches them at the second bond) ...
    d[0], n[0], ru[0], rv[0] = 12, 8, 64, 64
    d[1], n[1], ru[1], rv[1] = 10, 4, 64, 64
    # ... and 16 modes with stored rank 160: QR panels of 16 * 160 = 2560 rows.
    d[2], n[2], ru[2], rv[2] = 5, 16, 48, 64
    d[3], n[3], ru[3], rv[3] = 6, 16, 40, 80
    batch = T.build_sum_batch(d, n, ru, rv, 1000 + np.arange(count), 500000 + np.arange(count))
    assert max(max(t.ranks) for t in batch) == 192
    outs = T.round_batched(batch, eps=eps)
    for i in (0, 1, 2, 3, 20):
        x = O.OTT.from_cores(batch[i].cores)
        ref = O.round(x, eps, 0, 0)
        assert outs[i].ranks == ref.ranks, i
        assert dev_rel(outs[i], ref.cores()) <= 1e-12
    # The same TTs one by one through the single-TT path.
    for i in (2, 3):
        single = T.round(batch[i], eps=eps)
        assert single.ranks == outs[i].ranks
        assert dev_rel(single, outs[i].cores) <= 1e-12


def test_config2_full_size():
    A, x = cfg2_input(d=20, rA=8, rx=32)
    ref = O.matvec_round(A, x, 1e-8, 0, 0)
    Ad = T.TensorTrainMatrix.from_cores(
        [c.reshape(c.shape[0], 4, 4, c.shape[2]) for c in A.cores()])
    y = T.matvec_round(Ad, T.TensorTrain.from_cores(x.cores()), eps=1e-8)
    assert y.ranks == ref.ranks
    assert max(ref.ranks) == 256
    assert dev_rel(y, ref.cores()) <= 1e-12


def _tie_tt(n, sig, seed):
    """d = 2: y = U diag(sig) V^T, U (n x r) and V (n x r) orthonormal."""
    rng = np.random.default_rng(seed)
    r = len(sig)
    U, _ = np.linalg.qr(rng.standard_normal((n, r)))
    V, _ = np.linalg.qr(rng.standard_normal((n, r)))
    return [U.reshape(1, n, r), (np.diag(sig) @ V.T).reshape(r, n, 1)]


def _cut_gap(sig, r):
    """Relative gap between the last kept and the first dropped singular value."""
    return 1.0 if r >= len(sig) else (sig[r - 1] - sig[r]) / sig[r - 1]


def test_near_tie_exemption():
    """sigma_13 = sigma_14 and delta^2 = 2 sigma_13^2 to the last bit: the rank
    rule's choice between 13 and 14 rests on rounding, and a cut at 14 falls
    inside the tie (relative gap 0 <= 1e-3), where the kept singular vector of
    the tied pair is not unique.  Exempt cuts need not give equal ranks or
    equal tensors; both results must still meet the error bound and agree to
    2 eps.  A cut with a gap above 1e-3 must match exactly."""
    t = 1e-7
    sig = np.concatenate([2.0 ** -np.arange(13), [t, t]])
    cores = _tie_tt(32, sig, 5)
    x = O.OTT.from_cores(cores)
    dense = x.to_dense()
    base = np.sqrt(2.0) * t / np.linalg.norm(sig)
    exempt_seen = False
    for f in (1.0, 0.85, 1.1, 0.6):
        eps = f * base
        ref = O.round(x, eps, 0, 0)
        y = T.round(T.TensorTrain.from_cores(cores), eps=eps)
        for got in (y.to_dense(), ref.to_dense()):
            assert np.linalg.norm(got - dense) <= eps * np.linalg.norm(dense) * (1 + 1e-9)
        gaps = [_cut_gap(sig, r) for r in (y.ranks[1], ref.ranks[1])]
        if min(gaps) <= 1e-3:
            exempt_seen = True
            assert abs(y.ranks[1] - ref.ranks[1]) <= 1
            assert np.linalg.norm(y.to_dense() - ref.to_dense()) <= 2 * eps * np.linalg.norm(dense)
        else:
            assert y.ranks == ref.ranks
            assert rel_err(y.cores, ref) <= 1e-12
    assert exempt_seen  # f = 0.85 cuts inside the tie (rank 14)


def _mat_dev(o):
    return T.TensorTrainMatrix.from_cores(
        [c.reshape(c.shape[0], m, n, c.shape[2]) for c, m, n in zip(o.cores(), o.m, o.n)])


def test_round_tt_matrix_direct():
    a = O.OTT.random_matrix([2, 3, 4, 2, 3], [3, 2, 2, 4, 2], 6, 71)
    c = O.add(O.add(a, a), O.OTT.random_matrix([2, 3, 4, 2, 3], [3, 2, 2, 4, 2], 3, 72))
    for eps, cap, gauge in ((1e-10, 0, 0), (1e-10, 0, 1), (0.0, 4, 0)):
        ref = O.round(c, eps, cap, gauge)
        y = T.round(_mat_dev(c), eps=eps, max_rank=cap, gauge=gauge)
        assert y.ranks == ref.ranks
        got = [k.reshape(k.shape[0], -1, k.shape[-1]) for k in y.cores]
        assert rel_err(got, ref) <= 1e-12


def test_right_orthogonalize_tt_matrix():
    # Ranks that shrink: the last core (4, 2*2, 1) has rank at most 4, and the
    # stored bond 8 of random() before it (the last-bond quirk) drops to 4.
    a = O.OTT.random_matrix([2, 3, 2, 2], [2, 2, 3, 2], 8, 81)
    ref = a.copy()
    O.right_orthogonalize(ref)
    y = _mat_dev(a)
    T.right_orthogonalize(y)
    assert y.ranks == ref.ranks
    assert y.ranks != a.ranks
    cores = [k.reshape(k.shape[0], -1, k.shape[-1]) for k in y.cores]
    for k in range(1, len(cores)):
        M = cores[k].reshape(cores[k].shape[0], -1)
        assert np.linalg.norm(M @ M.T - np.eye(M.shape[0])) <= 1e-12
    dense = a.to_dense()
    got = O.OTT.from_cores(cores, a.m, a.n).to_dense()
    assert np.linalg.norm(got - dense) <= 1e-12 * np.linalg.norm(dense)
    assert rel_err(cores, ref) <= 1e-12
