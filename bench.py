#!/usr/bin/env python3
"""Benchmark of the TT-rounding hot path on one MI355X.

Workload (one step): plain FP64 ``matvec_round(A, x, eps=1e-8)`` of the
configuration the headline metric names -- TT-matrix A = random([4]^20,
[4]^20, 8, seed 3) with core k scaled by 1/sqrt(r_k m_k n_k), x =
random([4]^20, 32, seed 4), y = A x of ranks 256 rounded at eps = 1e-8.
Inputs are generated by the library's seeded generators (the reference's
``random`` semantics) and are resident in HBM before the timed region.

Prints one JSON line.  ``value`` is matvec_rounds per second on the one GPU;
``roofline`` covers the kernel class that takes the most device time in the
timed region (HIP events on the library's stream); ``cpu_baseline`` times the
FP64 CPU oracle (a restatement of the reference algorithm) on the same
inputs on the host.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "TT-rounds/s at rel.err<=1e-6, mixed+learned vs plain FP64 GPU; matvec_round ms"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
D, RA, RX, EPS = 20, 8, 32, 1e-8


def build_inputs(T):
    A = T.random_matrix([4] * D, [4] * D, RA, 3)
    rk = A.ranks
    for k in range(D):
        T.scale_core(A, k, 1.0 / math.sqrt(rk[k] * 16.0))
    x = T.random([4] * D, RX, 4)
    return A, x


def _f_qr(m: int, n: int) -> float:
    """Thin Householder QR with Q formed, GVL count (the larger dimension first)."""
    m, n = max(m, n), min(m, n)
    return 4.0 * m * n * n - 4.0 / 3.0 * n ** 3


def algorithmic_flops(A, x, y_ranks) -> float:
    """FP64 flops of the plain algorithm at this workload's shapes: the MPO
    apply, the right-orthogonalizing QR sweep (QR of each (n r' x r) core and
    the absorbing GEMM) and the SVD sweep (QR-condensed 22 N^3 SVD when
    rows >= 2 cols, else a 14 M N^2 + 8 N^3 thin SVD, plus the absorbing
    GEMM), with the output ranks of the run.  The same figure is credited
    whichever way the device reaches the result."""
    ra, rx = A.ranks, x.ranks
    m, n = A.row_shape, A.col_shape
    d = len(m)
    r = [ra[k] * rx[k] for k in range(d + 1)]
    f = sum(2.0 * ra[k] * m[k] * ra[k + 1] * n[k] * rx[k] * rx[k + 1] for k in range(d))
    reff = list(r)
    for k in range(d - 1, 0, -1):
        reff[k] = min(m[k] * reff[k + 1], r[k])
        f += _f_qr(m[k] * reff[k + 1], r[k]) + 2.0 * r[k - 1] * m[k - 1] * r[k] * reff[k]
    left = 1
    for k in range(d - 1):
        rows, cols = left * m[k], reff[k + 1]
        if rows >= 2 * cols:
            f += _f_qr(rows, cols) + 22.0 * cols ** 3 + 2.0 * rows * cols * cols
        else:
            big, small = max(rows, cols), min(rows, cols)
            f += 14.0 * big * small * small + 8.0 * small ** 3
        rnew = y_ranks[k + 1]
        f += 2.0 * rnew * cols * m[k + 1] * reff[k + 2]
        left = rnew
    return f


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE from two
    rocprofv3 counter runs of this benchmark, scaled by the same counters on
    a copy kernel of known size), from profiles/traffic.json; None when the
    kernel has not been measured."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    return None if rec is None else float(rec["bytes_per_launch"])


def _harness(fn, iters):
    """The reference benchmark harness (tests/bench/tt_bench_utils.hpp:38-67):
    max(3, min(32, iters / 10)) untimed calls, then `iters` timed calls;
    minimum and median in ms."""
    warm = max(3, min(32, iters // 10))
    for _ in range(warm):
        fn()
    ts = []
    for _ in range(iters):
        t0 = time.perf_counter()
        fn()
        ts.append((time.perf_counter() - t0) * 1e3)
    ts.sort()
    med = ts[len(ts) // 2] if len(ts) % 2 else 0.5 * (ts[len(ts) // 2 - 1] + ts[len(ts) // 2])
    return {"warmup": warm, "iters": iters, "min_ms": ts[0], "median_ms": med}


def _host_threads():
    """Threads for the batched CPU baseline: the host cores this process may
    run on, at most 16 (the CPU share of one GPU on the measuring hosts)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_leg(args):
    """One build of the FP64 CPU oracle (TT_ORACLE_VARIANT selects -O2 or
    -O3 -march=native) timed on the CPU only: the matvec_round workload with
    the reference harness, and bounded samples of the two batch workloads
    rounded by a std::thread pool.  Prints one JSON object."""
    import numpy as np

    import oracle as O

    O.set_svd_extended(False)  # timed in FP64, the reference's arithmetic
    A = O.OTT.random_matrix([4] * D, [4] * D, RA, 3)
    cores = [c / np.sqrt(c.shape[0] * 16.0) for c in A.cores()]
    A = O.OTT.from_cores(cores, [4] * D, [4] * D)
    x = O.OTT.random([4] * D, RX, 4)
    res = {"matvec_round": _harness(lambda: O.matvec_round(A, x, EPS, 0, 0), args.cpu_reps)}
    thr = _host_threads()
    rng = np.random.default_rng(2024)
    # Config 4 sample: the same construction (left-orthogonal Q_k
    # diag(10^(-j s)) cores from QR of Gaussian matrices, unit last core) with
    # numpy's generator instead of the device's.
    c4 = []
    for t in range(args.cpu_batch_cfg4):
        s = rng.uniform(0.05, 0.3)
        ranks = [1] + [min(4 ** k, 4 ** (32 - k), 128) for k in range(1, 32)] + [1]
        cs = []
        for k in range(32):
            G = rng.standard_normal((ranks[k] * 4, ranks[k + 1]))
            if k < 31:
                Q, _ = np.linalg.qr(G)
                cs.append((Q * 10.0 ** (-s * np.arange(ranks[k + 1])))[:, :].reshape(ranks[k], 4, ranks[k + 1]))
            else:
                cs.append((G / np.linalg.norm(G)).reshape(ranks[k], 4, 1))
        c4.append(O.OTT.from_cores(cs))
    # Config 3 sample: add(u, add(scale(u, 0.5), v)) with the reference's
    # random() (mt19937_64) for u and v, shapes drawn as in the GPU batch.
    c3 = []
    for t in range(args.cpu_batch_cfg3):
        d = int(rng.integers(8, 41))
        n = int(rng.choice([2, 4, 8]))
        u = O.OTT.random([n] * d, int(rng.integers(8, 65)), 1000 + t)
        v = O.OTT.random([n] * d, int(rng.integers(8, 65)), 501000 + t)
        c3.append(O.add(u, O.add(O.scale(u, 0.5), v)))
    for name, batch in (("config4", c4), ("config3", c3)):
        O.round_batch(batch[:1], 1e-6, 0, threads=1)  # warm-up of one TT
        t0 = time.perf_counter()
        O.round_batch(batch, 1e-6, 0, threads=thr)
        dt = time.perf_counter() - t0
        res[name] = {"rounds_per_s": len(batch) / dt, "tts": len(batch), "threads": thr, "seconds": dt}
    print(json.dumps(res))


def cpu_baseline(args, y_dev_cores):
    """The FP64 CPU oracle (the restated reference algorithm; the reference's
    Eigen build cannot be compiled here) timed on this host in two builds, as
    separate processes: -O2 without -march (the reference's flags) and
    -O3 -march=native.  Also returns the relative error of the GPU result
    against the oracle run with its extended-precision SVD accumulation."""
    import subprocess

    import numpy as np

    import oracle as O

    if not os.path.exists(O._LIB_PATH):
        O.build()
    legs = {}
    for variant in ("o2", "native"):
        env = dict(os.environ, TT_ORACLE_VARIANT="native" if variant == "native" else "")
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-leg", "--cpu-reps", str(args.cpu_reps),
               "--cpu-batch-cfg3", str(args.cpu_batch_cfg3), "--cpu-batch-cfg4", str(args.cpu_batch_cfg4)]
        print(f"[bench] CPU baseline ({variant}) ...", file=sys.stderr, flush=True)
        r = subprocess.run(cmd, env=env, capture_output=True, text=True, check=True)
        legs[variant] = json.loads(r.stdout.strip().splitlines()[-1])
    A = O.OTT.random_matrix([4] * D, [4] * D, RA, 3)
    cores = [c / np.sqrt(c.shape[0] * 16.0) for c in A.cores()]
    A = O.OTT.from_cores(cores, [4] * D, [4] * D)
    x = O.OTT.random([4] * D, RX, 4)
    O.set_svd_extended(True)
    yref = O.matvec_round(A, x, EPS, 0, 0)
    diff = O.axpby(1.0, O.OTT.from_cores(y_dev_cores), -1.0, yref)
    O.right_orthogonalize(diff)
    r = yref.copy()
    O.right_orthogonalize(r)
    rel = float(np.linalg.norm(diff.cores()[0]) / np.linalg.norm(r.cores()[0]))
    o2 = legs["o2"]["matvec_round"]
    try:
        nproc = os.cpu_count()
    except Exception:  # noqa: BLE001
        nproc = None
    return {
        "value": 1e3 / o2["median_ms"],
        "unit": "TT-rounds/s",
        "cores": 1,
        "kind": "port",
        "sample": f"matvec_round(A, x, eps=1e-8) of the d=20 workload by the FP64 CPU oracle (restated reference "
                  f"algorithm, g++ -O2, one thread), reference harness: {o2['warmup']} warm-up + {o2['iters']} "
                  f"timed calls, median {o2['median_ms']:.0f} ms, min {o2['min_ms']:.0f} ms",
        "host_nproc": nproc,
        "o2": legs["o2"],
        "o3_march_native": legs["native"],
        "batched_note": "config 3 / config 4 samples rounded by a std::thread pool (one TT per thread at a time, "
                        "threads = the host cores available, at most 16); inputs drawn with the same constructions "
                        "on the CPU (numpy / mt19937_64) rather than the device generators; one warm-up TT, one "
                        "timed pass",
    }, rel, yref.ranks


def _rel(T, a, b):
    """||a - b|| / ||b|| as ||core 0|| of the right-orthogonalized difference
    TT (QR on the device, no squaring; resolves about 1e-15)."""
    import numpy as np

    dif = T.sub(a, b)
    T.right_orthogonalize(dif)
    ref = b.copy()
    T.right_orthogonalize(ref)
    return float(np.linalg.norm(dif.cores[0]) / np.linalg.norm(ref.cores[0]))


def _timed(ctx, fn, reps=1):
    """(fn(), seconds): fn timed on the wall clock, the stream drained before
    and after; with reps > 1 the median of that many calls (the last call's
    result is returned)."""
    walls = []
    out = None
    for _ in range(reps):
        del out
        ctx.synchronize()
        t0 = time.perf_counter()
        out = fn()
        ctx.synchronize()
        walls.append(time.perf_counter() - t0)
    return out, sorted(walls)[len(walls) // 2]


def _kernel_ms(ctx, fn, classes=("qr", "jacobi", "gemm", "chol", "other")):
    """Per-class device milliseconds of one more call of fn, timed by the
    library's HIP events, and its Jacobi and rank-revealing counters."""
    ctx.prof_enable(True)
    out = fn()
    ctx.synchronize()
    ms = {k: ctx.prof_read(k)["ms"] for k in classes}
    extra = {"rank_revealing": ctx.rank_revealing_counts(), "jacobi_calls_sweeps": ctx.jacobi_sweeps()}
    ctx.prof_enable(False)
    del out
    return ms, extra


def _train_mlp(T, ctx, train, eps, steps):
    import numpy as np

    plain_tr = T.round_batched(train, eps=eps)
    X = T.bond_features(train, eps)
    y = np.array([p.ranks[k] / t.ranks[k] for t, p in zip(train, plain_tr) for k in range(1, t.d)], np.float32)
    mlp = T.RankMLP(X.shape[1], 64, 3, seed=1)
    ctx.synchronize()
    t0 = time.perf_counter()
    loss = mlp.train(X, y, steps=steps, lr=2e-3)
    ctx.synchronize()
    t_train = time.perf_counter() - t0
    # FP32 training, BF16 inference.
    mlp.set_bf16_inference(True)
    # The policy's gate: the learned rounding is used for this workload only
    # when it falls back on at most half of the training TTs (a fallback costs
    # the sketch plus the deterministic rounding).
    _, _, fb, _ = T.round_learned_batched(mlp, train, eps, seed=5, fallback="mixed")
    return mlp, {"train_bonds": int(len(y)), "steps": steps, "final_mse": loss, "train_s": t_train,
                 "inference": "bf16", "train_set_fallback_fraction": float(fb.mean()),
                 "gate_open": bool(fb.mean() <= 0.5)}


def _three_paths(T, ctx, batch, mlp, gate_open, eps, label):
    """Plain FP64, mixed and mixed + learned batched rounding of one
    workload: rounds/s, kernel time per class, sampled errors (against the
    input and against plain).  The learned rounding (randomized sketch
    rounding with the network's target ranks, mixed fallback) is timed on
    its own; "mixed + learned" is the policy: the learned rounding where its
    gate is open (it fell back on at most half of the training TTs), the
    mixed rounding otherwise."""
    import numpy as np

    n = len(batch)

    def learned_fn():
        return T.round_learned_batched(mlp, batch, eps, oversample=6, slack=0.1, seed=3, fallback="mixed")

    # One untimed pass of each path over the whole batch: allocator segments
    # and the GEMM tile autotuner's per-shape decisions are in place before
    # the timed passes, as in a long-running process.
    for _ in range(2):
        T.round_batched(batch, eps=eps)
        T.round_batched(batch, eps=eps, precision="mixed")
        learned_fn()
    ctx.synchronize()
    # Median of three timed passes of each path.
    plain, t_plain = _timed(ctx, lambda: T.round_batched(batch, eps=eps), reps=3)
    mixed, t_mixed = _timed(ctx, lambda: T.round_batched(batch, eps=eps, precision="mixed"), reps=3)
    (learned, err, fb, ranks), t_learned = _timed(ctx, learned_fn, reps=3)
    t_policy = t_learned if gate_open else t_mixed
    policy = learned if gate_open else mixed
    sample = list(range(0, n, max(1, n // 16)))
    res = {
        "plain_fp64_rounds_per_s": n / t_plain,
        "mixed_rounds_per_s": n / t_mixed,
        "learned_rounds_per_s": n / t_learned,
        "mixed_learned_policy": "learned" if gate_open else "mixed (learned gate closed)",
        "mixed_learned_rounds_per_s": n / t_policy,
        "speedup_mixed_vs_plain": t_plain / t_mixed,
        "speedup_learned_vs_plain": t_plain / t_learned,
        "speedup_mixed_learned_vs_plain": t_plain / t_policy,
        "learned_fallback_fraction": float(fb.mean()),
        "learned_max_discarded_rel_norm": float(err[fb == 0].max()) if (fb == 0).any() else None,
        "max_rel_err_vs_input_sampled": {
            "plain": max(_rel(T, plain[i], batch[i]) for i in sample),
            "mixed": max(_rel(T, mixed[i], batch[i]) for i in sample),
            "learned": max(_rel(T, learned[i], batch[i]) for i in sample),
            "mixed_learned": max(_rel(T, policy[i], batch[i]) for i in sample)},
        "max_rel_diff_vs_plain_sampled": {
            "mixed": max(_rel(T, mixed[i], plain[i]) for i in sample),
            "learned": max(_rel(T, learned[i], plain[i]) for i in sample),
            "mixed_learned": max(_rel(T, policy[i], plain[i]) for i in sample)},
        "mean_rank": {"plain": float(np.mean([np.mean(p.ranks[1:-1]) for p in plain])),
                      "mixed": float(np.mean([np.mean(p.ranks[1:-1]) for p in mixed])),
                      "learned_target": float(np.mean([np.mean(r[1:-1]) for r in ranks]))},
        "mixed_ranks_equal_plain": f"{sum(int(a.ranks == b.ranks) for a, b in zip(mixed, plain))}/{n}",
    }
    del plain, mixed, learned, policy
    res["kernel_ms"] = {}
    for name, fn in (("plain", lambda: T.round_batched(batch, eps=eps)),
                     ("mixed", lambda: T.round_batched(batch, eps=eps, precision="mixed")),
                     ("learned", lambda: learned_fn()[0])):
        ms, extra = _kernel_ms(ctx, fn)
        res["kernel_ms"][name] = ms
        if name == "plain":
            res["plain_jacobi_calls_sweeps"] = list(extra["jacobi_calls_sweeps"])
        if name == "mixed":
            res["mixed_certified_steps"] = {"tried": extra["rank_revealing"][0],
                                            "accepted": extra["rank_revealing"][1]}
    print(f"[bench] {label}: plain {res['plain_fp64_rounds_per_s']:.0f}, mixed {res['mixed_rounds_per_s']:.0f}, "
          f"learned {res['learned_rounds_per_s']:.0f}, mixed+learned {res['mixed_learned_rounds_per_s']:.0f} "
          f"rounds/s", file=sys.stderr, flush=True)
    return res


def qtt_section(T, ctx, args):
    """The large single-instance QTT workload at full size: d = 50, n = 2, x
    of rank 256 (Gaussian cores, left-orthogonalized, bond spectrum rescaled
    to 0.9^j), MPO of rank 16 (N(0, 1/(4 r)) cores), plain FP64 matvec_round
    and round at eps = 1e-10 (product ranks up to 4096).  One timed call each
    (each is tens of seconds); with --qtt-prof a second, event-timed
    matvec_round gives the time per kernel class."""
    import numpy as np

    d = 50
    x = T.random([2] * d, 256, 11)
    T.left_orthogonalize(x)
    cores = x.cores
    for k in range(d - 1):
        cores[k] = cores[k] * (0.9 ** np.arange(cores[k].shape[2]))[None, None, :]
    x = T.TensorTrain.from_cores(cores)
    A = T.random_matrix([2] * d, [2] * d, 16, 12)
    rk = A.ranks
    for k in range(d):
        T.scale_core(A, k, 1.0 / np.sqrt(rk[k] * 4.0))
    ctx.synchronize()
    t0 = time.perf_counter()
    y = T.matvec_round(A, x, eps=1e-10)
    ctx.synchronize()
    t_mvr = time.perf_counter() - t0
    yfull = T.matvec(A, x)
    ctx.synchronize()
    t0 = time.perf_counter()
    y2 = T.round(yfull, eps=1e-10)
    ctx.synchronize()
    t_round = time.perf_counter() - t0
    out = {"workload": "QTT d=50, n=2, x rank 256 (0.9^j bond spectrum), MPO rank 16, eps=1e-10, plain FP64",
           "matvec_round_s": t_mvr, "round_of_product_s": t_round, "product_max_rank": max(yfull.ranks),
           "out_max_rank": max(y.ranks), "out_ranks_equal": y.ranks == y2.ranks}
    del yfull, y2
    if args.qtt_prof:
        ctx.prof_enable(True)
        y = T.matvec_round(A, x, eps=1e-10)
        ctx.synchronize()
        out["kernel_ms"] = {k: ctx.prof_read(k)["ms"] for k in ("qr", "jacobi", "gemm", "gemm_splitk", "chol",
                                                                "chol_step", "apply", "other")}
        out["jacobi_calls_sweeps"] = list(ctx.jacobi_sweeps())
        ctx.prof_enable(False)
    print(f"[bench] QTT d=50: matvec_round {t_mvr:.1f} s, round {t_round:.1f} s", file=sys.stderr, flush=True)
    return out


def batched_section(T, ctx, args):
    """Batched workloads (config 4 and config 3): plain FP64, mixed, and
    mixed + learned batched rounding, and for config 4 the randomized
    rounding without the network (stored ranks as targets)."""
    import numpy as np

    out = {}
    eps = 1e-6
    rng = np.random.default_rng(2024)

    # ---- config 4: 2048 TTs, d=32, n=4, ranks <= 128, decaying spectra.
    n4 = args.cfg4_count
    s_train = rng.uniform(0.05, 0.3, args.train_tts)
    train = T.build_decay_batch(args.train_tts, 32, 4, 128, s_train, 10_000_000 + np.arange(args.train_tts))
    mlp4, mlp4_info = _train_mlp(T, ctx, train, eps, args.mlp_steps)
    del train
    s4 = rng.uniform(0.05, 0.3, n4)
    batch = T.build_decay_batch(n4, 32, 4, 128, s4, np.arange(n4))
    r4 = _three_paths(T, ctx, batch, mlp4, mlp4_info["gate_open"], eps, "config 4")
    # The method without the network: the same randomized rounding and
    # fallback with non-learned target ranks, the stored ranks (no truncation
    # but the final projection's) and fixed fractions of them plus 6.
    policies = {"stored_ranks": [t.ranks for t in batch]}
    for f in (0.6, 0.8):
        policies[f"fraction_{f}"] = [[1] + [min(t.ranks[k], int(np.ceil(f * t.ranks[k])) + 6)
                                            for k in range(1, t.d)] + [1] for t in batch]
    r4["non_learned_policies"] = {}
    for name, rk in policies.items():
        T.round_randomized_batched(batch, rk, eps, seed=3, fallback="mixed")
        (nolearn, err0, fb0), t0s = _timed(
            ctx, lambda: T.round_randomized_batched(batch, rk, eps, seed=3, fallback="mixed"))
        r4["non_learned_policies"][name] = {"rounds_per_s": n4 / t0s, "fallback_fraction": float(fb0.mean()),
                                            "mean_target_rank": float(np.mean([np.mean(r[1:-1]) for r in rk]))}
        del nolearn
    r4.update({"workload": f"{n4} TTs, d=32, n=4, ranks min(4^k, 4^(32-k), 128), cores Q_k diag(10^(-j s)), "
                           "s~U[0.05,0.3], eps=1e-6", "mlp": mlp4_info})
    out["config4"] = r4
    del batch

    # ---- config 3: 4096 TTs, d~U{8..40}, n in {2,4,8}, add(u, add(0.5u, v)).
    def sums(count, seed0):
        d = rng.integers(8, 41, count)
        nn = rng.choice([2, 4, 8], count)
        ru = rng.integers(8, 65, count)
        rv = rng.integers(8, 65, count)
        return T.build_sum_batch(d, nn, ru, rv, seed0 + np.arange(count), seed0 + 500000 + np.arange(count))

    train3 = sums(args.train_tts_cfg3, 50_000_000)
    mlp3, mlp3_info = _train_mlp(T, ctx, train3, eps, args.mlp_steps)
    del train3
    c3 = sums(args.cfg3_count, 1000)
    r3 = _three_paths(T, ctx, c3, mlp3, mlp3_info["gate_open"], eps, "config 3")
    r3.update({"workload": f"{args.cfg3_count} TTs, d~U{{8..40}}, n in {{2,4,8}}, add(u, add(0.5u, v)), "
                           "r_u, r_v ~ U{8..64}, eps=1e-6", "mlp": mlp3_info})
    out["config3"] = r3
    return out


def copy_calibration(mib: int) -> None:
    """Copies a float64 device buffer of `mib` MiB into another five times
    with torch's elementwise copy: each dispatch reads and writes exactly
    `mib` MiB.  The per-dispatch FETCH_SIZE / WRITE_SIZE that rocprofv3
    reports for it, divided by mib * 1024 KiB, is the factor by which those
    counters over- or under-count HBM bytes on this device."""
    import torch

    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x 5", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the copy kernel of known size used to calibrate rocprofv3's "
                         "FETCH_SIZE / WRITE_SIZE counters, then exit")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-reps", type=int, default=2,
                    help="timed matvec_round calls of the CPU baseline (after the harness's warm-up calls)")
    ap.add_argument("--cpu-batch-cfg3", type=int, default=128)
    ap.add_argument("--cpu-batch-cfg4", type=int, default=64)
    ap.add_argument("--cpu-leg", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--prof-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-qtt", action="store_true", help="skip the full-size QTT (d=50, ranks 4096) workload")
    ap.add_argument("--qtt-only", action="store_true")
    ap.add_argument("--qtt-prof", action="store_true")
    ap.add_argument("--batched-only", action="store_true",
                    help="skip the matvec_round workload and the CPU baseline (profiling the batched paths)")
    ap.add_argument("--cfg4-count", type=int, default=2048)
    ap.add_argument("--cfg3-count", type=int, default=4096)
    ap.add_argument("--train-tts", type=int, default=650)
    ap.add_argument("--train-tts-cfg3", type=int, default=850)
    ap.add_argument("--mlp-steps", type=int, default=1500)
    ap.add_argument("--plain-decay-only", type=int, default=0, metavar="COUNT",
                    help="only round COUNT decaying-spectrum TTs (d=32, n=4, ranks <= 128, eps=1e-6) by the plain "
                         "FP64 batched path, once untimed and once timed (a short run for counter profiles of the "
                         "batched QR and Jacobi kernels), then exit")
    ap.add_argument("--full-svd-steps", type=int, default=2,
                    help="steps timed again with the full-rank certificate, CholQR2 and speculative sweeps "
                         "off (every QR is Householder and every SVD step runs its SVD)")
    args = ap.parse_args()
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    if args.cpu_leg:
        cpu_leg(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the batched path would shard "
              "independent TTs across GPUs, which this library does not do.", file=sys.stderr)
        sys.exit(2)

    import tensor_train_amd as T

    ctx = T.Context.default(0)
    if args.plain_decay_only > 0:
        import numpy as np

        n = args.plain_decay_only
        s = np.random.default_rng(2024).uniform(0.05, 0.3, n)
        batch = T.build_decay_batch(n, 32, 4, 128, s, np.arange(n))
        T.round_batched(batch, eps=1e-6)
        out, t = _timed(ctx, lambda: T.round_batched(batch, eps=1e-6))
        print(json.dumps({"workload": f"plain FP64 batched rounding of {n} decaying-spectrum TTs (d=32, n=4, "
                                      "ranks <= 128, eps=1e-6)", "rounds_per_s": n / t}))
        return
    if args.batched_only:
        print(json.dumps({"batched": batched_section(T, ctx, args)}))
        return
    if args.qtt_only:
        print(json.dumps({"qtt_full": qtt_section(T, ctx, args)}))
        return
    A, x = build_inputs(T)
    for _ in range(args.warmup):
        y = T.matvec_round(A, x, eps=EPS)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        y = T.matvec_round(A, x, eps=EPS)
    ctx.synchronize()
    t1 = time.perf_counter()
    # Kernel timing in separate steps (the events add a little host work).
    ctx.prof_enable(True)
    for _ in range(args.prof_steps):
        y = T.matvec_round(A, x, eps=EPS)
    ctx.synchronize()
    stats = {k: ctx.prof_read(k) for k in ("qr", "jacobi", "gemm", "gemm_splitk", "gemm_chain", "gemm_multi",
                                           "gemm_grouped", "chol_step", "apply", "chol", "rinv", "other")}
    cert_tried, cert_held = ctx.rank_certificates()
    cholqr_tried, cholqr_used = ctx.cholqr_counts()
    spec_sweeps, spec_failed = ctx.speculation_counts()
    chain_blocks, chain_steps, chain_one_pass, chain_rejected = ctx.chain_stats()
    # Single-kernel classes: every timed launch is one kernel (the Jacobi
    # round kernel, the FP64 GEMM, the MPO core contraction, CholQR2's
    # Cholesky, the triangular inverse, the Gram chains' products, the
    # chained blocks' grouped products); the roofline is quoted for the one
    # with the most device time.
    single = {"jacobi_round": ("bj_round_kernel", ctx.prof_read("jacobi_round")),
              "gemm": ("gemm_f64_kernel", stats["gemm"]), "apply": ("matvec_core_batch_kernel", stats["apply"]),
              "gemm_splitk": ("gemm_f64_splitk_kernel", stats["gemm_splitk"]),
              "gemm_chain": ("gemm_chain_kernel", stats["gemm_chain"]),
              "gemm_grouped": ("gemm_f64_grouped_kernel", stats["gemm_grouped"]),
              "gemm_multi": ("gemm_f64_multi_kernel", stats["gemm_multi"]),
              "chol_step": ("chol_step_kernel", stats["chol_step"]),
              "chol": ("potrf_inv_kernel", stats["chol"]), "rinv": ("rinv_cert_kernel", stats["rinv"])}
    ctx.prof_enable(False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    full_svd_ms = None
    svd_every_ms = None
    if args.full_svd_steps > 0:
        # The truncating path's cost at this shape: the certificate off, so
        # every SVD step runs its Jacobi SVD (QR sweep, CholQR2, chained
        # blocks and speculation as by default).
        ctx.set_rank_certificate(False)
        y_full = T.matvec_round(A, x, eps=EPS)
        ctx.synchronize()
        t2 = time.perf_counter()
        for _ in range(args.full_svd_steps):
            y_full = T.matvec_round(A, x, eps=EPS)
        ctx.synchronize()
        svd_every_ms = (time.perf_counter() - t2) * 1e3 / args.full_svd_steps
        ctx.set_rank_certificate(True)
        assert y_full.ranks == y.ranks
        # The same with everything by Householder and Jacobi.
        ctx.set_rank_certificate(False)
        ctx.set_cholqr(False)
        ctx.set_speculation(False)
        y_full = T.matvec_round(A, x, eps=EPS)
        ctx.synchronize()
        t2 = time.perf_counter()
        for _ in range(args.full_svd_steps):
            y_full = T.matvec_round(A, x, eps=EPS)
        ctx.synchronize()
        full_svd_ms = (time.perf_counter() - t2) * 1e3 / args.full_svd_steps
        ctx.set_rank_certificate(True)
        ctx.set_cholqr(True)
        ctx.set_speculation(True)
        assert y_full.ranks == y.ranks
    alg_flops = algorithmic_flops(A, x, y.ranks)

    dom = max(single, key=lambda k: single[k][1]["ms"])
    kname, s = single[dom]
    intensity = s["flops"] / max(s["bytes"], 1.0)
    if intensity >= PEAK_FP64_TFLOPS * 1e12 / (PEAK_HBM_GBS * 1e9):
        achieved = s["flops"] / (s["ms"] * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64_TFLOPS}
    else:
        achieved = s["bytes"] / (s["ms"] * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS}
    traffic = measured_traffic(kname)
    roof.update({"traffic": traffic, "kernel": kname, "launches_per_step": s["launches"] / args.prof_steps,
                 "avg_launch_us": s["ms"] * 1e3 / max(s["launches"], 1),
                 "algorithmic_flops_per_launch": s["flops"] / max(s["launches"], 1),
                 "algorithmic_bytes_per_launch": s["bytes"] / max(s["launches"], 1)})

    out = {
        "metric": METRIC,
        "value": 1e3 / ms_per_step,
        "unit": "TT-rounds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded reference random() generators)",
        "config": {"workload": "plain FP64 matvec_round, TT-matrix [4x4]^20 rank 8 (scaled), x [4]^20 rank 32, "
                               "eps=1e-8, product ranks 256", "d": D, "eps": EPS, "y_ranks": y.ranks},
        "roofline": roof,
        "kernel_ms_per_step": {k: v["ms"] / args.prof_steps for k, v in stats.items()},
        "algorithmic_gflop_per_step": alg_flops / 1e9,
        "achieved_gflops_per_step": alg_flops / ms_per_step / 1e6,
        "speculative_sweeps": {"per_step": spec_sweeps / args.prof_steps,
                               "undone_per_step": spec_failed / args.prof_steps},
        "cholqr2": {"tall_qr_per_step": cholqr_tried / args.prof_steps,
                    "cholqr2_used_per_step": cholqr_used / args.prof_steps},
        "chained_blocks": {"per_step": chain_blocks / args.prof_steps,
                           "steps_per_step": chain_steps / args.prof_steps,
                           "one_pass_per_step": chain_one_pass / args.prof_steps,
                           "rejected_per_step": chain_rejected / args.prof_steps},
        "rank_certificate": {"svd_steps_per_step": cert_tried / args.prof_steps,
                             "certified_per_step": cert_held / args.prof_steps,
                             "ms_per_step_every_svd": svd_every_ms,
                             "ms_per_step_householder_and_svd": full_svd_ms},
    }
    if not args.no_batched:
        out["batched"] = batched_section(T, ctx, args)
    if not args.no_qtt:
        out["qtt_full"] = qtt_section(T, ctx, args)
    if not args.no_cpu:
        cb, rel, yref_ranks = cpu_baseline(args, y.cores)
        out["cpu_baseline"] = cb
        out["rel_err_vs_cpu"] = rel
        out["ranks_match_cpu"] = yref_ranks == y.ranks
    print(json.dumps(out))


if __name__ == "__main__":
    main()
