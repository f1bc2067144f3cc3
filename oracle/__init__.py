"""TEST INFRASTRUCTURE -- the FP64 CPU oracle, never part of the product path.

ctypes wrapper over ``oracle/_build/libtt_oracle.so`` (built from
``oracle/tt_oracle.cpp`` by ``oracle/Makefile``).  It restates the reference
algorithm of bdaszuta/tensor_train (``include/core/round/{tt_round,tt_soft_threshold,tt_als,tt_dmrg}.hpp``,
``include/core/detail/{orthogonalize,truncate,contract,streaming_apply,
fused_apply_core}.hpp``, ``include/core/apply/tt_matrix_apply.hpp``,
``include/core/algebra/{tt_inner,tt_ops}.hpp``,
``include/core/factory/{tt_factory,tt_matrix_factory,tt_svd,from_dense}.hpp``).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module, and only as the checker.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# TT_ORACLE_VARIANT=native selects the -O3 -march=native build (the CPU
# baseline's second line); the default is the -O2 build.
_VARIANT = os.environ.get("TT_ORACLE_VARIANT", "")
_LIB_PATH = os.path.join(_HERE, "_build", "libtt_oracle_native.so" if _VARIANT == "native" else "libtt_oracle.so")
_lib = None

_i32p = ctypes.POINTER(ctypes.c_int)
_f64p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


def build() -> None:
    """Compile the oracle shared library with its Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run oracle.build())")
        L = ctypes.CDLL(_LIB_PATH)
        L.ora_new.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p, _i32p, _f64p]
        L.ora_new.restype = _vp
        L.ora_free.argtypes = [_vp]
        L.ora_d.argtypes = [_vp]
        L.ora_ranks.argtypes = [_vp, _i32p]
        L.ora_modes.argtypes = [_vp, _i32p]
        L.ora_size.argtypes = [_vp]
        L.ora_size.restype = ctypes.c_longlong
        L.ora_data.argtypes = [_vp, _f64p]
        L.ora_random.argtypes = [ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_ulonglong]
        L.ora_random.restype = _vp
        L.ora_random_matrix.argtypes = [ctypes.c_int, _i32p, _i32p, ctypes.c_int, ctypes.c_ulonglong]
        L.ora_random_matrix.restype = _vp
        L.ora_copy.argtypes = [_vp]
        L.ora_copy.restype = _vp
        L.ora_axpby.argtypes = [ctypes.c_double, _vp, ctypes.c_double, _vp]
        L.ora_axpby.restype = _vp
        L.ora_scale.argtypes = [_vp, ctypes.c_double]
        L.ora_scale.restype = _vp
        L.ora_round.argtypes = [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.ora_round.restype = _vp
        L.ora_round_batch.argtypes = [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(_vp)]
        L.ora_round_batch.restype = None
        L.ora_right_orthogonalize.argtypes = [_vp]
        L.ora_left_orthogonalize.argtypes = [_vp]
        L.ora_matvec.argtypes = [_vp, _vp]
        L.ora_matvec.restype = _vp
        L.ora_matmat.argtypes = [_vp, _vp]
        L.ora_matmat.restype = _vp
        L.ora_matvec_round.argtypes = [_vp, _vp, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.ora_matvec_round.restype = _vp
        L.ora_matmat_round.argtypes = [_vp, _vp, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.ora_matmat_round.restype = _vp
        L.ora_inner.argtypes = [_vp, _vp]
        L.ora_inner.restype = ctypes.c_double
        L.ora_norm.argtypes = [_vp]
        L.ora_norm.restype = ctypes.c_double
        L.ora_dense_size.argtypes = [_vp]
        L.ora_dense_size.restype = ctypes.c_longlong
        L.ora_to_dense.argtypes = [_vp, _f64p]
        L.ora_qr_thin.argtypes = [_f64p, ctypes.c_int, ctypes.c_int, _f64p, _f64p]
        L.ora_svd_thin.argtypes = [_f64p, ctypes.c_int, ctypes.c_int, _f64p, _f64p, _f64p]
        L.ora_truncate_eps_rank.argtypes = [_f64p, ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.ora_round_als.argtypes = [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double, _vp,
                                    ctypes.c_int, ctypes.c_ulonglong, _i32p, _f64p, _i32p]
        L.ora_round_als.restype = _vp
        L.ora_round_dmrg.argtypes = [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double, _vp,
                                     ctypes.c_int, ctypes.c_ulonglong, _i32p, _f64p, _i32p]
        L.ora_round_dmrg.restype = _vp
        L.ora_set_svd_extended.argtypes = [ctypes.c_int]
        L.ora_set_svd_extended.restype = None
        L.ora_soft_threshold.argtypes = [_vp, ctypes.c_double]
        L.ora_soft_threshold.restype = _vp
        L.ora_from_dense.argtypes = [_f64p, ctypes.c_int, _i32p, ctypes.c_double, ctypes.c_int]
        L.ora_from_dense.restype = _vp
        L.ora_from_dense_qtt.argtypes = [_f64p, ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.ora_from_dense_qtt.restype = _vp
        L.ora_matrix_from_dense.argtypes = [_f64p, ctypes.c_int, _i32p, _i32p, ctypes.c_double, ctypes.c_int]
        L.ora_matrix_from_dense.restype = _vp
        _lib = L
    return _lib


def _ip(a):
    return a.ctypes.data_as(_i32p)


def _dp(a):
    return a.ctypes.data_as(_f64p)


class OTT:
    """Host TT (or TT-matrix) held by the oracle library."""

    def __init__(self, handle, m=None, n=None):
        self.h = handle
        self.m = None if m is None else list(m)
        self.n = None if n is None else list(n)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ora_free(self.h)
            self.h = None

    # -- construction ------------------------------------------------------
    @staticmethod
    def from_cores(cores, m=None, n=None):
        d = len(cores)
        modes = np.array([c.shape[1] for c in cores], dtype=np.int32)
        ranks = np.array([1] + [c.shape[2] for c in cores], dtype=np.int32) if d else np.array([1], np.int32)
        if d:
            ranks[0] = cores[0].shape[0]
        data = np.concatenate([np.ascontiguousarray(c, dtype=np.float64).ravel() for c in cores]) if d else np.zeros(1)
        mm = np.array(m, np.int32) if m is not None else None
        nn = np.array(n, np.int32) if n is not None else None
        h = lib().ora_new(d, _ip(modes), _ip(mm) if mm is not None else None,
                          _ip(nn) if nn is not None else None, _ip(ranks), _dp(data))
        return OTT(h, m, n)

    @staticmethod
    def random(shape, max_rank, seed):
        s = np.array(shape, np.int32)
        return OTT(lib().ora_random(len(shape), _ip(s), max_rank, seed))

    @staticmethod
    def random_matrix(row_shape, col_shape, max_rank, seed):
        m = np.array(row_shape, np.int32)
        n = np.array(col_shape, np.int32)
        return OTT(lib().ora_random_matrix(len(m), _ip(m), _ip(n), max_rank, seed), row_shape, col_shape)

    # -- inspection ----------------------------------------------------------
    @property
    def d(self):
        return lib().ora_d(self.h)

    @property
    def ranks(self):
        out = np.zeros(self.d + 1, np.int32)
        lib().ora_ranks(self.h, _ip(out))
        return [int(x) for x in out]

    @property
    def modes(self):
        out = np.zeros(max(self.d, 1), np.int32)
        lib().ora_modes(self.h, _ip(out))
        return [int(x) for x in out[: self.d]]

    def cores(self):
        data = np.zeros(max(int(lib().ora_size(self.h)), 1))
        lib().ora_data(self.h, _dp(data))
        r, md = self.ranks, self.modes
        out, off = [], 0
        for k in range(self.d):
            sz = r[k] * md[k] * r[k + 1]
            out.append(data[off:off + sz].reshape(r[k], md[k], r[k + 1]).copy())
            off += sz
        return out

    def to_dense(self):
        out = np.zeros(int(lib().ora_dense_size(self.h)))
        lib().ora_to_dense(self.h, _dp(out))
        return out

    def copy(self):
        return OTT(lib().ora_copy(self.h), self.m, self.n)


def axpby(alpha, a, beta, b):
    return OTT(lib().ora_axpby(alpha, a.h, beta, b.h), a.m, a.n)


def add(a, b):
    return axpby(1.0, a, 1.0, b)


def scale(a, alpha):
    return OTT(lib().ora_scale(a.h, alpha), a.m, a.n)


def round(a, eps=0.0, max_rank=0, gauge=0):  # noqa: A001 - mirrors the reference name
    return OTT(lib().ora_round(a.h, eps, max_rank, gauge), a.m, a.n)


def round_batch(tts, eps=0.0, max_rank=0, threads=1):
    """round(t, eps, max_rank) for every TT, on `threads` std::threads of the
    oracle (one TT per thread at a time)."""
    n = len(tts)
    arr = (_vp * n)(*[t.h for t in tts])
    outs = (_vp * n)()
    lib().ora_round_batch(arr, n, eps, max_rank, threads, outs)
    return [OTT(outs[i]) for i in range(n)]


def right_orthogonalize(a):
    lib().ora_right_orthogonalize(a.h)


def left_orthogonalize(a):
    lib().ora_left_orthogonalize(a.h)


def matvec(A, x):
    return OTT(lib().ora_matvec(A.h, x.h))


def matmat(A, B):
    return OTT(lib().ora_matmat(A.h, B.h), A.m, B.n)


def matvec_round(A, x, eps=0.0, max_rank=0, method=0):
    return OTT(lib().ora_matvec_round(A.h, x.h, eps, max_rank, method))


def matmat_round(A, B, eps=0.0, max_rank=0, method=0):
    return OTT(lib().ora_matmat_round(A.h, B.h, eps, max_rank, method), A.m, B.n)


def inner(a, b):
    return lib().ora_inner(a.h, b.h)


def norm(a):
    return lib().ora_norm(a.h)


def qr_thin(M):
    M = np.ascontiguousarray(M, dtype=np.float64)
    rows, cols = M.shape
    k = min(rows, cols)
    Q = np.zeros((rows, k))
    R = np.zeros((k, cols))
    lib().ora_qr_thin(_dp(M), rows, cols, _dp(Q), _dp(R))
    return Q, R


def svd_thin(M):
    M = np.ascontiguousarray(M, dtype=np.float64)
    rows, cols = M.shape
    k = min(rows, cols)
    U = np.zeros((rows, k))
    s = np.zeros(k)
    Vt = np.zeros((k, cols))
    lib().ora_svd_thin(_dp(M), rows, cols, _dp(U), _dp(s), _dp(Vt))
    return U, s, Vt


def frob_norm_apply(A, X):
    """frob_norm_apply (include/core/algebra/tt_matrix_frob_norm_apply.hpp:36-94):
    per core the product core (matvec_core / matmat_core) folded into the
    inner-product environment, i.e. norm(matvec(A, x)) / norm(matmat(A, B))
    with the same operation order; d = 0 gives 0."""
    if A.d == 0:
        return 0.0
    return norm(matmat(A, X) if X.m is not None else matvec(A, X))


def set_svd_extended(on: bool) -> None:
    """Accumulate the Jacobi SVD in long double (default) or in FP64.  The
    FP64 form is the reference's arithmetic, for timing; the extended form
    keeps the oracle's own rounding far below what it checks."""
    lib().ora_set_svd_extended(1 if on else 0)


def round_als(a, eps=0.0, max_rank=0, max_iters=20, tol=1e-12, warm=None, cold_random=False, seed=0):
    """ALS rounding (include/core/round/tt_als.hpp); returns (tt, iters_run,
    final_change, converged).  cold_random starts from random(shape,
    max_rank, seed) as the apply-then-ALS paths do."""
    it, cv = ctypes.c_int(), ctypes.c_int()
    ch = ctypes.c_double()
    h = lib().ora_round_als(a.h, eps, max_rank, max_iters, tol, warm.h if warm is not None else None,
                            1 if cold_random else 0, seed, ctypes.byref(it), ctypes.byref(ch), ctypes.byref(cv))
    return OTT(h, a.m, a.n), it.value, ch.value, bool(cv.value)


def matvec_round_als(A, x, eps=0.0, max_rank=0, max_iters=20, tol=1e-12, warm=None, seed=0):
    """matvec_als_round (include/core/round/tt_matvec_als_round.hpp:30-113)."""
    return round_als(matvec(A, x), eps, max_rank, max_iters, tol, warm, True, seed)


def matmat_round_als(A, B, eps=0.0, max_rank=0, max_iters=20, tol=1e-12, warm=None, seed=0):
    """matmat_als_round (include/core/round/tt_matmat_als_round.hpp:39-141)."""
    return round_als(matmat(A, B), eps, max_rank, max_iters, tol, warm, True, seed)


def round_dmrg(a, eps, max_rank=0, max_iters=20, tol=1e-12, warm=None, apply_path=False, seed=0):
    """Two-site DMRG rounding (include/core/round/tt_dmrg.hpp); returns (tt,
    iters_run, final_change, converged).  apply_path selects the starts of
    the apply-then-DMRG paths."""
    it, cv = ctypes.c_int(), ctypes.c_int()
    ch = ctypes.c_double()
    h = lib().ora_round_dmrg(a.h, eps, max_rank, max_iters, tol, warm.h if warm is not None else None,
                             1 if apply_path else 0, seed, ctypes.byref(it), ctypes.byref(ch), ctypes.byref(cv))
    return OTT(h, a.m, a.n), it.value, ch.value, bool(cv.value)


def matvec_round_dmrg(A, x, eps, max_rank=0, max_iters=20, tol=1e-12, warm=None, seed=0):
    """matvec_dmrg_round (include/core/round/tt_matvec_dmrg_round.hpp:29-105)."""
    return round_dmrg(matvec(A, x), eps, max_rank, max_iters, tol, warm, True, seed)


def matmat_round_dmrg(A, B, eps, max_rank=0, max_iters=20, tol=1e-12, warm=None, seed=0):
    """matmat_dmrg_round (include/core/round/tt_matmat_dmrg_round.hpp:38-130)."""
    return round_dmrg(matmat(A, B), eps, max_rank, max_iters, tol, warm, True, seed)


def soft_threshold(a, tau=0.0):
    """Singular-value soft thresholding (include/core/round/tt_soft_threshold.hpp)."""
    return OTT(lib().ora_soft_threshold(a.h, tau))


def from_dense(dense, shape, eps=1e-10, max_rank=0, method="svd", qtt_base=2):
    """TT-SVD of a row-major dense tensor (include/core/factory/tt_svd.hpp,
    from_dense.hpp)."""
    dense = np.ascontiguousarray(dense, dtype=np.float64).ravel()
    sh = np.array(shape, np.int32)
    if method == "qtt":
        return OTT(lib().ora_from_dense_qtt(_dp(dense), len(sh), _ip(sh), qtt_base, eps, max_rank))
    return OTT(lib().ora_from_dense(_dp(dense), len(sh), _ip(sh), eps, max_rank))


def matrix_from_dense(dense, row_shape, col_shape, eps=1e-10, max_rank=0):
    """TT-matrix from a row-major (prod m) x (prod n) matrix
    (include/core/factory/tt_matrix_factory.hpp:205-290)."""
    dense = np.ascontiguousarray(dense, dtype=np.float64).ravel()
    ms = np.array(row_shape, np.int32)
    ns = np.array(col_shape, np.int32)
    return OTT(lib().ora_matrix_from_dense(_dp(dense), len(ms), _ip(ms), _ip(ns), eps, max_rank), row_shape, col_shape)


def truncate_eps_rank(s, delta, max_rank=0):
    s = np.ascontiguousarray(s, dtype=np.float64)
    return int(lib().ora_truncate_eps_rank(_dp(s), len(s), delta, max_rank))
