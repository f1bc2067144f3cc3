"""tensor_train_amd -- TT rounding and MPO apply-and-round on MI355X (gfx950).

Host-side mirror of the reference's operator interface for the hot path
(bdaszuta/tensor_train: ``bind_py/tensor_train/__init__.py`` classes
``TensorTrain`` / ``TensorTrainMatrix`` / ``RoundOptions`` and the C++
functions ``round``, ``right_orthogonalize``, ``matvec_round``,
``matmat_round``, ``inner``, ``norm``).  Every call goes through the C ABI of
``libtt_amd.so`` (include/tt_amd.h) into hand-written HIP kernels; tensors
live in device memory between calls.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi
from ._capi import TTAError

__all__ = [
    "Context", "TensorTrain", "TensorTrainMatrix", "RoundOptions", "RoundResult", "round", "round_ex",
    "right_orthogonalize", "left_orthogonalize", "matvec", "matmat", "matvec_round", "matmat_round",
    "matvec_round_ex", "matmat_round_ex", "inner",
    "norm", "frob_inner", "frob_norm", "frob_norm_apply", "axpby", "add", "sub", "scale", "TTAError", "random", "random_matrix",
    "zeros", "ones", "random_host", "from_dense", "matrix_from_dense", "soft_threshold",
]

# round_method / round_gauge (include/core/round/round.hpp:33-53).
METHODS = {"svd": 0, "svd_streaming": 1, "svd_naive": 2, "als": 3, "dmrg": 4}
GAUGES = {"none": 0, "right_canonical": 1}


def _enum(table, v, what):
    if isinstance(v, str):
        if v not in table:
            raise ValueError(f"unknown {what} {v!r}; expected one of {sorted(table)}")
        return table[v]
    return int(v)


class Context:
    """One HIP stream + device pool on one GPU (``tta_ctx``)."""

    _default = {}

    def __init__(self, device: int = 0):
        self.lib = _capi.lib()
        h = ctypes.c_void_p()
        st = self.lib.tta_ctx_create(device, ctypes.byref(h))
        if st != 0:
            raise TTAError(st, f"tta_ctx_create(device={device}) failed (no usable GPU?)")
        self.h = h
        self.device = device

    @classmethod
    def default(cls, device: int = 0) -> "Context":
        if device not in cls._default:
            cls._default[device] = Context(device)
        return cls._default[device]

    def check(self, st):
        if st != 0:
            raise TTAError(st, self.lib.tta_last_error(self.h).decode(errors="replace"))

    def synchronize(self):
        self.check(self.lib.tta_synchronize(self.h))

    @property
    def stream(self) -> int:
        return int(self.lib.tta_stream(self.h) or 0)

    def prof_enable(self, on: bool = True):
        """Per-launch HIP-event timing of the library's kernels (resets totals)."""
        self.check(self.lib.tta_prof_enable(self.h, 1 if on else 0))

    def prof_read(self, kernel_class: str):
        """{launches, ms, flops, bytes} of one kernel class (qr, jacobi, gemm,
        apply, other): summed device time and the algorithmic work credited
        to those launches."""
        n = ctypes.c_longlong()
        ms, fl, by = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self.check(self.lib.tta_prof_read(self.h, _capi.KERNEL_CLASSES[kernel_class], ctypes.byref(n),
                                          ctypes.byref(ms), ctypes.byref(fl), ctypes.byref(by)))
        return {"launches": n.value, "ms": ms.value, "flops": fl.value, "bytes": by.value}

    def jacobi_sweeps(self):
        """(calls, sweeps) of the block-Jacobi SVD since the last prof_enable."""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        self.check(self.lib.tta_prof_jacobi_sweeps(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def rank_revealing_counts(self):
        """(tried, accepted) SVD steps of the mixed batched rounding's
        rank-revealing CholQR2 since the last prof_enable."""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        self.check(self.lib.tta_prof_rank_revealing(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def jacobi_unconverged(self) -> int:
        """Block-Jacobi problems still rotating at the sweep limit since the
        last prof_enable (0 when every SVD converged)."""
        a = ctypes.c_longlong()
        self.check(self.lib.tta_prof_jacobi_unconverged(self.h, ctypes.byref(a)))
        return a.value

    def set_rank_certificate(self, on: bool) -> None:
        """Full-rank certificate of the SVD steps (on by default): a step whose
        triangular factor provably keeps every singular value skips its SVD."""
        self.check(self.lib.tta_set_rank_certificate(self.h, 1 if on else 0))

    def set_cholqr(self, on: bool) -> None:
        """CholQR2 for the sweeps' tall QR factorizations (on by default; it
        falls back to Householder wherever it would lose accuracy)."""
        self.check(self.lib.tta_set_cholqr(self.h, 1 if on else 0))

    def set_autotune(self, on: bool) -> None:
        """Tile-size autotuner of the grouped GEMMs (on by default; identical
        results either way)."""
        self.check(self.lib.tta_set_autotune(self.h, 1 if on else 0))

    def autotune_stats(self):
        """(shape classes seen, decided, decided for the 32 x 32 tile)."""
        v = [ctypes.c_longlong() for _ in range(3)]
        self.check(self.lib.tta_prof_autotune(self.h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_graphs(self, on: bool) -> None:
        """Block-Jacobi sweeps captured into a HIP graph and replayed (off by
        default; identical results)."""
        self.check(self.lib.tta_set_graphs(self.h, 1 if on else 0))

    def set_jacobi_gram(self, on: bool) -> None:
        """Block-Jacobi rounds of few problems (65..512 columns) whose
        rotations are decided on the block pair's 16 x 16 Gram matrix (on by
        default; equal results up to rounding)."""
        self.check(self.lib.tta_set_jacobi_gram(self.h, 1 if on else 0))

    def set_speculation(self, on: bool) -> None:
        """Speculative sweeps (on by default): CholQR2 acceptance and the rank
        certificates decided on the device, one host read per sweep; a sweep
        with a failed decision is undone and run again checked."""
        self.check(self.lib.tta_set_speculation(self.h, 1 if on else 0))

    def set_chain(self, on: bool) -> None:
        """Chained sweep blocks inside speculative sweeps (on by default): runs
        of CholQR2-shaped steps factored together, with the Grams chained by
        products and the Choleskys batched off the critical path."""
        self.check(self.lib.tta_set_chain(self.h, 1 if on else 0))

    def chain_counts(self):
        """(blocks, steps) of chained sweep blocks since the last prof_enable."""
        return self.chain_stats()[:2]

    def chain_stats(self):
        """(blocks, steps, one_pass, rejected) of chained sweep blocks since
        the last prof_enable: blocks accepted, the steps they covered, those
        that needed no second pass, and blocks rejected (their steps then
        ran one by one)."""
        v = [ctypes.c_longlong() for _ in range(4)]
        self.check(self.lib.tta_prof_chain(self.h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def speculation_counts(self):
        """(sweeps, failed) speculative sweep counts since the last prof_enable."""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        self.check(self.lib.tta_prof_speculation(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def cholqr_counts(self):
        """(tried, used) CholQR2 counts since the last prof_enable."""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        self.check(self.lib.tta_prof_cholqr(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def cholqr_triangular_count(self) -> int:
        """CholQR2 factorizations accepted in the triangular form (condition
        numbers beyond the polar form's range) since the last prof_enable."""
        a = ctypes.c_longlong()
        self.check(self.lib.tta_prof_cholqr_triangular(self.h, ctypes.byref(a)))
        return a.value

    def rank_certificates(self):
        """(tried, held) certificate counts since the last prof_enable."""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        self.check(self.lib.tta_prof_rank_certificates(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def mem_stats(self):
        """{live_bytes, cached_bytes, device_allocs, trims} of the context's
        caching device allocator."""
        v = [ctypes.c_ulonglong() for _ in range(4)]
        self.check(self.lib.tta_mem_stats(self.h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("live_bytes", "cached_bytes", "device_allocs", "trims"), (x.value for x in v)))

    def mem_trim(self):
        """Returns the allocator's cached blocks to the device."""
        self.check(self.lib.tta_mem_trim(self.h))


class RoundOptions:
    """Options of ``round`` / ``matvec_round`` / ``matmat_round``.

    Same fields and defaults as ``round_options`` (round.hpp:55-105):
    eps=0.0 (no truncation), max_rank=0 (no cap), method="svd",
    gauge="none", max_iters=20, tol=1e-12, seed=0, warm_start=None (a
    TensorTrain, or a TensorTrainMatrix for matrix rounds: warm_start_tt /
    warm_start_tt_matrix).  method="als" and "dmrg" run the iterative
    engines.
    """

    def __init__(self, eps=0.0, max_rank=0, method="svd", gauge="none", max_iters=20, tol=1e-12, seed=0,
                 warm_start=None):
        self.eps = float(eps)
        self.max_rank = int(max_rank)
        self.method = _enum(METHODS, method, "round_method")
        self.gauge = _enum(GAUGES, gauge, "round_gauge")
        self.max_iters = int(max_iters)
        self.tol = float(tol)
        self.seed = int(seed)
        self.warm_start = warm_start

    @classmethod
    def merge(cls, opts=None, **kwargs) -> "RoundOptions":
        """RoundOptions / dict / kwargs merging of the reference wrapper
        (bind_py/tensor_train/__init__.py:603-656): kwargs override an
        options object; a dict is read as keyword arguments."""
        if isinstance(opts, RoundOptions):
            out = cls(opts.eps, opts.max_rank, opts.method, opts.gauge, opts.max_iters, opts.tol, opts.seed,
                      opts.warm_start)
            for k, v in kwargs.items():
                if k == "method":
                    v = _enum(METHODS, v, "round_method")
                elif k == "gauge":
                    v = _enum(GAUGES, v, "round_gauge")
                setattr(out, k, v)
            return out
        if isinstance(opts, dict):
            return cls(**opts)
        return cls(**kwargs)


class RoundResult:
    """round_result (round.hpp:112-119): iters_run, final_resid (the ALS
    final singular-value change), converged; the svd path reports {1, 0,
    True}."""

    def __init__(self, iters_run=1, final_resid=0.0, converged=True):
        self.iters_run = iters_run
        self.final_resid = final_resid
        self.converged = converged


class _DeviceTT:
    """Owner of a ``tta_tt`` handle."""

    def __init__(self, handle, ctx: Context):
        self._h = handle
        self._ctx = ctx

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._ctx.lib.tta_tt_free(self._ctx.h, h)
            except Exception:
                pass
            self._h = None

    # -- shape -------------------------------------------------------------
    @property
    def d(self) -> int:
        d = ctypes.c_int()
        m = ctypes.c_int()
        self._ctx.check(self._ctx.lib.tta_tt_dims(self._h, ctypes.byref(d), ctypes.byref(m)))
        return d.value

    def _shape_arrays(self):
        d = self.d
        modes = np.zeros(max(d, 1), np.int32)
        rm = np.zeros(max(d, 1), np.int32)
        cm = np.zeros(max(d, 1), np.int32)
        rk = np.zeros(d + 1, np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        self._ctx.check(self._ctx.lib.tta_tt_shape(self._h, modes.ctypes.data_as(ip), rm.ctypes.data_as(ip),
                                                   cm.ctypes.data_as(ip), rk.ctypes.data_as(ip)))
        return modes[:d], rm[:d], cm[:d], rk

    @property
    def ranks(self):
        return [int(x) for x in self._shape_arrays()[3]]

    @property
    def max_rank(self) -> int:
        return max(self.ranks)

    @property
    def num_params(self) -> int:
        return int(self._ctx.lib.tta_tt_numel(self._h))

    def _flat(self) -> np.ndarray:
        out = np.zeros(max(self.num_params, 1))
        self._ctx.check(self._ctx.lib.tta_tt_download(self._ctx.h, self._h,
                                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out[: self.num_params]

    def core_ptr(self, k: int) -> int:
        p = ctypes.POINTER(ctypes.c_double)()
        self._ctx.check(self._ctx.lib.tta_tt_core_ptr(self._h, k, ctypes.byref(p)))
        return ctypes.cast(p, ctypes.c_void_p).value or 0

    @staticmethod
    def _create(ctx, d, modes, ranks, row_modes=None, col_modes=None):
        ip = ctypes.POINTER(ctypes.c_int)
        h = ctypes.c_void_p()
        rk = np.ascontiguousarray(ranks, np.int32)
        if row_modes is None:
            md = np.ascontiguousarray(modes, np.int32)
            ctx.check(ctx.lib.tta_tt_create(ctx.h, d, md.ctypes.data_as(ip), rk.ctypes.data_as(ip), ctypes.byref(h)))
        else:
            rm = np.ascontiguousarray(row_modes, np.int32)
            cm = np.ascontiguousarray(col_modes, np.int32)
            ctx.check(ctx.lib.tta_ttm_create(ctx.h, d, rm.ctypes.data_as(ip), cm.ctypes.data_as(ip),
                                             rk.ctypes.data_as(ip), ctypes.byref(h)))
        return h

    def _upload(self, flat):
        flat = np.ascontiguousarray(flat, np.float64)
        self._ctx.check(self._ctx.lib.tta_tt_upload(self._ctx.h, self._h,
                                                    flat.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))


class TensorTrain(_DeviceTT):
    """A TT vector resident on the GPU; cores (r_k, n_k, r_{k+1}) row-major."""

    @classmethod
    def from_cores(cls, cores, ctx: Context | None = None) -> "TensorTrain":
        ctx = ctx or Context.default()
        cores = [np.ascontiguousarray(c, np.float64) for c in cores]
        d = len(cores)
        modes = [c.shape[1] for c in cores]
        ranks = [c.shape[0] for c in cores] + [cores[-1].shape[2]] if d else [1]
        h = cls._create(ctx, d, modes, ranks)
        t = cls(h, ctx)
        if d:
            t._upload(np.concatenate([c.ravel() for c in cores]))
        return t

    @property
    def shape(self):
        return [int(x) for x in self._shape_arrays()[0]]

    @property
    def cores(self):
        modes, _, _, rk = self._shape_arrays()
        flat = self._flat()
        out, off = [], 0
        for k in range(len(modes)):
            sz = int(rk[k] * modes[k] * rk[k + 1])
            out.append(flat[off:off + sz].reshape(int(rk[k]), int(modes[k]), int(rk[k + 1])))
            off += sz
        return out

    def to_dense(self) -> np.ndarray:
        """Dense reconstruction on the host (test helper, tt.hpp:145-198)."""
        cores = self.cores
        if not cores:
            return np.ones(1)
        T = cores[0].reshape(-1, cores[0].shape[2])
        for c in cores[1:]:
            T = (T @ c.reshape(c.shape[0], -1)).reshape(-1, c.shape[2])
        return T.ravel()

    def round(self, opts=None, **kwargs) -> "TensorTrain":
        return round(self, RoundOptions.merge(opts, **kwargs))

    def round_ex(self, opts=None, **kwargs):
        return round_ex(self, RoundOptions.merge(opts, **kwargs))

    def right_orthogonalize(self) -> "TensorTrain":
        """Return a copy with cores[1..d-1] right-orthogonal."""
        out = self.copy()
        right_orthogonalize(out)
        return out

    def copy(self):
        h = ctypes.c_void_p()
        self._ctx.check(self._ctx.lib.tta_tt_copy(self._ctx.h, self._h, ctypes.byref(h)))
        return type(self)(h, self._ctx)

    def norm(self) -> float:
        return norm(self)

    def dot(self, other) -> float:
        return inner(self, other)


class TensorTrainMatrix(_DeviceTT):
    """A TT-matrix on the GPU; cores (r_k, m_k, n_k, r_{k+1}) row-major."""

    @classmethod
    def from_cores(cls, cores, ctx: Context | None = None) -> "TensorTrainMatrix":
        ctx = ctx or Context.default()
        cores = [np.ascontiguousarray(c, np.float64) for c in cores]
        d = len(cores)
        ranks = [c.shape[0] for c in cores] + [cores[-1].shape[3]] if d else [1]
        h = cls._create(ctx, d, None, ranks, [c.shape[1] for c in cores], [c.shape[2] for c in cores])
        t = cls(h, ctx)
        if d:
            t._upload(np.concatenate([c.ravel() for c in cores]))
        return t

    @property
    def row_shape(self):
        return [int(x) for x in self._shape_arrays()[1]]

    @property
    def col_shape(self):
        return [int(x) for x in self._shape_arrays()[2]]

    @property
    def cores(self):
        _, rm, cm, rk = self._shape_arrays()
        flat = self._flat()
        out, off = [], 0
        for k in range(len(rm)):
            sz = int(rk[k] * rm[k] * cm[k] * rk[k + 1])
            out.append(flat[off:off + sz].reshape(int(rk[k]), int(rm[k]), int(cm[k]), int(rk[k + 1])))
            off += sz
        return out

    def round(self, opts=None, **kwargs) -> "TensorTrainMatrix":
        return round(self, RoundOptions.merge(opts, **kwargs))

    def matvec(self, x):
        return matvec(self, x)

    def matvec_round(self, x, opts=None, **kwargs):
        return matvec_round(self, x, RoundOptions.merge(opts, **kwargs))

    def matmat(self, b):
        return matmat(self, b)

    def matmat_round(self, b, opts=None, **kwargs):
        return matmat_round(self, b, RoundOptions.merge(opts, **kwargs))

    def right_orthogonalize(self):
        out = self.copy()
        right_orthogonalize(out)
        return out

    def copy(self):
        h = ctypes.c_void_p()
        self._ctx.check(self._ctx.lib.tta_tt_copy(self._ctx.h, self._h, ctypes.byref(h)))
        return type(self)(h, self._ctx)


def _out(ctx, fn, *args, cls=TensorTrain):
    h = ctypes.c_void_p()
    ctx.check(fn(ctx.h, *args, ctypes.byref(h)))
    return cls(h, ctx)


def _iterative(ctx, fn, args, o, cls):
    h = ctypes.c_void_p()
    it, cv = ctypes.c_int(), ctypes.c_int()
    ch = ctypes.c_double()
    warm = o.warm_start._h if o.warm_start is not None else None
    ctx.check(fn(ctx.h, *args, ctypes.c_int(o.max_iters), ctypes.c_double(o.tol), ctypes.c_ulonglong(o.seed), warm,
                 ctypes.byref(h), ctypes.byref(it), ctypes.byref(ch), ctypes.byref(cv)))
    return cls(h, ctx), RoundResult(it.value, ch.value, bool(cv.value))


def round_ex(a, opts=None, **kwargs):
    """round(a, round_options, round_result*) (include/core/round/round.hpp:230-425):
    returns (rounded, RoundResult)."""
    o = RoundOptions.merge(opts, **kwargs)
    return _iterative(a._ctx, a._ctx.lib.tta_round_iterative,
                      (a._h, ctypes.c_double(o.eps), o.max_rank, o.method, o.gauge), o, type(a))


def round(a, opts=None, **kwargs):  # noqa: A001 - the reference's name
    """round(a, round_options) (include/core/round/round.hpp:230-425)."""
    return round_ex(a, opts, **kwargs)[0]


def right_orthogonalize(a) -> None:
    """In place (tt_orthogonalize.hpp:26-36, tt_matrix_orthogonalize.hpp:34-85)."""
    a._ctx.check(a._ctx.lib.tta_right_orthogonalize(a._ctx.h, a._h))


def left_orthogonalize(a) -> None:
    a._ctx.check(a._ctx.lib.tta_left_orthogonalize(a._ctx.h, a._h))


def matvec(A, x):
    return _out(A._ctx, A._ctx.lib.tta_matvec, A._h, x._h)


def matmat(A, B):
    return _out(A._ctx, A._ctx.lib.tta_matmat, A._h, B._h, cls=TensorTrainMatrix)


def matvec_round_ex(A, x, opts=None, **kwargs):
    """matvec_round(A, x, round_options, round_result*) (include/core/round/matvec_round.hpp:42-91)."""
    o = RoundOptions.merge(opts, **kwargs)
    return _iterative(A._ctx, A._ctx.lib.tta_matvec_round_iterative,
                      (A._h, x._h, ctypes.c_double(o.eps), o.max_rank, o.method), o, TensorTrain)


def matvec_round(A, x, opts=None, **kwargs):
    """matvec_round (include/core/round/matvec_round.hpp:42-91)."""
    return matvec_round_ex(A, x, opts, **kwargs)[0]


def matmat_round_ex(A, B, opts=None, **kwargs):
    """matmat_round(A, B, round_options, round_result*) (include/core/round/matmat_round.hpp:41-94)."""
    o = RoundOptions.merge(opts, **kwargs)
    return _iterative(A._ctx, A._ctx.lib.tta_matmat_round_iterative,
                      (A._h, B._h, ctypes.c_double(o.eps), o.max_rank, o.method), o, TensorTrainMatrix)


def matmat_round(A, B, opts=None, **kwargs):
    """matmat_round (include/core/round/matmat_round.hpp:41-94)."""
    return matmat_round_ex(A, B, opts, **kwargs)[0]


def inner(a, b) -> float:
    out = ctypes.c_double()
    a._ctx.check(a._ctx.lib.tta_inner(a._ctx.h, a._h, b._h, ctypes.byref(out)))
    return out.value


def norm(a) -> float:
    out = ctypes.c_double()
    a._ctx.check(a._ctx.lib.tta_norm(a._ctx.h, a._h, ctypes.byref(out)))
    return out.value


frob_inner = inner
frob_norm = norm


def frob_norm_apply(A, X) -> float:
    """||A x||_F or ||A B||_F without the product tensor
    (include/core/algebra/tt_matrix_frob_norm_apply.hpp:36-94)."""
    out = ctypes.c_double()
    A._ctx.check(A._ctx.lib.tta_frob_norm_apply(A._ctx.h, A._h, X._h, ctypes.byref(out)))
    return out.value


def axpby(alpha, a, beta, b):
    return _out(a._ctx, a._ctx.lib.tta_axpby, ctypes.c_double(alpha), a._h, ctypes.c_double(beta), b._h,
                cls=type(a))


def add(a, b):
    return axpby(1.0, a, 1.0, b)


def sub(a, b):
    return axpby(1.0, a, -1.0, b)


def scale(a, alpha):
    return _out(a._ctx, a._ctx.lib.tta_scale, a._h, ctypes.c_double(alpha), cls=type(a))


def random_host(shape, max_rank, seed):
    """random(shape, max_rank, seed) cores on the host (tt_factory.hpp:141-207);
    needs no GPU."""
    L = _capi.lib()
    d = len(shape)
    md = np.ascontiguousarray(shape, np.int32)
    rk = np.zeros(d + 1, np.int32)
    n = ctypes.c_longlong()
    ip = ctypes.POINTER(ctypes.c_int)
    st = L.tta_random_host(d, md.ctypes.data_as(ip), max_rank, seed, rk.ctypes.data_as(ip), None, ctypes.byref(n))
    if st != 0:
        raise TTAError(st, "random: invalid shape")
    data = np.zeros(max(n.value, 1))
    L.tta_random_host(d, md.ctypes.data_as(ip), max_rank, seed, None,
                      data.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), None)
    out, off = [], 0
    for k in range(d):
        sz = int(rk[k] * md[k] * rk[k + 1])
        out.append(data[off:off + sz].reshape(int(rk[k]), int(md[k]), int(rk[k + 1])))
        off += sz
    return out


def random(shape, max_rank, seed, ctx: Context | None = None) -> TensorTrain:  # noqa: A001
    ctx = ctx or Context.default()
    md = np.ascontiguousarray(shape, np.int32)
    return _out(ctx, ctx.lib.tta_random, len(shape), md.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), max_rank,
                ctypes.c_ulonglong(seed))


def random_matrix(row_shape, col_shape, max_rank, seed, ctx: Context | None = None) -> TensorTrainMatrix:
    ctx = ctx or Context.default()
    rm = np.ascontiguousarray(row_shape, np.int32)
    cm = np.ascontiguousarray(col_shape, np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    return _out(ctx, ctx.lib.tta_random_matrix, len(rm), rm.ctypes.data_as(ip), cm.ctypes.data_as(ip), max_rank,
                ctypes.c_ulonglong(seed), cls=TensorTrainMatrix)


def _constant(shape, value, ctx):
    ctx = ctx or Context.default()
    md = np.ascontiguousarray(shape, np.int32)
    return _out(ctx, ctx.lib.tta_constant, len(shape), md.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                ctypes.c_double(value))


def zeros(shape, ctx: Context | None = None) -> TensorTrain:
    return _constant(shape, 0.0, ctx)


def ones(shape, ctx: Context | None = None) -> TensorTrain:
    return _constant(shape, 1.0, ctx)


def soft_threshold(a, tau=0.0):
    """soft_threshold(a, tau) (include/core/round/tt_soft_threshold.hpp:42-176)."""
    return _out(a._ctx, a._ctx.lib.tta_soft_threshold, a._h, ctypes.c_double(tau))


def from_dense(dense, shape=None, eps=1e-10, max_rank=0, method="svd", qtt_base=2,
               ctx: Context | None = None) -> TensorTrain:
    """TT of a dense tensor on the GPU: from_dense with from_dense_options
    {eps, max_rank, method ("svd" or "qtt"), qtt_base}
    (include/core/factory/from_dense.hpp:215-225).  ``shape`` defaults to
    dense.shape; the data is read row-major."""
    ctx = ctx or Context.default()
    arr = np.ascontiguousarray(dense, np.float64)
    sh = np.ascontiguousarray(arr.shape if shape is None else shape, np.int32)
    if int(np.prod(sh)) != arr.size:
        raise ValueError(f"from_dense: shape {list(sh)} does not match {arr.size} values")
    m = {"svd": 0, "qtt": 1}[method]
    return _out(ctx, ctx.lib.tta_from_dense, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(sh),
                sh.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.c_double(eps), max_rank, m, qtt_base)


def matrix_from_dense(dense, row_shape, col_shape, eps=1e-10, max_rank=0,
                      ctx: Context | None = None) -> TensorTrainMatrix:
    """TT-matrix of a dense (prod m) x (prod n) matrix on the GPU
    (matrix_from_dense, include/core/factory/tt_matrix_factory.hpp:205-290)."""
    ctx = ctx or Context.default()
    arr = np.ascontiguousarray(dense, np.float64)
    rm = np.ascontiguousarray(row_shape, np.int32)
    cm = np.ascontiguousarray(col_shape, np.int32)
    if int(np.prod(rm)) * int(np.prod(cm)) != arr.size:
        raise ValueError("matrix_from_dense: shapes do not match the data")
    ip = ctypes.POINTER(ctypes.c_int)
    return _out(ctx, ctx.lib.tta_matrix_from_dense, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(rm),
                rm.ctypes.data_as(ip), cm.ctypes.data_as(ip), ctypes.c_double(eps), max_rank, cls=TensorTrainMatrix)


def scale_core(t, k: int, alpha: float) -> None:
    """Multiply core k of t by alpha in place."""
    t._ctx.check(t._ctx.lib.tta_scale_core(t._ctx.h, t._h, k, ctypes.c_double(alpha)))


def qr_thin(M, ctx: Context | None = None):
    """Thin Householder QR on the GPU (qr_thin, tt_eigen_bridge.hpp:42-60)."""
    ctx = ctx or Context.default()
    M = np.ascontiguousarray(M, np.float64)
    rows, cols = M.shape
    k = min(rows, cols)
    Q = np.zeros((rows, k))
    R = np.zeros((k, cols))
    dp = ctypes.POINTER(ctypes.c_double)
    ctx.check(ctx.lib.tta_qr_thin(ctx.h, M.ctypes.data_as(dp), rows, cols, Q.ctypes.data_as(dp),
                                  R.ctypes.data_as(dp)))
    return Q, R


def svd_thin(M, ctx: Context | None = None):
    """Thin SVD on the GPU (svd_thin, tt_eigen_bridge.hpp:67-104)."""
    ctx = ctx or Context.default()
    M = np.ascontiguousarray(M, np.float64)
    rows, cols = M.shape
    k = min(rows, cols)
    U = np.zeros((rows, k))
    s = np.zeros(k)
    Vt = np.zeros((k, cols))
    dp = ctypes.POINTER(ctypes.c_double)
    ctx.check(ctx.lib.tta_svd_thin(ctx.h, M.ctypes.data_as(dp), rows, cols, U.ctypes.data_as(dp),
                                   s.ctypes.data_as(dp), Vt.ctypes.data_as(dp)))
    return U, s, Vt


PRECISIONS = {"fp64": 0, "mixed": 1, "learned": 2}


def round_batched(tts, eps=0.0, max_rank=0, precision="fp64"):
    """Round a list of device TTs in one batched call (tta_round_batched);
    precision "fp64" (plain) or "mixed" (rank-revealing CholQR2 steps)."""
    if not tts:
        return []
    ctx = tts[0]._ctx
    n = len(tts)
    arr_in = (ctypes.c_void_p * n)(*[t._h.value for t in tts])
    arr_out = (ctypes.c_void_p * n)()
    ctx.check(ctx.lib.tta_round_batched(ctx.h, n, arr_in, ctypes.c_double(eps), max_rank,
                                        _enum(PRECISIONS, precision, "precision"), arr_out))
    return [type(tts[i])(ctypes.c_void_p(arr_out[i]), ctx) for i in range(n)]


def build_sum_batch(d, n, ru, rv, seed_u, seed_v, ctx: Context | None = None):
    """c_i = add(u_i, add(scale(u_i, 0.5), v_i)) built on the device (Philox N(0,1) entries)."""
    ctx = ctx or Context.default()
    cnt = len(d)
    ip = ctypes.POINTER(ctypes.c_int)
    arrs = [np.ascontiguousarray(a, np.int32) for a in (d, n, ru, rv)]
    su = (ctypes.c_ulonglong * cnt)(*[int(x) for x in seed_u])
    sv = (ctypes.c_ulonglong * cnt)(*[int(x) for x in seed_v])
    out = (ctypes.c_void_p * cnt)()
    ctx.check(ctx.lib.tta_build_sum_batch(ctx.h, cnt, *[a.ctypes.data_as(ip) for a in arrs], su, sv, out))
    return [TensorTrain(ctypes.c_void_p(out[i]), ctx) for i in range(cnt)]


def build_decay_batch(count, d, n, rmax, s, seeds, ctx: Context | None = None):
    """Decaying-spectrum TTs (cores Q_k diag(10^(-j s_i)), unit-norm last core) built on the device."""
    ctx = ctx or Context.default()
    sv = np.ascontiguousarray(s, np.float64)
    sd = (ctypes.c_ulonglong * count)(*[int(x) for x in seeds])
    out = (ctypes.c_void_p * count)()
    ctx.check(ctx.lib.tta_build_decay_batch(ctx.h, count, d, n, rmax, sv.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                            sd, out))
    return [TensorTrain(ctypes.c_void_p(out[i]), ctx) for i in range(count)]


_FALLBACK = {"none": 0, "fp64": 1, "mixed": 2}


def _fallback_code(fallback):
    if fallback is True:
        return 1
    if fallback is False:
        return 0
    return _enum(_FALLBACK, fallback, "fallback")


def round_randomized_batched(tts, ranks, eps, seed=1, fallback="fp64"):
    """Randomized rounding with target ranks (list per TT of d+1 ints);
    TTs whose discarded norm exceeds eps are rounded again by the
    deterministic path `fallback` names ("fp64" plain, "mixed", or "none" to
    keep the randomized result; True / False mean "fp64" / "none").  Returns
    (outputs, rel_err array, fell_back array)."""
    if not tts:
        return [], np.zeros(0), np.zeros(0, np.int32)
    ctx = tts[0]._ctx
    n = len(tts)
    flat = np.ascontiguousarray(np.concatenate([np.asarray(r, np.int32) for r in ranks]), np.int32)
    arr_in = (ctypes.c_void_p * n)(*[t._h.value for t in tts])
    arr_out = (ctypes.c_void_p * n)()
    err = np.zeros(n)
    fb = np.zeros(n, np.int32)
    ctx.check(ctx.lib.tta_round_randomized_batched(ctx.h, n, arr_in, flat.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                   ctypes.c_double(eps), ctypes.c_ulonglong(seed),
                                                   _fallback_code(fallback),
                                                   arr_out, err.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                   fb.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
    return [type(tts[i])(ctypes.c_void_p(arr_out[i]), ctx) for i in range(n)], err, fb


def _f32(a):
    a = np.ascontiguousarray(a, np.float32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class RankMLP:
    """Rank-prediction MLP on the GPU (tta_mlp_*): n_in -> hidden x layers -> 1."""

    def __init__(self, n_in, hidden=64, layers=3, seed=0, ctx: Context | None = None):
        self._ctx = ctx or Context.default()
        h = ctypes.c_void_p()
        self._ctx.check(self._ctx.lib.tta_mlp_create(self._ctx.h, n_in, hidden, layers, seed, ctypes.byref(h)))
        self._h = h
        self.n_in = n_in

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._ctx.lib.tta_mlp_free(self._ctx.h, h)
            except Exception:
                pass
            self._h = None

    @property
    def num_params(self):
        return int(self._ctx.lib.tta_mlp_num_params(self._h))

    def get_params(self):
        out = np.zeros(self.num_params, np.float32)
        self._ctx.check(self._ctx.lib.tta_mlp_get_params(self._ctx.h, self._h, _f32(out)[1]))
        return out

    def set_params(self, p):
        a, ptr = _f32(p)
        self._ctx.check(self._ctx.lib.tta_mlp_set_params(self._ctx.h, self._h, ptr))

    def grad(self, X, y):
        Xa, Xp = _f32(X)
        ya, yp = _f32(y)
        g = np.zeros(self.num_params, np.float32)
        loss = ctypes.c_float()
        self._ctx.check(self._ctx.lib.tta_mlp_grad(self._ctx.h, self._h, Xp, yp, len(ya), _f32(g)[1],
                                                   ctypes.byref(loss)))
        return loss.value, g

    def train(self, X, y, steps=1000, lr=1e-3):
        Xa, Xp = _f32(X)
        ya, yp = _f32(y)
        loss = ctypes.c_float()
        self._ctx.check(self._ctx.lib.tta_mlp_train(self._ctx.h, self._h, Xp, yp, len(ya), steps, ctypes.c_float(lr),
                                                    ctypes.byref(loss)))
        return loss.value

    def set_bf16_inference(self, on=True):
        """Run inference (predict and the learned rounding's rank prediction)
        with BF16 weights, biases and activations and FP32 accumulation;
        training stays FP32."""
        self._ctx.check(self._ctx.lib.tta_mlp_set_bf16_inference(self._ctx.h, self._h, 1 if on else 0))

    def predict(self, X):
        Xa, Xp = _f32(X)
        out = np.zeros(Xa.shape[0], np.float32)
        self._ctx.check(self._ctx.lib.tta_mlp_predict(self._ctx.h, self._h, Xp, Xa.shape[0], _f32(out)[1]))
        return out


def bond_features(tts, eps):
    """Per-bond features (rows: bonds 1..d-1 of each TT in order)."""
    ctx = tts[0]._ctx
    nf = ctx.lib.tta_num_bond_features()
    nb = sum(t.d - 1 for t in tts)
    out = np.zeros((max(nb, 1), nf), np.float32)
    arr = (ctypes.c_void_p * len(tts))(*[t._h.value for t in tts])
    ctx.check(ctx.lib.tta_bond_features(ctx.h, len(tts), arr, ctypes.c_double(eps), _f32(out)[1]))
    return out[:nb]


def round_learned_batched(mlp, tts, eps, oversample=6, seed=1, slack=0.1, fallback="fp64"):
    """Learned-rank randomized rounding with deterministic fallback ("fp64"
    plain or "mixed"); target rank per bond = ceil(pred * r_k * (1 + slack))
    + oversample.  Returns (outputs, rel_err, fell_back, target ranks per
    TT)."""
    ctx = tts[0]._ctx
    n = len(tts)
    arr_in = (ctypes.c_void_p * n)(*[t._h.value for t in tts])
    arr_out = (ctypes.c_void_p * n)()
    err = np.zeros(n)
    fb = np.zeros(n, np.int32)
    rk = np.zeros(sum(t.d + 1 for t in tts), np.int32)
    ctx.check(ctx.lib.tta_round_learned_batched(ctx.h, mlp._h, n, arr_in, ctypes.c_double(eps), oversample,
                                                ctypes.c_double(slack), ctypes.c_ulonglong(seed),
                                                _fallback_code(fallback), arr_out,
                                                err.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                fb.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                rk.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
    ranks, off = [], 0
    for t in tts:
        ranks.append(rk[off:off + t.d + 1].tolist())
        off += t.d + 1
    return [type(tts[i])(ctypes.c_void_p(arr_out[i]), ctx) for i in range(n)], err, fb, ranks
