"""ctypes binding of ``libtt_amd.so`` (the C ABI declared in include/tt_amd.h).

The library is built in-tree into ``tensor_train_amd/_lib/libtt_amd.so`` by
``tensor_train_amd/csrc/Makefile``.  There is no fallback: a missing library
or a missing GPU raises.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libtt_amd.so")
CSRC = os.path.join(_HERE, "csrc")

_i32p = ctypes.POINTER(ctypes.c_int)
_f64p = ctypes.POINTER(ctypes.c_double)
_f32p = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p
_vpp = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes); mirrors include/tt_amd.h.
SIGNATURES = {
    "tta_ctx_create": (ctypes.c_int, [ctypes.c_int, _vpp]),
    "tta_ctx_destroy": (ctypes.c_int, [_vp]),
    "tta_last_error": (ctypes.c_char_p, [_vp]),
    "tta_synchronize": (ctypes.c_int, [_vp]),
    "tta_stream": (_vp, [_vp]),
    "tta_tt_create": (ctypes.c_int, [_vp, ctypes.c_int, _i32p, _i32p, _vpp]),
    "tta_ttm_create": (ctypes.c_int, [_vp, ctypes.c_int, _i32p, _i32p, _i32p, _vpp]),
    "tta_tt_free": (ctypes.c_int, [_vp, _vp]),
    "tta_tt_dims": (ctypes.c_int, [_vp, _i32p, _i32p]),
    "tta_tt_shape": (ctypes.c_int, [_vp, _i32p, _i32p, _i32p, _i32p]),
    "tta_tt_numel": (ctypes.c_longlong, [_vp]),
    "tta_tt_upload": (ctypes.c_int, [_vp, _vp, _f64p]),
    "tta_tt_download": (ctypes.c_int, [_vp, _vp, _f64p]),
    "tta_tt_core_ptr": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_f64p)]),
    "tta_tt_copy": (ctypes.c_int, [_vp, _vp, _vpp]),
    "tta_round": (ctypes.c_int, [_vp, _vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vpp]),
    "tta_right_orthogonalize": (ctypes.c_int, [_vp, _vp]),
    "tta_left_orthogonalize": (ctypes.c_int, [_vp, _vp]),
    "tta_matvec": (ctypes.c_int, [_vp, _vp, _vp, _vpp]),
    "tta_matmat": (ctypes.c_int, [_vp, _vp, _vp, _vpp]),
    "tta_matvec_round": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, _vpp]),
    "tta_matmat_round": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, _vpp]),
    "tta_round_iterative": (ctypes.c_int, [_vp, _vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_double, ctypes.c_ulonglong, _vp, _vpp, _i32p,
                                           _f64p, _i32p]),
    "tta_matvec_round_iterative": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_int, ctypes.c_double, ctypes.c_ulonglong, _vp, _vpp,
                                                  _i32p, _f64p, _i32p]),
    "tta_matmat_round_iterative": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_int, ctypes.c_double, ctypes.c_ulonglong, _vp, _vpp,
                                                  _i32p, _f64p, _i32p]),
    "tta_inner": (ctypes.c_int, [_vp, _vp, _vp, _f64p]),
    "tta_norm": (ctypes.c_int, [_vp, _vp, _f64p]),
    "tta_inner_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "tta_frob_norm_apply": (ctypes.c_int, [_vp, _vp, _vp, _f64p]),
    "tta_axpby": (ctypes.c_int, [_vp, ctypes.c_double, _vp, ctypes.c_double, _vp, _vpp]),
    "tta_scale": (ctypes.c_int, [_vp, _vp, ctypes.c_double, _vpp]),
    "tta_round_batched": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), ctypes.c_double,
                                         ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "tta_round_randomized_batched": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), _i32p,
                                                    ctypes.c_double, ctypes.c_ulonglong, ctypes.c_int,
                                                    ctypes.POINTER(ctypes.c_void_p), _f64p, _i32p]),
    "tta_mlp_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong, _vpp]),
    "tta_mlp_free": (ctypes.c_int, [_vp, _vp]),
    "tta_mlp_num_params": (ctypes.c_longlong, [_vp]),
    "tta_mlp_get_params": (ctypes.c_int, [_vp, _vp, _f32p]),
    "tta_mlp_set_params": (ctypes.c_int, [_vp, _vp, _f32p]),
    "tta_mlp_grad": (ctypes.c_int, [_vp, _vp, _f32p, _f32p, ctypes.c_int, _f32p, _f32p]),
    "tta_mlp_train": (ctypes.c_int, [_vp, _vp, _f32p, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_float, _f32p]),
    "tta_mlp_predict": (ctypes.c_int, [_vp, _vp, _f32p, ctypes.c_int, _f32p]),
    "tta_mlp_set_bf16_inference": (ctypes.c_int, [_vp, _vp, ctypes.c_int]),
    "tta_num_bond_features": (ctypes.c_int, []),
    "tta_bond_features": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), ctypes.c_double, _f32p]),
    "tta_round_learned_batched": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
                                                 ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_ulonglong,
                                                 ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), _f64p, _i32p, _i32p]),
    "tta_qr_thin": (ctypes.c_int, [_vp, _f64p, ctypes.c_int, ctypes.c_int, _f64p, _f64p]),
    "tta_svd_thin": (ctypes.c_int, [_vp, _f64p, ctypes.c_int, ctypes.c_int, _f64p, _f64p, _f64p]),
    "tta_selfcheck_wave_sum": (ctypes.c_int, [_vp, _f64p, ctypes.c_int, _i32p]),
    "tta_random_host": (ctypes.c_int, [ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_ulonglong, _i32p, _f64p,
                                       ctypes.POINTER(ctypes.c_longlong)]),
    "tta_random": (ctypes.c_int, [_vp, ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_ulonglong, _vpp]),
    "tta_random_matrix": (ctypes.c_int, [_vp, ctypes.c_int, _i32p, _i32p, ctypes.c_int, ctypes.c_ulonglong, _vpp]),
    "tta_constant": (ctypes.c_int, [_vp, ctypes.c_int, _i32p, ctypes.c_double, _vpp]),
    "tta_build_sum_batch": (ctypes.c_int, [_vp, ctypes.c_int, _i32p, _i32p, _i32p, _i32p,
                                           ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_ulonglong),
                                           ctypes.POINTER(ctypes.c_void_p)]),
    "tta_build_decay_batch": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p,
                                             ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_void_p)]),
    "tta_scale_core": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_double]),
    "tta_prof_enable": (ctypes.c_int, [_vp, ctypes.c_int]),
    "tta_prof_read": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), _f64p, _f64p, _f64p]),
    "tta_set_cholqr": (ctypes.c_int, [_vp, ctypes.c_int]),
    "tta_set_graphs": (ctypes.c_int, [_vp, ctypes.c_int]),
    "tta_set_jacobi_gram": (ctypes.c_int, [_vp, ctypes.c_int]),
    "tta_set_autotune": (ctypes.c_int, [_vp, ctypes.c_int]),
    "tta_prof_autotune": (ctypes.c_int, [_vp] + [ctypes.POINTER(ctypes.c_longlong)] * 3),
    "tta_set_speculation": (ctypes.c_int, [_vp, ctypes.c_int]),
    "tta_set_chain": (ctypes.c_int, [_vp, ctypes.c_int]),
    "tta_prof_chain": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                                      ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]),
    "tta_prof_cholqr_triangular": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_longlong)]),
    "tta_prof_speculation": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]),
    "tta_prof_cholqr": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]),
    "tta_set_rank_certificate": (ctypes.c_int, [_vp, ctypes.c_int]),
    "tta_prof_rank_certificates": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_longlong),
                                                  ctypes.POINTER(ctypes.c_longlong)]),
    "tta_prof_jacobi_sweeps": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]),
    "tta_prof_jacobi_unconverged": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_longlong)]),
    "tta_prof_rank_revealing": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_longlong),
                                               ctypes.POINTER(ctypes.c_longlong)]),
    "tta_soft_threshold": (ctypes.c_int, [_vp, _vp, ctypes.c_double, ctypes.POINTER(ctypes.c_void_p)]),
    "tta_from_dense": (ctypes.c_int, [_vp, _f64p, ctypes.c_int, _i32p, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "tta_matrix_from_dense": (ctypes.c_int, [_vp, _f64p, ctypes.c_int, _i32p, _i32p, ctypes.c_double, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_void_p)]),
    "tta_mem_stats": (ctypes.c_int, [_vp] + [ctypes.POINTER(ctypes.c_ulonglong)] * 4),
    "tta_mem_trim": (ctypes.c_int, [_vp]),
}

KERNEL_CLASSES = {"qr": 0, "jacobi": 1, "gemm": 2, "apply": 3, "other": 4, "jacobi_round": 5, "chol": 6, "rinv": 7,
                  "gemm_splitk": 8, "gemm_multi": 9, "chol_step": 10, "gemm_chain": 11,
                  "gemm_grouped": 12}

_lib = None


def build() -> None:
    """Compile libtt_amd.so for gfx950 with hipcc (no GPU needed)."""
    subprocess.run(["make", "-s", "-B", "-j8", "-C", CSRC], check=True)


def lib():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libtt_amd.so not built: {LIB_PATH} (run tensor_train_amd._capi.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class TTAError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[tta status {code}] {msg}")
        self.code = code


STATUS = {0: "OK", 1: "ARG", 2: "SHAPE", 3: "METHOD", 4: "HIP", 5: "NOMEM", 6: "UNSUPPORTED"}
