"""CPU tests of the boundary: the library loads, exports every entry point
include/tt_amd.h declares, the ctypes table matches the header, option
merging follows the reference wrapper, and the absence of a GPU is an error
(never a silent fallback)."""
import ctypes
import os
import re

import pytest

import tensor_train_amd as T
from tensor_train_amd import _capi

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "tt_amd.h")


def header_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tta_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_capi.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_ctypes_table_covers_header():
    assert sorted(_capi.SIGNATURES) == header_symbols()


def test_round_options_merging():
    o = T.RoundOptions(eps=1e-6, max_rank=4, method="svd", gauge="right_canonical")
    assert (o.eps, o.max_rank, o.method, o.gauge) == (1e-6, 4, 0, 1)
    m = T.RoundOptions.merge(o, eps=1e-3, method="svd_streaming")
    assert (m.eps, m.max_rank, m.method, m.gauge) == (1e-3, 4, 1, 1)
    d = T.RoundOptions.merge({"eps": 1e-4, "gauge": "none"})
    assert (d.eps, d.max_rank, d.method, d.gauge) == (1e-4, 0, 0, 0)
    k = T.RoundOptions.merge(None, max_rank=7)
    assert (k.eps, k.max_rank) == (0.0, 7)
    with pytest.raises(ValueError):
        T.RoundOptions(method="bogus")


def test_no_gpu_is_an_error():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(T.TTAError):
        T.Context(0)


def test_host_generator_size_query_and_errors():
    lib = _capi.lib()
    ip = ctypes.POINTER(ctypes.c_int)
    shape = (ctypes.c_int * 3)(4, 0, 4)
    n = ctypes.c_longlong()
    assert lib.tta_random_host(3, ctypes.cast(shape, ip), 8, 1, None, None, ctypes.byref(n)) == 2
    shape = (ctypes.c_int * 3)(4, 5, 6)
    rk = (ctypes.c_int * 4)()
    assert lib.tta_random_host(3, ctypes.cast(shape, ip), 8, 1, ctypes.cast(rk, ip), None, ctypes.byref(n)) == 0
    # Last bond keeps its forward cap (8), as random() sets r_d = 1 only at the end.
    assert list(rk) == [1, 4, 8, 1] and n.value == 4 * 4 + 4 * 5 * 8 + 8 * 6


def test_null_arguments_rejected():
    lib = _capi.lib()
    assert lib.tta_round(None, None, 0.0, 0, 0, 0, None) == 1
    assert lib.tta_tt_numel(None) == -1
