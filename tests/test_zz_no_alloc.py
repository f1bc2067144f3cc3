"""tta_inner_noalloc: inner(a, b) with its environments in the library's fixed
device scratch array and the result in a caller-provided device double.
Checked against tta_inner on two calls in a row (the second with smaller
operands, so it runs on what the first left in the array) to the
floating-point bar of the inner-product tests (relative to |a| |b|), with the
device pool's counters unchanged by the calls; operands beyond the array are
rejected with TTA_ERR_UNSUPPORTED."""
import ctypes

import numpy as np
import pytest

import tensor_train_amd as T

pytestmark = pytest.mark.gpu


def _out_buffer():
    """A one-element device tensor whose core is the caller's output double."""
    return T.TensorTrain.from_cores([np.zeros((1, 1, 1))])


def _noalloc(ctx, a, b, buf):
    st = ctx.lib.tta_inner_noalloc(ctx.h, a._h, b._h, ctypes.c_void_p(buf.core_ptr(0)))
    ctx.check(st)
    ctx.synchronize()
    return float(buf.cores[0].ravel()[0])


def test_inner_noalloc_matches_inner():
    ctx = T.Context.default(0)
    buf = _out_buffer()
    cases = [([4] * 12, 40, 24, 11), ([3, 2, 5], 3, 2, 12)]  # the second is smaller
    for shape, ra, rb, seed in cases:
        a = T.random(shape, ra, seed)
        b = T.random(shape, rb, seed + 100)
        ref = T.inner(a, b)
        before = ctx.mem_stats()
        got = _noalloc(ctx, a, b, buf)
        after = ctx.mem_stats()
        assert after["device_allocs"] == before["device_allocs"]
        assert after["live_bytes"] == before["live_bytes"]
        scale = T.norm(a) * T.norm(b)
        assert abs(got - ref) <= 1e-12 * scale, (shape, got, ref)
        # inner(a, a) = ||a||^2 through the same entry point.
        aa = _noalloc(ctx, a, a, buf)
        assert abs(aa - T.norm(a) ** 2) <= 1e-12 * T.norm(a) ** 2


def test_inner_noalloc_rejects_beyond_scratch():
    ctx = T.Context.default(0)
    buf = _out_buffer()
    # Ranks 1024 with 8 modes: 2 * 1024^2 + 1024 * 8 * 1024 doubles > 4 Mi.
    a = T.TensorTrain.from_cores([np.ones((1, 8, 1024)), np.ones((1024, 8, 1024)), np.ones((1024, 8, 1))])
    st = ctx.lib.tta_inner_noalloc(ctx.h, a._h, a._h, ctypes.c_void_p(buf.core_ptr(0)))
    assert st == 6  # TTA_ERR_UNSUPPORTED
    # A shape mismatch is an argument error, as tta_inner's.
    b = T.random([8, 8, 8], 4, 3)
    c = T.random([8, 8, 4], 4, 3)
    assert ctx.lib.tta_inner_noalloc(ctx.h, b._h, c._h, ctypes.c_void_p(buf.core_ptr(0))) == 2
