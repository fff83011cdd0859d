"""The allocation-free dot product (ganit_hip_dot_noalloc) against the context entry point ganit_hip_dot: the
same bits on two calls in a row, the second on a smaller input, in FP64 and FP32, on the context's stream and
on the null stream."""
import ctypes
import ctypes.util

import numpy as np
import pytest

from ganit_v2_amd import hip

pytestmark = pytest.mark.gpu


def _dot_ctx(ctx, vdt, n, a, b):
    out = ctypes.c_double()
    hip.check(hip.lib().ganit_hip_dot(ctx.h, vdt, n, a.ptr, b.ptr, ctypes.byref(out)), "dot")
    return out.value


def _device_synchronize():
    hip.lib()  # the HIP runtime is loaded with the library
    rt = ctypes.CDLL(ctypes.util.find_library("amdhip64") or "libamdhip64.so")
    assert rt.hipDeviceSynchronize() == 0


def _dot_noalloc(ctx, stream, vdt, n, a, b, out_dev):
    hip.check(hip.lib().ganit_hip_dot_noalloc(stream, vdt, n, a.ptr, b.ptr, out_dev.ptr), "dot_noalloc")
    _device_synchronize()  # the null stream does not order with the context's non-blocking stream
    return out_dev.numpy()[0]


@pytest.mark.parametrize("vdt", [hip.F64, hip.F32])
def test_dot_noalloc_matches_dot_on_two_calls(hipctx, vdt):
    rng = np.random.default_rng(17)
    out_dev = hipctx.empty(1, hip.F64)
    for n in (1_000_003, 70_001):  # the second call is smaller: stale partials must not leak in
        a_h, b_h = rng.standard_normal(n), rng.standard_normal(n)
        a, b = hipctx.vector(a_h, vdt), hipctx.vector(b_h, vdt)
        want = _dot_ctx(hipctx, vdt, n, a, b)
        got = _dot_noalloc(hipctx, hipctx.stream(), vdt, n, a, b, out_dev)
        assert got == want, (n, got, want)
        ar, br = a.numpy(), b.numpy()  # the values as stored (FP32-rounded for F32)
        assert abs(got - float(ar @ br)) <= 1e-13 * n * (np.abs(ar) @ np.abs(br))


def test_dot_noalloc_null_stream_and_edge_sizes(hipctx):
    out_dev = hipctx.empty(1, hip.F64)
    for n in (0, 1, 255, 256, 257, 262_145):
        a_h = np.arange(n, dtype=np.float64) % 7 - 3.0
        b_h = np.ones(n)
        a, b = hipctx.vector(a_h, hip.F64), hipctx.vector(b_h, hip.F64)
        got = _dot_noalloc(hipctx, None, hip.F64, n, a, b, out_dev)
        assert got == _dot_ctx(hipctx, hip.F64, n, a, b) == float(a_h.sum())  # integer-valued: exact


def test_dot_noalloc_rejects_bad_arguments(hipctx):
    L = hip.lib()
    out_dev = hipctx.empty(1, hip.F64)
    a = hipctx.vector(np.ones(4), hip.F64)
    assert L.ganit_hip_dot_noalloc(None, hip.F64, -1, a.ptr, a.ptr, out_dev.ptr) == 1
    assert L.ganit_hip_dot_noalloc(None, hip.F64, 4, a.ptr, a.ptr, None) == 1
    assert L.ganit_hip_dot_noalloc(None, 7, 4, a.ptr, a.ptr, out_dev.ptr) == 1
    assert b"dot" in L.ganit_hip_last_error() or b"dtype" in L.ganit_hip_last_error()
