"""ctypes view of libganit_hip.so (include/ganit_hip.h): device context, buffers, matrices, the
backend-concept kernels and the device-resident PDHG object.  Used by the tests, the benchmark and the
autograd layer; every call goes straight to the native library (there is no fallback)."""
from __future__ import annotations

import ctypes
import os
import weakref

import numpy as np

from . import _load

F64, F32, BF16 = 0, 1, 2
CHECK_AVG, CHECK_RAY, CHECK_FP64, CHECK_KAHAN, CHECK_DRAY, CHECK_NOUT = 1, 2, 4, 8, 16, 22
_NP = {F64: np.float64, F32: np.float32}
_SIZE = {F64: 8, F32: 4, BF16: 2}

_lib = None
_vp = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_d = ctypes.c_double


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    L = _load("libganit_hip.so")
    L.ganit_hip_last_error.restype = ctypes.c_char_p
    sig = {
        "ganit_hip_ctx_create": [_i, ctypes.POINTER(_vp)],
        "ganit_hip_ctx_destroy": [_vp],
        "ganit_hip_ctx_sync": [_vp],
        "ganit_hip_ctx_stream": [_vp, ctypes.POINTER(_vp)],
        "ganit_hip_ctx_device_info": [_vp, ctypes.c_char_p, _i, ctypes.POINTER(_i)],
        "ganit_hip_alloc": [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)],
        "ganit_hip_free": [_vp, _vp],
        "ganit_hip_upload": [_vp, _i, _vp, _i64, _vp],
        "ganit_hip_download": [_vp, _i, _vp, _i64, _vp],
        "ganit_hip_vec_fill": [_vp, _i, _vp, _i64, _d],
        "ganit_hip_vec_copy": [_vp, _i, _vp, _vp, _i64],
        "ganit_hip_vec_sub": [_vp, _i, _vp, _vp, _vp, _i64],
        "ganit_hip_vec_scale_into": [_vp, _i, _vp, _d, _vp, _i64],
        "ganit_hip_vec_convert": [_vp, _i, _vp, _i, _vp, _i64],
        "ganit_hip_csr_upload": [_vp, _i, _i, _vp, _vp, _vp, ctypes.c_uint, ctypes.POINTER(_vp)],
        "ganit_hip_csr_free": [_vp],
        "ganit_hip_csr_plan": [_vp, ctypes.POINTER(_i), _vp],
        "ganit_hip_csr_set_lanes": [_vp, _i],
        "ganit_hip_csr_autotune": [_vp, _vp, _i, _i, _i],
        "ganit_hip_spmv": [_vp, _vp, _i, _i, _vp, _vp],
        "ganit_hip_primal_step": [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _d, _vp, _vp, _vp],
        "ganit_hip_dual_step": [_vp, _i, _i, _vp, _vp, _vp, _vp, _d, _vp, _vp],
        "ganit_hip_primal_res2": [_vp, _i, _i, _vp, _vp, _vp, _vp, ctypes.POINTER(_d)],
        "ganit_hip_dual_res": [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, ctypes.POINTER(_d), ctypes.POINTER(_d)],
        "ganit_hip_row_dual_obj": [_vp, _i, _i, _vp, _vp, _vp, ctypes.POINTER(_d)],
        "ganit_hip_dot": [_vp, _i, _i64, _vp, _vp, ctypes.POINTER(_d)],
        "ganit_hip_diff2": [_vp, _i, _i64, _vp, _vp, ctypes.POINTER(_d)],
        "ganit_hip_dot_noalloc": [_vp, _i, _i64, _vp, _vp, _vp],
        "ganit_hip_pdhg_create": [_vp, _vp, ctypes.c_uint, ctypes.c_uint, ctypes.POINTER(_vp)],
        "ganit_hip_pdhg_destroy": [_vp],
        "ganit_hip_pdhg_plan": [_vp, _i, ctypes.POINTER(_i), _vp],
        "ganit_hip_pdhg_set_lanes": [_vp, _i, _i],
        "ganit_hip_pdhg_window_layout": [_vp, _i, ctypes.POINTER(_i)],
        "ganit_hip_pdhg_load": [_vp, _i, _vp, _vp, _vp, _vp],
        "ganit_hip_pdhg_handover": [_vp, _i, _i],
        "ganit_hip_pdhg_iterate": [_vp, _i, _i, _i, _d, _d],
        "ganit_hip_pdhg_check": [_vp, _i, _i, _i, _d, _vp],
        "ganit_hip_pdhg_restart": [_vp, _i, _i],
        "ganit_hip_pdhg_adopt_average": [_vp, _i],
        "ganit_hip_pdhg_store": [_vp, _i, _vp, _vp, _vp, _vp],
        "ganit_hip_pdhg_time_kernels": [_vp, _i, _i, _i, _d, _d, _vp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    _lib = L
    return L


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        raise RuntimeError(f"{what}: {lib().ganit_hip_last_error().decode()} (status {rc})")


def exported_symbols() -> list[str]:
    """The ganit_hip_* function names declared in include/ganit_hip.h."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    import re
    text = open(os.path.join(root, "include", "ganit_hip.h")).read()
    return sorted(set(re.findall(r"\b(ganit_hip_\w+)\s*\(", text)))


def _np_ptr(a: np.ndarray):
    return a.ctypes.data_as(_vp)


class Context:
    """A device context.  Matrices and PDHG objects made from it are released before the context itself
    when it closes, so no handle outlives the stream it was created on."""

    def __init__(self, device: int = 0):
        h = _vp()
        check(lib().ganit_hip_ctx_create(device, ctypes.byref(h)), "ctx_create")
        self.h = h
        self._bufs = []
        self._children = weakref.WeakSet()

    def _adopt(self, obj) -> None:
        self._children.add(obj)

    def close(self):
        if self.h:
            for child in list(self._children):
                child.release()
            for b in self._bufs:
                lib().ganit_hip_free(self.h, b)
            self._bufs = []
            lib().ganit_hip_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self):
        buf = ctypes.create_string_buffer(256)
        cu = _i()
        check(lib().ganit_hip_ctx_device_info(self.h, buf, 256, ctypes.byref(cu)), "device_info")
        return buf.value.decode(), cu.value

    def sync(self):
        check(lib().ganit_hip_ctx_sync(self.h), "sync")

    def stream(self) -> int:
        s = _vp()
        check(lib().ganit_hip_ctx_stream(self.h, ctypes.byref(s)), "stream")
        return s.value or 0

    def alloc(self, nbytes: int) -> int:
        p = _vp()
        check(lib().ganit_hip_alloc(self.h, max(int(nbytes), 1), ctypes.byref(p)), "alloc")
        self._bufs.append(p.value)
        return p.value

    def vector(self, host: np.ndarray, vdt: int = F64) -> "DeviceVector":
        host = np.ascontiguousarray(host, dtype=np.float64)
        v = DeviceVector(self, len(host), vdt)
        check(lib().ganit_hip_upload(self.h, vdt, _np_ptr(host), len(host), v.ptr), "upload")
        return v

    def empty(self, n: int, vdt: int = F64) -> "DeviceVector":
        return DeviceVector(self, n, vdt)


class DeviceVector:
    def __init__(self, ctx: Context, n: int, vdt: int):
        self.ctx, self.n, self.vdt = ctx, int(n), vdt
        self.ptr = ctx.alloc(self.n * _SIZE[vdt])

    def numpy(self) -> np.ndarray:
        out = np.zeros(self.n)
        check(lib().ganit_hip_download(self.ctx.h, self.vdt, self.ptr, self.n, _np_ptr(out)), "download")
        return out

    def fill(self, a: float):
        check(lib().ganit_hip_vec_fill(self.ctx.h, self.vdt, self.ptr, self.n, float(a)), "fill")


class DeviceCsr:
    def __init__(self, ctx: Context, rows: int, cols: int, ptr, idx, val, mask: int = 7):
        self.ctx, self.rows, self.cols = ctx, rows, cols
        self._ptr = np.ascontiguousarray(ptr, dtype=np.int32)
        self._idx = np.ascontiguousarray(idx, dtype=np.int32)
        self._val = np.ascontiguousarray(val, dtype=np.float64)
        h = _vp()
        check(lib().ganit_hip_csr_upload(ctx.h, rows, cols, _np_ptr(self._ptr), _np_ptr(self._idx),
                                         _np_ptr(self._val), mask, ctypes.byref(h)), "csr_upload")
        self.h = h
        ctx._adopt(self)

    def release(self):
        if getattr(self, "h", None):
            lib().ganit_hip_csr_free(self.h)
            self.h = None

    def __del__(self):
        self.release()

    def plan(self):
        lanes = _i()
        hist = np.zeros(9, dtype=np.int64)
        check(lib().ganit_hip_csr_plan(self.h, ctypes.byref(lanes), _np_ptr(hist)), "csr_plan")
        return lanes.value, hist

    def set_lanes(self, lanes: int):
        check(lib().ganit_hip_csr_set_lanes(self.h, lanes), "csr_set_lanes")

    def autotune(self, mdt: int = F64, vdt: int = F64, reps: int = 5) -> int:
        check(lib().ganit_hip_csr_autotune(self.ctx.h, self.h, mdt, vdt, reps), "csr_autotune")
        return self.plan()[0]

    def spmv(self, x: DeviceVector, y: DeviceVector, mdt: int = F64):
        check(lib().ganit_hip_spmv(self.ctx.h, self.h, mdt, x.vdt, x.ptr, y.ptr), "spmv")


class LpDesc(ctypes.Structure):
    _fields_ = [("m", _i), ("n", _i)] + [(k, _vp) for k in (
        "a_ptr", "a_idx", "a_val", "at_ptr", "at_idx", "at_val", "q_ptr", "q_idx", "q_val",
        "c", "l", "u", "lo", "hi", "rres", "cres")]


class DevicePdhg:
    """A scaled problem resident on the device (ganit_hip_pdhg_*)."""

    def __init__(self, ctx: Context, scaled: dict, mdt_mask: int = 1, vdt_mask: int = 1):
        self.ctx = ctx
        s = scaled
        self.m, self.n = s["m"], s["n"]
        keep = {}

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            keep[id(a)] = a
            return _np_ptr(a)

        d = LpDesc()
        d.m, d.n = self.m, self.n
        d.a_ptr, d.a_idx, d.a_val = arr(s["A"].ptr, np.int32), arr(s["A"].idx, np.int32), arr(s["A"].val, np.float64)
        d.at_ptr, d.at_idx, d.at_val = (arr(s["AT"].ptr, np.int32), arr(s["AT"].idx, np.int32),
                                        arr(s["AT"].val, np.float64))
        if s["Q"].nnz > 0:
            d.q_ptr, d.q_idx, d.q_val = arr(s["Q"].ptr, np.int32), arr(s["Q"].idx, np.int32), arr(s["Q"].val, np.float64)
        for k, key in (("c", "c"), ("l", "l"), ("u", "u"), ("lo", "lo"), ("hi", "hi"), ("rres", "Rres"),
                       ("cres", "Cres")):
            setattr(d, k, arr(s[key], np.float64))
        h = _vp()
        check(lib().ganit_hip_pdhg_create(ctx.h, ctypes.byref(d), mdt_mask, vdt_mask, ctypes.byref(h)), "pdhg_create")
        self.h = h
        ctx._adopt(self)

    def release(self):
        if getattr(self, "h", None):
            lib().ganit_hip_pdhg_destroy(self.h)
            self.h = None

    def __del__(self):
        self.release()

    def plan(self, which: int):
        lanes = _i()
        hist = np.zeros(9, dtype=np.int64)
        check(lib().ganit_hip_pdhg_plan(self.h, which, ctypes.byref(lanes), _np_ptr(hist)), "pdhg_plan")
        return lanes.value, hist

    def set_lanes(self, which: int, lanes: int):
        check(lib().ganit_hip_pdhg_set_lanes(self.h, which, lanes), "pdhg_set_lanes")

    def window_layout(self, which: int) -> int:
        """1 when matrix `which` (0 A, 1 A^T, 2 Q) stages its LDS windows swizzled, else 0."""
        sw = _i()
        check(lib().ganit_hip_pdhg_window_layout(self.h, which, ctypes.byref(sw)), "pdhg_window_layout")
        return sw.value

    def load(self, vdt, x, y, xr=None, yr=None):
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        xr = None if xr is None else np.ascontiguousarray(xr, np.float64)
        yr = None if yr is None else np.ascontiguousarray(yr, np.float64)
        check(lib().ganit_hip_pdhg_load(self.h, vdt, _np_ptr(x), _np_ptr(y), None if xr is None else _np_ptr(xr),
                                        None if yr is None else _np_ptr(yr)), "pdhg_load")

    def handover(self, src, dst):
        check(lib().ganit_hip_pdhg_handover(self.h, src, dst), "pdhg_handover")

    def iterate(self, mdt, vdt, iters, tau, sigma):
        check(lib().ganit_hip_pdhg_iterate(self.h, mdt, vdt, int(iters), float(tau), float(sigma)), "pdhg_iterate")

    def check(self, mdt, vdt, flags, inv_k=0.0) -> np.ndarray:
        out = np.zeros(CHECK_NOUT)
        check(lib().ganit_hip_pdhg_check(self.h, mdt, vdt, flags, float(inv_k), _np_ptr(out)), "pdhg_check")
        return out

    def restart(self, vdt, adopt):
        check(lib().ganit_hip_pdhg_restart(self.h, vdt, int(bool(adopt))), "pdhg_restart")

    def adopt_average(self, vdt):
        check(lib().ganit_hip_pdhg_adopt_average(self.h, vdt), "pdhg_adopt_average")

    def store(self, vdt):
        x, y, xr, yr = np.zeros(self.n), np.zeros(self.m), np.zeros(self.n), np.zeros(self.m)
        check(lib().ganit_hip_pdhg_store(self.h, vdt, _np_ptr(x), _np_ptr(y), _np_ptr(xr), _np_ptr(yr)), "pdhg_store")
        return x, y, xr, yr
