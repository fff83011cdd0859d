"""ctypes wrapper of oracle/liboracle.so -- the CPU restatement of the reference's PDHG path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg, and
only as the checker / CPU baseline.  Nothing in ganit_v2_amd/ imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "ganit")

_lib = None
_vp = ctypes.c_void_p
STATUS = ["OPTIMAL", "PRIMAL_INFEASIBLE", "ITERATION_LIMIT", "TIME_LIMIT", "NUMERICAL_ERROR"]
PRECISION = {"double": 0, "mixed": 1, "dynamic": 2}


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "port"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_lp_read_mps.restype = _vp
        L.oracle_lp_read_mps.argtypes = [ctypes.c_char_p]
        L.oracle_lp_from_triplets.restype = _vp
        L.oracle_lp_from_triplets.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_longlong] + [_vp] * 8 + [
            ctypes.c_double, ctypes.c_int, ctypes.c_longlong, _vp, _vp, _vp]
        L.oracle_lp_free.argtypes = [_vp]
        L.oracle_lp_dims.argtypes = [_vp, _vp]
        L.oracle_lp_export.argtypes = [_vp] + [_vp] * 12
        L.oracle_scaled_scalars.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp]
        L.oracle_scaled_export.argtypes = [_vp] * 19
        L.oracle_solve.argtypes = [_vp, _vp, _vp, _vp, _vp]
        L.oracle_solve_traced.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]
        L.oracle_spmv.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp]
        L.oracle_round_bf16.argtypes = [ctypes.c_longlong, _vp, _vp]
        L.oracle_kahan_sum.argtypes = [ctypes.c_longlong, _vp]
        L.oracle_kahan_sum.restype = ctypes.c_double
        L.oracle_primal_step.argtypes = [ctypes.c_int, ctypes.c_int] + [_vp] * 6 + [ctypes.c_double] + [_vp] * 3
        L.oracle_dual_step.argtypes = [ctypes.c_int, ctypes.c_int] + [_vp] * 4 + [ctypes.c_double] + [_vp] * 2
        L.oracle_kkt.argtypes = [_vp, _vp, _vp, _vp]
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


def set_threads(k: int) -> None:
    lib().oracle_set_threads(int(k))


def max_threads() -> int:
    return lib().oracle_max_threads()


class OracleLP:
    def __init__(self, h):
        if not h:
            raise RuntimeError(lib().oracle_last_error().decode())
        self.h = ctypes.c_void_p(h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.oracle_lp_free(self.h)
            self.h = None

    @classmethod
    def read_mps(cls, path: str) -> "OracleLP":
        return cls(lib().oracle_lp_read_mps(os.fsencode(path)))

    @classmethod
    def from_export(cls, d: dict) -> "OracleLP":
        """From the dict of ganit_v2_amd.Problem.export() (original-space data)."""
        A = d["A"]
        rows = np.repeat(np.arange(A.rows, dtype=np.int32), np.diff(A.ptr)).astype(np.int32)
        idx = np.ascontiguousarray(A.idx, np.int32)
        val = np.ascontiguousarray(A.val, np.float64)
        Q = d["Q"]
        if Q.nnz:
            qr = np.repeat(np.arange(Q.rows, dtype=np.int32), np.diff(Q.ptr)).astype(np.int32)
            qi = np.ascontiguousarray(Q.idx, np.int32)
            qv = np.ascontiguousarray(Q.val, np.float64)
        else:
            qr = qi = qv = None
        vecs = [np.ascontiguousarray(d[k], np.float64) for k in ("c", "l", "u", "lo", "hi")]
        h = lib().oracle_lp_from_triplets(A.rows, A.cols, len(val), _p(rows), _p(idx), _p(val), *[_p(v) for v in vecs],
                                          float(d["obj_const"]), int(bool(d["maximize"])), 0 if qv is None else len(qv),
                                          _p(qr), _p(qi), _p(qv))
        return cls(h)

    def dims(self):
        d = np.zeros(5, dtype=np.int64)
        lib().oracle_lp_dims(self.h, _p(d))
        return dict(m=int(d[0]), n=int(d[1]), nnz=int(d[2]), qnnz=int(d[3]), maximize=bool(d[4]))

    def export(self) -> dict:
        d = self.dims()
        m, n = d["m"], d["n"]
        out = dict(aptr=np.zeros(m + 1, np.int32), aidx=np.zeros(d["nnz"], np.int32), aval=np.zeros(d["nnz"]),
                   qptr=np.zeros(n + 1, np.int32), qidx=np.zeros(d["qnnz"], np.int32), qval=np.zeros(d["qnnz"]),
                   c=np.zeros(n), l=np.zeros(n), u=np.zeros(n), lo=np.zeros(m), hi=np.zeros(m))
        oc = np.zeros(1)
        lib().oracle_lp_export(self.h, *[_p(out[k]) for k in ("aptr", "aidx", "aval", "qptr", "qidx", "qval", "c", "l",
                                                               "u", "lo", "hi")], _p(oc))
        out["obj_const"] = float(oc[0])
        return out

    def scaled(self, ruiz=10, pc=1) -> dict:
        d = self.dims()
        m, n, nz, qz = d["m"], d["n"], d["nnz"], d["qnnz"]
        scal = np.zeros(9)
        lib().oracle_scaled_scalars(self.h, ruiz, pc, _p(scal))
        o = dict(aptr=np.zeros(m + 1, np.int32), aidx=np.zeros(nz, np.int32), aval=np.zeros(nz),
                 tptr=np.zeros(n + 1, np.int32), tidx=np.zeros(nz, np.int32), tval=np.zeros(nz),
                 qptr=np.zeros(n + 1, np.int32), qidx=np.zeros(qz, np.int32), qval=np.zeros(qz),
                 c=np.zeros(n), l=np.zeros(n), u=np.zeros(n), lo=np.zeros(m), hi=np.zeros(m), R=np.zeros(m),
                 C=np.zeros(n), Rres=np.zeros(m), Cres=np.zeros(n))
        keys = ["aptr", "aidx", "aval", "tptr", "tidx", "tval", "qptr", "qidx", "qval", "c", "l", "u", "lo", "hi", "R",
                "C", "Rres", "Cres"]
        lib().oracle_scaled_export(self.h, *[_p(o[k]) for k in keys])
        o.update(dict(zip(["bscale", "cscale", "normQ", "obj_const", "bnorm", "cnorm", "normA", "nnz_at", "maximize"],
                          scal.tolist())))
        return o

    def solve(self, tol=1e-4, max_iter=10_000_000, time_limit=3600.0, check_every=64, ruiz_iters=10,
              pock_chambolle=True, verbose=0, infeas_tol=1e-8, precision="double", stall_checks=12,
              stall_checks_l2=24, l1_matrix=0, kahan=0, eval_fp64=False, trace=False) -> dict:
        d = self.dims()
        opts = np.array([tol, max_iter, time_limit, check_every, ruiz_iters, int(bool(pock_chambolle)), verbose,
                         infeas_tol, PRECISION[precision], stall_checks, stall_checks_l2, l1_matrix, kahan,
                         int(bool(eval_fp64))], np.float64)
        out = np.zeros(13)
        x, y = np.zeros(d["n"]), np.zeros(d["m"])
        if trace:
            tr = np.zeros((200000, 5))
            rows = ctypes.c_int()
            rc = lib().oracle_solve_traced(self.h, _p(opts), _p(out), _p(x), _p(y), _p(tr), tr.shape[0],
                                           ctypes.byref(rows))
        else:
            rc = lib().oracle_solve(self.h, _p(opts), _p(out), _p(x), _p(y))
        if rc:
            raise RuntimeError(lib().oracle_last_error().decode())
        keys = ["status", "pobj", "dobj", "rel_p", "rel_d", "rel_gap", "iters", "fp32_iters", "l2_iters", "restarts",
                "setup_s", "solve_s", "norm_a"]
        r = dict(zip(keys, out.tolist()))
        r["status"] = STATUS[int(out[0])]
        for k in ("iters", "fp32_iters", "l2_iters", "restarts"):
            r[k] = int(r[k])
        r["x"], r["y"] = x, y
        if trace:
            r["trace"] = tr[: min(rows.value, tr.shape[0])].copy()
        return r

    def kkt(self, x_scaled, y_scaled) -> dict:
        out = np.zeros(12)
        rc = lib().oracle_kkt(self.h, _p(np.ascontiguousarray(x_scaled, np.float64)),
                              _p(np.ascontiguousarray(y_scaled, np.float64)), _p(out))
        if rc:
            raise RuntimeError(lib().oracle_last_error().decode())
        return dict(zip(["pobj", "dobj", "rel_p", "rel_d", "rel_gap", "err", "rp2", "rd2", "vobj", "robj", "cx", "xqx"],
                        out.tolist()))


def spmv(mdt, vdt, rows, cols, ptr, idx, val, x):
    y = np.zeros(rows)
    rc = lib().oracle_spmv(mdt, vdt, rows, cols, _p(np.ascontiguousarray(ptr, np.int32)),
                           _p(np.ascontiguousarray(idx, np.int32)), _p(np.ascontiguousarray(val, np.float64)),
                           _p(np.ascontiguousarray(x, np.float64)), _p(y))
    if rc:
        raise ValueError("unsupported dtype pair")
    return y


def round_bf16(a):
    a = np.ascontiguousarray(a, np.float64)
    out = np.zeros_like(a)
    lib().oracle_round_bf16(len(a), _p(a), _p(out))
    return out


def kahan_sum(a) -> float:
    """The oracle's Kahan-compensated FP32 sum (the reductions of its kahan option) of an array."""
    a = np.ascontiguousarray(a, np.float64)
    return float(lib().oracle_kahan_sum(len(a), _p(a)))


def primal_step(vdt, x, aty, qx, c, l, u, tau, sumx):
    n = len(x)
    xn, xbar, s = np.zeros(n), np.zeros(n), np.ascontiguousarray(sumx, np.float64).copy()
    f = lambda a: None if a is None else _p(np.ascontiguousarray(a, np.float64))
    lib().oracle_primal_step(vdt, n, f(x), f(aty), f(qx), f(c), f(l), f(u), float(tau), _p(xn), _p(xbar), _p(s))
    return xn, xbar, s


def dual_step(vdt, y, axbar, lo, hi, sigma, sumy):
    m = len(y)
    yn, s = np.zeros(m), np.ascontiguousarray(sumy, np.float64).copy()
    f = lambda a: _p(np.ascontiguousarray(a, np.float64))
    lib().oracle_dual_step(vdt, m, f(y), f(axbar), f(lo), f(hi), float(sigma), _p(yn), _p(s))
    return yn, s


def run_reference(path: str, *args: str, threads: int = 1, timeout: float = 600.0) -> dict:
    """Run the reference CLI built from its own sources (oracle/_ref/ganit) and parse its --json line."""
    import json
    if not os.path.exists(REF_BIN):
        raise FileNotFoundError(REF_BIN)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    p = subprocess.run([REF_BIN, path, "--json", "--quiet", *args], env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=timeout)
    line = [l for l in p.stdout.splitlines() if l.startswith("{")]
    if not line:
        raise RuntimeError(f"reference produced no JSON (rc {p.returncode}): {p.stderr[-500:]}")
    return json.loads(line[-1])
