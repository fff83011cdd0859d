"""ganit_v2_amd -- restarted-PDHG LP/QP engine for AMD Instinct MI355X (gfx950).

Python view of the two native libraries built in-tree by ``ganit_v2_amd/build.py``:

* ``libganit.so`` (include/ganit_host.h): MPS I/O, scaling, workload generators and the GPU solve, wrapped
  here by :class:`Problem`;
* ``libganit_hip.so`` (include/ganit_hip.h): the HIP kernels behind the C ABI, wrapped by
  :mod:`ganit_v2_amd.hip`.

There is no CPU fallback: if a library is missing, loading raises :class:`NativeLibraryMissing`.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "lib")
BIN_DIR = os.path.join(PKG_DIR, "bin")

STATUS = ["OPTIMAL", "PRIMAL_INFEASIBLE", "ITERATION_LIMIT", "TIME_LIMIT", "NUMERICAL_ERROR", "DUAL_INFEASIBLE"]
PRECISION = {"double": 0, "mixed": 1, "dynamic": 2}
VALUES = {"f32": 0, "bf16": 1, "bf16f32": 2}


class NativeLibraryMissing(RuntimeError):
    """A native library of this package has not been built (run ganit_v2_amd/build.py)."""


def _load(name: str) -> ctypes.CDLL:
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise NativeLibraryMissing(f"{path} is missing: build it with `python ganit_v2_amd/build.py`")
    return ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


_host = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p


def host_lib() -> ctypes.CDLL:
    """libganit.so with its prototypes declared."""
    global _host
    if _host is not None:
        return _host
    lib = _load("libganit.so")
    lib.ganit_last_error.restype = ctypes.c_char_p
    lib.ganit_lp_read_mps.restype = _vp
    lib.ganit_lp_read_mps.argtypes = [ctypes.c_char_p]
    lib.ganit_lp_from_triplets.restype = _vp
    lib.ganit_lp_from_triplets.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp, _vp,
                                           _vp, _vp, ctypes.c_double, ctypes.c_int, ctypes.c_int64, _vp, _vp, _vp]
    lib.ganit_lp_generate.restype = _vp
    lib.ganit_lp_generate.argtypes = [ctypes.c_int, _f64p, ctypes.c_uint64]
    lib.ganit_lp_free.argtypes = [_vp]
    lib.ganit_lp_dims.argtypes = [_vp, np.ctypeslib.ndpointer(dtype=np.int64)]
    lib.ganit_lp_export.argtypes = [_vp] + [_vp] * 12
    lib.ganit_lp_planted.argtypes = [_vp, _vp, _vp]
    lib.ganit_lp_write_mps.argtypes = [_vp, ctypes.c_char_p]
    lib.ganit_lp_scaled.argtypes = [_vp, ctypes.c_int, ctypes.c_int] + [_vp] * 19
    lib.ganit_lp_scaled_gpu.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [_vp] * 19
    lib.ganit_solve_gpu.argtypes = [_vp, _f64p, _f64p, _vp, _vp]
    lib.ganit_solver_create.restype = _vp
    lib.ganit_solver_create.argtypes = [_vp, _f64p]
    lib.ganit_solver_upload.argtypes = [_vp]
    lib.ganit_solver_run.argtypes = [_vp, _f64p, _vp, _vp]
    lib.ganit_solver_kernel_times.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _f64p]
    lib.ganit_solver_free.argtypes = [_vp]
    lib.ganit_solver_info.argtypes = [_vp, _f64p]
    lib.ganit_solver_index_bytes.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lib.ganit_solver_variant.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lib.ganit_solver_adjoint.argtypes = [_vp, _f64p, _f64p, _vp, _vp, _vp, _vp]
    lib.ganit_solver_set_solution.argtypes = [_vp, _f64p, _f64p, ctypes.c_int]
    lib.ganit_batch_create.restype = _vp
    lib.ganit_batch_create.argtypes = [_vp, ctypes.c_int, _f64p]
    lib.ganit_batch_upload.argtypes = [_vp]
    lib.ganit_batch_run.argtypes = [_vp, _f64p, _vp, _vp]
    lib.ganit_batch_stats.argtypes = [_vp, _f64p]
    lib.ganit_batch_free.argtypes = [_vp]
    lib.ganit_batch_set_bounds.argtypes = [_vp, ctypes.c_int, _f64p, _f64p]
    lib.ganit_batch_set_start.argtypes = [_vp, ctypes.c_int, _vp, _vp, ctypes.c_double]
    lib.ganit_batch_set_options.argtypes = [_vp, _f64p]
    lib.ganit_batch_weights.argtypes = [_vp, _f64p]
    lib.ganit_batch_adjoint.argtypes = [_vp, _f64p, _f64p, _vp, _vp, _vp, _vp]
    _host = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _check_host(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what}: {host_lib().ganit_last_error().decode()}")


@dataclass
class CSR:
    rows: int
    cols: int
    ptr: np.ndarray
    idx: np.ndarray
    val: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.ptr[-1]) if len(self.ptr) else 0

    def to_dense(self) -> np.ndarray:
        d = np.zeros((self.rows, self.cols))
        for i in range(self.rows):
            d[i, self.idx[self.ptr[i]:self.ptr[i + 1]]] += self.val[self.ptr[i]:self.ptr[i + 1]]
        return d

    def matvec(self, x: np.ndarray) -> np.ndarray:
        rows = np.repeat(np.arange(self.rows), np.diff(self.ptr))
        return np.bincount(rows, weights=self.val * x[self.idx], minlength=self.rows)


class Problem:
    """An LP/QP held by libganit.so (original space)."""

    def __init__(self, handle: int):
        if not handle:
            raise RuntimeError(host_lib().ganit_last_error().decode())
        self._h = ctypes.c_void_p(handle)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _host is not None:
            _host.ganit_lp_free(h)
            self._h = None

    # ---- construction
    @classmethod
    def read_mps(cls, path: str) -> "Problem":
        return cls(host_lib().ganit_lp_read_mps(os.fsencode(path)))

    @classmethod
    def generate(cls, kind: str, seed: int = 0, **p) -> "Problem":
        kinds = {
            "random_lp": (0, ["m", "n", "m_eq", "nnz_min", "nnz_max", "tighten_frac"],
                          dict(m=2000, n=5000, m_eq=1000, nnz_min=10, nnz_max=10, tighten_frac=0.0)),
            "blending": (1, ["crudes", "products", "specs", "periods"],
                         dict(crudes=60, products=16, specs=5, periods=365)),
            "portfolio": (2, ["assets", "factors", "gamma"], dict(assets=100000, factors=50, gamma=1.0)),
            "diff_qp": (3, ["n", "m_eq", "m_le", "density", "qdensity"],
                        dict(n=2000, m_eq=300, m_le=200, density=0.01, qdensity=0.001)),
            "batch_lp": (4, ["m_lo", "m_hi", "n_lo", "n_hi", "nnz_min", "nnz_max", "tighten_frac"],
                         dict(m_lo=200, m_hi=4000, n_lo=500, n_hi=10000, nnz_min=5, nnz_max=15,
                              tighten_frac=0.05)),
        }
        k, names, defaults = kinds[kind]
        vals = dict(defaults)
        for key, v in p.items():
            if key not in vals:
                raise TypeError(f"unknown parameter {key!r} for {kind}")
            vals[key] = v
        arr = np.array([float(vals[n]) for n in names])
        return cls(host_lib().ganit_lp_generate(k, arr, ctypes.c_uint64(seed)))

    @classmethod
    def from_triplets(cls, m, n, rows, cols, vals, c, l, u, lo, hi, obj_const=0.0, maximize=False,
                      qrows=None, qcols=None, qvals=None) -> "Problem":
        r = np.ascontiguousarray(rows, dtype=np.int32)
        cc = np.ascontiguousarray(cols, dtype=np.int32)
        v = np.ascontiguousarray(vals, dtype=np.float64)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (c, l, u, lo, hi)]
        if qvals is not None and len(qvals):
            qr = np.ascontiguousarray(qrows, dtype=np.int32)
            qc = np.ascontiguousarray(qcols, dtype=np.int32)
            qv = np.ascontiguousarray(qvals, dtype=np.float64)
            qn = len(qv)
        else:
            qr = qc = qv = None
            qn = 0
        h = host_lib().ganit_lp_from_triplets(int(m), int(n), len(v), _ptr(r), _ptr(cc), _ptr(v),
                                              *[_ptr(a) for a in arrs], float(obj_const), int(bool(maximize)), qn,
                                              _ptr(qr), _ptr(qc), _ptr(qv))
        return cls(h)

    # ---- inspection
    def dims(self) -> dict:
        d = np.zeros(5, dtype=np.int64)
        host_lib().ganit_lp_dims(self._h, d)
        return dict(m=int(d[0]), n=int(d[1]), nnz=int(d[2]), qnnz=int(d[3]), maximize=bool(d[4]))

    def export(self) -> dict:
        d = self.dims()
        m, n = d["m"], d["n"]
        A = CSR(m, n, np.zeros(m + 1, np.int32), np.zeros(d["nnz"], np.int32), np.zeros(d["nnz"]))
        Q = CSR(n, n, np.zeros(n + 1, np.int32), np.zeros(d["qnnz"], np.int32), np.zeros(d["qnnz"]))
        vec = {k: np.zeros(n) for k in ("c", "l", "u")}
        vec.update({k: np.zeros(m) for k in ("lo", "hi")})
        oc = np.zeros(1)
        host_lib().ganit_lp_export(self._h, _ptr(A.ptr), _ptr(A.idx), _ptr(A.val), _ptr(Q.ptr), _ptr(Q.idx),
                                   _ptr(Q.val), _ptr(vec["c"]), _ptr(vec["l"]), _ptr(vec["u"]), _ptr(vec["lo"]),
                                   _ptr(vec["hi"]), _ptr(oc))
        return dict(A=A, Q=Q, obj_const=float(oc[0]), maximize=d["maximize"], **vec)

    def planted(self):
        d = self.dims()
        x0, y0 = np.zeros(d["n"]), np.zeros(d["m"])
        _check_host(host_lib().ganit_lp_planted(self._h, _ptr(x0), _ptr(y0)), "planted")
        return x0, y0

    def write_mps(self, path: str) -> None:
        _check_host(host_lib().ganit_lp_write_mps(self._h, os.fsencode(path)), "write_mps")

    def scaled(self, ruiz_iters: int = 10, pock_chambolle: bool = True, on_gpu: bool = False, device: int = 0) -> dict:
        """The scaled problem exactly as the solver uploads it (scale_lp + estimate_norm); on_gpu: with the matrix
        work of both on GPU `device` (the same bits)."""
        d = self.dims()
        m, n = d["m"], d["n"]
        scal = np.zeros(9)
        # first call for the nnz of A^T and Q (same as A / Q)
        A = CSR(m, n, np.zeros(m + 1, np.int32), np.zeros(d["nnz"], np.int32), np.zeros(d["nnz"]))
        AT = CSR(n, m, np.zeros(n + 1, np.int32), np.zeros(d["nnz"], np.int32), np.zeros(d["nnz"]))
        Q = CSR(n, n, np.zeros(n + 1, np.int32), np.zeros(d["qnnz"], np.int32), np.zeros(d["qnnz"]))
        v = {k: np.zeros(n) for k in ("c", "l", "u", "C", "Cres")}
        v.update({k: np.zeros(m) for k in ("lo", "hi", "R", "Rres")})
        head = ((host_lib().ganit_lp_scaled_gpu, int(device)) if on_gpu else (host_lib().ganit_lp_scaled,))
        _check_host(head[0](
            self._h, int(ruiz_iters), int(bool(pock_chambolle)), *head[1:], _ptr(scal), _ptr(A.ptr), _ptr(A.idx),
            _ptr(A.val),
            _ptr(AT.ptr), _ptr(AT.idx), _ptr(AT.val), _ptr(Q.ptr), _ptr(Q.idx), _ptr(Q.val), _ptr(v["c"]),
            _ptr(v["l"]), _ptr(v["u"]), _ptr(v["lo"]), _ptr(v["hi"]), _ptr(v["R"]), _ptr(v["C"]), _ptr(v["Rres"]),
            _ptr(v["Cres"])), "scaled")
        keys = ["bscale", "cscale", "normQ", "obj_const", "bnorm", "cnorm", "normA", "nnz_at", "maximize"]
        out = dict(zip(keys, scal.tolist()))
        out.update(A=A, AT=AT, Q=Q, m=m, n=n, **v)
        return out

    # ---- solve
    def solve_gpu(self, return_solution=True, **opts) -> dict:
        """One solve_pdhg_gpu call (scaling, upload and solve; timing as the reference defines it)."""
        d = self.dims()
        out = np.zeros(18)
        x = np.zeros(d["n"]) if return_solution else None
        y = np.zeros(d["m"]) if return_solution else None
        _check_host(host_lib().ganit_solve_gpu(self._h, _options(**opts), out, _ptr(x), _ptr(y)), "solve_gpu")
        return _result(out, x, y)

    def solver(self, **opts) -> "Solver":
        """A staged solver: set-up now, upload() and run() repeatedly with the problem resident in HBM."""
        return Solver(self, **opts)


_RESULT_KEYS = ["status", "pobj", "dobj", "rel_p", "rel_d", "rel_gap", "iters", "fp32_iters", "l2_iters",
                "restarts", "setup_s", "solve_s", "norm_a", "upload_s", "l1_s", "l2_s", "l3_s", "checks"]


SETUP = {"auto": 0, "host": 1, "gpu": 2}


def _options(tol=1e-4, max_iter=10_000_000, time_limit=3600.0, check_every=64, ruiz_iters=10, pock_chambolle=True,
             verbose=0, infeas_tol=1e-8, precision="double", stall_checks=12, stall_checks_l2=24, values="f32",
             device=0, eval_fp64=False, kahan=False, detect_unbounded=False, setup="auto",
             polish=False, scale_q=False, tail_lanes=True, tail_fp64=False) -> np.ndarray:
    flags = (int(bool(kahan)) | (2 if detect_unbounded else 0) | (SETUP[setup] << 2) | (16 if polish else 0)
             | (32 if scale_q else 0) | (0 if tail_lanes else 64) | (128 if tail_fp64 else 0))
    return np.array([tol, max_iter, time_limit, check_every, ruiz_iters, int(bool(pock_chambolle)), verbose,
                     infeas_tol, PRECISION[precision], stall_checks, stall_checks_l2, VALUES[values], device,
                     int(bool(eval_fp64)), flags], dtype=np.float64)


def _result(out, x, y) -> dict:
    r = dict(zip(_RESULT_KEYS, out.tolist()))
    r["status"] = STATUS[int(out[0])]
    for k in ("iters", "fp32_iters", "l2_iters", "restarts", "checks"):
        r[k] = int(r[k])
    if x is not None:
        r["x"], r["y"] = x, y
    return r


class Solver:
    def __init__(self, problem: Problem, **opts):
        self._problem = problem
        self._dims = problem.dims()
        h = host_lib().ganit_solver_create(problem._h, _options(**opts))
        if not h:
            raise RuntimeError(host_lib().ganit_last_error().decode())
        self._h = ctypes.c_void_p(h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _host is not None:
            _host.ganit_solver_free(h)
            self._h = None

    def upload(self) -> None:
        _check_host(host_lib().ganit_solver_upload(self._h), "solver_upload")

    def run(self, return_solution=False) -> dict:
        out = np.zeros(18)
        x = np.zeros(self._dims["n"]) if return_solution else None
        y = np.zeros(self._dims["m"]) if return_solution else None
        _check_host(host_lib().ganit_solver_run(self._h, out, _ptr(x), _ptr(y)), "solver_run")
        return _result(out, x, y)

    def adjoint(self, gx, precision: str = "double", tol: float = 1e-8, max_iter: int = 20_000,
                method: str = "gmres", every: int = 0, mask_tol: float = -1.0) -> dict:
        """Gradient of L(x*) with respect to c, the active row bounds and diag Q at the last run()'s solution,
        given gx = dL/dx*: the adjoint of the PDHG iteration on the GPU (precision "double" or "single";
        method "gmres" = Krylov-accelerated, "fixed_point" = the plain transposed iteration)."""
        d = self._dims
        gx = np.ascontiguousarray(gx, dtype=np.float64)
        if gx.shape != (d["n"],):
            raise ValueError(f"gx must have shape ({d['n']},)")
        if method not in ("gmres", "fixed_point"):
            raise ValueError("method must be 'gmres' or 'fixed_point'")
        ao = np.array([1.0 if precision == "single" else 0.0, tol, max_iter, 0.0 if method == "gmres" else 1.0,
                       every, mask_tol], dtype=np.float64)
        dc, dq, db, info = np.zeros(d["n"]), np.zeros(d["n"]), np.zeros(d["m"]), np.zeros(5)
        _check_host(host_lib().ganit_solver_adjoint(self._h, gx, ao, _ptr(dc), _ptr(db), _ptr(dq), _ptr(info)),
                    "solver_adjoint")
        return dict(dc=dc, db=db, dq=dq, iters=int(info[0]), rel_change=float(info[1]), free_vars=int(info[2]),
                    active_rows=int(info[3]), seconds=float(info[4]))

    def set_solution(self, x, y, fp32_iterate: bool = False) -> None:
        """Make an original-space solution (x, y) found elsewhere (a Batch run) this solver's solution, as if
        run() had ended there, so that adjoint() differentiates at it."""
        d = self._dims
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        if x.shape != (d["n"],) or y.shape != (d["m"],):
            raise ValueError(f"x, y must have shapes ({d['n']},), ({d['m']},)")
        _check_host(host_lib().ganit_solver_set_solution(self._h, x, y, int(bool(fp32_iterate))), "set_solution")

    def _info(self):
        info = np.zeros(4)
        _check_host(host_lib().ganit_solver_info(self._h, info), "solver_info")
        return info

    def setup_seconds(self) -> float:
        """Seconds of this solver's set-up (scaling + norm estimate), outside solve_s as in the reference."""
        return float(self._info()[0])

    def setup_on_device(self) -> bool:
        """Whether the set-up's matrix work ran on the GPU."""
        return bool(self._info()[2])

    def polished(self) -> bool:
        """Whether the last run() kept a polished point (option polish=True)."""
        return bool(self._info()[3])

    def index_bytes(self, which: int) -> int:
        """Bytes per column index the fused kernels read for A (0), A^T (1), Q (2): 2 (compact) or 4."""
        b = ctypes.c_int(4)
        _check_host(host_lib().ganit_solver_index_bytes(self._h, int(which), ctypes.byref(b)), "index_bytes")
        return int(b.value)

    def variants(self) -> dict:
        """The kernel variant kept for each matrix: 0 = LDS-staged row blocks, else lanes per row (1..64)."""
        out = {}
        for which, name in ((0, "A"), (1, "AT")):
            v = ctypes.c_int(0)
            _check_host(host_lib().ganit_solver_variant(self._h, which, ctypes.byref(v)), "variant")
            out[name] = int(v.value)
        return out

    def kernel_times(self, level: int, reps: int = 50) -> dict:
        """Device ms of the fused iteration kernels at level 0 (L1), 1 (L2) or 2 (L3): primal / dual = each kernel
        with none of its data in L2 (as in a solve, where they alternate), iteration = both in solve order,
        primal_b2b / dual_b2b = each launched back to back (part of its data stays in L2), sweep = the 64 MiB
        streaming read the first two are measured against."""
        ms = np.zeros(6)
        _check_host(host_lib().ganit_solver_kernel_times(self._h, int(level), int(reps), ms), "kernel_times")
        return dict(zip(("primal", "dual", "iteration", "primal_b2b", "dual_b2b", "sweep"), map(float, ms)))


def solve_many(problems, threads: int = 8, **opts) -> list:
    """Solve independent problems concurrently from host threads, one staged solver (HIP context and stream)
    each: for small, launch-bound LPs/QPs the GPU overlaps the solves, so the batch takes about as long as its
    slowest member.  Returns one result dict per problem (with x and y), in order; each is that problem's own
    solve (bitwise whenever the per-solver autotuner picks the same kernel variants).  (For many LPs, Batch
    runs them in one set of block-diagonal kernels instead.)"""
    from concurrent.futures import ThreadPoolExecutor

    def one(p):
        s = p.solver(**opts)
        s.upload()
        return s.run(return_solution=True)

    problems = list(problems)
    with ThreadPoolExecutor(max(1, min(threads, len(problems) or 1))) as ex:
        return list(ex.map(one, problems))


_BATCH_STATS_KEYS = ["rows", "cols", "nnz", "blocks_a", "blocks_at", "upload_s", "solve_s", "certify_s", "l1_s",
                     "l2_s", "l3_s", "l1_instances", "l2_instances", "l3_instances", "max_iters", "total_iters",
                     "autotune_pick", "autotune_ms0", "autotune_ms1", "autotune_ms2", "autotune_ms3", "setup_s",
                     "tail_iters"]


class Batch:
    """Many LPs and convex QPs solved together by the block-diagonal batched kernels (one launch sequence for
    all).

    Options as Problem.solve_gpu (check_every must be even), plus tail_lanes (default True): once the running
    instances are few and small (at most 8 with 2^18 nonzeros: launch-bound), iterate them 8-32 lanes per row
    instead of on row blocks; tail_lanes=False keeps every row sum in row order (the CPU's bits) to the end.
    tail_fp64 (default False; mixed / dynamic): the reduced-precision levels end once the running instances are
    such a tail, and the ladder carries them on in FP64 (FP32 saves nothing on launch-bound iterations; a batch
    policy, not the reference's ladder rule).  run() returns one result dict per instance (solve_s is the whole
    batch's wall time) and stats() the batch-level figures of that run (tail_iters: batched iterations run
    lanes-per-row)."""

    def __init__(self, problems, **opts):
        self._problems = list(problems)
        if not self._problems:
            raise ValueError("empty batch")
        self._dims = [p.dims() for p in self._problems]
        arr = (ctypes.c_void_p * len(self._problems))(*[p._h for p in self._problems])
        h = host_lib().ganit_batch_create(ctypes.cast(arr, ctypes.c_void_p), len(self._problems), _options(**opts))
        if not h:
            raise RuntimeError(host_lib().ganit_last_error().decode())
        self._h = ctypes.c_void_p(h)

    def __len__(self):
        return len(self._problems)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _host is not None:
            _host.ganit_batch_free(h)
            self._h = None

    def upload(self) -> None:
        _check_host(host_lib().ganit_batch_upload(self._h), "batch_upload")

    def run(self, return_solution=False) -> list:
        K = len(self._problems)
        out = np.zeros(18 * K)
        N = sum(d["n"] for d in self._dims)
        M = sum(d["m"] for d in self._dims)
        x = np.zeros(N) if return_solution else None
        y = np.zeros(M) if return_solution else None
        _check_host(host_lib().ganit_batch_run(self._h, out, _ptr(x), _ptr(y)), "batch_run")
        res, xo, yo = [], 0, 0
        for k, d in enumerate(self._dims):
            xs = x[xo:xo + d["n"]] if x is not None else None
            ys = y[yo:yo + d["m"]] if y is not None else None
            res.append(_result(out[18 * k:18 * (k + 1)], xs, ys))
            xo += d["n"]
            yo += d["m"]
        return res

    def set_bounds(self, k: int, l, u) -> None:
        """Replace instance k's variable bounds (original space) in place: a branch-and-bound child of the same
        LP without rebuilding or re-uploading the batch."""
        n = self._dims[k]["n"]
        l = np.ascontiguousarray(l, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        if l.shape != (n,) or u.shape != (n,):
            raise ValueError(f"bounds of instance {k} must have shape ({n},)")
        _check_host(host_lib().ganit_batch_set_bounds(self._h, int(k), l, u), "batch_set_bounds")

    def set_start(self, k: int, x=None, y=None, w: float = 0.0) -> None:
        """Start instance k's next runs from (x, y) in original space (e.g. its parent's solution) with primal
        weight w (<= 0: its own initial weight); x=None returns it to the cold start."""
        d = self._dims[k]
        if x is None:
            _check_host(host_lib().ganit_batch_set_start(self._h, int(k), None, None, 0.0), "batch_set_start")
            return
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        if x.shape != (d["n"],) or y.shape != (d["m"],):
            raise ValueError(f"start of instance {k} must have shapes ({d['n']},), ({d['m']},)")
        _check_host(host_lib().ganit_batch_set_start(self._h, int(k), _ptr(x), _ptr(y), float(w)),
                    "batch_set_start")

    def set_options(self, **opts) -> None:
        """New tolerances, limits, stall rules or check interval for the next runs (precision, values, scaling
        and device must stay as built)."""
        _check_host(host_lib().ganit_batch_set_options(self._h, _options(**opts)), "batch_set_options")

    def weights(self) -> np.ndarray:
        """The primal weight each instance ended the last run with."""
        w = np.zeros(len(self._problems))
        _check_host(host_lib().ganit_batch_weights(self._h, w), "batch_weights")
        return w

    def adjoint(self, gx, precision: str = "double", tol: float = 1e-8, max_iter: int = 20_000, every: int = 0,
                mask_tol: float = -1.0) -> list:
        """Solver.adjoint (GMRES form) for every instance at its solution of the last run(), all instances in
        lockstep on the block-diagonal matrices: gx = a list of the instances' dL/dx* (or one (K, n) array when
        all n agree); returns one dict per instance, as Solver.adjoint (seconds: the whole call)."""
        ns = [d["n"] for d in self._dims]
        parts = [np.ascontiguousarray(g, dtype=np.float64).reshape(-1) for g in gx]
        if len(parts) != len(ns) or any(len(g) != n for g, n in zip(parts, ns)):
            raise ValueError("gx must hold one dL/dx* of length n_k per instance")
        g = np.concatenate(parts)
        ao = np.array([1.0 if precision == "single" else 0.0, tol, max_iter, 0.0, every, mask_tol], dtype=np.float64)
        N, M = sum(ns), sum(d["m"] for d in self._dims)
        dc, dq, db, info = np.zeros(N), np.zeros(N), np.zeros(M), np.zeros(5 * len(ns))
        _check_host(host_lib().ganit_batch_adjoint(self._h, g, ao, _ptr(dc), _ptr(db), _ptr(dq), _ptr(info)),
                    "batch_adjoint")
        out, xo, yo = [], 0, 0
        for k, d in enumerate(self._dims):
            n, m = d["n"], d["m"]
            i = info[5 * k:5 * k + 5]
            out.append(dict(dc=dc[xo:xo + n].copy(), db=db[yo:yo + m].copy(), dq=dq[xo:xo + n].copy(),
                            iters=int(i[0]), rel_change=float(i[1]), free_vars=int(i[2]), active_rows=int(i[3]),
                            seconds=float(i[4])))
            xo += n
            yo += m
        return out

    def stats(self) -> dict:
        st = np.zeros(24)
        _check_host(host_lib().ganit_batch_stats(self._h, st), "batch_stats")
        r = dict(zip(_BATCH_STATS_KEYS, st.tolist()))
        for k in ("rows", "cols", "nnz", "blocks_a", "blocks_at", "l1_instances", "l2_instances", "l3_instances",
                  "max_iters", "total_iters", "autotune_pick"):
            r[k] = int(r[k])
        return r


__all__ = ["Problem", "Solver", "Batch", "CSR", "host_lib", "NativeLibraryMissing", "STATUS", "LIB_DIR", "BIN_DIR"]
