"""Differentiable QP layer over the GPU PDHG engine.

    x*(c, b, q) = argmin_x  1/2 x'(Q_off + diag q) x + c'x   s.t.  A_i x = b_i (equality rows),
                                                                 A_i x <= b_i (<= rows), A_i x >= b_i (>= rows),
                                                                 l <= x <= u

The forward pass is a restarted-PDHG solve (libganit.so -> libganit_hip.so); the backward pass is the adjoint
PDHG iteration on the GPU (ganit_solver_adjoint): a transposed PDHG fixed-point iteration at the solution,
with the bound projections frozen, giving dL/dc, dL/db and dL/dq for a loss L(x*).  Everything crosses the
C ABI through ctypes; tensors are converted to float64 host arrays at the boundary.

    layer = QPLayer.from_problem(Problem.generate("diff_qp", seed=0))
    x = layer(c, b, q)              # torch tensors with requires_grad
    (w * x).sum().backward()        # c.grad, b.grad, q.grad

A batch (c, b, q of shape (B, n), (B, m), (B, n)) runs its B forward solves as one batched solve (Batch: the
instances packed block-diagonally, each following its single solve's trajectory) and its B adjoints as one batched
GMRES on that batch (Batch.adjoint).  With batch_forward=False the instances are solved concurrently from host
threads instead, each on its own HIP context and stream, and their adjoints run per instance.
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import Batch, Problem

ROW_EQ, ROW_LE, ROW_GE = 0, 1, 2


class QPLayer:
    """The fixed structure of a family of QPs (A, row senses, variable bounds, Q's off-diagonal part) and the
    solver options; the data (c, b, diag Q) is the layer's input."""

    def __init__(self, n, A_rows, A_cols, A_vals, senses, l, u, Q_rows=(), Q_cols=(), Q_vals=(),
                 precision="mixed", tol=1e-6, adjoint_precision="double", adjoint_tol=1e-8,
                 adjoint_max_iter=200_000, batch_forward=True, batch_adjoint=True, scale_q=True, **solve_opts):
        self.n = int(n)
        self.senses = np.asarray(senses, dtype=np.int8)
        self.m = len(self.senses)
        self.A = (np.asarray(A_rows, np.int32), np.asarray(A_cols, np.int32), np.asarray(A_vals, np.float64))
        qr, qc, qv = (np.asarray(Q_rows, np.int32), np.asarray(Q_cols, np.int32), np.asarray(Q_vals, np.float64))
        off = qr != qc
        self.Q_off = (qr[off], qc[off], qv[off])
        self.l = np.asarray(l, np.float64)
        self.u = np.asarray(u, np.float64)
        # scale_q: the column scaling also equilibrates Q ([Q A'; A 0]); on these QPs it takes the forward solve from
        # 1e5-1e7 PDHG iterations to a few hundred (the same optimum; the gradient is the solution map's either way)
        self.opts = dict(precision=precision, tol=tol, scale_q=scale_q, **solve_opts)
        self.adjoint_opts = dict(precision=adjoint_precision, tol=adjoint_tol, max_iter=adjoint_max_iter)
        # a (B, n) input: its B forward solves as one batched solve (else as concurrent single solves) ...
        self.batch_forward = bool(batch_forward)
        # ... and, after a batched forward, its B adjoints as one batched GMRES (else each instance's solver takes
        # its solution over and runs its own adjoint, concurrently)
        self.batch_adjoint = bool(batch_adjoint)
        self.last = None  # (result dict, adjoint dict) of the latest forward / backward, for inspection

    @classmethod
    def from_problem(cls, p: Problem, **opts) -> "QPLayer":
        """Structure of an existing problem; its c, b and diag Q are the natural first input."""
        d = p.export()
        A, Q = d["A"], d["Q"]
        rows = np.repeat(np.arange(A.rows, dtype=np.int32), np.diff(A.ptr))
        qrows = np.repeat(np.arange(Q.rows, dtype=np.int32), np.diff(Q.ptr))
        lo, hi = d["lo"], d["hi"]
        senses = np.empty(len(lo), np.int8)
        for i, (a, b) in enumerate(zip(lo, hi)):
            if a == b:
                senses[i] = ROW_EQ
            elif np.isneginf(a) and np.isfinite(b):
                senses[i] = ROW_LE
            elif np.isposinf(b) and np.isfinite(a):
                senses[i] = ROW_GE
            else:
                raise ValueError(f"row {i}: only equality and one-sided rows are differentiable here")
        return cls(A.cols, rows, A.idx, A.val, senses, d["l"], d["u"], qrows, Q.idx, Q.val, **opts)

    def data_of(self, p: Problem):
        """(c, b, diag Q) of a problem with this structure, as float64 tensors."""
        d = p.export()
        b = np.where(self.senses == ROW_GE, d["lo"], d["hi"])
        Q = d["Q"]
        q = np.zeros(self.n)
        for j in range(self.n):
            s, e = Q.ptr[j], Q.ptr[j + 1]
            hit = np.nonzero(Q.idx[s:e] == j)[0]
            if hit.size:
                q[j] = Q.val[s + hit[0]]
        return torch.from_numpy(d["c"].copy()), torch.from_numpy(b.copy()), torch.from_numpy(q)

    def problem(self, c: np.ndarray, b: np.ndarray, q: np.ndarray) -> Problem:
        lo = np.where(self.senses == ROW_LE, -np.inf, b)
        hi = np.where(self.senses == ROW_GE, np.inf, b)
        diag = np.arange(self.n, dtype=np.int32)
        qr = np.concatenate([self.Q_off[0], diag])
        qc = np.concatenate([self.Q_off[1], diag])
        qv = np.concatenate([self.Q_off[2], q])
        return Problem.from_triplets(self.m, self.n, *self.A, c, self.l, self.u, lo, hi, qrows=qr, qcols=qc,
                                     qvals=qv)

    def __call__(self, c: torch.Tensor, b: torch.Tensor, q: torch.Tensor) -> torch.Tensor:
        if c.dim() == 2:
            if not (b.dim() == 2 and q.dim() == 2 and len(b) == len(c) == len(q)):
                raise ValueError("a batch needs c, b, q of shapes (B, n), (B, m), (B, n)")
            return QPBatchFunction.apply(c, b, q, self)
        return QPFunction.apply(c, b, q, self)

    def threads(self, count: int) -> int:
        """Host threads for a batch of `count` instances (GANIT_QP_THREADS overrides; default up to 8)."""
        env = os.environ.get("GANIT_QP_THREADS")
        return max(1, min(count, int(env) if env else 8))

    def _forward_one(self, c: np.ndarray, b: np.ndarray, q: np.ndarray):
        p = self.problem(c, b, q)
        solver = p.solver(**self.opts)
        solver.upload()
        r = solver.run(return_solution=True)
        if r["status"] != "OPTIMAL":
            raise RuntimeError(f"QP layer forward solve ended {r['status']}")
        return solver, p, r


    def _forward_batch(self, C: np.ndarray, B: np.ndarray, Qd: np.ndarray):
        """The instances' forward solves as one batched solve (each follows its single solve's trajectory).
        Returns the batch, the problems and the results."""
        problems = [self.problem(c, b, q) for c, b, q in zip(C, B, Qd)]
        batch = Batch(problems, **self.opts)
        res = batch.run(return_solution=True)
        for k, r in enumerate(res):
            if r["status"] != "OPTIMAL":
                raise RuntimeError(f"QP layer forward solve {k} ended {r['status']}")
        return batch, problems, res

    def _adopt(self, problems, res):
        """A solver per instance that takes over that instance's batched solution, so that its own adjoint
        differentiates there."""
        low = self.opts.get("precision", "mixed") != "double"

        def adopt(k):
            solver = problems[k].solver(**self.opts)
            solver.upload()
            # an instance certified in the FP32 level never ran an FP64 iteration
            solver.set_solution(res[k]["x"], res[k]["y"], fp32_iterate=low and res[k]["fp32_iters"] == res[k]["iters"])
            return solver

        with ThreadPoolExecutor(self.threads(len(problems))) as ex:
            return list(ex.map(adopt, range(len(problems))))


def _host(t: torch.Tensor) -> np.ndarray:
    return np.ascontiguousarray(t.detach().to("cpu", torch.float64).numpy())


class QPFunction(torch.autograd.Function):
    """x* = QP(c, b, diag Q); backward by the adjoint PDHG kernel."""

    @staticmethod
    def forward(ctx, c, b, q, layer: QPLayer):
        solver, p, r = layer._forward_one(_host(c), _host(b), _host(q))
        ctx.solver, ctx.layer, ctx.problem = solver, layer, p
        layer.last = (r, None)
        return torch.from_numpy(r["x"]).to(device=c.device, dtype=c.dtype)

    @staticmethod
    def backward(ctx, gx):
        layer: QPLayer = ctx.layer
        a = ctx.solver.adjoint(_host(gx), **layer.adjoint_opts)
        layer.last = (layer.last[0] if layer.last else None, a)
        like = gx
        conv = lambda v: torch.from_numpy(v).to(device=like.device, dtype=like.dtype)
        return conv(a["dc"]), conv(a["db"]), conv(a["dq"]), None


__all__ = ["QPLayer", "QPFunction", "ROW_EQ", "ROW_LE", "ROW_GE"]


class QPBatchFunction(torch.autograd.Function):
    """Row k of x* = QP(c[k], b[k], diag q[k]); the B forward solves run as one batched solve (or concurrently),
    the B adjoints as one batched GMRES after a batched forward (or concurrently)."""

    @staticmethod
    def forward(ctx, c, b, q, layer: QPLayer):
        C, B, Qd = _host(c), _host(b), _host(q)
        ctx.batch, ctx.solvers = None, None
        if layer.batch_forward:
            batch, problems, res = layer._forward_batch(C, B, Qd)
            if layer.batch_adjoint:
                ctx.batch = batch
            else:
                ctx.solvers = layer._adopt(problems, res)
        else:
            with ThreadPoolExecutor(layer.threads(len(C))) as ex:
                outs = list(ex.map(layer._forward_one, C, B, Qd))
            ctx.solvers = [o[0] for o in outs]
            problems, res = [o[1] for o in outs], [o[2] for o in outs]
        ctx.problems = problems
        ctx.layer = layer
        layer.last = (res, None)
        x = np.stack([r["x"] for r in res])
        return torch.from_numpy(x).to(device=c.device, dtype=c.dtype)

    @staticmethod
    def backward(ctx, gx):
        layer: QPLayer = ctx.layer
        G = _host(gx)
        if ctx.batch is not None:
            adj = ctx.batch.adjoint(list(G), **layer.adjoint_opts)
        else:
            with ThreadPoolExecutor(layer.threads(len(G))) as ex:
                adj = list(ex.map(lambda sg: sg[0].adjoint(sg[1], **layer.adjoint_opts), zip(ctx.solvers, G)))
        layer.last = (layer.last[0] if layer.last else None, adj)
        conv = lambda key: torch.from_numpy(np.stack([a[key] for a in adj])).to(device=gx.device, dtype=gx.dtype)
        return conv("dc"), conv("db"), conv("dq"), None
