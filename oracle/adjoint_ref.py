"""Active-set KKT reference gradient of a convex QP's solution map (test infrastructure: the checker of the
adjoint kernel; nothing in the product imports it).

The reference has no differentiation code, so this is pinned by mathematics alone: the closed-form
derivatives of an equality-constrained QP (tests/test_oracle_adjoint.py) and FP64 central differences of the
oracle's own PDHG solves.

At a solution (x*, y*) of  min 1/2 x'Qx + c'x  s.t.  lo <= Ax <= hi,  l <= x <= u,  with the variables at a
bound (B) and the active rows (R: equality rows and rows with y*_i != 0) held fixed, the free variables F
satisfy the reduced KKT system
    Q_FF x_F + Q_FB x_B + c_F - A_RF' nu = 0,      A_RF x_F + A_RB x_B = bnd_R.
Differentiating, with K = [[Q_FF, -A_RF'], [A_RF, 0]]:  K [dx_F; dnu] = [-dc_F - dq_F o x_F; dbnd_R].
For L = w'x*, solve K' [xi; eta] = [w_F; 0]; then dL/dc_F = -xi, dL/dq_F = -xi o x_F, dL/dbnd_R = eta, and
every other component is zero.
"""
from __future__ import annotations

import numpy as np


def _dense(M, shape):
    if M is None:
        return np.zeros(shape)
    if hasattr(M, "to_dense"):
        return M.to_dense()
    return np.asarray(M, dtype=np.float64)


def active_set(x, y, l, u, lo, hi, bound_tol=1e-9):
    x = np.asarray(x, np.float64)
    at_l = np.isfinite(l) & (np.abs(x - l) <= bound_tol * (1.0 + np.abs(np.where(np.isfinite(l), l, 0.0))))
    at_u = np.isfinite(u) & (np.abs(x - u) <= bound_tol * (1.0 + np.abs(np.where(np.isfinite(u), u, 0.0))))
    free = ~(at_l | at_u)
    rows = (np.asarray(lo) == np.asarray(hi)) | (np.asarray(y) != 0.0)
    return free, rows


def active_set_gradient(A, Q, x, y, l, u, lo, hi, w, bound_tol=1e-9):
    """dL/dc (n), dL/d(active bound of each row) (m), dL/dq_jj (n) for L = w'x* at the solution (x, y).
    A, Q: anything with .to_dense() or a dense array; Q may be None."""
    x = np.asarray(x, np.float64)
    n = x.size
    Ad = _dense(A, (len(lo), n))
    Qd = _dense(Q, (n, n))
    free, rows = active_set(x, y, l, u, lo, hi, bound_tol)
    Fi = np.nonzero(free)[0]
    Ri = np.nonzero(rows)[0]
    nf, nr = Fi.size, Ri.size
    K = np.zeros((nf + nr, nf + nr))
    K[:nf, :nf] = Qd[np.ix_(Fi, Fi)]
    K[:nf, nf:] = -Ad[np.ix_(Ri, Fi)].T
    K[nf:, :nf] = Ad[np.ix_(Ri, Fi)]
    rhs = np.concatenate([np.asarray(w, np.float64)[Fi], np.zeros(nr)])
    try:
        sol = np.linalg.solve(K.T, rhs)
    except np.linalg.LinAlgError:
        sol = np.linalg.lstsq(K.T, rhs, rcond=None)[0]
    xi, eta = sol[:nf], sol[nf:]
    dc = np.zeros(n)
    dq = np.zeros(n)
    db = np.zeros(len(lo))
    dc[Fi] = -xi
    dq[Fi] = -xi * x[Fi]
    db[Ri] = eta
    return dict(dc=dc, db=db, dq=dq, free=int(nf), active_rows=int(nr))


def separable_qp_closed_form(q, c, b, w):
    """min 1/2 sum q_j x_j^2 + c'x  s.t.  sum x = b  (bounds inactive): x*, and the exact dL/dc, dL/db, dL/dq
    for L = w'x*.  x_j = (nu - c_j) / q_j with nu = (b + sum c/q) / sum 1/q."""
    q, c, w = (np.asarray(v, np.float64) for v in (q, c, w))
    S = np.sum(1.0 / q)
    nu = (b + np.sum(c / q)) / S
    x = (nu - c) / q
    Wq = np.sum(w / q)
    dc = Wq / (q * S) - w / q
    db = Wq / S
    dq = x / (q * S) * Wq - w * x / q
    return x, dict(dc=dc, db=np.array([db]), dq=dq)
