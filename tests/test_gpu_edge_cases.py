"""Edge cases of the problem data on the GPU path, each against the oracle on the same input: no constraints,
empty rows and columns, maximisation, ranged and free rows, free and one-sided variables, an objective
constant, a QP with a zero Hessian row, an infeasible system, ragged row lengths that exercise every SpMV
path (staged blocks, long rows, split rows) in one matrix, and the LP cases together as one batch."""
import numpy as np
import pytest

import ganit_v2_amd as g

pytestmark = pytest.mark.gpu

INF = np.inf


def _lp(m, n, trip, c, l, u, lo, hi, obj_const=0.0, maximize=False, q=None):
    r = np.array([t[0] for t in trip], np.int32)
    cc = np.array([t[1] for t in trip], np.int32)
    v = np.array([t[2] for t in trip], np.float64)
    kw = {}
    if q is not None:
        kw = dict(qrows=np.array([t[0] for t in q], np.int32), qcols=np.array([t[1] for t in q], np.int32),
                  qvals=np.array([t[2] for t in q], np.float64))
    return g.Problem.from_triplets(m, n, r, cc, v, np.asarray(c, float), np.asarray(l, float),
                                   np.asarray(u, float), np.asarray(lo, float), np.asarray(hi, float),
                                   obj_const=obj_const, maximize=maximize, **kw)


def _compare(oracle, p, precision="double", tol=1e-6):
    oracle.set_threads(2)
    o = oracle.OracleLP.from_export(p.export()).solve(tol=tol, time_limit=60)
    r = p.solve_gpu(tol=tol, time_limit=60, precision=precision)
    print(f"gpu {r['status']} {r['pobj']:.10g} it {r['iters']} | oracle {o['status']} {o['pobj']:.10g} "
          f"it {o['iters']}", flush=True)
    assert r["status"] == o["status"]
    if o["status"] == "OPTIMAL":
        assert abs(r["pobj"] - o["pobj"]) <= 100 * tol * (1 + abs(o["pobj"]))
        assert max(r["rel_p"], r["rel_d"], r["rel_gap"]) <= tol * (1 + 1e-7)
    return r, o


def test_no_constraints(oracle):
    """m = 0: only bounds; the solution sits at the bound the cost pushes against."""
    p = _lp(0, 4, [], [1.0, -2.0, 0.5, -1.0], [0, 0, -1, 0], [3, 5, 2, 1], [], [])
    r, _ = _compare(oracle, p)
    assert np.allclose(r["x"], [0, 5, -1, 1], atol=1e-5)


def test_empty_row_and_column(oracle):
    """A row with no entries (0 <= 0 <= 2) and a column in no row (bounded, cost pushes it to a bound)."""
    trip = [(0, 0, 1.0), (0, 1, 1.0), (2, 1, 1.0), (2, 3, 2.0)]
    p = _lp(3, 4, trip, [1, 2, -1, 1], [0, 0, 0, 0], [10, 10, 7, 10], [1, 0, 2], [INF, 2, 6])
    r, _ = _compare(oracle, p)
    assert abs(r["x"][2] - 7.0) < 1e-5


def test_maximize_ranged_and_free_rows(oracle):
    trip = [(0, 0, 1.0), (0, 1, 1.0), (1, 0, 1.0), (1, 1, -1.0), (2, 0, 3.0), (2, 2, 1.0)]
    # row 0 ranged [1, 4]; row 1 free; row 2 >= 2
    p = _lp(3, 3, trip, [1.0, 2.0, -0.5], [0, 0, 0], [3, 3, 10], [1, -INF, 2], [4, INF, INF], obj_const=5.0,
            maximize=True)
    _compare(oracle, p)


def test_free_and_one_sided_variables(oracle):
    trip = [(0, 0, 1.0), (0, 1, 1.0), (0, 2, 1.0), (1, 0, 1.0), (1, 2, -1.0)]
    # x0 free, x1 >= 0, x2 <= 4 (free below); x0 = x2, 2 x0 + x1 = 2: minimum 0 at x = (1, 0, 1)
    p = _lp(2, 3, trip, [1.0, 1.0, -1.0], [-INF, 0, -INF], [INF, INF, 4], [2, 0], [2, 0])
    r, _ = _compare(oracle, p)
    assert np.allclose(r["x"], [1, 0, 1], atol=1e-4)


def test_qp_with_zero_hessian_row(oracle):
    trip = [(0, 0, 1.0), (0, 1, 1.0), (0, 2, 1.0)]
    q = [(0, 0, 2.0), (2, 2, 1.0), (0, 2, 0.5), (2, 0, 0.5)]
    p = _lp(1, 3, trip, [-1.0, 0.3, -0.5], [0, 0, 0], [1, 1, 1], [1], [1], q=q)
    _compare(oracle, p)


def test_infeasible_equalities(oracle):
    """x0 + x1 = 1 and x0 + x1 = 3 with 0 <= x <= 10: primal infeasible."""
    trip = [(0, 0, 1.0), (0, 1, 1.0), (1, 0, 1.0), (1, 1, 1.0)]
    p = _lp(2, 2, trip, [1.0, 1.0], [0, 0], [10, 10], [1, 3], [1, 3])
    o = oracle.OracleLP.from_export(p.export()).solve(tol=1e-6, time_limit=60)
    r = p.solve_gpu(tol=1e-6, time_limit=60)
    assert r["status"] == o["status"] == "PRIMAL_INFEASIBLE"


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_every_spmv_path_in_one_matrix(oracle, precision):
    """Rows of 1, 40, 700 and 20,000 nonzeros side by side (staged blocks, a long row, a split row)."""
    rng = np.random.default_rng(3)
    n = 20000
    rows = []
    lengths = [1] * 50 + [40] * 50 + [700] * 3 + [20000]
    trip = []
    for i, k in enumerate(lengths):
        cols = rng.choice(n, size=k, replace=False)
        vals = rng.uniform(0.5, 1.5, size=k)
        trip += [(i, int(j), float(v)) for j, v in zip(cols, vals)]
        rows.append((cols, vals))
    m = len(lengths)
    x0 = rng.uniform(0, 1, n)
    b = np.array([float(np.dot(v, x0[c])) for c, v in rows])
    c = rng.uniform(0.1, 1.0, n)
    p = _lp(m, n, trip, c, np.zeros(n), np.full(n, 2.0), np.full(m, -INF), b)
    _compare(oracle, p, precision=precision, tol=1e-5)


def _unbounded(maximize=False):
    """min -x0 + x2 + x3 (or max x0 - x2 - x3) with x0 - x1 = 0, x2 + x3 <= 5, x0 + x2 >= 1, x >= 0: unbounded
    along d = (1, 1, 0, 0), which keeps every row in its cone (A d = (0, 0, 1)) and every bound (d >= 0)."""
    trip = [(0, 0, 1.0), (0, 1, -1.0), (1, 2, 1.0), (1, 3, 1.0), (2, 0, 1.0), (2, 2, 1.0)]
    sign = -1.0 if maximize else 1.0
    return _lp(3, 4, trip, [-sign, 0.0, sign, sign], [0] * 4, [INF] * 4, [0, -INF, 1], [0, 5, INF],
               maximize=maximize)


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("maximize", [False, True])
def test_unbounded_lp_certificate(oracle, precision, maximize):
    """With detect_unbounded the primal ray x - xr certifies dual infeasibility; without it the solve runs as the
    reference's would (no such test) and ends the way the oracle's does."""
    p = _unbounded(maximize)
    r = p.solve_gpu(tol=1e-6, max_iter=200_000, time_limit=60, precision=precision, detect_unbounded=True)
    print(f"{precision} max={maximize}: {r['status']} after {r['iters']} iterations", flush=True)
    assert r["status"] == "DUAL_INFEASIBLE"
    assert r["iters"] < 200_000
    o = oracle.OracleLP.from_export(p.export()).solve(tol=1e-6, max_iter=20_000, time_limit=60)
    q = p.solve_gpu(tol=1e-6, max_iter=20_000, time_limit=60, precision=precision)
    assert q["status"] == o["status"] != "DUAL_INFEASIBLE"


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_detect_unbounded_leaves_bounded_solves_unchanged(oracle, precision, monkeypatch):
    """On a bounded LP and an infeasible LP the extra test never fires: same status, objective and iterations.
    The per-matrix variant is pinned (no autotuning), so the two solves run the same kernels."""
    monkeypatch.setenv("GANIT_HIP_AUTOTUNE", "0")
    trip = [(0, 0, 1.0), (0, 1, 1.0), (1, 0, 1.0), (1, 1, -1.0), (2, 0, 3.0), (2, 2, 1.0)]
    bounded = _lp(3, 3, trip, [1.0, 2.0, -0.5], [0, 0, 0], [3, 3, 10], [1, -INF, 2], [4, INF, INF])
    trip = [(0, 0, 1.0), (0, 1, 1.0), (1, 0, 1.0), (1, 1, 1.0)]
    infeasible = _lp(2, 2, trip, [1.0, 1.0], [0, 0], [10, 10], [1, 3], [1, 3])
    for p in (bounded, infeasible, g.Problem.generate("random_lp", seed=2)):
        a = p.solve_gpu(tol=1e-6, time_limit=60, precision=precision)
        b = p.solve_gpu(tol=1e-6, time_limit=60, precision=precision, detect_unbounded=True)
        assert a["status"] == b["status"] != "DUAL_INFEASIBLE"
        assert a["iters"] == b["iters"] and a["pobj"] == b["pobj"]


@pytest.mark.parametrize("tail_lanes", [True, False])
def test_batch_of_edge_cases(oracle, tail_lanes):
    """The edge cases above as one batch -- an instance with no constraints, one with an empty row and column,
    a maximisation with ranged and free rows, free and one-sided variables, an infeasible system -- each
    instance against the oracle's solve of it alone.  Five tiny instances are a launch-bound tail from the first
    iteration, so tail_lanes=True runs them lanes-per-row throughout (the m = 0 instance has no rows of A there)
    and tail_lanes=False on row blocks."""
    probs = [
        _lp(0, 4, [], [1.0, -2.0, 0.5, -1.0], [0, 0, -1, 0], [3, 5, 2, 1], [], []),
        _lp(3, 4, [(0, 0, 1.0), (0, 1, 1.0), (2, 1, 1.0), (2, 3, 2.0)], [1, 2, -1, 1], [0, 0, 0, 0],
            [10, 10, 7, 10], [1, 0, 2], [INF, 2, 6]),
        _lp(3, 3, [(0, 0, 1.0), (0, 1, 1.0), (1, 0, 1.0), (1, 1, -1.0), (2, 0, 3.0), (2, 2, 1.0)], [1.0, 2.0, -0.5],
            [0, 0, 0], [3, 3, 10], [1, -INF, 2], [4, INF, INF], obj_const=5.0, maximize=True),
        _lp(2, 3, [(0, 0, 1.0), (0, 1, 1.0), (0, 2, 1.0), (1, 0, 1.0), (1, 2, -1.0)], [1.0, 1.0, -1.0],
            [-INF, 0, -INF], [INF, INF, 4], [2, 0], [2, 0]),
        _lp(2, 2, [(0, 0, 1.0), (0, 1, 1.0), (1, 0, 1.0), (1, 1, 1.0)], [1.0, 1.0], [0, 0], [10, 10], [1, 3],
            [1, 3]),
    ]
    tol = 1e-6
    b = g.Batch(probs, tol=tol, precision="double", time_limit=60, tail_lanes=tail_lanes)
    b.upload()
    rs = b.run()
    oracle.set_threads(2)
    for k, (p, r) in enumerate(zip(probs, rs)):
        o = oracle.OracleLP.from_export(p.export()).solve(tol=tol, time_limit=60)
        print(f"instance {k}: batch {r['status']} {r['pobj']:.10g} it {r['iters']} | oracle {o['status']} "
              f"{o['pobj']:.10g} it {o['iters']}", flush=True)
        assert r["status"] == o["status"]
        if not tail_lanes:  # row-order sums: the oracle's trajectory
            assert r["iters"] == o["iters"]
        if o["status"] == "OPTIMAL":
            assert abs(r["pobj"] - o["pobj"]) <= 100 * tol * (1 + abs(o["pobj"]))
            assert max(r["rel_p"], r["rel_d"], r["rel_gap"]) <= tol * (1 + 1e-7)
    if tail_lanes:
        assert b.stats()["tail_iters"] > 0
