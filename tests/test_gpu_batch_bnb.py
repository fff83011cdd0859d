"""Branch and bound on a resident batch: bound changes in place, warm starts from the parent's solution, and
option changes between runs.

The contract is the reference's warm-start test (tests/test_warm.cpp:8-38): after tightening one bound of a
fractional variable (upper bound to floor(x_j), or lower bound to ceil(x_j)), a warm re-solve must give the
status of a cold solve of the modified LP and its objective to 1e-6 (1 + |z|).  Here the children of a root
LP are instances of one batch: each gets its bounds with Batch.set_bounds (no rebuild) and the root's
solution and primal weight as its start (Batch.set_start); they are compared with a batch built from scratch
from the modified LPs.
"""
import numpy as np
import pytest

import ganit_v2_amd as g

pytestmark = pytest.mark.gpu

TOL = 1e-8


@pytest.fixture(autouse=True)
def _fixed_partition(monkeypatch):
    """Batches compared bit for bit must share their row-block partition (the FP64 check sums are added per
    block), so the timing-driven autotuner stays off here."""
    monkeypatch.setenv("GANIT_HIP_AUTOTUNE", "0")


def _root(seed=3):
    return g.Problem.generate("random_lp", seed=seed, m=300, n=800, m_eq=150, nnz_min=4, nnz_max=8)


def _rebuilt(p, l, u):
    """The same LP with other variable bounds, built from scratch (new scaling, new upload)."""
    e = p.export()
    A = e["A"]
    rows = np.repeat(np.arange(A.rows), np.diff(A.ptr))
    return g.Problem.from_triplets(A.rows, A.cols, rows, A.idx, A.val, e["c"], l, u, e["lo"], e["hi"],
                                   obj_const=e["obj_const"], maximize=e["maximize"])


def _children(p, x, count):
    """Up to `count` branchings (variable, direction) on fractional entries of x, as (l, u) pairs."""
    e = p.export()
    frac = x - np.floor(x)
    cand = [j for j in np.argsort(-np.minimum(frac, 1 - frac)) if 1e-3 < frac[j] < 1 - 1e-3]
    out = []
    for j in cand[: (count + 1) // 2]:
        for down in (True, False):
            l, u = e["l"].copy(), e["u"].copy()
            if down:
                u[j] = np.floor(x[j])
            else:
                l[j] = np.ceil(x[j])
            out.append((int(j), down, l, u))
    return out[:count]


def test_set_bounds_equals_a_rebuilt_batch():
    """Bounds replaced in place give the scaled problem a rebuild gives, bit for bit: the same iterations and
    objectives from the cold start."""
    root = _root()
    x = root.solve_gpu(precision="double", tol=1e-6)["x"]
    kids = _children(root, x, 6)
    b = g.Batch([root] * len(kids), precision="double", tol=1e-6)
    for k, (_, _, l, u) in enumerate(kids):
        b.set_bounds(k, l, u)
    inplace = b.run(return_solution=True)
    fresh = g.Batch([_rebuilt(root, l, u) for _, _, l, u in kids], precision="double", tol=1e-6).run(
        return_solution=True)
    for k, (a, c) in enumerate(zip(inplace, fresh)):
        assert a["status"] == c["status"], k
        assert a["iters"] == c["iters"] and a["restarts"] == c["restarts"], k
        assert a["pobj"] == c["pobj"], k
        assert np.array_equal(a["x"], c["x"]), k


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_warm_started_children_match_cold_solves(precision):
    root = _root()
    rb = g.Batch([root], precision="double", tol=TOL)
    r0 = rb.run(return_solution=True)[0]
    assert r0["status"] == "OPTIMAL"
    w0 = rb.weights()[0]
    assert w0 > 0
    kids = _children(root, r0["x"], 12)
    assert len(kids) >= 8
    warm = g.Batch([root] * len(kids), precision=precision, tol=TOL)
    for k, (_, _, l, u) in enumerate(kids):
        warm.set_bounds(k, l, u)
        warm.set_start(k, r0["x"], r0["y"], w0)
    rw = warm.run()
    rc = g.Batch([_rebuilt(root, l, u) for _, _, l, u in kids], precision=precision, tol=TOL).run()
    for k, ((j, down, _, _), a, c) in enumerate(zip(kids, rw, rc)):
        print(f"child {k} x[{j}] {'<= floor' if down else '>= ceil'}: warm {a['status']} it {a['iters']} "
              f"obj {a['pobj']:.10g} | cold {c['status']} it {c['iters']} obj {c['pobj']:.10g}", flush=True)
        assert a["status"] == c["status"], k
        if c["status"] == "OPTIMAL":
            assert max(a["rel_p"], a["rel_d"], a["rel_gap"]) <= TOL * (1 + 1e-7), k
            assert abs(a["pobj"] - c["pobj"]) <= 1e-6 * (1 + abs(c["pobj"])), k
    warm_it = [a["iters"] for a in rw]
    cold_it = [c["iters"] for c in rc]
    print(f"iterations: warm total {sum(warm_it)} median {np.median(warm_it)} | cold total {sum(cold_it)} "
          f"median {np.median(cold_it)}", flush=True)
    assert sum(warm_it) < sum(cold_it)
    assert np.median(warm_it) < np.median(cold_it)
    # back to the cold start: the rebuilt batch's run exactly
    for k in range(len(kids)):
        warm.set_start(k, None)
    again = warm.run()
    for a, c in zip(again, rc):
        assert a["iters"] == c["iters"] and a["pobj"] == c["pobj"]


@pytest.mark.parametrize("count", [1, 6])
def test_options_changed_between_runs_take_effect(count):
    """A batch run again with another tolerance follows the new tolerance, even where the launch lists are
    unchanged and a captured block of the first run could be replayed: the same result as a batch built with
    the new options."""
    probs = [_root(seed) for seed in range(10, 10 + count)]
    b = g.Batch(probs, precision="double", tol=1e-2)
    loose = b.run()
    b.set_options(precision="double", tol=1e-7, max_iter=200_000)
    tight = b.run()
    fresh = g.Batch(probs, precision="double", tol=1e-7, max_iter=200_000).run()
    for k, (lo, t, f) in enumerate(zip(loose, tight, fresh)):
        assert t["status"] == f["status"], k
        assert t["iters"] == f["iters"] and t["pobj"] == f["pobj"], k
        if t["status"] == "OPTIMAL":
            assert max(t["rel_p"], t["rel_d"], t["rel_gap"]) <= 1e-7 * (1 + 1e-7), k
            assert t["iters"] > lo["iters"], k


def test_set_options_keeps_the_built_precision():
    b = g.Batch([_root()], precision="double")
    with pytest.raises(RuntimeError):
        b.set_options(precision="mixed")
    with pytest.raises(RuntimeError):
        b.set_options(precision="double", check_every=63)


def test_set_bounds_rejects_crossed_bounds():
    p = _root()
    b = g.Batch([p], precision="double")
    e = p.export()
    l, u = e["l"].copy(), e["u"].copy()
    l[0], u[0] = 2.0, 1.0
    with pytest.raises(RuntimeError):
        b.set_bounds(0, l, u)
