"""Polish of PDHG points onto the face their active set names (option polish; the reference lists crossover /
polish as planned, README.md:195), against the reference's dual simplex objectives on the bundled Netlib LPs
(tests/golden/reference_runs.jsonl, `--method simplex` lines): a polished point must be FP64-certified to 1e-9 in
every KKT term and its objective must equal the simplex vertex's to 1e-9 relative.  Where no polished point passes,
the PDHG point is returned untouched."""
import numpy as np
import pytest

import ganit_v2_amd as g
from conftest import golden_runs, instance_path

pytestmark = pytest.mark.gpu


def _simplex(name):
    return next(r for r in golden_runs() if r["instance"] == name and r["args"][:2] == ["--method", "simplex"])


@pytest.mark.parametrize("name", ["afiro", "adlittle", "stair", "shell", "25fv47"])
def test_polish_reaches_the_simplex_vertex(name):
    ref = _simplex(name)
    assert ref["status"] == "OPTIMAL"
    p = g.Problem.read_mps(instance_path(name))
    s = p.solver(precision="double", tol=1e-8, polish=True, time_limit=120)
    s.upload()
    r = s.run(return_solution=True)
    err = abs(r["pobj"] - ref["pobj"]) / (1 + abs(ref["pobj"]))
    print(f"{name}: {r['status']} it {r['iters']} polished {s.polished()} obj {r['pobj']!r} | simplex "
          f"{ref['pobj']!r} rel err {err:.2e}", flush=True)
    assert r["status"] == "OPTIMAL" and s.polished()
    assert max(r["rel_p"], r["rel_d"], r["rel_gap"]) <= 1e-9
    assert err <= 1e-9
    # the vertex is primal feasible in the original space too
    e = p.export()
    ax = e["A"].matvec(r["x"])
    scale = 1 + np.abs(ax)
    assert np.all(ax >= e["lo"] - 1e-9 * scale) and np.all(ax <= e["hi"] + 1e-9 * scale)
    assert np.all(r["x"] >= e["l"] - 1e-9 * (1 + np.abs(e["l"]))) and np.all(r["x"] <= e["u"] + 1e-9 * (1 + np.abs(e["u"])))


def test_polish_off_leaves_the_solve_unchanged(monkeypatch):
    monkeypatch.setenv("GANIT_HIP_AUTOTUNE", "0")  # two solvers compared bit for bit: the same kernel variants
    p = g.Problem.read_mps(instance_path("adlittle"))
    a = p.solver(precision="double", tol=1e-6).run(return_solution=True)
    s = p.solver(precision="double", tol=1e-6, polish=False)
    b = s.run(return_solution=True)
    assert not s.polished()
    assert a["iters"] == b["iters"] and a["pobj"] == b["pobj"] and np.array_equal(a["x"], b["x"])


def test_polish_that_fails_returns_the_pdhg_point(monkeypatch):
    """etamacro: the active set read off a 1e-8 point leaves a duality gap above 1e-9 at every activity tolerance
    tried, so the PDHG point itself comes back (same iterations, objective and solution as without polish)."""
    monkeypatch.setenv("GANIT_HIP_AUTOTUNE", "0")  # two solvers compared bit for bit: the same kernel variants
    p = g.Problem.read_mps(instance_path("etamacro"))
    a = p.solver(precision="double", tol=1e-8, time_limit=120).run(return_solution=True)
    s = p.solver(precision="double", tol=1e-8, polish=True, time_limit=120)
    b = s.run(return_solution=True)
    print(f"etamacro polished {s.polished()}: obj {b['pobj']!r} vs {a['pobj']!r}", flush=True)
    if not s.polished():
        assert b["pobj"] == a["pobj"] and np.array_equal(b["x"], a["x"])
    else:
        assert abs(b["pobj"] - _simplex("etamacro")["pobj"]) <= 1e-9 * (1 + abs(b["pobj"]))


def test_polish_ignores_qps():
    p = g.Problem.read_mps(instance_path("port_tiny"))
    s = p.solver(precision="double", tol=1e-6, polish=True)
    r = s.run()
    assert r["status"] == "OPTIMAL" and not s.polished()
