"""The adjoint PDHG kernel (ganit_hip_pdhg_adjoint) and the torch.autograd.Function over it, against the
closed form of a separable QP, the active-set KKT reference gradient (oracle/adjoint_ref.py, itself pinned
by central differences of oracle solves) and FP64 central differences (h = 1e-6) of the oracle's solves."""
import numpy as np
import pytest
import torch

import ganit_v2_amd as g
from ganit_v2_amd.autograd import QPLayer
from oracle import adjoint_ref

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _solve(p, **opts):
    s = p.solver(**opts)
    s.upload()
    r = s.run(return_solution=True)
    assert r["status"] == "OPTIMAL", r["status"]
    return s, r


def test_adjoint_closed_form_separable_qp():
    rng = np.random.default_rng(3)
    n = 50
    q = rng.uniform(0.5, 2.0, n)
    c = rng.normal(size=n)
    w = rng.normal(size=n)
    b = 2.0
    x_exact, exact = adjoint_ref.separable_qp_closed_form(q, c, b, w)
    idx = np.arange(n, dtype=np.int32)
    p = g.Problem.from_triplets(1, n, np.zeros(n, np.int32), idx, np.ones(n), c, np.full(n, -100.0),
                                np.full(n, 100.0), [b], [b], qrows=idx, qcols=idx, qvals=q)
    s, r = _solve(p, tol=1e-10)
    assert np.max(np.abs(r["x"] - x_exact)) < 1e-7
    for method in ("gmres", "fixed_point"):
        a = s.adjoint(w, tol=1e-11, method=method, max_iter=200_000)
        print(method, "iters", a["iters"], "rel", a["rel_change"], flush=True)
        assert a["free_vars"] == n and a["active_rows"] == 1
        assert _rel(a["dc"], exact["dc"]) < 1e-6
        assert _rel(a["db"], exact["db"]) < 1e-6
        assert _rel(a["dq"], exact["dq"]) < 1e-6


@pytest.mark.parametrize("seed", [0, 1])
def test_adjoint_fp64_vs_active_set_gradient_c4(seed):
    """C4 shape: n = 2000, 300 equality + 200 <= rows, A 1% dense, Q = diag + 0.1% symmetric."""
    p = g.Problem.generate("diff_qp", seed=seed)
    d = p.export()
    s, r = _solve(p, tol=1e-9, time_limit=120)
    w = np.random.default_rng(100 + seed).normal(size=d["A"].cols)
    a = s.adjoint(w, tol=1e-10)
    ref = adjoint_ref.active_set_gradient(d["A"], d["Q"], r["x"], r["y"], d["l"], d["u"], d["lo"], d["hi"], w)
    errs = {k: _rel(a[k], ref[k]) for k in ("dc", "db", "dq")}
    print(f"seed {seed}: iters {a['iters']} free {a['free_vars']}/{ref['free']} active {a['active_rows']}/"
          f"{ref['active_rows']} rel errors {errs} adjoint {a['seconds']:.3f}s", flush=True)
    assert a["free_vars"] == ref["free"] and a["active_rows"] == ref["active_rows"]
    for k, e in errs.items():
        assert e < 1e-5, (k, e)


def test_adjoint_fp32_vs_active_set_gradient_c4():
    """FP32 forward (mixed ladder, FP64-certified) and FP32 adjoint; error stated against the FP64 reference."""
    p = g.Problem.generate("diff_qp", seed=2)
    d = p.export()
    s, r = _solve(p, tol=1e-6, precision="mixed", time_limit=120)
    w = np.random.default_rng(7).normal(size=d["A"].cols)
    a = s.adjoint(w, precision="single", tol=1e-5)
    _, r64 = _solve(p, tol=1e-9, time_limit=120)
    ref = adjoint_ref.active_set_gradient(d["A"], d["Q"], r64["x"], r64["y"], d["l"], d["u"], d["lo"], d["hi"], w)
    errs = {k: _rel(a[k], ref[k]) for k in ("dc", "db", "dq")}
    print(f"fp32 adjoint: iters {a['iters']} rel errors {errs}", flush=True)
    for k, e in errs.items():
        assert e < 1e-2, (k, e)


def test_adjoint_vs_oracle_central_differences(oracle):
    """Directional derivatives from FP64 central differences (h = 1e-6) of the oracle's CPU solves."""
    oracle.set_threads(8)
    p = g.Problem.generate("diff_qp", seed=5, n=300, m_eq=45, m_le=30, density=0.03, qdensity=0.005)
    d = p.export()
    s, r = _solve(p, tol=1e-10, time_limit=120)
    rng = np.random.default_rng(11)
    w = rng.normal(size=d["A"].cols)
    a = s.adjoint(w, tol=1e-11)
    h = 1e-6

    def loss(dd):
        # 1e-11: tight enough for h = 1e-6 differences, loose enough for the FP64 certification to pass
        o = oracle.OracleLP.from_export(dd).solve(tol=1e-11, time_limit=120)
        assert o["status"] == "OPTIMAL"
        return w @ o["x"]

    for trial in range(2):
        vc = rng.normal(size=d["A"].cols)
        vb = rng.normal(size=d["A"].rows)
        plus = dict(d, c=d["c"] + h * vc, lo=np.where(np.isfinite(d["lo"]), d["lo"] + h * vb, d["lo"]),
                    hi=d["hi"] + h * vb)
        minus = dict(d, c=d["c"] - h * vc, lo=np.where(np.isfinite(d["lo"]), d["lo"] - h * vb, d["lo"]),
                     hi=d["hi"] - h * vb)
        fd = (loss(plus) - loss(minus)) / (2 * h)
        an = a["dc"] @ vc + a["db"] @ vb
        print(f"trial {trial}: fd {fd:.10g} adjoint {an:.10g}", flush=True)
        assert abs(fd - an) <= 1e-5 * (1 + abs(fd))


def test_autograd_function_matches_adjoint():
    p = g.Problem.generate("diff_qp", seed=4, n=500, m_eq=75, m_le=50, density=0.02, qdensity=0.002)
    layer = QPLayer.from_problem(p, precision="double", tol=1e-9, adjoint_tol=1e-10)
    c, b, q = (t.clone().requires_grad_(True) for t in layer.data_of(p))
    w = torch.from_numpy(np.random.default_rng(9).normal(size=layer.n))
    x = layer(c, b, q)
    loss = (w * x).sum()
    loss.backward()
    d = p.export()
    r, a = layer.last
    ref = adjoint_ref.active_set_gradient(d["A"], d["Q"], r["x"], r["y"], d["l"], d["u"], d["lo"], d["hi"],
                                          w.numpy())
    assert np.array_equal(c.grad.numpy(), a["dc"]) and np.array_equal(q.grad.numpy(), a["dq"])
    assert np.array_equal(b.grad.numpy(), a["db"])
    for k, t in (("dc", c.grad), ("db", b.grad), ("dq", q.grad)):
        assert _rel(t.numpy(), ref[k]) < 1e-5, k
    # the layer's forward is the problem's own solution
    assert np.allclose(x.detach().numpy(), r["x"])


def test_set_solution_then_adjoint_matches_run_then_adjoint():
    """A solver that takes over another solver's original-space solution (Solver.set_solution) differentiates at
    the same point: its gradients equal those after its own run() to the adjoint's tolerance, and it refuses
    vectors of the wrong length."""
    p = g.Problem.generate("diff_qp", seed=5, n=400, m_eq=60, m_le=40, density=0.02, qdensity=0.002)
    s, r = _solve(p, precision="double", tol=1e-9)
    w = np.random.default_rng(12).normal(size=p.dims()["n"])
    a = s.adjoint(w, tol=1e-11)
    t = p.solver(precision="double", tol=1e-9)
    t.upload()
    t.set_solution(r["x"], r["y"])
    b = t.adjoint(w, tol=1e-11)
    for k in ("dc", "db", "dq"):
        assert _rel(b[k], a[k]) < 1e-7, k
    with pytest.raises(ValueError):
        t.set_solution(r["x"][:-1], r["y"])


@pytest.mark.parametrize("precision,tol,batch_forward", [("double", 1e-9, True), ("double", 1e-9, False),
                                                         ("mixed", 1e-6, True)])
def test_batched_layer_matches_single_instances(precision, tol, batch_forward):
    """A batch (B x n data) runs its forward solves as one batched solve (or concurrently) and its adjoints
    concurrently; each row's solution and gradients agree with the same instance through the single-instance
    layer."""
    p = g.Problem.generate("diff_qp", seed=4, n=500, m_eq=75, m_le=50, density=0.02, qdensity=0.002)
    layer = QPLayer.from_problem(p, precision=precision, tol=tol, adjoint_tol=1e-10, batch_forward=batch_forward)
    c0, b0, q0 = layer.data_of(p)
    rng = np.random.default_rng(3)
    k = 3
    C = torch.stack([c0 + 0.1 * torch.from_numpy(rng.normal(size=layer.n)) for _ in range(k)])
    B = torch.stack([b0] * k)
    Q = torch.stack([q0 * torch.from_numpy(rng.uniform(0.9, 1.1, layer.n)) for _ in range(k)])
    W = torch.from_numpy(rng.normal(size=(k, layer.n)))
    Cb, Bb, Qb = (t.clone().requires_grad_(True) for t in (C, B, Q))
    X = layer(Cb, Bb, Qb)
    assert X.shape == (k, layer.n)
    (W * X).sum().backward()
    # A mixed solve of these QPs ends on its FP32 iterate (a few hundred iterations with the Q-aware scaling): the
    # batch and the single solve take the same iterations and restarts, but the single solver's autotuned kernel
    # variant may sum a row in another order, so the two points agree to a few FP32 ulps, not to FP64 rounding.
    x_tol = 1e-7 if precision == "double" else 1e-6
    for i in range(k):
        c, b, q = (t[i].clone().requires_grad_(True) for t in (C, B, Q))
        x = layer(c, b, q)
        (W[i] * x).sum().backward()
        assert _rel(X[i].detach().numpy(), x.detach().numpy()) < x_tol, i
        for bt, st in ((Cb.grad[i], c.grad), (Bb.grad[i], b.grad), (Qb.grad[i], q.grad)):
            assert _rel(bt.numpy(), st.numpy()) < 1e-5, i


def test_solve_many_matches_single_solves():
    probs = [g.Problem.generate("random_lp", seed=s, m=400, n=900, m_eq=200) for s in range(6)]
    many = g.solve_many(probs, threads=6, tol=1e-6, precision="double")
    for p, r in zip(probs, many):
        s = p.solver(tol=1e-6, precision="double")
        s.upload()
        one = s.run(return_solution=True)
        assert r["status"] == one["status"]
        assert abs(r["pobj"] - one["pobj"]) <= 1e-5 * (1 + abs(one["pobj"]))


@pytest.mark.parametrize("seed", [0, 1, 2, 3, 4, 5])
def test_scale_q_adjoint_vs_active_set_gradient_c4(seed):
    """With the Q-aware column scaling (the QP layer's default) the C4 QPs converge in hundreds of PDHG
    iterations instead of 1e5-1e7; the solution map and its gradient are the same: FP64 gradients against the
    active-set KKT reference < 1e-6, FP32 (mixed forward + FP32 adjoint) < 1e-3, and the optimum agrees with the
    unscaled-Q solve's."""
    p = g.Problem.generate("diff_qp", seed=seed)
    d = p.export()
    s, r = _solve(p, tol=1e-9, time_limit=120, scale_q=True)
    assert r["iters"] < 20_000
    w = np.random.default_rng(100 + seed).normal(size=d["A"].cols)
    a = s.adjoint(w, tol=1e-10)
    ref = adjoint_ref.active_set_gradient(d["A"], d["Q"], r["x"], r["y"], d["l"], d["u"], d["lo"], d["hi"], w)
    errs = {k: _rel(a[k], ref[k]) for k in ("dc", "db", "dq")}
    s32, r32 = _solve(p, tol=1e-6, precision="mixed", time_limit=120, scale_q=True)
    a32 = s32.adjoint(w, precision="single", tol=1e-5)
    errs32 = {k: _rel(a32[k], ref[k]) for k in ("dc", "db", "dq")}
    print(f"seed {seed}: forward {r['iters']} / {r32['iters']} iterations, adjoint {a['iters']} / {a32['iters']}, "
          f"FP64 rel errors {errs}, FP32 rel errors {errs32}", flush=True)
    for k in errs:
        assert errs[k] < 1e-6, (k, errs[k])
        assert errs32[k] < 1e-3, (k, errs32[k])
    if seed in (0, 1):  # the others need 3e6+ iterations without it
        plain = p.solver(tol=1e-6, time_limit=60).run()
        assert plain["status"] == "OPTIMAL"
        assert abs(plain["pobj"] - r["pobj"]) <= 1e-5 * (1 + abs(r["pobj"]))


def test_batch_adjoint_vs_active_set_gradient_and_single_solves():
    """Batch.adjoint: every instance's GMRES adjoint in lockstep on the block-diagonal matrices.  Four C4-shape QPs
    and a smaller one in one batch (uneven instances): FP64 gradients against the active-set KKT reference < 1e-6
    and against each instance's own single-solver adjoint at the same solution < 1e-7; FP32 (mixed forward to 1e-6 +
    FP32 batched adjoint) against the reference gradient at the FP64 solution and against the single-solver FP32
    adjoint at the same FP32 solution < 1e-3."""
    probs = [g.Problem.generate("diff_qp", seed=s) for s in range(4)]
    probs.append(g.Problem.generate("diff_qp", seed=4, n=500, m_eq=75, m_le=50, density=0.02, qdensity=0.002))
    ws = [np.random.default_rng(100 + k).normal(size=p.dims()["n"]) for k, p in enumerate(probs)]
    bt = g.Batch(probs, precision="double", tol=1e-9, time_limit=120, scale_q=True)
    res = bt.run(return_solution=True)
    assert all(r["status"] == "OPTIMAL" for r in res)
    adj = bt.adjoint(ws, tol=1e-11)
    bt32 = g.Batch(probs, precision="mixed", tol=1e-6, time_limit=120, scale_q=True)
    res32 = bt32.run(return_solution=True)
    assert all(r["status"] == "OPTIMAL" for r in res32)
    adj32 = bt32.adjoint(ws, precision="single", tol=1e-5)
    for k, (p, r, a, r32, a32) in enumerate(zip(probs, res, adj, res32, adj32)):
        d = p.export()
        ref = adjoint_ref.active_set_gradient(d["A"], d["Q"], r["x"], r["y"], d["l"], d["u"], d["lo"], d["hi"], ws[k])
        s = p.solver(precision="double", tol=1e-9, scale_q=True)
        s.upload()
        s.set_solution(r["x"], r["y"])
        one = s.adjoint(ws[k], tol=1e-11)
        errs = {key: _rel(a[key], ref[key]) for key in ("dc", "db", "dq")}
        errs1 = {key: _rel(a[key], one[key]) for key in ("dc", "db", "dq")}
        errs32 = {key: _rel(a32[key], ref[key]) for key in ("dc", "db", "dq")}
        s32 = p.solver(precision="mixed", tol=1e-6, scale_q=True)
        s32.upload()
        s32.set_solution(r32["x"], r32["y"], fp32_iterate=r32["fp32_iters"] == r32["iters"])
        one32 = s32.adjoint(ws[k], precision="single", tol=1e-5)
        errs32_1 = {key: _rel(a32[key], one32[key]) for key in ("dc", "db", "dq")}
        print(f"instance {k}: FP32 vs single FP32 {errs32_1}", flush=True)
        print(f"instance {k}: products {a['iters']} / {a32['iters']} (single {one['iters']}), free {a['free_vars']}/"
              f"{ref['free']}, active {a['active_rows']}/{ref['active_rows']}; vs reference {errs}, vs single "
              f"{errs1}, FP32 vs reference {errs32}", flush=True)
        assert a["free_vars"] == one["free_vars"] and a["active_rows"] == one["active_rows"]
        for key in errs:
            assert errs[key] < 1e-6, (k, key, errs[key])
            assert errs1[key] < 1e-7, (k, key, errs1[key])
            assert errs32[key] < 1e-3, (k, key, errs32[key])
            assert errs32_1[key] < 1e-3, (k, key, errs32_1[key])
    # arguments are checked before anything runs
    with pytest.raises(ValueError):
        bt.adjoint(ws[:-1])
    with pytest.raises(RuntimeError):
        bt.adjoint(ws, precision="single")  # an FP64 batch has no FP32 working set


def test_batch_adjoint_instances_are_independent():
    """Each instance's GMRES in the batched adjoint runs on its own dot products, Hessenberg system and stopping
    test, so its result does not depend on the other instances: an instance solved beside others gives the same
    bits as in a batch of its own (here beside a larger QP and an LP, whose empty Q rows the batch carries).  The
    forward keeps its row blocks to the end (tail_lanes=False), so it too is each instance's own, bit for bit."""
    small = g.Problem.generate("diff_qp", seed=7, n=400, m_eq=60, m_le=40, density=0.03, qdensity=0.003)
    big = g.Problem.generate("diff_qp", seed=8, n=900, m_eq=120, m_le=80, density=0.02, qdensity=0.002)
    lp = g.Problem.generate("random_lp", seed=3, m=120, n=300, m_eq=60)
    rng = np.random.default_rng(5)
    ws = [rng.normal(size=p.dims()["n"]) for p in (small, big, lp)]
    together = g.Batch([small, big, lp], precision="double", tol=1e-8, scale_q=True, tail_lanes=False)
    r3 = together.run(return_solution=True)
    alone = g.Batch([small], precision="double", tol=1e-8, scale_q=True, tail_lanes=False)
    r1 = alone.run(return_solution=True)
    assert r3[0]["status"] == r1[0]["status"] == "OPTIMAL" and r3[1]["status"] == "OPTIMAL"
    assert np.array_equal(r3[0]["x"], r1[0]["x"])  # the forward is per instance too
    a3 = together.adjoint(ws, tol=1e-10)
    a1 = alone.adjoint(ws[:1], tol=1e-10)
    print(f"products {a3[0]['iters']} / {a1[0]['iters']}, big {a3[1]['iters']}, lp {a3[2]['iters']}", flush=True)
    assert a3[0]["iters"] == a1[0]["iters"]
    for key in ("dc", "db", "dq"):
        assert np.array_equal(a3[0][key], a1[0][key]), key
        assert np.all(np.isfinite(a3[1][key])) and np.all(np.isfinite(a3[2][key])), key
