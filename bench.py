"""Benchmark: time to 1e-4 relative KKT (FP64-certified) of restarted PDHG on the ~2.9M-nnz blending LP,
mixed precision (L1 FP32 -> L3 FP64) vs plain FP64, on one MI355X; plus the fused SpMV kernels' HBM GB/s.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--cpu-sample-s S]

A step is one full solve from the initial point with the scaled problem already resident in HBM (upload,
scaling and norm estimation happen before the timed region); "with_upload" reports a fresh solver's first solve,
whose solve_s includes the upload as the reference's does.  Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
WORKLOAD = dict(crudes=60, products=16, specs=5, periods=365)


def kernel_bytes(m, n, nnz, s_m, s_v, ib_at=4, ib_a=4, bound_vectors=2):
    """Algorithmic bytes per launch of the two fused iteration kernels, with the column-index width each
    matrix is read with (2 with compact indices, 4 otherwise).
    primal: A^T (row pointers + nnz index/value) + gathered y + x, c, [l, u,] sumx read + xn, xbar, sumx
            written (bound_vectors = how many of l, u differ between columns and are streamed: a bound that
            is the same in every column is a kernel argument)
    dual:   A   (row pointers + nnz index/value) + gathered xbar + y, lo, hi, sumy read + yn, sumy written"""
    primal = 4 * (n + 1) + nnz * (ib_at + s_m) + m * s_v + (6 + bound_vectors) * n * s_v
    dual = 4 * (m + 1) + nnz * (ib_a + s_m) + n * s_v + 6 * m * s_v
    return primal, dual


def streamed_bound_vectors(prob):
    """How many of the scaled column-bound vectors l, u the primal kernel streams (a vector whose entries are
    all the same bits is passed as a scalar instead)."""
    import numpy as np
    sc = prob.scaled()
    return sum(int(len(v) > 0 and not np.all(v.view(np.uint64) == v.view(np.uint64)[0])) for v in (sc["l"], sc["u"]))


def timed_kernels(solver, level, reps, m, n, nnz, s_m, s_v, bound_vectors):
    """HIP-event times of the fused kernels at `level` (0 L1, 2 L3) and their algorithmic GB/s: each kernel with
    none of its data in L2 (as in a solve, where the two alternate), the whole iteration in solve order, and
    each kernel back to back (part of its data then stays in L2 between launches)."""
    t = solver.kernel_times(level, reps)
    bp, bd = kernel_bytes(m, n, nnz, s_m, s_v, solver.index_bytes(1), solver.index_bytes(0), bound_vectors)
    return dict(primal_us=t["primal"] * 1e3, dual_us=t["dual"] * 1e3, iteration_us=t["iteration"] * 1e3,
                primal_gbs=bp / (t["primal"] * 1e6), dual_gbs=bd / (t["dual"] * 1e6),
                iteration_gbs=(bp + bd) / (t["iteration"] * 1e6), primal_bytes=bp, dual_bytes=bd,
                primal_b2b_us=t["primal_b2b"] * 1e3, dual_b2b_us=t["dual_b2b"] * 1e3,
                primal_b2b_gbs=bp / (t["primal_b2b"] * 1e6), dual_b2b_gbs=bd / (t["dual_b2b"] * 1e6),
                sweep_us=t["sweep"] * 1e3, sweep_gbs=(64 << 20) / (t["sweep"] * 1e6),
                index_bytes_at=solver.index_bytes(1), index_bytes_a=solver.index_bytes(0))


def reference_sequence_bytes(m, n, nnz, s_m, s_v):
    """B_iter of the reference's un-fused kernel sequence (2 SpMV + primal + dual)."""
    return 2 * nnz * (4 + s_m) + 4 * (m + n + 2) + s_v * (11 * n + 9 * m)


def load_traffic():
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def cpu_model():
    """The host CPU as the OS names it (/proc/cpuinfo's model name, with family / model / stepping, and the
    logical CPUs this process may use)."""
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" in line:
                    k, v = (t.strip() for t in line.split(":", 1))
                    info.setdefault(k, v)
    except OSError:
        pass
    name = info.get("model name", "unknown")
    fam = ", ".join(f"{k} {info[k]}" for k in ("cpu family", "model", "stepping") if k in info)
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return f"{name} ({fam}; {usable} logical CPUs usable)" if fam else f"{name} ({usable} logical CPUs usable)"


def reference_cpu_baseline(problem, threads, time_limit):
    """The reference's own CPU path: its CLI built from its sources (oracle/_ref/ganit), --device cpu
    --precision mixed with its defaults, solving the same LP written as MPS.  Returns None when the binary is
    not there (it is built from the reference's sources where they are present)."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "oracle", "_ref", "ganit")
    if not os.access(exe, os.X_OK):
        return None
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "workload.mps")
        problem.write_mps(path)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        out = subprocess.run([exe, path, "--device", "cpu", "--precision", "mixed", "--tol", "1e-4", "--time",
                              str(time_limit), "--json", "--quiet"], capture_output=True, text=True, env=env,
                             timeout=time_limit + 120)
        # one core, a bounded sample (640 iterations): seconds per iteration, for single-core comparisons
        one = subprocess.run([exe, path, "--device", "cpu", "--precision", "mixed", "--iters", "640", "--time",
                              str(time_limit), "--json", "--quiet"], capture_output=True, text=True,
                             env=dict(os.environ, OMP_NUM_THREADS="1"), timeout=time_limit + 120)
    one_lines = [l for l in one.stdout.splitlines() if l.startswith("{")]
    one_rate = None
    if one_lines:
        o = json.loads(one_lines[-1])
        one_rate = o["solve_s"] / max(o["iters"], 1)
    lines = [l for l in out.stdout.splitlines() if l.startswith("{")]
    if out.returncode not in (0, 2) or not lines:
        print(f"[bench] reference CPU run failed ({out.returncode}): {out.stderr[-300:]}", file=sys.stderr)
        return None
    r = json.loads(lines[-1])
    return {
        "value": r["solve_s"], "unit": "s", "cores": threads, "kind": "reference", "cpu_model": cpu_model(),
        "single_core_s_per_iter": one_rate,
        "sample": (f"the reference CLI built from its sources, --device cpu --precision mixed (its defaults: stall "
                   f"12, level-precision checks), full solve of the same LP written as MPS: {r['status']} in "
                   f"{r['iters']} iterations ({r['fp32_iters']} L1), solve_s as it reports it (read "
                   f"{r['read_s']:.2f} s and setup {r['setup_s']:.2f} s excluded)"),
        "status": r["status"], "iters": r["iters"], "pobj": r["pobj"], "dobj": r["dobj"],
    }


def cpu_baseline(problem, gpu_run, sample_s):
    """The CPU baseline: the reference's CPU path where its binary is built (reference_cpu_baseline), else the
    oracle (CPU restatement of the reference path, C++/OpenMP) on this host -- a bounded sample of the same
    solve, giving seconds per iteration at each level; the estimate is the GPU run's iteration split times
    those rates."""
    ref = reference_cpu_baseline(problem, os.cpu_count() or 1, max(30, int(4 * sample_s)))
    if ref is not None:
        return ref
    from oracle import pyoracle
    threads = os.cpu_count() or 1
    pyoracle.set_threads(threads)
    o = pyoracle.OracleLP.from_export(problem.export())
    rates = {}
    # L1 rate from a time-limited mixed run (its first phase is L1), L3 rate from a time-limited FP64 run
    r1 = o.solve(precision="mixed", time_limit=sample_s / 2, stall_checks=10 ** 9)
    rates["l1"] = r1["solve_s"] / max(r1["iters"], 1)
    r3 = o.solve(precision="double", time_limit=sample_s / 2)
    rates["l3"] = r3["solve_s"] / max(r3["iters"], 1)
    l1 = gpu_run["fp32_iters"]
    l3 = gpu_run["iters"] - l1
    est = l1 * rates["l1"] + l3 * rates["l3"]
    return {
        "value": est, "unit": "s", "cores": threads, "kind": "port",
        "sample": (f"oracle mixed run capped at {sample_s / 2:.0f}s ({r1['iters']} L1 iterations) and FP64 run "
                   f"capped at {sample_s / 2:.0f}s ({r3['iters']} iterations); estimate = GPU iteration split "
                   f"(L1 {l1}, L3 {l3}) x measured CPU s/iteration (L1 {rates['l1'] * 1e3:.2f} ms, "
                   f"L3 {rates['l3'] * 1e3:.2f} ms)"),
    }


def portfolio_ladder():
    """Factor-model portfolio QP (100,000 assets x 50 factors, tol 1e-6): FP64 vs the FP32-matrix ladder, each
    FP64-certified; one timed solve after a warm-up, problem resident."""
    import ganit_v2_amd as g
    p = g.Problem.generate("portfolio")
    d = p.dims()
    out = dict(workload="portfolio QP 100000 assets x 50 factors, seed 0, tol 1e-6", m=d["m"], n=d["n"], nnz=d["nnz"])
    runs = {}
    for name, cfg in (("fp64", dict(precision="double")), ("mixed", dict(precision="mixed")),
                      ("dynamic", dict(precision="dynamic")),
                      # extension: the column scaling also equilibrates Q (a different trajectory, same optimum)
                      ("fp64_scale_q", dict(precision="double", scale_q=True)),
                      ("mixed_scale_q", dict(precision="mixed", scale_q=True))):
        s = p.solver(tol=1e-6, time_limit=120, **cfg)
        s.upload()
        s.run()
        r = s.run()
        runs[name] = dict(solve_s=round(r["solve_s"], 5), status=r["status"], iters=r["iters"],
                          l1_iters=r["fp32_iters"], obj=r["pobj"],
                          certified_rel=[r["rel_p"], r["rel_d"], r["rel_gap"]])
        print(f"[bench] portfolio {name}: {runs[name]}", file=sys.stderr, flush=True)
    f = runs["fp64"]
    for name in ("mixed", "dynamic", "fp64_scale_q", "mixed_scale_q"):
        runs[name]["speedup_vs_fp64"] = round(f["solve_s"] / runs[name]["solve_s"], 3)
        runs[name]["obj_rel_err_vs_fp64"] = abs(runs[name]["obj"] - f["obj"]) / (1 + abs(f["obj"]))
    out["runs"] = runs
    return out


def batched_lps(count=512, sequential=True):
    """512 uneven random LPs (node-relaxation shapes): the batched kernels (FP64; FP32 + FP64 certification with
    the reference's ladder rules; the same with FP64 candidate checks and stall 60, the headline's configuration;
    and that with the FP32 level ended at the launch-bound tail) against sequential single-instance FP64 solves
    with each problem resident (sequential=False: the batched runs only)."""
    import numpy as np
    import ganit_v2_amd as g
    probs = [g.Problem.generate("batch_lp", seed=k) for k in range(count)]
    out = dict(workload=f"{count} random LPs, m log-U[200,4000], n log-U[500,10000], 5-15 nonzeros per column, "
                        f"5% tightened bounds, seeds 0..{count - 1}, tol 1e-4")
    res = {}
    configs = (("batched_fp64", dict(precision="double")), ("batched_mixed", dict(precision="mixed")),
               ("batched_mixed_fp64_checks", dict(precision="mixed", eval_fp64=True, stall_checks=60)),
               # a batch policy (not the reference's ladder rule): the FP32 level ends at the launch-bound tail
               ("batched_mixed_fp64_checks_tail_fp64",
                dict(precision="mixed", eval_fp64=True, stall_checks=60, tail_fp64=True)))
    for name, cfg in configs:
        b = g.Batch(probs, tol=1e-4, **cfg)
        b.upload()
        b.run()
        t = time.perf_counter()
        rs = b.run()
        wall = time.perf_counter() - t
        st = b.stats()
        res[name] = dict(solve_s=round(st["solve_s"], 4), wall_s=round(wall, 4), max_iters=st["max_iters"],
                         total_iters=st["total_iters"], l3_instances=st["l3_instances"],
                         tail_lanes_iters=st["tail_iters"],
                         statuses={k: sum(1 for r in rs if r["status"] == k) for k in {r["status"] for r in rs}})
        res[name + "_results"] = rs
        print(f"[bench] batched {name}: {res[name]}", file=sys.stderr, flush=True)
    if not sequential:
        for name, _ in configs:
            res.pop(name + "_results")
        out.update(res)
        return out
    seq, tsum = [], 0.0
    t = time.perf_counter()
    for p in probs:
        s = p.solver(precision="double", tol=1e-4)
        s.upload()
        r = s.run()
        seq.append(r)
        tsum += r["solve_s"]
    wall = time.perf_counter() - t
    res["sequential_fp64"] = dict(solve_s_sum=round(tsum, 4), wall_s=round(wall, 3))
    for name, _ in configs:
        rs = res.pop(name + "_results")
        errs = [abs(a["pobj"] - b["pobj"]) / (1 + abs(b["pobj"])) for a, b in zip(rs, seq)
                if a["status"] == b["status"] == "OPTIMAL"]
        both = [k for k, (a, b) in enumerate(zip(rs, seq)) if a["status"] == b["status"] == "OPTIMAL"]
        worst = both[int(np.argmax(errs))]
        res[name]["largest_obj_difference"] = dict(
            seed=worst, obj=rs[worst]["pobj"], obj_sequential_fp64=seq[worst]["pobj"], rel_err=max(errs),
            certified_rel=[rs[worst]["rel_p"], rs[worst]["rel_d"], rs[worst]["rel_gap"]],
            iters=rs[worst]["iters"], l1_iters=rs[worst]["fp32_iters"])
        res[name]["certified_all"] = all(max(r["rel_p"], r["rel_d"], r["rel_gap"]) <= 1e-4 * (1 + 1e-7)
                                         for r in rs if r["status"] == "OPTIMAL")
        res[name]["status_mismatches"] = sum(1 for a, b in zip(rs, seq) if a["status"] != b["status"])
        res[name]["same_iterations"] = sum(1 for a, b in zip(rs, seq) if a["iters"] == b["iters"])
        res[name]["obj_rel_err_vs_sequential_fp64"] = dict(max=max(errs), median=float(np.median(errs)))
        res[name]["speedup_vs_sequential_fp64"] = round(tsum / res[name]["solve_s"], 2)
    print(f"[bench] batched summary: {res}", file=sys.stderr, flush=True)
    out.update(res)
    return out


def diff_qps(count, threads=32, batched=True, scale_q=True, cd_h=1e-6):
    """Differentiable QPs (n=2000, 300 equality + 200 <= rows): the FP32 path -- forward (mixed ladder, tol 1e-6)
    + FP32 adjoint -- timed as one batch, then the FP64 path -- forward (tol 1e-9) + FP64 adjoint -- and each
    instance's FP32 gradient error against it; then each instance's FP64 gradient against FP64 central differences
    of the loss along the gradient's own unit direction in c, in b and in diag Q (step cd_h, the 6 perturbed QPs of
    every instance solved as one batch at tol 1e-11).  batched: the forwards of a phase run as one batched solve
    (ganit_v2_amd.Batch) and the adjoints as one batched GMRES on the same batch (Batch.adjoint); else every
    instance runs forward and adjoint in its own host thread.  scale_q: the column scaling also equilibrates Q (the
    QP layer's default)."""
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    import ganit_v2_amd as g
    from ganit_v2_amd.autograd import QPLayer
    probs = [g.Problem.generate("diff_qp", seed=seed) for seed in range(count)]
    ws = [np.random.default_rng(1000 + seed).normal(size=p.dims()["n"]) for seed, p in enumerate(probs)]

    def path(args, precision, tol, adj_precision, adj_tol):
        p, w = args
        s = p.solver(precision=precision, tol=tol, time_limit=300, scale_q=scale_q)
        s.upload()
        r = s.run(return_solution=True)
        a = s.adjoint(w, precision=adj_precision, tol=adj_tol) if r["status"] == "OPTIMAL" else None
        return r, a

    adj_walls = {}

    def batch_path(precision, tol, adj_precision, adj_tol):
        # the batch's clock is shared by its instances; 150 s keeps each phase's silent stretch short
        bt = g.Batch(probs, precision=precision, tol=tol, time_limit=150, scale_q=scale_q)
        res = bt.run(return_solution=True)
        n_opt = sum(r["status"] == "OPTIMAL" for r in res)
        t0 = time.perf_counter()
        adj = bt.adjoint(ws, precision=adj_precision, tol=adj_tol)
        adj_walls[precision] = time.perf_counter() - t0
        print(f"[bench] diff_qp {precision} batch: {n_opt}/{count} OPTIMAL, max {max(r['iters'] for r in res)} "
              f"iterations, forward {res[0]['solve_s']:.2f} s; batched adjoint {adj_walls[precision]:.3f} s, "
              f"max {max(a['iters'] for a in adj)} GMRES products", file=sys.stderr, flush=True)
        return [(r, a if r["status"] == "OPTIMAL" else None) for r, a in zip(res, adj)]

    nt = max(1, min(threads, count))
    t = time.perf_counter()
    if batched:
        f32 = batch_path("mixed", 1e-6, "single", 1e-5)
    else:
        with ThreadPoolExecutor(nt) as ex:
            f32 = list(ex.map(lambda a: path(a, "mixed", 1e-6, "single", 1e-5), zip(probs, ws)))
    wall32 = time.perf_counter() - t
    t = time.perf_counter()
    if batched:
        f64 = batch_path("double", 1e-9, "double", 1e-10)
    else:
        with ThreadPoolExecutor(nt) as ex:
            f64 = list(ex.map(lambda a: path(a, "double", 1e-9, "double", 1e-10), zip(probs, ws)))
    wall64 = time.perf_counter() - t
    # central differences of L = w'x* along the FP64 gradient's own direction in c, in b and in diag Q (unit
    # vectors; the largest change of L, so the difference quotient is not lost in the solves' tolerance), all
    # 6 count perturbed QPs solved at once
    t = time.perf_counter()
    layers = [QPLayer.from_problem(p) for p in probs]
    cd_probs, dirs = [], []
    for k, (p, lay) in enumerate(zip(probs, layers)):
        c, b, q = (v.numpy() for v in lay.data_of(p))
        a64 = f64[k][1]
        rng = np.random.default_rng(2000 + k)
        unit = lambda v: v / max(np.linalg.norm(v), 1e-300)
        dc, db, dq = ((unit(a64[key]) if a64 is not None and np.linalg.norm(a64[key]) > 0 else unit(rng.normal(size=n)))
                      for key, n in (("dc", c.size), ("db", b.size), ("dq", q.size)))
        dirs.append((dc, db, dq))
        for sgn in (1.0, -1.0):
            cd_probs += [lay.problem(c + sgn * cd_h * dc, b, q), lay.problem(c, b + sgn * cd_h * db, q),
                         lay.problem(c, b, q + sgn * cd_h * dq)]
    cd_res = g.Batch(cd_probs, precision="double", tol=1e-11, time_limit=300, scale_q=scale_q).run(
        return_solution=True)
    wall_cd = time.perf_counter() - t
    rel = lambda u, v: float(np.linalg.norm(u - v) / max(np.linalg.norm(v), 1e-300))
    rows = []
    for seed, ((r32, a32), (r64, a64)) in enumerate(zip(f32, f64)):
        row = dict(seed=seed, fwd32_status=r32["status"], fwd32_iters=r32["iters"], fwd64_status=r64["status"],
                   fwd64_iters=r64["iters"])
        if a32 is not None and a64 is not None:
            row.update(adj32_s=round(a32["seconds"], 3), adj64_s=round(a64["seconds"], 3), adj64_iters=a64["iters"],
                       grad_err_fp32={k: rel(a32[k], a64[k]) for k in ("dc", "db", "dq")})
        if a64 is not None:
            plus, minus = cd_res[6 * seed: 6 * seed + 3], cd_res[6 * seed + 3: 6 * seed + 6]
            cd = {}
            for j, key in enumerate(("dc", "db", "dq")):
                if plus[j]["status"] == minus[j]["status"] == "OPTIMAL":
                    fd = (ws[seed] @ plus[j]["x"] - ws[seed] @ minus[j]["x"]) / (2 * cd_h)
                    ad = float(a64[key] @ dirs[seed][j])
                    cd[key] = float(abs(ad - fd) / max(abs(fd), 1e-12))
            row["grad_err_fp64_vs_central_differences"] = cd
        rows.append(row)
    errs = [max(r["grad_err_fp32"].values()) for r in rows if "grad_err_fp32" in r]
    cds = [max(r["grad_err_fp64_vs_central_differences"].values()) for r in rows
           if r.get("grad_err_fp64_vs_central_differences")]
    out = dict(workload=f"{count} differentiable QPs n=2000, 300 E + 200 L rows, seeds 0..{count - 1}",
               forward="one batched solve per phase" if batched else "concurrent single solves",
               scale_q=scale_q, threads=nt, fp32_batch_wall_s=round(wall32, 3), fp64_check_wall_s=round(wall64, 3),
               central_difference_wall_s=round(wall_cd, 3), cd_step=cd_h,
               adjoint_wall_s={k: round(v, 4) for k, v in adj_walls.items()} if batched else None,
               optimal_fp32=sum(r["fwd32_status"] == "OPTIMAL" for r in rows),
               optimal_fp64=sum(r["fwd64_status"] == "OPTIMAL" for r in rows),
               fwd32_iters_max=max(r["fwd32_iters"] for r in rows), fwd64_iters_max=max(r["fwd64_iters"] for r in rows),
               grad_err_fp32_vs_fp64_max=max(errs) if errs else None,
               grad_err_fp32_vs_fp64_median=float(np.median(errs)) if errs else None,
               grad_err_fp64_vs_cd_max=max(cds) if cds else None,
               grad_err_fp64_vs_cd_median=float(np.median(cds)) if cds else None,
               cd_within_1e6=int(sum(c <= 1e-6 for c in cds)),
               cd_worst=(max(((v, r["seed"], k) for r in rows
                              for k, v in r.get("grad_err_fp64_vs_central_differences", {}).items()), default=None)),
               rows=rows)
    print(f"[bench] diff_qp {json.dumps({k: v for k, v in out.items() if k != 'rows'})}", file=sys.stderr, flush=True)
    return out


def reference_cli(path, threads, *args, timeout=600):
    """One run of the reference CLI built from its sources (oracle/_ref/ganit) on an MPS file; its --json line,
    or None where that binary is not built."""
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", "ganit")
    if not os.access(exe, os.X_OK):
        return None
    out = subprocess.run([exe, path, "--device", "cpu", "--json", "--quiet", *args], capture_output=True, text=True,
                         env=dict(os.environ, OMP_NUM_THREADS=str(threads)), timeout=timeout)
    lines = [l for l in out.stdout.splitlines() if l.startswith("{")]
    return json.loads(lines[-1]) if out.returncode in (0, 2) and lines else None


def random_lp_c0(threads):
    """The seeded random LP (m=2000 rows, half equalities, n=5000 columns, 10 nonzeros per column, seed 0) in
    FP64 to 1e-4: the GPU solve (problem resident; and a fresh solver's first run, whose solve_s includes its
    upload) against the reference CLI on the same LP, on this host's cores and on one core."""
    import statistics
    import tempfile
    import ganit_v2_amd as g
    p = g.Problem.generate("random_lp", seed=0)
    d = p.dims()
    s = p.solver(precision="double", tol=1e-4)
    first = s.run()
    runs = [s.run() for _ in range(5)]
    r = runs[-1]
    gpu = dict(status=r["status"], iters=r["iters"], pobj=r["pobj"],
               solve_s=round(statistics.median(x["solve_s"] for x in runs), 6),
               solve_s_with_upload=round(first["solve_s"], 6), setup_s=round(first["setup_s"], 6))
    out = dict(workload=f"random LP m={d['m']} n={d['n']} nnz={d['nnz']}, seed 0, FP64, tol 1e-4", gpu=gpu)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "random_lp_s0.mps")
        p.write_mps(path)
        for t in sorted({threads, 1}):
            ref = reference_cli(path, t, "--precision", "double", "--tol", "1e-4")
            if ref is None:
                continue
            out[f"reference_cpu_{t}_cores"] = dict(status=ref["status"], iters=ref["iters"], pobj=ref["pobj"],
                                                   solve_s=ref["solve_s"], setup_s=ref["setup_s"])
            out[f"gpu_speedup_vs_{t}_cores"] = round(ref["solve_s"] / gpu["solve_s"], 2)
            out[f"obj_rel_err_vs_reference_{t}_cores"] = abs(gpu["pobj"] - ref["pobj"]) / (1 + abs(ref["pobj"]))
    out["cpu_model"] = cpu_model()
    print(f"[bench] random LP (C0) {json.dumps(out)}", file=sys.stderr, flush=True)
    return out


def large_blending_kernels(periods, reps):
    """The fused iteration kernels on a blending LP large enough that an iteration's streams cannot stay in the
    256 MB Infinity Cache (60 crudes x 16 products x 5 specs x `periods` periods): HIP-event times of each fused
    kernel (back-to-back launches) and of whole iterations, at L1 and L3, and their algorithmic GB/s."""
    import ganit_v2_amd as g
    prob = g.Problem.generate("blending", seed=0, crudes=60, products=16, specs=5, periods=periods)
    d = prob.dims()
    m, n, nnz = d["m"], d["n"], d["nnz"]
    out = dict(workload=f"blending LP 60 x 16 x 5 x {periods} periods, seed 0", m=m, n=n, nnz=nnz)
    nbv = streamed_bound_vectors(prob)
    for level, (prec, sm, sv) in {"L1": ("mixed", 4, 4), "L3": ("double", 8, 8)}.items():
        s = prob.solver(precision=prec, tol=1e-4)
        s.upload()
        kv = timed_kernels(s, 0 if level == "L1" else 2, reps, m, n, nnz, sm, sv, nbv)
        out[level] = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in kv.items()}
        out[level]["iteration_bytes"] = kv["primal_bytes"] + kv["dual_bytes"]
        del s
    print(f"[bench] large blending kernels {json.dumps(out)}", file=sys.stderr, flush=True)
    return out


def _setup_seconds(prob, where):
    """Seconds of a solver's set-up (scale_lp + estimate_norm) with the matrix work on the GPU or the host."""
    return prob.solver(precision="double", tol=1e-4, verbose=0, setup=where).setup_seconds()


def kernels_only(prob, staged, a, m, n, nnz):
    """Back-to-back launches of the fused kernels at L1 and L3 (HIP events on the solver's stream), then one
    torch copy of 1 GiB: a plain copy kernel of known size whose FETCH_SIZE / WRITE_SIZE calibrate the
    counters.  Prints one JSON line."""
    import torch
    s_m_low = 2 if a.values == "bf16" else 4
    out = {}
    nbv = streamed_bound_vectors(prob)
    for level, (sm, sv, sv_obj) in {"L1": (s_m_low, 4, staged("mixed", values=a.values)),
                                    "L3": (8, 8, staged("double"))}.items():
        out[level] = timed_kernels(sv_obj, 0 if level == "L1" else 2, a.reps, m, n, nnz, sm, sv, nbv)
        del sv_obj
    src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
    src.fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    out["copy_bytes_each_way"] = src.numel() * 4
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample-s", type=float, default=20.0)
    ap.add_argument("--reps", type=int, default=50, help="launches per kernel for the HIP-event timing")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "dynamic", "double"])
    ap.add_argument("--values", default="f32", choices=["f32", "bf16"])
    ap.add_argument("--stall-checks", type=int, default=60,
                    help="checks without 20%% progress before leaving L1 (the reference's default is 12)")
    ap.add_argument("--eval", default="fp64", choices=["fp64", "level"],
                    help="evaluate L1 restart candidates in FP64 (default) or at L1 precision")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the random LP against the reference CLI, the larger blending LP's kernels, the "
                         "portfolio-QP ladder and the batched-LP comparison reported beside the headline "
                         "(--c4 0 skips the differentiable QPs)")
    ap.add_argument("--c4", type=int, default=64,
                    help="differentiable QPs (forward + adjoint) reported in extras.diff_qp (0: skip)")
    ap.add_argument("--c4-concurrent", action="store_true",
                    help="with --c4: forwards as concurrent single solves instead of one batched solve per phase")
    ap.add_argument("--periods", type=int, default=WORKLOAD["periods"],
                    help="periods of the blending LP (365: the 2.9M-nonzero headline workload); other values are "
                         "for --kernels-only profiler runs")
    ap.add_argument("--large-periods", type=int, default=1460,
                    help="periods of the blending LP whose kernels are timed in extras.large_blending (0: skip)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only time the fused iteration kernels (the launches the roofline figure uses), plus one "
                         "plain 1 GiB device copy as the counter calibration; for profiler runs")
    ap.add_argument("--batched-only", type=int, default=0, metavar="N",
                    help="only run the batched solves of N random LPs (no sequential comparison); for profiler runs")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (requested --gpus {a.gpus}); a single LP does not shard "
              "profitably across GPUs and multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)

    import ganit_v2_amd as g

    if a.batched_only > 0:
        print(json.dumps(batched_lps(a.batched_only, sequential=False)), flush=True)
        return
    t0 = time.time()
    if a.periods != WORKLOAD["periods"] and not a.kernels_only:
        print("--periods other than 365 is for --kernels-only runs", file=sys.stderr)
        sys.exit(2)
    prob = g.Problem.generate("blending", seed=0, **dict(WORKLOAD, periods=a.periods))
    d = prob.dims()
    m, n, nnz = d["m"], d["n"], d["nnz"]
    print(f"[bench] blending LP m={m} n={n} nnz={nnz} generated in {time.time() - t0:.1f}s", file=sys.stderr,
          flush=True)

    def staged(precision, **kw):
        s = prob.solver(precision=precision, tol=1e-4, verbose=0, **kw)
        s.upload()
        return s

    if a.kernels_only:
        kernels_only(prob, staged, a, m, n, nnz)
        return

    mixed = staged(a.precision, values=a.values, stall_checks=a.stall_checks, eval_fp64=(a.eval == "fp64"))
    for _ in range(a.warmup):
        mixed.run()
    t = time.perf_counter()
    runs = [mixed.run(return_solution=(i == a.steps - 1)) for i in range(a.steps)]
    wall = time.perf_counter() - t
    last = runs[-1]
    print(f"[bench] {a.precision}: {[round(r['solve_s'], 4) for r in runs]} s, status {last['status']}, "
          f"iters {last['iters']} (L1 {last['fp32_iters']})", file=sys.stderr, flush=True)

    # the reference's solve_s includes the device upload (pdhg_impl.hpp:273-274, 305): a fresh solver's first run,
    # whose solve_s also holds the upload of every precision and the autotuner's trials
    cold = prob.solver(precision=a.precision, tol=1e-4, verbose=0, values=a.values, stall_checks=a.stall_checks,
                       eval_fp64=(a.eval == "fp64"))
    rc = cold.run()
    cold_run = dict(solve_s_with_upload=round(rc["solve_s"], 5), upload_s=round(rc["upload_s"], 5),
                    setup_s=round(rc["setup_s"], 5), status=rc["status"], iters=rc["iters"])
    del cold
    # set-up (scaling + norm estimate; outside solve_s, as in the reference) on the GPU, the default at this size,
    # and on the host's cores -- bitwise the same scaled problem either way
    for where in ("gpu", "host"):
        t_set = []
        for _ in range(3):
            t_set.append(_setup_seconds(prob, where))
        cold_run[f"setup_s_{where}"] = round(min(t_set), 5)
    print(f"[bench] {a.precision} with upload: {cold_run}", file=sys.stderr, flush=True)

    fp64 = staged("double")
    fp64.run()
    f_runs = [fp64.run(return_solution=(i == a.steps - 1)) for i in range(a.steps)]
    f_last = f_runs[-1]
    print(f"[bench] fp64: {[round(r['solve_s'], 4) for r in f_runs]} s, status {f_last['status']}, "
          f"iters {f_last['iters']}", file=sys.stderr, flush=True)
    # other ladder configurations, one timed solve each after a warm-up (reported, not the headline)
    ladder = []
    for cfg in (dict(precision="mixed"), dict(precision="dynamic"), dict(precision="mixed", kahan=True),
                dict(precision="mixed", values="bf16", stall_checks=a.stall_checks, eval_fp64=True),
                dict(precision="mixed", values="bf16f32", stall_checks=12, stall_checks_l2=a.stall_checks,
                     eval_fp64=True)):
        sv = staged(**cfg)
        sv.run()
        r = sv.run()
        ladder.append(dict(cfg=cfg, solve_s=round(r["solve_s"], 5), status=r["status"], iters=r["iters"],
                           l1_iters=r["fp32_iters"], l2_iters=r["l2_iters"], obj=r["pobj"], variants=sv.variants()))
        print(f"[bench] ladder {ladder[-1]}", file=sys.stderr, flush=True)
        del sv
    t_mixed = sum(r["solve_s"] for r in runs) / len(runs)
    t_fp64 = sum(r["solve_s"] for r in f_runs) / len(f_runs)
    obj_err = abs(last["pobj"] - f_last["pobj"]) / (1 + abs(f_last["pobj"]))

    # per-kernel device times (HIP events on the solver's stream) at L1 and L3
    s_m_low = 2 if a.values == "bf16" else 4
    nbv = streamed_bound_vectors(prob)
    kern = {}
    for level, (sm, sv, solver) in {"L1": (s_m_low, 4, mixed), "L3": (8, 8, fp64)}.items():
        if level == "L1" and a.precision == "double":
            continue
        kern[level] = timed_kernels(solver, 0 if level == "L1" else 2, a.reps, m, n, nnz, sm, sv, nbv)
        kern[level]["ref_sequence_gbs"] = (reference_sequence_bytes(m, n, nnz, sm, sv)
                                           / (kern[level]["iteration_us"] * 1e3))
    print(f"[bench] kernels {json.dumps(kern)}", file=sys.stderr, flush=True)
    # the roofline entry is the kernel with the most device time per iteration of the headline level, timed as
    # it runs in a solve: launched with none of its own data in L2 (the two kernels alternate, and each one's
    # share of the matrix evicts the other's); the other fused kernel's figures, both kernels back to back,
    # and each level's whole-iteration rate ride along under "kernels"
    lead = kern.get("L1", kern["L3"])
    dom = "primal" if lead["primal_us"] >= lead["dual_us"] else "dual"
    level_dom = "L1" if "L1" in kern else "L3"
    traffic = load_traffic() or {}
    per_kernel = {}
    for lv, kv in kern.items():
        for which in ("primal", "dual"):
            per_kernel[f"k_iter_{which} ({lv})"] = dict(
                achieved=round(kv[f"{which}_gbs"], 1), frac=round(kv[f"{which}_gbs"] / PEAK_HBM_GBS, 4),
                us=round(kv[f"{which}_us"], 3), algorithmic_bytes=kv[f"{which}_bytes"],
                traffic=traffic.get(f"k_iter_{which}_{lv}"),
                back_to_back=dict(us=round(kv[f"{which}_b2b_us"], 3),
                                  achieved=round(kv[f"{which}_b2b_gbs"], 1),
                                  frac=round(kv[f"{which}_b2b_gbs"] / PEAK_HBM_GBS, 4)))
        per_kernel[f"stream read 64 MiB ({lv} run)"] = dict(
            achieved=round(kv["sweep_gbs"], 1), frac=round(kv["sweep_gbs"] / PEAK_HBM_GBS, 4),
            us=round(kv["sweep_us"], 3), algorithmic_bytes=64 << 20)
        per_kernel[f"iteration ({lv})"] = dict(achieved=round(kv["iteration_gbs"], 1),
                                               frac=round(kv["iteration_gbs"] / PEAK_HBM_GBS, 4),
                                               us=round(kv["iteration_us"], 3),
                                               algorithmic_bytes=kv["primal_bytes"] + kv["dual_bytes"])
    roof = {
        "bound": "hbm",
        "kernel": f"k_iter_{dom} ({level_dom})",
        "achieved": round(lead[f"{dom}_gbs"], 1),
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": round(lead[f"{dom}_gbs"] / PEAK_HBM_GBS, 4),
        "traffic": traffic.get(f"k_iter_{dom}_{level_dom}"),
        "kernels": per_kernel,
    }

    extras = {}
    if not a.no_extras:
        extras["random_lp_c0"] = random_lp_c0(os.cpu_count() or 1)
        if a.large_periods > 0:
            extras["large_blending"] = large_blending_kernels(a.large_periods, a.reps)
        extras["portfolio_qp"] = portfolio_ladder()
        extras["batched_lp"] = batched_lps()
    if a.c4 > 0:
        extras["diff_qp"] = diff_qps(a.c4, batched=not a.c4_concurrent)

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(prob, last, a.cpu_sample_s)
        if cpu and "pobj" in cpu:
            cpu["gpu_obj_rel_err_vs_reference"] = abs(last["pobj"] - cpu["pobj"]) / (1 + abs(cpu["pobj"]))
        print(f"[bench] cpu baseline {json.dumps(cpu)}", file=sys.stderr, flush=True)

    line = {
        "metric": "time-to-1e-4 rel KKT, mixed vs FP64 PDHG on 2.9M-nnz LP; SpMV HBM GB/s",
        "value": round(t_mixed, 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if a.precision != "double" else "f64",
        "data": "synthetic (seeded blending LP generator)",
        "config": {"workload": "blending LP 60 crudes x 16 products x 5 specs x 365 periods, seed 0",
                   "m": m, "n": n, "nnz": nnz, "precision": a.precision, "values": a.values, "tol": 1e-4,
                   "stall_checks": a.stall_checks, "eval": a.eval, "parallelism": "single GPU"},
        "status": last["status"],
        "iters": last["iters"], "l1_iters": last["fp32_iters"], "restarts": last["restarts"],
        "fp64_s": round(t_fp64, 5), "fp64_iters": f_last["iters"], "fp64_status": f_last["status"],
        "speedup_vs_fp64": round(t_fp64 / t_mixed, 3),
        "variants": {"mixed": mixed.variants(), "fp64": fp64.variants()},
        "with_upload": cold_run,
        "obj": last["pobj"], "obj_fp64": f_last["pobj"], "obj_rel_err_vs_fp64": obj_err,
        "certified_rel": [last["rel_p"], last["rel_d"], last["rel_gap"]],
        "ladder": ladder,
        "kernels": kern,
        "roofline": roof,
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
