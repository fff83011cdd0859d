"""Writes tests/golden/reference_runs.jsonl: the reference CLI's own --json results (plus the thread count and
arguments) on the bundled instances and on seeded random LPs.

The reference binary is oracle/_ref/ganit, compiled from the reference's C++ sources by `make -C oracle ref`
in a container where those sources exist.  Seeded random LPs are produced by this package's generator and
written as MPS files first (their sha256 is recorded so a fixture can be matched to its input).

    python tests/golden/make_reference_runs.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle  # noqa: E402
import ganit_v2_amd as g  # noqa: E402

INST = os.path.join(HERE, "instances")
LP_NAMES = ["afiro", "adlittle", "bgetam", "etamacro", "stair", "shell", "25fv47", "greenbea"]
QP_NAMES = ["port_tiny", "port_small", "qblend_small"]


def sha(path: str) -> str:
    return hashlib.sha256(open(path, "rb").read()).hexdigest()[:16]


def record(out, instance, path, args, threads, timeout=900):
    try:
        r = pyoracle.run_reference(path, *args, threads=threads, timeout=timeout)
    except Exception as e:  # a time-out or crash is recorded, not fatal
        r = {"error": str(e)[:200]}
    r.update(instance=instance, args=list(args), threads=threads)
    out.write(json.dumps(r) + "\n")
    out.flush()
    print(instance, args, threads, r.get("status"), r.get("pobj"), r.get("iters"), flush=True)


def main():
    dst = os.path.join(HERE, "reference_runs.jsonl")
    with open(dst, "w") as out:
        for name in LP_NAMES + QP_NAMES:
            path = os.path.join(INST, name + ".mps.gz")
            for prec in ("double", "mixed", "dynamic"):
                record(out, name, path, ["--device", "cpu", "--precision", prec, "--time", "120"], 1)
            record(out, name, path, ["--device", "cpu", "--time", "120"], 8)
            if name in ("afiro", "adlittle", "port_tiny", "bgetam", "stair"):
                record(out, name, path, ["--device", "cpu", "--tol", "1e-8", "--time", "120"], 1)
            if name in LP_NAMES:
                record(out, name, path, ["--method", "simplex", "--time", "300"], 1)
        with tempfile.TemporaryDirectory() as td:
            for seed in range(5):
                p = g.Problem.generate("random_lp", seed=seed)
                path = os.path.join(td, f"randlp_s{seed}.mps")
                p.write_mps(path)
                inst = f"random_lp_seed{seed}"
                meta = {"instance": inst, "generator": "random_lp", "seed": seed, "mps_sha256_16": sha(path)}
                out.write(json.dumps({"meta": meta}) + "\n")
                record(out, inst, path, ["--device", "cpu"], 1)
                record(out, inst, path, ["--device", "cpu"], 8)
                record(out, inst, path, ["--device", "cpu", "--precision", "mixed"], 1)
                record(out, inst, path, ["--device", "cpu", "--tol", "1e-8", "--time", "300"], 8)
                record(out, inst, path, ["--method", "simplex", "--time", "600"], 1)
        blending_runs(out, os.path.join(td_root(), "blending"))
        portfolio_runs(out, os.path.join(td_root(), "portfolio"))
    print("wrote", dst)


BLENDING = dict(crudes=60, products=16, specs=5, periods=365)


def td_root() -> str:
    return tempfile.gettempdir()


def blending_runs(out, workdir):
    """The 2.9M-nonzero blending LP (60 crudes x 16 products x 5 specs x 365 periods, seed 0), written as MPS
    (107 MB) in `workdir`: the reference CLI at tol 1e-4 in every precision at 1 and 8 threads, and in FP64 at
    tol 1e-8 (the accurate objective that 1e-4 solves are measured against)."""
    os.makedirs(workdir, exist_ok=True)
    p = g.Problem.generate("blending", seed=0, **BLENDING)
    path = os.path.join(workdir, "blending_s0.mps")
    p.write_mps(path)
    inst = "blending_s0"
    d = p.dims()
    meta = {"instance": inst, "generator": "blending", "seed": 0, **BLENDING, "m": d["m"], "n": d["n"],
            "nnz": d["nnz"], "mps_sha256_16": sha(path)}
    out.write(json.dumps({"meta": meta}) + "\n")
    for threads in (1, 8):
        for prec in ("double", "mixed", "dynamic"):
            record(out, inst, path, ["--device", "cpu", "--precision", prec, "--time", "600"], threads, timeout=900)
    record(out, inst, path, ["--device", "cpu", "--tol", "1e-8", "--time", "1200"], 8, timeout=1500)
    os.remove(path)


PORTFOLIO = dict(assets=100_000, factors=50)


def portfolio_runs(out, workdir):
    """The 5.1M-nonzero factor-model portfolio QP (100,000 assets x 50 factors, seed 0), written as QPS: the
    reference CLI at tol 1e-6 in every precision at 8 threads."""
    os.makedirs(workdir, exist_ok=True)
    p = g.Problem.generate("portfolio", seed=0, **PORTFOLIO)
    path = os.path.join(workdir, "portfolio_s0.mps")
    p.write_mps(path)
    inst = "portfolio_s0"
    d = p.dims()
    meta = {"instance": inst, "generator": "portfolio", "seed": 0, **PORTFOLIO, "m": d["m"], "n": d["n"],
            "nnz": d["nnz"], "qnnz": d["qnnz"], "mps_sha256_16": sha(path)}
    out.write(json.dumps({"meta": meta}) + "\n")
    for prec in ("double", "mixed", "dynamic"):
        record(out, inst, path, ["--device", "cpu", "--precision", prec, "--tol", "1e-6", "--time", "600"], 8,
               timeout=900)
    os.remove(path)


if __name__ == "__main__":
    if sys.argv[1:] == ["--append-blending"]:
        # add only the blending LP's runs to an existing file (the other runs take much longer to redo)
        with open(os.path.join(HERE, "reference_runs.jsonl"), "a") as f:
            blending_runs(f, os.path.join(td_root(), "blending"))
    elif sys.argv[1:] == ["--append-portfolio"]:
        with open(os.path.join(HERE, "reference_runs.jsonl"), "a") as f:
            portfolio_runs(f, os.path.join(td_root(), "portfolio"))
    else:
        main()
