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
    print("wrote", dst)


if __name__ == "__main__":
    main()
