"""Writes tests/golden/oracle_runs.jsonl: the oracle's (oracle/pdhg_oracle.cpp) runs of the headline blending LP
(60 crudes x 16 products x 5 specs x 365 periods, seed 0; the same LP as reference_runs.jsonl's blending_s0) in the
configurations that only the oracle has besides this package: the ladders with FP64 candidate evaluation
(eval_fp64), which the reference CLI does not offer.  These are fixtures of the checker, computed in this
container, so that the GPU test need not run a minute of CPU PDHG on the GPU host.

    python tests/golden/make_oracle_runs.py
"""
from __future__ import annotations

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle  # noqa: E402
import ganit_v2_amd as g  # noqa: E402

BLENDING = dict(crudes=60, products=16, specs=5, periods=365)
CONFIGS = {
    "mixed_fp64_checks_stall60": dict(precision="mixed", eval_fp64=True, stall_checks=60),
    "bf16_fp64_checks_stall60": dict(precision="mixed", eval_fp64=True, stall_checks=60, l1_matrix=1),
}
THREADS = 8


def main():
    p = g.Problem.generate("blending", seed=0, **BLENDING)
    lp = pyoracle.OracleLP.from_export(p.export())
    pyoracle.set_threads(THREADS)
    d = p.dims()
    dst = os.path.join(HERE, "oracle_runs.jsonl")
    with open(dst, "w") as out:
        for name, cfg in CONFIGS.items():
            r = lp.solve(tol=1e-4, **cfg)
            row = {k: r[k] for k in ("status", "pobj", "dobj", "rel_p", "rel_d", "rel_gap", "iters", "fp32_iters",
                                     "l2_iters", "restarts")}
            row.update(instance="blending_s0", config=name, options=dict(cfg, tol=1e-4), threads=THREADS,
                       m=d["m"], n=d["n"], nnz=d["nnz"])
            out.write(json.dumps(row) + "\n")
            print(name, row, flush=True)
    print("wrote", dst)


if __name__ == "__main__":
    main()
