import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
INSTANCES = os.path.join(GOLDEN, "instances")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X); run with -m gpu")


def instance_path(name: str) -> str:
    return os.path.join(INSTANCES, name + ".mps.gz")


def golden_runs():
    """The reference CLI's recorded results (tests/golden/reference_runs.jsonl)."""
    path = os.path.join(GOLDEN, "reference_runs.jsonl")
    rows = []
    with open(path) as f:
        for line in f:
            d = json.loads(line)
            if "meta" not in d and "error" not in d:
                rows.append(d)
    return rows


def golden(instance: str, args: list, threads: int):
    for r in golden_runs():
        if r["instance"] == instance and r["args"] == args and r["threads"] == threads:
            return r
    raise KeyError((instance, args, threads))


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    if not os.path.exists(pyoracle.LIB_PATH):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "port"], check=True)
    return pyoracle


@pytest.fixture(scope="session")
def hipctx():
    from ganit_v2_amd import hip
    ctx = hip.Context(0)
    yield ctx
    ctx.close()
