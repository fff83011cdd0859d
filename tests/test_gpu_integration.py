"""The drop-in integration: the reference's own CLI and driver (src/main.cpp + src/pdhg_impl.hpp, compiled
from the reference's sources into oracle/_ref/ganit_hip) with backend_cuda.cu replaced by HipBackend over
libganit_hip.so.  `ganit model.mps --device gpu` must run unchanged and agree with the reference CPU run."""
import json
import os
import subprocess

import pytest

from conftest import ROOT, golden, instance_path

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "oracle", "_ref", "ganit_hip")


def _run(name, *args):
    p = subprocess.run([EXE, instance_path(name), "--device", "gpu", "--json", "--quiet", *args],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    lines = [l for l in p.stdout.splitlines() if l.startswith("{")]
    assert lines, p.stderr
    return p.returncode, json.loads(lines[-1])


@pytest.mark.skipif(not os.path.exists(EXE), reason="oracle/_ref/ganit_hip is built only where the reference "
                                                    "sources exist")
@pytest.mark.parametrize("name", ["afiro", "adlittle", "bgetam", "25fv47", "port_tiny", "qblend_small"])
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_reference_driver_on_hip_backend(name, precision):
    ref = golden(name, ["--device", "cpu", "--precision", precision, "--time", "120"], 1)
    rc, d = _run(name, "--precision", precision, "--time", "120")
    assert d["device"].startswith("gpu:")
    assert d["status"] == ref["status"]
    assert rc == (0 if ref["status"] == "OPTIMAL" else 2)
    if ref["status"] == "OPTIMAL":
        assert max(d["rel_p"], d["rel_d"], d["rel_gap"]) <= 1e-4 * (1 + 1e-7)
        assert abs(d["pobj"] - ref["pobj"]) <= 5e-4 * (1 + abs(ref["pobj"]))
