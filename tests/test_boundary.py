"""The C-ABI boundary, without a GPU: both libraries load and export every declared symbol, errors come
back as status codes, the product never reaches the oracle, and the CLI keeps the reference's flags and
exit codes."""
import ctypes
import os
import re
import subprocess

import pytest

import ganit_v2_amd as g
from ganit_v2_amd import hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ganit_\w+)\s*\(", text)))


def test_hip_library_exports_every_declared_symbol():
    L = hip.lib()
    names = _declared("ganit_hip.h")
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing


def test_host_library_exports_every_declared_symbol():
    L = g.host_lib()
    names = _declared("ganit_host.h")
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing and len(names) >= 10


def test_library_code_objects_target_gfx950():
    data = open(os.path.join(g.LIB_DIR, "libganit_hip.so"), "rb").read()
    assert b"gfx950" in data


def test_hip_errors_are_status_codes():
    h = ctypes.c_void_p()
    rc = hip.lib().ganit_hip_ctx_create(9999, ctypes.byref(h))
    assert rc != 0
    assert len(hip.lib().ganit_hip_last_error()) > 0
    assert hip.lib().ganit_hip_csr_set_lanes(None, 8) == 1  # null handle -> GANIT_EINVAL


def test_missing_native_library_fails_loudly(monkeypatch, tmp_path):
    import ganit_v2_amd
    monkeypatch.setattr(ganit_v2_amd, "LIB_DIR", str(tmp_path))
    with pytest.raises(ganit_v2_amd.NativeLibraryMissing):
        ganit_v2_amd._load("libganit_hip.so")


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "ganit_v2_amd")
    offenders = []
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hpp", ".hip", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                if re.search(r"(import\s+oracle|from\s+oracle|pyoracle|liboracle|pdhg_oracle)", text):
                    offenders.append(f)
    assert not offenders
    # and the built libraries do not link it
    for so in ("libganit_hip.so", "libganit.so"):
        out = subprocess.run(["ldd", os.path.join(g.LIB_DIR, so)], stdout=subprocess.PIPE, text=True).stdout
        assert "oracle" not in out


def _cli(*args):
    return subprocess.run([os.path.join(g.BIN_DIR, "ganit"), *args], stdout=subprocess.PIPE,
                          stderr=subprocess.PIPE, text=True, timeout=60)


def test_cli_flags_and_exit_codes(tmp_path):
    assert _cli("--help").returncode == 0
    assert _cli().returncode == 1
    afiro = os.path.join(ROOT, "tests", "golden", "instances", "afiro.mps.gz")
    r = _cli(afiro, "--bogus")
    assert r.returncode == 1 and "unknown option" in r.stderr
    r = _cli(afiro, "--precision", "quad")
    assert r.returncode == 1
    r = _cli(afiro, "--device", "cpu")
    assert r.returncode == 1 and "GPU" in r.stderr
    r = _cli(str(tmp_path / "missing.mps"), "--quiet")
    assert r.returncode == 1 and "cannot open" in r.stderr


def test_cli_new_flags_parse(tmp_path):
    afiro = os.path.join(ROOT, "tests", "golden", "instances", "afiro.mps.gz")
    r = _cli(afiro, "--setup", "sideways")
    assert r.returncode == 1 and "--setup" in r.stderr
    # accepted flags reach the device check (this build runs on the GPU only)
    for flags in (["--polish"], ["--scale-q"], ["--setup", "gpu"], ["--setup", "host"], ["--setup", "auto"]):
        r = _cli(afiro, *flags, "--device", "cpu")
        assert r.returncode == 1 and "GPU" in r.stderr, flags
    assert "--polish" in _cli("--help").stdout and "--setup" in _cli("--help").stdout
