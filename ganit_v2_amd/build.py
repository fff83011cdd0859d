"""Build recipe for the in-tree native artefacts (no CMake; plain hipcc / g++ command lines).

    python -m ganit_v2_amd.build          # or: python ganit_v2_amd/build.py

Produces
    ganit_v2_amd/lib/libganit_hip.so   HIP kernels + C ABI (include/ganit_hip.h), gfx950 code objects
    ganit_v2_amd/lib/libganit.so       host library: MPS I/O, scaling, generators, PDHG driver
                                       (include/ganit_host.h); links libganit_hip.so
    ganit_v2_amd/bin/ganit             the command-line solver
Object files go to _build/ (git-ignored).  Every call recompiles.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")
OBJ = os.path.join(ROOT, "_build")

HIP_SOURCES = ["runtime.hip", "pdhg_device.hip", "batch.hip", "adjoint.hip", "setup.hip", "polish.hip"]
HOST_SOURCES = ["lp.cpp", "scaling.cpp", "generators.cpp", "driver.cpp", "batch.cpp", "capi.cpp"]

ARCH = os.environ.get("GANIT_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions are disabled for the whole library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# Host code: no FMA contraction, so the scaling reproduces the reference's roundings exactly.
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out:
        print(out, file=sys.stderr)


def build(verbose: bool = True) -> dict:
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        o = os.path.join(OBJ, s.replace(".hip", ".o"))
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
    for s in HOST_SOURCES + ["main.cpp"]:
        o = os.path.join(OBJ, "host_" + s.replace(".cpp", ".o"))
        jobs.append([CXX, *CXX_FLAGS, "-c", os.path.join(HOST, s), "-o", o])
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        list(ex.map(_run, jobs))
    hip_so = os.path.join(LIB, "libganit_hip.so")
    host_so = os.path.join(LIB, "libganit.so")
    exe = os.path.join(BIN, "ganit")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", hip_so,
          *[os.path.join(OBJ, s.replace(".hip", ".o")) for s in HIP_SOURCES]])
    _run([CXX, "-shared", "-fopenmp", "-o", host_so,
          *[os.path.join(OBJ, "host_" + s.replace(".cpp", ".o")) for s in HOST_SOURCES],
          f"-L{LIB}", "-lganit_hip", "-Wl,-rpath,$ORIGIN"])
    _run([CXX, "-fopenmp", "-o", exe, os.path.join(OBJ, "host_main.o"),
          f"-L{LIB}", "-lganit", "-lganit_hip", "-Wl,-rpath,$ORIGIN/../lib"])
    if verbose:
        print(f"built {hip_so}\nbuilt {host_so}\nbuilt {exe}")
    return {"hip": hip_so, "host": host_so, "exe": exe}


if __name__ == "__main__":
    build()
