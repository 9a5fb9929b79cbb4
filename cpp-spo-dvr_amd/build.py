"""Builds the in-tree shared libraries of the backend.

  lib/libspo_hip.so   csrc/*.hip, hipcc for gfx950 (the device backend, C ABI include/spo_hip.h)
  lib/libspo_host.so  host/*.cpp, g++ (set-up, time loop, analysis; C ABI include/spo_host.h), linked to the above

Usage: python build.py [--jobs N]. Every call compiles (no staleness check), as the build check expects.
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-result"]
CXX_FLAGS = ["-std=c++20", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fcx-fortran-rules", "-fopenmp",
             "-fPIC", "-Wall", "-Wextra", "-Wno-class-memaccess"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in proc.stdout.splitlines() if "packed-fp32-ops" not in line)
    if proc.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out)


def build(jobs: int = 8) -> None:
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    csrc = os.path.join(HERE, "csrc")
    host = os.path.join(HERE, "host")
    hip_sources = sorted(os.path.join(csrc, f) for f in os.listdir(csrc) if f.endswith(".hip"))
    host_sources = sorted(os.path.join(host, f) for f in os.listdir(host) if f.endswith(".cpp"))

    def hip_obj(src):
        obj = os.path.join(OBJ, os.path.basename(src) + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
        return obj

    def cxx_obj(src):
        obj = os.path.join(OBJ, os.path.basename(src) + ".o")
        _run(["g++", *CXX_FLAGS, "-c", src, "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as pool:
        hip_objs = list(pool.map(hip_obj, hip_sources))
        host_objs = list(pool.map(cxx_obj, host_sources))
    hip_lib = os.path.join(LIB, "libspo_hip.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", hip_lib + ".tmp", *hip_objs])
    os.replace(hip_lib + ".tmp", hip_lib)
    host_lib = os.path.join(LIB, "libspo_host.so")
    _run(["g++", "-shared", "-fopenmp", "-o", host_lib + ".tmp", *host_objs, f"-L{LIB}", "-lspo_hip",
          "-Wl,-rpath,$ORIGIN", "-Wl,--no-as-needed"])
    os.replace(host_lib + ".tmp", host_lib)


if __name__ == "__main__":
    jobs = int(sys.argv[sys.argv.index("--jobs") + 1]) if "--jobs" in sys.argv else 8
    build(jobs)
