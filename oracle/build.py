"""TEST INFRASTRUCTURE — builds the CPU oracle (_build/liboracle.so) against the product's host library.

The oracle is a restatement, in this repository, of the reference's FP64 OpenMP CPU backend; the reference's own
propagation path is not built (it needs FFTW, nlohmann_json and LAPACK, which this image lacks)."""

from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HOST_LIB_DIR = os.path.join(ROOT, "cpp-spo-dvr_amd", "lib")
FLAGS = ["-std=c++20", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fcx-fortran-rules", "-fopenmp", "-fPIC",
         "-Wall", "-Wextra"]


def build() -> None:
    out_dir = os.path.join(HERE, "_build")
    os.makedirs(out_dir, exist_ok=True)
    sources = sorted(os.path.join(HERE, f) for f in os.listdir(HERE) if f.endswith(".cpp"))
    lib = os.path.join(out_dir, "liboracle.so")
    cmd = ["g++", *FLAGS, "-shared", "-o", lib + ".tmp", *sources, f"-L{HOST_LIB_DIR}", "-lspo_host",
           "-Wl,-rpath,$ORIGIN/../../cpp-spo-dvr_amd/lib"]
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{proc.stdout}")
    os.replace(lib + ".tmp", lib)


if __name__ == "__main__":
    build()
