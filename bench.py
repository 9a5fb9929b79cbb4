#!/usr/bin/env python3
"""Throughput of the 6D split-operator propagation on one MI355X, in DVR grid-point-steps per second.

A grid point is one point of the 6D DVR product grid NZ * NR * NX * NY * n_theta * n_phi; a step is one time step
(one T.V pair of the reference's schedule, src/propagation/split_operator_schedule.cpp:7-13). The timed work is the
real time loop: analysis steps of steps_per_analysis time steps, each followed by the analysis surface update, the
on-device Balint-Kurti projection and FP64 S-matrix accumulation, and every dtAve the block densities. The wave
packet is synthetic (seeded pseudo-random values on the scattering grid, the Gaussian packet on the specular grid)
and resident in HBM before the timed region starts.

Printed: one JSON line with the throughput, the roofline of the dominant kernel (algorithmic bytes per launch over
its mean HIP-event duration in the timed region) and the CPU baseline (the FP64 OpenMP oracle on the same grid,
timed on this host over a bounded number of time steps).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload dz|hardwall] [--precision fp64|mixed]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cpp-spo-dvr_amd"))
sys.path.insert(0, ROOT)

METRIC = "6D SPO grid-point-steps/s, mixed vs FP64 on 1 GPU at |ΔP_react|≤1e-4"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "dz": ("DZ_6D", "H2/Cu(111) Dai-Light LEPS 6D: NZ 96 x NR 48 x 16 x 16, JMax 24 / MJMax 23 "
                    "(13 theta x 48 phi), C6v-reduced potential, specular grid 512"),
    "hardwall": ("HARDWALL_6D", "HardWallPES 6D: NZ 128 x NR 32 x 8 x 8, JMax 10 (6 theta x 22 phi)"),
}


def dvr_points(info) -> int:
    return info.nz * info.nr * info.nx * info.ny * info.n_theta * info.n_phi


def algorithmic_bytes(kernel: str, info, complex_bytes: int) -> float | None:
    """HBM bytes one launch of a hot kernel must move: one read and one write of its share of psi, plus the
    exp(iA) table for the interaction pass."""
    psi = info.n_rot * info.nx * info.ny * info.nr * info.nz * complex_bytes
    f_int = info.nz_potential / info.nz
    if kernel.startswith("zr_plane") or kernel.startswith("xy_tile"):
        return 2.0 * psi
    if kernel.startswith("interaction"):
        table = info.n_theta * info.n_phi * info.n_representatives * info.nz_potential * info.nr * complex_bytes
        return 2.0 * psi * f_int + table
    return None


def synthetic_wave(run, seed: int = 2024):
    import numpy as np
    rng = np.random.default_rng(seed)
    n = run.scattering_size
    s = np.empty(n, dtype=np.complex128)
    s.real = rng.random(n) - 0.5
    s.imag = rng.random(n) - 0.5
    s *= 1e-3
    return s, run.specular()


def cpu_baseline(path: str, sample_steps: int, wave) -> dict:
    """The CPU oracle (FP64 OpenMP restatement of the reference's CPU backend) on the same grid and wave packet:
    sample_steps full time steps (specular grid, T_Zr, T_xy, V), timed with a monotonic clock after one warm-up."""
    from oracle import oracle
    cpu = oracle.run(path, {"Precision": "fp64"})
    cpu.set_wave(*wave)
    cpu.kinetic_and_potential(False, False)
    t0 = time.perf_counter()
    for _ in range(sample_steps):
        cpu.kinetic_and_potential(False, False)
    dt = (time.perf_counter() - t0) / sample_steps
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    value = dvr_points(cpu.info) / dt
    cpu.close()
    return {"value": value, "unit": "grid-point-steps/s", "cores": cores, "kind": "port",
            "sample": f"{sample_steps} full time step(s) of the same grid after one warm-up step "
                      f"({dt:.2f} s per step, FP64 OpenMP oracle, {cores} threads)"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50, help="time steps timed (rounded up to whole analysis steps)")
    ap.add_argument("--warmup", type=int, default=10, help="untimed time steps before the timed region")
    ap.add_argument("--workload", default="dz", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--cpu-steps", type=int, default=1, help="time steps of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the propagation of one run is not sharded "
              "across GPUs in this build", file=sys.stderr)
        return 2

    import spodvr_amd
    from tests import cases

    name, description = WORKLOADS[args.workload]
    params = dict(getattr(cases, name))
    params["Precision"] = args.precision
    path = cases.write_case(f"bench_{args.workload}_{args.precision}", params)
    run = spodvr_amd.Run(path, device="hip")
    info = run.info
    spa = info.steps_per_analysis
    timed_analysis = max(1, -(-args.steps // spa))
    warm_analysis = max(1, -(-args.warmup // spa)) if args.warmup > 0 else 0
    steps = timed_analysis * spa
    wave = synthetic_wave(run)
    run.set_wave(*wave)
    if warm_analysis:
        run.steps(warm_analysis)
    run.kernel_timing(True)
    t0 = time.perf_counter()
    run.steps(timed_analysis)
    elapsed = time.perf_counter() - t0
    kernels = run.kernel_times()
    run.kernel_timing(False)

    points = dvr_points(info)
    value = points * steps / elapsed
    complex_bytes = 16 if args.precision == "fp64" else 8
    dominant = max(kernels.items(), key=lambda kv: kv[1][1]) if kernels else None
    roofline = None
    if dominant:
        kname, (launches, total_ms) = dominant
        avg_s = total_ms / launches / 1e3
        alg = algorithmic_bytes(kname, info, complex_bytes)
        if alg is not None:
            achieved = alg / avg_s / 1e9
            roofline = {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                        "algorithmic_bytes_per_launch": alg, "mean_launch_ms": round(total_ms / launches, 4),
                        "launches": launches}
    kernel_summary = {k: {"launches": n, "ms_per_time_step": round(ms / steps, 4)} for k, (n, ms) in kernels.items()}

    cpu = None
    if not args.no_cpu_baseline:
        run.close()
        cpu = cpu_baseline(path, args.cpu_steps, wave)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm_analysis * spa,
        "ms_per_step": elapsed / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "c64 storage / f64 phases and analysis",
        "data": "synthetic",
        "config": {"workload": description, "precision": args.precision, "dvr_points": points,
                   "psi_values": info.n_rot * info.nx * info.ny * info.nr * info.nz,
                   "steps_per_analysis": spa, "n_rot": info.n_rot, "nz_potential": info.nz_potential,
                   "n_representatives": info.n_representatives},
        "roofline": roofline,
        "kernels": kernel_summary,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
