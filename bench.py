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

By default both storage precisions run on the same workload: `value` is the mixed-precision throughput (complex64
wave, FP64 phases, analysis and sums), and the FP64 run is reported beside it with the ratio of the two.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload dz|hardwall] [--precision both|fp64|mixed]
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
    "batch12": ("BATCH_12", "12 DZ-LEPS runs batched in complex64: initial states (v, j, mJ) on the 96 x 48 x 16 x 16 "
                            "grid (JMax 24 / 23) and four at JMax 12 / 18 with NR 32; one hipGraph launch per "
                            "analysis step for the batch"),
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


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.txt")


def measured_traffic(kernel: str, workload: str, precision: str) -> float | None:
    """HBM bytes per launch of a kernel from the rocprofv3 FETCH_SIZE and WRITE_SIZE passes summarised in
    profiles/hbm_traffic.txt (each counter scaled by what it reads for a device copy of known size); None when the
    kernel, workload and precision have not been profiled."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        parts = line.split("#", 1)[0].split()
        if len(parts) == 5 and parts[:3] == [kernel, workload, precision]:
            return float(parts[3]) + float(parts[4])
    return None


def copy_calibration(n_bytes: int, reps: int = 10) -> None:
    """Device-to-device copies of n_bytes through torch (the reference copy kernel whose FETCH_SIZE and
    WRITE_SIZE calibrate the counters: each dispatch reads and writes exactly n_bytes)."""
    import torch
    a = torch.ones(n_bytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"copy_bytes": n_bytes, "reps": reps, "ms": dt * 1e3, "GB/s": 2 * n_bytes / dt / 1e9}), flush=True)


def synthetic_wave(run, seed: int = 2024):
    import numpy as np
    rng = np.random.default_rng(seed)
    n = run.scattering_size
    s = np.empty(n, dtype=np.complex128)
    s.real = rng.random(n) - 0.5
    s.imag = rng.random(n) - 0.5
    s *= 1e-3
    return s, run.specular()


def host_description() -> dict:
    """The CPU the baseline runs on: model, logical CPUs available to this process, NUMA nodes."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    node_dir = "/sys/devices/system/node"
    numa = len([d for d in os.listdir(node_dir) if d.startswith("node") and d[4:].isdigit()]) \
        if os.path.isdir(node_dir) else 1
    available = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"cpu_model": model, "logical_cpus": available, "numa_nodes": numa}


def cpu_baseline(path: str, sample_steps: int, wave) -> dict:
    """The CPU oracle (FP64 OpenMP restatement of the reference's CPU backend) on the same grid and wave packet:
    whole analysis steps (the reference's schedule, src/propagation/split_operator_schedule.cpp:7-13, followed by
    the analysis-line projection), at least sample_steps time steps, timed with a monotonic clock after one warm-up
    analysis step, as the reference's operator timing does (one warm-up call, then the mean of N;
    src/app/operator_timing.cpp:26-32)."""
    from oracle import oracle
    cpu = oracle.run(path, {"Precision": "fp64"})
    spa = cpu.info.steps_per_analysis
    n_analysis = max(1, -(-sample_steps // spa))
    cpu.set_wave(*wave)
    cpu.propagate()
    cpu.analyse(cpu.info.time_step * spa)
    t0 = time.perf_counter()
    for k in range(n_analysis):
        cpu.propagate()
        cpu.analyse(cpu.info.time_step * spa * (k + 2))
    dt = (time.perf_counter() - t0) / (n_analysis * spa)
    threads = int(os.environ.get("OMP_NUM_THREADS", 0)) or host_description()["logical_cpus"]
    value = dvr_points(cpu.info) / dt
    cpu.close()
    host = host_description()
    return {"value": value, "unit": "grid-point-steps/s", "cores": threads, "kind": "port",
            "sample": f"{n_analysis} analysis step(s) of {spa} time steps (schedule, BK projection and S-matrix sums) "
                      f"on the same grid and wave packet after one warm-up analysis step ({dt:.2f} s per time step, "
                      f"FP64 OpenMP oracle with its own Stockham FFT, {threads} threads)",
            "host": dict(host, omp_threads=threads,
                         omp_schedule="static for the kinetic passes and the analysis, dynamic,1 for the potential "
                                      "step (the reference's one-NUMA-node choice, src/core/loop_schedule.hpp:6-18)")}


def measure(path: str, workload: str, precision: str, steps: int, warmup: int, merge_half_steps: bool = False) -> dict:
    """One precision on the GPU: warm-up analysis steps, then the timed analysis steps with per-kernel HIP events.
    Every seam call returns with the device idle (the analysis step ends in a synchronous copy of the BK
    coefficients), so the wall clock brackets exactly the timed work. merge_half_steps: the schedule with the
    trailing and leading T(dt/2) of consecutive analysis steps merged into one T(dt) (SPO_HIP_MERGE_HALF_STEPS=1);
    otherwise the plain schedule of the reference."""
    import spodvr_amd
    os.environ.setdefault("SPO_HIP_TUNE_LOG", "1")  # the launch tuner's choices go to stderr
    os.environ["SPO_HIP_MERGE_HALF_STEPS"] = "1" if merge_half_steps else "0"
    try:
        run = spodvr_amd.Run(path, {"Precision": precision}, device="hip")
    finally:
        os.environ.pop("SPO_HIP_MERGE_HALF_STEPS", None)
    info = run.info
    spa = info.steps_per_analysis
    timed_analysis = max(1, -(-steps // spa))
    warm_analysis = max(1, -(-warmup // spa)) if warmup > 0 else 0
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
    run.close()
    n_steps = timed_analysis * spa
    complex_bytes = 16 if precision == "fp64" else 8
    # the dominant of the three hot passes (the specular grid's kernels run on a second stream, overlapped with
    # T_Zr and T_xy: their event intervals include that overlap and are not their cost)
    hot = {k: v for k, v in kernels.items() if algorithmic_bytes(k, info, complex_bytes) is not None}
    dominant = max(hot.items(), key=lambda kv: kv[1][1]) if hot else None
    roofline = None
    if dominant:
        kname, (launches, total_ms) = dominant
        alg = algorithmic_bytes(kname, info, complex_bytes)
        if alg is not None:
            achieved = alg / (total_ms / launches / 1e3) / 1e9
            roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": measured_traffic(kname, workload, precision), "kernel": kname,
                        "algorithmic_bytes_per_launch": alg, "mean_launch_ms": round(total_ms / launches, 4),
                        "launches": launches}
    # algorithmic bytes per time step: every launch of the three passes in the timed region (T_Zr and T_xy run
    # (n + 1) / n times per time step: the schedule's n + 1 kinetic blocks per analysis step of n time steps)
    step_bytes = sum((algorithmic_bytes(k, info, complex_bytes) or 0.0) * kernels[k][0]
                     for k in ("zr_plane", "xy_tile", "interaction") if k in kernels) / n_steps
    return {"info": info, "wave": wave, "steps": n_steps, "warmup": warm_analysis * spa,
            "value": dvr_points(info) * n_steps / elapsed, "ms_per_step": elapsed / n_steps * 1e3,
            "roofline": roofline, "step_roofline_frac": round(step_bytes / (elapsed / n_steps) / 1e9 / HBM_PEAK_GBS, 4),
            "kernels": {k: {"launches": n, "ms_per_time_step": round(ms / n_steps, 4)} for k, (n, ms) in kernels.items()}}


def measure_batch(steps: int, warmup: int) -> dict:
    """The 12-run batch on the GPU in graph mode (SPO_HIP_GRAPH=1: every analysis step of the batch is one launch of
    a hipGraph captured over all runs' streams); throughput summed over the runs. Each run starts from its own
    synthetic wave packet. The same number of steps is then timed for each run alone, one run after another (also
    in graph mode), for the batch's gain over running the members one by one."""
    import spodvr_amd
    from tests import cases
    os.environ["SPO_HIP_GRAPH"] = "1"
    runs, points = [], 0
    for name, params in cases.BATCH_12:
        run = spodvr_amd.Run(cases.write_case(f"bench_batch_{name}", params), device="hip")
        run.set_wave(*synthetic_wave(run, seed=len(runs) + 1))
        points += dvr_points(run.info)
        runs.append(run)
    spa = runs[0].info.steps_per_analysis
    timed = max(1, -(-steps // spa))
    try:
        if warmup > 0:
            spodvr_amd.run_batch(runs, max(1, -(-warmup // spa)))
        t0 = time.perf_counter()
        spodvr_amd.run_batch(runs, timed)
        elapsed = time.perf_counter() - t0
        alone = []
        for run in runs:
            spodvr_amd.run_batch([run], 1)  # warm-up of the single-run graph
            t0 = time.perf_counter()
            spodvr_amd.run_batch([run], timed)
            alone.append(time.perf_counter() - t0)
    finally:
        members = [{"name": name, "n_rot": r.info.n_rot, "nr": r.info.nr, "n_theta": r.info.n_theta,
                    "n_phi": r.info.n_phi, "dvr_points": dvr_points(r.info)} for (name, _), r in zip(cases.BATCH_12, runs)]
        for m, a in zip(members, alone):
            m["alone_ms_per_step"] = round(a / (timed * spa) * 1e3, 4)
        for r in runs:
            r.close()
        os.environ.pop("SPO_HIP_GRAPH", None)
    n_steps = timed * spa
    return {"value": points * n_steps / elapsed, "ms_per_step": elapsed / n_steps * 1e3, "steps": n_steps,
            "members": members, "dvr_points": points,
            "one_by_one": {"value": points * n_steps / sum(alone), "ms_per_step": sum(alone) / n_steps * 1e3}}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="time steps timed (rounded up to whole analysis steps)")
    ap.add_argument("--warmup", type=int, default=10, help="untimed time steps before the timed region")
    ap.add_argument("--workload", default="dz", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="both", choices=["both", "fp64", "mixed"],
                    help="both: the mixed-precision line with the FP64 run beside it")
    ap.add_argument("--cpu-steps", type=int, default=10, help="time steps of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-merged", action="store_true",
                    help="skip the merged-half-steps measure (profiling runs of the plain schedule alone)")
    ap.add_argument("--hardwall-steps", type=int, default=500,
                    help="time steps of the HardWall FP64 line beside the DZ line (0: none)")
    ap.add_argument("--hardwall-cpu-steps", type=int, default=100,
                    help="time steps of the HardWall CPU sample (its full run is 2000)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run device copies of BYTES (the counter calibration of the traffic figures)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the propagation of one run is not sharded "
              "across GPUs", file=sys.stderr)
        return 2

    from tests import cases

    if args.workload == "batch12":
        b = measure_batch(args.steps, args.warmup)
        print(json.dumps({
            "metric": METRIC, "value": b["value"], "unit": "grid-point-steps/s", "n_gpus": 1, "steps": b["steps"],
            "warmup": args.warmup, "ms_per_step": b["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32 (complex64 waves, FP64 phases and analysis)", "data": "synthetic",
            "config": {"workload": WORKLOADS["batch12"][1], "precision": "mixed", "runs": len(b["members"]),
                       "dvr_points_summed": b["dvr_points"], "members": b["members"]},
            "one_by_one": b["one_by_one"], "batch_over_one_by_one": b["value"] / b["one_by_one"]["value"],
            "roofline": None, "cpu_baseline": None}), flush=True)
        return 0

    name, description = WORKLOADS[args.workload]
    path = cases.write_case(f"bench_{args.workload}", dict(getattr(cases, name)))
    primary = "mixed" if args.precision in ("both", "mixed") else "fp64"
    fp64 = measure(path, args.workload, "fp64", args.steps, args.warmup) if args.precision in ("both", "fp64") else None
    mixed = measure(path, args.workload, "mixed", args.steps, args.warmup) if args.precision in ("both", "mixed") else None
    main_run = mixed if primary == "mixed" else fp64
    info = main_run["info"]
    # The same workload with merged half steps (opt-in; results equal the plain schedule's up to rounding): reported
    # beside the headline, which stays on the plain schedule.
    merged = None
    if not args.no_merged:
        m = measure(path, args.workload, primary, args.steps, args.warmup, merge_half_steps=True)
        merged = {"precision": primary, "value": m["value"], "ms_per_step": m["ms_per_step"],
                  "over_plain": m["value"] / main_run["value"], "kernels": m["kernels"]}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(path, args.cpu_steps, main_run["wave"])

    def brief(m):
        return None if m is None else {"value": m["value"], "ms_per_step": m["ms_per_step"], "roofline": m["roofline"],
                                        "step_roofline_frac": m["step_roofline_frac"], "kernels": m["kernels"]}

    # The HardWall workload (FP64, the reference's CPU-runnable case) beside the DZ line, with its own CPU sample.
    hardwall = None
    if args.workload == "dz" and args.hardwall_steps > 0:
        hname, hdesc = WORKLOADS["hardwall"]
        hpath = cases.write_case("bench_hardwall", dict(getattr(cases, hname)))
        h = measure(hpath, "hardwall", "fp64", args.hardwall_steps, args.warmup)
        hardwall = dict(brief(h), workload=hdesc, precision="fp64", steps=h["steps"], dvr_points=dvr_points(h["info"]))
        if not args.no_cpu_baseline:
            hardwall["cpu_baseline"] = cpu_baseline(hpath, args.hardwall_cpu_steps, h["wave"])

    line = {
        "metric": METRIC,
        "value": main_run["value"],
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": main_run["steps"],
        "warmup": main_run["warmup"],
        "ms_per_step": main_run["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (complex64 wave, FP64 phases and analysis)" if primary == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": description, "precision": primary, "dvr_points": dvr_points(info),
                   "psi_values": info.n_rot * info.nx * info.ny * info.nr * info.nz,
                   "steps_per_analysis": info.steps_per_analysis, "n_rot": info.n_rot,
                   "nz_potential": info.nz_potential, "n_representatives": info.n_representatives},
        "roofline": main_run["roofline"],
        "mixed_over_fp64": (mixed["value"] / fp64["value"]) if (mixed and fp64) else None,
        "fp64": brief(fp64),
        "mixed": brief(mixed),
        "merged_half_steps": merged,
        "cpu_baseline": cpu,
        "hardwall": hardwall,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
