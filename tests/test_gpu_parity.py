"""HIP backend against the CPU oracle, from the same seeded pseudo-random wave packet (the reference's
CPU-vs-GPU check, tests/cuda/test_cuda_backend.cpp:94-106: max |a - b| / max |a| per stage, 1e-13 there).

Every test prints the case, the seed and the stage before touching the GPU."""

import sys

import numpy as np
import pytest

from tests import cases
from tests.util import pseudo_random, relative_difference

pytestmark = pytest.mark.gpu

TOL = 1e-12

CASE_PARAMS = {
    "smoke": lambda: cases.write_case("smoke_dailight", cases.SMOKE_DAILIGHT),
    "generic": lambda: cases.write_case("smoke_generic", cases.SMOKE_GENERIC),
    "dz_angular": lambda: cases.write_case("dz_angular", cases.DZ_ANGULAR),
    "jmax12_angular": lambda: cases.write_case("jmax12_angular", cases.JMAX12_ANGULAR),
    "jmax18_angular": lambda: cases.write_case("jmax18_angular", cases.JMAX18_ANGULAR),
    "hardwall_short": lambda: cases.golden_input("hardwall_short"),
    "cu111_dailight_short": lambda: cases.golden_input("cu111_dailight_short"),
    "cu111_dailight_short_sym": lambda: cases.golden_input("cu111_dailight_short_sym"),
    "cu111_dailight_oblique": lambda: cases.golden_input("cu111_dailight_oblique"),
}


def _pair(case, overrides=None, interaction=None):
    """A HIP run and an oracle run of one case. interaction="lds" makes the HIP backend use its LDS-contraction
    interaction kernel instead of the MFMA one (SPO_HIP_INTERACTION is read when the backend is created). Both
    runs take the potential grid from the host build, so that these tests compare the propagation alone: the
    device build of the grid agrees with it to rounding (test_gpu_potential.py), but at the strongly repulsive
    points (V of hundreds of hartree) a last-bit difference in V is a phase difference of ~1e-11 after V dt."""
    import os
    import spodvr_amd
    from oracle import oracle
    path = CASE_PARAMS[case]()
    saved = {k: os.environ.pop(k, None) for k in ("SPO_HIP_INTERACTION", "SPO_HIP_POTENTIAL")}
    if interaction:
        os.environ["SPO_HIP_INTERACTION"] = interaction
    os.environ["SPO_HIP_POTENTIAL"] = "host"
    try:
        gpu = spodvr_amd.Run(path, overrides, device="hip")
    finally:
        for k, v in saved.items():
            os.environ.pop(k, None)
            if v is not None:
                os.environ[k] = v
    cpu = oracle.run(path, overrides)
    return gpu, cpu


def _randomise(gpu, cpu, seed):
    s = pseudo_random(gpu.scattering_size, seed)
    p = pseudo_random(gpu.specular_size, seed + 1000)
    gpu.set_wave(s, p)
    cpu.set_wave(s, p)
    return s, p


@pytest.mark.parametrize("case", list(CASE_PARAMS))
@pytest.mark.parametrize("half,kinetic_only", [(False, False), (True, False), (True, True)])
def test_one_t_block_matches_oracle(case, half, kinetic_only):
    seed = 1 + 2 * int(half) + int(kinetic_only)
    print(f"case {case} seed {seed} half {half} kinetic_only {kinetic_only}", flush=True)
    gpu, cpu = _pair(case)
    s0, p0 = _randomise(gpu, cpu, seed)
    gpu.kinetic_and_potential(half, kinetic_only)
    cpu.kinetic_and_potential(half, kinetic_only)
    sg, pg = gpu.get_wave()
    sc, pc = cpu.get_wave()
    ds, dp = relative_difference(sc, sg), relative_difference(pc, pg)
    print(f"  scattering {ds:.3e} specular {dp:.3e}", flush=True)
    assert ds < TOL and dp < TOL
    assert relative_difference(s0, sc) > 1e-3 and relative_difference(p0, pc) > 1e-3


@pytest.mark.parametrize("case", ["smoke", "generic", "dz_angular", "cu111_dailight_short_sym",
                                  "cu111_dailight_oblique"])
def test_analysis_step_matches_oracle(case):
    print(f"case {case} seed 11 analysis step", flush=True)
    gpu, cpu = _pair(case)
    _randomise(gpu, cpu, 11)
    gpu.propagate()
    cpu.propagate()
    bg, bc = gpu.analyse(10.0), cpu.analyse(10.0)
    db = relative_difference(bc, bg)
    dsum = relative_difference(cpu.s_sums(), gpu.s_sums())
    dens = relative_difference(cpu.block_densities(), gpu.block_densities())
    print(f"  bk {db:.3e} sums {dsum:.3e} densities {dens:.3e}", flush=True)
    assert db < TOL and dsum < TOL and dens < TOL
    iz = gpu.info.iz_analysis
    assert relative_difference(cpu.z_plane(iz), gpu.z_plane(iz)) < TOL
    assert relative_difference(cpu.site_column(3), gpu.site_column(3)) < TOL


@pytest.mark.parametrize("case", ["smoke", "generic", "dz_angular", "jmax12_angular", "jmax18_angular",
                                  "cu111_dailight_short_sym", "cu111_dailight_oblique"])
def test_mixed_precision_t_block_against_fp64_oracle(case):
    """The complex64 storage mode against the FP64 oracle on one full T block from the same wave packet:
    single-precision rounding only (max |a - b| / max |a| < 1e-5)."""
    print(f"case {case} seed 21 mixed precision", flush=True)
    gpu, cpu = _pair(case, {"Precision": "mixed"})
    _randomise(gpu, cpu, 21)
    gpu.kinetic_and_potential(False, False)
    cpu.kinetic_and_potential(False, False)
    sg, pg = gpu.get_wave()
    sc, pc = cpu.get_wave()
    ds, dp = relative_difference(sc, sg), relative_difference(pc, pg)
    print(f"  scattering {ds:.3e} specular {dp:.3e}", flush=True)
    assert ds < 1e-5 and dp < 1e-5
    assert ds > 1e-9  # the mode really stores complex64


@pytest.mark.parametrize("case", ["smoke", "dz_angular", "jmax12_angular", "jmax18_angular",
                                  "cu111_dailight_short_sym", "cu111_dailight_oblique"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("interaction", ["lds", "mfma_staged", "mfma", "buf"])
def test_interaction_variants_match_oracle(case, precision, interaction):
    """Every interaction kernel (vector-unit contractions; MFMA contractions with staged phi FFTs, with the fused
    register phi step (the FP64 default), and with buffer addressing (the complex64 default)) on one full T
    block."""
    print(f"case {case} seed 31 interaction {interaction} {precision}", flush=True)
    gpu, cpu = _pair(case, {"Precision": precision}, interaction=interaction)
    _randomise(gpu, cpu, 31)
    gpu.kinetic_and_potential(False, False)
    cpu.kinetic_and_potential(False, False)
    sg, pg = gpu.get_wave()
    sc, pc = cpu.get_wave()
    ds, dp = relative_difference(sc, sg), relative_difference(pc, pg)
    print(f"  scattering {ds:.3e} specular {dp:.3e}", flush=True)
    tol = TOL if precision == "fp64" else 1e-5
    assert ds < tol and dp < tol


def test_bf16x3_contraction_t_block_against_fp64_oracle():
    """The split-BF16x3 theta contraction (SPO_HIP_CONTRACTION=bf16x3, complex64, the 48-point phi step of the DZ
    angular basis) against the FP64 oracle on one full T block: L enters with about 16 significant bits, so the bound
    is ten times the FP32-MFMA path's."""
    import os
    print("case dz_angular seed 61 mixed bf16x3", flush=True)
    old = os.environ.get("SPO_HIP_CONTRACTION")
    os.environ["SPO_HIP_CONTRACTION"] = "bf16x3"
    try:
        gpu, cpu = _pair("dz_angular", {"Precision": "mixed"})
    finally:
        os.environ.pop("SPO_HIP_CONTRACTION", None)
        if old is not None:
            os.environ["SPO_HIP_CONTRACTION"] = old
    _randomise(gpu, cpu, 61)
    gpu.kinetic_and_potential(False, False)
    cpu.kinetic_and_potential(False, False)
    sg, pg = gpu.get_wave()
    sc, pc = cpu.get_wave()
    ds, dp = relative_difference(sc, sg), relative_difference(pc, pg)
    print(f"  scattering {ds:.3e} specular {dp:.3e}", flush=True)
    assert ds < 1e-4 and dp < 1e-4
    assert ds > 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_two_runs_are_bitwise_identical(precision):
    """Determinism (the reference's CUDA test TwoRunsAreBitwiseIdentical, tests/cuda/test_cuda_backend.cpp:388):
    no atomics and fixed reduction orders, so two runs of the same input agree bit for bit."""
    import spodvr_amd
    print(f"case smoke two runs {precision}", flush=True)
    path = CASE_PARAMS["smoke"]()
    out = []
    for _ in range(2):
        run = spodvr_amd.Run(path, {"Precision": precision}, device="hip")
        run.set_wave(pseudo_random(run.scattering_size, 41), pseudo_random(run.specular_size, 42))
        run.steps(3)
        s, p = run.get_wave()
        out.append((s, p, run.bk_records(), run.s_sums(), run.block_densities()))
        run.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case,settings", [("dz_angular", ["interaction=16", "interaction=#2", "xy=32", "xy=64"]),
                                           ("hardwall_short", ["zr=256", "zr=512", "xy=32", "xy=64"]),
                                           ("dz", ["xy=8", "xy=16", "xy=32"] + [f"zr=#{k}" for k in range(1, 7)])])
def test_launch_variants_match_the_default(case, settings, precision):
    """The launch variants (block sizes, tile widths) against the default variant after a full T block: the same
    arithmetic up to the compiler's multiply-add contraction, so within the parity bar (FP64 1e-13, complex64 1e-6
    relative)."""
    import os
    import spodvr_amd
    # dz: the 16 x 16 (x, y) grid of the DZ workload, whose T_xy tiles are 8, 16 or 32 izr wide
    path = cases.write_case("dz", cases.DZ_6D) if case == "dz" else CASE_PARAMS[case]()
    tol = 1e-13 if precision == "fp64" else 1e-6
    if case == "dz_angular" and precision == "mixed":
        settings = settings + ["interaction=#3"]  # the complex64 default kernel has a fourth (8 x 6 phi radices)
    wave = None
    reference = None
    for setting in ["default"] + settings:
        print(f"case {case} seed 51 {precision} SPO_HIP_TUNE={setting}", flush=True)
        if setting != "default":
            os.environ["SPO_HIP_TUNE"] = setting
        try:
            run = spodvr_amd.Run(path, {"Precision": precision}, device="hip")
        finally:
            os.environ.pop("SPO_HIP_TUNE", None)
        if wave is None:
            wave = (pseudo_random(run.scattering_size, 51), pseudo_random(run.specular_size, 52))
        run.set_wave(*wave)
        run.kinetic_and_potential(False, False)
        result = run.get_wave()
        run.close()
        if reference is None:
            reference = result
            continue
        ds = relative_difference(reference[0], result[0])
        dp = relative_difference(reference[1], result[1])
        print(f"  {setting}: scattering {ds:.3e} specular {dp:.3e}", flush=True)
        assert ds < tol and dp < tol, setting


@pytest.mark.parametrize("case", ["smoke", "dz_angular"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_energy_resolved_sums_match_oracle(case, precision):
    """The device-side energy-resolved sums (two sites, two energies) against the oracle's accumulator after three
    analysis steps from the same wave packet."""
    from tests.test_energy_resolved import SITES
    print(f"case {case} seed 71 energy-resolved {precision}", flush=True)
    gpu, cpu = _pair(case, dict(SITES, Precision=precision))
    _randomise(gpu, cpu, 71)
    gpu.steps(3)
    cpu.steps(3)
    (sg, pg), (sc, pc) = gpu.energy_resolved(), cpu.energy_resolved()
    ds, dp = relative_difference(sc, sg), relative_difference(pc, pg)
    print(f"  sums {ds:.3e} specular {dp:.3e}", flush=True)
    tol = TOL if precision == "fp64" else 1e-5
    assert ds < tol and dp < tol
