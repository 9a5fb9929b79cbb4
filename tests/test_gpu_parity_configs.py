"""HIP backend against the CPU oracle at the two benchmark workloads' own shapes, so that the kernel instantiations
the benchmark times are the ones checked:

  hardwall_6d  HardWallPES, NZ 128 x NR 32, 8 x 8 (x, y), JMax 10 (6 theta x 22 phi): zr_reg_kernel<128, 32, ...>,
               xy_reg_kernel<8, 8, ...>, the MFMA interaction kernel with a 22-point phi step;
  dz_6d        Dai-Light H2/Cu(111), NZ 96 x NR 48, 16 x 16, JMax 24 / MJMax 23 (13 theta x 48 phi), C6v-reduced
               potential: zr_reg_kernel<96, 48, 3, 16, 8, 12, ...>, xy_reg_kernel<16, 16, {16, 32}>, the MFMA
               interaction kernel with the 12 x 4 phi step.

Same measure as the reference's CPU-vs-GPU check (tests/cuda/test_cuda_backend.cpp:94-106): max |a - b| / max |a|
after one T block (full, half, kinetic only) and after one analysis step from the same seeded wave packet, 1e-12 in
FP64 and 1e-5 for the complex64 storage mode against the FP64 oracle. Both sides use the host-built potential grid
(the device build is held to it separately). The oracle and the wave packet are built once per workload."""

import os

import numpy as np
import pytest

from tests import cases
from tests.util import pseudo_random, relative_difference

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12
TOL_MIXED = 1e-5

WORKLOADS = {"hardwall_6d": cases.HARDWALL_6D, "dz_6d": cases.DZ_6D}


def _hip_run(path, overrides=None):
    import spodvr_amd
    saved = os.environ.pop("SPO_HIP_POTENTIAL", None)
    os.environ["SPO_HIP_POTENTIAL"] = "host"
    try:
        return spodvr_amd.Run(path, overrides, device="hip")
    finally:
        os.environ.pop("SPO_HIP_POTENTIAL", None)
        if saved is not None:
            os.environ["SPO_HIP_POTENTIAL"] = saved


class Workload:
    def __init__(self, name):
        from oracle import oracle
        self.name = name
        self.path = cases.write_case(name, WORKLOADS[name])
        print(f"workload {name}: oracle set-up", flush=True)
        self.cpu = oracle.run(self.path)
        print(f"workload {name}: HIP FP64 set-up", flush=True)
        self.gpu = _hip_run(self.path)
        i = self.gpu.info
        print(f"workload {name}: {i.nz} x {i.nr} Z r, {i.nx} x {i.ny} xy, n_rot {i.n_rot}, "
              f"{i.n_theta} theta x {i.n_phi} phi, {i.n_representatives} representatives", flush=True)
        self.scattering = pseudo_random(self.gpu.scattering_size, 1)
        self.specular = pseudo_random(self.gpu.specular_size, 2)

    def reset(self, *runs):
        for r in runs:
            r.set_wave(self.scattering, self.specular)

    def close(self):
        self.gpu.close()
        self.cpu.close()


@pytest.fixture(scope="module", params=list(WORKLOADS))
def workload(request):
    w = Workload(request.param)
    yield w
    w.close()


def _compare(cpu, gpu):
    sc, pc = cpu.get_wave()
    sg, pg = gpu.get_wave()
    return relative_difference(sc, sg), relative_difference(pc, pg), sc, pc


@pytest.mark.parametrize("half,kinetic_only", [(False, False), (True, False), (True, True)])
def test_t_block_matches_oracle(workload, half, kinetic_only):
    w = workload
    print(f"{w.name} fp64 T block half {half} kinetic_only {kinetic_only} (seed 1)", flush=True)
    w.reset(w.cpu, w.gpu)
    w.gpu.kinetic_and_potential(half, kinetic_only)
    w.cpu.kinetic_and_potential(half, kinetic_only)
    ds, dp, sc, pc = _compare(w.cpu, w.gpu)
    print(f"  scattering {ds:.3e} specular {dp:.3e}", flush=True)
    assert ds < TOL_FP64 and dp < TOL_FP64
    # the block really moved the wave packet
    assert relative_difference(w.scattering, sc) > 1e-3 and relative_difference(w.specular, pc) > 1e-3


def test_analysis_step_matches_oracle(workload):
    w = workload
    print(f"{w.name} fp64 analysis step (seed 1)", flush=True)
    w.reset(w.cpu, w.gpu)
    w.gpu.propagate()
    w.cpu.propagate()
    bg, bc = w.gpu.analyse(10.0), w.cpu.analyse(10.0)
    db = relative_difference(bc, bg)
    dsum = relative_difference(w.cpu.s_sums(), w.gpu.s_sums())
    dens = relative_difference(w.cpu.block_densities(), w.gpu.block_densities())
    iz = w.gpu.info.iz_analysis
    dplane = relative_difference(w.cpu.z_plane(iz), w.gpu.z_plane(iz))
    dcol = relative_difference(w.cpu.site_column(5), w.gpu.site_column(5))
    ds, dp, _, _ = _compare(w.cpu, w.gpu)
    print(f"  bk {db:.3e} sums {dsum:.3e} densities {dens:.3e} plane {dplane:.3e} column {dcol:.3e} "
          f"wave {ds:.3e} specular {dp:.3e}", flush=True)
    for d in (db, dsum, dens, dplane, dcol, ds, dp):
        assert d < TOL_FP64


def test_mixed_precision_matches_fp64_oracle(workload):
    """The complex64 storage mode on one full T block and then one analysis step, against the FP64 oracle from the
    same wave packet: single-precision rounding only."""
    w = workload
    print(f"{w.name} mixed: HIP complex64 set-up", flush=True)
    mixed = _hip_run(w.path, {"Precision": "mixed"})
    try:
        print(f"{w.name} mixed T block (seed 1)", flush=True)
        w.reset(w.cpu, mixed)
        mixed.kinetic_and_potential(False, False)
        w.cpu.kinetic_and_potential(False, False)
        ds, dp, _, _ = _compare(w.cpu, mixed)
        print(f"  scattering {ds:.3e} specular {dp:.3e}", flush=True)
        assert ds < TOL_MIXED and dp < TOL_MIXED
        assert ds > 1e-9  # the mode really stores complex64
        print(f"{w.name} mixed analysis step (seed 1)", flush=True)
        w.reset(w.cpu, mixed)
        mixed.propagate()
        w.cpu.propagate()
        bg, bc = mixed.analyse(10.0), w.cpu.analyse(10.0)
        ds, dp, _, _ = _compare(w.cpu, mixed)
        db = relative_difference(bc, bg)
        print(f"  wave {ds:.3e} specular {dp:.3e} bk {db:.3e}", flush=True)
        assert ds < TOL_MIXED and dp < TOL_MIXED and db < TOL_MIXED
    finally:
        mixed.close()


def test_kinetic_half_steps_compose_and_keep_the_norm(workload):
    """The reference's ScatteringKinetic.ConservesTheNormAndHalfStepsComposeToFullSteps
    (tests/test_propagation.cpp:99-118) on the device at the workload's shape: a kinetic-only block (specular grid,
    T_Zr, T_xy) is unitary, and two half blocks equal one full block (exp(-i T dt / 2)^2 = exp(-i T dt) in
    exact arithmetic)."""
    w = workload
    print(f"{w.name} fp64 kinetic half steps compose (seed 1)", flush=True)
    w.reset(w.gpu)
    w.gpu.kinetic_and_potential(False, True)
    full, full_spec = w.gpu.get_wave()
    w.reset(w.gpu)
    w.gpu.kinetic_and_potential(True, True)
    w.gpu.kinetic_and_potential(True, True)
    halves, halves_spec = w.gpu.get_wave()
    norm0 = np.vdot(w.scattering, w.scattering).real
    norm1 = np.vdot(full, full).real
    d = relative_difference(full, halves)
    dspec = relative_difference(full_spec, halves_spec)
    print(f"  norm ratio - 1 {norm1 / norm0 - 1:.3e}, full vs halves {d:.3e} (specular {dspec:.3e})", flush=True)
    assert abs(norm1 / norm0 - 1.0) < 1e-13
    assert d < 1e-12 and dspec < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_parallel_plane_wave_gets_its_kinetic_phase_on_device(precision):
    """tests/test_propagation.cpp:179-206 on the device at the DZ workload's (x, y) grid (16 x 16, hexagonal cell)
    with a parallel momentum (EParallel 0.05 eV at 30 degrees, so the momentum shift of T_xy is active): the plane
    wave exp(i (k_par + G) . R) comes back multiplied by exp(-i |k_par + G|^2 dt / 2M), the free-particle formula on
    the lattice's reciprocal vectors."""
    from tests import closed_forms as cf
    print(f"dz_6d plane wave with EParallel 0.05 eV at 30 degrees, {precision}", flush=True)
    path = cases.write_case("dz_6d", cases.DZ_6D)
    run = _hip_run(path, {"EParallel": 0.05, "PhiParallel": 30.0, "Precision": precision})
    try:
        i = run.info
        cell = run.array("derived.cell")
        values, k2 = cf.plane_wave(cell, i.nx, i.ny, 1, -1)
        wave = cf.host_wave(values[None, :, None], i.n_rot, i.nx * i.ny, i.nz * i.nr)
        run.set_wave(wave, np.zeros(run.specular_size, complex))
        run.kinetic_and_potential(False, True)
        s, _ = run.get_wave()
        err = float(np.abs(s - np.exp(-1j * k2 / (2 * cell[3]) * cell[6]) * wave).max())
        print(f"  |K|^2 {k2:.4f}: max error {err:.2e}", flush=True)
        assert err < (1e-12 if precision == "fp64" else 2e-6)
    finally:
        run.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_hard_wall_has_no_angular_coupling_on_device(precision):
    """At the HardWall workload's shape (no angular dependence of V) a wave packet in one rotational state, with an
    empty specular grid, stays in that state through a full T block with V on the device, with one factor per
    (Z, r) point at every site: the MFMA theta contractions and the phi FFTs of the interaction kernel compose to the
    identity."""
    print(f"hardwall_6d single rotational state through a T block, {precision}", flush=True)
    path = cases.write_case("hardwall_6d", cases.HARDWALL_6D)
    run = _hip_run(path, {"Precision": precision})
    try:
        i = run.info
        n_xy, n_zr = i.nx * i.ny, i.nz * i.nr
        irot = 5 if i.initial_channel != 5 else 6
        rng = np.random.default_rng(7)
        column = np.zeros((n_zr, n_xy, i.n_rot), complex)
        column[:, :, irot] = rng.standard_normal((n_zr, 1)) + 1j * rng.standard_normal((n_zr, 1))
        run.set_wave(column.reshape(-1), np.zeros(run.specular_size, complex))
        run.kinetic_and_potential(False, False)
        s, _ = run.get_wave()
        s = s.reshape(n_zr, n_xy, i.n_rot)
        leak = float(np.abs(np.delete(s, irot, axis=2)).max() / np.abs(s[:, :, irot]).max())
        kept = s[:, :, irot]
        spread = float(np.abs(kept - kept[:, :1]).max() / np.abs(kept).max())
        print(f"  state {irot}: leak {leak:.2e}, site spread {spread:.2e}", flush=True)
        tol = 1e-13 if precision == "fp64" else 1e-6
        assert leak < tol and spread < tol
    finally:
        run.close()
