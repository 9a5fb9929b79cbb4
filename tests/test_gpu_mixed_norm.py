"""The complex64 storage mode keeps the norm without a drift: at the DZ and HardWall workloads' shapes, with the
absorbers moved off the grid (so every step is unitary up to the projection onto the specular grid), a seeded wave
packet that fills the whole grid is propagated through 40 full time steps (specular grid, T_Zr, T_xy, V) once with
an FP64 wave and once in complex64, and the norms (the sums of the block densities, in FP64) are compared.

Rounding a factor that the step applies every time (a kinetic phase, a twiddle, exp(i A), the Legendre matrix) to
complex64 changes what it does to the norm by up to ~6e-8, the same at every step, so such an error grows linearly
with the step count (2e4 steps: ~1e-3 of a probability), while the rounding of the wave itself only adds a random
walk. Measured on MI355X after 40 steps: 1.6e-7 (DZ) and 1.1e-7 (HardWall) relative; with the factor tables'
remainders dropped (plain complex64 factors) 3.7e-6 and 3.7e-6. The bound is twice the larger measured value."""

import numpy as np
import pytest

from tests import cases
from tests.util import pseudo_random

pytestmark = pytest.mark.gpu

STEPS = 40
BOUND = 3.2e-7  # relative norm difference after STEPS time steps
NO_ABSORBERS = {"ZOptStart": 100.0, "ZOptEnd": 101.0, "ROptStart": 100.0, "ROptEnd": 101.0}


def _norm_after_steps(path, precision, scattering, specular):
    import spodvr_amd
    run = spodvr_amd.Run(path, {"Precision": precision}, device="hip")
    try:
        run.set_wave(scattering, specular)
        start = float(run.block_densities().sum())
        for _ in range(STEPS):
            run.kinetic_and_potential(False, False)
        return start, float(run.block_densities().sum())
    finally:
        run.close()


@pytest.mark.parametrize("workload", ["dz_6d", "hardwall_6d"])
def test_complex64_norm_follows_fp64(workload):
    params = dict({"dz_6d": cases.DZ_6D, "hardwall_6d": cases.HARDWALL_6D}[workload], **NO_ABSORBERS)
    path = cases.write_case(f"{workload}_no_absorbers", params)
    import spodvr_amd
    probe = spodvr_amd.Run(path, device="hip")
    n_s, n_p = probe.scattering_size, probe.specular_size
    probe.close()
    scattering = pseudo_random(n_s, 3)
    specular = np.zeros(n_p, complex)
    print(f"{workload}: {STEPS} time steps from seeded wave 3 without absorbers, FP64 then complex64", flush=True)
    s64, n64 = _norm_after_steps(path, "fp64", scattering, specular)
    s32, n32 = _norm_after_steps(path, "mixed", scattering, specular)
    rel = abs(n32 - n64) / n64
    print(f"  start {s64:.10e} (complex64 {s32:.10e}); after: FP64 {n64:.12e} complex64 {n32:.12e}; "
          f"relative difference {rel:.3e} ({rel / STEPS:.2e} per step)", flush=True)
    assert rel < BOUND
