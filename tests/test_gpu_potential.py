"""The potential grid built on the GPU (spo_hip_potential_grid, the default set-up of a HIP run) against the
reference's own values (gate 1: pes_reference.txt, 1e-13 of max(|a|, |b|, 1 eV), tests/pes_gate.hpp:24-28), against
the host build of the same grid, and the device surfaces against the host ones at scattered points.

Every test prints its case before touching the GPU."""

import os
import sys

import numpy as np
import pytest

from tests import cases, fixtures

pytestmark = pytest.mark.gpu

FLOOR = 1.0 / 27.2114  # 1 eV in hartree: near V = 0 the LEPS terms cancel and a pure ratio means nothing
TOL = 1e-13


def _note(*what):
    print("potential:", *what, file=sys.stderr, flush=True)


def _run(path, where):
    """A HIP run whose potential grid is built on the device (where="device") or on the host."""
    import spodvr_amd
    old = os.environ.get("SPO_HIP_POTENTIAL")
    os.environ["SPO_HIP_POTENTIAL"] = where
    try:
        return spodvr_amd.Run(path, device="hip")
    finally:
        if old is None:
            os.environ.pop("SPO_HIP_POTENTIAL", None)
        else:
            os.environ["SPO_HIP_POTENTIAL"] = old


def _relative(a, b, floor):
    return float(np.max(np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), floor)))


@pytest.mark.parametrize("case", ["cu111_dailight_short_sym", "cu111_dailight_oblique"])
def test_device_grid_matches_fortran(case):
    _note("gate 1", case)
    run = _run(cases.golden_input(case), "device")
    try:
        f = np.load(fixtures.case_dir(case) + "/pes.npz")
        nzpot, nr, nx, ny, nth, nph = (int(x) for x in f["dims"])
        info = run.info
        assert (info.nz_potential, info.nr, info.nx, info.ny, info.n_theta, info.n_phi) == (nzpot, nr, nx, ny, nth, nph)
        values = run.array("potential.values")
        rep = run.array("symmetry.representative_of").astype(int)
        op = run.array("symmetry.operation_of").astype(int)
        phi_index = run.array("symmetry.phi_index").astype(int).reshape(-1, nph)
        n_tp, n_rep = nth * nph, info.n_representatives
        idx = f["index"]
        it, ip, ix, iy, iz, ir = (idx[:, k] for k in range(6))
        ixy = iy + ny * ix
        izr = ir + nr * iz
        mine = values[phi_index[op[ixy], ip] + nph * it + n_tp * (rep[ixy] + n_rep * izr)]
        assert _relative(mine, f["value"], FLOOR) <= TOL
    finally:
        run.close()


@pytest.mark.parametrize("case", ["hardwall_short", "cu111_dailight_short", "dz_angular", "dz"])
def test_device_grid_matches_host_grid(case):
    path = {"hardwall_short": lambda: cases.golden_input("hardwall_short"),
            "cu111_dailight_short": lambda: cases.golden_input("cu111_dailight_short"),
            "dz_angular": lambda: cases.write_case("dz_angular", cases.DZ_ANGULAR),
            "dz": lambda: cases.write_case("dz", cases.DZ_6D)}[case]()
    import time
    _note("device vs host grid", case)
    t0 = time.perf_counter()
    device = _run(path, "device")
    t1 = time.perf_counter()
    try:
        dv = device.array("potential.values")
    finally:
        device.close()
    t2 = time.perf_counter()
    host = _run(path, "host")
    t3 = time.perf_counter()
    try:
        hv = host.array("potential.values")
    finally:
        host.close()
    _note(f"{case}: {dv.size} values; run set-up {t1 - t0:.2f} s with the device grid, {t3 - t2:.2f} s with the host grid")
    assert dv.shape == hv.shape and dv.size > 0
    assert _relative(dv, hv, FLOOR) <= TOL


@pytest.mark.parametrize("system_id", ["H2Cu111DaiLight", "HardWallPES"])
def test_device_surface_matches_host_surface(system_id):
    import spodvr_amd
    _note("device surface at scattered points", system_id)
    rng = np.random.default_rng(7)
    n = 3000
    pts = np.column_stack([rng.uniform(0, 4.84, n), rng.uniform(0, 4.84, n), rng.uniform(0.05, 18.0, n),
                           rng.uniform(0.6, 6.0, n), rng.uniform(0, np.pi, n), rng.uniform(0, 2 * np.pi, n)])
    # the switch points of the two surfaces exactly
    pts[0, 2], pts[1, 2], pts[2, 2], pts[3, 2] = 15.0, 7.0, 0.1778, 15.0 + 5e-13
    dev = spodvr_amd.device_pes_values(system_id, pts)
    host = np.array([spodvr_amd.pes_value(system_id, *p) for p in pts])
    assert _relative(dev, host, 1.0) <= TOL


def test_device_surface_rejects_unknown_system():
    import spodvr_amd
    with pytest.raises(spodvr_amd.SpoError):
        spodvr_amd.device_pes_values("H2Cu111SRP48", np.zeros((1, 6)))
