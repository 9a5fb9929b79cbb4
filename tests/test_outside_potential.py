"""spo_run_create_with_potential (a potential grid built outside the set-up, e.g. by the surrogate), on the CPU
oracle: the set-up's own grid passed back gives bitwise the run created without it, a changed grid changes the
propagation, and a grid of the wrong size or with a value that is not finite is refused. The grid axes and the
unit factor that a caller needs to build such a grid are read back (spo_run_array)."""

import numpy as np
import pytest

import spodvr_amd
from oracle import oracle
from tests import cases


def _oracle_run(path, potential=None):
    return spodvr_amd.Run(path, None, device="external", factory=oracle.factory(), potential=potential)


def _two_steps(run):
    run.set_wave(*_wave(run))
    run.steps(2)
    return run.bk_records(), run.get_wave()[0]


def _wave(run):
    from tests.util import pseudo_random
    return pseudo_random(run.scattering_size, 5), pseudo_random(run.specular_size, 6)


def test_own_grid_passed_back_is_the_same_run():
    path = cases.write_case("smoke_outside_potential", cases.SMOKE_DAILIGHT)
    plain = _oracle_run(path)
    try:
        grid = plain.array("potential.values")
        bk_a, psi_a = _two_steps(plain)
    finally:
        plain.close()
    again = _oracle_run(path, potential=grid)
    try:
        assert np.array_equal(again.array("potential.values"), grid)
        bk_b, psi_b = _two_steps(again)
    finally:
        again.close()
    assert np.array_equal(bk_a, bk_b) and np.array_equal(psi_a, psi_b)
    changed = _oracle_run(path, potential=grid + 1e-3)  # a constant shift of 1 mhartree: a phase, visible in psi
    try:
        _, psi_c = _two_steps(changed)
    finally:
        changed.close()
    assert not np.array_equal(psi_a, psi_c)


def test_bad_grids_are_refused():
    path = cases.write_case("smoke_outside_potential", cases.SMOKE_DAILIGHT)
    probe = _oracle_run(path)
    grid = probe.array("potential.values")
    probe.close()
    with pytest.raises(spodvr_amd.SpoError, match="values for a potential grid"):
        _oracle_run(path, potential=grid[:-1])
    bad = grid.copy()
    bad[7] = np.nan
    with pytest.raises(spodvr_amd.SpoError, match="not finite"):
        _oracle_run(path, potential=bad)


def test_grid_axes_for_an_outside_build():
    path = cases.write_case("smoke_outside_potential", cases.SMOKE_DAILIGHT)
    run = _oracle_run(path)
    try:
        i = run.info
        x, y, z, r = (run.array(f"grids.{k}") for k in "xyzr")
        assert (len(x), len(y), len(z), len(r)) == (i.nx, i.ny, i.nz, i.nr)
        assert run.array("units.hartree_in_ev")[0] == pytest.approx(27.2114)
        n = i.n_theta * i.n_phi * i.n_representatives * i.nz_potential * i.nr
        assert run.array("potential.values").size == n
    finally:
        run.close()


def test_surrogate_grid_points_are_the_setups():
    """surrogate.grid_points lists the geometries of a run's potential grid in its own order: the CRP at those
    geometries (QD wrapper, host), converted with the set-up's factor and clamped to VMax, is the set-up's grid on
    every row below the switch to the gas phase (crp_cu111_srp48_small, 2000 random entries)."""
    from spodvr_amd import surrogate
    path, over = cases.golden_input("crp_cu111_srp48_small"), cases.golden_overrides("crp_cu111_srp48_small")
    run = spodvr_amd.Run(path, over, device="external", factory=oracle.factory())
    crp = spodvr_amd.CrpSurface(over["CRPlibDataSet"])
    try:
        pts, below = surrogate.grid_points(run)
        grid = run.array("potential.values")
        assert len(pts) == grid.size and 0.9 < below.mean() < 1.0
        k = np.random.default_rng(0).choice(np.flatnonzero(below), 2000, replace=False)
        expect = np.minimum(crp.qd(pts[k]), 10.0) * (1.0 / run.array("units.hartree_in_ev")[0])
        assert np.max(np.abs(grid[k] - expect)) < 1e-15
    finally:
        crp.close()
        run.close()
