"""Closed forms that pin the set-up shared by the HIP backend and the CPU oracle at the workloads' own shapes (JMax 24
with MJMax 23 on a 16 x 16 C6v grid; JMax 10 on 8 x 8), and the oracle's propagation steps. The Legendre matrix,
the theta grid and the symmetry maps are the arrays the GPU receives (spo_run_array); they are checked against
numpy / scipy and the mathematics, not against the code that made them. Every check also fails on a planted error of
a stated size, so that none of them passes vacuously."""

import numpy as np
import pytest

from tests import cases
from tests import closed_forms as cf
from tests.util import pseudo_random, relative_difference

ANGULAR = {
    "dz_jmax24_16x16": cases.DZ_SETUP,
    "jmax10_8x8": cases.DZ_SETUP_8X8,
    "hardwall_6d": cases.HARDWALL_6D,
    "odd_jmax23": dict(cases.DZ_SETUP_8X8, JMax=23, MJMax=23, JIn=1, nX=4, nY=4),
    "heteronuclear_jmax9": dict(cases.DZ_SETUP_8X8, JMax=9, MJMax=9, HomoNuc=False, Mass2=2.0141, nX=4, nY=4),
}


def _setup_run(params, name, overrides=None):
    from oracle import oracle
    return oracle.run(cases.write_case(name, params), overrides)


@pytest.fixture(scope="module", params=list(ANGULAR))
def angular(request):
    run = _setup_run(ANGULAR[request.param], request.param)
    a = dict(name=request.param, params=ANGULAR[request.param], thetas=run.array("angular.thetas"),
             weights=run.array("angular.weights"), legendre=run.array("angular.legendre"),
             n_phi=run.info.n_phi, blocks=cf.blocks_of(run.array("basis.j"), run.array("basis.mj")))
    run.close()
    return a


def _j_max(a):
    return max(max(js) for _, js in a["blocks"])


def test_theta_grid_is_the_gauss_legendre_rule(angular):
    a = angular
    homo = a["params"].get("HomoNuc", True)
    err = cf.gauss_error(a["thetas"], a["weights"], a["params"]["JMax"], homo)
    print(f"{a['name']}: {len(a['thetas'])} theta, gauss error {err:.2e}")
    assert err < 1e-14
    bad = a["weights"].copy()
    bad[len(bad) // 2] *= 1 + 1e-12
    assert cf.gauss_error(a["thetas"], bad, a["params"]["JMax"], homo) > 1e-14


def test_legendre_matrix_matches_scipy(angular):
    a = angular
    err = cf.legendre_error(a["legendre"], a["thetas"], a["weights"], a["n_phi"], a["blocks"])
    print(f"{a['name']}: legendre matrix {len(a['legendre'])} values, error {err:.2e}")
    assert err < 1e-13
    bad = a["legendre"].copy()
    bad[int(np.argmax(np.abs(bad)))] *= -1  # a sign slip in one entry
    assert cf.legendre_error(bad, a["thetas"], a["weights"], a["n_phi"], a["blocks"]) > 1e-3


def test_angular_round_trip_is_identity(angular):
    a = angular
    err = cf.round_trip_error(a["legendre"], len(a["thetas"]), a["n_phi"], a["blocks"])
    print(f"{a['name']}: round trip error {err:.2e}")
    assert err < 1e-13
    bad = a["legendre"].copy()
    bad[int(np.argmax(np.abs(bad)))] *= 1 + 1e-10
    assert cf.round_trip_error(bad, len(a["thetas"]), a["n_phi"], a["blocks"]) > 1e-11


def test_cos_squared_matrix_elements_are_exact(angular):
    a = angular
    err = cf.cos2_error(a["legendre"], a["thetas"], a["n_phi"], a["blocks"], _j_max(a))
    print(f"{a['name']}: cos^2 matrix error {err:.2e}")
    assert err < 1e-13
    bad = a["thetas"].copy()
    bad[1] += 1e-9  # a theta node off by 1e-9 rad
    assert cf.cos2_error(a["legendre"], bad, a["n_phi"], a["blocks"], _j_max(a)) > 1e-11


@pytest.mark.parametrize("name", ["dz_jmax24_16x16", "jmax10_8x8"])
def test_symmetry_reduced_grid_equals_the_full_grid(name):
    """The potential grid stored on the representatives and read through the maps (operation per site, phi index
    per operation) equals the grid evaluated at every site (UseSymmetry false)."""
    params = ANGULAR[name]
    sym = _setup_run(params, name + "_sym")
    full = _setup_run(dict(params, UseSymmetry=False), name + "_full")
    i = sym.info
    n_xy = i.nx * i.ny
    assert full.info.n_representatives == n_xy and i.n_representatives < n_xy
    v_sym, v_full = sym.array("potential.values"), full.array("potential.values")
    rep, op = sym.array("symmetry.representative_of"), sym.array("symmetry.operation_of")
    phi = sym.array("symmetry.phi_index")
    sym.close()
    full.close()
    err = cf.symmetry_error(v_sym, v_full, rep, op, phi, i.n_theta, i.n_phi, n_xy)
    print(f"{name}: {i.n_representatives} representatives of {n_xy}, symmetric vs full {err:.2e}")
    assert err < 1e-13
    # a wrong phi map of one operation, or a wrong operation of one site, is seen
    bad_phi = phi.copy().reshape(-1, i.n_phi)
    k = int(op[np.flatnonzero(op != 0)[0]])
    bad_phi[k] = np.roll(bad_phi[k], 1)
    assert cf.symmetry_error(v_sym, v_full, rep, op, bad_phi.reshape(-1), i.n_theta, i.n_phi, n_xy) > 1e-3
    bad_op = op.copy()
    site = int(np.flatnonzero(op != 0)[0])
    bad_op[site] = 0
    assert cf.symmetry_error(v_sym, v_full, rep, bad_op, phi, i.n_theta, i.n_phi, n_xy) > 1e-3


# ---- the oracle's propagation steps ----------------------------------------------------------------------------

@pytest.mark.parametrize("name,params", [("dz_jmax24_16x16", cases.DZ_SETUP), ("hardwall_6d", cases.HARDWALL_6D)])
def test_kinetic_step_keeps_the_norm_and_half_steps_compose(name, params):
    """tests/test_propagation.cpp:99-118 on the oracle: a kinetic-only block is unitary, and two half blocks equal
    a full block."""
    run = _setup_run(params, name + "_kin")
    s0 = pseudo_random(run.scattering_size, 3)
    p0 = pseudo_random(run.specular_size, 4)
    run.set_wave(s0, p0)
    run.kinetic_and_potential(False, True)
    full, full_p = run.get_wave()
    run.set_wave(s0, p0)
    run.kinetic_and_potential(True, True)
    run.kinetic_and_potential(True, True)
    halves, halves_p = run.get_wave()
    run.close()
    norm = abs(np.vdot(full, full).real / np.vdot(s0, s0).real - 1.0)
    d, dp = relative_difference(full, halves), relative_difference(full_p, halves_p)
    print(f"{name}: norm {norm:.2e}, full vs halves {d:.2e} / {dp:.2e}")
    assert norm < 1e-13 and d < 1e-12 and dp < 1e-12
    halves[12345] *= 1 + 1e-10
    assert relative_difference(full, halves) > 1e-12


def test_free_gaussian_spreads_analytically_in_z():
    """tests/test_propagation.cpp:119-141 on the oracle: a Gaussian in Z, the same at every (r, irot, x, y), after
    100 kinetic-only time steps equals the free-particle solution (r and x, y act trivially on a constant). With
    a = 2 bohr^-2 at Z0 = 4 the Gaussian is below 1e-21 at both ends of the Z grid [-1, 9.08] and its spectrum below
    1e-20 at the grid's largest momentum, so the periodic grid represents it to rounding."""
    run = _setup_run(cases.SMOKE_DAILIGHT, "smoke_gauss")
    i = run.info
    z = run.array("grids.z")
    cell = run.array("derived.cell")
    mass, dt = cell[3], cell[6]
    n_steps = 100
    column = np.repeat(cf.free_gaussian(z, 0.0, mass, a=2.0, z0=4.0), i.nr)[:, None, None]
    run.set_wave(cf.host_wave(column, i.n_rot, i.nx * i.ny, i.nz * i.nr), np.zeros(run.specular_size, complex))
    for _ in range(n_steps):
        run.kinetic_and_potential(False, True)
    s, _ = run.get_wave()
    run.close()
    expected = cf.host_wave(np.repeat(cf.free_gaussian(z, n_steps * dt, mass, a=2.0, z0=4.0), i.nr)[:, None, None],
                            i.n_rot, i.nx * i.ny, i.nz * i.nr)
    err = float(np.abs(s - expected).max())
    print(f"free Gaussian after {n_steps} steps: max error {err:.2e}")
    assert err < 1e-12
    # a mass 1e-6 off is seen
    wrong = cf.free_gaussian(z, n_steps * dt, mass * (1 + 1e-6), a=2.0, z0=4.0)
    assert float(np.abs(np.repeat(wrong, i.nr) - expected.reshape(i.nz * i.nr, -1)[:, 0]).max()) > 1e-12


@pytest.mark.parametrize("name,params", [("dz_jmax24_16x16", cases.DZ_SETUP), ("hardwall_6d", cases.HARDWALL_6D)])
def test_parallel_plane_wave_gets_its_kinetic_phase(name, params):
    """tests/test_propagation.cpp:179-206 on the oracle: with a parallel momentum (EParallel 0.05 eV at 30 degrees)
    the plane wave exp(i (k_par + G) . R) is an eigenfunction of the (x, y) step (shift, FFT, kinetic factor, FFT,
    unshift) with eigenvalue exp(-i |k_par + G|^2 dt / 2M); here the eigenvalue is the free-particle formula on the
    lattice's reciprocal vectors, not the table the step uses."""
    run = _setup_run(params, name + "_pw", {"EParallel": 0.05, "PhiParallel": 30.0})
    i = run.info
    cell = run.array("derived.cell")
    mass, dt = cell[3], cell[6]
    assert np.hypot(cell[7], cell[8]) > 0.0
    for n, m in [(0, 0), (1, -1), (-2, 1)]:
        values, k2 = cf.plane_wave(cell, i.nx, i.ny, n, m)
        wave = cf.host_wave(values[None, :, None], i.n_rot, i.nx * i.ny, i.nz * i.nr)
        run.set_wave(wave, np.zeros(run.specular_size, complex))
        run.kinetic_and_potential(False, True)
        s, _ = run.get_wave()
        factor = np.exp(-1j * k2 / (2 * mass) * dt)
        err = float(np.abs(s - factor * wave).max())
        print(f"{name}: G = ({n}, {m}), |K|^2 {k2:.4f}, max error {err:.2e}")
        assert err < 1e-12
    run.close()
    wrong = np.exp(-1j * k2 * (1 + 1e-9) / (2 * mass) * dt)
    assert float(np.abs(s - wrong * wave).max()) > 1e-12


def test_hard_wall_potential_has_no_angular_coupling():
    """A potential without angular dependence (HardWallPES: V(Z, r) only) makes the interaction step's
    FBR -> DVR -> x exp(i A) -> DVR -> FBR a phase per (Z, r) point: a wave packet in one rotational state, away from
    the initial channel and with an empty specular grid, stays in that state through a full T block with V, with the
    same factor at every (x, y) site. The DZ surface, which does couple, moves part of it to other states."""
    for name, params, coupled in [("hardwall_6d", cases.HARDWALL_6D, False),
                                  ("dz_jmax24_16x16", cases.DZ_SETUP, True)]:
        run = _setup_run(params, name + "_rot")
        i = run.info
        n_xy, n_zr = i.nx * i.ny, i.nz * i.nr
        irot = 5 if i.initial_channel != 5 else 6
        rng = np.random.default_rng(7)
        column = np.zeros((n_zr, n_xy, i.n_rot), complex)
        column[:, :, irot] = (rng.standard_normal((n_zr, 1)) + 1j * rng.standard_normal((n_zr, 1)))
        wave = column.reshape(-1)
        run.set_wave(wave, np.zeros(run.specular_size, complex))
        run.kinetic_and_potential(False, False)
        s, _ = run.get_wave()
        run.close()
        s = s.reshape(n_zr, n_xy, i.n_rot)
        leak = float(np.abs(np.delete(s, irot, axis=2)).max() / np.abs(s[:, :, irot]).max())
        kept = s[:, :, irot]
        spread = float(np.abs(kept - kept[:, :1]).max() / np.abs(kept).max())
        print(f"{name}: state {irot}, leak into other states {leak:.2e}, site spread {spread:.2e}")
        if coupled:
            assert leak > 1e-3
        else:
            assert leak < 1e-13 and spread < 1e-13
