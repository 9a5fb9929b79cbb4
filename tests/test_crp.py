"""The CRP potential (SystemID CRPlibrary; host/crp.cpp) against the reference's Fortran CRP driver outputs
(tests/crplib/reference/<case>/, extracted bit for bit by tests/golden/make_crp_fixtures.py), against the symmetry
the surface must have, and through the propagation set-up against the Fortran program's potential grid
(crp_cu111_srp48_small/pes_reference.txt, gate 1) and gas-phase states (debug_asyp.dat, gate 2).

Tolerances (tests/crplib/test_molecule_surface.cpp:1-12, :35-62):
  * data sets whose cuts hold the interpolation function (CutsInterpolation T; the impulse cases): bitwise;
  * data sets of total energies (CutsInterpolation F): the Fortran reduces the cuts with a vectorised cosine, a few
    ulp from the scalar one; results agree to 2e-14 of max(unit, |value|) (the reference measures up to 1.9e-14 on
    the SRP48 data set);
  * atom-surface values: bitwise; their gradients to 1e-15 of max(1, |value|): the sines of the Fourier basis enter
    only the derivatives, and this machine's libm rounds a few of them an ulp away from the one that made the
    fixtures (the reference compares them bitwise on its own machine).
"""

import json
import math
import os

import numpy as np
import pytest

import spodvr_amd
from tests import cases, fixtures

CRP_DIR = os.path.join(cases.GOLDEN_DIR, "crp")
CASES = ["cu111_srp48", "c3v_pairwise", "c6v_pairwise", "c3v_impulse_15", "c6v_impulse_5", "c4v_pairwise",
         "c4v_impulse_6", "c4v_negative_spacing"]
BOHR_PER_ANGSTROM = 1.0 / 0.529177249  # the CRP library's factor (src/crplib/units.cpp:13-18)
EV_PER_HARTREE = 27.2113961


def data_set(case):
    return os.path.join(CRP_DIR, case, "dataset")


def cuts_interpolation(case):
    return open(os.path.join(data_set(case), "CRPData.6D")).read().split()[1].upper().startswith(("T", ".T"))


@pytest.fixture(scope="module", params=CASES)
def crp_case(request):
    f = dict(np.load(os.path.join(CRP_DIR, request.param, "points.npz")))
    surface = spodvr_amd.CrpSurface(data_set(request.param))
    yield request.param, f, surface
    surface.close()


def assert_fortran(mine, ref, bitwise, unit=1.0, rel=2e-14):
    mine, ref = np.asarray(mine, dtype=np.float64), np.asarray(ref, dtype=np.float64)
    if bitwise:
        bad = np.flatnonzero(mine.view(np.uint64) != ref.view(np.uint64))
        assert bad.size == 0, (bad[:5], mine.ravel()[bad[:5]], ref.ravel()[bad[:5]])
    else:
        err = np.abs(mine - ref) / np.maximum(unit, np.abs(ref))
        assert err.max() <= rel, err.max()


def test_parameters(crp_case):
    case, f, s = crp_case
    p = s.parameters()
    assert p["point_group"] == str(f["point_group"])
    lattice, lattice_b, z_asy, skew = f["wrapper_params"]
    assert p["lattice_constant"] == lattice == lattice_b
    assert p["z_asymptotic"] == z_asy
    assert p["skewing_angle"] == skew
    if "zso" in f:
        assert (p["zso_min"], p["zso_max"]) == tuple(f["zso"])


def test_atom_surface(crp_case):
    case, f, s = crp_case
    if "atom_in" not in f:
        pytest.skip("no atom-surface points in this case")
    out = s.atom(f["atom_in"])
    assert_fortran(out[:, 0], f["atom_out"][:, 0], bitwise=True)
    assert_fortran(out[:, 1:], f["atom_out"][:, 1:], bitwise=False, rel=1e-15)


def test_molecule_surface(crp_case):
    case, f, s = crp_case
    atoms = f["molecule_in"][:, [1, 2, 0, 4, 5, 3]]  # the driver's order is (Za, Xa, Ya, Zb, Xb, Yb)
    out = s.molecule(atoms)
    assert_fortran(out, f["molecule_out"], cuts_interpolation(case))


def test_molecule_surface_in_angstrom_and_ev(crp_case):
    case, f, s = crp_case
    if "molecule13_in" not in f:
        pytest.skip("no angstrom / eV points in this case")
    out = s.molecule(f["molecule13_in"][:, [1, 2, 0, 4, 5, 3]], spodvr_amd.CRP_ANGSTROM, spodvr_amd.CRP_EV)
    unit = np.array([EV_PER_HARTREE] + [EV_PER_HARTREE * (1 / BOHR_PER_ANGSTROM) ** -1] * 6)
    assert_fortran(out, f["molecule13_out"], cuts_interpolation(case), unit=unit)


def test_qd_wrapper(crp_case):
    case, f, s = crp_case
    assert_fortran(s.qd(f["wrapper_in"]), f["wrapper_out"], cuts_interpolation(case), unit=EV_PER_HARTREE)


def test_gas_phase_and_switch(crp_case):
    case, f, s = crp_case
    if "gasphase_in" not in f:
        pytest.skip("no gas-phase points in this case")
    assert_fortran(s.gas_phase(f["gasphase_in"]), f["gasphase_out"], bitwise=True)


def test_molecule_geometry(crp_case):
    case, f, s = crp_case
    if "cambcord_in" not in f:
        pytest.skip("no geometry points in this case")
    assert_fortran(spodvr_amd.CrpSurface.geometry(f["cambcord_in"]), f["cambcord_out"], bitwise=True)


# ---------------------------------------------------------------------------------------------------------------
# Symmetry of the surface: the C3v (C6v) point group about a top site, the lattice period and the exchange of the
# two identical atoms leave V unchanged. A wrong site, table column or rotated azimuth breaks one of these.

def _skewed(x, y, gamma=math.pi / 3):
    v = y / math.sin(gamma)
    return x - v * math.cos(gamma), v


def _cartesian(u, v, gamma=math.pi / 3):
    return u + v * math.cos(gamma), v * math.sin(gamma)


@pytest.mark.parametrize("case", ["cu111_srp48", "c6v_pairwise", "c4v_pairwise"])
def test_surface_symmetry(case):
    s = spodvr_amd.CrpSurface(data_set(case))
    a = s.parameters()["lattice_constant"]
    square = s.parameters()["point_group"] == "C4v"
    gamma = math.pi / 2 if square else math.pi / 3
    turn = math.pi / 2 if square else 2 * math.pi / 3  # the smallest rotation of the point group's C3 / C4 axis
    rng = np.random.default_rng(7)
    base, images = [], []
    for _ in range(40):
        u, v = rng.uniform(0, a, 2)
        z, r = rng.uniform(1.5, 4.0), rng.uniform(1.0, 2.0)
        theta, phi = rng.uniform(0.1, math.pi - 0.1), rng.uniform(0, 2 * math.pi)
        x, y = _cartesian(u, v, gamma)
        point = (u, v, z, r, theta, phi)
        c, sn = math.cos(turn), math.sin(turn)
        rotated = (*_skewed(c * x - sn * y, sn * x + c * y, gamma), z, r, theta, phi + turn)
        mirrored = (*_skewed(-x, y, gamma), z, r, theta, math.pi - phi)  # the plane x = 0 (hcp and fcc for C3v)
        shifted = (u + a, v - a, z, r, theta, phi)
        exchanged = (u, v, z, r, math.pi - theta, phi + math.pi)
        for image in (rotated, mirrored, shifted, exchanged):
            base.append(point)
            images.append(image)
    v0, v1 = s.qd(np.array(base)), s.qd(np.array(images))
    np.testing.assert_allclose(v1, v0, rtol=1e-10, atol=1e-10)
    if case == "c6v_pairwise":  # C6v: also the rotation by 60 degrees
        pts = np.array(base[::4])
        rot = []
        for (u, v, z, r, theta, phi) in pts:
            x, y = _cartesian(u, v)
            c, sn = math.cos(math.pi / 3), math.sin(math.pi / 3)
            rot.append((*_skewed(c * x - sn * y, sn * x + c * y), z, r, theta, phi + math.pi / 3))
        np.testing.assert_allclose(s.qd(np.array(rot)), s.qd(pts), rtol=1e-10, atol=1e-10)


def test_unusable_data_set_is_an_input_error(tmp_path):
    with pytest.raises(spodvr_amd.SpoError) as e:
        spodvr_amd.CrpSurface(str(tmp_path / "missing"))
    assert e.value.code == 2 and "CRPData" in str(e.value)


# ---------------------------------------------------------------------------------------------------------------
# Through the propagation set-up: SystemID CRPlibrary with the SRP48 data set on the Fortran program's grid.

@pytest.fixture(scope="module")
def crp_setup():
    from oracle import oracle
    run = oracle.run(cases.golden_input("crp_cu111_srp48_small"), {"CRPlibDataSet": data_set("cu111_srp48")})
    yield run
    run.close()


def test_crp_potential_grid_matches_fortran(crp_setup):
    run = crp_setup
    f = np.load(fixtures.case_dir("crp_cu111_srp48_small") + "/pes.npz")
    nzpot, nr, nx, ny, nth, nph = (int(x) for x in f["dims"])
    info = run.info
    assert (info.nz_potential, info.nr, info.nx, info.ny, info.n_theta, info.n_phi) == (nzpot, nr, nx, ny, nth, nph)
    np.testing.assert_allclose(run.array("angular.thetas"), f["thetas"], rtol=0, atol=1e-15)
    values = run.array("potential.values")
    rep = run.array("symmetry.representative_of").astype(int)
    op = run.array("symmetry.operation_of").astype(int)
    phi_index = run.array("symmetry.phi_index").astype(int).reshape(-1, nph)
    n_tp, n_rep = nth * nph, info.n_representatives
    # gate 1 relative to max(|a|, |b|, 1 eV) (tests/pes_gate.hpp:24-28); 3e-13 for this data set of total energies,
    # where the cut reduction's cosine differs from the Fortran's vectorised one (tests/test_regression_pes.cpp:70-71,
    # :97-103: up to 1.5e-13 there)
    floor = 1.0 / 27.2114
    # The Fortran writes its values before ImposeVMax clips them; the gate clips the reference values
    # (tests/pes_gate.hpp:44-49). VMax in eV, 27.2114 eV per hartree (src/core/unit_conversion.hpp).
    v_max = json.load(open(cases.golden_input("crp_cu111_srp48_small")))["VMax"] / 27.2114
    worst, clipped = 0.0, 0
    for (it, ip, ix, iy, iz, ir), v in zip(f["index"], f["value"]):
        ixy = iy + ny * ix
        mine = values[phi_index[op[ixy], ip] + nph * it + n_tp * (rep[ixy] + n_rep * (ir + nr * iz))]
        clipped += v > v_max
        v = min(v, v_max)
        worst = max(worst, abs(mine - v) / max(abs(v), abs(mine), floor))
    assert worst <= 3e-13, worst
    assert 0 < clipped < len(f["value"])  # both branches of the clip are covered


def test_crp_rovibrational_states_match_fortran(crp_setup):
    run = crp_setup
    f = np.load(fixtures.case_dir("crp_cu111_srp48_small") + "/gas.npz")
    energies = run.array("gas.energies")
    jmax1 = len(energies) // run.info.nr
    for v, j, e in zip(f["v"], f["j"], f["energy"]):  # tests/test_regression_pes.cpp:205-223
        mine = energies[j + jmax1 * v]
        assert abs(mine - e) / max(abs(e), abs(mine)) <= (1e-12 if v < 3 else 1e-10), (v, j, mine, e)
