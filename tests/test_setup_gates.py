"""The set-up against the reference's fixtures (gates 1 and 2 of its validation): the PES on the DVR grid through
the symmetry maps (pes_reference.txt: 1e-13 of max(|a|, |b|, 1 eV) there, tests/pes_gate.hpp:24-28), the rovibrational states of the Fourier grid
Hamiltonian (debug_asyp.dat, energies to 1e-12 there) and the analysis limits the Fortran prints."""

import numpy as np
import pytest

from tests import cases, fixtures
import spodvr_amd
from oracle import oracle


@pytest.fixture(scope="module", params=["cu111_dailight_short_sym", "cu111_dailight_oblique"])
def setup_run(request):
    run = oracle.run(cases.golden_input(request.param))
    yield request.param, run
    run.close()


def test_potential_grid_matches_fortran(setup_run):
    case, run = setup_run
    f = np.load(fixtures.case_dir(case) + "/pes.npz")
    nzpot, nr, nx, ny, nth, nph = (int(x) for x in f["dims"])
    info = run.info
    assert (info.nz_potential, info.nr, info.nx, info.ny, info.n_theta, info.n_phi) == (nzpot, nr, nx, ny, nth, nph)
    np.testing.assert_allclose(run.array("angular.thetas"), f["thetas"], rtol=0, atol=1e-15)
    np.testing.assert_allclose(run.array("angular.phis"), f["phis"], rtol=0, atol=1e-15)
    assert run.info.n_representatives == int(f["symmetry"][1]) or not f["symmetry"][0]
    values = run.array("potential.values")
    rep = run.array("symmetry.representative_of").astype(int)
    op = run.array("symmetry.operation_of").astype(int)
    phi_index = run.array("symmetry.phi_index").astype(int).reshape(-1, nph)
    n_tp, n_rep = nth * nph, info.n_representatives
    floor = 1.0 / 27.2114  # 1 eV in hartree: near V = 0 the LEPS terms cancel and a pure ratio means nothing
    worst = 0.0
    for (it, ip, ix, iy, iz, ir), v in zip(f["index"], f["value"]):
        ixy = iy + ny * ix
        izr = ir + nr * iz
        mine = values[phi_index[op[ixy], ip] + nph * it + n_tp * (rep[ixy] + n_rep * izr)]
        worst = max(worst, abs(mine - v) / max(abs(v), abs(mine), floor))
    assert worst <= 1e-13, worst


def test_rovibrational_states_match_fortran(setup_run):
    case, run = setup_run
    f = np.load(fixtures.case_dir(case) + "/gas.npz")
    energies = run.array("gas.energies")
    wf = run.array("gas.wave_functions")
    nr = run.info.nr
    jmax1 = len(energies) // nr
    for v, j, e, vec in zip(f["v"], f["j"], f["energy"], f["vectors"]):
        k = j + jmax1 * v
        assert abs(energies[k] - e) <= 1e-12, (v, j, energies[k], e)
        if v < 3:  # bound states; the eigenvectors of the continuum levels are not unique to rounding
            np.testing.assert_allclose(wf[nr * k:nr * (k + 1)], vec, rtol=0, atol=1e-8)


@pytest.mark.parametrize("case", ["hardwall_short", "cu111_dailight_short", "cu111_dailight_short_sym",
                                  "cu111_dailight_oblique", "cu111_dailight_short_lowz"])
def test_analysis_limits_match_fortran(case):
    run = oracle.run(cases.golden_input(case))
    stdout = fixtures.load_stdout(case)
    for name, mine in zip(["jana", "nviba", "nana", "nstata", "inchna"], run.array("analysis.limits")):
        assert stdout["limits"][name] == [int(mine)], (case, name)
    f = np.load(fixtures.case_dir(case) + "/fort25.npz")
    te = run.array("energies.total")
    for b in f["blocks"]:
        text = str(f["total_energy"][b])
        assert abs(te[b] - fixtures.fortran_float(text)) <= fixtures.half_unit(text)


def test_pes_values_are_the_reference_formulas():
    """Spot values of the two potentials at closed-form points: the hard wall is the H2 Rydberg V(r) beyond
    Z = 7 (and V(re) = 0 there); Dai-Light is the H2 LEPS curve U_M + A_M at and beyond zasymp = 15."""
    assert spodvr_amd.pes_value("HardWallPES", 0.0, 0.0, 8.0, 1.40, 1.0, 0.0) == pytest.approx(0.0, abs=1e-14)
    d, a1, a2, a3, a4, re = 4.8286, 2.282, 1.555, 0.7533, 2.23, 1.40
    r = 2.0 - re
    vr = -d * (1 + a1 * r + a2 * r ** 2 + a3 * r ** 3) * np.exp(-a4 * r) + d
    assert spodvr_amd.pes_value("HardWallPES", 1.0, 2.0, 9.0, 2.0, 0.3, 0.7) == pytest.approx(vr, rel=1e-14)
    # Dai-Light beyond zasymp: the molecular LEPS term only, independent of (x, y, theta, phi).
    v1 = spodvr_amd.pes_value("H2Cu111DaiLight", 0.0, 0.0, 16.0, 1.4, 0.5, 0.0)
    v2 = spodvr_amd.pes_value("H2Cu111DaiLight", 1.3, 2.1, 20.0, 1.4, 1.2, 2.0)
    assert v1 == v2
    # At r = r_M the Coulomb and exchange integrals are D/4/(1+delta) (1 - 4 delta) and D/4/(1+delta) (-5 - 2 delta)
    dm, delta = 0.1744, -0.15
    um = dm / 4 / (1 + delta) * ((3 + delta) - (2 + 6 * delta))
    am = dm / 4 / (1 + delta) * ((1 + 3 * delta) - (6 + 2 * delta))
    assert v1 == pytest.approx((um + am) * 27.2114, rel=1e-13)
