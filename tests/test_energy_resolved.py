"""Energy-resolved sums at chosen sites (EnergyResolvedSites / EnergyResolvedEnergies, src/analysis/energy_resolved.cpp
of the reference): the oracle's accumulator against the sum written out by hand from the site columns and the
specular wave after every analysis step (the check of tests/test_energy_resolved.cpp:93 of the reference), and the
input checks of src/input/validate_outputs.cpp:87-110."""

import numpy as np
import pytest

import spodvr_amd
from tests import cases
from tests.util import pseudo_random
from oracle import oracle

HARTREE_EV = 27.2114  # the reference's legacy energy unit (src/core/unit_conversion.hpp)

SITES = {"PlotSites": [{"name": "top", "x": 0.0, "y": 0.0}, {"name": "hollow", "x": 0.25, "y": 0.5}],
         "EnergyResolvedSites": ["hollow", "top"], "EnergyResolvedEnergies": [0.3, 0.45]}


def test_oracle_sums_match_the_hand_sum():
    run = oracle.run(cases.write_case("er_smoke", cases.SMOKE_DAILIGHT), SITES)
    i = run.info
    run.set_wave(pseudo_random(run.scattering_size, 61), pseudo_random(run.specular_size, 62))
    e0 = run.array("gas.energies")[0]  # E(v = 0, j = 0), the initial state
    totals = [(e + HARTREE_EV * e0) / HARTREE_EV for e in SITES["EnergyResolvedEnergies"]]
    p = cases.SMOKE_DAILIGHT
    dt_ana = i.steps_per_analysis * p["dtStep"]
    ixy = [2 + 4 * 1, 0]  # hollow: x 0.25 -> ix 1, y 0.5 -> iy 2 (ixy = iy + ny ix); top: 0
    expect = np.zeros((2, 2, i.nz * i.nr, i.n_rot), complex)
    expect_sp = np.zeros((2, i.nz_specular * i.nr), complex)
    for k in range(1, 4):
        run.steps(1)
        t = p["tEnd"] + (k * i.steps_per_analysis - p["nStep"]) * p["dtStep"]
        for s, site in enumerate(ixy):
            col = run.site_column(site).reshape(i.n_rot, i.nz * i.nr).T  # [izr, irot]
            for e, tot in enumerate(totals):
                expect[s, e] += col * np.exp(1j * tot * t) * dt_ana
        spec = run.specular()
        for e, tot in enumerate(totals):
            expect_sp[e] += spec * np.exp(1j * tot * t) * dt_ana
    sums, spec = run.energy_resolved()
    assert sums.shape == (2, 2, i.nz, i.nr, i.n_rot) and spec.shape == (2, i.nz_specular, i.nr)
    scale = np.abs(expect).max()
    assert np.abs(sums.reshape(expect.shape) - expect).max() <= 1e-13 * scale
    assert np.abs(spec.reshape(expect_sp.shape) - expect_sp).max() <= 1e-13 * np.abs(expect_sp).max()
    assert np.abs(sums).max() > 0 and np.abs(spec).max() > 0
    run.close()


def test_no_sums_without_the_keys():
    run = oracle.run(cases.write_case("er_none", cases.SMOKE_DAILIGHT))
    assert run.energy_resolved() is None
    run.close()


@pytest.mark.parametrize("change,key", [
    ({"EnergyResolvedEnergies": []}, "EnergyResolved"),
    ({"EnergyResolvedSites": ["bridge"]}, "EnergyResolvedSites"),
    ({"EnergyResolvedSites": ["top", "top"]}, "EnergyResolvedSites"),
    ({"PlotSites": [{"name": "top", "x": 0.1, "y": 0.0}]}, "PlotSites"),
    ({"PlotSites": [{"name": "top", "x": 1.0, "y": 0.0}]}, "PlotSites"),
    ({"EnergyResolvedEnergies": [0.1]}, "EnergyResolvedEnergies"),
    ({"EnergyResolvedEnergies": [0.7]}, "EnergyResolvedEnergies"),
])
def test_invalid_energy_resolved_keys_are_refused(change, key):
    params = dict(cases.SMOKE_DAILIGHT, **SITES)
    params.update(change)
    with pytest.raises(spodvr_amd.SpoError, match=key):
        oracle.run(cases.write_case("er_invalid", params))
