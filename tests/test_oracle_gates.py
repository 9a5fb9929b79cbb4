"""The CPU oracle against the reference's Fortran-gated regression cases (gates 3 and 4 of its validation:
tests/test_regression_propagation.cpp and tests/test_regression_hardwall.cpp of the reference): the averages of
every dtAve, the BK coefficients of every analysis step, the analysis limits, the fixed specular line and the
on-the-fly S-matrix rows of fort.25. This pins the oracle before it is trusted as the checker of the HIP backend.

hardwall_short runs in the default CPU suite (about 50 s on 8 threads); the four Dai-Light cases take about three
minutes each and run with SPODVR_SLOW=1."""

import os

import numpy as np
import pytest

from tests import cases, fixtures
from oracle import oracle

SLOW = os.environ.get("SPODVR_SLOW") == "1"
DAILIGHT = ["cu111_dailight_short", "cu111_dailight_short_sym", "cu111_dailight_oblique",
            "cu111_dailight_short_lowz"]


@pytest.fixture(scope="module")
def hardwall():
    run = oracle.run(cases.golden_input("hardwall_short"))
    out = fixtures.full_run(run)
    yield run, out
    run.close()


def test_hardwall_short_passes_the_fortran_gates(hardwall):
    run, out = hardwall
    report = fixtures.check_gates(run, out, "hardwall_short")
    assert report["fort25_rows"] > 1000


def test_gates_fail_on_a_planted_difference(hardwall):
    """The comparisons are not vacuous: a difference of 1e-9 of the BK gate scale in one coefficient, one printed unit
    in one average, or 1e-8 in one S-matrix element fails them."""
    run, out = hardwall
    f = fixtures.load_bk("hardwall_short")
    bk = out["bk"].copy()
    k = int(f["indices"][int(np.argmax(np.abs(f["records"]).max(axis=0)))])
    bk[bk.shape[0] // 2, k] += 1e-9 * max(float(f["scale"]), 1e-4)
    diff, scale = fixtures.compare_bk(bk, "hardwall_short")
    assert diff > 1e-10 * scale
    averages = out["averages"].copy()
    text = fixtures.read_averages("hardwall_short")[3][2]
    averages[3, 2] += 2.0 * fixtures.half_unit(text)
    assert fixtures.compare_averages(averages, "hardwall_short") > 1.0
    results = {key: (v.copy() if isinstance(v, np.ndarray) else v) for key, v in out["results"].items()}
    f25 = np.load(os.path.join(fixtures.case_dir("hardwall_short"), "fort25.npz"))
    block, v, j, mj, od, n, m = (int(x) for x in f25["rows"][0])
    state = [(int(a), int(b), int(c), int(d)) for a, b, c, d in zip(
        run.array("state.j"), run.array("state.mj"), run.array("state.n"), run.array("state.m"))].index((j, mj, n, m))
    results["s"][block, state + run.info.n_states * v] += 1e-8
    assert fixtures.compare_fort25(run, results, "hardwall_short")["failed"] >= 1


@pytest.mark.slow
@pytest.mark.skipif(not SLOW, reason="minutes of CPU per case; run with SPODVR_SLOW=1")
@pytest.mark.parametrize("case", DAILIGHT)
def test_dailight_cases_pass_the_fortran_gates(case):
    run = oracle.run(cases.golden_input(case))
    report = fixtures.check_gates(run, fixtures.full_run(run), case)
    assert report["fort25_phase_compared"] > 0
    run.close()
