"""The HIP backend through whole runs of the reference's Fortran-gated cases: the averages of every dtAve, the BK
coefficients of every analysis step (1e-10 of the record scale), the analysis limits, the fixed specular line and the
fort.25 S-matrix rows, all against the golden fixtures (tests/test_regression_propagation.cpp of the reference runs
the same gates on its CPU and CUDA backends). The SCM case runs the CRP surface of the SRP48 data set plus the SCM
correction of the rescaled example slab, its potential grid built on the GPU. The complex64 storage mode runs the same cases and is held to the
reaction and channel probabilities of the FP64 run (1e-4 absolute)."""

import numpy as np
import pytest

from tests import cases, fixtures

pytestmark = pytest.mark.gpu

CASES = ["hardwall_short", "cu111_dailight_short", "cu111_dailight_short_sym", "cu111_dailight_oblique",
         "cu111_dailight_short_lowz", "scm_cu111_srp48_small"]

_fp64 = {}


@pytest.mark.parametrize("case", CASES)
def test_hip_run_passes_the_fortran_gates(case):
    import spodvr_amd
    print(f"case {case} fp64 full run", flush=True)
    run = spodvr_amd.Run(cases.golden_input(case), cases.golden_overrides(case) or None, device="hip")
    out = fixtures.full_run(run)
    report = fixtures.check_gates(run, out, case)
    print(f"  {report}", flush=True)
    _fp64[case] = out["results"]
    run.close()


@pytest.mark.parametrize("case", ["hardwall_short", "cu111_dailight_short_sym", "cu111_dailight_oblique",
                                  "scm_cu111_srp48_small"])
def test_mixed_precision_run_matches_fp64_probabilities(case):
    import spodvr_amd
    print(f"case {case} mixed full run", flush=True)
    if case not in _fp64:
        run = spodvr_amd.Run(cases.golden_input(case), cases.golden_overrides(case) or None, device="hip")
        _fp64[case] = fixtures.full_run(run)["results"]
        run.close()
    run = spodvr_amd.Run(cases.golden_input(case), {"Precision": "mixed", **cases.golden_overrides(case)},
                         device="hip")
    mixed = fixtures.full_run(run)["results"]
    ref = _fp64[case]
    dp = float(np.abs(mixed["p"] - ref["p"]).max())
    dr = float(np.abs(mixed["p_reaction"] - ref["p_reaction"]).max())
    print(f"  max |dP| {dp:.3e} max |dP_reaction| {dr:.3e}", flush=True)
    assert dp <= 1e-4 and dr <= 1e-4
    assert dp > 0.0  # the run really used complex64 storage
    run.close()
