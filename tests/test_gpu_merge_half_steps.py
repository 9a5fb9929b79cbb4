"""Merged half steps (SPO_HIP_MERGE_HALF_STEPS=1): the scattering wave's trailing T(dt/2) of one analysis step and
the leading T(dt/2) of the next run as one T(dt), and the analysis reads its Z row from a T(dt/2) of that row alone.
The results are the plain schedule's up to rounding: in FP64 the BK records, S-matrix sums, block densities and the
wave after six analysis steps are within 1e-12 of the largest value of the same steps without merging, and not
bitwise equal (the merged schedule really ran); in complex64 the merged run is as close to the FP64 run as the plain
one is. A whole gated HardWall run with merged half steps passes the Fortran gates, and two merged runs are bitwise
equal. Every test prints its case first."""

import os

import numpy as np
import pytest

from tests import cases, fixtures
from tests.util import pseudo_random, relative_difference

pytestmark = pytest.mark.gpu


def _case_input(case):
    if case == "dz_angular":
        return cases.write_case("dz_angular", cases.DZ_ANGULAR), {}
    return cases.golden_input(case), cases.golden_overrides(case)


def _run(case, precision, merge):
    import spodvr_amd
    path, overrides = _case_input(case)
    old = os.environ.get("SPO_HIP_MERGE_HALF_STEPS")
    os.environ["SPO_HIP_MERGE_HALF_STEPS"] = "1" if merge else "0"
    try:
        return spodvr_amd.Run(path, dict(overrides, Precision=precision), device="hip")
    finally:
        if old is None:
            os.environ.pop("SPO_HIP_MERGE_HALF_STEPS", None)
        else:
            os.environ["SPO_HIP_MERGE_HALF_STEPS"] = old


def _steps_from_seeded_wave(case, precision, merge, steps, seed=71):
    run = _run(case, precision, merge)
    try:
        run.set_wave(pseudo_random(run.scattering_size, seed), pseudo_random(run.specular_size, seed + 1))
        run.steps(steps)
        out = {"bk": run.bk_records(), "sums": run.s_sums()}
        out["densities"] = run.block_densities()
        out["scattering"], out["specular"] = run.get_wave()
        return out
    finally:
        run.close()


@pytest.mark.parametrize("case", ["hardwall_short", "cu111_dailight_short_sym", "dz_angular"])
def test_fp64_merged_half_steps_match_the_plain_schedule(case):
    print(f"case {case} seed 71 fp64: 6 analysis steps, plain then merged", flush=True)
    plain = _steps_from_seeded_wave(case, "fp64", False, 6)
    merged = _steps_from_seeded_wave(case, "fp64", True, 6)
    diffs = {k: relative_difference(merged[k], plain[k]) for k in plain}
    print("  " + " ".join(f"{k} {v:.2e}" for k, v in diffs.items()), flush=True)
    assert all(v < 1e-12 for v in diffs.values()), diffs
    assert diffs["bk"] > 0.0, diffs  # the merged schedule ran (the same rounding would mean it did not)


@pytest.mark.parametrize("case", ["cu111_dailight_oblique", "dz_angular"])
def test_complex64_merged_half_steps_keep_the_accuracy(case):
    """In complex64 the merged and the plain schedule differ by their roundings (measured ~1e-6 after six analysis
    steps of a seeded wave that fills the grid): both are held to the FP64 plain run, the merged one within 1.5x
    the plain one's error (plus 1e-7 of the largest value)."""
    print(f"case {case} seed 71: 6 analysis steps, FP64 plain, complex64 plain and merged", flush=True)
    ref = _steps_from_seeded_wave(case, "fp64", False, 6)
    plain = _steps_from_seeded_wave(case, "mixed", False, 6)
    merged = _steps_from_seeded_wave(case, "mixed", True, 6)
    for k in ref:
        e_plain = relative_difference(plain[k], ref[k])
        e_merged = relative_difference(merged[k], ref[k])
        print(f"  {k}: plain {e_plain:.2e} merged {e_merged:.2e} "
              f"(merged vs plain {relative_difference(merged[k], plain[k]):.2e})", flush=True)
        assert e_merged <= 1.5 * e_plain + 1e-7, (k, e_plain, e_merged)
    assert relative_difference(merged["bk"], plain["bk"]) > 0.0


def test_merged_half_steps_pass_the_fortran_gates():
    case = "hardwall_short"
    print(f"case {case} fp64 full run with merged half steps", flush=True)
    run = _run(case, "fp64", True)
    try:
        out = fixtures.full_run(run)
        report = fixtures.check_gates(run, out, case)
        print(f"  {report}", flush=True)
    finally:
        run.close()


def test_merged_runs_are_bitwise_identical():
    print("case cu111_dailight_short_sym seed 71 fp64: two merged runs of 4 analysis steps", flush=True)
    a = _steps_from_seeded_wave("cu111_dailight_short_sym", "fp64", True, 4)
    b = _steps_from_seeded_wave("cu111_dailight_short_sym", "fp64", True, 4)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_switch_values_other_than_0_and_1_are_refused():
    import spodvr_amd
    print("case hardwall_short: SPO_HIP_MERGE_HALF_STEPS=2 at run creation", flush=True)
    path, overrides = _case_input("hardwall_short")
    old = os.environ.get("SPO_HIP_MERGE_HALF_STEPS")
    os.environ["SPO_HIP_MERGE_HALF_STEPS"] = "2"
    try:
        with pytest.raises(spodvr_amd.SpoError, match="SPO_HIP_MERGE_HALF_STEPS must be 0 or 1"):
            spodvr_amd.Run(path, overrides or None, device="hip")
    finally:
        if old is None:
            os.environ.pop("SPO_HIP_MERGE_HALF_STEPS", None)
        else:
            os.environ["SPO_HIP_MERGE_HALF_STEPS"] = old
