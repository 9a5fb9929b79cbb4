"""Native restart files on the HIP backend: a split run is bitwise the unsplit one (FP64 and mixed precision), and
a run checkpointed on the CPU oracle continues on the GPU to the oracle's unsplit result (the files do not depend
on the backend that wrote them).

Every test prints its case before touching the GPU."""

import sys

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu

FIRST = {"tEnd": 40.0, "nStep": 20}
SECOND = {"NewCalc": False, "tStart": 40.0, "tEnd": 80.0, "nStep": 20}


def _note(*what):
    print("checkpoint:", *what, file=sys.stderr, flush=True)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_split_run_on_the_gpu_is_bitwise_the_unsplit_run(tmp_path, precision):
    import spodvr_amd
    _note("split run, SMOKE_DAILIGHT 40 steps at 20,", precision)
    path = cases.write_case("restart_smoke", cases.SMOKE_DAILIGHT, str(tmp_path))
    full = spodvr_amd.Run(path, {"Precision": precision}, device="hip")
    full.steps(full.info.analysis_steps_total)
    wave, sums, bk, res = full.get_wave(), full.s_sums(), full.bk_records(), full.results()
    full.close()
    first = spodvr_amd.Run(path, dict(FIRST, Precision=precision), device="hip")
    first.steps(first.info.analysis_steps_total)
    first.write_checkpoint(str(tmp_path / "first"))
    first.close()
    second = spodvr_amd.Run(path, dict(SECOND, Precision=precision, RestartDir=str(tmp_path / "first")), device="hip")
    second.steps(second.info.analysis_steps_total)
    sg, pg = second.get_wave()
    assert np.array_equal(sg, wave[0]) and np.array_equal(pg, wave[1])
    assert np.array_equal(second.s_sums(), sums)
    assert np.array_equal(second.bk_records(), bk)
    r2 = second.results()
    for key in ("s", "p", "p_reaction"):
        assert np.array_equal(r2[key], res[key]), key
    second.close()


def test_oracle_checkpoint_continues_on_the_gpu(tmp_path):
    import os
    import spodvr_amd
    from oracle import oracle
    _note("oracle checkpoint at 20 steps, continued on the GPU, SMOKE_DAILIGHT")
    path = cases.write_case("restart_smoke", cases.SMOKE_DAILIGHT, str(tmp_path))
    full = oracle.run(path)
    full.steps(full.info.analysis_steps_total)
    wave, bk = full.get_wave(), full.bk_records()
    full.close()
    first = oracle.run(path, FIRST)
    first.steps(first.info.analysis_steps_total)
    first.write_checkpoint(str(tmp_path / "first"))
    first.close()
    os.environ["SPO_HIP_POTENTIAL"] = "host"  # the oracle's potential grid, to compare at 1e-12
    try:
        second = spodvr_amd.Run(path, dict(SECOND, RestartDir=str(tmp_path / "first")), device="hip")
    finally:
        os.environ.pop("SPO_HIP_POTENTIAL", None)
    second.steps(second.info.analysis_steps_total)
    sg, pg = second.get_wave()
    # against the whole wave's largest value: the scattering part of this short run is ~1e-11 of the specular
    # part, and the two backends agree to rounding of the latter
    scale = max(np.abs(wave[0]).max(), np.abs(wave[1]).max())
    assert max(np.abs(sg - wave[0]).max(), np.abs(pg - wave[1]).max()) / scale < 1e-12
    b2 = second.bk_records()
    n_first = len(b2) // 2
    assert np.array_equal(b2[:n_first], bk[:n_first])  # the earlier records as written
    # the BK gate's measure: 1e-10 max(max |c|, 1e-4) (tests/regression_gate.hpp:19-45)
    assert np.abs(b2[n_first:] - bk[n_first:]).max() <= 1e-10 * max(np.abs(bk).max(), 1e-4)
    second.close()


def test_split_run_through_fortran_files_on_the_gpu(tmp_path):
    import spodvr_amd
    _note("split run through fort.9 / fort.21, SMOKE_DAILIGHT FP64")
    path = cases.write_case("restart_smoke", cases.SMOKE_DAILIGHT, str(tmp_path))
    full = spodvr_amd.Run(path, device="hip")
    full.steps(full.info.analysis_steps_total)
    wave, sums = full.get_wave(), full.s_sums()
    full.close()
    first = spodvr_amd.Run(path, dict(FIRST, CheckpointFormat="fortran"), device="hip")
    first.steps(first.info.analysis_steps_total)
    first.write_checkpoint(str(tmp_path / "first"))
    first.close()
    second = spodvr_amd.Run(path, dict(SECOND, RestartDir=str(tmp_path / "first"), RestartFormat="fortran"),
                            device="hip")
    second.steps(second.info.analysis_steps_total)
    sg, pg = second.get_wave()
    assert np.array_equal(sg, wave[0]) and np.array_equal(pg, wave[1])
    assert np.array_equal(second.s_sums(), sums)
    second.close()
