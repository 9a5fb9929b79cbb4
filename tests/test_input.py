"""Input validation of the set-up the backend is fed from (src/input/validate.cpp:48-66 of the reference): grid
lengths that are not FFT sizes, nX /= nY under a square or hexagonal point group, and unsupported options are
refused before any device work, with the key named in the message."""

import pytest

import spodvr_amd
from tests import cases
from oracle import oracle


@pytest.mark.parametrize("key,value", [("nZ", 17), ("nZSpecular", 130), ("nR", 13), ("nX", 9), ("nY", 1)])
def test_invalid_fft_sizes_are_refused(key, value):
    params = dict(cases.SMOKE_DAILIGHT, **{key: value})
    with pytest.raises(spodvr_amd.SpoError, match=key):
        oracle.run(cases.write_case("invalid_" + key, params))


def test_unequal_xy_grids_are_refused_for_c6v():
    params = dict(cases.SMOKE_DAILIGHT, nX=4, nY=8)
    with pytest.raises(spodvr_amd.SpoError, match="nY"):
        oracle.run(cases.write_case("invalid_nxny", params))


def test_projection_operator_other_than_2_is_not_supported():
    params = dict(cases.SMOKE_DAILIGHT, ProjOp=1)
    with pytest.raises(spodvr_amd.SpoError):
        oracle.run(cases.write_case("invalid_projop", params))


def test_unknown_precision_is_refused():
    params = dict(cases.SMOKE_DAILIGHT, Precision="fp16")
    with pytest.raises(spodvr_amd.SpoError, match="Precision"):
        oracle.run(cases.write_case("invalid_precision", params))


# The restart and time keys (src/input/validate.cpp:90-103): each refusal names its key, status 2.
@pytest.mark.parametrize("overrides,key,message", [
    ({"tStart": 10.0, "tEnd": 90.0}, "tStart", "This is a new calculation but tStart /= 0."),
    ({"NewCalc": False, "RestartDir": "somewhere"}, "tStart", "This is a restart but tStart = 0."),
    ({"NewCalc": False, "tStart": 10.0, "tEnd": 90.0}, "RestartDir", "RestartDir is not given"),
    ({"tEnd": 81.0}, "tEnd", "tEnd /= tStart + nStep*dtStep."),
    ({"RestartFormat": "hdf5"}, "RestartFormat", "must be"),
    ({"CheckpointFormat": "all"}, "CheckpointFormat", "must be"),
])
def test_restart_and_time_keys_are_checked(overrides, key, message):
    params = dict(cases.SMOKE_DAILIGHT, **overrides)
    with pytest.raises(spodvr_amd.SpoError) as e:
        oracle.run(cases.write_case("invalid_restart_" + key, params))
    assert e.value.code == 2 and key in str(e.value) and message in str(e.value)


@pytest.mark.parametrize("spelling", ["h2cu111dailight", "H2CU111DAILIGHT", "H2Cu111Dailight"])
def test_system_id_is_matched_without_case(spelling):
    """The reference compares SystemIDs without case (src/input/json_input.cpp:419-420): any spelling runs, with
    the same potential as the canonical one."""
    canonical = oracle.run(cases.write_case("sysid_canonical", cases.SMOKE_DAILIGHT))
    other = oracle.run(cases.write_case("sysid_" + spelling, dict(cases.SMOKE_DAILIGHT, SystemID=spelling)))
    import numpy as np
    assert np.array_equal(canonical.array("potential.values"), other.array("potential.values"))
    canonical.close()
    other.close()
