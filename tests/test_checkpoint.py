"""Native restart files (checkpoint.bin, bk_coefficients.bin) on the CPU oracle: a split run is bitwise the
unsplit one (the reference's tests/test_restart.cpp:241), the files carry format version 2 with the reference's
35-entry parameter list (:571-611), and mismatched or damaged files are refused with the reference's messages
(:295, :332, :923-1040)."""

import os
import shutil
import struct

import numpy as np
import pytest

import spodvr_amd
from oracle import oracle
from tests import cases

# The parameter names of format version 2 without low-Z absorbers, in file order (tests/test_restart.cpp:574-609).
FORMAT_2 = ["ZStart", "dZ", "nZ", "nZSpecular", "RStart", "dR", "nR", "nX", "nY", "JMax", "MJMax",
            "number of phi points (MJMax, nPhisOptimal)", "Mass1", "Mass2", "HomoNuc", "dtStep", "dtAna", "Zana",
            "Z0", "EMinWP", "EMaxWP", "EParallel", "PhiParallel", "VIn", "JIn", "MJIn", "ZOptStart", "ZOptEnd",
            "ROptStart", "ROptEnd", "ZSpOptStart", "ZSpOptEnd", "ZOptFactor (effective)", "ROptFactor (effective)",
            "ZSpOptFactor (effective)"]

FIRST = {"tEnd": 40.0, "nStep": 20}
SECOND = {"NewCalc": False, "tStart": 40.0, "tEnd": 80.0, "nStep": 20}


@pytest.fixture(scope="module")
def split(tmp_path_factory):
    """The unsplit 40-step run, and its first 20 steps with their restart files."""
    base = tmp_path_factory.mktemp("restart")
    path = cases.write_case("restart_smoke", cases.SMOKE_DAILIGHT, str(base))
    full = oracle.run(path)
    full.steps(full.info.analysis_steps_total)
    whole = {"wave": full.get_wave(), "sums": full.s_sums(), "bk": full.bk_records(), "results": full.results()}
    full.close()
    first = oracle.run(path, FIRST)
    first.steps(first.info.analysis_steps_total)
    first_dir = base / "first"
    first.write_checkpoint(str(first_dir))
    first.close()
    return path, first_dir, whole


def _restart(path, directory, **extra):
    return oracle.run(path, dict(SECOND, RestartDir=str(directory), **extra))


def _error(path, directory, **extra):
    with pytest.raises(spodvr_amd.SpoError) as e:
        _restart(path, directory, **extra)
    return e.value


def test_split_run_is_bitwise_the_unsplit_run(split):
    path, first_dir, whole = split
    second = _restart(path, first_dir)
    second.steps(second.info.analysis_steps_total)
    sg, pg = second.get_wave()
    assert np.array_equal(sg, whole["wave"][0]) and np.array_equal(pg, whole["wave"][1])
    assert np.array_equal(second.s_sums(), whole["sums"])
    assert np.array_equal(second.bk_records(), whole["bk"])
    results = second.results()
    for key in ("s", "p", "p_reaction"):
        assert np.array_equal(results[key], whole["results"][key]), key
    second.close()


def _read_header(data):
    """(magic, version, system id, [(name, value)], offset after the header) of a native file."""
    magic = data[:16]
    version, = struct.unpack_from("<I", data, 16)
    n, = struct.unpack_from("<Q", data, 20)
    system_id = data[28:28 + n].decode()
    off = 28 + n
    count, = struct.unpack_from("<Q", data, off)
    off += 8
    params = []
    for _ in range(count):
        n, = struct.unpack_from("<Q", data, off)
        name = data[off + 8:off + 8 + n].decode()
        value, = struct.unpack_from("<d", data, off + 8 + n)
        params.append((name, value))
        off += 8 + n + 8
    return magic, version, system_id, params, off


def test_files_have_format_version_2_and_its_parameter_list(split):
    path, first_dir, _ = split
    first = oracle.run(path, FIRST)
    info = first.info
    first.close()
    for name, magic in (("checkpoint.bin", b"SPODVR-CHECKPNT\0"), ("bk_coefficients.bin", b"SPODVR-BKCOEFFS\0")):
        data = (first_dir / name).read_bytes()
        m, version, system_id, params, off = _read_header(data)
        assert (m, version, system_id) == (magic, 2, "H2Cu111DaiLight"), name
        assert [p[0] for p in params] == FORMAT_2, name
        values = dict(params)
        assert values["nZ"] == 64 and values["dtStep"] == 2.0 and values["MJMax"] == 4 and values["JMax"] == 4
        assert values["number of phi points (MJMax, nPhisOptimal)"] == info.n_phi
    data = (first_dir / "checkpoint.bin").read_bytes()
    *_, off = _read_header(data)
    time, = struct.unpack_from("<d", data, off)
    shape = struct.unpack_from("<5i", data, off + 8)
    n_xy, n_zr = info.nx * info.ny, info.nr * info.nz
    assert time == 40.0 and shape == (info.n_rot, n_xy, n_zr, info.nr, info.nz_specular)
    n_values = info.nr * info.nz_specular + info.n_rot * n_xy * n_zr
    assert len(data) == off + 8 + 20 + 4 + 4 + 1 + 8 + 16 * n_values + 8
    assert struct.unpack_from("<Q", data, len(data) - 8)[0] == 0x454E444F464350


def test_mismatching_input_is_an_error(split):
    path, first_dir, _ = split
    e = _error(path, first_dir, dtStep=1.0, nStep=40)
    assert e.code == 4 and "dtStep is 1 in the input but 2 in checkpoint" in str(e) and "same parameters" in str(e)


def test_tstart_must_be_the_checkpoint_time(split):
    path, first_dir, _ = split
    e = _error(path, first_dir, tStart=42.0, tEnd=82.0)
    assert e.code == 4 and "tStart = 42" in str(e) and "is at t = 40" in str(e)


def test_restart_without_restart_dir_is_an_input_error(split):
    path, _, _ = split
    with pytest.raises(spodvr_amd.SpoError) as e:
        oracle.run(path, SECOND)
    assert e.value.code == 2 and "RestartDir" in str(e.value)


def _patched(first_dir, tmp_path, name, patch):
    d = tmp_path / "patched"
    shutil.copytree(first_dir, d)
    data = bytearray((d / name).read_bytes())
    patch(data)
    (d / name).write_bytes(bytes(data))
    return d


@pytest.mark.parametrize("name", ["checkpoint.bin", "bk_coefficients.bin"])
def test_format_version_is_checked(split, tmp_path, name):
    path, first_dir, _ = split
    d = _patched(first_dir, tmp_path, name, lambda b: struct.pack_into("<I", b, 16, 3))
    e = _error(path, d)
    assert e.code == 4 and f"unsupported format version 3 in" in str(e) and name in str(e)


def test_end_marker_is_checked(split, tmp_path):
    path, first_dir, _ = split

    def flip(b):
        b[-1] ^= 1
    d = _patched(first_dir, tmp_path, "checkpoint.bin", flip)
    e = _error(path, d)
    assert e.code == 4 and "is corrupt (bad end marker)" in str(e)


def test_truncated_checkpoint_is_an_error(split, tmp_path):
    path, first_dir, _ = split
    d = _patched(first_dir, tmp_path, "checkpoint.bin", lambda b: b.__delitem__(slice(len(b) - 100, len(b))))
    e = _error(path, d)
    assert e.code == 4 and "ends too early" in str(e)


def test_system_id_is_compared_without_case(split, tmp_path):
    path, first_dir, whole = split

    def lower(b):
        b[28:28 + len("H2Cu111DaiLight")] = b"h2cu111dailight"
    d = _patched(first_dir, tmp_path, "checkpoint.bin", lower)
    run = _restart(path, d)
    run.close()


def test_analysis_surface_index_is_checked(split, tmp_path):
    path, first_dir, _ = split
    data = (first_dir / "checkpoint.bin").read_bytes()
    *_, off = _read_header(data)
    d = _patched(first_dir, tmp_path, "checkpoint.bin", lambda b: struct.pack_into("<i", b, off + 28, 4096))
    e = _error(path, d)
    assert e.code == 4 and "analysis surface index 4096 is outside the specular grid" in str(e)


# ---- the Fortran program's layout (fort.9 / fort.21 written, fort.8 / fort.22 or those read back) ----

def _records(data):
    """The gfortran records of a sequential unformatted file (int32 markers around each)."""
    pos, out = 0, []
    while pos < len(data):
        n, = struct.unpack_from("<i", data, pos)
        out.append(data[pos + 4:pos + 4 + n])
        assert struct.unpack_from("<i", data, pos + 4 + n)[0] == n
        pos += 8 + n
    return out


@pytest.mark.parametrize("case", ["hardwall_short", "cu111_dailight_short", "cu111_dailight_oblique",
                                  "cu111_dailight_short_sym", "cu111_dailight_short_lowz"])
def test_fortran_bk_header_is_the_fortran_programs(case, tmp_path):
    """The ten header records of fort.21 against those the Fortran program wrote for the same input: sizes,
    labels and framing exactly, the reals to 1e-12 relative."""
    from tests import fixtures
    run = oracle.run(cases.golden_input(case), {"CheckpointFormat": "fortran"})
    run.write_checkpoint(str(tmp_path))
    run.close()
    mine = _records((tmp_path / "fort.21").read_bytes())
    ref = _records(open(fixtures.case_dir(case) + "/fort21_header.bin", "rb").read())
    assert len(mine) == 10 and [len(r) for r in mine] == [len(r) for r in ref]
    assert struct.unpack("<4i", mine[0])[1:] == struct.unpack("<4i", ref[0])[1:]  # nstatw, ndftow, nrotw
    np.testing.assert_allclose(struct.unpack("<6d", mine[1]), struct.unpack("<6d", ref[1]), rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(struct.unpack("<3d", mine[2][:24]), struct.unpack("<3d", ref[2][:24]), rtol=1e-12,
                               atol=1e-14)
    assert mine[2][24:] == ref[2][24:]  # JIn, MJIn, initial channel
    for k in range(3, 9):
        assert mine[k] == ref[k], f"label record {k + 1}"
    np.testing.assert_allclose(np.frombuffer(mine[9], "<f8"), np.frombuffer(ref[9], "<f8"), rtol=1e-12, atol=1e-14)


def test_split_run_through_fortran_files_is_bitwise_the_unsplit_run(split, tmp_path):
    path, _, whole = split
    first = oracle.run(path, dict(FIRST, CheckpointFormat="both"))
    first.steps(first.info.analysis_steps_total)
    first.write_checkpoint(str(tmp_path))
    first.close()
    assert {"fort.9", "fort.21", "checkpoint.bin", "bk_coefficients.bin"} <= set(os.listdir(tmp_path))
    recs = _records((tmp_path / "fort.21").read_bytes())
    assert struct.unpack("<4i", recs[0])[0] == 4 and len(recs) == 10 + 4
    second = oracle.run(path, dict(SECOND, RestartDir=str(tmp_path), RestartFormat="fortran"))
    second.steps(second.info.analysis_steps_total)
    sg, pg = second.get_wave()
    assert np.array_equal(sg, whole["wave"][0]) and np.array_equal(pg, whole["wave"][1])
    assert np.array_equal(second.s_sums(), whole["sums"])
    assert np.array_equal(second.bk_records(), whole["bk"])
    second.close()


def test_fortran_restart_prefers_fort8_and_fort22(split, tmp_path):
    """A Fortran restart reads fort.8 / fort.22 when they are there (the names the Fortran program reads), before
    fort.9 / fort.21; a wrong fort.8 is therefore the one refused."""
    path, _, _ = split
    first = oracle.run(path, dict(FIRST, CheckpointFormat="fortran"))
    first.steps(first.info.analysis_steps_total)
    first.write_checkpoint(str(tmp_path))
    first.close()
    (tmp_path / "fort.8").write_bytes(struct.pack("<iii", 4, 3, 4))  # nproj = 3
    e = _error(path, tmp_path, RestartFormat="fortran")
    assert e.code == 4 and "ProjOp is 2 but 3 in" in str(e) and "fort.8" in str(e)


def test_truncated_fortran_bk_file_is_refused(split, tmp_path):
    """fort.21 announces its record count in record 1; a file cut short (a killed run) is refused, as the reference
    reads exactly that many records (src/io/fortran_files.cpp:177-183)."""
    path, _, _ = split
    first = oracle.run(path, dict(FIRST, CheckpointFormat="fortran"))
    first.steps(first.info.analysis_steps_total)
    first.write_checkpoint(str(tmp_path))
    first.close()
    data = (tmp_path / "fort.21").read_bytes()
    recs = _records(data)
    assert struct.unpack("<4i", recs[0])[0] == len(recs) - 10 == 4
    last = len(recs[-1]) + 8
    (tmp_path / "fort.21").write_bytes(data[:-last])
    e = _error(path, tmp_path, RestartFormat="fortran")
    assert e.code == 4 and "BK record 4 of 4" in str(e) and "fort.21" in str(e)


SCM_CASE = "scm_cu111_srp48_small"

# The SCM identity the reference appends for SystemID CRPDistrtSurface (src/io/checkpoint.cpp:87-106).
SCM_IDENTITY = ["PosFileSlabNum", "SCMCutoff", "EnableSurfExpan", "PosFilepath (a_T)", "PosFilepath (atoms used)",
                "PosFilepath (rms displacement)", "PosFilepath (layer 1 mean shift x)",
                "PosFilepath (layer 1 mean shift y)", "PosFilepath (layer 1 mean shift z)",
                "PosFilepath (highest atom)"] + [f"SCMParFilepath (layer {l}, P{k})" for l in (1, 2) for k in range(1, 19)]


@pytest.fixture(scope="module")
def scm_checkpoint(tmp_path_factory):
    """The native restart files of the SCM case after its first analysis step (t = 50)."""
    d = tmp_path_factory.mktemp("scm_restart")
    over = dict(cases.golden_overrides(SCM_CASE), nStep=20, tEnd=50.0)
    run = oracle.run(cases.golden_input(SCM_CASE), over)
    run.steps(1)
    run.write_checkpoint(str(d))
    run.close()
    return d, dict(over, NewCalc=False, RestartDir=str(d), tStart=50.0, tEnd=100.0)


def test_scm_files_carry_the_scm_identity(scm_checkpoint):
    d, restart = scm_checkpoint
    run = oracle.run(cases.golden_input(SCM_CASE), restart)  # the same set-up continues
    run.close()
    _, _, system_id, params, _ = _read_header((d / "checkpoint.bin").read_bytes())
    names = [p[0] for p in params]
    assert system_id == "CRPDistrtSurface"
    assert names == FORMAT_2 + SCM_IDENTITY
    values = dict(params)
    assert values["PosFileSlabNum"] == 1 and values["SCMCutoff"] == 16.0 and values["EnableSurfExpan"] == 1.0
    assert values["PosFilepath (atoms used)"] > 0 and values["PosFilepath (rms displacement)"] > 0
    assert values["PosFilepath (a_T)"] == pytest.approx(4.8896663475)


@pytest.mark.parametrize("key,value,text", [("SCMCutoff", 12.0, "SCMCutoff is 12 in the input but 16"),
                                            ("EnableSurfExpan", False, "EnableSurfExpan is 0 in the input but 1")])
def test_scm_restart_on_another_slab_set_up_is_refused(scm_checkpoint, key, value, text):
    d, over = scm_checkpoint
    with pytest.raises(spodvr_amd.SpoError) as e:
        oracle.run(cases.golden_input(SCM_CASE), dict(over, **{key: value}))
    assert e.value.code == 4 and text in str(e.value)
