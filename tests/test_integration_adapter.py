"""The reference-side adapter (cpp-spo-dvr_amd/integration/reference_hip_backend.cpp: a spodvr::PropagationBackend
of the reference program over the C ABI of include/spo_hip.h) compiles against the reference's own headers:
g++ -std=c++20 -fsyntax-only with the reference's src/ on the include path, as the reference's hot-path translation
units do with this compiler. Skipped where the reference is not mounted (it is not on the GPU machine). Nothing is
linked or run: the check is that the adapter's use of the reference's types (BackendInput, PropagationTables,
AngularTransform, PotentialGrid, GridSymmetry, WavePacket) and of the C ABI is well-formed."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFERENCE_SRC = "/root/reference/src"
ADAPTER = os.path.join(ROOT, "cpp-spo-dvr_amd", "integration", "reference_hip_backend.cpp")

pytestmark = pytest.mark.skipif(not os.path.isdir(REFERENCE_SRC) or shutil.which("g++") is None,
                                reason="the reference's sources (or g++) are not available here")


def _syntax_check(source, include_reference=True):
    cmd = ["g++", "-std=c++20", "-fsyntax-only", "-fopenmp", "-Wall", "-Wextra", f"-I{os.path.join(ROOT, 'include')}"]
    if include_reference:
        cmd.append(f"-I{REFERENCE_SRC}")
    return subprocess.run(cmd + [source], capture_output=True, text=True)


def test_adapter_compiles_against_the_reference_headers():
    proc = _syntax_check(ADAPTER)
    assert proc.returncode == 0, proc.stderr
    assert "warning" not in proc.stderr, proc.stderr


def test_the_check_depends_on_the_reference_headers():
    """Without the reference's include path the same command fails: the check does read the reference's types."""
    proc = _syntax_check(ADAPTER, include_reference=False)
    assert proc.returncode != 0
    assert "No such file" in proc.stderr and ".hpp" in proc.stderr


def test_a_planted_type_mistake_is_caught(tmp_path):
    """The adapter with one seam signature changed (z_plane returning a vector instead of the reference's span) no
    longer compiles: the override is checked against the reference's PropagationBackend."""
    text = open(ADAPTER).read()
    planted = text.replace("std::span<const cplx> z_plane(int iz) const override",
                           "std::vector<cplx> z_plane(int iz) const override")
    assert planted != text
    path = tmp_path / "planted.cpp"
    path.write_text(planted)
    proc = _syntax_check(str(path))
    assert proc.returncode != 0
