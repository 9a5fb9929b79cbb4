"""The C-ABI libraries load without a GPU and export every function the public headers declare: include/spo_hip.h
from libspo_hip.so (the device backend) and include/spo_host.h from libspo_host.so (the run driver)."""

import ctypes
import os
import re

import pytest

import spodvr_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b(spo_\w+)\s*\(", re.M)


def declared(header: str) -> list[str]:
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = [m.group(1) for m in DECL.finditer(text) if not m.group(0).lstrip().startswith(("typedef", "#"))]
    return sorted(set(names))


@pytest.mark.parametrize("header,library", [("spo_hip.h", "libspo_hip.so"), ("spo_host.h", "libspo_host.so")])
def test_library_exports_every_declared_function(header, library):
    names = declared(header)
    assert len(names) >= 10, names
    lib = ctypes.CDLL(spodvr_amd.library_path(library))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_device_count_reports_without_a_device():
    """spo_hip_device_count answers on a machine with no GPU (0 there) instead of failing."""
    lib = ctypes.CDLL(spodvr_amd.library_path("libspo_hip.so"))
    count = ctypes.c_int(-1)
    status = lib.spo_hip_device_count(ctypes.byref(count))
    assert status == 0 and count.value >= 0


def test_hip_run_without_a_device_fails_loudly():
    """On a machine without a GPU the HIP device is an error, never a silent CPU fallback."""
    lib = ctypes.CDLL(spodvr_amd.library_path("libspo_hip.so"))
    count = ctypes.c_int(0)
    lib.spo_hip_device_count(ctypes.byref(count))
    if count.value > 0:
        pytest.skip("a GPU is present")
    from tests import cases
    with pytest.raises(spodvr_amd.SpoError):
        spodvr_amd.Run(cases.write_case("smoke", cases.SMOKE_DAILIGHT), device="hip")
