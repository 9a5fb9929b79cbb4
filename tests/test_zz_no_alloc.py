"""spo_hip_pes_values_static, the entry point that allocates no device memory (its points and values pass through a
fixed device array), against spo_hip_pes_values on the same points: bitwise equal on two calls in a row, the second
with fewer points than the first (so stale values of the first call would show), for both device surfaces; a call
beyond the array's 65536 points is refused with SPO_HIP_E_ARGUMENT before anything runs."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 16
SURFACES = {"HardWallPES": 1, "H2Cu111DaiLight": 2}


def _lib():
    import spodvr_amd
    spodvr_amd.host_library()  # loads libspo_hip.so with it
    lib = ctypes.CDLL(spodvr_amd.library_path("libspo_hip.so"))
    for fn in ("spo_hip_pes_values", "spo_hip_pes_values_static"):
        getattr(lib, fn).argtypes = [ctypes.c_int, ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p]
    lib.spo_hip_last_error.restype = ctypes.c_char_p
    return lib


def _points(n, seed):
    rng = np.random.default_rng(seed)
    p = np.empty((n, 6))
    p[:, 0] = rng.uniform(0.0, 4.84, n)      # x
    p[:, 1] = rng.uniform(0.0, 4.19, n)      # y
    p[:, 2] = rng.uniform(0.5, 12.0, n)      # Z
    p[:, 3] = rng.uniform(0.8, 4.0, n)       # r
    p[:, 4] = rng.uniform(0.0, np.pi, n)     # theta
    p[:, 5] = rng.uniform(0.0, 2 * np.pi, n)  # phi
    return p


def _call(lib, fn, pes, pts):
    out = np.full(len(pts), np.nan)
    status = getattr(lib, fn)(pes, len(pts), pts.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p))
    assert status == 0, lib.spo_hip_last_error().decode()
    return out


@pytest.mark.parametrize("system_id", list(SURFACES))
def test_static_entry_point_matches_the_allocating_one(system_id):
    lib = _lib()
    pes = SURFACES[system_id]
    for call, n in enumerate((MAX_POINTS, 3001)):
        print(f"{system_id}: call {call + 1}, {n} points", flush=True)
        pts = _points(n, 11 + call)
        ref = _call(lib, "spo_hip_pes_values", pes, pts)
        got = _call(lib, "spo_hip_pes_values_static", pes, pts)
        assert np.all(np.isfinite(ref))
        assert np.array_equal(got, ref), f"max diff {np.abs(got - ref).max():.3e}"


def test_static_entry_point_refuses_more_points_than_its_array():
    lib = _lib()
    pts = np.zeros((MAX_POINTS + 1, 6))
    out = np.zeros(MAX_POINTS + 1)
    status = lib.spo_hip_pes_values_static(2, len(pts), pts.ctypes.data_as(ctypes.c_void_p),
                                           out.ctypes.data_as(ctypes.c_void_p))
    assert status == 2
    assert "at most 65536" in lib.spo_hip_last_error().decode()
    assert not out.any()
