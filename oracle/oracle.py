"""TEST INFRASTRUCTURE — loader of the CPU oracle (liboracle.so) for tests/, the smoke check and bench.py's
CPU-baseline leg. The oracle is the checker of the HIP backend, never part of the product path."""

from __future__ import annotations

import ctypes
import os

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "_build", "liboracle.so")

_lib = None


def library() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (python oracle/build.py)")
        import spodvr_amd  # noqa: F401  (loads libspo_host.so globally first)
        spodvr_amd.host_library()
        lib = ctypes.CDLL(LIB_PATH)
        lib.oracle_backend_factory.restype = ctypes.c_void_p
        lib.oracle_backend_factory.argtypes = [ctypes.c_void_p]
        lib.oracle_fft.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        _lib = lib
    return _lib


def factory():
    """The C function pointer to pass as spodvr_amd.Run(..., device='external', factory=...)."""
    return library().oracle_backend_factory


def run(input_path: str, overrides=None):
    import spodvr_amd
    return spodvr_amd.Run(input_path, overrides, device="external", factory=factory())


def fft(x: np.ndarray, sign: int) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    if library().oracle_fft(a.ctypes.data_as(ctypes.c_void_p), a.size, sign):
        raise RuntimeError("oracle_fft failed")
    return a
