"""Shared helpers of the tests: paths, the seeded pseudo-random wave packets and the comparison measure."""

from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "cpp-spo-dvr_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pseudo_random(n: int, seed: int = 12345) -> np.ndarray:
    """n complex values, real and imaginary parts uniform in [-1, 1), from the 64-bit LCG
    x <- 6364136223846793005 x + 1442695040888963407 (the generator of the reference's GPU tests,
    tests/cuda/cuda_test_setup.hpp:53-63), drawn real part first."""
    a = np.uint64(6364136223846793005)
    c = np.uint64(1442695040888963407)
    # Jump-ahead in blocks keeps this vectorised: x_{k} = A_k x_0 + C_k for k = 1 .. 2n.
    m = 2 * n
    out = np.empty(m, dtype=np.uint64)
    x = np.uint64(seed)
    with np.errstate(over="ignore"):
        block = min(m, 4096)
        mult = np.empty(block, dtype=np.uint64)
        add = np.empty(block, dtype=np.uint64)
        am, cm = np.uint64(1), np.uint64(0)
        for k in range(block):
            am = am * a
            cm = cm * a + c
            mult[k], add[k] = am, cm
        for start in range(0, m, block):
            cnt = min(block, m - start)
            out[start:start + cnt] = mult[:cnt] * x + add[:cnt]
            x = out[start + cnt - 1]
    vals = (out >> np.uint64(11)).astype(np.float64) / float(1 << 53) * 2.0 - 1.0
    return vals[0::2] + 1j * vals[1::2]


def relative_difference(a: np.ndarray, b: np.ndarray) -> float:
    """max |a - b| / max |a| (the reference's GPU-test measure); NaN if either holds a NaN."""
    a = np.asarray(a)
    b = np.asarray(b)
    if np.isnan(a).any() or np.isnan(b).any():
        return float("nan")
    scale = np.abs(a).max()
    return float(np.abs(a - b).max() / scale) if scale > 0 else float(np.abs(b).max())
