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
    m = 2 * n
    if m == 0:
        return np.zeros(0, dtype=np.complex128)
    out = _lcg_states(m, seed)
    np.right_shift(out, np.uint64(11), out=out)
    vals = out.view(np.int64).astype(np.float64)  # < 2^53: exact, and the signed conversion is the fast one
    del out
    vals *= 2.0 / float(1 << 53)
    vals -= 1.0
    return vals.view(np.complex128)  # (re, im) interleaved: draws 2k and 2k + 1


_LCG_A = 6364136223846793005
_LCG_C = 1442695040888963407


def _jump_tables(block: int, a: int, c: int):
    """x_k = A_k x_0 + C_k (mod 2^64) for k = 1 .. block: the multipliers and increments of k steps of the LCG
    x <- a x + c."""
    mult = np.empty(block, dtype=np.uint64)
    add = np.empty(block, dtype=np.uint64)
    am, cm = 1, 0
    mask = (1 << 64) - 1
    for k in range(block):
        am = (am * a) & mask
        cm = (cm * a + c) & mask
        mult[k], add[k] = am, cm
    return mult, add


def _lcg_states(m: int, seed: int) -> np.ndarray:
    """The states x_1 .. x_m of the LCG from x_0 = seed. Two levels of jump-ahead keep it vectorised: the states at
    every block start come from the LCG of `block` steps (itself jumped ahead), and each block from its start."""
    block = min(m, 4096)
    mult, add = _jump_tables(block, _LCG_A, _LCG_C)
    n_blocks = (m + block - 1) // block
    # Block starts s_j = x_{j block}: the LCG s <- A_block s + C_block from s_0 = seed.
    starts = np.empty(n_blocks, dtype=np.uint64)
    starts[0] = np.uint64(seed)
    if n_blocks > 1:
        inner = min(n_blocks - 1, 4096)
        bm, ba = _jump_tables(inner, int(mult[-1]), int(add[-1]))
        x = np.uint64(seed)
        with np.errstate(over="ignore"):
            for j0 in range(1, n_blocks, inner):
                cnt = min(inner, n_blocks - j0)
                starts[j0:j0 + cnt] = bm[:cnt] * x + ba[:cnt]
                x = starts[j0 + cnt - 1]
    out = np.empty(n_blocks * block, dtype=np.uint64)
    view = out.reshape(n_blocks, block)
    with np.errstate(over="ignore"):
        rows = 1 << 12
        for j0 in range(0, n_blocks, rows):
            s = starts[j0:j0 + rows, None]
            np.multiply(mult[None, :], s, out=view[j0:j0 + rows])
            view[j0:j0 + rows] += add[None, :]
    return out[:m]


def relative_difference(a: np.ndarray, b: np.ndarray) -> float:
    """max |a - b| / max |a| (the reference's GPU-test measure); NaN if either holds a NaN."""
    a = np.asarray(a)
    b = np.asarray(b)
    if np.isnan(a).any() or np.isnan(b).any():
        return float("nan")
    scale = np.abs(a).max()
    return float(np.abs(a - b).max() / scale) if scale > 0 else float(np.abs(b).max())
