"""The potential surrogate's inputs and targets: six features of a molecular geometry that are invariant under the
C3v operations about a top site (acting on position and azimuth together), the lattice translations and the
exchange of the two identical atoms, and the scaled energy the network learns.

  f1 = (Z - 4) / 2, f2 = (r - 1.6) / 0.6                         height and bond length
  f3 = cos 2 theta                                                 orientation to the normal (exchange-invariant)
  f4 = (cos g1 + cos g2 + cos g3) / 3                              first star of the surface lattice, g_k = G_k . R
  f5 = (sin g1 + sin g2 + sin g3) / 3                              its odd part, which C3v (not C6v) keeps
  f6 = sin^2 theta sum_k cos 2 (phi - alpha_k) cos g_k / 3         the azimuth coupled to the lattice
with (u, v) the skewed coordinates, g1 = 2 pi u / a, g2 = 2 pi v / a, g3 = -(g1 + g2), and alpha_k the directions of
the reciprocal vectors G_k (-30, 90 and 210 degrees for the 60-degree cell). The target is V in eV clipped to
[V_LOW, V_HIGH]: above it the propagation's potential is capped, and the surrogate's error is judged only where the
reference surface is below V_HIGH."""

from __future__ import annotations

import math

import numpy as np

V_LOW, V_HIGH = -1.0, 4.0  # eV
ALPHAS = np.radians([-30.0, 90.0, 210.0])


def features(points: np.ndarray, lattice_constant: float) -> np.ndarray:
    """points (n, 6) = u v Z r theta phi -> (n, 6) float32."""
    p = np.asarray(points, dtype=np.float64)
    u, v, z, r, th, ph = (p[:, k] for k in range(6))
    g1 = 2 * math.pi * u / lattice_constant
    g2 = 2 * math.pi * v / lattice_constant
    g3 = -(g1 + g2)
    s2 = np.sin(th) ** 2
    f = np.empty((len(p), 6), dtype=np.float64)
    f[:, 0] = (z - 4.0) / 2.0
    f[:, 1] = (r - 1.6) / 0.6
    f[:, 2] = np.cos(2 * th)
    f[:, 3] = (np.cos(g1) + np.cos(g2) + np.cos(g3)) / 3.0
    f[:, 4] = (np.sin(g1) + np.sin(g2) + np.sin(g3)) / 3.0
    f[:, 5] = s2 * sum(np.cos(2 * (ph - a)) * np.cos(gk) for a, gk in zip(ALPHAS, (g1, g2, g3))) / 3.0
    return f.astype(np.float32)


def target(v_ev: np.ndarray) -> np.ndarray:
    return np.clip(np.asarray(v_ev, dtype=np.float64), V_LOW, V_HIGH).astype(np.float32)


def sample_points(parameters: dict, n: int, seed: int, z_range=(0.5, 8.0), r_range=(0.6, 3.0)) -> np.ndarray:
    """n random geometries over the cell (u, v), Z, r and the orientation sphere (theta uniform in cos theta)."""
    rng = np.random.default_rng(seed)
    a = parameters["lattice_constant"]
    return np.column_stack([rng.uniform(0, a, n), rng.uniform(0, a, n), rng.uniform(*z_range, n),
                            rng.uniform(*r_range, n), np.arccos(rng.uniform(-1, 1, n)),
                            rng.uniform(0, 2 * math.pi, n)])


def train(net, crp, n: int = 1 << 19, seed: int = 1, schedule=((2e-3, 6000), (1e-3, 6000), (3e-4, 6000))):
    """Trains `net` (spodvr_amd.Mlp) on n CRP energies at random geometries (the CRP evaluated on the GPU);
    returns the batch losses of every step."""
    pts = sample_points(crp.parameters(), n, seed)
    x = features(pts, crp.parameters()["lattice_constant"])
    t = target(crp.device_qd(pts))
    return np.concatenate([net.train(x, t, steps=steps, lr=lr) for lr, steps in schedule])


def grid_values(points: np.ndarray, slab: int, lattice_constant: float, net, reference, stride: int = 16,
                tolerance: float = 0.01):
    """V in eV at `points` (n, 6), ordered slab by slab (`slab` points each, e.g. one (Z, r) pair of the potential
    grid): the network where it is trusted, `reference` (the CRP on the GPU) elsewhere. A slab is trusted when the
    network agrees with the reference within `tolerance` eV at every `stride`-th of its points and none of those
    lies above V_HIGH; the checked points keep the reference's values. Returns (values, served, trusted): served
    marks the points the network gave, trusted the slabs."""
    n = len(points)
    if n % slab:
        raise ValueError("the points are not a whole number of slabs")
    y = net.predict(features(points, lattice_constant)).astype(np.float64)
    checked = (np.arange(n // slab)[:, None] * slab + np.arange(0, slab, stride)[None, :]).ravel()
    exact = reference(points[checked])
    err = np.abs(y[checked] - exact).reshape(n // slab, -1)
    trusted = (err.max(axis=1) <= tolerance) & (exact.reshape(n // slab, -1).max(axis=1) < V_HIGH)
    values = y
    values[checked] = exact
    redo = np.flatnonzero(np.repeat(~trusted, slab))
    if redo.size:
        values[redo] = reference(points[redo])
    served = np.repeat(trusted, slab)
    served[checked] = False
    return values, served, trusted
