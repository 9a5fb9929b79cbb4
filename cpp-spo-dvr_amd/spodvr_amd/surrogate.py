"""The potential surrogate's inputs and targets: six features of a molecular geometry that are invariant under the
C3v operations about a top site (acting on position and azimuth together), the lattice translations and the
exchange of the two identical atoms, and the energy the network learns.

  f1 = (Z - 4) / 2, f2 = (r - 1.6) / 0.6                         height and bond length
  f3 = cos 2 theta                                                 orientation to the normal (exchange-invariant)
  f4 = (cos g1 + cos g2 + cos g3) / 3                              first star of the surface lattice, g_k = G_k . R
  f5 = (sin g1 + sin g2 + sin g3) / 3                              its odd part, which C3v (not C6v) keeps
  f6 = sin^2 theta sum_k cos 2 (phi - alpha_k) cos g_k / 3         the azimuth coupled to the lattice
with (u, v) the skewed coordinates, g1 = 2 pi u / a, g2 = 2 pi v / a, g3 = -(g1 + g2), and alpha_k the directions of
the reciprocal vectors G_k (-30, 90 and 210 degrees for the 60-degree cell).

The network learns the CRP's own smooth part: the remainder R = V - w(Z) (V_atom(A) + V_atom(B)) of the
molecule-surface energy after the two atom-surface energies (the corrugation-reducing split of the CRP method), in
eV, where V is the QD wrapper's energy (r and Z clamped into the data, the free molecule at its last tabulated r as
the zero) and w fades the atom terms out across the switch to the gas phase (atom_weight). The atom terms carry the
steep, site-specific repulsion, which six features cannot resolve; R varies slowly over the cell. The surrogate's V
is R + w (V_atom(A) + V_atom(B)), clamped to [V_MIN, V_MAX] as the propagation clamps its grid (keys ImposeVMin /
VMin and ImposeVMax / VMax, defaults -5.5 and 10 eV)."""

from __future__ import annotations

import math

import numpy as np

V_MIN, V_MAX = -5.5, 10.0  # eV: the propagation's default clamp of the potential grid
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


def clamp_points(points: np.ndarray, parameters: dict) -> np.ndarray:
    """The geometries with r and Z clamped into the data set's range as the QD wrapper clamps them (the energy is
    constant beyond): the network sees the geometry the CRP evaluates."""
    p = np.array(points, dtype=np.float64)
    delta = 1e-12
    p[:, 3] = np.clip(p[:, 3], parameters["r_min"] + delta, parameters["r_max"] - delta)
    p[:, 2] = np.maximum(p[:, 2], parameters["z_min"] + delta)
    return p


def atom_positions(points: np.ndarray, parameters: dict):
    """The two atoms (bohr, Cartesian) of each geometry u v Z r theta phi, as the QD wrapper places them: r and Z
    clamped into the data set's range, (x, y) = (u + v cos g, v sin g) with g the cell's skewing angle."""
    p = np.asarray(points, dtype=np.float64)
    u, v, z, r, th, ph = (p[:, k] for k in range(6))
    delta = 1e-12
    r = np.clip(r, parameters["r_min"] + delta, parameters["r_max"] - delta)
    z = np.maximum(z, parameters["z_min"] + delta)
    g = parameters["skewing_angle"]
    centre = np.column_stack([u + v * math.cos(g), v * math.sin(g), z])
    half = 0.5 * r[:, None] * np.column_stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    return centre - half, centre + half


def atom_weight(z: np.ndarray, parameters: dict) -> np.ndarray:
    """1 below the molecule surface's switch to the gas phase (Z <= zso_min), 0 above it (Z >= zso_max, where the
    molecule's energy no longer depends on the atoms), cos^2 between: the atom terms fade out with it, so that the
    remainder stays continuous and the atoms are never evaluated beyond their data."""
    lo, hi = parameters["zso_min"], parameters["zso_max"]
    s = np.clip((np.asarray(z, dtype=np.float64) - lo) / (hi - lo), 0.0, 1.0)
    return np.where(s < 1.0, np.cos(0.5 * math.pi * s) ** 2, 0.0)  # exactly 0 from zso_max on


def atom_energies(crp, points: np.ndarray) -> np.ndarray:
    """w(Z) (V_atom(A) + V_atom(B)) in eV at each geometry (the CRP's atom-surface PES on the host; w from
    atom_weight at the clamped Z)."""
    from . import CRP_EV
    par = crp.parameters()
    z = np.maximum(np.asarray(points, dtype=np.float64)[:, 2], par["z_min"] + 1e-12)
    w = atom_weight(z, par)
    out = np.zeros(len(z))
    near = np.flatnonzero(w > 0.0)
    if near.size:
        a, b = atom_positions(np.asarray(points)[near], par)
        out[near] = w[near] * (crp.atom(a, energy_unit=CRP_EV)[:, 0] + crp.atom(b, energy_unit=CRP_EV)[:, 0])
    return out


def target(v_ev: np.ndarray, atoms_ev: np.ndarray) -> np.ndarray:
    """The remainder the network learns: V minus atom_energies, eV."""
    return (np.asarray(v_ev, dtype=np.float64) - np.asarray(atoms_ev, dtype=np.float64)).astype(np.float32)


def sample_points(parameters: dict, n: int, seed: int, z_range=None, r_range=None) -> np.ndarray:
    """n random geometries over the cell (u, v), Z, r and the orientation sphere (theta uniform in cos theta). By
    default Z and r cover the data set's range: Z from its lowest value to 1 bohr past the gas-phase asymptote, r
    over the tabulated bond lengths (the QD wrapper clamps outside both)."""
    rng = np.random.default_rng(seed)
    a = parameters["lattice_constant"]
    if z_range is None:
        z_range = (parameters["z_min"], parameters["z_asymptotic"] + 1.0)
    if r_range is None:
        r_range = (parameters["r_min"], parameters["r_max"])
    return np.column_stack([rng.uniform(0, a, n), rng.uniform(0, a, n), rng.uniform(*z_range, n),
                            rng.uniform(*r_range, n), np.arccos(rng.uniform(-1, 1, n)),
                            rng.uniform(0, 2 * math.pi, n)])


def train(net, crp, n: int = 1 << 21, seed: int = 1, schedule=((2e-3, 12000), (1e-3, 12000), (3e-4, 12000))):
    """Trains `net` (spodvr_amd.Mlp) on the remainders of n CRP energies at random geometries (the molecule energy
    from the CRP on the GPU, the atom energies on the host); returns the batch losses of every step."""
    par = crp.parameters()
    pts = sample_points(par, n, seed)
    x = features(clamp_points(pts, par), par["lattice_constant"])
    t = target(crp.device_qd(pts), atom_energies(crp, pts))
    return np.concatenate([net.train(x, t, steps=steps, lr=lr) for lr, steps in schedule])


def grid_values(points: np.ndarray, slab: int, net, crp, stride: int = 16, tolerance: float = 0.01,
                v_min: float = V_MIN, v_max: float = V_MAX):
    """V in eV at `points` (n, 6), ordered slab by slab (`slab` points each, e.g. one (Z, r) pair of the potential
    grid), clamped to [v_min, v_max]: the network (its remainder plus the two atom energies) where it is trusted,
    the CRP on the GPU elsewhere. A slab is trusted when the network's clamped value agrees with the CRP's within
    `tolerance` eV at every `stride`-th of its points; the checked points keep the CRP's values. Returns (values,
    served, trusted): served marks the points the network gave, trusted the slabs."""
    n = len(points)
    if n % slab:
        raise ValueError("the points are not a whole number of slabs")
    par = crp.parameters()
    y = net.predict(features(clamp_points(points, par), par["lattice_constant"])).astype(np.float64)
    y = np.clip(y + atom_energies(crp, points), v_min, v_max)
    checked = (np.arange(n // slab)[:, None] * slab + np.arange(0, slab, stride)[None, :]).ravel()
    exact = np.clip(crp.device_qd(points[checked]), v_min, v_max)
    err = np.abs(y[checked] - exact).reshape(n // slab, -1)
    trusted = err.max(axis=1) <= tolerance
    values = y
    values[checked] = exact
    redo = np.flatnonzero(np.repeat(~trusted, slab))
    if redo.size:
        values[redo] = np.clip(crp.device_qd(points[redo]), v_min, v_max)
    served = np.repeat(trusted, slab)
    served[checked] = False
    return values, served, trusted


def grid_points(run) -> tuple[np.ndarray, np.ndarray]:
    """The geometries (u v Z r theta phi) of a run's potential grid in the order of its array("potential.values")
    ((theta, phi) fastest, then the representative sites, then izr = iz NR + ir), and the mask of the entries at or
    below ZminSwitch (all of them without UseSwitching): the rows the surface itself fills, before the switch to the
    gas phase blends them."""
    i = run.info
    x, y, z, r = (run.array(f"grids.{k}") for k in "xyzr")
    thetas, phis = run.array("angular.thetas"), run.array("angular.phis")
    reps = run.array("symmetry.representatives").astype(int)
    ix, iy = reps // i.ny, reps % i.ny
    iz, ir = np.divmod(np.arange(i.nz_potential * i.nr), i.nr)
    n_tp, n_rep = len(thetas) * len(phis), len(reps)
    izr = np.repeat(np.arange(len(iz)), n_rep * n_tp)
    irep = np.tile(np.repeat(np.arange(n_rep), n_tp), len(iz))
    itp = np.tile(np.arange(n_tp), len(iz) * n_rep)
    it, ip = np.divmod(itp, len(phis))
    pts = np.column_stack([x[ix[irep]], y[iy[irep]], z[iz[izr]], r[ir[izr]], thetas[it], phis[ip]])
    switch = run.array("switch.z")
    below = np.ones(len(pts), bool) if switch.size == 0 else pts[:, 2] <= switch[0]
    return pts, below


def propagation_grid(run, net, crp, tolerance: float = 0.01, stride: int = 16):
    """A potential grid for spodvr_amd.Run(..., potential=...), built from the network on the rows below the switch
    to the gas phase: the (Z, r) slabs the network is trusted on (grid_values, checked against the CRP at every
    `stride`-th point within `tolerance` eV) take its values in hartree, and every other entry keeps the value of the
    run's own CRP-built grid exactly (untrusted slabs, the checked points, the switched rows). Returns (grid,
    served): served marks the entries the network gave."""
    pts, below = grid_points(run)
    grid = run.array("potential.values").copy()
    i = run.info
    slab = i.n_theta * i.n_phi * i.n_representatives
    rows = np.flatnonzero(below)
    if rows.size == 0:
        return grid, np.zeros(len(grid), bool)
    values, served_rows, _ = grid_values(pts[rows], slab, net, crp, stride=stride, tolerance=tolerance)
    served = np.zeros(len(grid), bool)
    served[rows] = served_rows
    grid[served] = values[served_rows] / run.array("units.hartree_in_ev")[0]
    return grid, served
