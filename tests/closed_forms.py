"""Closed-form checks of the set-up and of the propagation steps, computed independently of the code under test
(numpy / scipy, or the mathematics alone). Each returns the worst deviation, so that a test can assert it is small
and that a planted error of a given size makes it large.

The reference's known-answer tests these restate: tests/test_angular.cpp:126 (DVR values of single basis
functions), :151 (round trip = identity), :167 (cos^2 theta matrix elements); tests/test_propagation.cpp:99 (the
kinetic step keeps the norm, half steps compose), :119 (a free Gaussian spreads analytically), :179 (a parallel plane
wave gets its kinetic phase), :209 (a constant potential is a phase)."""

from __future__ import annotations

import numpy as np


# ---- angular set-up --------------------------------------------------------------------------------------------

def gauss_error(thetas, weights, j_max, homonuclear):
    """The theta grid against numpy's Gauss-Legendre rule of j_max + 1 points (src/grids/angular_grid.cpp:37-64): for
    a homonuclear molecule the nodes with cos(theta) >= 0 with doubled weights (theta = pi/2 counted once), else the
    whole rule. Returns max(|cos theta - x|, |w - w_ref|)."""
    x, w = np.polynomial.legendre.leggauss(j_max + 1)
    order = np.argsort(-x)
    x, w = x[order], w[order]
    if homonuclear:
        keep = x >= -1e-14
        x, w = x[keep], np.where(np.abs(x[keep]) < 1e-14, w[keep], 2.0 * w[keep])
    if len(x) != len(thetas):
        return np.inf
    return float(max(np.abs(np.cos(thetas) - x).max(), np.abs(np.asarray(weights) - w).max()))


def blocks_of(basis_j, basis_mj):
    """The mJ blocks of the rotational basis: [(m, [J ...])] in basis order (src/grids/rotational_basis.cpp:11-26)."""
    out = []
    for j, m in zip(basis_j.astype(int), basis_mj.astype(int)):
        if not out or out[-1][0] != m:
            out.append((m, []))
        out[-1][1].append(j)
    return out


def split_legendre(legendre, blocks, n_theta):
    """The (mJ block, theta, J) array as one (n_theta, n_J) matrix per block."""
    mats, at = [], 0
    for _, js in blocks:
        n = n_theta * len(js)
        mats.append(np.asarray(legendre[at:at + n]).reshape(n_theta, len(js)))
        at += n
    if at != len(legendre):
        raise ValueError("Legendre array does not match the basis")
    return mats


def legendre_error(legendre, thetas, weights, n_phi, blocks):
    """The matrix sqrt(w_t / n_phi) P~_J^m(theta_t) against scipy's spherical harmonics: with the reference's
    normalisation P~_J^m = sqrt(J + 1/2) sqrt((J - |m|)! / (J + |m|)!) P_J^|m| and the Condon-Shortley phase for
    m > 0 (src/grids/associated_legendre.cpp:20-47), P~_J^m(theta) = sqrt(2 pi) Y_J^m(theta, 0) for every m.
    Returns max |L - L_ref| / max |L_ref| over the blocks."""
    from scipy.special import sph_harm_y
    mats = split_legendre(legendre, blocks, len(thetas))
    worst, scale = 0.0, 0.0
    for (m, js), mat in zip(blocks, mats):
        for k, j in enumerate(js):
            ref = np.sqrt(2 * np.pi) * sph_harm_y(j, m, np.asarray(thetas), 0.0).real
            ref = ref * np.sqrt(np.asarray(weights) / n_phi)
            worst = max(worst, float(np.abs(mat[:, k] - ref).max()))
            scale = max(scale, float(np.abs(ref).max()))
    return worst / scale


def round_trip_error(legendre, n_theta, n_phi, blocks):
    """FBR -> DVR -> FBR is the identity (tests/test_angular.cpp:151): the phi FFT and its unscaled inverse give
    n_phi, so n_phi L_m^T L_m = 1 for every mJ block. Returns max |n_phi L^T L - 1|."""
    worst = 0.0
    for mat in split_legendre(legendre, blocks, n_theta):
        g = n_phi * mat.T @ mat
        worst = max(worst, float(np.abs(g - np.eye(len(g))).max()))
    return worst


def cos2_reference(j_values, m):
    """<J m| cos^2 theta |J' m> of orthonormal associated Legendre functions: cos theta couples J to J +- 1 with
    a_J = sqrt(((J + 1)^2 - m^2) / ((2 J + 1) (2 J + 3))); the square of that tridiagonal matrix, built over
    J = |m| .. max + 2, restricted to j_values."""
    am = abs(m)
    top = max(j_values) + 2
    js = list(range(am, top + 1))
    x = np.zeros((len(js), len(js)))
    for i, j in enumerate(js[:-1]):
        a = np.sqrt(((j + 1) ** 2 - m * m) / ((2 * j + 1) * (2 * j + 3)))
        x[i, i + 1] = x[i + 1, i] = a
    x2 = x @ x
    idx = [js.index(j) for j in j_values]
    return x2[np.ix_(idx, idx)]


def cos2_error(legendre, thetas, n_phi, blocks, j_max):
    """Multiplying by cos^2 theta on the DVR grid and going back gives the exact matrix elements while the Gauss rule
    integrates the product (J + J' + 2 <= 2 JMax + 1) (tests/test_angular.cpp:167)."""
    c2 = np.cos(np.asarray(thetas)) ** 2
    worst = 0.0
    for (m, js), mat in zip(blocks, split_legendre(legendre, blocks, len(thetas))):
        got = n_phi * mat.T @ (c2[:, None] * mat)
        ref = cos2_reference(js, m)
        for a, ja in enumerate(js):
            for b, jb in enumerate(js):
                if ja + jb + 2 <= 2 * j_max + 1:
                    worst = max(worst, abs(got[a, b] - ref[a, b]))
    return float(worst)


# ---- symmetry-reduced potential grid ---------------------------------------------------------------------------

def symmetry_error(v_sym, v_full, rep_of, op_of, phi_index, n_theta, n_phi, n_xy):
    """V on the symmetry representatives through the maps against V at every site (src/symmetry/grid_symmetry.hpp:
    60-70): V_full[phi + n_phi theta, ixy, izr] = V_sym[phi_index[op(ixy)][phi] + n_phi theta, rep(ixy), izr].
    Returns max |difference| / max(|V|, 1 eV in hartree)."""
    n_tp = n_theta * n_phi
    n_rep = len(v_sym) // (n_tp * (len(v_full) // (n_tp * n_xy)))
    full = np.asarray(v_full).reshape(-1, n_xy, n_tp)
    sym = np.asarray(v_sym).reshape(-1, n_rep, n_tp)
    maps = np.asarray(phi_index, dtype=int).reshape(-1, n_phi)
    tp = np.arange(n_tp)
    th, ph = tp // n_phi, tp % n_phi
    worst = 0.0
    for ixy in range(n_xy):
        op, rep = int(op_of[ixy]), int(rep_of[ixy])
        gathered = sym[:, rep, maps[op][ph] + n_phi * th]
        d = np.abs(full[:, ixy, :] - gathered) / np.maximum(np.abs(full[:, ixy, :]), 1.0 / 27.2114)
        worst = max(worst, float(d.max()))
    return worst


# ---- propagation ------------------------------------------------------------------------------------------------

def free_gaussian(z, t, mass, a=1.0, z0=5.0):
    """psi(Z, t) = (2a/pi)^(1/4) / sqrt(1 + i b) exp(-a (Z - Z0)^2 / (1 + i b)), b = 2 a t / M."""
    d = 1.0 + 2j * a * t / mass
    return (2 * a / np.pi) ** 0.25 / np.sqrt(d) * np.exp(-a * (z - z0) ** 2 / d)


def plane_wave(cell, nx, ny, n, m):
    """exp(i K . R) on the (x, y) grid, K = k_parallel + n b1 + m b2 with b1, b2 the reciprocal vectors of the cell
    a1 = lx (1, 0), a2 = ly (cos s, sin s); grid points R = ix dx (1, 0) + iy dy (cos s, sin s); ixy = iy + ny ix.
    Returns (values per ixy, |K|^2)."""
    lx, ly, skew, _, dx, dy, _, kpx, kpy = cell
    b1 = 2 * np.pi / lx * np.array([1.0, -np.cos(skew) / np.sin(skew)])
    b2 = 2 * np.pi / (ly * np.sin(skew)) * np.array([0.0, 1.0])
    k = np.array([kpx, kpy]) + n * b1 + m * b2
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    rx = ix * dx + iy * dy * np.cos(skew)
    ry = iy * dy * np.sin(skew)
    return np.exp(1j * (k[0] * rx + k[1] * ry)).reshape(-1), float(k @ k)


def host_wave(column_values, n_rot, n_xy, n_zr):
    """A host-layout scattering wave (irot, ixy, izr; irot fastest) from values[izr, ixy, irot]."""
    return np.ascontiguousarray(np.broadcast_to(column_values, (n_zr, n_xy, n_rot))).reshape(-1).astype(np.complex128)
