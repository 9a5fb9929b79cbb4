"""The surrogate network's CPU oracle (oracle/mlp.py): its backward pass against central finite differences of its
forward pass (with the bf16 rounding switched off, the loss is smooth), and the bf16 rounding against its
definition (nearest even, ties to even)."""

import numpy as np

import oracle.mlp as mlp


def test_gradient_matches_finite_differences(monkeypatch):
    monkeypatch.setattr(mlp, "bf16", lambda a: np.asarray(a, dtype=np.float64))
    p = mlp.initial_parameters(1).astype(np.float64)
    rng = np.random.default_rng(0)
    x = rng.normal(size=(48, 6))
    t = rng.normal(size=48)
    _, g = mlp.loss_and_gradient(p, x, t)
    # parameters of every block: W1, b1 (zero), W2, b2, W3, b3, W4, b4
    offsets = np.cumsum([0] + [int(np.prod(s)) for s in mlp.SHAPES])
    for k in range(len(mlp.SHAPES)):
        for i in rng.choice(np.arange(offsets[k], offsets[k + 1]), size=min(4, offsets[k + 1] - offsets[k]),
                            replace=False):
            q = p.copy()
            q[i] += 1e-6
            lp, _ = mlp.loss_and_gradient(q, x, t)
            q[i] -= 2e-6
            lm, _ = mlp.loss_and_gradient(q, x, t)
            assert abs((lp - lm) / 2e-6 - g[i]) <= 1e-7 + 1e-6 * abs(g[i]), (k, i)


def test_bf16_rounding():
    one_ulp = 2.0 ** -7  # bf16 has 8 significant bits
    x = np.array([1.0, 1.0 + one_ulp / 2, 1.0 + 3 * one_ulp / 2, 1.0 + one_ulp / 4, -2.5, 3.0e-3], dtype=np.float32)
    r = mlp.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0  # a tie rounds to the even significand
    assert r[2] == 1.0 + 2 * one_ulp  # a tie rounds to the even significand
    assert r[3] == 1.0
    assert r[4] == -2.5
    assert abs(r[5] - 3.0e-3) <= 3.0e-3 * 2.0 ** -8


def test_surrogate_features_are_invariant():
    """The six features under the rotation by 120 degrees about a top site, the mirror x -> -x (the plane through
    the hcp and fcc sites), the exchange of the atoms and a lattice translation (spodvr_amd/surrogate.py)."""
    import math

    from spodvr_amd import surrogate
    a = 4.9
    pts = surrogate.sample_points({"lattice_constant": a}, 200, 1, z_range=(0.5, 8.0), r_range=(0.6, 3.0))

    def cart(u, v):
        return u + v * 0.5, v * math.sqrt(3) / 2

    def skew(x, y):
        v = y / (math.sqrt(3) / 2)
        return x - 0.5 * v, v

    images = [[], [], [], []]
    c, s = math.cos(2 * math.pi / 3), math.sin(2 * math.pi / 3)
    for (u, v, z, r, th, ph) in pts:
        x, y = cart(u, v)
        images[0].append((*skew(c * x - s * y, s * x + c * y), z, r, th, ph + 2 * math.pi / 3))
        images[1].append((*skew(-x, y), z, r, th, math.pi - ph))
        images[2].append((u, v, z, r, math.pi - th, ph + math.pi))
        images[3].append((u + a, v - a, z, r, th, ph))
    f0 = surrogate.features(pts, a)
    for im in images:
        np.testing.assert_allclose(surrogate.features(np.array(im), a), f0, rtol=0, atol=1e-6)


def test_surrogate_geometry_is_the_qd_wrappers():
    """The surrogate's atoms (surrogate.atom_positions, r and Z clamped as the QD wrapper clamps them) give the QD
    wrapper's energy through the molecule-surface PES: molecule(A, B) in eV minus the free molecule's asymptote equals
    qd(u, v, Z, r, theta, phi), at random geometries that include clamped ones (SRP48 data set, host CRP). A wrong
    skew, a swapped component or a missing clamp moves the energies by far more than the 1e-9 eV allowed."""
    import os

    import spodvr_amd
    from spodvr_amd import surrogate
    from tests import cases
    crp = spodvr_amd.CrpSurface(os.path.join(cases.GOLDEN_DIR, "crp", "cu111_srp48", "dataset"))
    try:
        par = crp.parameters()
        pts = surrogate.sample_points(par, 2000, seed=4, z_range=(-1.0, 7.0), r_range=(0.4, 5.0))
        a, b = surrogate.atom_positions(pts, par)
        mol = crp.molecule(np.column_stack([a, b]), energy_unit=spodvr_amd.CRP_EV)[:, 0]
        qd = crp.qd(pts)
        shift = mol - qd  # the asymptote, the same constant at every geometry
        assert np.max(np.abs(shift - np.median(shift))) < 1e-9
        clamped = surrogate.clamp_points(pts, par)
        assert np.array_equal(surrogate.atom_positions(clamped, par)[0], a)
        assert np.allclose(crp.qd(clamped), qd, rtol=0, atol=1e-12)
    finally:
        crp.close()


def test_surrogate_atom_weight():
    """The fade of the atom terms across the switch to the gas phase: 1 up to zso_min, exactly 0 from zso_max on
    (the atoms are not evaluated there), continuous and decreasing between."""
    from spodvr_amd import surrogate
    par = {"zso_min": 6.4, "zso_max": 7.6}
    z = np.linspace(5.0, 9.0, 4001)
    w = surrogate.atom_weight(z, par)
    assert np.all(w[z <= 6.4] == 1.0) and np.all(w[z >= 7.6] == 0.0)
    assert np.all(np.diff(w) <= 0.0) and np.max(np.abs(np.diff(w))) < 2e-3
