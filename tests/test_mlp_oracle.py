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
    pts = surrogate.sample_points({"lattice_constant": a}, 200, 1)

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
