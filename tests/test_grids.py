"""Host parts of the molecular grids (hesship.grids): the Treutler-Ahlrichs radial rule, the
Lebedev angular rule, the radii adjustment and the synthetic water clusters."""
import math

import numpy as np
import pytest

from hesship import grids, synth


@pytest.mark.parametrize("a", [0.3, 1.0, 8.0])
def test_radial_rule_integrates_gaussians(a):
    r, w = grids.treutler_ahlrichs(75, 8)
    assert np.all(np.diff(r) > 0) and np.all(w > 0)
    # int e^{-a r^2} d^3r = (pi / a)^{3/2}
    assert (w * np.exp(-a * r * r)).sum() == pytest.approx((math.pi / a) ** 1.5, rel=1e-9)


def test_lebedev_rule_and_adjustment():
    x, w = grids.lebedev(302)
    assert x.shape == (302, 3) and w.sum() == pytest.approx(1.0, abs=1e-14)
    np.testing.assert_allclose(np.linalg.norm(x, axis=1), 1.0, atol=1e-14)
    # exact for polynomials up to degree 29: <z^4> = 1/5, <x^2 y^2 z^2> = 1/105
    assert (w * x[:, 2] ** 4).sum() == pytest.approx(0.2, abs=1e-14)
    assert (w * (x[:, 0] * x[:, 1] * x[:, 2]) ** 2).sum() == pytest.approx(1 / 105, abs=1e-14)
    a = grids.treutler_radii_adjust([8, 1, 1, 6])
    np.testing.assert_allclose(a, -a.T, atol=1e-15)
    assert np.all(np.abs(a) <= 0.5) and a[0, 1] != 0.0 and a[1, 2] == 0.0


def test_water_cluster_geometry():
    z, R = synth.water_cluster(8, seed=3)
    assert z.tolist() == [8, 1, 1] * 8
    ang = R * 0.52917721092
    for m in range(8):
        o, h1, h2 = ang[3 * m], ang[3 * m + 1], ang[3 * m + 2]
        assert np.linalg.norm(h1 - o) == pytest.approx(0.96, abs=1e-12)
        c = np.dot(h1 - o, h2 - o) / 0.96 ** 2
        assert math.degrees(math.acos(c)) == pytest.approx(104.5, abs=1e-9)
    oxy = ang[0::3]
    d = np.linalg.norm(oxy[:, None] - oxy[None], axis=-1)
    assert d[np.triu_indices(8, 1)].min() >= 2.7
