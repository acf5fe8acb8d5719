"""The oracle's state measurements (messung.rs) against closed forms: the cut of
a regular prism with a plane is its polygon's perimeter, seam gaps are the
pair distances, the mean velocity of a uniform motion is its speed."""
import math

import numpy as np

import oracle as O


def prisma(n_ecken, radius, hoehe, ringe):
    """A triangulated open prism over the regular n-gon, axis z, z in [-hoehe, 0]."""
    w = 2.0 * math.pi * np.arange(n_ecken) / n_ecken
    x = []
    for r in range(ringe + 1):
        z = -hoehe * r / ringe
        x += [[radius * math.cos(a), radius * math.sin(a), z] for a in w]
    tri = []
    for r in range(ringe):
        for k in range(n_ecken):
            a, b = r * n_ecken + k, r * n_ecken + (k + 1) % n_ecken
            c, d = a + n_ecken, b + n_ecken
            tri += [[a, b, d], [a, d, c]]
    return np.array(x), np.array(tri, dtype=np.uint32)


def test_umfang_ist_polygonumfang():  # messung.rs:10-31
    for n_ecken in (7, 57, 200):
        x, tri = prisma(n_ecken, 0.0925, 0.2, 8)
        for z in (-0.1037, -0.05, -0.1999):
            u = O.umfang_in_ebene(x, tri, [0.0, 0.0, z], [0.0, 0.0, 2.0])
            assert abs(u - 2.0 * n_ecken * 0.0925 * math.sin(math.pi / n_ecken)) < 1e-12
    # a tilted plane cuts an ellipse-like polygon: every crossing triangle counts once
    x, tri = prisma(4, 1.0, 1.0, 1)
    assert O.umfang_in_ebene(x, tri, [0.0, 0.0, 5.0], [0.0, 0.0, 1.0]) == 0.0


def test_nahtspalt_und_geschwindigkeit():  # messung.rs:57-64, 84-87, 113-116
    x = np.array([[0.0, 0.0, 0.0], [0.003, 0.004, 0.0], [1.0, 1.0, 1.0], [1.0, 1.0, 1.001]])
    assert O.nahtspalt(x, np.array([[0, 1], [2, 3]])) == 0.005
    assert O.nahtspalt(x, np.zeros((0, 2), dtype=np.uint32)) == 0.0
    v = np.tile([0.3, -0.4, 1.2], (4, 1))
    assert abs(O.mittlere_geschwindigkeit(v, np.array([1.0, 2.0, 3.0, 4.0])) - 1.3) < 1e-15
    v[0] = 0.0
    m = np.array([1.0, 1.0, 1.0, 1.0])
    assert abs(O.mittlere_geschwindigkeit(v, m) - 1.3 * math.sqrt(0.75)) < 1e-15
