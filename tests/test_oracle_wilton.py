"""Pins the oracle's Wilton panel integral (integrals/panel.rs:147-208).

Checks: the reference's own known answers (crates/pex/tests/analytic.rs:14-71),
the textbook closed form for an axis-aligned rectangle (independent of the
edge decomposition), adaptive quadrature on tiny panels, additivity and
orientation independence.
"""
import math

import numpy as np
import pytest
from scipy import integrate

from oracle import oracle as O

UNIT_SQ = [[-0.5, -0.5, 0.0], [0.5, -0.5, 0.0], [0.5, 0.5, 0.0], [-0.5, 0.5, 0.0]]


def test_unit_square_self_integral():
    # analytic.rs:14-27: 4 asinh(1) at the centre, 1e-12 relative
    got = O.potential_integral(UNIT_SQ, [0, 0, 0])
    assert abs(got - 4.0 * math.asinh(1.0)) <= 1e-12 * 4.0 * math.asinh(1.0)


def test_far_field_is_monopole():
    # analytic.rs:29-42
    got = O.potential_integral(UNIT_SQ, [0, 0, 50.0])
    assert abs(got - 1 / 50.0) <= 1e-3 / 50.0


def test_offplane_matches_midpoint_quadrature():
    # analytic.rs:44-71: 400 x 280 midpoint rule on [0,1]x[0,0.7], 5e-4
    verts = [[0, 0, 0], [1, 0, 0], [1, 0.7, 0], [0, 0.7, 0]]
    p = np.array([0.3, 0.2, 0.4])
    got = O.potential_integral(verts, p)
    nx, ny = 400, 280
    dx, dy = 1.0 / nx, 0.7 / ny
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dy
    X, Y = np.meshgrid(x, y, indexing="ij")
    q = np.sum(dx * dy / np.sqrt((X - p[0]) ** 2 + (Y - p[1]) ** 2 + p[2] ** 2))
    assert abs(got - q) <= 5e-4 * abs(q)


def rect_closed_form(x1, x2, y1, y2, p):
    """Integral of 1/R over [x1,x2]x[y1,y2] (z=0) seen from p: the textbook
    corner sum of F(u,v,w) = u asinh(v/sqrt(u^2+w^2)) + v asinh(u/sqrt(v^2+w^2))
    - w atan(u v / (w R)), R = sqrt(u^2+v^2+w^2)."""
    w = p[2]

    def F(u, v):
        R = math.sqrt(u * u + v * v + w * w)
        a = u * math.asinh(v / math.sqrt(u * u + w * w)) if (u * u + w * w) > 0 else 0.0
        b = v * math.asinh(u / math.sqrt(v * v + w * w)) if (v * v + w * w) > 0 else 0.0
        c = w * math.atan(u * v / (w * R)) if w != 0 else 0.0
        return a + b - c

    s = 0.0
    for xs, sx in ((x2, 1), (x1, -1)):
        for ys, sy in ((y2, 1), (y1, -1)):
            s += sx * sy * F(xs - p[0], ys - p[1])
    return s


@pytest.mark.parametrize("seed", range(6))
def test_rectangle_matches_corner_formula(seed):
    rng = np.random.default_rng(seed)
    for _ in range(50):
        x1, y1 = rng.uniform(-2, 2, 2)
        x2, y2 = x1 + rng.uniform(0.01, 3), y1 + rng.uniform(0.01, 3)
        kind = rng.integers(3)
        if kind == 0:   # off-plane anywhere
            p = [rng.uniform(-4, 4), rng.uniform(-4, 4), rng.uniform(-2, 2)]
        elif kind == 1:  # in-plane inside the panel (self-type)
            p = [rng.uniform(x1, x2), rng.uniform(y1, y2), 0.0]
        else:           # in-plane outside the panel
            p = [x2 + rng.uniform(0.01, 2), rng.uniform(-4, 4), 0.0]
        verts = [[x1, y1, 0], [x2, y1, 0], [x2, y2, 0], [x1, y2, 0]]
        got = O.potential_integral(verts, p)
        ref = rect_closed_form(x1, x2, y1, y2, p)
        assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref)), (verts, p, got, ref)


def test_orientation_and_plane_independence():
    rng = np.random.default_rng(7)
    for _ in range(30):
        verts = np.array([[0, 0, 0], [1.3, 0, 0], [1.3, 0.4, 0], [0, 0.4, 0]], dtype=float)
        p = rng.uniform(-1, 2, 3)
        base = O.potential_integral(verts, p)
        # reversed vertex order
        assert abs(O.potential_integral(verts[::-1], p) - base) <= 1e-13 * abs(base)
        # rigid rotation of panel and point
        A = np.linalg.qr(rng.normal(size=(3, 3)))[0]
        t = rng.normal(size=3)
        got = O.potential_integral(verts @ A.T + t, A @ p + t)
        assert abs(got - base) <= 1e-12 * abs(base)


def test_additivity_quad_is_two_triangles():
    rng = np.random.default_rng(11)
    for _ in range(30):
        q = np.array([[0, 0, 0], [1.0, 0.1, 0], [1.2, 0.9, 0], [-0.1, 0.7, 0]])
        p = rng.uniform(-1, 2, 3)
        t1 = [q[0], q[1], q[2]]
        t2 = [q[0], q[2], q[3]]
        whole = O.potential_integral(q, p)
        parts = O.potential_integral(t1, p) + O.potential_integral(t2, p)
        assert abs(whole - parts) <= 1e-12 * abs(whole)


def test_triangle_matches_adaptive_quadrature():
    rng = np.random.default_rng(3)
    for _ in range(5):
        a, b, c = rng.uniform(-1, 1, (3, 3))
        n = np.cross(b - a, c - a)
        p = (a + b + c) / 3 + 0.3 * n / np.linalg.norm(n) + rng.uniform(-0.2, 0.2, 3)

        def f(v, u):
            r = a + u * (b - a) + v * (c - a)
            return 1.0 / np.linalg.norm(r - p)

        val, _ = integrate.dblquad(f, 0, 1, 0, lambda u: 1 - u, epsabs=1e-13, epsrel=1e-12)
        val *= np.linalg.norm(n)  # Jacobian = 2 * area
        got = O.potential_integral([a, b, c], p)
        assert abs(got - val) <= 1e-9 * abs(val)


def test_triangle_repeated_vertex_equals_triangle():
    a, b, c = np.array([0, 0, 0.0]), np.array([1, 0.2, 0.1]), np.array([0.3, 0.9, -0.2])
    p = np.array([0.4, 0.3, 0.5])
    assert O.potential_integral([a, b, c, c], p) == pytest.approx(O.potential_integral([a, b, c], p), rel=1e-15)


def test_panel_props_closed_forms():
    # centroid/area of a rectangle and of a triangle given as a quad (panel.rs:29-67)
    v = np.array([[[0, 0, 0], [2, 0, 0], [2, 1, 0], [0, 1, 0]],
                  [[0, 0, 1], [3, 0, 1], [0, 3, 1], [0, 3, 1]]], dtype=float)
    cen, ar, nrm = O.panel_props(v)
    assert np.allclose(cen[0], [1, 0.5, 0]) and ar[0] == pytest.approx(2.0)
    assert np.allclose(cen[1], [1, 1, 1]) and ar[1] == pytest.approx(4.5)
    assert np.allclose(nrm[0], [0, 0, 1])
