"""Pins the oracle's ground plane by images (cap_dense_image, cap_fmm_image;
no reference counterpart: the reference meshes a finite substrate plate,
bridge.rs:346-379) against closed forms and independent routes:

* a sphere of radius a, centre at height h over a grounded plane: the
  textbook series C = 4 pi eps0 a sinh(alpha) sum_n 1 / sinh(n alpha),
  cosh(alpha) = h / a (a sign error in the image term lowers C below the
  isolated sphere's instead of raising it);
* the image system equals the plain dense solve of the panels and their
  mirrors as separate conductors, C_ij - C_ij* (linearity, through the
  solver pinned by the reference's own tests);
* a finite meshed plate under the sphere converges to the image result as
  it grows (the reference's substrate model in its limit);
* the FMM form agrees with the dense one to the FMM's error."""
import math

import numpy as np
import pytest

from oracle import oracle as O

EPS0 = 8.8541878128e-12


def sphere_over_plane(a, h):
    al = math.acosh(h / a)
    terms = [1 / math.sinh(n * al) for n in range(1, int(40 / al) + 2)]  # to ~e^-40
    return 4 * math.pi * EPS0 * a * math.sinh(al) * sum(terms)


@pytest.mark.parametrize("h", [1.5, 2.0, 4.0])
def test_sphere_over_plane_series(h):
    sph = O.mesh_icosphere((0.3, -0.2, h), 1.0, 3)
    C = O.cap_dense_image(sph, np.zeros(len(sph)), 1, 0.0)[0, 0]
    exact = sphere_over_plane(1.0, h)
    assert abs(C / exact - 1) < 0.01
    # the isolated sphere of the same mesh is below it: the plane raises C
    assert C > O.cap_dense(sph, np.zeros(len(sph)), 1)[0, 0] * (1 + 0.5 / h - 0.1 / h ** 2)


def test_image_equals_mirrored_conductors():
    rng = np.random.default_rng(3)
    a = O.mesh_icosphere((0.0, 0.0, 1.2), 0.5, 2)
    b = O.mesh_icosphere((1.4, 0.3, 0.9), 0.4, 1)
    v = np.concatenate([a, b])
    c = np.r_[np.zeros(len(a)), np.ones(len(b))].astype(np.uint32)
    zg = float(rng.uniform(-0.3, 0.2))
    Ci = O.cap_dense_image(v, c, 2, zg)
    m = v.copy()
    m[:, :, 2] = 2 * zg - m[:, :, 2]
    C4 = O.cap_dense(np.concatenate([v, m]), np.r_[c, c + 2].astype(np.uint32), 4)
    assert np.abs(Ci - (C4[:2, :2] - C4[:2, 2:])).max() < 1e-12 * np.abs(Ci).max()
    # Maxwell structure with a grounded plane: positive diagonal, negative
    # coupling, each row sums to the (positive) capacitance to the plane
    assert Ci[0, 0] > 0 and Ci[0, 1] < 0 and np.all(Ci.sum(axis=1) > 0)
    # (collocation makes C symmetric only to the discretisation error)
    assert abs(Ci[0, 1] - Ci[1, 0]) < 1e-3 * abs(Ci[0, 1])


def test_finite_plate_converges_to_the_plane():
    sph = O.mesh_icosphere((0, 0, 2.0), 1.0, 2)
    Ci = O.cap_dense_image(sph, np.zeros(len(sph)), 1, 0.0)[0, 0]
    err = []
    for side, nx in ((8, 16), (16, 24), (24, 32)):
        pl = O.mesh_plate(0, 0, 0, side, side, nx, nx)
        C = O.cap_dense(np.concatenate([sph, pl]),
                        np.r_[np.zeros(len(sph)), np.ones(len(pl))].astype(np.uint32), 2)
        err.append(abs(C[0, 0] / Ci - 1))
    assert err[0] > err[1] > err[2] and err[2] < 1e-3


def test_fmm_image_follows_dense():
    sph = O.mesh_icosphere((0, 0, 1.5), 1.0, 3)
    z = np.zeros(len(sph))
    Cd = O.cap_dense_image(sph, z, 1, 0.0)[0, 0]
    Cf, it, rs = O.cap_fmm_image(sph, z, 1, 0.0, order=4, max_level=3, tol=1e-10)
    assert rs[0] <= 1e-10 and abs(Cf[0, 0] / Cd - 1) < 5e-3


def test_panels_must_lie_above_the_plane():
    sph = O.mesh_icosphere((0, 0, 0.5), 1.0, 1)
    with pytest.raises(RuntimeError):
        O.cap_dense_image(sph, np.zeros(len(sph)), 1, 0.0)
    with pytest.raises(RuntimeError):
        O.cap_fmm_image(sph, np.zeros(len(sph)), 1, 0.0)
