"""The oracle's restatement of cap::solve_dielectric (cap/dielectric.rs:
47-151, field_integral of integrals/panel.rs:213-253) pinned by the
reference's own tests (crates/pex/tests/dielectric.rs:8-59): the
dielectric-coated sphere against its closed form and the vacuum limit, plus
mesh convergence and the orientation rule of oriented_normal."""
import numpy as np

from oracle import oracle as O

EPS0 = 8.8541878128e-12


def coated_sphere(subdiv, a=1.0, b=2.0, er=4.0, flip=False):
    vc = O.mesh_icosphere((0, 0, 0), a, subdiv)
    vd = O.mesh_icosphere((0, 0, 0), b, subdiv)
    if flip:  # reversed vertex order: the interface normal points inward first
        vd = vd[:, [2, 1, 0, 3]] if np.all(vd[:, 2] == vd[:, 3]) else vd[:, ::-1]
    n1, n2 = len(vc), len(vd)
    v = np.concatenate([vc, vd])
    role = np.r_[np.zeros(n1), np.ones(n2)].astype(int)
    cond = np.zeros(n1 + n2, dtype=np.uint32)
    eps = np.r_[np.full(n1, er), np.zeros(n2)]
    eo = np.r_[np.zeros(n1), np.ones(n2)]
    ei = np.r_[np.zeros(n1), np.full(n2, er)]
    return v, role, cond, eps, eo, ei, np.zeros((n1 + n2, 3))


def exact(a=1.0, b=2.0, er=4.0):
    return 4 * np.pi * EPS0 / ((1 / er) * (1 / a - 1 / b) + 1 / b)


def test_coated_sphere_closed_form_and_convergence():
    e2 = abs(O.solve_dielectric(*coated_sphere(2), 1)[0, 0] - exact()) / exact()
    e3 = abs(O.solve_dielectric(*coated_sphere(3), 1)[0, 0] - exact()) / exact()
    assert e2 < 0.05 and e3 < 0.05  # dielectric.rs:44 (the reference's own bound)
    assert e3 < e2  # finer mesh closer (4.9 % -> 3.3 %; flat panels inside the spheres)


def test_vacuum_limit_equals_the_conductor_solve():
    vc = O.mesh_icosphere((0, 0, 0), 1.0, 2)
    n = len(vc)
    z = np.zeros(n)
    C = O.solve_dielectric(vc, np.zeros(n, int), np.zeros(n, np.uint32), np.ones(n), z, z,
                           np.zeros((n, 3)), 1)
    Cd = O.cap_dense(vc, np.zeros(n, np.uint32), 1)
    assert abs(C[0, 0] - Cd[0, 0]) <= 1e-13 * Cd[0, 0]
    assert abs(C[0, 0] - 4 * np.pi * EPS0) / (4 * np.pi * EPS0) < 0.03   # dielectric.rs:58


def test_interface_orientation_follows_the_reference_point():
    C = O.solve_dielectric(*coated_sphere(2), 1)
    Cf = O.solve_dielectric(*coated_sphere(2, flip=True), 1)
    assert abs(C[0, 0] - Cf[0, 0]) <= 1e-12 * C[0, 0]


def test_interpolation_derivative_weights_are_exact_below_the_order():
    # d/dx of sum_m w_m(x) p(x_m) is p'(x) for every polynomial of degree < n
    rng = np.random.default_rng(5)
    for n in (3, 4, 6, 8):
        xm = O.cheb_nodes(n)
        c = rng.standard_normal(n)
        p = np.polynomial.Polynomial(c)
        for x in (-1.0, -0.37, 0.0, 0.81, 1.0):
            dw = O.interp_dweights(x, n)
            assert abs(dw @ p(xm) - p.deriv()(x)) <= 1e-11 * np.abs(c).sum() * n * n
            # and the derivative of the interpolation weights themselves
            h = 1e-6
            fd = (O.interp_weights(x + h, n) - O.interp_weights(x - h, n)) / (2 * h)
            assert np.abs(dw - fd).max() < 1e-6 * max(1.0, np.abs(dw).max())


def test_fmm_form_follows_the_dense_solve():
    # the interface rows on the FMM operator: near field as the dense rows,
    # far field the normal derivative of the local expansion
    p = coated_sphere(2)
    Cd = O.solve_dielectric(*p, 1)[0, 0]
    Cf = O.solve_dielectric_fmm(*p, 1, order=4, max_level=2, tol=1e-12)[0, 0]
    assert abs(Cf - Cd) <= 5e-3 * Cd  # 3.2e-3: centroid lumping of the far field
    assert abs(Cf - exact()) <= 0.05 * exact()
    Cff = O.solve_dielectric_fmm(*coated_sphere(2, flip=True), 1, order=4, max_level=2, tol=1e-12)
    assert abs(Cff[0, 0] - Cf) <= 1e-10 * Cf
