"""The grounded plane by images on the device (gp_cap_opts.ground =
GP_GROUND_IMAGE; no reference counterpart, the reference meshes a finite
substrate plate, bridge.rs:346-379) against the oracle's image solves, the
sphere-over-plane series and the finite-plate model it replaces."""
import math

import numpy as np
import pytest

import gpurify_amd as G
from gpurify_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu
EPS0 = 8.8541878128e-12


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def image_opts(zg, **kw):
    return G.default_opts(ground=G.GP_GROUND_IMAGE, ground_z=zg, **kw)


def two_spheres(subdiv=2):
    a = O.mesh_icosphere((0.0, 0.0, 1.2), 0.5, subdiv)
    b = O.mesh_icosphere((1.4, 0.3, 0.9), 0.4, subdiv)
    return np.concatenate([a, b]), np.r_[np.zeros(len(a)), np.ones(len(b))].astype(np.uint32)


def test_dense_image_matches_oracle(ctx):
    v, c = two_spheres()
    zg = -0.15
    Cd, rep = ctx.cap_solve(v, c, 2, image_opts(zg, method=G.GP_METHOD_DENSE))
    assert rep["method_used"] == G.GP_METHOD_DENSE
    assert rel(Cd, O.cap_dense_image(v, c, 2, zg)) < 1e-11
    # the plane raises every self capacitance over free space
    Cf, _ = ctx.cap_solve(v, c, 2, G.default_opts(method=G.GP_METHOD_DENSE))
    assert np.all(np.diag(Cd) > np.diag(Cf))


def test_sphere_over_plane_series(ctx):
    h = 2.0
    sph = O.mesh_icosphere((0, 0, h), 1.0, 3)  # 1,280 panels: the blocked LU path
    al = math.acosh(h)
    exact = 4 * math.pi * EPS0 * math.sinh(al) * sum(1 / math.sinh(n * al) for n in range(1, 40))
    z = np.zeros(len(sph), dtype=np.uint32)
    Cd, _ = ctx.cap_solve(sph, z, 1, image_opts(0.0, method=G.GP_METHOD_DENSE))
    assert abs(Cd[0, 0] / exact - 1) < 0.01
    assert rel(Cd, O.cap_dense_image(sph, z, 1, 0.0)) < 1e-11
    Cm, rm = ctx.cap_solve(sph, z, 1, image_opts(0.0, method=G.GP_METHOD_FMM, fmm_order=4,
                                                 max_level=3, tol=1e-10))
    assert rm["converged"] and abs(Cm[0, 0] / Cd[0, 0] - 1) < 5e-3


@pytest.mark.parametrize("prec", [G.GP_PREC_FP64, G.GP_PREC_MIXED])
def test_fmm_image_matches_oracle(ctx, prec):
    v, c = two_spheres(3)
    zg = -0.15
    o = image_opts(zg, method=G.GP_METHOD_FMM, fmm_order=4, max_level=3, tol=1e-10,
                   precision=prec)
    Cg, rep = ctx.cap_solve(v, c, 2, o)
    Co, _, _ = O.cap_fmm_image(v, c, 2, zg, order=4, max_level=3, tol=1e-10)
    assert rep["converged"] == 1
    assert rel(Cg, Co) < (1e-9 if prec == G.GP_PREC_FP64 else 1e-6)
    assert np.max(np.abs(Cg - Co) / np.abs(Co)) < 1e-6


def test_bridge_image_replaces_the_plate(ctx):
    """gp_solve_capacitance with the plane at the plate's height: the wires'
    Maxwell block equals the oracle's image solve of the wire panels, the
    plane's row and column hold minus the row and column sums, and the
    result sits close to the finite-plate model (0.15-0.2 % on the self
    terms: the plate's finite extent)."""
    groups = W.two_wires()
    o = G.default_opts(ground=G.GP_GROUND_IMAGE)
    for refine in (1, 2):  # 128 wire panels: dense LU; 512: the FMM(4, 4, 1e-8) branch
        caf, Ci, rep = ctx.solve_capacitance(groups, refine=refine, opts=o)
        v, c, e = O.mesh_groups(groups, refine=refine)
        real = c < 2
        zg = float(v[~real][0, 0, 2])
        assert rep["n_panels"] == int(real.sum())
        if rep["n_panels"] < 500:
            assert rep["method_used"] == G.GP_METHOD_DENSE
            Co = O.cap_dense_image(v[real], c[real], 2, zg) * e[0]
            assert rel(Ci[:2, :2], Co) < 1e-10
        else:
            assert rep["method_used"] == G.GP_METHOD_FMM
            Co = O.cap_fmm_image(v[real], c[real], 2, zg, order=4, max_level=4, tol=1e-8)[0] * e[0]
            assert np.max(np.abs(Ci[:2, :2] - Co) / np.abs(Co)) < 1e-6
            Cd = O.cap_dense_image(v[real], c[real], 2, zg) * e[0]
            assert rel(Ci[:2, :2], Cd) < 5e-3  # the FMM's own error
        assert np.allclose(Ci[:2, 2], -Ci[:2, :2].sum(axis=1), rtol=1e-12)
        assert np.allclose(Ci[2, :2], -Ci[:2, :2].sum(axis=0), rtol=1e-12)
        assert abs(Ci[2, 2] - Ci[:2, :2].sum()) < 1e-12 * abs(Ci[2, 2])
        cafp, Cp, _ = ctx.solve_capacitance(groups, refine=refine)
        for k in (0, 1):
            assert abs(caf[k] / cafp[k] - 1) < 5e-3 and caf[k] == Ci[k, k] * 1e18


def test_batched_image_equals_single_solves(ctx):
    batch = W.net_batch(n_nets=6, seed=11, pmin=150, pmax=3000)
    meshes = W.net_batch_meshes(batch, lambda g, r: G.mesh_groups(g, refine=r, max_panels=10 ** 7))
    # each net without its plate, over the plane z = 0 (the sky130 stack
    # starts at 300 nm)
    probs = []
    for v, c, _ in meshes:
        keep = c < c.max()
        probs.append((v[keep], c[keep], int(c[keep].max()) + 1))
    o = image_opts(0.0, tol=1e-11, max_level=0, leaf_target=32)
    Cb, reps = ctx.cap_solve_batched(probs, o)
    assert {r["method_used"] for r in reps} == {G.GP_METHOD_DENSE, G.GP_METHOD_FMM}
    for (v, c, nc), C1b, r in zip(probs, Cb, reps):
        C1, _ = ctx.cap_solve(v, c, nc, o)
        assert r["converged"] == 1 and rel(C1b, C1) < 1e-8
        if len(v) < 500:
            assert rel(C1b, O.cap_dense_image(v, c, nc, 0.0)) < 1e-11


def test_ground_is_validated(ctx):
    v, c = two_spheres(1)
    with pytest.raises(G.GpError) as e:  # a vertex below the plane
        ctx.cap_solve(v, c, 2, image_opts(0.8, method=G.GP_METHOD_DENSE))
    assert e.value.code == G._lib.GP_ERR_INVALID_INPUT
    with pytest.raises(G.GpError) as e:
        ctx.cap_solve(v, c, 2, G.default_opts(ground=7))
    assert e.value.code == G._lib.GP_ERR_INVALID_INPUT
    with pytest.raises(G.GpError) as e:  # no adjoint with a ground plane
        ctx.cap_jacobian(v, c, 2, [], image_opts(-1.0, method=G.GP_METHOD_DENSE))
    assert e.value.code == G._lib.GP_ERR_INVALID_INPUT
