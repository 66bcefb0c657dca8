"""The committed golden fixtures of tests/golden/ (written by
tests/golden/make_golden.py): the oracle and the product mesher reproduce them
on the CPU, the device reproduces them on the GPU.

Expected values come from closed forms where one exists (Wilton known answers,
tests/analytic.rs:14-42 and the rectangle corner formula) and from the oracle
otherwise; the oracle-made ones are pinned here against independent checks
too (numpy brute force for the dense point sum, the reference's FMM error
bounds of crates/pex/tests/fmm.rs:37-60, cap_validation.rs closed forms)."""
import hashlib
import json
import math
import os

import numpy as np
import pytest

from gpurify_amd import workloads as W

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EPS0 = 8.8541878128e-12


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def load_json(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.abs(a - b).max() / np.abs(b).max())


def panel_hash(v, c):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(v, dtype="<f8").tobytes())
    h.update(np.ascontiguousarray(c, dtype="<u4").tobytes())
    return h.hexdigest()


# --------------------------------------------------------------------- CPU


def test_wilton_known_answers():
    from oracle import oracle as O

    cases = load_json("wilton_kat.json")
    assert abs(cases[0]["expected"] - 4 * math.asinh(1.0)) < 1e-15
    for c in cases:
        got = O.potential_integral(np.array(c["verts"], float), np.array(c["p"], float))
        assert abs(got - c["expected"]) <= 1e-12 * abs(c["expected"]), c["what"]
        if "monopole" in c:  # analytic.rs:29-42: the far field of a unit-area panel
            assert abs(got - c["monopole"]) / c["monopole"] < 1e-6


def test_c1_reference_mesh_is_reproduced_bit_for_bit():
    """The product mesher (host C++) and the oracle mesher both give the
    fixture's 173 panels, exactly."""
    import gpurify_amd as G
    from oracle import oracle as O

    g = load("c1_r1.npz")
    for mesh in (G.mesh_groups, O.mesh_groups):
        v, c, e = mesh(W.two_wires(), refine=1)
        assert np.array_equal(v, g["verts"]) and np.array_equal(c, g["cond"])
        assert np.array_equal(e, g["eps"])
    gv3, gc3, _ = G.mesh_groups(W.two_wires(), refine=3)
    assert panel_hash(gv3, gc3) == str(load("c1_r3.npz")["sha256"])


def test_c1_oracle_reproduces_the_fixtures():
    from oracle import oracle as O

    g = load("c1_r1.npz")
    caf, Cm, n = O.solve_capacitance(W.two_wires(), refine=1)
    assert n == 173 and rel(Cm, g["C"]) < 1e-12
    assert rel([caf[0], caf[1]], g["cap_af"]) < 1e-12
    # Maxwell structure of the fixture: symmetric off-diagonal block,
    # positive diagonal, negative couplings, rows summing to >= 0
    C = g["C"]
    assert C[0, 1] < 0 and C[0, 0] > 0 and abs(C[0, 1] - C[1, 0]) < 1e-12 * C[0, 0]
    assert np.all(C[:2].sum(axis=1) > 0)
    g3 = load("c1_r3.npz")
    caf3, C3, n3 = O.solve_capacitance(W.two_wires(), refine=3)
    assert n3 == int(g3["n_panels"]) and rel(C3, g3["C_fmm"]) < 1e-12
    # the FMM branch (bridge.rs:403) sits within the FMM's own error of the
    # dense LU on the same panels (fmm_cap.rs:7-27 bounds it by 5e-3)
    assert rel(g3["C_fmm"], g3["C_dense"]) < 5e-3


def test_lcg_fixture_dense_sum_and_fmm_bounds():
    from oracle import oracle as O

    g = load("lcg_fmm.npz")
    for n, seed in ((900, 12345), (800, 999)):
        P = g[f"pts{n}"]
        assert np.array_equal(O.lcg_points(n, seed), P)
        q = g[f"q{n}"]
        D = np.linalg.norm(P[:, None, :] - P[None, :, :], axis=2)
        np.fill_diagonal(D, np.inf)
        brute = (q[None, :] / D).sum(axis=1)  # numpy, independent of the oracle
        assert rel(g[f"dense{n}"], brute) < 1e-13
    err = lambda k, n: (np.linalg.norm(g[k] - g[f"dense{n}"]) / np.linalg.norm(g[f"dense{n}"]))
    assert err("fmm900_n5_L3", 900) < 1e-3  # fmm.rs:37-47
    assert err("fmm800_n5_L3", 800) < err("fmm800_n3_L3", 800) and err("fmm800_n5_L3", 800) < 1e-4
    for k in g.files:  # the oracle still produces every stored matvec
        if k.startswith("fmm"):
            n = int(k[3:6])
            order, lvl = int(k.split("_n")[1][0]), int(k.split("_L")[1])
            y = O.fmm_matvec(g[f"pts{n}"], g[f"q{n}"], order, lvl)
            assert rel(y, g[k]) < 1e-12, k


def test_sphere_and_plate_fixtures():
    from oracle import oracle as O

    s = load_json("sphere_plate.json")
    four_pi_eps0 = 4 * math.pi * EPS0
    assert abs(s["sphere_C_dense"] - four_pi_eps0) / four_pi_eps0 < 0.03  # cap_validation.rs:9-28
    assert abs(s["sphere_C_fmm_order4_depth3"] - s["sphere_C_dense"]) / s["sphere_C_dense"] < 5e-3
    sph = O.mesh_icosphere((0, 0, 0), 1.0, 2)
    assert rel(O.cap_dense(sph, np.zeros(len(sph)), 1)[0, 0], s["sphere_C_dense"]) < 1e-12
    Cp = np.array(s["plate_C_dense"])
    ratio = abs(Cp[0, 1]) / (EPS0 * s["plate_side"] ** 2 / s["plate_gap"])
    assert 0.95 < ratio < 1.35  # cap_validation.rs:47-77


# --------------------------------------------------------------------- GPU


@pytest.mark.gpu
def test_device_dense_solve_reproduces_c1(ctx):
    import gpurify_amd as G

    g = load("c1_r1.npz")
    Cd, rep = ctx.cap_solve(g["verts"], g["cond"], 3, G.default_opts(method=G.GP_METHOD_DENSE))
    assert rel(Cd * g["eps"][0], g["C"]) < 1e-11
    caf, Cm, rep = ctx.solve_capacitance(W.two_wires(), refine=1)
    assert rep["n_panels"] == 173 and rel(Cm, g["C"]) < 1e-11
    assert rel([caf[0], caf[1]], g["cap_af"]) < 1e-11


@pytest.mark.gpu
def test_device_fmm_branch_reproduces_c1_refine3(ctx):
    """bridge.rs:403: FMM(4, 4, 1e-8) with the CPQR-truncated M2L on 1,557
    panels; cap_af within 1e-6 of the fixture (the GMRES stopping point and
    the Gram-Schmidt variant may move it by the solve tolerance)."""
    g = load("c1_r3.npz")
    caf, Cm, rep = ctx.solve_capacitance(W.two_wires(), refine=3)
    assert rep["n_panels"] == int(g["n_panels"])
    assert rel([caf[0], caf[1]], g["cap_af_fmm"]) < 1e-6
    assert rel(Cm, g["C_fmm"]) < 1e-6


@pytest.mark.gpu
def test_device_reproduces_sphere_and_plate(ctx):
    import gpurify_amd as G
    from oracle import oracle as O

    s = load_json("sphere_plate.json")
    sph = O.mesh_icosphere((0, 0, 0), 1.0, 2)
    z = np.zeros(len(sph), dtype=np.uint32)
    Cd, _ = ctx.cap_solve(sph, z, 1, G.default_opts(method=G.GP_METHOD_DENSE))
    assert rel(Cd[0, 0], s["sphere_C_dense"]) < 1e-11
    Cf, rf = ctx.cap_solve(sph, z, 1, G.default_opts(method=G.GP_METHOD_FMM, fmm_order=4,
                                                     max_level=3, tol=1e-8))
    assert rf["converged"] and rel(Cf[0, 0], s["sphere_C_fmm_order4_depth3"]) < 1e-6
    pl = np.concatenate([O.mesh_plate(0, 0, 0, s["plate_side"], s["plate_side"], 10, 10),
                         O.mesh_plate(0, 0, s["plate_gap"], s["plate_side"], s["plate_side"],
                                      10, 10)])
    Cp, _ = ctx.cap_solve(pl, np.repeat([0, 1], 100).astype(np.uint32), 2,
                          G.default_opts(method=G.GP_METHOD_DENSE))
    assert rel(Cp, s["plate_C_dense"]) < 1e-11
