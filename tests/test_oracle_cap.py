"""Pins the oracle's capacitance solvers and bridge mesher
(cap/solver.rs, cap/fmm_solver.rs, linalg.rs, krylov.rs, bridge.rs) with the
reference's validation tests (crates/pex/tests/cap_validation.rs, fmm_cap.rs)
and invariants (scale covariance, translation, Maxwell symmetry)."""
import math

import numpy as np
import pytest

from gpurify_amd import workloads as W
from oracle import oracle as O

EPS0 = 8.8541878128e-12
FOUR_PI_EPS0 = 4 * math.pi * EPS0


def sphere(subdiv, r=1.0):
    return O.mesh_icosphere((0, 0, 0), r, subdiv)


def test_icosphere_counts_and_radius():
    for s, n in ((0, 20), (1, 80), (2, 320)):
        v = sphere(s)
        assert len(v) == n
        assert np.allclose(np.linalg.norm(v[:, :3], axis=2), 1.0)


def test_isolated_sphere_self_capacitance():
    # cap_validation.rs:9-28
    v = sphere(2)
    C = O.cap_dense(v, np.zeros(len(v)), 1)
    assert abs(C[0, 0] - FOUR_PI_EPS0) / FOUR_PI_EPS0 < 0.03


def test_sphere_converges_with_refinement():
    # cap_validation.rs:30-45
    err = lambda s: abs(O.cap_dense(sphere(s), np.zeros(len(sphere(s))), 1)[0, 0] - FOUR_PI_EPS0)
    assert err(2) < err(1)


def test_parallel_plate_capacitor():
    # cap_validation.rs:47-77
    side, gap = 0.2, 0.02
    v = np.concatenate([O.mesh_plate(0, 0, 0, side, side, 10, 10), O.mesh_plate(0, 0, gap, side, side, 10, 10)])
    cond = np.repeat([0, 1], 100)
    C = O.cap_dense(v, cond, 2)
    ratio = abs(C[0, 1]) / (EPS0 * side * side / gap)
    assert 0.95 < ratio < 1.35
    assert abs(C[0, 1] - C[1, 0]) / abs(C[0, 1]) < 1e-6
    assert C[0, 1] < 0 < C[0, 0]


def test_direct_and_iterative_agree():
    # cap_validation.rs:79-90
    v = sphere(1)
    cd = O.cap_dense(v, np.zeros(len(v)), 1, method=0)[0, 0]
    ci = O.cap_dense(v, np.zeros(len(v)), 1, method=1)[0, 0]
    assert abs(cd - ci) / cd < 1e-6


def test_fmm_capacitance_matches_dense_and_analytic():
    # crates/pex/tests/fmm_cap.rs:7-27
    v = sphere(2)
    cd = O.cap_dense(v, np.zeros(len(v)), 1)[0, 0]
    cf = O.cap_fmm(v, np.zeros(len(v)), 1, order=4, max_level=3, tol=1e-8)[0][0, 0]
    assert abs(cf - cd) / cd < 5e-3
    assert abs(cf - FOUR_PI_EPS0) / FOUR_PI_EPS0 < 0.03


def test_lu_matches_numpy_and_detects_singular():
    rng = np.random.default_rng(0)
    A = rng.normal(size=(60, 60)) + 8 * np.eye(60)
    B = rng.normal(size=(3, 60))  # column-major RHS blocks
    X = B.copy()
    assert O.lib().or_lu_solve(60, O._p(np.ascontiguousarray(A)), 3, O._p(X)) == 0
    assert np.allclose(A @ X.T, B.T, rtol=1e-12, atol=1e-12)
    S = np.zeros((4, 4))
    assert O.lib().or_lu_solve(4, O._p(S), 1, O._p(np.zeros(4))) == -1


def test_dense_solve_equals_numpy_solve_of_oracle_P():
    groups = W.two_wires()
    v, c, _ = O.mesh_groups(groups)
    C, P = O.cap_dense(v, c, 3, want_p=True)
    # P is symmetric-ish but not symmetric; check against numpy on the same P
    _, ar, _ = O.panel_props(v)
    B = np.stack([(c == j).astype(float) for j in range(3)], 1)
    S = np.linalg.solve(P, B)
    Cn = np.stack([[np.sum(S[c == i, j] * ar[c == i]) for j in range(3)] for i in range(3)])
    assert np.allclose(C, Cn, rtol=1e-11, atol=0)
    # P entries: diagonal of a rectangle equals the closed form
    from tests.test_oracle_wilton import rect_closed_form  # noqa
    k_e = 1 / FOUR_PI_EPS0
    x1, y1 = v[0, 0, :2]
    x2, y2 = v[0, 2, :2]
    cen, _, _ = O.panel_props(v[:1])
    lo, hi = min(x1, x2), max(x1, x2)
    ylo, yhi = min(y1, y2), max(y1, y2)
    ref = k_e * rect_closed_form(lo, hi, ylo, yhi, [cen[0, 0], cen[0, 1], 0.0])
    assert P[0, 0] == pytest.approx(ref, rel=1e-12)


def test_scale_translation_and_maxwell_structure():
    groups = W.two_wires()
    v, c, _ = O.mesh_groups(groups)
    C = O.cap_dense(v, c, 3)
    lam, t = 3.7, np.array([1e-5, -2e-5, 3e-6])
    C2 = O.cap_dense(v * lam, c, 3)
    C3 = O.cap_dense(v + t, c, 3)
    assert np.allclose(C2, lam * C, rtol=1e-11, atol=0)
    assert np.allclose(C3, C, rtol=1e-9, atol=0)
    assert np.all(np.diag(C) > 0) and np.all(C[~np.eye(3, dtype=bool)] < 0)
    assert np.allclose(C, C.T, rtol=2e-2)  # collocation is only approximately symmetric
    assert np.allclose(C[0, 0], C[1, 1], rtol=1e-9)  # mirror-symmetric wires


def test_bridge_panel_counts_and_closed_surfaces():
    groups = W.two_wires()
    for r, n in ((1, 173), (3, 1557), (4, 2768)):
        v, c, e = O.mesh_groups(groups, refine=r)
        assert len(v) == n
        _, ar, nrm = O.panel_props(v)
        # each wire is a closed box: area = 2(wl + wt + lt), normals sum to 0
        for k in (0, 1):
            w, l, t = 140e-9, 10000e-9, 200e-9
            assert ar[c == k].sum() == pytest.approx(2 * (w * l + w * t + l * t), rel=1e-12)
            assert np.allclose((nrm[c == k] * ar[c == k, None]).sum(0), 0, atol=1e-24)
        assert np.all(e == 3.9)


def test_process_stack_heuristic():
    t, h, k = O.process_stack(np.zeros(7), np.zeros(7), np.zeros(7))
    assert list(h) == [0, 300, 600, 900, 1200, 1500, 1800] and np.all(t == 200) and np.all(k == 3.9)
    st = W.sky130_stack()
    assert st["met1"] == (600.0, 200.0, 3.9) and st["met2"] == (900.0, 200.0, 3.9)
    # explicit values pin the layer (process_stack.rs:53-61)
    t, h, k = O.process_stack([0, 350, 0], [0, 700, 0], [0, 4.2, 0])
    assert list(h) == [0, 700, 1150] and list(t) == [200, 350, 200] and list(k) == [3.9, 4.2, 3.9]


def test_bridge_eps_scaling_and_cap_af():
    groups = W.two_wires()
    caf, C, n = O.solve_capacitance(groups)
    g1 = [dict(g, k=1.0) for g in groups]
    caf1, C1, _ = O.solve_capacitance(g1)
    assert np.allclose(C, 3.9 * C1, rtol=1e-14)
    assert caf[0] == pytest.approx(C[0, 0] * 1e18, rel=1e-15)


def test_bridge_rules_and_errors():
    groups = W.two_wires()
    # reference rule: < 500 panels dense, else FMM(4,4,1e-8)
    caf_d, Cd, n1 = O.solve_capacitance(groups, refine=1)
    caf_f, Cf, n3 = O.solve_capacitance(groups, refine=3)
    assert n1 == 173 and n3 == 1557
    assert abs(caf_f[0] - caf_d[0]) / caf_d[0] < 0.05  # refinement changes C by a few %
    with pytest.raises(RuntimeError, match="-3"):
        O.solve_capacitance(groups, refine=4, max_panels=2000)
    big = [dict(net=0, height_nm=600, thickness_nm=200, k=3.9,
                rects=[[i * 3, j * 3, i * 3 + 2, j * 3 + 2] for i in range(300) for j in range(300)])]
    with pytest.raises(RuntimeError, match="-2"):
        O.mesh_groups(big)
