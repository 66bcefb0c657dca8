"""cap::solve_dielectric on the device (gp_cap_solve_dielectric: interface
rows by k_dielectric_rows, dense FP64 LU) against the oracle restatement of
cap/dielectric.rs:47-151 on the same panels."""
import numpy as np
import pytest

import gpurify_amd as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu

EPS0 = 8.8541878128e-12


def coated_sphere(subdiv, a=1e-6, b=2e-6, er=4.0):
    vc = O.mesh_icosphere((0, 0, 0), a, subdiv)
    vd = O.mesh_icosphere((0, 0, 0), b, subdiv)
    n1, n2 = len(vc), len(vd)
    v = np.concatenate([vc, vd])
    role = np.r_[np.zeros(n1), np.ones(n2)].astype(int)
    cond = np.zeros(n1 + n2, dtype=np.uint32)
    eps = np.r_[np.full(n1, er), np.zeros(n2)]
    eo = np.r_[np.zeros(n1), np.ones(n2)]
    ei = np.r_[np.zeros(n1), np.full(n2, er)]
    return v, role, cond, eps, eo, ei, np.zeros((n1 + n2, 3))


def test_coated_sphere_matches_the_oracle(ctx):
    prob = coated_sphere(2)
    print(f"coated sphere: {len(prob[0])} panels", flush=True)
    Cg, rep = ctx.cap_solve_dielectric(*prob, 1)
    Co = O.solve_dielectric(*prob, 1)
    assert abs(Cg[0, 0] - Co[0, 0]) <= 1e-10 * Co[0, 0]
    exact = 4 * np.pi * EPS0 / ((1 / 4.0) * (1 / 1e-6 - 1 / 2e-6) + 1 / 2e-6)
    assert abs(Cg[0, 0] - exact) / exact < 0.05


def test_large_system_takes_the_blocked_lu(ctx):
    """2,560 panels: above the one-work-group LU, factored by rocSOLVER."""
    prob = coated_sphere(3)
    print(f"coated sphere: {len(prob[0])} panels (blocked LU)", flush=True)
    Cg, rep = ctx.cap_solve_dielectric(*prob, 1)
    Co = O.solve_dielectric(*prob, 1)
    assert abs(Cg[0, 0] - Co[0, 0]) <= 1e-10 * Co[0, 0]


def test_two_conductors_in_a_dielectric_box_match_the_oracle(ctx):
    # two spheres inside a dielectric cube (eps 3.9 inside, 1 outside), the
    # cube's faces as interface panels with their reference point inside
    s1 = O.mesh_icosphere((-1.5e-6, 0, 0), 1e-6, 1)
    s2 = O.mesh_icosphere((1.5e-6, 0, 0), 0.7e-6, 1)
    faces = []
    L, m = 3.5e-6, 4
    g = np.linspace(-L, L, m + 1)
    for ax in range(3):
        for side in (-L, L):
            for i in range(m):
                for j in range(m):
                    q = np.zeros((4, 3))
                    a0, a1 = [(ax + 1) % 3, (ax + 2) % 3]
                    for k, (u, w) in enumerate([(g[i], g[j]), (g[i + 1], g[j]),
                                                (g[i + 1], g[j + 1]), (g[i], g[j + 1])]):
                        q[k, ax], q[k, a0], q[k, a1] = side, u, w
                    faces.append(q)
    box = np.array(faces)
    v = np.concatenate([s1, s2, box])
    n1, n2, nb = len(s1), len(s2), len(box)
    role = np.r_[np.zeros(n1 + n2), np.ones(nb)].astype(int)
    cond = np.r_[np.zeros(n1), np.ones(n2), np.zeros(nb)].astype(np.uint32)
    eps = np.r_[np.full(n1 + n2, 3.9), np.zeros(nb)]
    eo = np.r_[np.zeros(n1 + n2), np.ones(nb)]
    ei = np.r_[np.zeros(n1 + n2), np.full(nb, 3.9)]
    ref = np.zeros((len(v), 3))
    print(f"dielectric box: {len(v)} panels", flush=True)
    Cg, _ = ctx.cap_solve_dielectric(v, role, cond, eps, eo, ei, ref, 2)
    Co = O.solve_dielectric(v, role, cond, eps, eo, ei, ref, 2)
    assert np.abs(Cg - Co).max() <= 1e-10 * np.abs(Co).max()
    assert Cg[0, 1] < 0 and Cg[0, 0] > 0 and abs(Cg[0, 1] - Cg[1, 0]) < 0.02 * abs(Cg[0, 1])


def test_interface_with_equal_permittivities_is_rejected(ctx):
    v, role, cond, eps, eo, ei, ref = coated_sphere(1)
    ei = np.where(role == 1, 1.0, ei)
    with pytest.raises(G.GpError, match="INVALID_INPUT"):
        ctx.cap_solve_dielectric(v, role, cond, eps, eo, ei, ref, 1)


def test_list_file_problem_on_the_device(ctx, tmp_path):
    """A FastCap list file (C conductor file + D interface file, translated)
    read by parse_lst (cap/listfile.rs:35-115) and solved on the device."""
    from gpurify_amd.fastcap import parse_lst, write_qui
    (tmp_path / "s.qui").write_text(write_qui(O.mesh_icosphere((0, 0, 0), 1e-6, 2),
                                              np.zeros(320, np.uint32)))
    (tmp_path / "b.qui").write_text(write_qui(O.mesh_icosphere((0, 0, 0), 2e-6, 2),
                                              np.zeros(320, np.uint32)))
    (tmp_path / "p.lst").write_text("C s.qui 4.0 1e-6 0 0\nD b.qui 1.0 4.0 0 0 0 1e-6 0 0\n")
    g = parse_lst(str(tmp_path / "p.lst"))
    args = (g["vertices"], g["role"], g["conductor"], g["eps"], g["eps_out"], g["eps_in"], g["ref"], 1)
    print(f"list file: {len(g['role'])} panels", flush=True)
    Cg, _ = ctx.cap_solve_dielectric(*args)
    Co = O.solve_dielectric(*args)
    assert abs(Cg[0, 0] - Co[0, 0]) <= 1e-10 * Co[0, 0]


@pytest.mark.parametrize("subdiv,bound", [(2, 1.2e-2), (3, 4e-3)])
def test_fmm_form_matches_the_oracle_and_the_dense_solve(ctx, subdiv, bound):
    """gp_cap_solve_dielectric_fmm: the interface rows on the FMM operator
    (near field exact, far field the normal derivative of the local
    expansion). The same operator in the oracle agrees to the solve
    tolerance; against the dense solve the far field's centroid lumping
    leaves 0.9 % at 640 panels and 0.24 % at 2,560 (measured on the oracle,
    independent of the Chebyshev order)."""
    prob = coated_sphere(subdiv)
    print(f"coated sphere, FMM form: {len(prob[0])} panels", flush=True)
    o = G.default_opts(method=G.GP_METHOD_FMM, fmm_order=4, max_level=3, tol=1e-12)
    Cf, rep = ctx.cap_solve_dielectric(*prob, 1, opts=o)
    assert rep["converged"] == 1 and rep["method_used"] == G.GP_METHOD_FMM
    assert rep["max_level"] == 3
    Cof = O.solve_dielectric_fmm(*prob, 1, order=4, max_level=3, tol=1e-12)
    assert abs(Cf[0, 0] - Cof[0, 0]) <= 1e-9 * Cof[0, 0]
    Cd, _ = ctx.cap_solve_dielectric(*prob, 1)
    assert abs(Cf[0, 0] - Cd[0, 0]) <= bound * Cd[0, 0]
    exact = 4 * np.pi * EPS0 / ((1 / 4.0) * (1 / 1e-6 - 1 / 2e-6) + 1 / 2e-6)
    assert abs(Cf[0, 0] - exact) / exact < 0.05  # dielectric.rs:44


def test_fmm_form_two_conductors_in_a_dielectric_box(ctx):
    s1 = O.mesh_icosphere((-1.5e-6, 0, 0), 1e-6, 2)
    s2 = O.mesh_icosphere((1.5e-6, 0, 0), 0.7e-6, 2)
    L, m = 3.5e-6, 8
    g = np.linspace(-L, L, m + 1)
    faces = []
    for ax in range(3):
        a0, a1 = (ax + 1) % 3, (ax + 2) % 3
        for side in (-L, L):
            for i in range(m):
                for j in range(m):
                    q = np.zeros((4, 3))
                    for k, (u, w) in enumerate([(g[i], g[j]), (g[i + 1], g[j]),
                                                (g[i + 1], g[j + 1]), (g[i], g[j + 1])]):
                        q[k, ax], q[k, a0], q[k, a1] = side, u, w
                    faces.append(q)
    box = np.array(faces)
    v = np.concatenate([s1, s2, box])
    n1, n2, nb = len(s1), len(s2), len(box)
    role = np.r_[np.zeros(n1 + n2), np.ones(nb)].astype(int)
    cond = np.r_[np.zeros(n1), np.ones(n2), np.zeros(nb)].astype(np.uint32)
    eps = np.r_[np.full(n1 + n2, 3.9), np.zeros(nb)]
    eo = np.r_[np.zeros(n1 + n2), np.ones(nb)]
    ei = np.r_[np.zeros(n1 + n2), np.full(nb, 3.9)]
    ref = np.zeros((len(v), 3))
    args = (v, role, cond, eps, eo, ei, ref, 2)
    print(f"dielectric box, FMM form: {len(v)} panels", flush=True)
    o = G.default_opts(method=G.GP_METHOD_FMM, fmm_order=4, max_level=3, tol=1e-12)
    Cf, rep = ctx.cap_solve_dielectric(*args, opts=o)
    assert rep["converged"] == 1 and rep["max_level"] == 3
    Cof = O.solve_dielectric_fmm(*args, order=4, max_level=3, tol=1e-12)
    assert np.abs(Cf - Cof).max() <= 1e-9 * np.abs(Cof).max()
    Cd, _ = ctx.cap_solve_dielectric(*args)
    assert np.abs(Cf - Cd).max() <= 1e-2 * np.abs(Cd).max()  # 2.8e-3 on the oracle


def test_fmm_form_rejects_mixed_precision(ctx):
    o = G.default_opts(method=G.GP_METHOD_FMM, precision=G.GP_PREC_MIXED)
    with pytest.raises(G.GpError, match="INVALID_INPUT"):
        ctx.cap_solve_dielectric(*coated_sphere(1), 1, opts=o)
