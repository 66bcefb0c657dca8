"""The drop-in bridge entry (gp_solve_capacitance, bridge.rs:307-443) on the
device: the empty layout, layered-k decks (the conductor-role dielectric
solve, bridge.rs:409-430) through the FMM operator, and the fused leaf kernel
against its two-launch form."""
import numpy as np
import pytest

import gpurify_amd as G
from gpurify_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def relmax(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_no_groups_returns_no_nets(ctx):
    # bridge.rs:311-313: an empty group map is Ok with nothing added
    caf, C, rep = ctx.solve_capacitance([])
    assert caf == {}


def layered_wires(k1=3.9, k2=7.0):
    """The two C1 wires, the second moved up to met2 with its own k."""
    g = W.two_wires()
    st = W.sky130_stack()
    g[1] = dict(g[1], height_nm=st["met2"][0], thickness_nm=st["met2"][1], k=k2)
    g[0] = dict(g[0], k=k1)
    return g


def test_layered_k_reference_rule_and_fmm(ctx):
    groups = layered_wires()
    # the reference refuses 500 panels or more on a layered-k deck
    with pytest.raises(G.GpError) as ei:
        ctx.solve_capacitance(groups, refine=3)
    assert ei.value.code == G._lib.GP_ERR_PANEL_CAPACITY
    # below the limit: the dense conductor-role solve equals the oracle's
    caf_d, Cd, _ = ctx.solve_capacitance(groups)
    caf_o, Co, _ = O.solve_capacitance(groups)
    assert relmax(Cd, Co) < 1e-11
    # GP_METHOD_FMM lifts the limit: the same potential coefficients, the
    # charges weighted by each panel's permittivity.  Every conductor here has
    # one permittivity (net 0: k1, net 1: k2, the plate: the lowest group's),
    # so C = diag(eps) C_vacuum of the oracle's FMM solve on the same panels.
    o = G.default_opts(method=G.GP_METHOD_FMM, tol=1e-11)
    caf, C, rep = ctx.solve_capacitance(groups, refine=3, opts=o)
    assert rep["method_used"] == G.GP_METHOD_FMM and rep["n_panels"] > 500
    v, c, eps = G.mesh_groups(groups, refine=3)
    Cv, _, _ = O.cap_fmm(v, c, 3, order=4, max_level=4, tol=1e-11, m2l_mode=0)
    e_cond = np.array([eps[c == i][0] for i in range(3)])
    assert all(np.all(eps[c == i] == e_cond[i]) for i in range(3))
    assert relmax(C, e_cond[:, None] * Cv) < 1e-9
    for net in (0, 1):
        assert caf[net] == pytest.approx(C[net, net] * 1e18, rel=1e-15)


@pytest.mark.parametrize("order,ncols", [(3, 40), (4, 97), (5, 33), (5, 64)])
def test_fused_leaf_kernel_matches_two_launches(ctx, order, ncols):
    # FP32 blocks of RHS: the finest L2L + L2P in one kernel (fuse_leaf = 1)
    # against k_l2l_tiled + k_l2p_wave (fuse_leaf = 0); depth 3 leaves hold
    # tens of panels, column counts are not multiples of 32
    v, c, _ = G.mesh_groups(W.two_wires(), refine=3)
    X = np.random.default_rng(ncols).normal(size=(len(v), ncols))
    o = G.default_opts(method=G.GP_METHOD_FMM, fmm_order=order, max_level=3,
                       m2l_inner=G.GP_M2L_FP32, m2l_rank=0)
    Yf = ctx.fmm_apply(v, X, o, precision=G.GP_PREC_MIXED)
    o.fuse_leaf = 0
    Yu = ctx.fmm_apply(v, X, o, precision=G.GP_PREC_MIXED)
    e = np.linalg.norm(Yf - Yu, axis=0) / np.linalg.norm(Yu, axis=0)
    assert e.max() < 1e-6
