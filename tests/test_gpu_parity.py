"""Device parity: every result of libgpurify's HIP path against the oracle
on the same inputs (called through the C ABI)."""
import math

import numpy as np
import pytest

import gpurify_amd as G
from gpurify_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

EPS0 = 8.8541878128e-12


def relmax(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_device_is_gfx950(ctx):
    assert "gfx950" in ctx.info()


def test_dense_c1_reference_mesh(ctx):
    # bridge path at the reference mesh (173 panels -> dense LU, bridge.rs:399-401)
    groups = W.two_wires()
    caf, C, rep = ctx.solve_capacitance(groups)
    caf_o, Co, n = O.solve_capacitance(groups)
    assert rep["n_panels"] == n == 173 and rep["method_used"] == G.GP_METHOD_DENSE
    assert relmax(C, Co) < 1e-11
    for net in caf_o:
        assert caf[net] == pytest.approx(caf_o[net], rel=1e-11)


def test_dense_sphere_and_plates(ctx):
    v = O.mesh_icosphere((0, 0, 0), 1.0, 2)
    C, rep = ctx.cap_solve(v, np.zeros(len(v)), 1, G.default_opts(method=G.GP_METHOD_DENSE))
    Co = O.cap_dense(v, np.zeros(len(v)), 1)
    assert relmax(C, Co) < 1e-11
    assert abs(C[0, 0] - 4 * math.pi * EPS0) / (4 * math.pi * EPS0) < 0.03
    side, gap = 0.2, 0.02
    v = np.concatenate([O.mesh_plate(0, 0, 0, side, side, 10, 10), O.mesh_plate(0, 0, gap, side, side, 10, 10)])
    cond = np.repeat([0, 1], 100)
    C, _ = ctx.cap_solve(v, cond, 2, G.default_opts(method=G.GP_METHOD_DENSE))
    assert relmax(C, O.cap_dense(v, cond, 2)) < 1e-11


@pytest.mark.parametrize("order,level", [(3, 3), (4, 4), (5, 4), (6, 3)])
def test_fmm_operator_matches_oracle(ctx, order, level):
    v, c, _ = G.mesh_groups(W.two_wires(), refine=3)
    n = len(v)
    rng = np.random.default_rng(order)
    X = rng.normal(size=(n, 3))
    # default: the reference's CPQR-truncated M2L (bbfmm.rs:132-174)
    o = G.default_opts(method=G.GP_METHOD_FMM, fmm_order=order, max_level=level)
    Y = ctx.fmm_apply(v, X, o)
    # opt-in: the full M2L matrices
    Yu = ctx.fmm_apply(v, X, G.default_opts(method=G.GP_METHOD_FMM, fmm_order=order,
                                            max_level=level, m2l_compress=0))
    for j in range(3):
        yr = O.fmm_cap_apply(v, X[:, j], order, level, m2l_mode=0)
        assert np.linalg.norm(Y[:, j] - yr) / np.linalg.norm(yr) < 1e-12
        yo = O.fmm_cap_apply(v, X[:, j], order, level, m2l_mode=1)
        assert np.linalg.norm(Yu[:, j] - yo) / np.linalg.norm(yo) < 1e-12


def test_fmm_operator_with_skew_panels_matches_oracle(ctx):
    """Quadrilaterals whose vertices leave their plane (a panel file's skew
    quads) take the per-entry closed form in the near field; flat ones the
    plane-frame form: both equal the oracle's wilton() to 1e-12."""
    v, c, _ = G.mesh_groups(W.two_wires(), refine=3)
    v = v.copy()
    rng = np.random.default_rng(5)
    skew = rng.random(len(v)) < 0.3
    # lift one vertex of 30 % of the panels off the plane by 1 % of the panel size
    size = np.linalg.norm(v[:, 2] - v[:, 0], axis=1)
    v[skew, 2, :] += 0.01 * size[skew, None] * rng.normal(size=(skew.sum(), 3))
    x = rng.normal(size=(len(v), 2))
    o = G.default_opts(method=G.GP_METHOD_FMM, fmm_order=4, max_level=4)
    Y = ctx.fmm_apply(v, x, o)
    for j in range(2):
        yr = O.fmm_cap_apply(v, x[:, j], 4, 4, m2l_mode=0)
        assert np.linalg.norm(Y[:, j] - yr) / np.linalg.norm(yr) < 1e-12


@pytest.mark.parametrize("order,ncols", [(4, 16), (4, 40), (5, 97), (3, 20), (6, 32), (6, 97)])
def test_fmm_operator_low_precision_close(ctx, order, ncols):
    # ncols > 16 spreads the columns over several 16-wide MFMA row tiles
    v, c, _ = G.mesh_groups(W.two_wires(), refine=3)
    X = np.random.default_rng(order).normal(size=(len(v), ncols))
    o = G.default_opts(method=G.GP_METHOD_FMM, fmm_order=order, m2l_inner=G.GP_M2L_FP32,
                       m2l_rank=0)
    Y64 = ctx.fmm_apply(v, X, o)
    Y32 = ctx.fmm_apply(v, X, o, precision=G.GP_PREC_MIXED)
    e32 = np.linalg.norm(Y32 - Y64, axis=0) / np.linalg.norm(Y64, axis=0)
    assert e32.max() < 1e-5
    o.m2l_inner = G.GP_M2L_BF16
    Yb = ctx.fmm_apply(v, X, o, precision=G.GP_PREC_MIXED)
    nrm = np.linalg.norm(Y64, axis=0)
    e = np.linalg.norm(Yb - Y64, axis=0) / nrm
    # BF16 interaction matrices and multipoles (8-bit mantissa)
    assert 1e-7 < e.min() and e.max() < 5e-3
    # The same BF16-rounded operands through the FP32 VALU kernel: only the
    # FP32 summation order differs, so a tile or lane mapping error in the
    # matrix-core kernel shows at the size of the far field itself.
    o.m2l_inner = G.GP_M2L_BF16_VALU
    Yv = ctx.fmm_apply(v, X, o, precision=G.GP_PREC_MIXED)
    ex = np.linalg.norm(Yb - Yv, axis=0) / nrm
    assert ex.max() < 2e-5 and ex.max() < 0.05 * e.min()


@pytest.mark.parametrize("order", [3, 4, 5, 6])
def test_single_rhs_bf16_m2l_matches_valu(ctx, order):
    # one RHS: the target-tiled matrix-core M2L against the same BF16-rounded
    # operands on the FP32 VALU, on a mesh whose octree has partial
    # interaction lists (tiles with absent sources)
    v, c, _ = G.mesh_groups(W.two_wires(), refine=3)
    x = np.random.default_rng(order).normal(size=(len(v), 1))
    o = G.default_opts(method=G.GP_METHOD_FMM, fmm_order=order)
    Y64 = ctx.fmm_apply(v, x, o)
    o.m2l_inner = G.GP_M2L_BF16
    Yb = ctx.fmm_apply(v, x, o, precision=G.GP_PREC_MIXED)
    o.m2l_inner = G.GP_M2L_BF16_VALU
    Yv = ctx.fmm_apply(v, x, o, precision=G.GP_PREC_MIXED)
    nrm = np.linalg.norm(Y64)
    eb = np.linalg.norm(Yb - Y64) / nrm
    assert 1e-7 < eb < 5e-3
    assert np.linalg.norm(Yb - Yv) / nrm < min(2e-5, 0.05 * eb)


@pytest.mark.parametrize("order", [3, 4, 5])
def test_single_rhs_fp16_near_values(ctx, order):
    # one RHS, mixed: the inner near field as FP16 scaled per row (11
    # significant bits per entry) against FP32 values, with the FP32 M2L so
    # only the near field differs.  Rounding each entry to 2^-12 relative
    # moves the operator by ~1e-4; a wrong scale, a lost row or a misread
    # vector moves it by the near field itself (~1)
    v, c, _ = G.mesh_groups(W.two_wires(), refine=3)
    x = np.random.default_rng(order).normal(size=(len(v), 1))
    o = G.default_opts(method=G.GP_METHOD_FMM, fmm_order=order, m2l_inner=G.GP_M2L_FP32)
    Y64 = ctx.fmm_apply(v, x, o)
    o.near_inner = G.GP_NEAR_FP32
    Y32 = ctx.fmm_apply(v, x, o, precision=G.GP_PREC_MIXED)
    o.near_inner = G.GP_NEAR_FP16
    Y16 = ctx.fmm_apply(v, x, o, precision=G.GP_PREC_MIXED)
    nrm = np.linalg.norm(Y64)
    e32 = np.linalg.norm(Y32 - Y64) / nrm
    e16 = np.linalg.norm(Y16 - Y64) / nrm
    assert e32 < 1e-5 and 1e-7 < e16 < 1e-3
    # entry by entry: each output row within the FP16 rounding of its own row
    assert np.max(np.abs(Y16 - Y32)) < 2e-3 * np.max(np.abs(Y64))


@pytest.mark.parametrize("order,ncols", [(5, 97), (6, 32), (5, 1), (6, 1), (5, 16)])
def test_fmm_low_rank_operator(ctx, order, ncols):
    # rank-64 projection of the M2L matrices (inner operator of a block of
    # RHS): the far-field matrices keep <= 7.7e-5 of their Frobenius norm
    # (orders 5 and 6); one RHS keeps the full rank
    v, c, _ = G.mesh_groups(W.two_wires(), refine=3)
    X = np.random.default_rng(order + ncols).normal(size=(len(v), ncols))
    # (FP32 near values for one RHS as for a block: the bounds are FP32's)
    o = G.default_opts(method=G.GP_METHOD_FMM, fmm_order=order, m2l_inner=G.GP_M2L_FP32,
                       m2l_rank=0, near_inner=G.GP_NEAR_FP32)
    Y64 = ctx.fmm_apply(v, X, o)
    nrm = np.linalg.norm(Y64, axis=0)
    Yf = ctx.fmm_apply(v, X, o, precision=G.GP_PREC_MIXED)
    ef = np.linalg.norm(Yf - Y64, axis=0) / nrm
    o.m2l_rank = 64
    Yl = ctx.fmm_apply(v, X, o, precision=G.GP_PREC_MIXED)
    el = np.linalg.norm(Yl - Y64, axis=0) / nrm
    # the full FP32 operator (near tiles, upward/downward passes) to FP32
    # rounding; the projection shows above it and stays below its bound
    assert ef.max() < 1e-5
    if ncols == 1:
        assert np.array_equal(Yl, Yf)
    else:
        assert el.max() < 1e-4 and el.min() > 3 * ef.max()
    o.m2l_inner = G.GP_M2L_BF16
    Yb = ctx.fmm_apply(v, X, o, precision=G.GP_PREC_MIXED)
    eb = np.linalg.norm(Yb - Y64, axis=0) / nrm
    if ncols > 1:
        # a block of RHS takes split-BF16 products (hi*hi + hi*lo + lo*hi):
        # the FP32 low-rank operator to ~2^-16, the truncation unchanged
        assert eb.max() < 1e-4
        assert (np.linalg.norm(Yb - Yl, axis=0) / nrm).max() < 2e-5
    else:
        # one RHS: plain BF16 target tiles, against the same BF16-rounded
        # low-rank operands on the FP32 VALU
        assert eb.max() < 5e-3
        o.m2l_inner = G.GP_M2L_BF16_VALU
        Yv = ctx.fmm_apply(v, X, o, precision=G.GP_PREC_MIXED)
        ex = np.linalg.norm(Yb - Yv, axis=0) / nrm
        assert ex.max() < 2e-5 and ex.max() < 0.05 * eb.min()
    # rank 32 (the default): the expansion of levels >= 3 runs inside L2L
    # (locals = U Lt + L2L(parent) in one pass).  A lost or misplaced U Lt
    # term moves the far field itself (~10 % of the operator here); the
    # truncation of rank 32 stays far below that
    o.m2l_rank, o.m2l_inner = 32, G.GP_M2L_FP32
    Y32 = ctx.fmm_apply(v, X, o, precision=G.GP_PREC_MIXED)
    o.m2l_inner = G.GP_M2L_BF16
    Y32b = ctx.fmm_apply(v, X, o, precision=G.GP_PREC_MIXED)
    e32 = np.linalg.norm(Y32 - Y64, axis=0) / nrm
    if ncols == 1:
        assert np.array_equal(Y32, Yf)
    else:
        assert ef.max() < e32.min() and e32.max() < 1e-2
        assert (np.linalg.norm(Y32b - Y32, axis=0) / nrm).max() < 2e-5


@pytest.mark.parametrize("order,rank", [(5, 64), (6, 64), (6, 128), (5, 32)])
def test_mixed_low_rank_matches_fp64(ctx, order, rank):
    v, c, _ = G.mesh_groups(W.two_wires(), refine=3)
    common = dict(method=G.GP_METHOD_FMM, tol=1e-11, fmm_order=order)
    C64, _ = ctx.cap_solve(v, c, 3, G.default_opts(**common))
    Cm, rep = ctx.cap_solve(v, c, 3, G.default_opts(precision=G.GP_PREC_MIXED, m2l_rank=rank,
                                                   **common))
    assert rep["converged"] == 1 and np.max(np.abs(Cm - C64) / np.abs(C64)) < 1e-6


def test_m2l_matrix_cores_match_valu(ctx):
    v, c, _ = G.mesh_groups(W.two_wires(), refine=3)
    X = np.random.default_rng(5).normal(size=(len(v), 97))
    for order, level, nc in ((4, 4, 40), (5, 4, 97), (5, 4, 16), (3, 3, 40), (6, 3, 32),
                             (6, 3, 97)):
        o = G.default_opts(method=G.GP_METHOD_FMM, fmm_order=order, max_level=level)
        Ym = ctx.fmm_apply(v, X[:, :nc], o)
        o.m2l_outer = G.GP_M2L_FP64_VALU
        Yv = ctx.fmm_apply(v, X[:, :nc], o)
        err = np.linalg.norm(Ym - Yv, axis=0) / np.linalg.norm(Yv, axis=0)
        assert err.max() < 1e-13, (order, nc)


@pytest.mark.parametrize("order,ncols", [(4, 3), (4, 16), (4, 20), (4, 40), (4, 97), (3, 3),
                                         (5, 3), (6, 3), (2, 3)])
def test_fmm_single_and_block_rhs_agree(ctx, order, ncols):
    # block kernels (near-field tiles of 16 / 32 / 64 RHS columns, far field
    # over RHS blocks) against the single-column kernels (FP64 M2L over
    # 64-target tiles at orders <= 4, grouped by offset above), FP64 and FP32
    v, c, _ = G.mesh_groups(W.two_wires(), refine=3)
    X = np.random.default_rng(ncols).normal(size=(len(v), ncols))
    # (full rank: one RHS never takes the projected operator; FP32 near values
    # for one RHS as for the block)
    o = G.default_opts(method=G.GP_METHOD_FMM, m2l_inner=G.GP_M2L_FP32, fmm_order=order,
                       m2l_rank=0, near_inner=G.GP_NEAR_FP32)
    Yb = ctx.fmm_apply(v, X, o)
    Yb32 = ctx.fmm_apply(v, X, o, precision=G.GP_PREC_MIXED)
    for j in sorted({0, ncols // 2, ncols - 1}):
        Ys = ctx.fmm_apply(v, X[:, j:j + 1], o)
        nrm = np.linalg.norm(Yb[:, j])
        assert np.linalg.norm(Ys[:, 0] - Yb[:, j]) / nrm < 1e-13
        Ys32 = ctx.fmm_apply(v, X[:, j:j + 1], o, precision=G.GP_PREC_MIXED)
        assert np.linalg.norm(Ys32[:, 0] - Yb32[:, j]) / nrm < 2e-6


def test_fmm_solve_fp64_matches_oracle(ctx):
    v, c, _ = G.mesh_groups(W.two_wires(), refine=3)
    o = G.default_opts(method=G.GP_METHOD_FMM, tol=1e-11)
    C, rep = ctx.cap_solve(v, c, 3, o)
    Co, it, rs = O.cap_fmm(v, c, 3, order=4, max_level=4, tol=1e-11, m2l_mode=0)
    assert rep["converged"] == 1 and rep["final_rel_residual"] <= 1e-11
    assert relmax(C, Co) < 1e-9
    # the full M2L matrices (opt-in) against the oracle's uncompressed mode
    Cu, _ = ctx.cap_solve(v, c, 3, G.default_opts(method=G.GP_METHOD_FMM, tol=1e-11,
                                                  m2l_compress=0))
    Cuo, _, _ = O.cap_fmm(v, c, 3, order=4, max_level=4, tol=1e-11, m2l_mode=1)
    assert relmax(Cu, Cuo) < 1e-9


@pytest.mark.parametrize("order", [4, 5])
def test_fmm_single_rhs_blocks_match(ctx, order):
    # one RHS at a time runs the single-column kernels (P2M / L2P per panel,
    # matrix-core M2L grouped by offset, FP64 and BF16)
    v, c, _ = G.mesh_groups(W.two_wires(), refine=3)
    common = dict(method=G.GP_METHOD_FMM, tol=1e-11, fmm_order=order)
    Cb, _ = ctx.cap_solve(v, c, 3, G.default_opts(**common))
    Cs, _ = ctx.cap_solve(v, c, 3, G.default_opts(rhs_block=1, **common))
    assert relmax(Cs, Cb) < 1e-9
    Cm, rep = ctx.cap_solve(v, c, 3, G.default_opts(rhs_block=1, precision=G.GP_PREC_MIXED,
                                                   **common))
    assert rep["converged"] == 1 and np.max(np.abs(Cm - Cb) / np.abs(Cb)) < 1e-6


@pytest.mark.parametrize("basis,passes", [(0, 1), (0, 2), (1, 1), (1, 2)])
def test_mixed_matches_fp64_within_1e6(ctx, basis, passes):
    v, c, _ = G.mesh_groups(W.two_wires(), refine=3)
    C64, _ = ctx.cap_solve(v, c, 3, G.default_opts(method=G.GP_METHOD_FMM, tol=1e-11))
    Cm, rep = ctx.cap_solve(v, c, 3, G.default_opts(method=G.GP_METHOD_FMM, tol=1e-11,
                                                   precision=G.GP_PREC_MIXED, inner_basis=basis,
                                                   inner_passes=passes))
    assert rep["converged"] == 1 and rep["refine_steps"] >= 1
    assert np.max(np.abs(Cm - C64) / np.abs(C64)) < 1e-6


def test_bridge_fmm_rule_matches_oracle(ctx):
    # 1557 panels -> FMM(4,4,1e-8) under the reference rule
    groups = W.two_wires()
    caf, C, rep = ctx.solve_capacitance(groups, refine=3)
    caf_o, Co, n = O.solve_capacitance(groups, refine=3)
    assert rep["method_used"] == G.GP_METHOD_FMM and n == rep["n_panels"]
    # the same CPQR-truncated operator on both sides; GMRES tol 1e-8 (CGS2
    # here, MGS in the oracle) leaves differences of that order
    for net in caf_o:
        assert caf[net] == pytest.approx(caf_o[net], rel=1e-6)


def test_strict_nonconvergence_is_reported(ctx):
    v, c, _ = G.mesh_groups(W.two_wires(), refine=3)
    o = G.default_opts(method=G.GP_METHOD_FMM, tol=1e-14, restart=2, max_restarts=1, strict=1)
    with pytest.raises(G.GpError) as ei:
        ctx.cap_solve(v, c, 3, o)
    assert ei.value.code == G._lib.GP_ERR_NOT_CONVERGED
    o.strict = 0
    C, rep = ctx.cap_solve(v, c, 3, o)
    assert rep["converged"] == 0 and np.all(np.isfinite(C))


def test_invalid_inputs_fail_closed(ctx):
    v, c, _ = G.mesh_groups(W.two_wires())
    with pytest.raises(G.GpError) as ei:
        ctx.cap_solve(v, c, 2)  # conductor id 2 out of range
    assert ei.value.code == G._lib.GP_ERR_INVALID_INPUT
    bad = v.copy()
    bad[5, 0, 0] = np.nan
    with pytest.raises(G.GpError) as ei:
        ctx.cap_solve(bad, c, 3)
    assert ei.value.code == G._lib.GP_ERR_INVALID_INPUT
    with pytest.raises(G.GpError) as ei:
        ctx.cap_solve(v, c, 3, G.default_opts(method=G.GP_METHOD_FMM, fmm_order=9))
    assert ei.value.code == G._lib.GP_ERR_INVALID_INPUT
