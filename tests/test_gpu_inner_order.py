"""Mixed solves whose inner operator runs its far field at a lower Chebyshev
order than the FP64 operator (gp_cap_opts.inner_order): the FP64 outer
residual keeps fmm_order, so the result must agree with the FP64 solve of the
fmm_order operator to the solve tolerance, single RHS, blocks of RHS (the
low-rank split-BF16 M2L at the inner order) and batches alike."""
import numpy as np
import pytest

import gpurify_amd as G
from gpurify_amd import workloads as W

pytestmark = pytest.mark.gpu


def small_batch(n, pmax, seed=11):
    batch = W.net_batch(n_nets=n, seed=seed, pmin=150, pmax=pmax)
    meshes = W.net_batch_meshes(batch, lambda g, r: G.mesh_groups(g, refine=r, max_panels=10 ** 7))
    return [(v, c, int(c.max()) + 1) for v, c, _ in meshes]


@pytest.mark.parametrize("inner", [2, 3])
def test_single_solve_inner_order_matches_fp64(ctx, inner):
    v, c, _ = G.mesh_groups(W.two_wires(), refine=3)
    common = dict(method=G.GP_METHOD_FMM, fmm_order=4, max_level=4, tol=1e-10)
    C64, r64 = ctx.cap_solve(v, c, 3, G.default_opts(**common))
    o = G.default_opts(precision=G.GP_PREC_MIXED, inner_order=inner, **common)
    Cm, rm = ctx.cap_solve(v, c, 3, o)
    assert rm["converged"] == 1 and r64["converged"] == 1
    assert np.abs(Cm - C64).max() / np.abs(C64).max() < 1e-8
    assert np.max(np.abs(Cm - C64) / np.abs(C64)) < 1e-6  # every entry
    # the same solve with the inner operator at fmm_order
    Cs, _ = ctx.cap_solve(v, c, 3, G.default_opts(precision=G.GP_PREC_MIXED, **common))
    assert np.abs(Cs - Cm).max() / np.abs(C64).max() < 1e-8


def test_block_of_rhs_inner_order_matches_fp64(ctx):
    """The bus at the reference mesh, order 5, 97 RHS in one block: the
    inner operator's low-rank BF16 M2L runs at order 4."""
    v, c, _ = G.mesh_groups(W.bus_crossing(), refine=1)
    nc = int(c.max()) + 1
    common = dict(method=G.GP_METHOD_FMM, fmm_order=5, max_level=0, leaf_target=64, tol=1e-12)
    C64, r64 = ctx.cap_solve(v, c, nc, G.default_opts(**common))
    Cm, rm = ctx.cap_solve(v, c, nc, G.default_opts(precision=G.GP_PREC_MIXED, inner_order=4,
                                                    **common))
    Cs, rs = ctx.cap_solve(v, c, nc, G.default_opts(precision=G.GP_PREC_MIXED, **common))
    assert rm["converged"] == 1 and r64["converged"] == 1 and rs["converged"] == 1
    # entries relative to their rows' and columns' self terms (the weakest
    # couplings are 1e-5 of them and carry the solve tolerance times the
    # conditioning, whatever the inner operator)
    d = np.sqrt(np.outer(np.diag(C64), np.diag(C64)))
    assert np.max(np.abs(Cm - C64) / d) < 1e-9
    assert np.max(np.abs(Cm - Cs) / d) < 1e-9
    strong = np.abs(C64) >= 1e-3 * d
    assert np.max(np.abs(Cm - C64)[strong] / np.abs(C64)[strong]) < 1e-6


def test_batched_inner_order_matches_fp64(ctx):
    probs = small_batch(10, 4000)
    common = dict(method=G.GP_METHOD_FMM, tol=1e-10, rhs_count=1, max_level=0, leaf_target=32)
    C64, _ = ctx.cap_solve_batched(probs, G.default_opts(**common))
    for inner in (3, 4):
        Cm, reps = ctx.cap_solve_batched(probs, G.default_opts(precision=G.GP_PREC_MIXED,
                                                               inner_order=inner, **common))
        for a, b, rep in zip(Cm, C64, reps):
            assert rep["converged"] == 1
            assert np.max(np.abs(a[:, 0] - b[:, 0]) / np.abs(b[:, 0])) < 1e-6


@pytest.mark.parametrize("inner,ncols", [(2, 1), (3, 1), (3, 20), (4, 1)])
def test_inner_operator_is_the_lower_order_operator(ctx, inner, ncols):
    """The FP32 inner operator of an order-5 system with inner_order k is
    the order-k operator on the same tree (its tables, weights and CPQR
    truncation at order k): it equals the FP64 order-k operator to FP32
    rounding, one RHS (target tiles, packed tree passes) and blocks alike."""
    v, c, _ = G.mesh_groups(W.two_wires(), refine=3)
    X = np.random.default_rng(inner).normal(size=(len(v), ncols))
    common = dict(method=G.GP_METHOD_FMM, max_level=4, m2l_inner=G.GP_M2L_FP32, m2l_rank=0)
    Yk = ctx.fmm_apply(v, X, G.default_opts(fmm_order=inner, **common))
    Yi = ctx.fmm_apply(v, X, G.default_opts(fmm_order=5, inner_order=inner, **common),
                       precision=G.GP_PREC_MIXED)
    Y5 = ctx.fmm_apply(v, X, G.default_opts(fmm_order=5, **common))
    nrm = np.linalg.norm(Yk)
    assert np.linalg.norm(Yi - Yk) / nrm < 3e-6
    assert np.linalg.norm(Yi - Y5) / nrm > 1e-6  # (it is not the order-5 operator)


def test_inner_order_is_validated(ctx):
    v, c, _ = G.mesh_groups(W.two_wires(), refine=3)
    for bad in (1, 5):
        o = G.default_opts(method=G.GP_METHOD_FMM, fmm_order=4, precision=G.GP_PREC_MIXED,
                           inner_order=bad)
        with pytest.raises(G.GpError) as e:
            ctx.cap_solve(v, c, 3, o)
        assert e.value.code == G._lib.GP_ERR_INVALID_INPUT
