"""Batched per-net solves (gp_cap_solve_batched) on the device: every
problem of a batch must come out as its own single solve does, in FP64 and
mixed precision, with the reference's dense/FMM rule per problem."""
import numpy as np
import pytest

import gpurify_amd as G
from gpurify_amd import workloads as W

pytestmark = pytest.mark.gpu


def small_batch(n, pmax, seed=11):
    batch = W.net_batch(n_nets=n, seed=seed, pmin=150, pmax=pmax)
    meshes = W.net_batch_meshes(batch, lambda g, r: G.mesh_groups(g, refine=r, max_panels=10 ** 7))
    return [(v, c, int(c.max()) + 1) for v, c, _ in meshes]


@pytest.fixture(scope="module")
def probs():
    ps = small_batch(10, 4000)
    assert sum(len(v) >= 500 for v, _, _ in ps) >= 5 and sum(len(v) < 500 for v, _, _ in ps) >= 1
    return ps


def test_batched_fmm_equals_single_solves(ctx, probs):
    o = G.default_opts(method=G.GP_METHOD_FMM, tol=1e-11, rhs_count=1, max_level=0,
                       leaf_target=32)
    Cs, reps = ctx.cap_solve_batched(probs, o)
    for (v, c, nc), Cb, rep in zip(probs, Cs, reps):
        C1, r1 = ctx.cap_solve(v, c, nc, o)
        assert rep["converged"] == 1 and rep["method_used"] == G.GP_METHOD_FMM
        assert rep["max_level"] == r1["max_level"] and rep["n_boxes"] == r1["n_boxes"]
        # same tree, same operator: the solutions agree to the solve tolerance
        np.testing.assert_allclose(Cb[:, 0], C1[:, 0], rtol=1e-9)
        assert np.all(Cb[:, 1:] == 0.0)  # columns outside rhs_count stay zero


def test_batched_mixed_matches_fp64_within_1e6(ctx, probs):
    common = dict(method=G.GP_METHOD_FMM, tol=1e-10, rhs_count=1, max_level=0, leaf_target=32)
    C64, _ = ctx.cap_solve_batched(probs, G.default_opts(**common))
    Cm, reps = ctx.cap_solve_batched(probs, G.default_opts(precision=G.GP_PREC_MIXED, **common))
    for a, b, rep in zip(Cm, C64, reps):
        assert rep["converged"] == 1 and rep["refine_steps"] >= 1
        assert np.max(np.abs(a[:, 0] - b[:, 0]) / np.abs(b[:, 0])) < 1e-6


def test_batched_auto_rule_and_all_columns(ctx, probs):
    # AUTO: problems below 500 panels go dense (bridge.rs:397-404), as in gp_cap_solve
    o = G.default_opts(tol=1e-11)
    Cs, reps = ctx.cap_solve_batched(probs, o)
    for (v, c, nc), Cb, rep in zip(probs, Cs, reps):
        want = G.GP_METHOD_DENSE if len(v) < 500 else G.GP_METHOD_FMM
        assert rep["method_used"] == want
        C1, _ = ctx.cap_solve(v, c, nc, o)
        np.testing.assert_allclose(Cb, C1, rtol=1e-8, atol=1e-9 * np.abs(C1).max())


def test_batched_rejects_bad_problem(ctx, probs):
    bad = list(probs[:3])
    v, c, nc = bad[1]
    bad[1] = (v, c, nc - 1)  # conductor id out of range
    with pytest.raises(G.GpError) as ei:
        ctx.cap_solve_batched(bad)
    assert ei.value.code == G._lib.GP_ERR_INVALID_INPUT


@pytest.mark.parametrize("prec", [G.GP_PREC_FP64, G.GP_PREC_MIXED])
def test_skipping_converged_problems_changes_nothing(ctx, probs, prec):
    # the operator applies of a batched GMRES step skip the problems whose
    # system has converged; problems share no box or tile, so every result
    # must come out bit for bit as with every problem applied in every step
    common = dict(method=G.GP_METHOD_FMM, tol=1e-10, rhs_count=1, max_level=0, leaf_target=32,
                  precision=prec)
    Cs, rs = ctx.cap_solve_batched(probs, G.default_opts(skip_converged=1, **common))
    Ca, ra = ctx.cap_solve_batched(probs, G.default_opts(skip_converged=0, **common))
    assert len({r["iterations"] for r in rs}) > 1  # the systems converge at different steps
    for a, b, x, y in zip(Cs, Ca, rs, ra):
        assert x["converged"] == 1 and x["iterations"] == y["iterations"]
        assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", [G.GP_PREC_FP64, G.GP_PREC_MIXED])
def test_groups_set_up_beside_the_solve_agree_with_one_group(ctx, probs, prec):
    # batch_groups = 3: three stacked operators, each set up on a second host
    # thread while the one before solves, the dense problems solving during
    # the first set-up; every problem keeps its own tree and operator, so
    # the solutions agree to the solve tolerance
    common = dict(tol=1e-11, rhs_count=1, max_level=0, leaf_target=32, precision=prec)
    C1, r1 = ctx.cap_solve_batched(probs, G.default_opts(batch_groups=1, **common))
    C3, r3 = ctx.cap_solve_batched(probs, G.default_opts(batch_groups=3, **common))
    for a, b, x, y, (v, _, _) in zip(C3, C1, r3, r1, probs):
        assert x["converged"] == 1 and x["method_used"] == y["method_used"]
        assert x["n_boxes"] == y["n_boxes"] and x["near_nnz"] == y["near_nnz"]
        assert x["m2l_pairs"] == y["m2l_pairs"]
        np.testing.assert_allclose(a[:, 0], b[:, 0], rtol=1e-8 if prec == G.GP_PREC_FP64 else 1e-6)
    # each problem reports its own share of the operator
    fm = [r for r in r1 if r["method_used"] == G.GP_METHOD_FMM]
    assert len({r["near_nnz"] for r in fm}) == len(fm)


def test_groups_under_the_profiler(ctx, probs):
    # the set-up thread's kernel records join the calling thread's profile
    ctx.profile(True)
    try:
        ctx.cap_solve_batched(probs, G.default_opts(batch_groups=3, tol=1e-10, rhs_count=1,
                                                    max_level=0, leaf_target=32))
        prof = ctx.profile_report()
    finally:
        ctx.profile(False)
    assert prof["near_assemble"]["launches"] == 3


def test_groups_report_the_failing_problem_by_its_batch_position(ctx, probs):
    o = G.default_opts(method=G.GP_METHOD_FMM, batch_groups=3, tol=1e-14, restart=5,
                       max_restarts=1, strict=1, rhs_count=1, max_level=0, leaf_target=32)
    with pytest.raises(G.GpError) as ei:
        ctx.cap_solve_batched(probs, o)
    assert ei.value.code == G._lib.GP_ERR_NOT_CONVERGED
    assert "for problem" in str(ei.value)


def test_fp16_near_values_reach_the_fp64_answer(ctx, probs):
    # the inner operator of one-RHS batches on FP16 near tiles or FP32 (the
    # default): both refine to the same FP64 operator's solution
    common = dict(method=G.GP_METHOD_FMM, tol=1e-10, rhs_count=1, max_level=0, leaf_target=32)
    C64, _ = ctx.cap_solve_batched(probs, G.default_opts(**common))
    mixed = dict(common, precision=G.GP_PREC_MIXED)
    C16, r16 = ctx.cap_solve_batched(probs, G.default_opts(near_inner=G.GP_NEAR_FP16, **mixed))
    C32, r32 = ctx.cap_solve_batched(probs, G.default_opts(near_inner=G.GP_NEAR_FP32, **mixed))
    for a, b, c, x, y in zip(C16, C32, C64, r16, r32):
        assert x["converged"] == 1 and y["converged"] == 1
        assert x["refine_steps"] <= y["refine_steps"] + 1
        assert np.max(np.abs(a[:, 0] - c[:, 0]) / np.abs(c[:, 0])) < 1e-6
        assert np.max(np.abs(b[:, 0] - c[:, 0]) / np.abs(c[:, 0])) < 1e-6


def test_near_inner_is_validated(ctx, probs):
    with pytest.raises(G.GpError, match="INVALID_INPUT"):
        ctx.cap_solve_batched(probs[:2], G.default_opts(precision=G.GP_PREC_MIXED, near_inner=7,
                                                        method=G.GP_METHOD_FMM))
