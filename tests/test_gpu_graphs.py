"""Captured inner GMRES steps (HIP graphs, replayed from the second
refinement step on): the same results as the kernels launched one by one.
Each graph replays several times, with the solver's own eager work (the
FP64 residual, the restart bookkeeping) and synchronisations between the
replays, and every solve is checked against its eager twin."""
import numpy as np
import pytest

import gpurify_amd as G
from gpurify_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_block_solve_graphs_match_eager(ctx):
    v, c, _ = G.mesh_groups(W.two_wires(), refine=3)
    common = dict(method=G.GP_METHOD_FMM, tol=1e-11, precision=G.GP_PREC_MIXED)
    Ce, re = ctx.cap_solve(v, c, 3, G.default_opts(step_graphs=0, **common))
    for _ in range(2):  # the second solve re-captures on a fresh workspace
        Cg, rg = ctx.cap_solve(v, c, 3, G.default_opts(step_graphs=1, **common))
        assert rg["refine_steps"] >= 3 and rg["converged"] == 1
        assert (rg["iterations"], rg["refine_steps"]) == (re["iterations"], re["refine_steps"])
        assert np.array_equal(Cg, Ce)


@pytest.mark.parametrize("rhs_block", [1, 0])
def test_single_rhs_and_batch_graphs_match_eager(ctx, rhs_block):
    v, c, _ = G.mesh_groups(W.two_wires(), refine=3)
    common = dict(method=G.GP_METHOD_FMM, tol=1e-11, precision=G.GP_PREC_MIXED,
                  rhs_block=rhs_block)
    Ce, _ = ctx.cap_solve(v, c, 3, G.default_opts(step_graphs=0, **common))
    Cg, _ = ctx.cap_solve(v, c, 3, G.default_opts(step_graphs=1, **common))
    assert np.array_equal(Cg, Ce)
    # a batch of independent problems (segmented GMRES in lockstep)
    probs = [G.mesh_groups(g, refine=r)[:2] for g, r in ((W.two_wires(), 3), (W.two_wires(), 2))]
    probs = [(pv, pc, int(pc.max()) + 1) for pv, pc in probs]
    bo = dict(method=G.GP_METHOD_FMM, tol=1e-11, precision=G.GP_PREC_MIXED)
    Be, _ = ctx.cap_solve_batched(probs, G.default_opts(step_graphs=0, **bo))
    Bg, reps = ctx.cap_solve_batched(probs, G.default_opts(step_graphs=1, **bo))
    assert all(r["converged"] for r in reps)
    for a, b in zip(Bg, Be):
        assert np.array_equal(a, b)
