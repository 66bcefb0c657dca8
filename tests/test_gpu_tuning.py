"""The context's kernel-variant autotuner: every variant of every tunable
kernel computes the same operator (checked variant by variant against
variant 0), and the tuned choice is recorded per shape."""
import numpy as np
import pytest

import gpurify_amd as G
from gpurify_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("order", [3, 4, 5])
def test_every_variant_applies_the_same_operator(ctx, order):
    v, c, _ = G.mesh_groups(W.two_wires(), refine=3)
    rng = np.random.default_rng(order)
    X = rng.normal(size=(len(v), 97))
    x1 = X[:, :1].copy()
    X40 = X[:, :40].copy()
    o = G.default_opts(method=G.GP_METHOD_FMM, fmm_order=order, m2l_inner=G.GP_M2L_BF16)
    ref = {}
    try:
        for var in range(12):  # every variant (kernels with fewer clamp)
            ctx.tune_force(var)
            out = {
                "fp64_1": ctx.fmm_apply(v, x1, o),
                "fp64_97": ctx.fmm_apply(v, X, o),
                "fp64_40": ctx.fmm_apply(v, X40, o),
                "mixed_1": ctx.fmm_apply(v, x1, o, precision=G.GP_PREC_MIXED),
                "mixed_97": ctx.fmm_apply(v, X, o, precision=G.GP_PREC_MIXED),
            }
            if var == 0:
                ref = out
                continue
            for k, y in out.items():
                # near-field variants differ in summation order only; the M2L
                # variants in which work-group takes which target, and the
                # sibling pairs in the order of the sources
                bound = 1e-13 if k.startswith("fp64") else 1e-6
                assert rel(y, ref[k]) < bound, (var, k)
    finally:
        ctx.tune_force(-1)


def test_autotuner_records_its_choices(ctx):
    v, c, _ = G.mesh_groups(W.two_wires(), refine=3)
    X = np.random.default_rng(3).normal(size=(len(v), 97))
    o = G.default_opts(method=G.GP_METHOD_FMM, fmm_order=5, m2l_inner=G.GP_M2L_BF16, m2l_rank=0)
    ctx.tune_force(0)
    Y0 = ctx.fmm_apply(v, X, o, precision=G.GP_PREC_MIXED)
    ctx.tune_force(-1)  # clears the cache: the next launches tune
    Y1 = ctx.fmm_apply(v, X, o, precision=G.GP_PREC_MIXED)
    Y2 = ctx.fmm_apply(v, X, o, precision=G.GP_PREC_MIXED)
    assert rel(Y1, Y0) < 1e-6 and rel(Y2, Y1) == 0.0
    rep = ctx.tune_report()
    kernels = {e["kernel"] for e in rep}
    assert {"near_spmm_f32", "m2l_bf16_pipe"} <= kernels
    for e in rep:
        assert len(e["ms"]) >= 2 and 0 <= e["variant"] < len(e["ms"])
        assert e["ms"][e["variant"]] == min(e["ms"])
    # one entry per (kernel, shape): a repeated shape reuses the choice
    keys = [(e["kernel"], tuple(e["key"])) for e in rep]
    assert len(keys) == len(set(keys))
