"""gp_cap_solve_dense_noalloc (the dense solve in the library's fixed device
workspace) against gp_cap_solve's dense path, through the C ABI: the same C
on two calls in a row, the second with a smaller input, and the size limits."""
import numpy as np
import pytest

import gpurify_amd as G
from gpurify_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_noalloc_matches_dense_entry_on_consecutive_calls(ctx):
    # 1st: the C1 reference mesh, 173 panels, 3 conductors
    v1, c1, _ = G.mesh_groups(W.two_wires())
    # 2nd, smaller: a 60-panel icosphere-free plate pair, 2 conductors
    v2 = np.concatenate([O.mesh_plate(0, 0, 0, 0.2, 0.2, 5, 6), O.mesh_plate(0, 0, 0.02, 0.2, 0.2, 5, 6)])
    c2 = np.repeat([0, 1], 30)
    od = G.default_opts(method=G.GP_METHOD_DENSE)
    for v, c, nc in ((v1, c1, 3), (v2, c2, 2), (v1, c1, 3)):
        print(f"noalloc: {len(v)} panels, {nc} conductors", flush=True)
        Cn, rn = ctx.cap_solve_dense_noalloc(v, c, nc)
        Cd, rd = ctx.cap_solve(v, c, nc, od)
        assert rn["method_used"] == G.GP_METHOD_DENSE and rn["n_panels"] == len(v)
        assert np.array_equal(Cn, Cd)  # the same arithmetic in the same order
        assert np.abs(Cn - O.cap_dense(v, c, nc)).max() / np.abs(Cd).max() < 1e-11


def test_noalloc_limits_and_errors(ctx):
    v = O.mesh_icosphere((0, 0, 0), 1.0, 3)  # 1280 panels > 512
    with pytest.raises(G.GpError) as ei:
        ctx.cap_solve_dense_noalloc(v, np.zeros(len(v)), 1)
    assert ei.value.code == G._lib.GP_ERR_INVALID_INPUT and "at most 512" in ei.value.message
    v, c, _ = G.mesh_groups(W.two_wires())
    with pytest.raises(G.GpError) as ei:
        ctx.cap_solve_dense_noalloc(v, c, 65)
    assert ei.value.code == G._lib.GP_ERR_INVALID_INPUT
    with pytest.raises(G.GpError) as ei:
        ctx.cap_solve_dense_noalloc(v, c, 2)  # conductor id 2 out of range
    assert ei.value.code == G._lib.GP_ERR_INVALID_INPUT
    # a valid call after the failures still matches
    Cn, _ = ctx.cap_solve_dense_noalloc(v, c, 3)
    Cd, _ = ctx.cap_solve(v, c, 3, G.default_opts(method=G.GP_METHOD_DENSE))
    assert np.array_equal(Cn, Cd)
