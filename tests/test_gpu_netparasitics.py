"""The whole NetParasitics of extract_quasistatic (bridge.rs:73-147):
r_ohm from gp_solve_resistance and cap_af from gp_solve_capacitance on a
window of the routing block, against the oracle's restatement of both halves
(bridge.rs:180-299 through the henry nodal solve; bridge.rs:307-443 with the
reference's FMM(4, 4, 1e-8) rule); and the drop-in capacitance entry at the
bus's full size against the panel-level solve of the same mesh."""
import numpy as np
import pytest

import gpurify_amd as G
from gpurify_amd import workloads as W
from oracle import oracle as O
from oracle import resistance as OR

pytestmark = pytest.mark.gpu

SHEET = {"met1": 0.125, "met2": 0.125, "met3": 0.047}  # ohm/sq


def polygons(groups):
    out = []
    for g in groups:
        for x0, y0, x1, y1 in np.asarray(g["rects"]).reshape(-1, 4):
            out.append(dict(net=g["net"], xy=[(x0, y0), (x1, y0), (x1, y1), (x0, y1)],
                            sheet_res=SHEET[g["layer"]], thickness_nm=g["thickness_nm"],
                            height_nm=g["height_nm"]))
    return out


def test_net_parasitics_routing_window(ctx):
    groups = W.routing_block(size=12000)[0]
    r_dev = G.solve_resistance(polygons(groups))
    r_or = OR.solve_resistance(polygons(groups))
    assert r_dev == r_or  # bit for bit
    caf, C, rep = ctx.solve_capacitance(groups)  # the reference rule
    caf_o, Co, n = O.solve_capacitance(groups)
    print(f"window 12 um: {n} panels, {len(caf)} nets, method {rep['method_used']}", flush=True)
    assert rep["n_panels"] == n and rep["method_used"] == G.GP_METHOD_FMM
    assert set(caf) == set(caf_o) == set(r_dev)
    worst = max(abs(caf[k] - caf_o[k]) / caf_o[k] for k in caf)
    print(f"  cap_af vs oracle: worst {worst:.2e}", flush=True)
    assert worst < 1e-6
    parasitics = {k: (r_dev[k], caf[k]) for k in sorted(caf)}
    assert all(r > 0 and c > 0 for r, c in parasitics.values())


def test_drop_in_entry_at_bus_scale(ctx):
    # gp_solve_capacitance itself (mesher, plate, solve, eps, cap_af) on the
    # 211k-panel bus with the benchmark's solver options
    groups = W.bus_crossing()
    o = G.default_opts(method=G.GP_METHOD_FMM, precision=G.GP_PREC_MIXED, fmm_order=5,
                       max_level=0, tol=1e-12)
    caf, C, rep = ctx.solve_capacitance(groups, refine=4, max_panels=10 ** 9, opts=o)
    v, c, eps = G.mesh_groups(groups, refine=4, max_panels=10 ** 9)
    assert rep["n_panels"] == len(v) == 211136 and np.all(eps == 3.9)
    Cp, rp = ctx.cap_solve(v, c, int(c.max()) + 1, o)
    assert rep["converged"] == 1 and rp["converged"] == 1
    for net in range(96):
        assert caf[net] == pytest.approx(3.9 * Cp[net, net] * 1e18, rel=1e-12)
    np.testing.assert_allclose(C, 3.9 * Cp, rtol=1e-12, atol=1e-12 * np.abs(Cp).max())
