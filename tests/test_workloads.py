"""The seeded synthetic layouts: deterministic, well-formed (no two nets
overlap on one layer, so no coincident panels), and of the advertised size."""
import numpy as np
import pytest

import gpurify_amd as G
from gpurify_amd import workloads as W


def _overlap(a, b):
    return a[0] < b[2] and b[0] < a[2] and a[1] < b[3] and b[1] < a[3]


def test_routing_block_is_deterministic_and_non_overlapping():
    g1, d1 = W.routing_block(size=30000, seed=6)
    g2, d2 = W.routing_block(size=30000, seed=6)
    assert d1 == d2 and len(g1) == len(g2)
    assert all((a["rects"] == b["rects"]).all() and a["net"] == b["net"] for a, b in zip(g1, g2))
    for layer in ("met1", "met2", "met3"):
        rs = [g["rects"][0] for g in g1 if g["layer"] == layer]
        for i in range(len(rs)):
            for j in range(i + 1, len(rs)):
                assert not _overlap(rs[i], rs[j]), (layer, rs[i], rs[j])
    nets = sorted(g["net"] for g in g1)
    assert nets == list(range(len(g1)))  # one segment per net, nets 0..K-1
    for d in d1.values():
        assert 0.15 < d < 0.45


def test_routing_block_full_size_panel_count():
    g, dens = W.routing_block()
    assert 5000 < len(g) < 8000
    assert all(0.25 < d < 0.35 for d in dens.values())
    v, c, _ = G.mesh_groups(g, refine=3, max_panels=10 ** 9)
    assert 1.8e6 < len(v) < 2.8e6  # ~2M panels
    assert int(c.max()) == len(g)  # the substrate plate is the last conductor


def test_wire_params_velocities():
    from gpurify_amd import sensitivity as S
    v, c, _ = G.mesh_groups(W.two_wires(), refine=2)
    params, names = S.wire_params(v, c, [0, 1])
    assert names[:3] == [(0, "width"), (0, "spacing"), (0, "height")]
    idx, dv = params[0]
    assert set(np.unique(c[idx])) == {0}
    pts = v[idx]
    y = pts[:, :, 1]
    # the two side faces move apart by 1/2 each per unit width
    assert np.allclose(dv[:, :, 1][np.isclose(y, y.max())], 0.5)
    assert np.allclose(dv[:, :, 1][np.isclose(y, y.min())], -0.5)
    assert np.all(dv[:, :, [0, 2]] == 0.0)
    moved = S.move(v, params, np.eye(len(params))[0] * 2e-9)
    w0 = np.ptp(v[idx][:, :, 1])
    assert np.ptp(moved[idx][:, :, 1]) == pytest.approx(w0 + 2e-9, rel=1e-9)
