"""DC resistance half of extract_quasistatic: the product's closed-form chain
(gp_solve_resistance, host C++) against the oracle's full restatement of the
reference path (bridge.rs:180-299 through the henry nodal direct solve at
f = 0, henry/solver.rs:336-584, with num-complex 0.4.6 arithmetic), bit for
bit; the oracle pinned by closed forms and known answers.

The reference's own tests hold no numeric resistance of the quasi-static
path (crates/pex/tests/bridge.rs checks it qualitatively), so the pins are:
the equivalent rectangle of analytical/mod.rs:571-576 (exact for rectangles
and ideal constant-width bends), R = sheet * L / W, the error messages of
rectilinear_metrics, and the map-entry rules of bridge.rs:97-129.
Host-only: nothing here needs a GPU.
"""
import time

import numpy as np
import pytest

import gpurify_amd as G
from oracle import resistance as O

MET1 = dict(sheet_res=0.125, thickness_nm=360.0, height_nm=600.0)


def rect(x0, y0, x1, y1):
    return [(x0, y0), (x1, y0), (x1, y1), (x0, y1)]


def l_shape(x0, y0, a, b, w):
    """Constant-width L: a long along x, b tall along y, width w."""
    return [(x0, y0), (x0 + a, y0), (x0 + a, y0 + b), (x0 + a - w, y0 + b),
            (x0 + a - w, y0 + w), (x0, y0 + w)]


# ---------------------------------------------------------------- known answers
def test_equivalent_rectangle_known_answers():
    assert O.rectilinear_metrics(rect(0, 0, 1000, 140)) == (1000.0, 140.0)
    assert O.rectilinear_metrics(rect(0, 0, 140, 1000)) == (1000.0, 140.0)
    # an ideal bend: area 280000, perimeter 4280 -> the 2000 x 140 rectangle
    assert O.rectilinear_metrics(l_shape(0, 0, 1000, 1140, 140)) == (2000.0, 140.0)


def test_sheet_resistance_closed_form():
    for (x1, y1) in [(1000, 140), (50000, 500), (140, 7000)]:
        r = O.solve_resistance([dict(net=1, xy=rect(0, 0, x1, y1), **MET1)])[1]
        L, W = max(x1, y1), min(x1, y1)
        assert r == pytest.approx(0.125 * L / W, rel=1e-14)
        assert G.solve_resistance([dict(net=1, xy=rect(0, 0, x1, y1), **MET1)])[1] == r
    # dbu scaling leaves the squares count unchanged
    a = G.solve_resistance([dict(net=1, xy=rect(0, 0, 1000, 140), **MET1)], dbu_nm=5.0)[1]
    assert a == pytest.approx(0.125 * 1000 / 140, rel=1e-14)


@pytest.mark.parametrize("ring,msg", [
    ([(0, 0), (10, 0), (10, 10)], "polygon has 3 vertices; at least four are required"),
    ([(0, 0), (20, 0), (20, 10), (30, 10), (30, 20), (10, 20), (10, 10), (0, 10)], None),  # a Z
    ([(0, 0), (20, 0), (20, 10), (10, 10), (10, -10), (0, -10)],
     "self-intersecting polygon is unsupported"),  # a figure 8
    ([(0, 0), (20, 0), (20, 10), (5, 10), (5, -10), (15, -10), (15, 5), (0, 5)],
     "self-intersecting polygon is unsupported"),
    ([(0, 0), (10, 0), (10, 0), (10, 10), (0, 10)], "zero-length edge at vertex 1"),
    ([(0, 0), (10, 0), (12, 10), (0, 10)], "non-Manhattan edge at vertex 1"),
])
def test_metric_errors_match_the_reference_messages(ring, msg):
    q = [dict(net=4, xy=ring, **MET1)]
    if msg is None:
        assert G.solve_resistance(q) == O.solve_resistance(q)
        return
    with pytest.raises(O.ResError, match=msg):
        O.solve_resistance(q)
    with pytest.raises(G.GpError, match=msg):
        G.solve_resistance(q)


def test_map_entries_follow_the_polygon_loop():
    polys = [
        dict(net=1, xy=rect(0, 0, 1000, 140), **MET1),                       # port
        dict(net=1, xy=rect(0, 0, 140, 140), via_res=4.5, sheet_res=0.0,
             thickness_nm=100.0, height_nm=0.0, volume_conductor=0),       # via only
        dict(net=2, xy=rect(0, 0, 500, 500), sheet_res=0.0, thickness_nm=100.0,
             height_nm=0.0),                                               # volume, no R
        dict(net=3, xy=rect(0, 0, 500, 500), sheet_res=0.0, thickness_nm=100.0,
             height_nm=0.0, volume_conductor=0),                           # no entry
        dict(net=5, xy=rect(0, 0, 900, 140), has_pex=0, **MET1),           # no deck params
    ]
    got = G.solve_resistance(polys)
    assert got == O.solve_resistance(polys)
    assert set(got) == {1, 2} and got[2] == 0.0
    assert got[1] == 4.5 + O.solve_resistance(polys[:1])[1]  # vias first, then ports


def test_invalid_process_and_dbu():
    with pytest.raises(G.GpError, match="sheet-resistance volume"):
        G.solve_resistance([dict(net=1, xy=rect(0, 0, 10, 10), sheet_res=-1.0,
                                 thickness_nm=100.0, height_nm=0.0)])
    with pytest.raises(G.GpError, match="database unit"):
        G.solve_resistance([dict(net=1, xy=rect(0, 0, 10, 10), **MET1)], dbu_nm=0.0)


# ---------------------------------------------------------------- product vs oracle
def _random_layout(rng, n):
    polys = []
    for _ in range(n):
        x0, y0 = (int(v) for v in rng.integers(-5000, 5000, 2))
        w = int(rng.choice([140, 170, 230, 500]))
        ln = int(rng.integers(w + 1, 20000))
        kind = rng.integers(4)
        if kind == 0:
            ring = rect(x0, y0, x0 + ln, y0 + w)
        elif kind == 1:
            ring = rect(x0, y0, x0 + w, y0 + ln)
        elif kind == 2:
            ring = l_shape(x0, y0, ln, int(rng.integers(2 * w, 9000)), w)
        else:
            ring = rect(x0, y0, x0 + int(rng.integers(1, 3000)), y0 + int(rng.integers(1, 3000)))
        if rng.random() < 0.5:
            ring = ring[::-1]  # either orientation
        layer = rng.integers(3)
        q = dict(net=int(rng.integers(0, 4)), xy=ring,
                 sheet_res=[0.125, 0.047, 12.2][layer] if rng.random() < 0.85 else 0.0,
                 via_res=4.5 if rng.random() < 0.2 else 0.0,
                 thickness_nm=[360.0, 360.0, 100.0][layer],
                 height_nm=[600.0, 900.0, 0.0][layer],
                 volume_conductor=int(rng.random() < 0.9), has_pex=int(rng.random() < 0.95))
        polys.append(q)
    return polys


def test_closed_form_equals_the_full_nodal_solve_bit_for_bit():
    rng = np.random.default_rng(11)
    for trial in range(60):
        polys = _random_layout(rng, int(rng.integers(1, 10)))
        dbu = float(rng.choice([1.0, 0.5, 5.0]))
        assert G.solve_resistance(polys, dbu) == O.solve_resistance(polys, dbu), trial


def test_routing_block_polygons_closed_form_and_rate():
    """Every rectangle of the seeded 200 um routing block (met1-met3), one
    port each: the squares count sheet * L / W per polygon, and the host rate."""
    from gpurify_amd import workloads as W
    groups, _ = W.routing_block()
    sheet = {600.0: 0.125, 900.0: 0.125, 1200.0: 0.047}
    polys = []
    for g in groups:
        for x0, y0, x1, y1 in np.asarray(g["rects"]).reshape(-1, 4).tolist():
            polys.append(dict(net=g["net"], xy=rect(x0, y0, x1, y1), sheet_res=sheet[g["height_nm"]],
                              thickness_nm=g["thickness_nm"], height_nm=g["height_nm"]))
    t = time.perf_counter()
    got = G.solve_resistance(polys)
    dt = time.perf_counter() - t
    want = {}
    for q in polys:
        (x0, y0), (x1, y1) = q["xy"][0], q["xy"][2]
        L, Wd = max(x1 - x0, y1 - y0), min(x1 - x0, y1 - y0)
        want[q["net"]] = want.get(q["net"], 0.0) + q["sheet_res"] * L / Wd
    assert set(got) == set(want)
    assert max(abs(got[k] - want[k]) / want[k] for k in want) < 1e-13
    print(f"{len(polys)} polygons, {len(got)} nets: {dt * 1e3:.1f} ms "
          f"({len(polys) / dt / 1e6:.2f}M polygons/s with the Python marshalling, two passes)")
