"""Seeded synthetic layouts for the capacitance workloads (DBU = 1 nm).

Each workload is a list of volume groups ``{net, height_nm, thickness_nm, k,
rects}`` in the reference's (net, layer) order (bridge.rs:91, 131-141), on
the sky130 deck's PEX layers with its heuristic process stack.
"""
from __future__ import annotations

import numpy as np

# sky130 PEX layers in LayerId order, i.e. sorted by (GDS layer, datatype)
# (pdks/sky130.json "layers"/"pex"; core/src/io/pdk.rs:489-500).  None of
# them sets thickness_nm / height_nm / dielectric_k.
SKY130_PEX_LAYERS = [("poly", 66, 20), ("li", 67, 20), ("met1", 68, 20), ("met2", 69, 20),
                     ("met3", 70, 20), ("met4", 71, 20), ("met5", 72, 20)]


def sky130_stack():
    """{layer: (height_nm, thickness_nm, k)} by the heuristic of
    analytical/process_stack.rs:31-76: 200 nm thick, 100 nm apart, k = 3.9,
    stacked bottom-up from z = 0 in LayerId order."""
    out, z = {}, 0.0
    for name, _, _ in SKY130_PEX_LAYERS:
        t, h = 200.0, z
        z = h + t + 100.0
        out[name] = (h, t, 3.9)
    return out


def _group(net, layer, rects, stack):
    h, t, k = stack[layer]
    return {"net": int(net), "height_nm": h, "thickness_nm": t, "k": k, "layer": layer,
            "rects": np.asarray(rects, dtype=np.int32).reshape(-1, 4)}


def sort_groups(groups):
    order = {name: i for i, (name, _, _) in enumerate(SKY130_PEX_LAYERS)}
    return sorted(groups, key=lambda g: (g["net"], order[g["layer"]]))


def two_wires():
    """2 parallel met1 wires, 10 um long, 0.14 um wide, 0.14 um apart."""
    st = sky130_stack()
    return sort_groups([_group(0, "met1", [0, 0, 10000, 140], st),
                        _group(1, "met1", [0, 280, 10000, 420], st)])


def bus_crossing(n_m1=64, n_m2=32, length=50000, wmin=140, wmax=500, seed_m1=2, seed_m2=3):
    """A met1 bus of n_m1 wires along x (length nm long, widths and spacings
    uniform in [wmin, wmax] nm, seed_m1) crossed by n_m2 met2 wires along y
    spanning the bus (same distribution, seed_m2), centred in x.  One net per
    wire: met1 wires are nets 0..n_m1-1, met2 wires follow."""
    st = sky130_stack()
    groups = []
    r1 = np.random.default_rng(seed_m1)
    y = 0
    for i in range(n_m1):
        w = int(r1.integers(wmin, wmax + 1))
        groups.append(_group(i, "met1", [0, y, length, y + w], st))
        y += w + int(r1.integers(wmin, wmax + 1))
    bus_h = y
    r2 = np.random.default_rng(seed_m2)
    ws = [(int(r2.integers(wmin, wmax + 1)), int(r2.integers(wmin, wmax + 1))) for _ in range(n_m2)]
    total = sum(w + s for w, s in ws)
    x = max(0, (length - total) // 2)
    for j, (w, s) in enumerate(ws):
        groups.append(_group(n_m1 + j, "met2", [x, 0, x + w, bus_h], st))
        x += w + s
    return sort_groups(groups)


def crossing_bundle(n=16, width=140, space=140, length=20000, seed=5):
    """n met1 wires along x crossed by n met2 wires along y (W = S = 140 nm)."""
    st = sky130_stack()
    groups = []
    pitch = width + space
    span = n * pitch - space
    off = (length - span) // 2
    for i in range(n):
        y0 = off + i * pitch
        groups.append(_group(i, "met1", [0, y0, length, y0 + width], st))
    for j in range(n):
        x0 = off + j * pitch
        groups.append(_group(n + j, "met2", [x0, 0, x0 + width, length], st))
    return sort_groups(groups)


def routing_block(size=200000, layers=("met1", "met2", "met3"), wmin=140, wmax=1000,
                  lmin=1000, lmax=20000, n_rhs=32, seed=6):
    """Random Manhattan routing over a size x size nm block: on each layer
    (met1 and met3 along x, met2 along y) tracks of width U[wmin, wmax]
    separated by U[wmin, wmax] carry segments of length U[lmin, lmax]
    separated by U[wmin, 4/3 length] gaps, about 30% area density per layer.
    Every segment is a net; n_rhs segments drawn at random are nets
    0..n_rhs-1 (the right-hand sides solved), the others follow in
    generation order.  Returns (groups, density per layer)."""
    st = sky130_stack()
    rng = np.random.default_rng(seed)
    segs, dens = [], {}
    for li, layer in enumerate(layers):
        horiz = li % 2 == 0
        area, pos = 0, 0
        while True:
            w = int(rng.integers(wmin, wmax + 1))
            if pos + w > size:
                break
            t = int(rng.integers(0, lmax + 1)) - lmax // 2
            while True:
                ln = int(rng.integers(lmin, lmax + 1))
                gap = int(rng.integers(wmin, (4 * ln) // 3 + 1))
                a, b = max(t, 0), min(t + ln, size)
                if a >= size:
                    break
                if b - a >= lmin // 2:
                    r = [a, pos, b, pos + w] if horiz else [pos, a, pos + w, b]
                    segs.append((layer, r))
                    area += (b - a) * w
                t += ln + gap
            pos += w + int(rng.integers(wmin, wmax + 1))
        dens[layer] = area / float(size) ** 2
    pick = rng.choice(len(segs), size=min(n_rhs, len(segs)), replace=False)
    order = list(pick) + [i for i in range(len(segs)) if i not in set(pick.tolist())]
    groups = [_group(net, segs[i][0], segs[i][1], st) for net, i in enumerate(order)]
    return sort_groups(groups), dens


def net_batch(n_nets=4096, seed=4, kmax=12, widths=(140, 200, 300, 500), lmin=1000,
              lmax=10000, pmin=200, pmax=20000):
    """Independent nets for the batched per-net solve.  Net i: K ~ U{1..kmax}
    Manhattan segments as a random walk alternating x (met1) and y (met2),
    widths drawn from `widths` nm, lengths U[lmin, lmax] nm, and a target
    panel count log-uniform in [pmin, pmax].  Returns [(groups, target)],
    each net alone as net 0 (its substrate plate is conductor 1)."""
    st = sky130_stack()
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_nets):
        k = int(rng.integers(1, kmax + 1))
        x, y = 0, 0
        rects = {"met1": [], "met2": []}
        horiz = bool(rng.integers(0, 2))
        for _s in range(k):
            w = int(rng.choice(widths))
            ln = int(rng.integers(lmin, lmax + 1)) * (1 if rng.integers(0, 2) else -1)
            if horiz:
                x1 = x + ln
                rects["met1"].append([min(x, x1), y - w // 2, max(x, x1), y - w // 2 + w])
                x = x1
            else:
                y1 = y + ln
                rects["met2"].append([x - w // 2, min(y, y1), x - w // 2 + w, max(y, y1)])
                y = y1
            horiz = not horiz
        target = float(np.exp(rng.uniform(np.log(pmin), np.log(pmax))))
        groups = [_group(0, layer, rs, st) for layer, rs in rects.items() if rs]
        out.append((sort_groups(groups), target))
    return out


def net_batch_meshes(batch, mesh):
    """Mesh every net of net_batch at the refinement r >= 1 whose panel count
    is closest to its target (mesh = the bridge mesher, mesh(groups, refine)
    -> (verts, cond, ...)).  Returns [(verts, cond, r)]."""
    out = []
    for groups, target in batch:
        v1, c1 = mesh(groups, 1)[:2]
        r = max(1, int(round(np.sqrt(target / len(v1)))))
        if r > 1:
            v1, c1 = mesh(groups, r)[:2]
        out.append((v1, c1, r))
    return out


WORKLOADS = {
    "two_wires_met1": two_wires,
    "bus64x32": bus_crossing,
    "bundle16x16": crossing_bundle,
    "routing200": lambda: routing_block()[0],
}
