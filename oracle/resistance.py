"""Pure-Python restatement of the DC resistance half of extract_quasistatic.

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker of the
product's ``gp_solve_resistance``; never shipped or measured.  It follows the
reference path without shortcuts, for small layouts:

* ``rectilinear_metrics`` — crates/pex/src/analytical/mod.rs:577-628, with
  ``poly_self_intersects`` of crates/core/src/geometry/ops.rs:216-298 as an
  all-pairs test (the reference's sweep only skips pairs whose spans along the
  sweep axis are disjoint, which cannot cross properly).
* ``solve_resistance`` — crates/pex/src/bridge.rs:180-299: one segment per
  sheet-resistance polygon between two fresh nodes, one port each.
* the henry nodal direct solve at f = 0 — henry/solver.rs:162-198
  (filaments, one per segment: filament_partition(total, 1, r) = (0, total)),
  336-584 (components, grounds at the ports' second nodes, Zb = R + j*w*L with
  w = 0, LU of Zb, Yn = A Zb^-1 A^T, LU of Yn, unit port currents) and
  linalg.rs:172-258 (partial-pivot Doolittle LU and its solve).  The partial
  inductances L only enter as j*0*L: they are taken as +0 here (finite for
  distinct filaments; a zero of either sign changes no real part).
* complex arithmetic — the third-party crate num-complex 0.4.6 (Cargo.lock),
  absent here: Complex64 +, -, * in the textbook forms and division as
  published, ((a c + b d) / (c^2 + d^2), (b c - a d) / (c^2 + d^2)).
* extract_quasistatic's polygon loop — bridge.rs:97-129 (via resistance,
  map entries of volume conductors) before solve_resistance.
"""
from __future__ import annotations

import math

METRES_PER_NM = 1e-9


class ResError(ValueError):
    pass


# ------------------------------------------------------------ num-complex 0.4.6
def c_sub(x, y):
    return (x[0] - y[0], x[1] - y[1])


def c_mul(x, y):
    return (x[0] * y[0] - x[1] * y[1], x[0] * y[1] + x[1] * y[0])


def c_div(x, y):
    (a, b), (c, d) = x, y
    ns = c * c + d * d
    return ((a * c + b * d) / ns, (b * c - a * d) / ns)


def c_mag(x):
    return math.hypot(x[0], x[1])


# ------------------------------------------------------------ linalg.rs:172-258
def lu_factor(a):
    n = len(a)
    lu = [row[:] for row in a]
    piv = list(range(n))
    for k in range(n):
        p, mx = k, c_mag(lu[k][k])
        for i in range(k + 1, n):
            m = c_mag(lu[i][k])
            if m > mx:
                mx, p = m, i
        if mx == 0.0:
            raise ResError(f"singular at {k}")
        if p != k:
            lu[k], lu[p] = lu[p], lu[k]
            piv[k], piv[p] = piv[p], piv[k]
        pivot = lu[k][k]
        for i in range(k + 1, n):
            row = lu[i]
            f = c_div(row[k], pivot)
            row[k] = f
            for j in range(k + 1, n):
                row[j] = c_sub(row[j], c_mul(f, lu[k][j]))
    return lu, piv


def lu_solve(lu, piv, b):
    n = len(lu)
    x = [b[piv[i]] for i in range(n)]
    for i in range(n):
        s = x[i]
        for j in range(i):
            s = c_sub(s, c_mul(lu[i][j], x[j]))
        x[i] = s
    for i in reversed(range(n)):
        s = x[i]
        for j in range(i + 1, n):
            s = c_sub(s, c_mul(lu[i][j], x[j]))
        x[i] = c_div(s, lu[i][i])
    return x


# ------------------------------------------------------------ geometry
def _orient(ax, ay, bx, by, cx, cy):
    return (bx - ax) * (cy - ay) - (by - ay) * (cx - ax)


def self_intersects(ring):
    n = len(ring)
    if n < 4:
        return False
    edges = [(ring[i], ring[(i + 1) % n]) for i in range(n)]
    for i in range(n):
        (ax0, ay0), (ax1, ay1) = edges[i]
        if (ax0, ay0) == (ax1, ay1):
            continue
        for j in range(n):
            if j == i or j == (i + 1) % n or i == (j + 1) % n:
                continue
            (bx0, by0), (bx1, by1) = edges[j]
            if (bx0, by0) == (bx1, by1):
                continue
            d1 = _orient(bx0, by0, bx1, by1, ax0, ay0)
            d2 = _orient(bx0, by0, bx1, by1, ax1, ay1)
            d3 = _orient(ax0, ay0, ax1, ay1, bx0, by0)
            d4 = _orient(ax0, ay0, ax1, ay1, bx1, by1)
            if (d1 > 0) != (d2 > 0) and (d3 > 0) != (d4 > 0) and 0 not in (d1, d2, d3, d4):
                return True
    return False


def rectilinear_metrics(ring):
    """analytical/mod.rs:577-628 -> (length_nm, width_nm)."""
    n = len(ring)
    if n < 4:
        raise ResError(f"polygon has {n} vertices; at least four are required")
    if self_intersects(ring):
        raise ResError("self-intersecting polygon is unsupported")
    area2, perim = 0, 0.0
    for i in range(n):
        (x0, y0), (x1, y1) = ring[i], ring[(i + 1) % n]
        dx, dy = x1 - x0, y1 - y0
        if dx == 0 and dy == 0:
            raise ResError(f"zero-length edge at vertex {i}")
        if dx != 0 and dy != 0:
            raise ResError(f"non-Manhattan edge at vertex {i}")
        perim += float(abs(dx) + abs(dy))
        area2 += x0 * y1 - x1 * y0
    area = float(abs(area2)) / 2.0
    if not math.isfinite(area) or area <= 0.0 or perim <= 0.0:
        raise ResError("polygon has zero or non-finite area/perimeter")
    s = perim / 2.0
    disc = s * s - 4.0 * area
    if disc < -(s * s * 1e-12):
        raise ResError("area/perimeter cannot form a rectilinear equivalent rectangle")
    root = math.sqrt(max(disc, 0.0))
    length, width = (s + root) / 2.0, (s - root) / 2.0
    if width <= 0.0 or not math.isfinite(length):
        raise ResError("polygon has no positive equivalent conductor width")
    return length, width


# ------------------------------------------------------------ bridge.rs + henry
def solve_resistance(polys, dbu_nm=1.0):
    """{net: r_ohm} as extract_quasistatic leaves it (polygon dicts as the
    product binding takes them)."""
    out = {}
    for q in polys:  # bridge.rs:97-129
        if not q.get("has_pex", 1):
            continue
        if q.get("via_res", 0.0) != 0.0:
            out[q["net"]] = out.get(q["net"], 0.0) + q["via_res"]
        if q.get("volume_conductor", 1):
            out.setdefault(q["net"], 0.0)
    xy_scale = dbu_nm * METRES_PER_NM
    nodes, segs, ports, port_nets = [], [], [], []
    for q in polys:  # bridge.rs:187-273
        if not q.get("has_pex", 1) or q.get("sheet_res", 0.0) == 0.0:
            continue
        ring = [tuple(int(v) for v in p) for p in q["xy"]]
        ln, wn = rectilinear_metrics(ring)
        length, width = ln * xy_scale, wn * xy_scale
        thickness = q["thickness_nm"] * METRES_PER_NM
        sigma = 1.0 / (q["sheet_res"] * thickness)
        if not (math.isfinite(length) and math.isfinite(width) and math.isfinite(sigma)
                and length > 0 and width > 0 and sigma > 0):
            raise ResError("invalid process value sheet-resistance volume")
        xs, ys = [p[0] for p in ring], [p[1] for p in ring]
        cx = 0.5 * (float(min(xs)) + float(max(xs))) * xy_scale
        cy = 0.5 * (float(min(ys)) + float(max(ys))) * xy_scale
        cz = (q["height_nm"] + 0.5 * q["thickness_nm"]) * METRES_PER_NM
        if max(xs) - min(xs) >= max(ys) - min(ys):
            p1, p2 = (cx - 0.5 * length, cy, cz), (cx + 0.5 * length, cy, cz)
        else:
            p1, p2 = (cx, cy - 0.5 * length, cz), (cx, cy + 0.5 * length, cz)
        i1, i2 = len(nodes), len(nodes) + 1
        nodes += [p1, p2]
        segs.append((i1, i2, width, thickness, sigma))
        ports.append((i1, i2))
        port_nets.append(q["net"])
    if not segs:
        return out
    # build_filaments (solver.rs:162-198), one filament per segment
    r, n1, n2 = [], [], []
    for i1, i2, w, h, sigma in segs:
        d = [nodes[i2][k] - nodes[i1][k] for k in range(3)]
        length = math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
        r.append(length / (sigma * (w * h)))
        n1.append(i1)
        n2.append(i2)
    nb, nn = len(r), len(nodes)
    # components (union-find) and grounds (solver.rs:355-385)
    parent = list(range(nn))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for a, b in zip(n1, n2):
        ra, rb = find(a), find(b)
        if ra != rb:
            parent[rb] = ra
    comp_id, comp = {}, []
    for n in range(nn):
        comp.append(comp_id.setdefault(find(n), len(comp_id)))
    ground = [None] * len(comp_id)
    for _, c in ports:
        if ground[comp[c]] is None:
            ground[comp[c]] = c
    for n in range(nn):
        if ground[comp[n]] is None:
            ground[comp[n]] = n
    is_ground = [False] * nn
    for g in ground:
        is_ground[g] = True
    red, nred = [None] * nn, 0
    for n in range(nn):
        if not is_ground[n]:
            red[n], nred = nred, nred + 1
    # f = 0: Zb = R (diag) + j * 0 * L
    zero = (0.0, 0.0)
    zb = [[(r[i] if i == j else 0.0, 0.0 * 0.0) for j in range(nb)] for i in range(nb)]
    lu, piv = lu_factor(zb)
    yn = [[zero] * nred for _ in range(nred)]
    for n in range(nn):
        if is_ground[n]:
            continue
        rhs = [zero] * nb
        for b in range(nb):
            if n1[b] == n:
                rhs[b] = (rhs[b][0] + 1.0, rhs[b][1] + 0.0)
            if n2[b] == n:
                rhs[b] = (rhs[b][0] - 1.0, rhs[b][1] - 0.0)
        x = lu_solve(lu, piv, rhs)
        rn = red[n]
        for b in range(nb):
            b1, b2 = n1[b], n2[b]
            if not is_ground[b1]:
                yn[red[b1]][rn] = (yn[red[b1]][rn][0] + x[b][0], yn[red[b1]][rn][1] + x[b][1])
            if not is_ground[b2]:
                yn[red[b2]][rn] = c_sub(yn[red[b2]][rn], x[b])
    ylu, ypiv = lu_factor(yn)
    for p, (a, c) in enumerate(ports):
        iinj = [zero] * nred
        if not is_ground[a]:
            iinj[red[a]] = (iinj[red[a]][0] + 1.0, iinj[red[a]][1] + 0.0)
        if not is_ground[c]:
            iinj[red[c]] = (iinj[red[c]][0] - 1.0, iinj[red[c]][1] - 0.0)
        v = lu_solve(ylu, ypiv, iinj)
        va = zero if is_ground[a] else v[red[a]]
        vc = zero if is_ground[c] else v[red[c]]
        ohm = c_sub(va, vc)[0]  # resistance(0) = Re Z (solver.rs:95-105)
        if not math.isfinite(ohm):
            raise ResError("non-finite resistance")
        out[port_nets[p]] = out.get(port_nets[p], 0.0) + ohm
    return out
