#!/usr/bin/env python3
"""Writes the golden fixtures of the capacitance path into tests/golden/.

Inputs are the reference's own test inputs where it has them and seeded
synthetic layouts otherwise; expected outputs come from the oracle (the CPU
FP64 restatement of the reference algorithm under oracle/), except where a
closed form fixes the value, which is then written from the closed form.

  wilton_kat.json   potential_integral known answers (integrals/panel.rs:147-208):
                    the unit square at its centre (4*asinh 1, tests/analytic.rs:14-27),
                    off-centre and off-plane points of rectangles (the textbook corner
                    formula), a triangle, and the far monopole limit
  c1_r1.npz         the two-wire met1 layout at the reference mesh: the 173 panels of
                    bridge.rs:445-652 + cap/mesh.rs:12-43 (vertices, conductor, eps),
                    the 3 x 3 Maxwell C (F, eps applied) of the dense LU branch and
                    cap_af of both nets (bridge.rs:397-443)
  c1_r3.npz         the same layout at refine 3 (1,557 panels: the FMM branch of the
                    bridge rule, FMM(4, 4, 1e-8) with the CPQR-truncated M2L), C of
                    that branch and of the dense LU on the same panels; a SHA-256 of
                    the panel bytes instead of the panels
  lcg_fmm.npz       the point sets of crates/pex/tests/fmm.rs:21-60 (900 points seed
                    12345, 800 points seed 999, their charges), the dense O(N^2)
                    apply and the oracle's FMM matvec at orders 3, 4, 5 and depths 3, 4
  sphere_plate.json the 320-panel icosphere (unit radius, cap_validation.rs:9-28):
                    dense and FMM(4, 3) C; the 10 x 10 parallel-plate pair
                    (cap_validation.rs:47-77): 2 x 2 dense C

Run from the repository root after building the oracle (make -C oracle):
    python tests/golden/make_golden.py
"""
import hashlib
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from gpurify_amd import workloads as W  # noqa: E402  (layout generators, pure Python)
from oracle import oracle as O  # noqa: E402


def rect(x0, y0, x1, y1, z=0.0):
    return [[x0, y0, z], [x1, y0, z], [x1, y1, z], [x0, y1, z]]


def corner_formula(a, b, x, y, z):
    """Integral of 1/|p - r'| over [0, a] x [0, b] in z = 0 at p = (x, y, z):
    the sum over the corners (u, v) of F(u, v) with the signs of a mixed
    difference, F = u asinh(v / sqrt(u^2 + z^2)) + v asinh(u / sqrt(v^2 + z^2))
    - z atan(u v / (z R))."""
    def F(u, v):
        R = math.sqrt(u * u + v * v + z * z)
        t = 0.0
        if u != 0.0:
            t += u * math.asinh(v / math.hypot(u, z))
        if v != 0.0:
            t += v * math.asinh(u / math.hypot(v, z))
        if z != 0.0:
            t -= z * math.atan(u * v / (z * R))
        return t
    u0, u1, v0, v1 = -x, a - x, -y, b - y
    return F(u1, v1) - F(u0, v1) - F(u1, v0) + F(u0, v0)


def wilton_kat():
    cases = []
    cases.append({"what": "unit square at its centre (analytic.rs:14-27)",
                  "verts": rect(-0.5, -0.5, 0.5, 0.5), "p": [0, 0, 0],
                  "expected": 4 * math.asinh(1.0), "source": "closed form 4 asinh 1"})
    for (a, b, x, y, z) in [(1.0, 1.0, 0.3, 0.2, 0.0), (2.0, 0.5, 0.1, 0.4, 0.25),
                            (0.14, 3.0, 0.5, -0.2, 0.07), (1.0, 1.0, 2.5, 3.0, -1.5),
                            (0.3, 0.3, 0.15, 0.15, 1e-3)]:
        cases.append({"what": f"rectangle {a} x {b}, point ({x}, {y}, {z})",
                      "verts": rect(0, 0, a, b), "p": [x, y, z],
                      "expected": corner_formula(a, b, x, y, z),
                      "source": "textbook rectangle corner formula"})
    tri = [[0, 0, 0], [1, 0, 0], [0.2, 0.9, 0], [0.2, 0.9, 0]]
    for p in ([0.3, 0.3, 0.0], [0.4, 0.2, 0.3], [-0.5, 1.2, 0.8]):
        cases.append({"what": f"triangle (repeated fourth vertex), point {p}", "verts": tri,
                      "p": p, "expected": O.potential_integral(np.array(tri), np.array(p)),
                      "source": "oracle"})
    far = [1e3, -2e3, 5e2]
    cases.append({"what": "unit square seen from far away (analytic.rs:29-42)",
                  "verts": rect(-0.5, -0.5, 0.5, 0.5), "p": far,
                  "expected": O.potential_integral(np.array(rect(-0.5, -0.5, 0.5, 0.5)),
                                                   np.array(far)),
                  "monopole": 1.0 / float(np.linalg.norm(far)), "source": "oracle"})
    for c in cases:  # every case, whatever its source, also through the oracle
        c["oracle"] = O.potential_integral(np.array(c["verts"], float), np.array(c["p"], float))
    return cases


def panel_hash(v, c):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(v, dtype="<f8").tobytes())
    h.update(np.ascontiguousarray(c, dtype="<u4").tobytes())
    return h.hexdigest()


def main():
    O.set_threads(min(8, os.cpu_count() or 1))
    with open(os.path.join(HERE, "wilton_kat.json"), "w") as f:
        json.dump(wilton_kat(), f, indent=1)

    groups = W.two_wires()
    v, c, e = O.mesh_groups(groups, refine=1)
    caf, Cm, n = O.solve_capacitance(groups, refine=1)
    assert n == len(v) == 173
    np.savez(os.path.join(HERE, "c1_r1.npz"), verts=v, cond=c.astype(np.uint32), eps=e, C=Cm,
             cap_af=np.array([caf[0], caf[1]]))

    v3, c3, e3 = O.mesh_groups(groups, refine=3)
    caf3, C3, n3 = O.solve_capacitance(groups, refine=3)  # the bridge rule: FMM(4, 4, 1e-8)
    Cd3 = O.cap_dense(v3, c3, 3) * e3[0]  # same panels through the dense LU
    np.savez(os.path.join(HERE, "c1_r3.npz"), n_panels=np.int64(n3), sha256=panel_hash(v3, c3),
             C_fmm=C3, cap_af_fmm=np.array([caf3[0], caf3[1]]), C_dense=Cd3, eps=e3[0])

    p1 = O.lcg_points(900, 12345)
    q1 = np.array([(i % 13) - 6.0 for i in range(900)])
    p2 = O.lcg_points(800, 999)
    q2 = np.array([(i * 7 % 11) - 5.0 for i in range(800)])
    out = {"pts900": p1, "q900": q1, "dense900": O.dense_p2p(p1, q1),
           "pts800": p2, "q800": q2, "dense800": O.dense_p2p(p2, q2)}
    for order in (3, 4, 5):
        for lvl in (3, 4):
            out[f"fmm900_n{order}_L{lvl}"] = O.fmm_matvec(p1, q1, order, lvl)
            out[f"fmm800_n{order}_L{lvl}"] = O.fmm_matvec(p2, q2, order, lvl)
    np.savez(os.path.join(HERE, "lcg_fmm.npz"), **out)

    sph = O.mesh_icosphere((0, 0, 0), 1.0, 2)
    z = np.zeros(len(sph))
    side, gap = 0.2, 0.02
    pl = np.concatenate([O.mesh_plate(0, 0, 0, side, side, 10, 10),
                         O.mesh_plate(0, 0, gap, side, side, 10, 10)])
    res = {"sphere_panels": len(sph), "sphere_radius": 1.0,
           "sphere_C_dense": float(O.cap_dense(sph, z, 1)[0, 0]),
           "sphere_C_fmm_order4_depth3": float(O.cap_fmm(sph, z, 1, order=4, max_level=3,
                                                          tol=1e-8)[0][0, 0]),
           "plate_side": side, "plate_gap": gap, "plate_n": 10,
           "plate_C_dense": O.cap_dense(pl, np.repeat([0, 1], 100), 2).tolist()}
    with open(os.path.join(HERE, "sphere_plate.json"), "w") as f:
        json.dump(res, f, indent=1)
    print("wrote", sorted(x for x in os.listdir(HERE) if not x.endswith(".py")))


if __name__ == "__main__":
    main()
