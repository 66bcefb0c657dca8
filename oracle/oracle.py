"""ctypes view of the CPU FP64 oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``.  It is the checker, never the
thing measured as the product.  The C++ behind it restates the reference
quasi-static capacitance path (crates/pex of UW-ASIC/GPurify); every C++
routine cites the reference file:line it follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_d = C.POINTER(C.c_double)
_u32 = C.POINTER(C.c_uint32)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)
_int = C.POINTER(C.c_int)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc); returns the .so path."""
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.or_potential_integral.restype = C.c_double
        L.or_potential_integral.argtypes = [_d, C.c_int, _d]
        L.or_field_integral.restype = C.c_double
        L.or_field_integral.argtypes = [_d, C.c_int, _d, _d]
        L.or_tri7.argtypes = [_d]
        L.or_solve_dielectric_fmm.argtypes = [C.c_int, _d, _int, _u32, _d, _d, _d, _d, C.c_int,
                                              C.c_int, C.c_int, C.c_double, C.c_int, _d]
        L.or_interp_dweights.argtypes = [C.c_double, C.c_int, _d]
        L.or_lowrank_compress.restype = C.c_int
        L.or_lowrank_compress.argtypes = [_d, C.c_int, C.c_int, C.c_double, _d, _d]
        L.or_m2l_eps.restype = C.c_double
        L.or_m2l_eps.argtypes = [C.c_int]
        L.or_m2l_matrix.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _d]
        L.or_panel_props.argtypes = [C.c_int, _d, _d, _d, _d, _d, _d]
        L.or_assemble_p.argtypes = [C.c_int, _d, _d]
        L.or_lu_solve.argtypes = [C.c_int, _d, C.c_int, _d]
        L.or_cap_solve_dense.argtypes = [C.c_int, _d, _u32, C.c_int, C.c_int, _d, _d]
        L.or_solve_dielectric.argtypes = [C.c_int, _d, _int, _u32, _d, _d, _d, _d, C.c_int, _d]
        L.or_cap_solve_fmm.argtypes = [C.c_int, _d, _u32, C.c_int, C.c_int, C.c_int,
                                       C.c_double, C.c_int, _d, _int, _d]
        L.or_cap_solve_dense_image.argtypes = [C.c_int, _d, _u32, C.c_int, C.c_double, _d]
        L.or_cap_solve_fmm_image.argtypes = [C.c_int, _d, _u32, C.c_int, C.c_double, C.c_int,
                                             C.c_int, C.c_double, C.c_int, _d, _int, _d]
        L.or_cap_solve_fmm_cols.argtypes = [C.c_int, _d, _u32, C.c_int, C.c_int, C.c_int,
                                            C.c_double, C.c_int, C.c_int, C.c_int, _d, _int, _d]
        L.or_fmm_cap_apply.argtypes = [C.c_int, _d, C.c_int, C.c_int, C.c_int, _d, _d]
        L.or_fmm_matvec.argtypes = [C.c_int, _d, _d, C.c_int, C.c_int, C.c_int, C.c_int, _d]
        L.or_dense_p2p.argtypes = [C.c_int, _d, _d, _d]
        L.or_lcg_points.argtypes = [C.c_int, C.c_uint64, _d]
        L.or_cheb_nodes.argtypes = [C.c_int, _d]
        L.or_interp_weights.argtypes = [C.c_double, C.c_int, _d]
        L.or_octree_stats.argtypes = [C.c_int, _d, C.c_int, _d]
        L.or_fmm_near_pattern.restype = C.c_long
        L.or_fmm_near_pattern.argtypes = [C.c_int, _d, C.c_int, C.c_int, _i64, _u32]
        L.or_mesh_plate.argtypes = [C.c_double] * 5 + [C.c_int, C.c_int, _d]
        L.or_mesh_icosphere.argtypes = [C.c_double] * 4 + [C.c_int, _d]
        L.or_process_stack.argtypes = [C.c_int, _d, _d, _d, _d, _d, _d]
        L.or_mesh_groups.argtypes = [C.c_int, _u32, _d, _d, _d, _i64, _i32, C.c_double,
                                     C.c_int, C.c_int, _int, _d, _u32, _d]
        L.or_solve_capacitance.argtypes = [C.c_int, _u32, _d, _d, _d, _i64, _i32, C.c_double,
                                           C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                           C.c_double, _int, _u32, _d, _d, _int]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_get_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a, t=_d):
    return a.ctypes.data_as(t) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n: int) -> None:
    lib().or_set_threads(int(n))


def get_threads() -> int:
    return int(lib().or_get_threads())


def potential_integral(verts, p) -> float:
    v = _f64(verts).reshape(-1, 3)
    return float(lib().or_potential_integral(_p(v), v.shape[0], _p(_f64(p))))


def tri7():
    """The 7-point triangle rule (panel.rs:213-221): rows (l0, l1, l2, weight)."""
    t = np.zeros((7, 4))
    lib().or_tri7(_p(t))
    return t


def lowrank_compress(a, eps):
    """LowRankApprox::compress (linalg.rs:277-381): (U, Vt, k) with A ~ U @ Vt."""
    a = _f64(a)
    m, n = a.shape
    kk = min(m, n)
    u = np.zeros(m * kk)
    vt = np.zeros(kk * n)
    k = lib().or_lowrank_compress(_p(a), m, n, float(eps), _p(u), _p(vt))
    return u[:m * k].reshape(m, k), vt[:k * n].reshape(k, n), k


def m2l_eps(order) -> float:
    return float(lib().or_m2l_eps(int(order)))


def m2l_matrix(order, off):
    n3 = order ** 3
    g = np.zeros((n3, n3))
    lib().or_m2l_matrix(int(order), int(off[0]), int(off[1]), int(off[2]), _p(g))
    return g


def field_integral(verts, p, n_obs) -> float:
    v = _f64(verts).reshape(-1, 3)
    return float(lib().or_field_integral(_p(v), v.shape[0], _p(_f64(p)), _p(_f64(n_obs))))


def panel_props(verts4):
    v = _f64(verts4).reshape(-1, 4, 3)
    n = v.shape[0]
    cx, cy, cz, ar = (np.empty(n) for _ in range(4))
    nrm = np.empty((n, 3))
    lib().or_panel_props(n, _p(v), _p(cx), _p(cy), _p(cz), _p(ar), _p(nrm))
    return np.stack([cx, cy, cz], axis=1), ar, nrm


def assemble_p(verts4):
    v = _f64(verts4).reshape(-1, 4, 3)
    n = v.shape[0]
    P = np.empty((n, n))
    lib().or_assemble_p(n, _p(v), _p(P))
    return P


def cap_dense(verts4, cond, ncond, method=0, want_p=False):
    v = _f64(verts4).reshape(-1, 4, 3)
    n = v.shape[0]
    c = np.ascontiguousarray(cond, dtype=np.uint32)
    Cm = np.zeros((ncond, ncond))
    P = np.empty((n, n)) if want_p else None
    rc = lib().or_cap_solve_dense(n, _p(v), _p(c, _u32), ncond, method, _p(Cm), _p(P))
    if rc:
        raise RuntimeError(f"oracle dense solve failed ({rc})")
    return (Cm, P) if want_p else Cm


def solve_dielectric(verts4, role, cond, eps, eps_out, eps_in, ref, ncond):
    """cap/dielectric.rs:47-151 (C in farads, permittivities relative): role 0
    conductor (cond, eps = eps_surrounding), 1 dielectric interface (eps_out,
    eps_in, ref = a point on the inner side, (n, 3))."""
    v = _f64(verts4).reshape(-1, 4, 3)
    n = v.shape[0]
    r = np.ascontiguousarray(role, dtype=np.int32)
    c = np.ascontiguousarray(cond, dtype=np.uint32)
    e, eo, ei = _f64(eps), _f64(eps_out), _f64(eps_in)
    rf = _f64(ref).reshape(-1, 3)
    Cm = np.zeros((ncond, ncond))
    rc = lib().or_solve_dielectric(n, _p(v), r.ctypes.data_as(_int), _p(c, _u32), _p(e), _p(eo),
                                   _p(ei), _p(rf), ncond, _p(Cm))
    if rc:
        raise RuntimeError(f"oracle dielectric solve failed ({rc})")
    return Cm


def solve_dielectric_fmm(verts4, role, cond, eps, eps_out, eps_in, ref, ncond, order=4,
                         max_level=4, tol=1e-10, m2l_mode=0):
    """The dielectric problem of solve_dielectric on the FMM operator (no
    reference counterpart): C in farads."""
    v = _f64(verts4).reshape(-1, 4, 3)
    n = v.shape[0]
    r = np.ascontiguousarray(role, dtype=np.int32)
    c = np.ascontiguousarray(cond, dtype=np.uint32)
    e, eo, ei = _f64(eps), _f64(eps_out), _f64(eps_in)
    rf = _f64(ref).reshape(-1, 3)
    Cm = np.zeros((ncond, ncond))
    rc = lib().or_solve_dielectric_fmm(n, _p(v), r.ctypes.data_as(_int), _p(c, _u32), _p(e),
                                       _p(eo), _p(ei), _p(rf), ncond, order, max_level, tol,
                                       m2l_mode, _p(Cm))
    if rc:
        raise RuntimeError(f"oracle FMM dielectric solve failed ({rc})")
    return Cm


def interp_dweights(x, n):
    """d/dx of the 1-D Chebyshev interpolation weights at x."""
    w = np.zeros(n)
    lib().or_interp_dweights(float(x), int(n), _p(w))
    return w


def cap_fmm(verts4, cond, ncond, order=4, max_level=4, tol=1e-8, m2l_mode=0):
    v = _f64(verts4).reshape(-1, 4, 3)
    n = v.shape[0]
    c = np.ascontiguousarray(cond, dtype=np.uint32)
    Cm = np.zeros((ncond, ncond))
    it = np.zeros(ncond, dtype=np.int32)
    rs = np.zeros(ncond)
    rc = lib().or_cap_solve_fmm(n, _p(v), _p(c, _u32), ncond, order, max_level, tol, m2l_mode,
                                _p(Cm), it.ctypes.data_as(_int), _p(rs))
    if rc:
        raise RuntimeError(f"oracle FMM solve failed ({rc})")
    return Cm, it, rs


def cap_dense_image(verts4, cond, ncond, zg):
    """Conductors over a grounded plane z = zg by images, dense LU (no
    reference counterpart): C in farads, the plane as the reference."""
    v = _f64(verts4).reshape(-1, 4, 3)
    c = np.ascontiguousarray(cond, dtype=np.uint32)
    Cm = np.zeros((ncond, ncond))
    rc = lib().or_cap_solve_dense_image(v.shape[0], _p(v), _p(c, _u32), ncond, float(zg), _p(Cm))
    if rc:
        raise RuntimeError(f"oracle image dense solve failed ({rc})")
    return Cm


def cap_fmm_image(verts4, cond, ncond, zg, order=4, max_level=4, tol=1e-8, m2l_mode=0):
    """The image problem of cap_dense_image on the FMM operator of the panels
    and their mirrors (antisymmetric right-hand sides): (C, iters, resid)."""
    v = _f64(verts4).reshape(-1, 4, 3)
    c = np.ascontiguousarray(cond, dtype=np.uint32)
    Cm = np.zeros((ncond, ncond))
    it = np.zeros(ncond, dtype=np.int32)
    rs = np.zeros(ncond)
    rc = lib().or_cap_solve_fmm_image(v.shape[0], _p(v), _p(c, _u32), ncond, float(zg), order,
                                      max_level, tol, m2l_mode, _p(Cm), it.ctypes.data_as(_int),
                                      _p(rs))
    if rc:
        raise RuntimeError(f"oracle image FMM solve failed ({rc})")
    return Cm, it, rs


def cap_fmm_cols(verts4, cond, ncond, j0, nj, order=4, max_level=4, tol=1e-8, m2l_mode=0):
    """FMM solve of conductor columns [j0, j0+nj): (C, iters, (setup_s, solve_s))."""
    v = _f64(verts4).reshape(-1, 4, 3)
    n = v.shape[0]
    c = np.ascontiguousarray(cond, dtype=np.uint32)
    Cm = np.zeros((ncond, ncond))
    it = np.zeros(max(1, nj), dtype=np.int32)
    sec = np.zeros(2)
    rc = lib().or_cap_solve_fmm_cols(n, _p(v), _p(c, _u32), ncond, order, max_level, tol,
                                     m2l_mode, j0, nj, _p(Cm), it.ctypes.data_as(_int), _p(sec))
    if rc:
        raise RuntimeError(f"oracle FMM solve failed ({rc})")
    return Cm, it, (float(sec[0]), float(sec[1]))


def fmm_cap_apply(verts4, sigma, order=4, max_level=4, m2l_mode=0):
    v = _f64(verts4).reshape(-1, 4, 3)
    n = v.shape[0]
    s = _f64(sigma)
    y = np.empty(n)
    lib().or_fmm_cap_apply(n, _p(v), order, max_level, m2l_mode, _p(s), _p(y))
    return y


def fmm_matvec(pts, q, order, max_level, with_p2p=True, m2l_mode=0):
    P = _f64(pts).reshape(-1, 3)
    y = np.empty(P.shape[0])
    lib().or_fmm_matvec(P.shape[0], _p(P), _p(_f64(q)), order, max_level, int(with_p2p),
                        m2l_mode, _p(y))
    return y


def dense_p2p(pts, q):
    P = _f64(pts).reshape(-1, 3)
    y = np.empty(P.shape[0])
    lib().or_dense_p2p(P.shape[0], _p(P), _p(_f64(q)), _p(y))
    return y


def lcg_points(n, seed):
    pts = np.empty((n, 3))
    lib().or_lcg_points(n, seed, _p(pts))
    return pts


def cheb_nodes(n):
    v = np.empty(n)
    lib().or_cheb_nodes(n, _p(v))
    return v


def interp_weights(x, n):
    w = np.empty(n)
    lib().or_interp_weights(float(x), n, _p(w))
    return w


def octree_stats(pts, max_level):
    P = _f64(pts).reshape(-1, 3)
    out = np.empty(6)
    lib().or_octree_stats(P.shape[0], _p(P), max_level, _p(out))
    keys = ["nodes", "leaves", "leaf_points", "interaction_pairs", "min_sep", "near_pairs"]
    return dict(zip(keys, out.tolist()))


def fmm_near_pattern(verts4, order, max_level):
    v = _f64(verts4).reshape(-1, 4, 3)
    n = v.shape[0]
    nnz = lib().or_fmm_near_pattern(n, _p(v), order, max_level, None, None)
    off = np.empty(n + 1, dtype=np.int64)
    idx = np.empty(nnz, dtype=np.uint32)
    lib().or_fmm_near_pattern(n, _p(v), order, max_level, _p(off, _i64), _p(idx, _u32))
    return off, idx


def mesh_plate(cx, cy, z0, lx, ly, nx, ny):
    n = lib().or_mesh_plate(cx, cy, z0, lx, ly, nx, ny, None)
    v = np.empty((n, 4, 3))
    lib().or_mesh_plate(cx, cy, z0, lx, ly, nx, ny, _p(v))
    return v


def mesh_icosphere(center, r, subdiv):
    n = lib().or_mesh_icosphere(center[0], center[1], center[2], r, subdiv, None)
    v = np.empty((n, 4, 3))
    lib().or_mesh_icosphere(center[0], center[1], center[2], r, subdiv, _p(v))
    return v


def process_stack(thick, height, k):
    t_in, h_in, k_in = _f64(thick), _f64(height), _f64(k)
    n = t_in.shape[0]
    t, h, kk = np.empty(n), np.empty(n), np.empty(n)
    lib().or_process_stack(n, _p(t_in), _p(h_in), _p(k_in), _p(t), _p(h), _p(kk))
    return t, h, kk


def _groups_arrays(groups):
    """groups: list of dicts {net, height_nm, thickness_nm, k, rects: (m,4) int}."""
    net = np.array([g["net"] for g in groups], dtype=np.uint32)
    h = np.array([g["height_nm"] for g in groups], dtype=np.float64)
    t = np.array([g["thickness_nm"] for g in groups], dtype=np.float64)
    k = np.array([g["k"] for g in groups], dtype=np.float64)
    counts = [len(g["rects"]) for g in groups]
    off = np.zeros(len(groups) + 1, dtype=np.int64)
    off[1:] = np.cumsum(counts)
    rects = (np.concatenate([np.asarray(g["rects"], dtype=np.int32).reshape(-1, 4) for g in groups])
             if groups else np.zeros((0, 4), np.int32))
    return net, h, t, k, off, np.ascontiguousarray(rects, dtype=np.int32)


def mesh_groups(groups, xy_scale=1e-9, refine=1, max_panels=0):
    net, h, t, k, off, rects = _groups_arrays(groups)
    n = C.c_int(0)
    L = lib()
    args = [len(groups), _p(net, _u32), _p(h), _p(t), _p(k), _p(off, _i64), _p(rects, _i32),
            xy_scale, refine, max_panels, C.byref(n)]
    rc = L.or_mesh_groups(*args, None, None, None)
    if rc:
        raise RuntimeError(f"oracle mesher failed ({rc}, {n.value} panels)")
    v = np.empty((n.value, 4, 3))
    cond = np.empty(n.value, dtype=np.uint32)
    eps = np.empty(n.value)
    rc = L.or_mesh_groups(*args, _p(v), _p(cond, _u32), _p(eps))
    if rc:
        raise RuntimeError(f"oracle mesher failed ({rc})")
    return v, cond, eps


def solve_capacitance(groups, xy_scale=1e-9, refine=1, max_panels=0, mode=0, order=4,
                      max_level=4, tol=1e-8):
    """bridge.rs:307-443 restated: returns ({net: cap_af}, C with eps applied, n_panels)."""
    net, h, t, k, off, rects = _groups_arrays(groups)
    nn_max = len(set(int(x) for x in net))
    nets = np.zeros(nn_max, dtype=np.uint32)
    caf = np.zeros(nn_max)
    Cm = np.zeros((nn_max + 1) * (nn_max + 1))
    n_nets = C.c_int(0)
    n_pan = C.c_int(0)
    rc = lib().or_solve_capacitance(len(groups), _p(net, _u32), _p(h), _p(t), _p(k),
                                    _p(off, _i64), _p(rects, _i32), xy_scale, refine, max_panels,
                                    mode, order, max_level, tol, C.byref(n_nets), _p(nets, _u32),
                                    _p(caf), _p(Cm), C.byref(n_pan))
    if rc:
        raise RuntimeError(f"oracle solve_capacitance failed ({rc})")
    return ({int(a): float(b) for a, b in zip(nets, caf)}, Cm.reshape(nn_max + 1, nn_max + 1),
            n_pan.value)
