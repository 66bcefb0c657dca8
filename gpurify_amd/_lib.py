"""ctypes binding of libgpurify (include/gpurify.h).

The library is loaded from this package directory (built in-tree by
``gpurify_amd/csrc/Makefile``).  There is no fallback: when the shared
library is missing, importing the binding raises, and every solve goes
through the HIP kernels or fails with the library's status code.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# GPURIFY_LIB_VARIANT=san loads the host-sanitizer build (libgpurify_san.so,
# `make -C gpurify_amd/csrc sanitize`), for the host-code tests under ASan
LIB_PATH = os.path.join(_HERE, "libgpurify_san.so" if os.environ.get("GPURIFY_LIB_VARIANT") == "san"
                        else "libgpurify.so")

GP_OK = 0
GP_ERR_INVALID_INPUT = 1
GP_ERR_NO_DEVICE = 2
GP_ERR_OOM = 3
GP_ERR_NOT_CONVERGED = 4
GP_ERR_NON_FINITE = 5
GP_ERR_INTERNAL = 6
GP_ERR_GRID_CAPACITY = 7
GP_ERR_PANEL_CAPACITY = 8
GP_ERR_SINGULAR = 9
GP_ERR_EMPTY = 10

ERROR_NAMES = {
    0: "OK", 1: "INVALID_INPUT", 2: "NO_DEVICE", 3: "OOM", 4: "NOT_CONVERGED",
    5: "NON_FINITE", 6: "INTERNAL", 7: "GRID_CAPACITY", 8: "PANEL_CAPACITY", 9: "SINGULAR",
    10: "EMPTY",
}

GP_PREC_FP64 = 0
GP_PREC_MIXED = 1
GP_METHOD_AUTO = 0
GP_METHOD_DENSE = 1
GP_METHOD_FMM = 2
GP_M2L_FP64 = 0
GP_M2L_FP32 = 1
GP_M2L_BF16 = 2
GP_M2L_FP64_VALU = 3
GP_M2L_BF16_VALU = 4
GP_BASIS_FP32 = 0
GP_BASIS_FP64 = 1
GP_NEAR_FP32 = 0
GP_NEAR_FP16 = 1
GP_GROUND_NONE = 0
GP_GROUND_IMAGE = 1


class GpPanels(C.Structure):
    _fields_ = [("n_panels", C.c_uint32), ("n_conductors", C.c_uint32),
                ("vertices", C.POINTER(C.c_double)), ("conductor", C.POINTER(C.c_uint32))]


class GpCapOpts(C.Structure):
    _fields_ = [("precision", C.c_uint32), ("method", C.c_uint32), ("fmm_order", C.c_uint32),
                ("max_level", C.c_uint32), ("leaf_target", C.c_uint32), ("restart", C.c_uint32),
                ("max_restarts", C.c_uint32), ("rhs_block", C.c_uint32), ("tol", C.c_double),
                ("inner_tol", C.c_double), ("max_refine", C.c_uint32), ("m2l_inner", C.c_uint32),
                ("m2l_outer", C.c_uint32), ("strict", C.c_uint32), ("dense_limit", C.c_uint32),
                ("rhs_first", C.c_uint32), ("rhs_count", C.c_uint32),
                ("inner_passes", C.c_uint32), ("inner_basis", C.c_uint32),
                ("m2l_rank", C.c_uint32), ("step_graphs", C.c_uint32),
                ("m2l_compress", C.c_uint32), ("fuse_leaf", C.c_uint32),
                ("skip_converged", C.c_uint32), ("batch_groups", C.c_uint32),
                ("near_inner", C.c_uint32), ("inner_order", C.c_uint32),
                ("ground", C.c_uint32), ("ground_z", C.c_double),
                ("m2l_tile_order", C.c_uint32)]


class GpParamMap(C.Structure):
    _fields_ = [("n_params", C.c_uint32), ("param_off", C.POINTER(C.c_uint32)),
                ("panel", C.POINTER(C.c_uint32)), ("dverts", C.POINTER(C.c_double))]


class GpPanelRole(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("conductor", C.c_uint32), ("eps_surrounding", C.c_double),
                ("eps_out", C.c_double), ("eps_in", C.c_double), ("reference", C.c_double * 3)]


class GpCapReport(C.Structure):
    _fields_ = [("n_panels", C.c_uint32), ("n_conductors", C.c_uint32),
                ("method_used", C.c_uint32), ("precision_used", C.c_uint32),
                ("iterations", C.c_uint32), ("matvecs", C.c_uint32),
                ("refine_steps", C.c_uint32), ("converged", C.c_uint32),
                ("max_level", C.c_uint32), ("n_leaves", C.c_uint32), ("n_boxes", C.c_uint32),
                ("fmm_order", C.c_uint32), ("near_nnz", C.c_uint64), ("m2l_pairs", C.c_uint64),
                ("final_rel_residual", C.c_double), ("seconds_total", C.c_double),
                ("seconds_setup", C.c_double), ("seconds_solve", C.c_double),
                ("seconds_near", C.c_double), ("seconds_far", C.c_double),
                ("seconds_orth", C.c_double), ("seconds_upload", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class GpVolumeGroup(C.Structure):
    _fields_ = [("net", C.c_uint32), ("n_rects", C.c_uint32), ("height_nm", C.c_double),
                ("thickness_nm", C.c_double), ("dielectric_k", C.c_double),
                ("rects", C.POINTER(C.c_int32))]


# Every symbol include/gpurify.h declares (checked by the CPU test suite).
EXPORTS = [
    "gp_ctx_create", "gp_ctx_destroy", "gp_last_error", "gp_device_ready", "gp_device_info",
    "gp_version", "gp_cap_opts_default", "gp_cap_solve", "gp_cap_solve_batched", "gp_cap_jacobian",
    "gp_fmm_apply", "gp_mesh_groups",
    "gp_solve_capacitance", "gp_profile_enable", "gp_profile_report", "gp_tune_force",
    "gp_tune_report", "gp_qui_parse", "gp_solve_resistance", "gp_cap_solve_dielectric",
    "gp_m2l_operator", "gp_m2l_factors", "gp_cap_solve_dense_noalloc", "gp_device_synchronize",
    "gp_cap_solve_dielectric_fmm",
]

_lib = None


def load():
    """Load libgpurify.so; raises OSError (no fallback) when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"libgpurify.so not built at {LIB_PATH}; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.gp_ctx_create.argtypes = [C.POINTER(vp), C.c_uint32]
    L.gp_ctx_create.restype = C.c_int
    L.gp_ctx_destroy.argtypes = [vp]
    L.gp_ctx_destroy.restype = None
    L.gp_last_error.argtypes = [vp]
    L.gp_last_error.restype = C.c_char_p
    L.gp_device_ready.argtypes = []
    L.gp_device_ready.restype = C.c_int
    L.gp_device_synchronize.argtypes = [vp]
    L.gp_device_synchronize.restype = C.c_int
    L.gp_device_info.argtypes = [vp, C.c_char_p, C.c_size_t]
    L.gp_device_info.restype = C.c_int
    L.gp_version.argtypes = []
    L.gp_version.restype = C.c_char_p
    L.gp_cap_opts_default.argtypes = [C.POINTER(GpCapOpts)]
    L.gp_cap_opts_default.restype = None
    L.gp_cap_solve.argtypes = [vp, C.POINTER(GpPanels), C.POINTER(GpCapOpts),
                               C.POINTER(C.c_double), C.POINTER(GpCapReport)]
    L.gp_cap_solve.restype = C.c_int
    L.gp_cap_solve_batched.argtypes = [vp, C.POINTER(GpPanels), C.c_uint32, C.POINTER(GpCapOpts),
                                       C.POINTER(C.c_double), C.POINTER(GpCapReport)]
    L.gp_cap_solve_batched.restype = C.c_int
    L.gp_cap_jacobian.argtypes = [vp, C.POINTER(GpPanels), C.POINTER(GpCapOpts),
                                  C.POINTER(GpParamMap), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double), C.POINTER(GpCapReport)]
    L.gp_cap_jacobian.restype = C.c_int
    L.gp_fmm_apply.argtypes = [vp, C.POINTER(GpPanels), C.POINTER(GpCapOpts), C.c_uint32,
                               C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.gp_fmm_apply.restype = C.c_int
    L.gp_mesh_groups.argtypes = [C.POINTER(GpVolumeGroup), C.c_uint32, C.c_double, C.c_uint32,
                                 C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_double),
                                 C.POINTER(C.c_uint32), C.POINTER(C.c_double)]
    L.gp_mesh_groups.restype = C.c_int
    L.gp_solve_capacitance.argtypes = [vp, C.POINTER(GpVolumeGroup), C.c_uint32, C.c_double,
                                       C.c_uint32, C.c_uint32, C.POINTER(GpCapOpts),
                                       C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                       C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(GpCapReport)]
    L.gp_solve_capacitance.restype = C.c_int
    L.gp_profile_enable.argtypes = [vp, C.c_int]
    L.gp_profile_enable.restype = C.c_int
    L.gp_profile_report.argtypes = [vp, C.c_char_p, C.c_size_t]
    L.gp_profile_report.restype = C.c_int
    L.gp_cap_solve_dielectric.argtypes = [vp, C.POINTER(GpPanels), C.POINTER(GpPanelRole),
                                          C.POINTER(C.c_double), C.POINTER(GpCapReport)]
    L.gp_cap_solve_dielectric.restype = C.c_int
    L.gp_cap_solve_dielectric_fmm.argtypes = [vp, C.POINTER(GpPanels), C.POINTER(GpPanelRole),
                                              C.POINTER(GpCapOpts), C.POINTER(C.c_double),
                                              C.POINTER(GpCapReport)]
    L.gp_cap_solve_dielectric_fmm.restype = C.c_int
    L.gp_qui_parse.argtypes = [C.c_char_p, C.c_size_t, C.c_double, C.POINTER(C.c_uint32),
                               C.POINTER(C.c_uint32), C.POINTER(C.c_size_t), C.POINTER(C.c_double),
                               C.POINTER(C.c_uint32), C.POINTER(C.c_uint8), C.c_char_p]
    L.gp_qui_parse.restype = C.c_int
    L.gp_solve_resistance.argtypes = [C.POINTER(GpResPolygon), C.c_uint32, C.c_double,
                                      C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                      C.POINTER(C.c_double)]
    L.gp_solve_resistance.restype = C.c_int
    L.gp_cap_solve_dense_noalloc.argtypes = [vp, C.POINTER(GpPanels), C.POINTER(C.c_double),
                                             C.POINTER(GpCapReport)]
    L.gp_cap_solve_dense_noalloc.restype = C.c_int
    L.gp_m2l_operator.argtypes = [C.c_uint32, C.c_int32, C.c_int32, C.c_int32, C.c_uint32,
                                  C.POINTER(C.c_double), C.POINTER(C.c_uint32)]
    L.gp_m2l_operator.restype = C.c_int
    L.gp_m2l_factors.argtypes = [C.c_uint32, C.c_int32, C.c_int32, C.c_int32,
                                 C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 C.POINTER(C.c_uint32)]
    L.gp_m2l_factors.restype = C.c_int
    L.gp_tune_force.argtypes = [vp, C.c_int]
    L.gp_tune_force.restype = C.c_int
    L.gp_tune_report.argtypes = [vp, C.c_char_p, C.c_size_t]
    L.gp_tune_report.restype = C.c_int
    _lib = L
    return L


class GpError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {message}")
        self.code = code
        self.message = message


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _uptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def default_opts(**kw) -> GpCapOpts:
    o = GpCapOpts()
    load().gp_cap_opts_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def device_ready() -> bool:
    return bool(load().gp_device_ready())


def version() -> str:
    return load().gp_version().decode()


class Context:
    """One device context (gp_ctx).  Raises GpError(NO_DEVICE) without a gfx950."""

    def __init__(self, device: int = 0):
        L = load()
        h = C.c_void_p()
        rc = L.gp_ctx_create(C.byref(h), device)
        if rc != GP_OK:
            raise GpError(rc, L.gp_last_error(None).decode())
        self._h = h
        self._L = L

    def close(self):
        if getattr(self, "_h", None):
            self._L.gp_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != GP_OK:
            raise GpError(rc, self._L.gp_last_error(self._h).decode())

    def profile(self, enable: bool = True):
        self._check(self._L.gp_profile_enable(self._h, 1 if enable else 0))

    def profile_report(self) -> dict:
        import json
        buf = C.create_string_buffer(1 << 16)
        self._check(self._L.gp_profile_report(self._h, buf, 1 << 16))
        return json.loads(buf.value.decode())

    def tune_force(self, variant: int):
        """Pin every tunable kernel to `variant` (< 0: autotune again)."""
        self._check(self._L.gp_tune_force(self._h, int(variant)))

    def tune_report(self) -> list:
        import json
        buf = C.create_string_buffer(1 << 16)
        self._check(self._L.gp_tune_report(self._h, buf, 1 << 16))
        return json.loads(buf.value.decode())

    def synchronize(self):
        """Wait for all work on the context's device (gp_device_synchronize)."""
        self._check(self._L.gp_device_synchronize(self._h))

    def info(self) -> str:
        buf = C.create_string_buffer(256)
        self._check(self._L.gp_device_info(self._h, buf, 256))
        return buf.value.decode()

    def cap_solve(self, verts4, cond, ncond, opts: GpCapOpts | None = None):
        """Maxwell C (ncond x ncond, F, vacuum) and the solve report."""
        v = np.ascontiguousarray(verts4, dtype=np.float64).reshape(-1, 4, 3)
        c = np.ascontiguousarray(cond, dtype=np.uint32)
        p = GpPanels(v.shape[0], ncond, _dptr(v), _uptr(c))
        Cm = np.zeros((ncond, ncond))
        rep = GpCapReport()
        o = opts if opts is not None else default_opts()
        self._check(self._L.gp_cap_solve(self._h, C.byref(p), C.byref(o), _dptr(Cm), C.byref(rep)))
        return Cm, rep.as_dict()

    def cap_solve_dense_noalloc(self, verts4, cond, ncond):
        """The dense FP64 solve in the library's fixed device workspace
        (gp_cap_solve_dense_noalloc): C (ncond x ncond, F, vacuum) and report."""
        v = np.ascontiguousarray(verts4, dtype=np.float64).reshape(-1, 4, 3)
        c = np.ascontiguousarray(cond, dtype=np.uint32)
        p = GpPanels(v.shape[0], ncond, _dptr(v), _uptr(c))
        Cm = np.zeros((ncond, ncond))
        rep = GpCapReport()
        self._check(self._L.gp_cap_solve_dense_noalloc(self._h, C.byref(p), _dptr(Cm),
                                                       C.byref(rep)))
        return Cm, rep.as_dict()

    def cap_solve_dielectric(self, verts4, role, cond, eps, eps_out, eps_in, ref, ncond,
                             opts: GpCapOpts | None = None):
        """cap::solve_dielectric on the device: C (ncond x ncond, F,
        permittivities applied), dense FP64, or with opts the FMM form.
        role[k] 0 = conductor (cond[k], eps[k] surrounding), 1 = dielectric
        interface (eps_out[k], eps_in[k], ref[k] a point on the inner side)."""
        v = np.ascontiguousarray(verts4, dtype=np.float64).reshape(-1, 4, 3)
        n = v.shape[0]
        c = np.ascontiguousarray(np.where(np.asarray(role) == 0, cond, 0), dtype=np.uint32)
        roles = (GpPanelRole * max(1, n))()
        ref = np.asarray(ref, dtype=np.float64).reshape(-1, 3)
        for k in range(n):
            r = roles[k]
            r.kind = int(role[k])
            r.conductor = int(c[k])
            r.eps_surrounding = float(eps[k])
            r.eps_out = float(eps_out[k])
            r.eps_in = float(eps_in[k])
            r.reference[:] = [float(x) for x in ref[k]]
        p = GpPanels(n, ncond, _dptr(v), _uptr(c))
        Cm = np.zeros((ncond, ncond))
        rep = GpCapReport()
        if opts is None:
            self._check(self._L.gp_cap_solve_dielectric(self._h, C.byref(p), roles, _dptr(Cm),
                                                         C.byref(rep)))
        else:  # the FMM form (gp_cap_solve_dielectric_fmm)
            self._check(self._L.gp_cap_solve_dielectric_fmm(self._h, C.byref(p), roles,
                                                             C.byref(opts), _dptr(Cm),
                                                             C.byref(rep)))
        return Cm, rep.as_dict()

    def cap_jacobian(self, verts4, cond, ncond, params, opts: GpCapOpts | None = None):
        """C (ncond x ncond) and J = dC/dtheta (n_params x ncond x ncond).
        params: [(panel indices, dverts (m x 12 or m x 4 x 3), metres per
        unit theta)]."""
        v = np.ascontiguousarray(verts4, dtype=np.float64).reshape(-1, 4, 3)
        c = np.ascontiguousarray(cond, dtype=np.uint32)
        p = GpPanels(v.shape[0], ncond, _dptr(v), _uptr(c))
        off = np.zeros(len(params) + 1, dtype=np.uint32)
        for i, (idx, _) in enumerate(params):
            off[i + 1] = off[i] + len(idx)
        pan = np.ascontiguousarray(np.concatenate([np.asarray(i, dtype=np.uint32) for i, _ in params])
                                   if params else np.zeros(0, np.uint32), dtype=np.uint32)
        dv = np.ascontiguousarray(np.concatenate([np.asarray(d, dtype=np.float64).reshape(-1, 12)
                                                  for _, d in params])
                                  if params else np.zeros((0, 12)), dtype=np.float64)
        pm = GpParamMap(len(params), _uptr(off), _uptr(pan), _dptr(dv))
        Cm = np.zeros((ncond, ncond))
        Jm = np.zeros((max(len(params), 1), ncond, ncond))
        rep = GpCapReport()
        o = opts if opts is not None else default_opts()
        self._check(self._L.gp_cap_jacobian(self._h, C.byref(p), C.byref(o), C.byref(pm), _dptr(Cm),
                                            _dptr(Jm), C.byref(rep)))
        return Cm, Jm[:len(params)], rep.as_dict()

    def cap_solve_batched(self, problems, opts: GpCapOpts | None = None):
        """problems: [(verts4, cond, ncond)] -> ([C_p], [report_p])."""
        keep, arr = [], (GpPanels * len(problems))()
        for i, (verts4, cond, ncond) in enumerate(problems):
            v = np.ascontiguousarray(verts4, dtype=np.float64).reshape(-1, 4, 3)
            c = np.ascontiguousarray(cond, dtype=np.uint32)
            keep += [v, c]
            arr[i] = GpPanels(v.shape[0], ncond, _dptr(v), _uptr(c))
        sizes = [nc * nc for _, _, nc in problems]
        Cflat = np.zeros(sum(sizes))
        reps = (GpCapReport * len(problems))()
        o = opts if opts is not None else default_opts()
        self._check(self._L.gp_cap_solve_batched(self._h, arr, len(problems), C.byref(o),
                                                  _dptr(Cflat), reps))
        out, off = [], 0
        for (_, _, nc), sz in zip(problems, sizes):
            out.append(Cflat[off:off + sz].reshape(nc, nc))
            off += sz
        return out, [r.as_dict() for r in reps]

    def fmm_apply(self, verts4, X, opts: GpCapOpts | None = None, precision=GP_PREC_FP64):
        v = np.ascontiguousarray(verts4, dtype=np.float64).reshape(-1, 4, 3)
        n = v.shape[0]
        X = np.ascontiguousarray(X, dtype=np.float64).reshape(n, -1)
        nr = X.shape[1]
        Y = np.zeros_like(X)
        c = np.zeros(n, dtype=np.uint32)
        p = GpPanels(n, 1, _dptr(v), _uptr(c))
        o = opts if opts is not None else default_opts(method=GP_METHOD_FMM)
        self._check(self._L.gp_fmm_apply(self._h, C.byref(p), C.byref(o), nr, precision, _dptr(X),
                                         _dptr(Y)))
        return Y

    def solve_capacitance(self, groups, xy_scale=1e-9, refine=1, max_panels=0,
                          opts: GpCapOpts | None = None):
        """bridge solve_capacitance: ({net: cap_af}, C with eps applied, report)."""
        arr, keep = _group_array(groups)
        nn_max = len({int(g["net"]) for g in groups})
        nets = np.zeros(nn_max, dtype=np.uint32)
        caf = np.zeros(nn_max)
        Cm = np.zeros((nn_max + 1, nn_max + 1))
        nn = C.c_uint32(0)
        rep = GpCapReport()
        o = opts if opts is not None else default_opts()
        self._check(self._L.gp_solve_capacitance(self._h, arr, len(groups), xy_scale, refine,
                                                 max_panels, C.byref(o), C.byref(nn),
                                                 _uptr(nets), _dptr(caf), _dptr(Cm),
                                                 C.byref(rep)))
        del keep
        return {int(a): float(b) for a, b in zip(nets, caf)}, Cm, rep.as_dict()


def _group_array(groups):
    keep = []
    arr = (GpVolumeGroup * max(1, len(groups)))()
    for i, g in enumerate(groups):
        r = np.ascontiguousarray(np.asarray(g["rects"], dtype=np.int32).reshape(-1, 4))
        keep.append(r)
        arr[i].net = int(g["net"])
        arr[i].n_rects = r.shape[0]
        arr[i].height_nm = float(g["height_nm"])
        arr[i].thickness_nm = float(g["thickness_nm"])
        arr[i].dielectric_k = float(g["k"])
        arr[i].rects = r.ctypes.data_as(C.POINTER(C.c_int32))
    return arr, keep


def mesh_groups(groups, xy_scale=1e-9, refine=1, max_panels=0):
    """Product mesher (host C++): (verts (n,4,3), cond, eps).  Raises GpError."""
    L = load()
    arr, keep = _group_array(groups)
    n = C.c_uint32(0)
    rc = L.gp_mesh_groups(arr, len(groups), xy_scale, refine, max_panels, C.byref(n), None, None,
                          None)
    if rc != GP_OK:
        raise GpError(rc, L.gp_last_error(None).decode())
    v = np.zeros((n.value, 4, 3))
    c = np.zeros(n.value, dtype=np.uint32)
    e = np.zeros(n.value)
    rc = L.gp_mesh_groups(arr, len(groups), xy_scale, refine, max_panels, C.byref(n), _dptr(v),
                          _uptr(c), _dptr(e))
    if rc != GP_OK:
        raise GpError(rc, L.gp_last_error(None).decode())
    del keep
    return v, c, e


def parse_qui(text, scale=1.0):
    """FastCap quickif geometry (cap/geometry.rs:56-135) through the product
    parser: dict with vertices (n,4,3) (a triangle repeats its last vertex),
    conductor (n,) uint32, n_vertices (n,) uint8 and names (list of str).
    Raises GpError (GP_ERR_INVALID_INPUT, "line <n>: ...") on a bad line."""
    L = load()
    raw = text.encode() if isinstance(text, str) else bytes(text)
    n, nc, nl = C.c_uint32(0), C.c_uint32(0), C.c_size_t(0)
    rc = L.gp_qui_parse(raw, len(raw), scale, C.byref(n), C.byref(nc), C.byref(nl), None, None,
                        None, None)
    if rc != GP_OK:
        raise GpError(rc, L.gp_last_error(None).decode())
    v = np.zeros((n.value, 4, 3))
    c = np.zeros(n.value, dtype=np.uint32)
    nv = np.zeros(n.value, dtype=np.uint8)
    names = C.create_string_buffer(max(1, nl.value))
    rc = L.gp_qui_parse(raw, len(raw), scale, C.byref(n), C.byref(nc), C.byref(nl), _dptr(v),
                        _uptr(c), nv.ctypes.data_as(C.POINTER(C.c_uint8)), names)
    if rc != GP_OK:
        raise GpError(rc, L.gp_last_error(None).decode())
    out = names.raw[:nl.value].split(b"\0")[:-1] if nl.value else []
    return {"vertices": v, "conductor": c, "n_vertices": nv, "names": [x.decode() for x in out]}


def m2l_operator(order, off, compressed=True):
    """One M2L interaction matrix as the FMM operator applies it
    (gp_m2l_operator, host only): (G (n^3 x n^3), rank)."""
    L = load()
    n3 = int(order) ** 3
    G = np.zeros((n3, n3))
    rk = C.c_uint32(0)
    rc = L.gp_m2l_operator(int(order), int(off[0]), int(off[1]), int(off[2]), int(bool(compressed)),
                           _dptr(G), C.byref(rk))
    if rc != GP_OK:
        raise GpError(rc, L.gp_last_error(None).decode())
    return G, int(rk.value)


def m2l_factors(order, off):
    """The factors of the truncated M2L block (gp_m2l_factors, host only):
    (U (n^3 x rank), Vt (rank x n^3), rank)."""
    L = load()
    n3 = int(order) ** 3
    U = np.zeros((n3, n3))
    Vt = np.zeros((n3, n3))
    rk = C.c_uint32(0)
    rc = L.gp_m2l_factors(int(order), int(off[0]), int(off[1]), int(off[2]), _dptr(U), _dptr(Vt),
                          C.byref(rk))
    if rc != GP_OK:
        raise GpError(rc, L.gp_last_error(None).decode())
    k = int(rk.value)
    return U[:, :k].copy(), Vt[:k].copy(), k


class GpResPolygon(C.Structure):
    _fields_ = [("net", C.c_uint32), ("n_vertices", C.c_uint32), ("xy", C.POINTER(C.c_int32)),
                ("has_pex", C.c_uint32), ("volume_conductor", C.c_uint32),
                ("sheet_res_ohm_sq", C.c_double), ("via_res_ohm", C.c_double),
                ("thickness_nm", C.c_double), ("height_nm", C.c_double)]


def solve_resistance(polys, dbu_nm=1.0):
    """r_ohm per net of extract_quasistatic's resistance half (host C++,
    gp_solve_resistance).  polys: dicts with net, xy ((n,2) DBU ring), and
    optional has_pex (1), volume_conductor (1), sheet_res (0), via_res (0),
    thickness_nm, height_nm.  Returns {net: r_ohm}; raises GpError."""
    L = load()
    keep, arr = [], (GpResPolygon * max(1, len(polys)))()
    for i, q in enumerate(polys):
        xy = np.ascontiguousarray(np.asarray(q["xy"], dtype=np.int32).reshape(-1))
        keep.append(xy)
        a = arr[i]
        a.net = int(q["net"])
        a.n_vertices = len(xy) // 2
        a.xy = xy.ctypes.data_as(C.POINTER(C.c_int32))
        a.has_pex = int(q.get("has_pex", 1))
        a.volume_conductor = int(q.get("volume_conductor", 1))
        a.sheet_res_ohm_sq = float(q.get("sheet_res", 0.0))
        a.via_res_ohm = float(q.get("via_res", 0.0))
        a.thickness_nm = float(q.get("thickness_nm", 0.0))
        a.height_nm = float(q.get("height_nm", 0.0))
    n = C.c_uint32(0)
    rc = L.gp_solve_resistance(arr, len(polys), dbu_nm, C.byref(n), None, None)
    if rc != GP_OK:
        raise GpError(rc, L.gp_last_error(None).decode())
    nets = np.zeros(n.value, dtype=np.uint32)
    r = np.zeros(n.value)
    rc = L.gp_solve_resistance(arr, len(polys), dbu_nm, C.byref(n), _uptr(nets), _dptr(r))
    if rc != GP_OK:
        raise GpError(rc, L.gp_last_error(None).decode())
    del keep
    return {int(k): float(v) for k, v in zip(nets, r)}
