"""FastCap quickif (.qui) input: the product parser (gp_qui_parse, host C++)
against the oracle restatement of cap::geometry::parse (crates/pex/src/
quasistatic/cap/geometry.rs:56-135), the oracle pinned by known answers
written out by hand from that grammar, and the writer's round trip.

The reference's own tests hold no .qui file or parser test, so the known
answers below are the pin: each case names the geometry.rs lines it
exercises.  Host-only (no device): the parser needs the library, not a GPU.
"""
import math

import numpy as np
import pytest

import gpurify_amd as G
from gpurify_amd import workloads as W
from gpurify_amd.fastcap import write_qui
from oracle import qui as Q

SQUARE = "0 0 0  1 0 0  1 1 0  0 1 0"
TRI = "0 0 1  1 0 1  0 1 1"


def product(text, scale=1.0):
    try:
        return G.parse_qui(text, scale)
    except G.GpError as e:
        return str(e)


def oracle(text, scale=1.0):
    try:
        return Q.parse(text, scale)
    except Q.QuiError as e:
        return e


def same(text, scale=1.0):
    """Product == oracle: panels bit for bit (a triangle repeats v3), indices,
    names, or the same error line and message."""
    p, o = product(text, scale), oracle(text, scale)
    if isinstance(o, Q.QuiError):
        assert isinstance(p, str) and p.endswith(str(o)), (text, p, o)
        return o
    assert not isinstance(p, str), (text, p)
    panels, names = o
    assert p["names"] == names
    assert len(p["conductor"]) == len(panels)
    for k, (verts, ci) in enumerate(panels):
        assert p["conductor"][k] == ci and p["n_vertices"][k] == len(verts)
        full = verts + [verts[2]] * (4 - len(verts))
        want = np.array(full)
        got = p["vertices"][k]
        assert np.array_equal(np.isnan(got), np.isnan(want))
        assert np.array_equal(got[~np.isnan(got)], want[~np.isnan(want)])
    return o


# ---------------------------------------------------------------- known answers
def test_known_answer_interning_and_names():
    # geometry.rs:62-77 (first-seen interning, "cond<c>"), 120-129 (N before
    # and after interning), 88 (comment kinds), 105-119 (triangles)
    text = ("0 title line\n* comment\n% another\n"
            f"N 7 gnd\nQ 7 {SQUARE}\nT -3 {TRI}\nQ 7 {SQUARE}\n"
            f"Q 9223372036854775807 {SQUARE}\nN -3 top\n")
    panels, names = Q.parse(text)
    assert names == ["gnd", "top", "cond9223372036854775807"]
    assert [ci for _, ci in panels] == [0, 1, 0, 2]
    assert panels[1][0] == [[0.0, 0.0, 1.0], [1.0, 0.0, 1.0], [0.0, 1.0, 1.0]]
    same(text)


def test_known_answer_line_handling():
    # geometry.rs:79-88: lines() (CRLF), trim, the FIRST CHARACTER picks the
    # kind ("Qx" is a Q line, "q"/"t"/"n" lower case too), extra tokens ignored
    text = f"  \r\n\tqx 2 {SQUARE} 99 junk\r\nt 2 {TRI}\nn 2 net_b extra\n"
    panels, names = Q.parse(text, 2.0)
    assert names == ["net_b"] and len(panels) == 2
    assert panels[0][0][2] == [2.0, 2.0, 0.0]  # scale multiplies every coordinate
    same(text, 2.0)


@pytest.mark.parametrize("text,line,msg", [
    (f"Q 1.0 {SQUARE}\n", 1, "Q: missing conductor number"),            # i64 parse
    (f"\nT\n", 2, "T: missing conductor number"),
    (f"Q 99999999999999999999 {SQUARE}\n", 1, "Q: missing conductor number"),  # i64 overflow
    ("Q 1 0 0 0\n", 1, "expected 12 numbers, got 3"),                      # geometry.rs:42-43
    (f"T 1 0 0 1 1 0 1 0 1\n", 1, "expected 9 numbers, got 8"),
    (f"Q 1 0x10 {SQUARE}\n", 1, "bad number '0x10'"),                      # no hex floats
    (f"Q 1 1_0 {SQUARE}\n", 1, "bad number '1_0'"),                        # no underscores
    (f"Q 1 1e {SQUARE}\n", 1, "bad number '1e'"),
    (f"Q 1 . {SQUARE}\n", 1, "bad number '.'"),
    ("0 ok\nC sub.qui 1 0 0 0\n", 2, "C (sub-file include) not yet supported"),
    ("0 ok\n\n  X 1 2 3  \n", 3, "unrecognized line: X 1 2 3"),
])
def test_known_answer_errors(text, line, msg):
    e = same(text)
    assert isinstance(e, Q.QuiError) and (e.line, e.msg) == (line, msg)


def test_known_answer_number_forms():
    # Rust f64 grammar: sign, "1.", ".5", exponents, inf / infinity / nan
    text = "Q 1 +1. .5 -2e-3  1E+2 inf -Infinity  NaN 0 0  1e999 -0.0 7\n"
    (verts, _), = Q.parse(text)[0]
    flat = [x for v in verts for x in v]
    assert flat[:4] == [1.0, 0.5, -0.002, 100.0]
    assert flat[4] == math.inf and flat[5] == -math.inf and math.isnan(flat[6])
    assert flat[9] == math.inf and math.copysign(1.0, flat[10]) == -1.0
    same(text)


# ---------------------------------------------------------------- product vs oracle
def _fuzz_line(rng):
    kinds = ["Q", "T", "N", "0", "*", "%", "C", "q", "t", "n", "X", ""]
    nums = ["0", "1", "-2.5", "1e-6", ".5", "3.", "inf", "nan", "1e", "0x1", "+7", "-0", "1_0",
            "12345678901234567890", "2e308", "abc"]
    k = kinds[rng.integers(len(kinds))]
    if k == "":
        return " " * int(rng.integers(3))
    toks = [k + ("x" if rng.random() < 0.1 else "")]
    toks += [str(int(rng.integers(-3, 5))) if rng.random() < 0.8 else nums[rng.integers(len(nums))]]
    want = {"Q": 12, "q": 12, "T": 9, "t": 9}.get(k, 1)
    m = want + int(rng.integers(-1, 3)) if rng.random() < 0.3 else want
    for _ in range(max(0, m)):
        toks.append(nums[rng.integers(len(nums))] if rng.random() < 0.05
                    else repr(float(rng.normal())))
    return " ".join(toks)


def test_product_matches_oracle_on_seeded_files():
    rng = np.random.default_rng(2024)
    errors = 0
    for _ in range(400):
        lines = [_fuzz_line(rng) for _ in range(int(rng.integers(1, 12)))]
        sep = "\r\n" if rng.random() < 0.2 else "\n"
        if isinstance(same(sep.join(lines) + (sep if rng.random() < 0.5 else "")), Q.QuiError):
            errors += 1
    assert 50 < errors < 390  # both outcomes exercised


# ---------------------------------------------------------------- writer round trip
def test_round_trip_of_the_mesher_panels_is_bit_exact():
    v, c, _ = G.mesh_groups(W.two_wires(), refine=2)
    names = ["net0", "net1", "substrate"]
    text = write_qui(v, c, names=names)
    g = G.parse_qui(text)
    assert np.array_equal(g["vertices"], v) and np.array_equal(g["conductor"], c)
    assert g["names"] == names and set(g["n_vertices"].tolist()) == {4}
    panels, onames = Q.parse(text)
    assert onames == names and len(panels) == len(v)


def test_triangles_round_trip():
    rng = np.random.default_rng(5)
    v = rng.normal(size=(20, 4, 3))
    v[::3, 3] = v[::3, 2]  # every third panel a triangle
    c = np.repeat(np.arange(4), 5).astype(np.uint32)
    g = G.parse_qui(write_qui(v, c))
    assert np.array_equal(g["vertices"], v) and np.array_equal(g["conductor"], c)
    assert np.array_equal(g["n_vertices"] == 3, np.arange(20) % 3 == 0)
    assert g["names"] == [f"cond{i}" for i in range(1, 5)]


# ---------------------------------------------------------------- list files
def _coated_sphere_files(tmp_path, tr=(0.0, 0.0, 0.0)):
    from oracle import oracle as O
    (tmp_path / "sphere.qui").write_text(write_qui(O.mesh_icosphere((0, 0, 0), 1.0, 2),
                                                   np.zeros(320, np.uint32)))
    (tmp_path / "shell.qui").write_text(write_qui(O.mesh_icosphere((0, 0, 0), 2.0, 2),
                                                  np.zeros(320, np.uint32)))
    t = " ".join(repr(x) for x in tr)
    lst = tmp_path / "coated.lst"
    lst.write_text(f"* coated sphere (crates/pex/tests/dielectric.rs:8-45)\n"
                   f"C sphere.qui 4.0 {t}\n\n% shell: vacuum outside, eps 4 inside\n"
                   f"D shell.qui 1.0 4.0 0 0 0 {t} +\n")
    return str(lst)


def test_list_file_roles_and_translation(tmp_path):
    # listfile.rs:56-109: C panels are conductors of eps outerperm, numbered
    # "<file>:<local>" per C line; D panels interfaces with the reference
    # point translated with them; a trailing +/- flag is ignored
    from gpurify_amd.fastcap import parse_lst
    g = parse_lst(_coated_sphere_files(tmp_path, tr=(1.0, -2.0, 0.5)))
    assert g["names"] == ["sphere.qui:0"]
    assert (g["role"] == 0).sum() == 320 and (g["role"] == 1).sum() == 320
    assert np.all(g["eps"][g["role"] == 0] == 4.0)
    assert np.all(g["eps_out"][g["role"] == 1] == 1.0) and np.all(g["eps_in"][g["role"] == 1] == 4.0)
    assert np.allclose(g["ref"][g["role"] == 1], [1.0, -2.0, 0.5])
    assert np.allclose(g["vertices"][g["role"] == 0].mean(axis=(0, 1)), [1.0, -2.0, 0.5], atol=1e-12)


def test_list_file_coated_sphere_solves_to_the_closed_form(tmp_path):
    from gpurify_amd.fastcap import parse_lst
    from oracle import oracle as O
    g = parse_lst(_coated_sphere_files(tmp_path))
    C = O.solve_dielectric(g["vertices"], g["role"], g["conductor"], g["eps"], g["eps_out"],
                           g["eps_in"], g["ref"], 1)
    eps0 = 8.8541878128e-12
    exact = 4 * np.pi * eps0 / ((1 / 4.0) * (1 / 1.0 - 1 / 2.0) + 1 / 2.0)
    assert abs(C[0, 0] - exact) / exact < 0.05


@pytest.mark.parametrize("body,msg", [
    ("X a.qui 1\n", "line 1: unrecognized list line: X a.qui 1"),
    ("C\n", "line 1: C: missing file"),
    ("C sphere.qui 4.0 0 0\n", "line 1: expected number for ztrans"),
    ("D shell.qui 1.0 0x4 0 0 0 0 0 0\n", "line 1: expected number for inner permittivity"),
    ("C missing.qui 1 0 0 0\n", "cannot read"),
])
def test_list_file_errors(tmp_path, body, msg):
    from gpurify_amd.fastcap import ListFileError, parse_lst
    _coated_sphere_files(tmp_path)
    p = tmp_path / "bad.lst"
    p.write_text(body)
    with pytest.raises(ListFileError, match=msg):
        parse_lst(str(p))
