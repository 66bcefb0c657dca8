"""FastCap quickif (.qui) panel files: input through the product parser, output
for regression fixtures.

``parse_qui`` is the library's reader (``gp_qui_parse``, the grammar of
cap::geometry::parse, crates/pex/src/quasistatic/cap/geometry.rs:56-135).
``write_qui`` writes a panel set in the same grammar with shortest
round-trip decimal coordinates, so ``parse_qui(write_qui(...))`` gives the
same panels bit for bit.  The reference has no writer; this one emits only
line kinds its reader accepts (0 title, N names, Q and T panels).
``parse_lst`` reads a FastCap list file (cap/listfile.rs:35-115): C lines
(a .qui conductor file, outer permittivity, translation) and D lines (a .qui
dielectric-interface file, outer and inner permittivity, inner reference
point, translation), giving the panels and roles ``Context.
cap_solve_dielectric`` takes.
"""
from __future__ import annotations

import os
import re

import numpy as np

from ._lib import parse_qui  # noqa: F401


def write_qui(vertices, conductor, names=None, n_vertices=None, title="gpurify panels") -> str:
    """vertices (n,4,3) or (n,12), conductor (n,) 0-based indices; conductor i
    is written as FastCap number i + 1, named by names[i] when given.  A panel
    with n_vertices 3 (or, without n_vertices, whose last two vertices are
    equal) is written as a T line.  The reader numbers conductors by first
    appearance, so indices survive the round trip when each conductor's
    first panel comes after those of the lower indices (the mesher's order)."""
    v = np.asarray(vertices, dtype=np.float64).reshape(-1, 4, 3)
    c = np.asarray(conductor).reshape(-1)
    if len(c) != len(v):
        raise ValueError("vertices and conductor lengths differ")
    if "\n" in title or "\r" in title:
        raise ValueError("the title must be one line")
    out = [f"0 {title}"]
    if names is not None:
        for i, nm in enumerate(names):
            if not nm or any(ch.isspace() for ch in nm):
                raise ValueError(f"conductor name {nm!r} must be one non-empty word")
            out.append(f"N {i + 1} {nm}")
    for k in range(len(v)):
        tri = (n_vertices[k] == 3) if n_vertices is not None else bool(np.all(v[k, 2] == v[k, 3]))
        nv = 3 if tri else 4
        coords = " ".join(repr(float(x)) for x in v[k, :nv].reshape(-1))
        out.append(f"{'T' if tri else 'Q'} {int(c[k]) + 1} {coords}")
    return "\n".join(out) + "\n"


_F64 = re.compile(r"[+-]?([0-9]+\.?[0-9]*|\.[0-9]+)([eE][+-]?[0-9]+)?\Z")


class ListFileError(ValueError):
    """ListError (cap/listfile.rs:19-33)."""


def _getf(tok, lineno, what):
    """getf, listfile.rs:30-33: the Rust f64 grammar (no hex, no '_')."""
    if tok is not None:
        body = tok[1:] if tok[:1] in "+-" else tok
        if body.lower() in ("inf", "infinity", "nan") or _F64.match(tok):
            return float(tok)
    raise ListFileError(f"line {lineno}: expected number for {what}")


def parse_lst(path):
    """listfile.rs:35-115: dict with vertices (n,4,3), role (0 conductor, 1
    dielectric), conductor, eps (surrounding, conductors), eps_out, eps_in,
    ref ((n,3), dielectric panels), names (one per conductor, "<file>:<c>").
    Sub-files resolve against the list file's directory; each C line's
    conductors are new ones (numbered by first appearance in its file)."""
    try:
        with open(path) as f:
            text = f.read()
    except OSError as e:
        raise ListFileError(f"cannot read '{path}': {e}") from None
    base = os.path.dirname(path)
    verts, role, cond, eps, eo, ei, ref, names = [], [], [], [], [], [], [], []
    cond_map, file_counter = {}, 0

    def load(file, lineno):
        p = os.path.join(base, file)
        try:
            with open(p) as f:
                t = f.read()
        except OSError as e:
            raise ListFileError(f"cannot read '{p}': {e}") from None
        try:
            return parse_qui(t, 1.0)
        except Exception as e:  # GpError from the reader
            raise ListFileError(f"geometry parse error in '{file}': {e} (line {lineno} of list)")

    lines = text.split("\n")
    for lineno0, raw in enumerate(lines):
        lineno = lineno0 + 1
        line = raw[:-1].strip() if raw.endswith("\r") else raw.strip()
        if not line or line.startswith("*") or line.startswith("%"):
            continue
        tok = line.split()
        at = lambda i: tok[i] if i < len(tok) else None  # noqa: E731
        kind = tok[0]
        if kind in ("C", "c"):
            if len(tok) < 2:
                raise ListFileError(f"line {lineno}: C: missing file")
            outer = _getf(at(2), lineno, "outer permittivity")
            tr = [_getf(at(3 + k), lineno, ("xtrans", "ytrans", "ztrans")[k]) for k in range(3)]
            g = load(tok[1], lineno)
            file_counter += 1
            for k in range(len(g["conductor"])):
                local = int(g["conductor"][k])
                key = (file_counter, local)
                if key not in cond_map:
                    cond_map[key] = len(names)
                    names.append(f"{tok[1]}:{local}")
                verts.append(g["vertices"][k] + np.array(tr))
                role.append(0)
                cond.append(cond_map[key])
                eps.append(outer)
                eo.append(0.0)
                ei.append(0.0)
                ref.append((0.0, 0.0, 0.0))
        elif kind in ("D", "d"):
            if len(tok) < 2:
                raise ListFileError(f"line {lineno}: D: missing file")
            outer = _getf(at(2), lineno, "outer permittivity")
            inner = _getf(at(3), lineno, "inner permittivity")
            rp = [_getf(at(4 + k), lineno, ("xref", "yref", "zref")[k]) for k in range(3)]
            tr = [_getf(at(7 + k), lineno, ("xtrans", "ytrans", "ztrans")[k]) for k in range(3)]
            g = load(tok[1], lineno)
            for k in range(len(g["conductor"])):
                verts.append(g["vertices"][k] + np.array(tr))
                role.append(1)
                cond.append(0)
                eps.append(0.0)
                eo.append(outer)
                ei.append(inner)
                ref.append(tuple(rp[d] + tr[d] for d in range(3)))
        else:
            raise ListFileError(f"line {lineno}: unrecognized list line: {line}")
    return {"vertices": np.array(verts).reshape(-1, 4, 3), "role": np.array(role, dtype=np.int32),
            "conductor": np.array(cond, dtype=np.uint32), "eps": np.array(eps),
            "eps_out": np.array(eo), "eps_in": np.array(ei),
            "ref": np.array(ref, dtype=np.float64).reshape(-1, 3), "names": names}
