"""FastCap quickif (.qui) panel files: input through the product parser, output
for regression fixtures.

``parse_qui`` is the library's reader (``gp_qui_parse``, the grammar of
cap::geometry::parse, crates/pex/src/quasistatic/cap/geometry.rs:56-135).
``write_qui`` writes a panel set in the same grammar with shortest
round-trip decimal coordinates, so ``parse_qui(write_qui(...))`` gives the
same panels bit for bit.  The reference has no writer; this one emits only
line kinds its reader accepts (0 title, N names, Q and T panels).
"""
from __future__ import annotations

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
