"""Pure-Python restatement of the reference's FastCap quickif (.qui) reader.

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker of the
product parser (``gp_qui_parse`` in libgpurify), never shipped or measured.
It follows ``cap::geometry::parse`` of crates/pex/src/quasistatic/cap/
geometry.rs:56-135 line by line; the numeric token grammars are those of
Rust's ``str::parse::<i64>`` and ``str::parse::<f64>`` (decimal only, no
underscores or hexadecimal, ``inf``/``infinity``/``nan`` in any case).
"""
from __future__ import annotations

import re

_I64 = re.compile(r"[+-]?[0-9]+\Z")
_F64 = re.compile(r"[+-]?([0-9]+\.?[0-9]*|\.[0-9]+)([eE][+-]?[0-9]+)?\Z")
_SPECIAL = {"inf", "infinity", "nan"}


class QuiError(ValueError):
    """ParseError::At(line, message) (geometry.rs:35-39)."""

    def __init__(self, line, msg):
        super().__init__(f"line {line}: {msg}")
        self.line, self.msg = line, msg


def _i64(tok):
    if not _I64.match(tok):
        return None
    v = int(tok)
    return v if -(2 ** 63) <= v < 2 ** 63 else None


def _f64(tok):
    body = tok[1:] if tok[:1] in "+-" and tok else tok
    if body.lower() in _SPECIAL:
        return float(tok)
    if not _F64.match(tok):
        return None
    return float(tok)


def _floats(tokens, n, lineno):
    """parse_floats, geometry.rs:41-52."""
    if len(tokens) < n:
        raise QuiError(lineno, f"expected {n} numbers, got {len(tokens)}")
    out = []
    for t in tokens[:n]:
        v = _f64(t)
        if v is None:
            raise QuiError(lineno, f"bad number '{t}'")
        out.append(v)
    return out


def parse(text: str, scale: float = 1.0):
    """geometry.rs:56-135: (panels, names); a panel is (vertex list, index)."""
    panels, names = [], []
    idx_of, names_by_number = {}, {}

    def intern(num):  # geometry.rs:62-77
        if num in idx_of:
            return idx_of[num]
        i = len(names)
        idx_of[num] = i
        names.append(names_by_number.get(num, f"cond{num}"))
        return i

    lines = text.split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    for lineno0, raw in enumerate(lines):
        lineno = lineno0 + 1
        if raw.endswith("\r"):
            raw = raw[:-1]
        line = raw.strip()
        if not line:
            continue
        first = line[0]
        tokens = line.split()
        if first in "0*%":
            continue
        if first in "QqTt":
            quad = first in "Qq"
            c = _i64(tokens[1]) if len(tokens) > 1 else None
            if c is None:
                raise QuiError(lineno, ("Q" if quad else "T") + ": missing conductor number")
            f = _floats(tokens[2:], 12 if quad else 9, lineno)
            verts = [[x * scale for x in f[3 * k:3 * k + 3]] for k in range(4 if quad else 3)]
            panels.append((verts, intern(c)))
        elif first in "Nn":
            if len(tokens) > 2:
                c = _i64(tokens[1])
                if c is not None:
                    names_by_number[c] = tokens[2]
                    if c in idx_of:
                        names[idx_of[c]] = tokens[2]
        elif first in "Cc":
            raise QuiError(lineno, "C (sub-file include) not yet supported")
        else:
            raise QuiError(lineno, f"unrecognized line: {line}")
    return panels, names
