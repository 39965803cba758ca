"""Minimal MPS reader (fixed and free format) producing a :class:`cuopt_amd.generators.LP`.

Semantics follow the reference parser (cpp/libmps_parser/src/mps_parser.cpp):
the right-hand side given on the objective row is the negated objective offset
(mps_parser.cpp:1098), ``UP`` with a negative value and no other bound sets the
lower bound to -inf (mps_parser.cpp:1382-1390), ``MI``/``PL``/``FR``/``FX``/``BV``
as in mps_parser.cpp:1403-1430, and RANGES follow the usual convention
(E row: [b, b+|R|] for R > 0, [b-|R|, b] for R < 0; L row: [b-|R|, b];
G row: [b, b+|R|]).  Default variable bounds are [0, +inf).
"""

from __future__ import annotations

import numpy as np

from .generators import LP


class MPSError(ValueError):
    pass


def read_mps(path: str) -> LP:
    with open(path, "r") as f:
        lines = f.read().splitlines()
    section = None
    maximize = False
    obj_row = None
    row_names: list[str] = []
    row_type: dict[str, str] = {}
    col_index: dict[str, int] = {}
    cols: list[str] = []
    entries: list[tuple[int, int, float]] = []  # (row, col, val)
    cost: dict[int, float] = {}
    rhs: dict[str, float] = {}
    ranges: dict[str, float] = {}
    lb: dict[int, float] = {}
    ub: dict[int, float] = {}
    bounded: set[int] = set()
    offset = 0.0
    row_id: dict[str, int] = {}

    def col_of(name: str) -> int:
        if name not in col_index:
            col_index[name] = len(cols)
            cols.append(name)
        return col_index[name]

    for raw in lines:
        if not raw.strip() or raw.lstrip().startswith("*"):
            continue
        if not raw[0].isspace():
            tok = raw.split()
            head = tok[0].upper()
            if head == "NAME":
                section = "NAME"
            elif head == "OBJSENSE":
                section = "OBJSENSE"
                if len(tok) > 1:
                    maximize = tok[1].upper().startswith("MAX")
            elif head in ("ROWS", "COLUMNS", "RHS", "RANGES", "BOUNDS"):
                section = head
            elif head == "ENDATA":
                break
            elif section == "OBJSENSE":
                maximize = head.startswith("MAX")
            else:
                raise MPSError(f"unknown section line: {raw!r}")
            continue
        tok = raw.split()
        if section == "OBJSENSE":
            maximize = tok[0].upper().startswith("MAX")
        elif section == "ROWS":
            t, name = tok[0].upper(), tok[1]
            if t == "N":
                if obj_row is None:
                    obj_row = name
                continue
            if t not in ("E", "L", "G"):
                raise MPSError(f"bad row type {t}")
            row_type[name] = t
            row_id[name] = len(row_names)
            row_names.append(name)
        elif section == "COLUMNS":
            if len(tok) >= 3 and tok[1].strip("'").upper() == "MARKER":
                continue
            j = col_of(tok[0])
            pairs = tok[1:]
            if len(pairs) % 2:
                raise MPSError(f"bad COLUMNS line: {raw!r}")
            for r, v in zip(pairs[0::2], pairs[1::2]):
                val = float(v)
                if r == obj_row:
                    cost[j] = cost.get(j, 0.0) + val
                elif r in row_id:
                    entries.append((row_id[r], j, val))
                else:
                    raise MPSError(f"unknown row {r}")
        elif section in ("RHS", "RANGES"):
            pairs = tok[1:] if len(tok) % 2 == 1 else tok
            for r, v in zip(pairs[0::2], pairs[1::2]):
                val = float(v)
                if section == "RHS":
                    if r == obj_row:
                        offset = -val
                    elif r in row_id:
                        rhs[r] = val
                    else:
                        raise MPSError(f"unknown row {r}")
                else:
                    if r not in row_id:
                        raise MPSError(f"unknown row {r}")
                    ranges[r] = val
        elif section == "BOUNDS":
            bt = tok[0].upper()
            if bt in ("FR", "MI", "PL", "BV"):
                name = tok[2] if len(tok) >= 3 else tok[1]
                val = None
            else:
                if len(tok) < 3:
                    raise MPSError(f"bad BOUNDS line: {raw!r}")
                name, val = (tok[2], float(tok[3])) if len(tok) >= 4 else (tok[1], float(tok[2]))
            j = col_of(name)
            if bt == "LO" or bt == "LI":
                lb[j] = val
            elif bt == "UP" or bt == "UI":
                ub[j] = val
                if j not in bounded and val < 0:
                    lb[j] = -np.inf
            elif bt == "FX":
                lb[j] = val
                ub[j] = val
            elif bt == "FR":
                lb[j] = -np.inf
                ub[j] = np.inf
            elif bt == "MI":
                lb[j] = -np.inf
            elif bt == "PL":
                ub[j] = np.inf
            elif bt == "BV":
                lb[j] = 0.0
                ub[j] = 1.0
            else:
                raise MPSError(f"unsupported bound type {bt}")
            bounded.add(j)
    m, n = len(row_names), len(cols)
    c = np.zeros(n)
    for j, v in cost.items():
        c[j] = v
    var_lb = np.zeros(n)
    var_ub = np.full(n, np.inf)
    for j, v in lb.items():
        var_lb[j] = v
    for j, v in ub.items():
        var_ub[j] = v
    con_lb = np.empty(m)
    con_ub = np.empty(m)
    for i, name in enumerate(row_names):
        b = rhs.get(name, 0.0)
        t = row_type[name]
        r = ranges.get(name)
        if t == "E":
            lo, hi = b, b
            if r is not None:
                lo, hi = (b, b + abs(r)) if r > 0 else (b - abs(r), b)
        elif t == "L":
            lo, hi = -np.inf, b
            if r is not None:
                lo = b - abs(r)
        else:
            lo, hi = b, np.inf
            if r is not None:
                hi = b + abs(r)
        con_lb[i], con_ub[i] = lo, hi
    if entries:
        e = np.array(entries, dtype=np.float64)
        rows, cc, vals = e[:, 0].astype(np.int64), e[:, 1].astype(np.int64), e[:, 2]
    else:
        rows = cc = np.zeros(0, dtype=np.int64)
        vals = np.zeros(0)
    key = rows * max(n, 1) + cc
    order = np.argsort(key, kind="stable")
    rows, cc, vals = rows[order], cc[order], vals[order]
    ptr = np.zeros(m + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=m), out=ptr[1:])
    return LP(
        m=m, n=n, row_offsets=ptr.astype(np.int32), col_indices=cc.astype(np.int32),
        values=vals.astype(np.float64), c=c, con_lb=con_lb, con_ub=con_ub,
        var_lb=var_lb, var_ub=var_ub, maximize=maximize, offset=offset, name=path.rsplit("/", 1)[-1],
    )
