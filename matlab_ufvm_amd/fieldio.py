"""OpenFOAM ASCII field files: the reference's time-directory output and
restart input (host side).

`format_field` / `write_field` produce the layout of
``src/utilities/postProcess/cfdWriteOpenFoamField.m:11-160`` (header block,
``dimensions``, ``internalField nonuniform List<scalar|vector>`` with one entry
per line -- cells for vol fields, interior faces for surface fields -- then a
``boundaryField`` entry per patch with its type and, where the patch carries
one, its value as ``uniform`` or ``nonuniform List``).  Numbers are written
either as MATLAB's ``num2str`` writes them (``exact=False``: the reference's
own files, about five significant digits) or with 17 significant digits
(``exact=True``: lossless, so a run restarts from its files bit for bit).

`read_field` parses such files (the reference's readers,
``src/utilities/IO/cfdReadTimeDirectory.m:13-133`` and ``readNonuniformList.m``,
read the same structure).  `write_time` / `read_time` write and read a SIMPLE
case's U, p and mdot_f as a time directory (``cfdWriteResults.m:21-35``).
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass, field

import numpy as np

HEADER = r"""/*--------------------------------*- C++ -*----------------------------------*\\
| =========                 |                                                 |
| \\      /  F ield         | OpenFOAM: The Open Source CFD Toolbox           |
|  \\    /   O peration     | Version:  v1706                                 |
|   \\  /    A nd           | Web:      www.OpenFOAM.com                      |
|    \\/     M anipulation  |                                                 |
\\*---------------------------------------------------------------------------*/"""

SEPARATOR = "// * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * //"
FOOTER = "// ************************************************************************* //"

DIMENSIONS = {"U": "[0 1 -1 0 0 0 0]", "p": "[0 2 -2 0 0 0 0]", "mdot_f": "[1 0 -1 0 0 0 0]"}


def num2str(x: float) -> str:
    """A real scalar as MATLAB's num2str writes it: integers in full, other
    values with %.Ng, N = max(ceil(log10(|x|)), 1) + 4 significant digits."""
    x = float(x)
    if x == 0.0:
        return "0"
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    if x == round(x) and abs(x) < 1e15:
        return "%d" % int(x)
    digits = max(math.ceil(math.log10(abs(x))), 1) + 4
    return "%.*g" % (digits, x)


def _fmt(x: float, exact: bool) -> str:
    return repr(float(x)) if exact else num2str(x)


def _fmt_entry(v, exact: bool) -> str:
    v = np.asarray(v, dtype=np.float64)
    if v.ndim == 0:
        return _fmt(v, exact)
    return "(" + " ".join(_fmt(c, exact) for c in v) + ")"


@dataclass
class BoundaryEntry:
    type: str
    kind: str | None = None          # None (no value), "uniform" or "nonuniform"
    value: np.ndarray | None = None  # uniform: one entry; nonuniform: [nfaces] or [nfaces, 3]


@dataclass
class FoamField:
    name: str
    cls: str                                  # volScalarField, volVectorField, surfaceScalarField
    location: str
    dimensions: str
    internal: np.ndarray
    boundary: dict = field(default_factory=dict)   # patch name -> BoundaryEntry (file order)


def _list_lines(values, exact: bool, kind: str):
    values = np.asarray(values, dtype=np.float64)
    out = [f"nonuniform List<{kind}>", str(values.shape[0]), "("]
    out += [_fmt_entry(v, exact) for v in values]
    out += [")", ";"]
    return out


def format_field(f: FoamField, exact: bool = False) -> str:
    """The text of an OpenFOAM field file in the reference's layout."""
    kind = "vector" if f.cls == "volVectorField" else "scalar"
    lines = [HEADER, "FoamFile", "{", "    version     2.0;", "    format      ascii;",
             f"    class       {f.cls};", f'    location    "{f.location}";',
             f"    object      {f.name};", "}", SEPARATOR, "",
             f"dimensions      {f.dimensions}", "", ""]
    body = _list_lines(f.internal, exact, kind)
    lines.append("internalField   " + body[0])
    lines += body[1:]
    lines += ["", "boundaryField", "{"]
    for pname, be in f.boundary.items():
        lines += [f"    {pname}", "    {", f"        type             {be.type};"]
        if be.kind == "uniform":
            lines.append(f"        value            uniform {_fmt_entry(be.value, exact)};")
        elif be.kind == "nonuniform":
            vb = _list_lines(be.value, exact, kind)
            lines.append("        value            " + vb[0])
            lines += vb[1:]
        lines.append("    }")
    lines += ["}", "", FOOTER]
    return "\n".join(lines) + "\n"


def write_field(path: str, f: FoamField, exact: bool = False) -> None:
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    with open(path, "w") as fh:
        fh.write(format_field(f, exact))


_NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?|[-+]?Inf|NaN"


def _parse_values(text: str, vector: bool):
    if vector:
        rows = re.findall(r"\(([^()]*)\)", text)
        return np.array([[float(t) for t in r.split()] for r in rows], dtype=np.float64)
    return np.array([float(t) for t in re.findall(_NUM, text)], dtype=np.float64)


def _parse_value(spec: str, vector: bool):
    """'uniform X' / 'nonuniform List<...> n ( ... )' -> (kind, array)."""
    spec = spec.strip()
    if spec.startswith("uniform"):
        v = _parse_values(spec[len("uniform"):], vector)
        return "uniform", (v[0] if vector else v.reshape(()) if v.size == 1 else v)
    m = re.match(r"nonuniform\s+List<\w+>\s*(\d+)\s*\((.*)\)\s*$", spec, re.S)
    if not m:
        raise ValueError(f"cannot parse field value: {spec[:60]!r}")
    n = int(m.group(1))
    vals = _parse_values(m.group(2), vector)
    if vals.shape[0] != n:
        raise ValueError(f"list declares {n} entries, holds {vals.shape[0]}")
    return "nonuniform", vals


def read_field(path: str) -> FoamField:
    text = open(path).read()
    head = re.search(r"FoamFile\s*\{(.*?)\}", text, re.S).group(1)
    def get(key, default=None):
        m = re.search(key + r"\s+\"?([^;\"]*)\"?;", head)
        return m.group(1).strip() if m else default

    # a 0 directory's files may omit "location"
    cls, loc, name = get("class"), get("location", ""), get("object")
    vector = cls == "volVectorField"
    dims = re.search(r"dimensions\s+(\[[^\]]*\])", text).group(1)
    istart = text.index("internalField")
    bstart = text.index("boundaryField", istart)
    ispec = text[istart + len("internalField"):bstart].strip()
    ispec = ispec[:ispec.rindex(";")]
    _, internal = _parse_value(ispec, vector)
    boundary = {}
    body = text[bstart + len("boundaryField"):]
    body = body[body.index("{") + 1:]
    pos = 0
    while True:
        m = re.compile(r"\s*(\w+)\s*\{", re.S).match(body, pos)
        if not m:
            break
        # the patch block ends at the first "}" at its own depth
        depth, j = 1, m.end()
        while depth:
            if body[j] == "{":
                depth += 1
            elif body[j] == "}":
                depth -= 1
            j += 1
        block = body[m.end():j - 1]
        btype = re.search(r"type\s+(\w+)\s*;", block).group(1)
        vm = re.search(r"value\s+(.*);\s*$", block, re.S)
        if vm:
            kind, val = _parse_value(vm.group(1), vector)
            boundary[m.group(1)] = BoundaryEntry(btype, kind, val)
        else:
            boundary[m.group(1)] = BoundaryEntry(btype)
        pos = j
    return FoamField(name, cls, loc, dims, internal, boundary)


# ------------------------------------------------------------------- a case

def _entry(values):
    """uniform when every boundary value is the same, else nonuniform"""
    v = np.asarray(values, dtype=np.float64)
    if v.shape[0] and np.all(v == v[0]):
        return "uniform", v[0].copy()
    return "nonuniform", v.copy()


def case_fields(case, time: str = "0"):
    """U, p and mdot_f of a SIMPLE case as FoamFields: boundary types from the
    case's boundary conditions (U fixedValue with its value, zeroGradient,
    empty; mdot_f 'calculated' with its face values)."""
    m = case.mesh
    N, Fi = m.n_cells, m.n_interior
    U, p, mdot = case.field("U"), case.field("p"), case.field("mdot_f")
    fU, fp = {}, {}
    fm = {}
    for pt in m.patches:
        ubc, pbc = case.bcs[pt.name][0], case.bcs[pt.name][2]
        bel = slice(N + pt.start - Fi, N + pt.start - Fi + pt.nfaces)
        if ubc == "fixedValue":
            fU[pt.name] = BoundaryEntry("fixedValue", *_entry(U[bel]))
        else:
            fU[pt.name] = BoundaryEntry(ubc)
        fp[pt.name] = BoundaryEntry(pbc)
        # the reference writes mdot_f's patches as 'calculated' with their
        # face values, empty ones included
        fm[pt.name] = BoundaryEntry("calculated",
                                    *_entry(mdot[pt.start:pt.start + pt.nfaces]))
    return [FoamField("U", "volVectorField", time, DIMENSIONS["U"], U[:N].copy(), fU),
            FoamField("p", "volScalarField", time, DIMENSIONS["p"], p[:N].copy(), fp),
            FoamField("mdot_f", "surfaceScalarField", time, DIMENSIONS["mdot_f"],
                      mdot[:Fi].copy(), fm)]


def write_time(case, case_dir: str, time, exact: bool = True) -> str:
    """Writes U, p and mdot_f of `case` to case_dir/<time>/ (cfdWriteResults);
    exact=True keeps every bit, so read_time restarts the run exactly."""
    t = num2str(time) if not isinstance(time, str) else time
    d = os.path.join(case_dir, t)
    for f in case_fields(case, t):
        write_field(os.path.join(d, f.name), f, exact)
    return d


def read_time(case, case_dir: str, time) -> None:
    """Loads U, p and mdot_f of case_dir/<time>/ into `case` (restart).  Cell
    and face values come from the files; the boundary values of patches
    without a value entry (zeroGradient, empty) take the owner cell's value."""
    t = num2str(time) if not isinstance(time, str) else time
    m = case.mesh
    N, Fi, Fb = m.n_cells, m.n_interior, m.n_bfaces
    owner_b = np.asarray(m.owner[Fi:], dtype=np.int64)
    out = {}
    for name in ("U", "p", "mdot_f"):
        f = read_field(os.path.join(case_dir, t, name))
        if name == "mdot_f":
            full = np.zeros(m.n_faces)
            full[:Fi] = f.internal
        else:
            shape = (N + Fb, 3) if name == "U" else (N + Fb,)
            full = np.zeros(shape)
            full[:N] = f.internal
            full[N:] = f.internal[owner_b]
        for pt in m.patches:
            be = f.boundary.get(pt.name)
            if be is None or be.kind is None:
                continue
            sl = (slice(pt.start, pt.start + pt.nfaces) if name == "mdot_f"
                  else slice(N + pt.start - Fi, N + pt.start - Fi + pt.nfaces))
            full[sl] = be.value
        out[name] = full
    for name, v in out.items():
        case.set_field(name, v)
