"""OpenFOAM ASCII polyMesh reader and writer (host side).

Reads ``constant/polyMesh/{points,faces,owner,neighbour,boundary}`` into 0-based
numpy arrays.  Behaviour follows ``src/utilities/IO/cfdReadPolyMesh.m:25-133``
(points as the node table, faces as node lists, ``owner``/``neighbour`` as the
face-to-cell maps, the number of cells taken as the largest owner/neighbour
index, boundary patches as ``{name, type, startFace, nFaces}``), with indices
kept 0-based instead of the reference's +1 shift (``:62,79,91,124``).
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Patch:
    name: str
    type: str
    start: int          # 0-based first face index
    nfaces: int


@dataclass
class PolyMesh:
    points: np.ndarray                  # [nP, 3] float64
    face_ptr: np.ndarray                # [F+1] int64 offsets into face_nodes
    face_nodes: np.ndarray              # [sum] int32 node ids (0-based)
    owner: np.ndarray                   # [F] int32
    neighbour: np.ndarray               # [F_i] int32
    patches: list = field(default_factory=list)
    n_cells: int = 0

    @property
    def n_faces(self) -> int:
        return int(self.owner.shape[0])

    @property
    def n_interior_faces(self) -> int:
        return int(self.neighbour.shape[0])


def _strip_header(text: str) -> str:
    # drop C/C++ comments and the FoamFile{...} dictionary
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    m = re.search(r"FoamFile\s*\{", text)
    if m:
        depth, i = 1, m.end()
        while depth:
            if text[i] == "{":
                depth += 1
            elif text[i] == "}":
                depth -= 1
            i += 1
        text = text[i:]
    return text


def _list_body(text: str):
    text = _strip_header(text)
    m = re.search(r"(\d+)\s*\(", text)
    if not m:
        raise ValueError("no OpenFOAM list found")
    n = int(m.group(1))
    end = text.rfind(")")
    return n, text[m.end():end]


def read_points(path: str) -> np.ndarray:
    with open(path) as f:
        n, body = _list_body(f.read())
    vals = np.array(body.replace("(", " ").replace(")", " ").split(), dtype=np.float64)
    if vals.size != 3 * n:
        raise ValueError(f"points: expected {3*n} values, got {vals.size}")
    return vals.reshape(n, 3)


def read_faces(path: str):
    with open(path) as f:
        n, body = _list_body(f.read())
    toks = body.replace("(", " ").replace(")", " ").split()
    vals = np.array(toks, dtype=np.int64)
    ptr = np.zeros(n + 1, dtype=np.int64)
    nodes = []
    i = 0
    for k in range(n):
        cnt = int(vals[i])
        nodes.append(vals[i + 1:i + 1 + cnt])
        ptr[k + 1] = ptr[k] + cnt
        i += 1 + cnt
    if i != vals.size:
        raise ValueError("faces: trailing tokens")
    return ptr, np.concatenate(nodes).astype(np.int32)


def read_labels(path: str) -> np.ndarray:
    with open(path) as f:
        n, body = _list_body(f.read())
    vals = np.array(body.split(), dtype=np.int64)
    if vals.size != n:
        raise ValueError(f"labels: expected {n}, got {vals.size}")
    return vals.astype(np.int32)


def read_boundary(path: str) -> list:
    with open(path) as f:
        text = _strip_header(f.read())
    m = re.search(r"(\d+)\s*\(", text)
    body = text[m.end():text.rfind(")")]
    patches = []
    for pm in re.finditer(r"(\w+)\s*\{([^}]*)\}", body):
        name, d = pm.group(1), pm.group(2)
        typ = re.search(r"\btype\s+(\w+)\s*;", d).group(1)
        nf = int(re.search(r"\bnFaces\s+(\d+)\s*;", d).group(1))
        sf = int(re.search(r"\bstartFace\s+(\d+)\s*;", d).group(1))
        patches.append(Patch(name, typ, sf, nf))
    return patches


def read_polymesh(case_dir: str) -> PolyMesh:
    """Read ``<case_dir>/constant/polyMesh`` (cfdReadPolyMesh.m:18-24 file set)."""
    d = os.path.join(case_dir, "constant", "polyMesh")
    pts = read_points(os.path.join(d, "points"))
    ptr, nodes = read_faces(os.path.join(d, "faces"))
    own = read_labels(os.path.join(d, "owner"))
    nei = read_labels(os.path.join(d, "neighbour"))
    patches = read_boundary(os.path.join(d, "boundary"))
    n_cells = int(max(own.max(), nei.max() if nei.size else -1)) + 1
    return PolyMesh(pts, ptr, nodes, own, nei, patches, n_cells)


# ------------------------------------------------------------------- writer

_BANNER = """/*--------------------------------*- C++ -*----------------------------------*\\
  =========                 |
  \\\\      /  F ield         | OpenFOAM: The Open Source CFD Toolbox
   \\\\    /   O peration     |
    \\\\  /    A nd           |
     \\\\/     M anipulation  |
\\*---------------------------------------------------------------------------*/"""


def _foam_header(cls: str, obj: str, note: str = "") -> str:
    lines = [_BANNER, "FoamFile", "{", "    version     2.0;", "    format      ascii;",
             f"    class       {cls};"]
    if note:
        lines.append(f'    note        "{note}";')
    lines += ['    location    "constant/polyMesh";', f"    object      {obj};", "}",
              "// * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * //",
              ""]
    return "\n".join(lines) + "\n"


_FOOTER = "\n// ************************************************************************* //\n"


def write_polymesh(pm: PolyMesh, case_dir: str) -> None:
    """Write ``<case_dir>/constant/polyMesh/{points,faces,owner,neighbour,
    boundary}`` in OpenFOAM ASCII (the file set read_polymesh reads; points
    with 17 significant digits, so a write-read round trip is exact)."""
    d = os.path.join(case_dir, "constant", "polyMesh")
    os.makedirs(d, exist_ok=True)
    note = (f"nPoints:{pm.points.shape[0]}  nCells:{pm.n_cells}  nFaces:{pm.n_faces}  "
            f"nInternalFaces:{pm.n_interior_faces}")
    with open(os.path.join(d, "points"), "w") as f:
        f.write(_foam_header("vectorField", "points"))
        f.write(f"\n{pm.points.shape[0]}\n(\n")
        f.writelines("(%r %r %r)\n" % tuple(float(c) for c in p) for p in pm.points)
        f.write(")\n" + _FOOTER)
    with open(os.path.join(d, "faces"), "w") as f:
        f.write(_foam_header("faceList", "faces"))
        f.write(f"\n{pm.n_faces}\n(\n")
        for k in range(pm.n_faces):
            nodes = pm.face_nodes[pm.face_ptr[k]:pm.face_ptr[k + 1]]
            f.write(f"{nodes.size}(" + " ".join(str(int(v)) for v in nodes) + ")\n")
        f.write(")\n" + _FOOTER)
    for name, arr in (("owner", pm.owner), ("neighbour", pm.neighbour)):
        with open(os.path.join(d, name), "w") as f:
            f.write(_foam_header("labelList", name, note))
            f.write(f"\n{arr.shape[0]}\n(\n")
            f.writelines(f"{int(v)}\n" for v in arr)
            f.write(")\n" + _FOOTER)
    with open(os.path.join(d, "boundary"), "w") as f:
        f.write(_foam_header("polyBoundaryMesh", "boundary"))
        f.write(f"\n{len(pm.patches)}\n(\n")
        for p in pm.patches:
            f.write(f"    {p.name}\n    {{\n        type            {p.type};\n"
                    f"        nFaces          {p.nfaces};\n        startFace       {p.start};\n"
                    f"    }}\n")
        f.write(")\n" + _FOOTER)
