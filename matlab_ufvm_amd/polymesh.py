"""OpenFOAM ASCII polyMesh reader (host side).

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
