"""Finite-volume mesh geometry (host side, the input contract of the GPU path).

``process_mesh`` restates ``src/utilities/IO/cfdProcessOpenFoamMesh.m:15-241``
and the element lists of ``src/utilities/IO/cfdReadPolyMesh.m:139-178`` as
vectorised numpy over faces and cells, keeping the reference's per-face and
per-cell floating-point operation order so that the result reproduces the
reference's processed geometry, including:

* the face centre as the plain node mean, then the triangle fan about it
  (``cfdProcessOpenFoamMesh.m:23-52``);
* the element "centre" (pyramid apex) taken as the mean centroid of GLOBAL
  faces ``1..nFaces(e)`` rather than the element's own faces
  (``cfdProcessOpenFoamMesh.m:69-75``).  With planar faces the pyramid sums
  do not depend on the apex, but the faces of a jittered hex are warped, and
  then the centroid error grows with the apex's distance from the cell: on a
  128^3 cavity jittered by 15% of a cell, cells far from global face 1 get
  centroids displaced by several cells and some boundary faces get negative
  ``gDiff``.  ``apex="reference"`` (the default) keeps this for parity;
  ``apex="own"`` uses the mean centroid of the element's own faces, the
  centre the reference's comment describes;
* interior ``CN, eCN, gDiff, T, gf`` (``:108-132``) and boundary
  ``CN, gDiff = A^2/(CN.Sf), eCN, T, gf = 1, walldist`` (``:134-155``);
* ``cconn{e}`` = neighbours in increasing interior-face order (``:235-241``,
  built at ``cfdReadPolyMesh.m:152-163``).

All indices are 0-based.  The connectivity is flattened to CSR
(``cconn_ptr``/``cconn``) in exactly the reference's ``cconn`` order, which is
the order every bit-exact GPU kernel sums in.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .polymesh import PolyMesh


@dataclass
class Mesh:
    n_cells: int
    n_faces: int
    n_interior: int
    points: np.ndarray
    face_ptr: np.ndarray
    face_nodes: np.ndarray
    owner: np.ndarray               # [F] int32
    neighbour: np.ndarray           # [F_i] int32
    patches: list
    # face geometry
    face_centroid: np.ndarray = None   # [F,3]
    Sf: np.ndarray = None              # [F,3]
    area: np.ndarray = None            # [F]
    CN: np.ndarray = None              # [F,3]
    eCN: np.ndarray = None             # [F,3]
    gDiff: np.ndarray = None           # [F]
    T: np.ndarray = None               # [F,3]
    gf: np.ndarray = None              # [F]
    walldist: np.ndarray = None        # [F]
    # cell geometry
    volume: np.ndarray = None          # [N]
    centroid: np.ndarray = None        # [N,3]
    # element face lists in reference order (interior faces first, then boundary)
    cell_face_ptr: np.ndarray = None
    cell_faces: np.ndarray = None
    cell_face_sign: np.ndarray = None
    # cconn as CSR in reference order
    cconn_ptr: np.ndarray = None       # [N+1] int32
    cconn: np.ndarray = None           # [nnz] int32
    cconn_face: np.ndarray = None      # [nnz] int32 interior face of each entry
    own_nnz: np.ndarray = None         # [F_i] CSR slot of neighbour in owner's row
    nei_nnz: np.ndarray = None         # [F_i] CSR slot of owner in neighbour's row
    extra: dict = field(default_factory=dict)

    @property
    def n_bfaces(self) -> int:
        return self.n_faces - self.n_interior

    @property
    def csize(self) -> np.ndarray:
        return np.diff(self.cconn_ptr).astype(np.int32)

    def patch(self, name: str):
        for p in self.patches:
            if p.name == name:
                return p
        raise KeyError(name)


def _mag(v: np.ndarray) -> np.ndarray:
    # cfdMagnitude.m:3 -- sqrt(dot(V',V')): ((x*x + y*y) + z*z)
    return np.sqrt((v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]) + v[..., 2] * v[..., 2])


def _dot(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]


def _cross(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    c = np.empty_like(a)
    c[..., 0] = a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1]
    c[..., 1] = a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2]
    c[..., 2] = a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]
    return c


def face_geometry(points, face_ptr, face_nodes):
    """Face centroid, area vector and area (cfdProcessOpenFoamMesh.m:15-58)."""
    F = face_ptr.shape[0] - 1
    cnt = np.diff(face_ptr)
    centroid = np.zeros((F, 3))
    Sf = np.zeros((F, 3))
    area = np.zeros(F)
    for k in np.unique(cnt):
        idx = np.nonzero(cnt == k)[0]
        nodes = face_nodes[face_ptr[idx][:, None] + np.arange(k)[None, :]]   # [m,k]
        P = points[nodes]                                                    # [m,k,3]
        centre = np.zeros((idx.size, 3))
        for j in range(k):
            centre = centre + P[:, j]
        centre = centre / k
        c = np.zeros((idx.size, 3))
        s = np.zeros((idx.size, 3))
        a = np.zeros(idx.size)
        for t in range(k):
            p1 = centre
            p2 = P[:, t]
            p3 = P[:, (t + 1) % k]
            lc = ((p1 + p2) + p3) / 3
            lS = 0.5 * _cross(p2 - p1, p3 - p1)
            la = _mag(lS)
            c = c + la[:, None] * lc
            s = s + lS
            a = a + la
        centroid[idx] = c / a[:, None]
        Sf[idx] = s
        area[idx] = a
    return centroid, Sf, area


def element_lists(n_cells, owner, neighbour, n_faces):
    """Per-element face list, face sign and cconn in reference order
    (cfdReadPolyMesh.m:139-178): interior faces appended in face order to the
    owner (+1, neighbour appended to iNeighbours) and to the neighbour (-1,
    owner appended), then boundary faces in face order (+1)."""
    Fi = neighbour.shape[0]
    fi = np.arange(Fi, dtype=np.int64)
    fb = np.arange(Fi, n_faces, dtype=np.int64)
    cells = np.concatenate([owner[:Fi], neighbour, owner[Fi:]]).astype(np.int64)
    faces = np.concatenate([fi, fi, fb])
    sign = np.concatenate([np.ones(Fi), -np.ones(Fi), np.ones(fb.size)]).astype(np.int8)
    isb = np.concatenate([np.zeros(2 * Fi, np.int8), np.ones(fb.size, np.int8)])
    other = np.concatenate([neighbour, owner[:Fi], -np.ones(fb.size, np.int64)]).astype(np.int64)
    order = np.lexsort((faces, isb, cells))
    cells, faces, sign, isb, other = cells[order], faces[order], sign[order], isb[order], other[order]
    cf_ptr = np.zeros(n_cells + 1, np.int64)
    np.add.at(cf_ptr, cells + 1, 1)
    cf_ptr = np.cumsum(cf_ptr)
    inter = isb == 0
    cc_cells = cells[inter]
    cconn = other[inter].astype(np.int32)
    cconn_face = faces[inter].astype(np.int32)
    cc_ptr = np.zeros(n_cells + 1, np.int64)
    np.add.at(cc_ptr, cc_cells + 1, 1)
    cc_ptr = np.cumsum(cc_ptr)
    # slots: owner's entry for face f is where (cell=owner, face=f) sits in the CSR
    pos = np.arange(cconn.size, dtype=np.int64)
    own_nnz = np.empty(Fi, np.int64)
    nei_nnz = np.empty(Fi, np.int64)
    is_owner_entry = sign[inter] > 0
    own_nnz[cconn_face[is_owner_entry]] = pos[is_owner_entry]
    nei_nnz[cconn_face[~is_owner_entry]] = pos[~is_owner_entry]
    return (cf_ptr, faces.astype(np.int32), sign, cc_ptr.astype(np.int32), cconn, cconn_face,
            own_nnz.astype(np.int32), nei_nnz.astype(np.int32))


def process_mesh(pm: PolyMesh, apex: str = "reference") -> Mesh:
    """cfdProcessOpenFoamMesh restated (see module docstring)."""
    if apex not in ("reference", "own"):
        raise ValueError("apex must be 'reference' or 'own'")
    N, F, Fi = pm.n_cells, pm.n_faces, pm.n_interior_faces
    owner = pm.owner.astype(np.int32)
    nei = pm.neighbour.astype(np.int32)
    fc, Sf, area = face_geometry(pm.points, pm.face_ptr, pm.face_nodes)
    cf_ptr, cfaces, csign, cc_ptr, cconn, cconn_face, own_nnz, nei_nnz = element_lists(N, owner, nei, F)

    # element volume / centroid (cfdProcessOpenFoamMesh.m:62-103)
    nf = np.diff(cf_ptr)
    if apex == "reference":
        prefix = np.cumsum(fc, axis=0)    # sequential sums of global face centroids 1..k
        centre = prefix[nf - 1] / nf[:, None]
    else:
        centre = np.zeros((N, 3))
        for k in range(int(nf.max())):
            cells = np.nonzero(nf > k)[0]
            centre[cells] = centre[cells] + fc[cfaces[cf_ptr[cells] + k]]
        centre = centre / nf[:, None]
    vcs = np.zeros((N, 3))
    vs = np.zeros(N)
    maxnf = int(nf.max())
    for k in range(maxnf):
        has = nf > k
        cells = np.nonzero(has)[0]
        f = cfaces[cf_ptr[cells] + k]
        s = csign[cf_ptr[cells] + k].astype(np.float64)
        S = Sf[f] * s[:, None]
        Cf = fc[f] - centre[cells]
        lv = _dot(S, Cf) / 3
        lc = 0.75 * fc[f] + 0.25 * centre[cells]
        vcs[cells] = vcs[cells] + lc * lv[:, None]
        vs[cells] = vs[cells] + lv
    centroid = vcs / vs[:, None]
    volume = vs

    CN = np.zeros((F, 3))
    eCN = np.zeros((F, 3))
    gDiff = np.zeros(F)
    T = np.zeros((F, 3))
    gf = np.zeros(F)
    wd = np.zeros(F)
    # interior (:108-132)
    o, n = owner[:Fi], nei
    nfv = Sf[:Fi] / area[:Fi, None]
    cn = centroid[n] - centroid[o]
    e = cn / _mag(cn)[:, None]
    E = area[:Fi, None] * e
    CN[:Fi] = cn
    eCN[:Fi] = e
    gDiff[:Fi] = _mag(E) / _mag(cn)
    T[:Fi] = Sf[:Fi] - E
    Cf = fc[:Fi] - centroid[o]
    fF = centroid[n] - fc[:Fi]
    gf[:Fi] = _dot(Cf, nfv) / (_dot(Cf, nfv) + _dot(fF, nfv))
    # boundary (:134-155)
    ob = owner[Fi:]
    cnb = fc[Fi:] - centroid[ob]
    CN[Fi:] = cnb
    gDiff[Fi:] = area[Fi:] * area[Fi:] / _dot(cnb, Sf[Fi:])
    eb = cnb / _mag(cnb)[:, None]
    eCN[Fi:] = eb
    T[Fi:] = Sf[Fi:] - area[Fi:, None] * eb
    gf[Fi:] = 1.0
    wd[Fi:] = _dot(cnb, Sf[Fi:]) / _mag(Sf[Fi:])

    return Mesh(N, F, Fi, pm.points, pm.face_ptr, pm.face_nodes, owner, nei, list(pm.patches),
                fc, Sf, area, CN, eCN, gDiff, T, gf, wd, volume, centroid,
                cf_ptr, cfaces, csign, cc_ptr, cconn, cconn_face, own_nnz, nei_nnz)
