"""Synthetic OpenFOAM-ordered meshes for the benchmark and test cases.

* ``box_hex(nx, ny, nz, ...)``: a unit box of hexahedra written in OpenFOAM
  polyMesh conventions -- interior faces in upper-triangular order (owner <
  neighbour, sorted by (owner, neighbour)), face normals pointing from owner to
  neighbour, boundary faces grouped by patch with outward normals.  Optional
  seeded node jitter (interior nodes in all directions, boundary nodes only
  tangentially so the box stays a box), and an optional seeded random cell
  permutation followed by reverse Cuthill-McKee renumbering, after which the
  faces are re-sorted back to upper-triangular order.
* ``box_tet(n, ...)``: every hex of an n^3 grid split into 6 tetrahedra around
  its main diagonal (Kuhn/Freudenthal split, conforming across hexes).

The reference reads such meshes from ``constant/polyMesh``
(``src/utilities/IO/cfdReadPolyMesh.m:25-211``); these generators produce the
same ``PolyMesh`` record directly.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import reverse_cuthill_mckee

from .polymesh import Patch, PolyMesh

# side name -> (axis, 0 = low side / 1 = high side)
SIDES = {"x0": (0, 0), "x1": (0, 1), "y0": (1, 0), "y1": (1, 1), "z0": (2, 0), "z1": (2, 1)}

CAVITY_3D = [("movingWall", "wall", ["y1"]), ("fixedWalls", "wall", ["x0", "x1", "y0", "z0", "z1"])]
CAVITY_2D = [("movingWall", "wall", ["y1"]), ("fixedWalls", "wall", ["x0", "x1", "y0"]),
             ("frontAndBack", "empty", ["z0", "z1"])]
POISSON_3D = [("dirichlet", "patch", ["x0"]), ("walls", "wall", ["x1", "y0", "y1", "z0", "z1"])]


def _node_id(i, j, k, nx, ny):
    return i + (nx + 1) * (j + (ny + 1) * k)


def _jitter_nodes(nx, ny, nz, lx, ly, lz, amp, seed, planar_z=False):
    I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    I, J, K = (a.transpose(2, 1, 0).reshape(-1) for a in (I, J, K))   # i fastest
    h = np.array([lx / nx, ly / ny, lz / nz])
    pts = np.stack([I * h[0], J * h[1], K * h[2]], axis=1).astype(np.float64)
    if amp > 0:
        rng = np.random.default_rng(seed)
        d = rng.uniform(-amp, amp, size=pts.shape) * h[None, :]
        # boundary nodes move only tangentially
        d[(I == 0) | (I == nx), 0] = 0.0
        d[(J == 0) | (J == ny), 1] = 0.0
        d[(K == 0) | (K == nz), 2] = 0.0
        if planar_z:
            d[:, 2] = 0.0
        pts = pts + d
    return pts


def _hex_faces(nx, ny, nz):
    """Interior faces (owner, neighbour, 4 nodes) and boundary faces per side."""
    N = nx * ny * nz
    c = np.arange(N, dtype=np.int64)
    i = c % nx
    j = (c // nx) % ny
    k = c // (nx * ny)

    def quad(axis, i, j, k):
        # quad on the plane at the low side of cell (i,j,k) along axis, normal +axis
        if axis == 0:
            return np.stack([_node_id(i, j, k, nx, ny), _node_id(i, j + 1, k, nx, ny),
                             _node_id(i, j + 1, k + 1, nx, ny), _node_id(i, j, k + 1, nx, ny)], 1)
        if axis == 1:
            return np.stack([_node_id(i, j, k, nx, ny), _node_id(i, j, k + 1, nx, ny),
                             _node_id(i + 1, j, k + 1, nx, ny), _node_id(i + 1, j, k, nx, ny)], 1)
        return np.stack([_node_id(i, j, k, nx, ny), _node_id(i + 1, j, k, nx, ny),
                         _node_id(i + 1, j + 1, k, nx, ny), _node_id(i, j + 1, k, nx, ny)], 1)

    own, nei, nodes = [], [], []
    strides = (1, nx, nx * ny)
    ijk = (i, j, k)
    dims = (nx, ny, nz)
    for ax in range(3):
        m = ijk[ax] < dims[ax] - 1
        cc = c[m]
        own.append(cc)
        nei.append(cc + strides[ax])
        ii, jj, kk = i[m].copy(), j[m].copy(), k[m].copy()
        (ii, jj, kk)[ax][:] += 1
        nodes.append(quad(ax, ii, jj, kk))
    own = np.concatenate(own)
    nei = np.concatenate(nei)
    nodes = np.concatenate(nodes)
    order = np.lexsort((nei, own))
    own, nei, nodes = own[order], nei[order], nodes[order]

    bfaces = {}
    for name, (ax, side) in SIDES.items():
        m = ijk[ax] == (0 if side == 0 else dims[ax] - 1)
        cc = c[m]
        ii, jj, kk = i[m].copy(), j[m].copy(), k[m].copy()
        if side == 1:
            (ii, jj, kk)[ax][:] += 1
        q = quad(ax, ii, jj, kk)
        if side == 0:
            q = q[:, ::-1]          # outward = -axis
        bfaces[name] = (cc, q)
    return own, nei, nodes, bfaces


def _assemble(points, own, nei, nodes, bfaces, patch_spec, n_cells, perm=None):
    """Build a PolyMesh; optionally renumber cells with new_id = perm[old]."""
    bown, bnodes, patches = [], [], []
    start = own.shape[0]
    for name, typ, sides in patch_spec:
        cc = np.concatenate([bfaces[s][0] for s in sides])
        qq = np.concatenate([bfaces[s][1] for s in sides])
        bown.append(cc)
        bnodes.append(qq)
        patches.append(Patch(name, typ, start, int(cc.size)))
        start += cc.size
    if perm is not None:
        o2, n2 = perm[own], perm[nei]
        swap = o2 > n2
        own = np.where(swap, n2, o2)
        nei = np.where(swap, o2, n2)
        nodes = nodes.copy()
        nodes[swap] = nodes[swap][:, ::-1]
        order = np.lexsort((nei, own))
        own, nei, nodes = own[order], nei[order], nodes[order]
        for p in range(len(bown)):
            bown[p] = perm[bown[p]]
            o = np.argsort(bown[p], kind="stable")
            bown[p], bnodes[p] = bown[p][o], bnodes[p][o]
    k = nodes.shape[1]
    all_nodes = np.concatenate([nodes] + bnodes).astype(np.int32)
    F = all_nodes.shape[0]
    owner = np.concatenate([own] + bown).astype(np.int32)
    return PolyMesh(points, np.arange(F + 1, dtype=np.int64) * k, all_nodes.reshape(-1), owner,
                    nei.astype(np.int32), patches, n_cells)


def rcm_permutation(n_cells, own, nei, seed):
    """new_id[old]: seeded random shuffle followed by reverse Cuthill-McKee."""
    rng = np.random.default_rng(seed)
    shuffle = rng.permutation(n_cells)               # new1 = shuffle[old]
    a, b = shuffle[own], shuffle[nei]
    A = sp.coo_matrix((np.ones(a.size, np.int8), (a, b)), shape=(n_cells, n_cells)).tocsr()
    A = A + A.T
    order = reverse_cuthill_mckee(A.tocsr(), symmetric_mode=True)   # order[new2] = new1
    inv = np.empty(n_cells, np.int64)
    inv[order] = np.arange(n_cells)
    return inv[shuffle]


def box_hex(nx, ny, nz, patch_spec=None, lengths=(1.0, 1.0, 1.0), jitter=0.0, seed=0,
            renumber_seed=None) -> PolyMesh:
    if patch_spec is None:
        patch_spec = CAVITY_2D if nz == 1 else CAVITY_3D
    planar = any(t == "empty" for _, t, _ in patch_spec)
    pts = _jitter_nodes(nx, ny, nz, *lengths, jitter, seed, planar_z=planar)
    own, nei, nodes, bfaces = _hex_faces(nx, ny, nz)
    N = nx * ny * nz
    perm = None
    if renumber_seed is not None:
        perm = rcm_permutation(N, own, nei, renumber_seed)
    return _assemble(pts, own, nei, nodes, bfaces, patch_spec, N, perm)


# Kuhn split: 6 tets sharing the hex diagonal (0,0,0)-(1,1,1); local vertex = (dx,dy,dz)
_KUHN = []
for _p in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
    v = [np.zeros(3, int)]
    for ax in _p:
        w = v[-1].copy()
        w[ax] = 1
        v.append(w)
    _KUHN.append(v)


def box_tet(n, patch_spec=POISSON_3D, jitter=0.0, seed=0, renumber_seed=None) -> PolyMesh:
    nx = ny = nz = n
    pts = _jitter_nodes(nx, ny, nz, 1.0, 1.0, 1.0, jitter, seed)
    H = nx * ny * nz
    c = np.arange(H, dtype=np.int64)
    i, j, k = c % nx, (c // nx) % ny, c // (nx * ny)
    tets = []
    for verts in _KUHN:
        tets.append(np.stack([_node_id(i + v[0], j + v[1], k + v[2], nx, ny) for v in verts], 1))
    tets = np.stack(tets, 1).reshape(-1, 4)          # cell = 6*hex + t
    N = tets.shape[0]
    # positively oriented tets: det > 0
    P = pts[tets]
    det = np.einsum("ij,ij->i", np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), P[:, 3] - P[:, 0])
    neg = det < 0
    tets[neg] = tets[neg][:, [0, 2, 1, 3]]
    # faces opposite each vertex, oriented outward: for a positive tet (a,b,c,d),
    # outward faces are (b,c,d), (a,d,c), (a,b,d), (a,c,b)
    loc = np.array([[1, 2, 3], [0, 3, 2], [0, 1, 3], [0, 2, 1]])
    fn = tets[:, loc].reshape(-1, 3)
    fc = np.repeat(np.arange(N, dtype=np.int64), 4)
    key = np.sort(fn, axis=1)
    _, inv, cnt = np.unique(key, axis=0, return_inverse=True, return_counts=True)
    inv = inv.reshape(-1)
    order = np.argsort(inv, kind="stable")
    sinv = inv[order]
    first = np.ones(sinv.size, bool)
    first[1:] = sinv[1:] != sinv[:-1]
    # interior: pairs
    pair_start = np.nonzero(first & (cnt[sinv] == 2))[0]
    fa, fb = order[pair_start], order[pair_start + 1]
    ca, cb = fc[fa], fc[fb]
    sw = ca > cb
    own = np.where(sw, cb, ca)
    nei = np.where(sw, ca, cb)
    nodes = np.where(sw[:, None], fn[fb], fn[fa])     # outward from owner
    o = np.lexsort((nei, own))
    own, nei, nodes = own[o], nei[o], nodes[o]
    single = order[np.nonzero(first & (cnt[sinv] == 1))[0]]
    bc, bn = fc[single], fn[single]
    cen = pts[bn].mean(axis=1)
    tol = 1e-9
    bfaces = {}
    for name, (ax, side) in SIDES.items():
        target = 0.0 if side == 0 else 1.0
        m = np.abs(cen[:, ax] - target) < tol
        oo = np.argsort(bc[m], kind="stable")
        bfaces[name] = (bc[m][oo], bn[m][oo])
    perm = None
    if renumber_seed is not None:
        perm = rcm_permutation(N, own, nei, renumber_seed)
    return _assemble(pts, own, nei, nodes, bfaces, patch_spec, N, perm)
