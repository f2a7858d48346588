"""Synthetic workloads of the benchmark and tests (inputs, not the solve path).

* ``cavity_mesh`` / ``poisson_tet_mesh``: the jittered, renumbered meshes of
  the 3-D lid-driven-cavity and tetrahedral pressure-Poisson cases.
* ``channel_mesh`` / ``channel_bcs``: a jittered, renumbered 3-D channel with
  an inlet, a fixed-pressure outlet, walls of each velocity condition and a
  symmetry plane (every boundary branch of the SIMPLE path at once).
* ``poisson_system``: the pressure-Poisson operator of a mesh in the
  reference's coefficient layout -- per face a coefficient C_f = Gamma*gDiff_f,
  scattered in face order as uFVM/src/fvm/assemble/
  cfdAssembleIntoGlobalMatrixFaceFluxes.m:25-55 does (owner: ac += C,
  anb -= C; neighbour: ac += C, anb -= C; Dirichlet boundary faces: ac += C),
  i.e. the sign convention of the pressure-correction matrix
  (cfdAssembleMdotTerm.m:158-159: diagonal positive, neighbours negative).
* ``gaussian_rhs``: b_i = V_i * xi_i with seeded standard-normal xi, minus
  its mean.
"""
from __future__ import annotations

import numpy as np

from .mesh import Mesh, process_mesh
from .meshgen import CAVITY_2D, CAVITY_3D, POISSON_3D, box_hex, box_tet

CHANNEL = [("inlet", "inlet", ["x0"]), ("outlet", "outlet", ["x1"]),
           ("lowerWall", "wall", ["y0"]), ("upperWall", "wall", ["y1"]),
           ("front", "symmetry", ["z0"]), ("back", "wall", ["z1"])]


def cavity_mesh(n: int, nz: int | None = None, jitter: float = 0.15, seed: int = 2026,
                renumber_seed: int | None = 2027, apex: str = "reference") -> Mesh:
    nz = n if nz is None else nz
    spec = CAVITY_2D if nz == 1 else CAVITY_3D
    lengths = (1.0, 1.0, 1.0 / n if nz == 1 else 1.0)
    return process_mesh(box_hex(n, n, nz, spec, lengths=lengths, jitter=jitter, seed=seed,
                                renumber_seed=renumber_seed), apex=apex)


def channel_mesh(nx: int, ny: int, nz: int, jitter: float = 0.15, seed: int = 2031,
                 renumber_seed: int | None = 2032, apex: str = "reference") -> Mesh:
    """[0, 2] x [0, 1] x [0, 1] channel of nx x ny x nz hexahedra"""
    return process_mesh(box_hex(nx, ny, nz, CHANNEL, lengths=(2.0, 1.0, 1.0), jitter=jitter,
                                seed=seed, renumber_seed=renumber_seed), apex=apex)


def channel_bcs(u_in=(1.0, 0.0, 0.0), p_out=0.0, lower="noSlip", upper="slip",
                back="fixedValue"):
    """{patch: (U bc, U value, p bc, p value)} of channel_mesh: fixedValue
    inlet with zeroGradient p, zeroGradient outlet at fixed p, walls with the
    given U conditions and zeroGradient p, a symmetry plane"""
    z = (0.0, 0.0, 0.0)
    return {"inlet": ("fixedValue", tuple(u_in), "zeroGradient", 0.0),
            "outlet": ("zeroGradient", z, "fixedValue", float(p_out)),
            "lowerWall": (lower, z, "zeroGradient", 0.0),
            "upperWall": (upper, z, "zeroGradient", 0.0),
            "front": ("symmetry", z, "symmetry", 0.0),
            "back": (back, z, "zeroGradient", 0.0)}


def poisson_tet_mesh(n: int, jitter: float = 0.15, seed: int = 2028,
                     renumber_seed: int | None = 2029) -> Mesh:
    return process_mesh(box_tet(n, POISSON_3D, jitter=jitter, seed=seed,
                                renumber_seed=renumber_seed))


def poisson_system(mesh: Mesh, dirichlet=("dirichlet",), gamma: float = 1.0):
    """(ptr, col, ac, anb) of the pressure-Poisson operator on ``mesh``."""
    Fi = mesh.n_interior
    C = gamma * mesh.gDiff
    anb = np.zeros(mesh.cconn.size)
    anb[mesh.own_nnz] = -C[:Fi]
    anb[mesh.nei_nnz] = -C[:Fi]
    # diagonal: contributions in increasing face index, owner and neighbour alike
    cells = [mesh.owner[:Fi], mesh.neighbour]
    faces = [np.arange(Fi), np.arange(Fi)]
    for p in mesh.patches:
        if p.name in dirichlet:
            cells.append(mesh.owner[p.start:p.start + p.nfaces])
            faces.append(np.arange(p.start, p.start + p.nfaces))
    cells = np.concatenate(cells)
    faces = np.concatenate(faces)
    order = np.argsort(faces, kind="stable")
    ac = np.zeros(mesh.n_cells)
    np.add.at(ac, cells[order], C[faces[order]])
    return mesh.cconn_ptr.copy(), mesh.cconn.copy(), ac, anb


def gaussian_rhs(mesh: Mesh, seed: int = 2029) -> np.ndarray:
    rng = np.random.default_rng(seed)
    b = mesh.volume * rng.standard_normal(mesh.n_cells)
    return b - b.mean()
