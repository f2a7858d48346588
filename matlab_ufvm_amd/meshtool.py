"""Standalone mesh tool: an OpenFOAM case's constant/polyMesh read and
processed into the finite-volume geometry of the GPU path's input contract
(``uFVM/src/utilities/IO/cfdReadPolyMesh.m:25-211`` then
``cfdProcessOpenFoamMesh.m:15-241``), summarised and optionally saved as an
``.npz`` of the processed arrays (0-based, cconn as CSR in reference order).

    python -m matlab_ufvm_amd.meshtool CASE_DIR [--apex reference|own] [--out mesh.npz]
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

from .mesh import Mesh, process_mesh
from .polymesh import read_polymesh

# processed arrays written by save_npz (the C ABI's ufvm_mesh_desc plus walldist)
ARRAYS = ("owner", "neighbour", "face_centroid", "Sf", "area", "CN", "eCN", "gDiff", "T", "gf",
          "walldist", "volume", "centroid", "cconn_ptr", "cconn", "own_nnz", "nei_nnz")


def summary(m: Mesh) -> dict:
    vol = m.volume
    return {"cells": int(m.n_cells), "faces": int(m.n_faces),
            "interior_faces": int(m.n_interior), "boundary_faces": int(m.n_bfaces),
            "offdiagonal_nnz": int(m.cconn.size),
            "patches": [{"name": p.name, "type": p.type, "start": int(p.start),
                         "faces": int(p.nfaces)} for p in m.patches],
            "volume": {"total": float(vol.sum()), "min": float(vol.min()),
                       "max": float(vol.max())},
            "negative_gDiff_faces": int(np.count_nonzero(m.gDiff < 0))}


def save_npz(m: Mesh, path: str) -> None:
    arrays = {k: getattr(m, k) for k in ARRAYS}
    arrays["patch_start"] = np.array([p.start for p in m.patches], np.int32)
    arrays["patch_nfaces"] = np.array([p.nfaces for p in m.patches], np.int32)
    arrays["patch_names"] = np.array([p.name for p in m.patches])
    arrays["patch_types"] = np.array([p.type for p in m.patches])
    np.savez(path, n_cells=m.n_cells, n_faces=m.n_faces, n_interior=m.n_interior, **arrays)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("case_dir", help="OpenFOAM case (reads CASE_DIR/constant/polyMesh)")
    ap.add_argument("--apex", choices=["reference", "own"], default="reference",
                    help="pyramid apex of the cell volumes: the reference's global-face mean "
                         "(bit-exact) or the cell's own faces")
    ap.add_argument("--out", help="write the processed arrays to this .npz")
    args = ap.parse_args(argv)
    m = process_mesh(read_polymesh(args.case_dir), apex=args.apex)
    print(json.dumps(summary(m), indent=1))
    if args.out:
        save_npz(m, args.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
