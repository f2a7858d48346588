"""Write tests/golden/cavity1_mesh.npz from the reference tutorial's own data.

Inputs (data files held by the reference):
  tutorials/basic/cavity1/constant/polyMesh/{points,faces,owner,neighbour,boundary}
      -- the OpenFOAM ASCII mesh of the 20x20x1 cavity;
  tutorials/basic/cavity1/constant/polyMesh/theMesh.mat
      -- the reference's processed mesh struct (written by
         cfdGenerateInitialResults.m:16-23), read with scipy.io.loadmat as plain
         data (MAT v5 numeric arrays; nothing in the file is executed).

The npz stores the raw mesh arrays (inputs) and the reference's processed
geometry (expected outputs), 0-based.  Run from the repository root:

    python tests/golden/make_cavity_mesh_fixture.py
"""
import os
import sys

import numpy as np
import scipy.io as sio

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from matlab_ufvm_amd.polymesh import read_polymesh  # noqa: E402

CASE = "/root/reference/uFVM/tutorials/basic/cavity1"


def main():
    pm = read_polymesh(CASE)
    m = sio.loadmat(os.path.join(CASE, "constant/polyMesh/theMesh.mat"),
                    squeeze_me=False, struct_as_record=True)["theMesh"]
    faces = m["faces"][0, 0][0]
    elems = m["elements"][0, 0][0]
    F = faces.shape[0]
    N = elems.shape[0]
    Fi = int(m["numberOfInteriorFaces"][0, 0][0, 0])

    def fvec(name):
        return np.array([np.asarray(faces[i][name], dtype=np.float64).reshape(3) for i in range(F)])

    def fsc(name, lo=0, hi=F):
        return np.array([float(np.asarray(faces[i][name]).reshape(-1)[0]) for i in range(lo, hi)])

    out = dict(
        points=pm.points, face_ptr=pm.face_ptr, face_nodes=pm.face_nodes,
        owner=pm.owner, neighbour=pm.neighbour,
        patch_names=np.array([p.name for p in pm.patches]),
        patch_types=np.array([p.type for p in pm.patches]),
        patch_start=np.array([p.start for p in pm.patches]),
        patch_nfaces=np.array([p.nfaces for p in pm.patches]),
        # reference-processed geometry
        ref_face_centroid=fvec("centroid"), ref_Sf=fvec("Sf"), ref_area=fsc("area"),
        ref_CN=fvec("CN"), ref_gDiff=fsc("gDiff"), ref_T=fvec("T"), ref_gf=fsc("gf"),
        ref_walldist=fsc("walldist"),
        ref_eCN=fvec("eCN"),
        ref_iOwnerNeighbourCoef=fsc("iOwnerNeighbourCoef", 0, Fi).astype(np.int32) - 1,
        ref_iNeighbourOwnerCoef=fsc("iNeighbourOwnerCoef", 0, Fi).astype(np.int32) - 1,
        ref_volume=np.array([float(np.asarray(elems[i]["volume"]).reshape(-1)[0]) for i in range(N)]),
        ref_centroid=np.array([np.asarray(elems[i]["centroid"], dtype=np.float64).reshape(3)
                               for i in range(N)]),
    )
    cconn = m["cconn"][0, 0]
    csize = np.asarray(m["csize"][0, 0]).reshape(-1).astype(np.int32)
    flat = [np.asarray(cconn[i, 0]).reshape(-1).astype(np.int32) - 1 for i in range(N)]
    out["ref_csize"] = csize
    out["ref_cconn"] = np.concatenate(flat)
    ifaces = [np.asarray(elems[i]["iFaces"]).reshape(-1).astype(np.int32) - 1 for i in range(N)]
    out["ref_iFaces"] = np.concatenate(ifaces)
    out["ref_nfaces"] = np.array([a.size for a in ifaces], np.int32)
    out["ref_faceSign"] = np.concatenate(
        [np.asarray(elems[i]["faceSign"]).reshape(-1).astype(np.int8) for i in range(N)])
    dst = os.path.join(HERE, "cavity1_mesh.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, os.path.getsize(dst), "bytes")


if __name__ == "__main__":
    main()
