"""Host-side mesh geometry: pinned bit-for-bit to the reference's own processed
mesh (tutorials/basic/cavity1/constant/polyMesh/theMesh.mat, converted to
tests/golden/cavity1_mesh.npz by tests/golden/make_cavity_mesh_fixture.py),
plus closed-form invariants of the synthetic meshes."""
import os

import numpy as np
import pytest

from matlab_ufvm_amd.mesh import process_mesh
from matlab_ufvm_amd.meshgen import box_hex, box_tet
from matlab_ufvm_amd.polymesh import Patch, PolyMesh

GOLD = os.path.join(os.path.dirname(__file__), "golden", "cavity1_mesh.npz")


@pytest.fixture(scope="module")
def cavity1():
    z = np.load(GOLD)
    patches = [Patch(str(a), str(b), int(c), int(d)) for a, b, c, d in
               zip(z["patch_names"], z["patch_types"], z["patch_start"], z["patch_nfaces"])]
    pm = PolyMesh(z["points"], z["face_ptr"], z["face_nodes"], z["owner"], z["neighbour"],
                  patches, int(z["owner"].max()) + 1)
    return z, process_mesh(pm)


@pytest.mark.parametrize("ours,ref", [
    ("face_centroid", "ref_face_centroid"), ("Sf", "ref_Sf"), ("area", "ref_area"),
    ("CN", "ref_CN"), ("eCN", "ref_eCN"), ("gDiff", "ref_gDiff"), ("T", "ref_T"),
    ("gf", "ref_gf"), ("walldist", "ref_walldist"), ("volume", "ref_volume"),
    ("centroid", "ref_centroid")])
def test_geometry_bitexact_vs_reference(cavity1, ours, ref):
    z, m = cavity1
    a, b = getattr(m, ours), z[ref]
    assert a.shape == b.shape
    assert np.array_equal(a, b), f"{ours}: {np.count_nonzero(a != b)} entries differ"


def test_apex_noise_reproduced(cavity1):
    # the element-centre apex quirk leaves z-noise of 8.67e-19 in CN of face 1
    z, m = cavity1
    assert m.CN[0, 2] == z["ref_CN"][0, 2] != 0.0


def test_connectivity_bitexact_vs_reference(cavity1):
    z, m = cavity1
    assert np.array_equal(m.cconn, z["ref_cconn"])
    assert np.array_equal(m.csize, z["ref_csize"])
    assert np.array_equal(m.cell_faces, z["ref_iFaces"])
    assert np.array_equal(m.cell_face_sign, z["ref_faceSign"])
    Fi = m.n_interior
    assert np.array_equal(m.own_nnz - m.cconn_ptr[m.owner[:Fi]], z["ref_iOwnerNeighbourCoef"])
    assert np.array_equal(m.nei_nnz - m.cconn_ptr[m.neighbour], z["ref_iNeighbourOwnerCoef"])


def _closed_cell_sums(m):
    s = np.zeros((m.n_cells, 3))
    np.add.at(s, m.owner, m.Sf)
    np.add.at(s, m.neighbour, -m.Sf[:m.n_interior])
    return s


@pytest.mark.parametrize("kind", ["hex", "tet"])
def test_generated_mesh_invariants(kind):
    if kind == "hex":
        pm = box_hex(7, 6, 5, jitter=0.15, seed=3, renumber_seed=4)
        n_cells = 7 * 6 * 5
    else:
        pm = box_tet(4, jitter=0.1, seed=5, renumber_seed=6)
        n_cells = 6 * 64
    m = process_mesh(pm)
    assert m.n_cells == n_cells
    Fi = m.n_interior
    # upper-triangular face order: owner < neighbour, sorted by (owner, neighbour)
    assert np.all(m.owner[:Fi] < m.neighbour)
    key = m.owner[:Fi].astype(np.int64) * m.n_cells + m.neighbour
    assert np.all(np.diff(key) > 0)
    # closed cells, positive volumes filling the unit box
    assert np.abs(_closed_cell_sums(m)).max() < 1e-14
    assert m.volume.min() > 0
    assert abs(m.volume.sum() - 1.0) < 1e-12
    # interior normals point from owner to neighbour; boundary normals outward
    d = m.centroid[m.neighbour] - m.centroid[m.owner[:Fi]]
    assert np.all(np.einsum("ij,ij->i", d, m.Sf[:Fi]) > 0)
    db = m.face_centroid[Fi:] - m.centroid[m.owner[Fi:]]
    assert np.all(np.einsum("ij,ij->i", db, m.Sf[Fi:]) > 0)
    # cconn is the neighbour list in increasing face order -> sorted ascending here
    for i in range(m.n_cells):
        row = m.cconn[m.cconn_ptr[i]:m.cconn_ptr[i + 1]]
        assert np.all(np.diff(row) > 0)


def test_uniform_closed_forms():
    # uniform grid: gf = 1/2, T = 0, gDiff = A/h, volume = h^3
    m = process_mesh(box_hex(4, 4, 4))
    h = 0.25
    Fi = m.n_interior
    assert np.allclose(m.gf[:Fi], 0.5, atol=1e-15)
    assert np.abs(m.T[:Fi]).max() < 1e-15
    assert np.allclose(m.gDiff[:Fi], h * h / h, rtol=1e-14)
    assert np.allclose(m.volume, h ** 3, rtol=1e-14)
    # boundary: gDiff = A^2/(CN.Sf) = A/(h/2)
    assert np.allclose(m.gDiff[Fi:], h * h / (h / 2), rtol=1e-14)
    assert np.allclose(m.walldist[Fi:], h / 2, rtol=1e-14)


def test_reference_apex_degrades_warped_meshes():
    # The reference's pyramid apex (mean centroid of global faces 1..nFaces(e),
    # cfdProcessOpenFoamMesh.m:69-75) makes the centroid error of a warped-face
    # cell grow with the cell's distance from face 1.  On a 32^3 cavity jittered
    # by 15% some boundary faces already get CN.Sf < 0 (negative gDiff); the
    # element's own faces as apex keep every one positive, and the centroid it
    # gives stays inside the cell's node bounding box.
    from matlab_ufvm_amd.workloads import cavity_mesh
    ref = cavity_mesh(32, jitter=0.15, seed=2026, renumber_seed=2027)
    own = cavity_mesh(32, jitter=0.15, seed=2026, renumber_seed=2027, apex="own")
    Fi = ref.n_interior
    assert (ref.gDiff[Fi:] < 0).any()
    assert (own.gDiff[Fi:] > 0).all() and (own.gDiff[:Fi] > 0).all()
    # volumes do not depend on the apex (closed fan-triangulated surfaces)
    assert np.allclose(ref.volume, own.volume, rtol=1e-9)
    fid = np.repeat(np.arange(own.n_faces), np.diff(own.face_ptr))
    pts = own.points[own.face_nodes]
    inner = fid < Fi
    cells = np.concatenate([own.owner[fid], own.neighbour[fid[inner]]])
    p = np.concatenate([pts, pts[inner]])
    lo = np.full((own.n_cells, 3), np.inf)
    hi = np.full((own.n_cells, 3), -np.inf)
    np.minimum.at(lo, cells, p)
    np.maximum.at(hi, cells, p)
    assert np.all(own.centroid >= lo) and np.all(own.centroid <= hi)
    assert not (np.all(ref.centroid >= lo) and np.all(ref.centroid <= hi))
    # a uniform (planar-face) mesh is apex-independent up to rounding
    u = box_hex(5, 4, 3)
    assert np.allclose(process_mesh(u).centroid, process_mesh(u, apex="own").centroid,
                       rtol=0, atol=1e-15)
