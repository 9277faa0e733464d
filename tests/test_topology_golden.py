"""Integer topology pinned to the reference's own outputs (CPU).

``tests/golden/topology_vectors.npz`` holds what the reference's
``computeEdges`` / ``computeFaces`` / ``computeFacesEdges`` (``mesh.py:18-125``),
``computeConnectivityDOFS`` (``hvfem.py:15-98``), ``computeBoundaryFaces`` /
``computeBoundaryEdges`` / ``computeBoundaries`` (``mesh.py:149-321``) and
``computeElementOrientation`` (``hvfem.py:122-220``) return on the
reference's test mesh (9453 tetrahedra) and on a 3^3 Kuhn grid; it is
written by ``tests/golden/make_topology_vectors.py``.  Both the oracle and
the product's host topology (``petgem_amd.mesh.Topology``) must reproduce
them exactly.

Independently of any fixture, a conforming H(curl) discretisation has a
tangential trace that is single-valued across every interior face: the
field of an arbitrary DOF vector, evaluated from the two elements that share
a face, must have the same tangential component there.  A wrong face owner,
face code or edge flag breaks that (shown by planting a wrong face code in
a copy of the oracle's orientations).
"""
import hashlib
import os

import numpy as np
import pytest

from oracle import fem, topology as otopo
from petgem_amd.mesh import Topology, read_gmsh_v2

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KEYS = ("elems_edges", "edges_nodes", "elems_faces", "faces_nodes", "faces_edges")


@pytest.fixture(scope="module")
def ref():
    return np.load(os.path.join(GOLDEN, "topology_vectors.npz"))


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i8").tobytes()).hexdigest()


def _check_dofs(ref, tag, p, dofs):
    if f"{tag}_dofs_p{p}" in ref.files:
        assert np.array_equal(dofs, ref[f"{tag}_dofs_p{p}"])
    else:
        rows = ref[f"{tag}_dofs_p{p}_rows"]
        assert np.array_equal(dofs[rows], ref[f"{tag}_dofs_p{p}_sample"])
        assert _digest(dofs) == str(ref[f"{tag}_dofs_p{p}_sha256"])


def test_reader_matches_reference_input(ref):
    _, en, _ = read_gmsh_v2(os.path.join(GOLDEN, "test_mesh.msh"))
    assert np.array_equal(en, ref["mesh_elems_nodes"])


@pytest.mark.parametrize("tag", ["mesh", "kuhn3"])
@pytest.mark.parametrize("p", [1, 2, 3])
def test_oracle_topology_equals_reference(ref, tag, p):
    en = ref[f"{tag}_elems_nodes"].astype(np.int64)
    o = otopo.build(en, p)
    for k in KEYS:
        assert np.array_equal(o[k], ref[f"{tag}_{k}"]), k
    _check_dofs(ref, tag, p, o["dofs"])
    assert o["ndof"] == int(ref[f"{tag}_ndof_p{p}"])
    assert np.array_equal(o["boundary_dofs"], ref[f"{tag}_boundary_dofs_p{p}"])
    assert np.array_equal(o["edge_orient"], ref[f"{tag}_edge_orient"])
    assert np.array_equal(o["face_orient"], ref[f"{tag}_face_orient"])


@pytest.mark.parametrize("tag", ["mesh", "kuhn3"])
@pytest.mark.parametrize("p", [1, 2, 3])
def test_product_topology_equals_reference(ref, tag, p):
    en = ref[f"{tag}_elems_nodes"].astype(np.int64)
    pts = np.zeros((int(en.max()) + 1, 3))
    T = Topology(pts, en, p)
    for k in KEYS:
        assert np.array_equal(getattr(T, k), ref[f"{tag}_{k}"]), k
    _check_dofs(ref, tag, p, T.dofs)
    assert T.ndof == int(ref[f"{tag}_ndof_p{p}"])
    assert np.array_equal(T.boundary_faces, ref[f"{tag}_boundary_faces"])
    assert np.array_equal(T.boundary_edges, ref[f"{tag}_boundary_edges"])
    assert np.array_equal(T.boundary_dofs, ref[f"{tag}_boundary_dofs_p{p}"])


def interior_face_pairs(elems_faces):
    """(face, lower element, higher element) for every face shared by two elements."""
    ef = np.asarray(elems_faces)
    flat = ef.ravel()
    order = np.argsort(flat, kind="stable")
    f_sorted = flat[order]
    same = np.nonzero(f_sorted[1:] == f_sorted[:-1])[0]
    return f_sorted[same], order[same] // 4, order[same + 1] // 4


def tangential_jump(coords, dofs, eo, fo, faces_nodes, points, p, faces, e1, e2, x):
    """Largest |n x (E_e1 - E_e2)| / max|E| over points of the listed faces."""
    lam = np.array([[1, 1, 1], [4, 1, 1], [1, 4, 2], [2, 2, 5]], dtype=np.float64)
    lam /= lam.sum(axis=1, keepdims=True)
    worst, scale = 0.0, 0.0
    for f, a, b in zip(faces, e1, e2):
        tri = points[faces_nodes[f]]
        nrm = np.cross(tri[1] - tri[0], tri[2] - tri[0])
        nrm /= np.linalg.norm(nrm)
        xyz = lam @ tri
        vals = []
        for e in (a, b):
            xi = fem.xyz_to_reference(coords[e], xyz)
            N, _ = fem.physical_basis(coords[e], eo[e], fo[e], p, xi)      # (P,n,3)
            vals.append(np.einsum("pia,i->pa", N, x[dofs[e]]))
        jump = np.cross(vals[0] - vals[1], nrm)
        worst = max(worst, np.abs(jump).max())
        scale = max(scale, np.abs(vals[0]).max(), np.abs(vals[1]).max())
    return worst / scale


@pytest.mark.parametrize("p", [1, 2, 3])
def test_tangential_continuity_on_reference_mesh(ref, p):
    pts, en, _ = read_gmsh_v2(os.path.join(GOLDEN, "test_mesh.msh"))
    T = Topology(pts, en, p)
    eo = ref["mesh_edge_orient"].astype(np.int64)
    fo = ref["mesh_face_orient"].astype(np.int64)
    faces, e1, e2 = interior_face_pairs(T.elems_faces)
    rng = np.random.default_rng(11 + p)
    pick = rng.choice(faces.size, 60, replace=False)
    x = rng.standard_normal(T.ndof) + 1j * rng.standard_normal(T.ndof)
    args = (T.coords, T.dofs, eo, fo, T.faces_nodes, T.points, p, faces[pick], e1[pick], e2[pick])
    assert tangential_jump(*args, x) <= 1e-11
    if p >= 2:
        # a face code read wrongly in one element shows up as a jump on that face
        bad = fo.copy()
        k = int(np.nonzero(T.elems_faces[e2[pick[0]]] == faces[pick[0]])[0][0])
        bad[e2[pick[0]], k] = (bad[e2[pick[0]], k] + 1) % 6
        args_bad = (T.coords, T.dofs, eo, bad) + args[4:]
        assert tangential_jump(*args_bad, x) > 1e-3
    # a flipped edge flag breaks continuity at every order
    bad_e = eo.copy()
    k = int(np.nonzero(T.elems_faces[e2[pick[1]]] == faces[pick[1]])[0][0])
    bad_e[e2[pick[1]], otopo.FACE_EDGE_TRIPLES[k][0]] ^= 1
    args_bad = (T.coords, T.dofs, bad_e, fo) + args[4:]
    assert tangential_jump(*args_bad, x) > 1e-3
