"""Mesh topology, DOF numbering and orientations -- oracle restatement.

Follows the reference rules literally (row-unique with first occurrence,
owner = first element in element order) rather than any shortcut:

* ``edges``: ``petgem/mesh.py:18-53`` + ``petgem/vectors.py:15-64`` -- edge id is
  the rank of the sorted node pair among all unique sorted pairs;
  ``edges_nodes`` rows are sorted pairs.
* ``faces``: ``petgem/mesh.py:56-89`` -- face id is the rank of the sorted
  node triple; ``faces_nodes`` keeps the first occurrence's node order.
* ``faces_edges``: ``petgem/mesh.py:92-125`` + ``petgem/vectors.py:81-132`` --
  owner element = lowest element index holding the face; the face's edges
  are the owner's local face-edge triple (0,1,2),(0,4,3),(1,5,4),(2,5,3).
* ``dof_connectivity``: ``petgem/hvfem.py:15-98``.
* ``orientations``: ``petgem/hvfem.py:122-220``.
* ``boundary_dofs``: ``petgem/mesh.py:149-321`` -- faces with one adjacent
  element, their edges, and the DOFs on both (returned as a sorted set).

Pinned by the reference's own golden values ``tests/test_mesh.py:31-36`` on
``tests/data/test_mesh.msh`` (counts and element 0's edges and faces), and
array-for-array by the outputs of the reference's ``computeEdges``,
``computeFaces``, ``computeFacesEdges``, ``computeConnectivityDOFS``,
``computeBoundaries`` and ``computeElementOrientation`` on that mesh and on a
Kuhn grid (golden vectors, ``tests/test_topology_golden.py``), plus the
tangential-continuity check of assembled fields across shared faces.
"""
import numpy as np

LOCAL_EDGES = ((0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3))
LOCAL_FACES = ((0, 1, 2), (0, 1, 3), (1, 2, 3), (0, 2, 3))
FACE_EDGE_TRIPLES = ((0, 1, 2), (0, 4, 3), (1, 5, 4), (2, 5, 3))
# (k1*10 + k2) -> face orientation code, hvfem.py:199-211
FACE_CODE = {12: 0, 31: 1, 23: 2, 32: 3, 13: 4, 21: 5}


def _unique_rows(rows):
    """np.unique over rows (lexicographic), returning first-occurrence index and inverse."""
    rows = np.ascontiguousarray(rows)
    view = rows.view([("", rows.dtype)] * rows.shape[1]).ravel()
    _, first, inv = np.unique(view, return_index=True, return_inverse=True)
    return first, inv.ravel()


def edges(elems_nodes):
    en = np.asarray(elems_nodes, dtype=np.int64)
    ne = en.shape[0]
    pairs = np.empty((ne * 6, 2), dtype=np.int64)
    for k, (a, b) in enumerate(LOCAL_EDGES):
        pairs[k::6, 0] = en[:, a]
        pairs[k::6, 1] = en[:, b]
    srt = np.sort(pairs, axis=1)
    first, inv = _unique_rows(srt)
    edges_nodes = np.sort(pairs[first], axis=1)
    return inv.reshape(ne, 6), edges_nodes


def faces(elems_nodes):
    en = np.asarray(elems_nodes, dtype=np.int64)
    ne = en.shape[0]
    tri = np.empty((ne * 4, 3), dtype=np.int64)
    for k, f in enumerate(LOCAL_FACES):
        tri[k::4] = en[:, list(f)]
    srt = np.sort(tri, axis=1)
    first, inv = _unique_rows(srt)
    return inv.reshape(ne, 4), tri[first]


def faces_edges(elems_faces, elems_edges, n_faces):
    ef = np.asarray(elems_faces)
    ee = np.asarray(elems_edges)
    owner = np.full(n_faces, -1, dtype=np.int64)
    owner_local = np.zeros(n_faces, dtype=np.int64)
    # walk elements in descending order so the lowest element index wins
    for e in range(ef.shape[0] - 1, -1, -1):
        for k in range(4):
            owner[ef[e, k]] = e
            owner_local[ef[e, k]] = k
    fe = np.empty((n_faces, 3), dtype=np.int64)
    for f in range(n_faces):
        e, k = owner[f], owner_local[f]
        fe[f] = ee[e, list(FACE_EDGE_TRIPLES[k])]
    return fe


def dof_counts(p):
    return p, p * (p - 1), p * (p - 1) * (p - 2) // 2


def dof_connectivity(elems_edges, elems_faces, p):
    ee = np.asarray(elems_edges, dtype=np.int64)
    ef = np.asarray(elems_faces, dtype=np.int64)
    ne = ee.shape[0]
    n_edges = int(ee.max()) + 1
    n_faces = int(ef.max()) + 1
    de, df, dv = dof_counts(p)
    dof_edges = np.arange(n_edges * de, dtype=np.int64).reshape(n_edges, de)
    off = n_edges * de
    dof_faces = off + np.arange(n_faces * df, dtype=np.int64).reshape(n_faces, df)
    off += n_faces * df
    dof_vol = off + np.arange(ne * dv, dtype=np.int64).reshape(ne, dv)
    off += ne * dv
    rows = []
    for e in range(ne):
        r = []
        for j in range(6):
            r.extend(dof_edges[ee[e, j]])
        for j in range(4):
            r.extend(dof_faces[ef[e, j]])
        r.extend(dof_vol[e])
        rows.append(r)
    dofs = np.array(rows, dtype=np.int64).reshape(ne, -1)
    return dofs, dof_edges, dof_faces, off


def orientations(elems_nodes, elems_edges, edges_nodes, elem_faces_edges):
    """Edge flags (E,6) and face codes (E,4).

    ``elem_faces_edges`` is (E,4,3): for each element face, the owner's edge triple
    (the ``facesEdges`` scratch array, ``preprocessing.py:191-194``)."""
    en = np.asarray(elems_nodes)
    ee = np.asarray(elems_edges)
    edn = np.asarray(edges_nodes)
    fe = np.asarray(elem_faces_edges)
    ne = en.shape[0]
    eo = np.zeros((ne, 6), dtype=np.int64)
    fo = np.zeros((ne, 4), dtype=np.int64)
    for e in range(ne):
        for i, (a, b) in enumerate(LOCAL_EDGES):
            g = edn[ee[e, i]]
            if en[e, a] == g[1] and en[e, b] == g[0]:
                eo[e, i] = 1
        for i, tr in enumerate(FACE_EDGE_TRIPLES):
            ge = ee[e, list(tr)]
            k1 = k2 = 0
            for k in range(3):
                if ge[0] == fe[e, i, k]:
                    k1 = k + 1
                if ge[1] == fe[e, i, k]:
                    k2 = k + 1
            fo[e, i] = FACE_CODE.get(k1 * 10 + k2, 0)
    return eo, fo


def boundary_faces(elems_faces, n_faces):
    cnt = np.bincount(np.asarray(elems_faces).ravel(), minlength=n_faces)
    return np.nonzero(cnt == 1)[0]


def boundary_dofs(elems_nodes, elems_faces, faces_nodes, edges_nodes, dof_edges, dof_faces):
    n_faces = faces_nodes.shape[0]
    bf = boundary_faces(elems_faces, n_faces)
    lookup = {tuple(r): i for i, r in enumerate(np.asarray(edges_nodes).tolist())}
    be = set()
    for f in bf:
        a, b, c = faces_nodes[f]
        for u, v in ((a, b), (b, c), (c, a)):
            be.add(lookup[(min(u, v), max(u, v))])
    be = np.array(sorted(be), dtype=np.int64)
    parts = [dof_edges[be].ravel()]
    if dof_faces.shape[1] > 0:
        parts.append(dof_faces[bf].ravel())
    return np.sort(np.concatenate(parts))


def build(elems_nodes, p):
    """All scratch-level arrays for a mesh, in the reference's layout."""
    ee, edn = edges(elems_nodes)
    ef, fn = faces(elems_nodes)
    fe = faces_edges(ef, ee, fn.shape[0])
    dofs, dof_e, dof_f, ndof = dof_connectivity(ee, ef, p)
    efe = fe[ef]                       # (E,4,3) facesEdges per element
    eo, fo = orientations(elems_nodes, ee, edn, efe)
    bd = boundary_dofs(elems_nodes, ef, fn, edn, dof_e, dof_f)
    return dict(elems_edges=ee, edges_nodes=edn, elems_faces=ef, faces_nodes=fn,
                faces_edges=fe, elem_faces_edges=efe, dofs=dofs, ndof=ndof,
                edge_orient=eo, face_orient=fo, boundary_dofs=bd)
