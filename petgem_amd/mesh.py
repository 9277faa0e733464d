"""Mesh input and topology for the CSEM hot path (host side, vectorised numpy).

Produces exactly the arrays the reference's preprocessing writes into its
scratch directory (``petgem/preprocessing.py:75-312``), so that the device
path sees the same inputs the reference ``Solver`` reads:

=====================  ==========================================  ==========================
array                  meaning                                     reference
=====================  ==========================================  ==========================
``elems_nodes``        (E,4) node ids, file order                  ``preprocessing.py:105-118``
``coords``             (E,4,3) node coordinates per element        ``preprocessing.py:84-101``
``elems_edges``        (E,6) edge ids = rank of sorted node pair   ``mesh.py:18-53``
``edges_nodes``        (nEdges,2) sorted node pairs                ``mesh.py:47-50``
``elems_faces``        (E,4) face ids = rank of sorted triple      ``mesh.py:56-89``
``faces_edges``        (nFaces,3) owner's local face-edge triple   ``mesh.py:92-125``
``dofs``               (E,n) global DOFs per element               ``hvfem.py:15-98``
``boundary_dofs``      PEC DOFs (edges+faces on boundary faces)    ``mesh.py:149-321``
=====================  ==========================================  ==========================

The numbering rules are the reference's; the implementation is a sort-based
one (int64 keys + ``np.unique``) that handles millions of elements in
seconds.  Owner of a face = lowest element index containing it, which is
the first occurrence in element-major order.
"""
import numpy as np

LOCAL_EDGES = np.array([(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)], dtype=np.int64)
LOCAL_FACES = np.array([(0, 1, 2), (0, 1, 3), (1, 2, 3), (0, 2, 3)], dtype=np.int64)
FACE_EDGE_TRIPLES = np.array([(0, 1, 2), (0, 4, 3), (1, 5, 4), (2, 5, 3)], dtype=np.int64)


class MeshError(ValueError):
    pass


def read_gmsh_v2(path):
    """Read an ASCII gmsh 2.2 file: tetrahedra (type 4) only.

    Returns (points (nNodes,3) f64, elems_nodes (E,4) int64 0-based in file
    order of the node block, physical tag (E,) int64 1-based).  Mirrors what
    the reference takes from ``meshio.read`` (``preprocessing.py:75-81``):
    the first cell block, which is the tetrahedra for the meshes it uses.
    """
    with open(path, "r") as f:
        text = f.read().split("\n")
    it = iter(range(len(text)))
    nodes_tag = None
    pts = None
    tets, phys = [], []
    i = 0
    n = len(text)
    while i < n:
        line = text[i].strip()
        if line == "$MeshFormat":
            ver = text[i + 1].split()
            if not ver[0].startswith("2"):
                raise MeshError(f"{path}: gmsh format {ver[0]} not supported (need 2.x ASCII)")
            if len(ver) > 1 and ver[1] != "0":
                raise MeshError(f"{path}: binary gmsh files are not supported")
            i += 2
        elif line == "$Nodes":
            cnt = int(text[i + 1])
            arr = np.array([text[i + 2 + k].split() for k in range(cnt)], dtype=np.float64)
            nodes_tag = arr[:, 0].astype(np.int64)
            pts = arr[:, 1:4].copy()
            i += cnt + 2
        elif line == "$Elements":
            cnt = int(text[i + 1])
            for k in range(cnt):
                t = text[i + 2 + k].split()
                if t[1] != "4":
                    continue
                ntag = int(t[2])
                phys.append(int(t[3]) if ntag > 0 else 0)
                tets.append([int(v) for v in t[3 + ntag:3 + ntag + 4]])
            i += cnt + 2
        else:
            i += 1
    if pts is None or not tets:
        raise MeshError(f"{path}: no nodes or no tetrahedra found")
    tag_to_idx = np.full(int(nodes_tag.max()) + 1, -1, dtype=np.int64)
    tag_to_idx[nodes_tag] = np.arange(nodes_tag.shape[0])
    en = tag_to_idx[np.asarray(tets, dtype=np.int64)]
    if (en < 0).any():
        raise MeshError(f"{path}: element references an undefined node")
    return pts, en, np.asarray(phys, dtype=np.int64)


def _rank_rows(keys_rows):
    """Lexicographic rank of rows (unique), first occurrence and inverse."""
    rows = np.ascontiguousarray(keys_rows)
    if rows.shape[1] <= 3 and rows.max(initial=0) < (1 << 20):
        sh = 21
        k = rows[:, 0].copy()
        for c in range(1, rows.shape[1]):
            k = (k << sh) | rows[:, c]
        _, first, inv = np.unique(k, return_index=True, return_inverse=True)
    else:
        view = rows.view([("", rows.dtype)] * rows.shape[1]).ravel()
        _, first, inv = np.unique(view, return_index=True, return_inverse=True)
    return first, inv.ravel()


def dof_counts(p):
    """DOFs per edge, face, volume and element for uniform order p (``hvfem.py:44-48``)."""
    return p, p * (p - 1), p * (p - 1) * (p - 2) // 2, p * (p + 2) * (p + 3) // 2


class Topology:
    """Numbered mesh entities and DOFs for one tetrahedral mesh and order p."""

    def __init__(self, points, elems_nodes, p):
        if p not in (1, 2, 3):
            raise MeshError(f"basis order {p} not supported (1, 2 or 3)")
        self.p = p
        self.points = np.asarray(points, dtype=np.float64)
        en = np.asarray(elems_nodes, dtype=np.int64)
        if en.ndim != 2 or en.shape[1] != 4:
            raise MeshError("elems_nodes must be (E,4)")
        self.elems_nodes = en
        ne = en.shape[0]
        self.n_elems = ne
        self.coords = self.points[en]                               # (E,4,3)

        # edges: rank of sorted node pair
        pairs = en[:, LOCAL_EDGES]                                  # (E,6,2)
        sp = np.sort(pairs.reshape(-1, 2), axis=1)
        first, inv = _rank_rows(sp)
        self.elems_edges = inv.reshape(ne, 6)
        self.edges_nodes = sp[first]
        self.n_edges = self.edges_nodes.shape[0]

        # faces: rank of sorted triple; faces_nodes keeps first occurrence order
        tri = en[:, LOCAL_FACES].reshape(-1, 3)
        st = np.sort(tri, axis=1)
        first, inv = _rank_rows(st)
        self.elems_faces = inv.reshape(ne, 4)
        self.faces_nodes = tri[first]
        self.n_faces = self.faces_nodes.shape[0]
        owner, owner_local = first // 4, first % 4
        self.faces_edges = self.elems_edges[owner[:, None], FACE_EDGE_TRIPLES[owner_local]]

        de, df, dv, n = dof_counts(p)
        self.n_dof_elem = n
        self.dof_edge0 = 0
        self.dof_face0 = self.n_edges * de
        self.dof_vol0 = self.dof_face0 + self.n_faces * df
        self.ndof = self.dof_vol0 + ne * dv
        cols = [self.elems_edges[:, j:j + 1] * de + np.arange(de) for j in range(6)]
        if df:
            cols += [self.dof_face0 + self.elems_faces[:, j:j + 1] * df + np.arange(df) for j in range(4)]
        if dv:
            cols.append(self.dof_vol0 + np.arange(ne)[:, None] * dv + np.arange(dv))
        self.dofs = np.concatenate(cols, axis=1)

        # boundary: faces with a single adjacent element, their edges, their DOFs
        cnt = np.bincount(self.elems_faces.ravel(), minlength=self.n_faces)
        self.boundary_faces = np.nonzero(cnt == 1)[0]
        bfo = first[self.boundary_faces]
        be = self.elems_edges[(bfo // 4)[:, None], FACE_EDGE_TRIPLES[bfo % 4]]
        self.boundary_edges = np.unique(be.ravel())
        parts = [(self.boundary_edges[:, None] * de + np.arange(de)).ravel()]
        if df:
            parts.append((self.dof_face0 + self.boundary_faces[:, None] * df + np.arange(df)).ravel())
        self.boundary_dofs = np.sort(np.concatenate(parts))

    # per-element views in the scratch-file layout (solver.py:193-227)
    @property
    def elem_edges_nodes(self):
        return self.edges_nodes[self.elems_edges]                   # (E,6,2)

    @property
    def elem_faces_edges(self):
        return self.faces_edges[self.elems_faces]                   # (E,4,3)

    def signed_volumes6(self):
        c = self.coords
        J = c[:, 1:4, :] - c[:, 0:1, :]
        return np.linalg.det(J)
