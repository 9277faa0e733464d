"""Element matrices, global assembly, BC, source, receivers and direct solve -- oracle.

* ``jacobian`` -- ``petgem/hvfem.py:101-119`` (rows x_k - x_0; signed det as
  in ``hvfem.py:265``).
* ``element_matrices`` -- ``petgem/hvfem.py:223-316``: per quadrature point,
  N_real = J^-1 N_ref, M += w (N_real)^T diag(sh,sh,sv) N_real,
  curl_real = J^T c / det, K += w curl_real . curl_real, then both x det.
* ``assemble`` -- ``petgem/solver.py:191-235``: A_e = K_e - i w mu0 M_e added into
  the union-of-cliques pattern (explicit zeros kept), mu0 = 4 pi 1e-7.
* ``apply_dirichlet`` -- ``petgem/solver.py:552-577``: zero rows and columns of the
  boundary DOFs, unit diagonal, b = 0 there (pattern kept).
* ``source_vector`` -- ``petgem/solver.py:247-316`` with
  ``computeSourceVectorRotation`` ``hvfem.py:2303-2344``,
  ``tetrahedronXYZToXiEtaZeta`` ``hvfem.py:2347-2490`` and
  ``computeBasisFunctions`` ``hvfem.py:2493-2553``.
* ``locate`` -- the ``Delaunay.find_simplex(bruteforce=True, tol=1e-12)`` rule of
  ``preprocessing.py:414-420`` / ``postprocessing.py:532-539``: lowest-index
  element whose barycentric coordinates are all within [-tol, 1+tol].
* ``receiver_fields`` -- ``petgem/postprocessing.py:479-616``.
* ``solve_direct`` -- SuperLU (the exact FP64 solution of the same system).

Pinning: p=1 element matrices against closed-form Whitney integrals; p=1..3
against golden vectors from the reference's ``computeElementalMatrices``;
assembly pattern against the closed-form nnz counts; the whole pipeline
against SuperLU residuals and the analytic behaviour of the fields.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import basis as B
from . import quadrature as Q

MU0 = 4.0 * np.pi * 1e-7


def jacobian(nodes):
    J = np.asarray(nodes[1:4] - nodes[0], dtype=np.float64)
    return J, np.linalg.inv(J), np.linalg.det(J)


def element_matrices(nodes, edge_orient, face_orient, sigma, p):
    """Reference-form M_e, K_e (n x n) for one element."""
    J, iJ, det = jacobian(np.asarray(nodes, dtype=np.float64))
    pts, w = Q.rule_for_order(p)
    S, C = B.shape_functions(pts, edge_orient, face_orient, p)     # (Q,n,3)
    er = np.array([sigma[0], sigma[0], sigma[1]], dtype=np.float64)
    n = S.shape[1]
    M = np.zeros((n, n))
    K = np.zeros((n, n))
    for q in range(pts.shape[0]):
        Nr = S[q] @ iJ.T                 # rows: (J^-1 N_ref)^T
        Cr = (C[q] @ J) / det            # rows: c^T J / det
        M += w[q] * (Nr * er) @ Nr.T
        K += w[q] * Cr @ Cr.T
    return M * det, K * det


def all_element_matrices(coords, edge_orient, face_orient, sigma, p):
    """M, K for every element (E,n,n); basis cached per orientation class."""
    coords = np.asarray(coords, dtype=np.float64)
    ne = coords.shape[0]
    pts, w = Q.rule_for_order(p)
    n = B.ndof_element(p)
    cache = {}
    M = np.empty((ne, n, n))
    K = np.empty((ne, n, n))
    J = coords[:, 1:4, :] - coords[:, 0:1, :]
    iJ = np.linalg.inv(J)
    det = np.linalg.det(J)
    keys = [tuple(a) + tuple(b) for a, b in zip(np.asarray(edge_orient).tolist(),
                                                   np.asarray(face_orient).tolist())]
    for e in range(ne):
        k = keys[e]
        if k not in cache:
            cache[k] = B.shape_functions(pts, k[:6], k[6:], p)
        S, C = cache[k]
        er = np.array([sigma[e][0], sigma[e][0], sigma[e][1]])
        Nr = np.einsum("qia,ba->qib", S, iJ[e])
        Cr = np.einsum("qia,ab->qib", C, J[e]) / det[e]
        M[e] = np.einsum("q,qia,qja->ij", w, Nr * er, Nr) * det[e]
        K[e] = np.einsum("q,qia,qja->ij", w, Cr, Cr) * det[e]
    return M, K


def csr_pattern(dofs, ndof):
    """Sorted-column CSR pattern of the union of element DOF cliques."""
    dofs = np.asarray(dofs, dtype=np.int64)
    n = dofs.shape[1]
    r = np.repeat(dofs, n, axis=1).ravel()
    c = np.tile(dofs, (1, n)).ravel()
    key = np.unique(r * ndof + c)
    rows, cols = key // ndof, key % ndof
    indptr = np.zeros(ndof + 1, dtype=np.int64)
    np.add.at(indptr, rows + 1, 1)
    return np.cumsum(indptr), cols


def assemble(dofs, M, K, frequency, ndof):
    """Global complex A = sum_e (K_e - i w mu0 M_e) on the clique pattern."""
    omega = 2.0 * np.pi * frequency
    const = 1j * omega * MU0
    dofs = np.asarray(dofs, dtype=np.int64)
    n = dofs.shape[1]
    Ae = K - const * M
    r = np.repeat(dofs, n, axis=1).ravel()
    c = np.tile(dofs, (1, n)).ravel()
    A = sp.coo_matrix((Ae.ravel(), (r, c)), shape=(ndof, ndof)).tocsr()
    A.sort_indices()
    return A


def apply_dirichlet(A, b, bdofs):
    A = A.tocsr(copy=True)
    mask = np.zeros(A.shape[0], dtype=bool)
    mask[np.asarray(bdofs)] = True
    rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
    kill = mask[rows] | mask[A.indices]
    A.data[kill] = 0.0
    diag = kill & (rows == A.indices) & mask[rows]
    A.data[diag] = 1.0
    b = np.array(b, dtype=np.complex128, copy=True)
    b[mask] = 0.0
    return A, b


def source_rotation(azimuth, dip):
    """Unit dipole direction R(az) R(dip) e_x (``hvfem.py:2303-2344``)."""
    a = np.deg2rad(azimuth)
    d = np.deg2rad(dip)
    return np.array([np.cos(a) * np.cos(d), np.sin(a) * np.cos(d), np.sin(d)])


def xyz_to_reference(nodes, x):
    """Reference coordinates (xi, eta, zeta) of physical points x (P,3)."""
    J, iJ, _ = jacobian(np.asarray(nodes, dtype=np.float64))
    return (np.atleast_2d(x) - nodes[0]) @ iJ


def physical_basis(nodes, edge_orient, face_orient, p, xi):
    """N_real (P,n,3) and curl_real (P,n,3) at reference points ``xi``."""
    J, iJ, det = jacobian(np.asarray(nodes, dtype=np.float64))
    S, C = B.shape_functions(xi, edge_orient, face_orient, p)
    return S @ iJ.T, (C @ J) / det


def locate(coords, pts, tol=1e-12):
    """Lowest element index containing each point (barycentric tolerance), -1 if none."""
    coords = np.asarray(coords, dtype=np.float64)
    pts = np.atleast_2d(pts)
    T = coords[:, 0:3, :] - coords[:, 3:4, :]          # rows v_i - v_3
    Ti = np.linalg.inv(np.transpose(T, (0, 2, 1)))     # maps x - v3 -> (c0,c1,c2)
    out = np.full(pts.shape[0], -1, dtype=np.int64)
    for i, x in enumerate(pts):
        c = np.einsum("eij,ej->ei", Ti, x[None, :] - coords[:, 3, :])
        c3 = 1.0 - c.sum(axis=1)
        cc = np.concatenate([c, c3[:, None]], axis=1)
        inside = np.all((cc >= -tol) & (cc <= 1.0 + tol), axis=1)
        idx = np.nonzero(inside)[0]
        if idx.size:
            out[i] = idx[0]
    return out


def source_vector(coords, dofs, edge_orient, face_orient, p, ndof, source, element):
    """Point-dipole RHS b[dofs_e] += i w mu0 (I L R) . N_real(x_s)."""
    omega = 2.0 * np.pi * source["frequency"]
    const = 1j * omega * MU0
    rot = source_rotation(source["azimuth"], source["dip"])
    field = source["current"] * source["length"] * rot
    nodes = coords[element]
    xi = xyz_to_reference(nodes, np.asarray(source["position"], dtype=np.float64))
    Nr, _ = physical_basis(nodes, edge_orient[element], face_orient[element], p, xi)
    b = np.zeros(ndof, dtype=np.complex128)
    np.add.at(b, dofs[element], const * (Nr[0] @ field))
    return b


def receiver_fields(x, coords, dofs, edge_orient, face_orient, p, frequency, receivers, elements):
    """(nRx,6) complex: E = sum x_j N_j, H = sum x_j curl N_j / (i w mu0)."""
    omega = 2.0 * np.pi * frequency
    const = 1j * omega * MU0
    out = np.zeros((len(elements), 6), dtype=np.complex128)
    for i, e in enumerate(elements):
        nodes = coords[e]
        xi = xyz_to_reference(nodes, receivers[i])
        Nr, Cr = physical_basis(nodes, edge_orient[e], face_orient[e], p, xi)
        xe = x[dofs[e]]
        out[i, 0:3] = xe @ Nr[0]
        out[i, 3:6] = (xe @ Cr[0]) / const
    return out


def solve_direct(A, b):
    lu = spla.splu(A.tocsc())
    return lu.solve(b)


def relres(A, x, b):
    return np.linalg.norm(b - A @ x) / np.linalg.norm(b)


def forward(topo_like, coords, sigma, p, source, receivers):
    """The full reference pipeline on CPU; returns a dict of intermediate results.

    ``topo_like`` provides ``dofs``, ``ndof``, ``boundary_dofs``,
    ``edge_orient``, ``face_orient``.
    """
    eo, fo = topo_like["edge_orient"], topo_like["face_orient"]
    dofs, ndof = topo_like["dofs"], topo_like["ndof"]
    M, K = all_element_matrices(coords, eo, fo, sigma, p)
    A = assemble(dofs, M, K, source["frequency"], ndof)
    src_el = int(locate(coords, np.asarray(source["position"])[None, :])[0])
    if src_el < 0:
        raise ValueError("source outside the mesh")
    b = source_vector(coords, dofs, eo, fo, p, ndof, source, src_el)
    Abc, bbc = apply_dirichlet(A, b, topo_like["boundary_dofs"])
    x = solve_direct(Abc, bbc)
    rx_el = locate(coords, receivers)
    keep = rx_el >= 0
    f = receiver_fields(x, coords, dofs, eo, fo, p, source["frequency"], receivers[keep], rx_el[keep])
    return dict(M=M, K=K, A=A, b=b, Abc=Abc, bbc=bbc, x=x, src_el=src_el,
                rx_el=rx_el, fields=f, relres=relres(Abc, x, bbc))
