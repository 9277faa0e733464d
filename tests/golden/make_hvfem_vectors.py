"""Write golden input/output vectors from the reference's own ``petgem/hvfem.py``.

Run once, in the development container only (the reference repository is
not available where the tests run):

    python tests/golden/make_hvfem_vectors.py /root/reference

Why an import is needed: the hierarchical high-order basis (family
interleaving, signs, orientation embedding of ``shape3DETet``,
``hvfem.py:319-1052``) is not fixed by any closed form or invariant -- any
basis of the same Nedelec space passes those -- so for p >= 2 only the
reference itself can pin the element matrices and the basis values.

``hvfem.py`` needs only numpy and ``Print`` from the package's ``common``
module (which would pull in PETSc/MPI), so the module is loaded from its
file with a stand-in ``petgem.common`` that provides ``Print``; the removed
numpy aliases ``np.int`` / ``np.float`` / ``np.complex`` used by the
reference are mapped to the Python builtins for the duration of this script.
Nothing from the reference is copied into the repository: only the numeric
vectors are saved, to ``tests/golden/hvfem_vectors.npz``.
"""
import importlib.util
import os
import sys
import types

import numpy as np


def load_hvfem(ref_root):
    for name, val in (("int", int), ("float", float), ("complex", complex)):
        if not hasattr(np, name):
            setattr(np, name, val)
    pkg = types.ModuleType("petgem")
    pkg.__path__ = []
    common = types.ModuleType("petgem.common")

    class Print:
        @staticmethod
        def master(*a, **k):
            raise RuntimeError("reference reported an error: %r" % (a,))

    common.Print = Print
    sys.modules["petgem"] = pkg
    sys.modules["petgem.common"] = common
    path = os.path.join(ref_root, "petgem", "hvfem.py")
    spec = importlib.util.spec_from_file_location("petgem.hvfem", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def main(ref_root):
    hv = load_hvfem(ref_root)
    rng = np.random.default_rng(20261202)
    out = {}

    # quadrature tables as the reference maps them
    for order in (2, 4, 6):
        pts, w = hv.compute3DGaussPoints(order)
        out[f"gauss{order}_points"] = pts
        out[f"gauss{order}_weights"] = w

    # reference-element shape functions at random interior points, for
    # every edge flag and every face code (edge e uses flag bit, face f code)
    X = rng.dirichlet(np.ones(4), size=6)[:, 1:4]
    out["shape_points"] = X
    orient_sets = []
    for k in range(6):
        eo = np.array([(k >> b) & 1 for b in range(6)], dtype=int)
        eo = (eo + np.arange(6) * k) % 2
        fo = np.array([(k + f) % 6 for f in range(4)], dtype=int)
        orient_sets.append((eo, fo))
    out["shape_edge_orient"] = np.array([o[0] for o in orient_sets])
    out["shape_face_orient"] = np.array([o[1] for o in orient_sets])
    for p in (1, 2, 3):
        n = p * (p + 2) * (p + 3) // 2
        S = np.zeros((len(orient_sets), X.shape[0], n, 3))
        C = np.zeros_like(S)
        for a, (eo, fo) in enumerate(orient_sets):
            for q in range(X.shape[0]):
                nr, sh, cu = hv.shape3DETet(X[q], np.ones(11, dtype=int) * p, eo, fo)
                assert nr == n
                S[a, q] = sh[:, :n].T
                C[a, q] = cu[:, :n].T
        out[f"shape_p{p}"] = S
        out[f"curl_p{p}"] = C

    # element matrices on elements of random shape with random orientations
    ne = 5
    nodes = rng.uniform(-500.0, 500.0, size=(ne, 4, 3))
    for e in range(ne):
        J = nodes[e, 1:] - nodes[e, 0]
        if np.linalg.det(J) < 0:
            nodes[e, [2, 3]] = nodes[e, [3, 2]]
    sig = rng.uniform(0.01, 3.3, size=(ne, 2))
    eo = rng.integers(0, 2, size=(ne, 6))
    fo = rng.integers(0, 6, size=(ne, 4))
    out["elem_nodes"] = nodes
    out["elem_sigma"] = sig
    out["elem_edge_orient"] = eo
    out["elem_face_orient"] = fo
    for p in (1, 2, 3):
        Ms, Ks = [], []
        for e in range(ne):
            jac, ijac = hv.computeJacobian(nodes[e])
            M, K = hv.computeElementalMatrices(eo[e], fo[e], jac, ijac, p, sig[e])
            Ms.append(M)
            Ks.append(K)
        out[f"elem_M_p{p}"] = np.array(Ms)
        out[f"elem_K_p{p}"] = np.array(Ks)

    # orientation from connectivity (computeElementOrientation): elements of the
    # reference test mesh are produced by the tests themselves; here a synthetic
    # permutation set covering every face code
    en = np.array([[0, 1, 2, 3], [3, 1, 0, 2], [2, 3, 1, 0]])
    out["orient_nodes"] = en

    # physical-space basis at points and the xyz -> reference map
    pts = []
    for e in range(ne):
        lam = rng.dirichlet(np.ones(4), size=3)
        pts.append(lam @ nodes[e])
    pts = np.array(pts)
    out["basis_points"] = pts
    xi_all = np.zeros((ne, 3, 3))
    for p in (1, 2, 3):
        B, CB = [], []
        for e in range(ne):
            jac, ijac = hv.computeJacobian(nodes[e])
            xi = hv.tetrahedronXYZToXiEtaZeta(nodes[e], pts[e])
            xi_all[e] = xi
            b, cb = hv.computeBasisFunctions(eo[e], fo[e], jac, ijac, p, xi)
            B.append(b)
            CB.append(cb)
        out[f"basis_p{p}"] = np.array(B)
        out[f"curlbasis_p{p}"] = np.array(CB)
    out["basis_xi"] = xi_all

    rot = []
    angles = np.array([[0.0, 0.0], [30.0, 0.0], [0.0, 45.0], [110.0, -20.0]])
    for az, dip in angles:
        rot.append(hv.computeSourceVectorRotation(az, dip))
    out["rot_angles"] = angles
    out["rot_vectors"] = np.array(rot)

    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hvfem_vectors.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, sum(v.nbytes for v in out.values()), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
