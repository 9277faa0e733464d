"""Hierarchical Nedelec (first kind) H(curl) basis on the unit tetrahedron -- oracle.

Restatement of ``shape3DETet`` and its helpers, ``petgem/hvfem.py:319-1052``
(Fuentes, Keith, Demkowicz & Nagaraj 2015, "Orientation embedded high order
shape functions for the exact sequence elements of all shapes"), evaluated
for many points at once (leading axis = point).  Only uniform order
p = 1..3 is supported (the reference accepts up to 6).

DOF ordering of the element (``hvfem.py:372-453``):
  * 6 edges x p functions, edge local pairs (0,1),(1,2),(0,2),(0,3),(1,3),(2,3),
    Legendre degree ascending;
  * 4 faces x p(p-1) functions, faces (0,1,2),(0,1,3),(1,2,3),(0,2,3), two
    families interleaved (family 0 at even, family 1 at odd offsets);
  * p(p-1)(p-2)/2 interior ("bubble") functions, three families interleaved.

Orientation: edge flag o in {0,1} swaps the edge's two affine coordinates
(``OrientE`` ``hvfem.py:791-822``); face code o in 0..5 permutes the face's
three coordinates (``OrientTri`` ``hvfem.py:825-878``).

Pinning: p=1 against the closed-form Whitney functions; for every p the
span is checked (exact sequence: curl of gradients vanish, dimension and
rank of the curl map), and the specific hierarchical functions (the part
that only the reference can pin) against golden vectors written from the
reference's own ``hvfem.py`` (``tests/golden/make_hvfem_vectors.py``).
"""
import numpy as np

EDGE_VERTS = ((0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3))
FACE_VERTS = ((0, 1, 2), (0, 1, 3), (1, 2, 3), (0, 2, 3))
# OrientTri permutations, hvfem.py:849-872
FACE_PERMS = ((0, 1, 2), (1, 2, 0), (2, 0, 1), (0, 2, 1), (1, 0, 2), (2, 1, 0))


def ndof_element(p):
    return p * (p + 2) * (p + 3) // 2


def _cross(a, b):
    return np.cross(a, b)


def _legendre(x, t, nord):
    """Shifted scaled Legendre P_i(x; t), i = 0..nord-1 (``PolyLegendre`` hvfem.py:629-664).

    P_0 = 1, P_1 = 2x - t, (i+1) P_{i+1} = (2i+1)(2x-t) P_i - i t^2 P_{i-1}.
    Returns a list of arrays.
    """
    out = [np.ones_like(x)]
    if nord >= 2:
        y = 2.0 * x - t
        out.append(y)
        tt = t * t
        for i in range(1, nord - 1):
            nxt = (2.0 * i + 1.0) * y * out[i] - float(i) * tt * out[i - 1]
            out.append(nxt / float(i + 1))
    return out


def _jacobi(x, t, nord, minalpha):
    """Shifted scaled Jacobi table P^{alpha_i}_j(x; t) (``PolyJacobi`` hvfem.py:733-788).

    Row i has alpha_i = minalpha + 2i; entries (i, j) are filled for
    j <= nord - i (plus column 1 for every row i < nord).  Returns a dict
    keyed by (i, j).
    """
    P = {}
    alpha = [minalpha + 2.0 * i for i in range(nord + 1)]
    for i in range(nord + 1):
        P[(i, 0)] = np.ones_like(x)
    if nord >= 1:
        y = 2.0 * x - t
        for i in range(nord):
            P[(i, 1)] = y + alpha[i] * x
    if nord >= 2:
        y = 2.0 * x - t
        tt = t * t
        for i in range(nord - 1):
            al = alpha[i]
            aa = al * al
            for j in range(2, nord - i + 1):
                ai = 2.0 * j * (j + al) * (2.0 * j + al - 2.0)
                bi = 2.0 * j + al - 1.0
                ci = (2.0 * j + al) * (2.0 * j + al - 2.0)
                di = 2.0 * (j + al - 1.0) * (j - 1.0) * (2.0 * j + al)
                val = bi * (ci * y + aa * t) * P[(i, j - 1)] - di * tt * P[(i, j - 2)]
                P[(i, j)] = val / ai
    return P


def _int_jacobi(x, t, nord, minalpha):
    """Integrated Jacobi L, plain Jacobi P and t-derivative part R
    (``PolyIJacobi`` hvfem.py:667-730).  Dicts keyed by (i, j-1), j >= 1."""
    pt = _jacobi(x, t, nord, minalpha)
    alpha = [minalpha + 2.0 * i for i in range(nord)]
    zero = np.zeros_like(x)
    L, P, R = {}, {}, {}
    for i in range(nord):
        for j in range(nord):
            P[(i, j)] = pt.get((i, j), zero)
            L[(i, j)] = zero
            R[(i, j)] = zero
        L[(i, 0)] = x
    if nord >= 2:
        tt = t * t
        for i in range(nord - 1):
            al = alpha[i]
            for j in range(2, nord - i + 1):
                tia = 2.0 * j + al
                tiam1 = tia - 1.0
                tiam2 = tia - 2.0
                ai = (j + al) / (tiam1 * tia)
                bi = al / (tiam2 * tia)
                ci = (j - 1.0) / (tiam2 * tiam1)
                L[(i, j - 1)] = ai * pt[(i, j)] + bi * t * pt[(i, j - 1)] - ci * tt * pt[(i, j - 2)]
                R[(i, j - 1)] = -(j - 1.0) * (pt[(i, j - 1)] + t * pt[(i, j - 2)]) / tiam2
    return L, P, R


def _hom_int_jacobi(s, ds, nord, minalpha, idec):
    """Homogenised integrated Jacobi and gradients (``HomIJacobi`` hvfem.py:583-626).

    s = (s0, s1) arrays (P,), ds = (ds0, ds1) arrays (P,3).
    idec=True means s0 + s1 == 1 (then t = 1 and only the x-derivative).
    """
    if idec:
        L, P, _ = _int_jacobi(s[1], np.ones_like(s[1]), nord, minalpha)
        D = {}
        for i in range(nord):
            for j in range(nord - i):
                D[(i, j)] = P[(i, j)][:, None] * ds[1]
    else:
        L, P, R = _int_jacobi(s[1], s[0] + s[1], nord, minalpha)
        ds01 = ds[0] + ds[1]
        D = {}
        for i in range(nord):
            for j in range(nord - i):
                D[(i, j)] = P[(i, j)][:, None] * ds[1] + R[(i, j)][:, None] * ds01
    return L, D


def _anc_edge(s, ds, nord):
    """Edge ancillary functions E_i = P_i(s1; s0+s1) (s0 grad s1 - s1 grad s0),
    curl E_i = (i+2) P_i grad s0 x grad s1, i = 0..nord-1 (``AncEE`` hvfem.py:467-513,
    general case Idec = False)."""
    hp = _legendre(s[1], s[0] + s[1], nord)
    whi = s[0][:, None] * ds[1] - s[1][:, None] * ds[0]
    cwhi = _cross(ds[0], ds[1])
    E = [hp[i][:, None] * whi for i in range(nord)]
    C = [(i + 2.0) * hp[i][:, None] * cwhi for i in range(nord)]
    return E, C


def _anc_tri(s, ds, nord):
    """Triangle ancillary functions ETri[(j, k-1)] for j >= 0, k >= 1, j + k <= nord-1
    (``AncETri`` hvfem.py:516-566, Idec = False)."""
    E, C = _anc_edge(s[0:2], ds[0:2], nord - 1)
    sl = (s[0] + s[1], s[2])
    dsl = (ds[0] + ds[1], ds[2])
    L, D = _hom_int_jacobi(sl, dsl, nord - 1, 1, False)
    ET, CT = {}, {}
    for i in range(nord):
        for j in range(0, i):
            k = i - j
            ET[(j, k - 1)] = E[j] * L[(j, k - 1)][:, None]
            CT[(j, k - 1)] = L[(j, k - 1)][:, None] * C[j] + _cross(D[(j, k - 1)], E[j])
    return ET, CT


def affine_coords(X):
    """Barycentric coordinates and their (constant) gradients (``AffineTetrahedron``)."""
    X = np.atleast_2d(np.asarray(X, dtype=np.float64))
    npt = X.shape[0]
    lam = [1.0 - X[:, 0] - X[:, 1] - X[:, 2], X[:, 0].copy(), X[:, 1].copy(), X[:, 2].copy()]
    g = np.array([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    dlam = [np.broadcast_to(g[a], (npt, 3)).copy() for a in range(4)]
    return lam, dlam


def edge_functions(X, edge, orient, p):
    """The p functions of local edge ``edge`` with orientation flag ``orient``.
    Returns (shape, curl) arrays of shape (P, p, 3)."""
    lam, dlam = affine_coords(X)
    a, b = EDGE_VERTS[edge]
    if orient:
        a, b = b, a
    E, C = _anc_edge((lam[a], lam[b]), (dlam[a], dlam[b]), p)
    return np.stack(E, axis=1), np.stack(C, axis=1)


def face_functions(X, face, orient, p):
    """The p(p-1) functions of local face ``face`` with orientation code ``orient``
    (families interleaved).  Returns (shape, curl) of shape (P, p(p-1), 3)."""
    lam, dlam = affine_coords(X)
    npt = lam[0].shape[0]
    nf = p * (p - 1)
    S = np.zeros((npt, nf, 3))
    Cc = np.zeros((npt, nf, 3))
    if nf == 0:
        return S, Cc
    verts = FACE_VERTS[face]
    perm = FACE_PERMS[orient]
    g = [verts[perm[0]], verts[perm[1]], verts[perm[2]]]
    for fam in range(2):
        abc = [g[(fam + 0) % 3], g[(fam + 1) % 3], g[(fam + 2) % 3]]
        ET, CT = _anc_tri([lam[v] for v in abc], [dlam[v] for v in abc], p)
        m = fam
        for k in range(1, p):
            for r in range(0, k):
                pp = k - r
                S[:, m] = ET[(r, pp - 1)]
                Cc[:, m] = CT[(r, pp - 1)]
                m += 2
    return S, Cc


def bubble_functions(X, p):
    """The p(p-1)(p-2)/2 interior functions (three families interleaved)."""
    lam, dlam = affine_coords(X)
    npt = lam[0].shape[0]
    nb = p * (p - 1) * (p - 2) // 2
    S = np.zeros((npt, nb, 3))
    Cc = np.zeros((npt, nb, 3))
    if nb == 0:
        return S, Cc
    for fam in range(3):
        abcd = [(fam + 0) % 4, (fam + 1) % 4, (fam + 2) % 4, (fam + 3) % 4]
        abc, d = abcd[0:3], abcd[3]
        ET, CT = _anc_tri([lam[v] for v in abc], [dlam[v] for v in abc], p - 1)
        Lb, Db = _hom_int_jacobi((1.0 - lam[d], lam[d]), (-dlam[d], dlam[d]), p - 2, 2, True)
        m = fam
        for j in range(2, p):
            for k in range(1, j):
                for r in range(0, k):
                    pp = k - r
                    q = j - k
                    hb = Lb[(k - 1, q - 1)][:, None]
                    S[:, m] = ET[(r, pp - 1)] * hb
                    Cc[:, m] = hb * CT[(r, pp - 1)] + _cross(Db[(k - 1, q - 1)], ET[(r, pp - 1)])
                    m += 3
    return S, Cc


def shape_functions(X, edge_orient, face_orient, p):
    """Reference-element shape functions and curls for one element.

    ``X``: (P,3) points in the unit tetrahedron; ``edge_orient`` (6,),
    ``face_orient`` (4,).  Returns (shape, curl), each (P, n, 3) with
    n = p(p+2)(p+3)/2, in the reference DOF order.
    """
    X = np.atleast_2d(np.asarray(X, dtype=np.float64))
    parts_s, parts_c = [], []
    for e in range(6):
        s, c = edge_functions(X, e, int(edge_orient[e]), p)
        parts_s.append(s)
        parts_c.append(c)
    for f in range(4):
        s, c = face_functions(X, f, int(face_orient[f]), p)
        parts_s.append(s)
        parts_c.append(c)
    s, c = bubble_functions(X, p)
    parts_s.append(s)
    parts_c.append(c)
    S = np.concatenate(parts_s, axis=1)
    C = np.concatenate(parts_c, axis=1)
    assert S.shape[1] == ndof_element(p)
    return S, C
