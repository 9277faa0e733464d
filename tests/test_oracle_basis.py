"""Pinning of the oracle's quadrature and hierarchical basis (CPU).

* quadrature: exactness on monomials up to the rule order (closed form
  a! b! c! / (a+b+c+3)!) and the reference's own mapped tables;
* basis: golden vectors written from the reference's ``shape3DETet``
  (``tests/golden/make_hvfem_vectors.py``), the closed-form Whitney
  functions at p = 1, and structural invariants at every p (the exact
  sequence: curls of gradient fields vanish, rank of the curl map).
A dropped term, a sign error or a swapped family in the restatement fails
at least one of these.
"""
from math import factorial

import numpy as np
import pytest

from oracle import basis as B
from oracle import quadrature as Q


@pytest.mark.parametrize("order", [2, 4, 6])
def test_quadrature_exact_on_monomials(order):
    pts, w = Q.gauss_points(order)
    for a in range(order + 1):
        for b in range(order + 1 - a):
            for c in range(order + 1 - a - b):
                exact = factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + 3)
                got = np.sum(w * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c)
                assert abs(got - exact) <= 2e-14, (a, b, c, got, exact)


@pytest.mark.parametrize("order", [2, 4, 6])
def test_quadrature_matches_reference_tables(golden, order):
    pts, w = Q.gauss_points(order)
    assert np.array_equal(pts, golden[f"gauss{order}_points"])
    assert np.array_equal(w, golden[f"gauss{order}_weights"])


@pytest.mark.parametrize("p", [1, 2, 3])
def test_basis_matches_reference_golden(golden, p):
    X = golden["shape_points"]
    for a in range(golden["shape_edge_orient"].shape[0]):
        S, C = B.shape_functions(X, golden["shape_edge_orient"][a], golden["shape_face_orient"][a], p)
        assert np.abs(S - golden[f"shape_p{p}"][a]).max() <= 1e-14
        assert np.abs(C - golden[f"curl_p{p}"][a]).max() <= 1e-14


def test_p1_is_whitney():
    """p = 1: N_e = l_a grad l_b - l_b grad l_a, curl = 2 grad l_a x grad l_b."""
    rng = np.random.default_rng(0)
    X = rng.dirichlet(np.ones(4), size=20)[:, 1:]
    lam = np.stack([1 - X.sum(1), X[:, 0], X[:, 1], X[:, 2]], axis=1)
    g = np.array([[-1.0, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    for flip in (0, 1):
        S, C = B.shape_functions(X, [flip] * 6, [0] * 4, 1)
        for e, (a, b) in enumerate(B.EDGE_VERTS):
            if flip:
                a, b = b, a
            whit = lam[:, a:a + 1] * g[b] - lam[:, b:b + 1] * g[a]
            assert np.allclose(S[:, e], whit, atol=1e-15)
            assert np.allclose(C[:, e], 2 * np.cross(g[a], g[b]), atol=1e-15)


def _mass_and_curl_gram(p, eo, fo):
    pts, w = Q.rule_for_order(p)
    S, C = B.shape_functions(pts, eo, fo, p)
    M = np.einsum("q,qia,qja->ij", w, S, S)
    K = np.einsum("q,qia,qja->ij", w, C, C)
    return M, K


@pytest.mark.parametrize("p", [1, 2, 3])
def test_exact_sequence_rank(p):
    """dim N_p = p(p+2)(p+3)/2; kernel of curl = gradients of P_p minus constants."""
    n = B.ndof_element(p)
    M, K = _mass_and_curl_gram(p, [0, 1, 0, 1, 1, 0], [0, 3, 5, 1])
    assert np.linalg.matrix_rank(M, tol=1e-12 * np.abs(M).max()) == n  # independent functions
    grad_dim = (p + 1) * (p + 2) * (p + 3) // 6 - 1
    assert np.linalg.matrix_rank(K, tol=1e-10 * np.abs(K).max()) == n - grad_dim


@pytest.mark.parametrize("p", [2, 3])
def test_orientation_permutes_span_only(p):
    """Changing an orientation changes the basis but not the space: the Gram
    matrices of the two bases have the same eigenvalues of M^-1 K."""
    M1, K1 = _mass_and_curl_gram(p, [0] * 6, [0] * 4)
    M2, K2 = _mass_and_curl_gram(p, [1, 0, 1, 0, 1, 1], [2, 4, 1, 5])
    e1 = np.sort(np.linalg.eigvals(np.linalg.solve(M1, K1)).real)
    e2 = np.sort(np.linalg.eigvals(np.linalg.solve(M2, K2)).real)
    assert np.allclose(e1, e2, atol=1e-9 * e1.max())
