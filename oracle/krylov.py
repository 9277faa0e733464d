"""Krylov solvers on CPU (numpy/scipy) -- oracle and CPU baseline.

The reference solves with PETSc's KSP (``petgem/solver.py:584-590``; the
examples use ``-ksp_type gmres -pc_type sor -ksp_rtol 1e-8``,
``doc/source/_downloads/petsc.opts:2-4``).  PETSc is not available here, so
these are restatements of the textbook methods:

* ``cocg`` -- conjugate orthogonal CG for complex symmetric A = A^T
  (van der Vorst & Melissen 1990), bilinear products u^T v, left
  Jacobi preconditioner, stopping on the true-norm residual
  ||b - A x|| / ||b||.
* ``gmres`` -- restarted GMRES(m) with modified Gram-Schmidt and Givens
  rotations (Saad & Schultz 1986), right Jacobi preconditioning.

Pinned by exact SuperLU solutions of the same systems (residual and
solution agreement), see ``tests/test_oracle_fem.py``.
"""
import numpy as np


def cocg(A, b, diag_inv=None, tol=1e-8, maxit=10000, x0=None):
    b = np.asarray(b, dtype=np.complex128)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.complex128)
    if diag_inv is None:
        diag_inv = np.ones_like(b)
    nb = np.linalg.norm(b)
    if nb == 0.0:
        return x, 0, 0.0
    r = b - A @ x
    z = diag_inv * r
    p = z.copy()
    rho = r @ z
    it = 0
    res = np.linalg.norm(r) / nb
    while res > tol and it < maxit:
        q = A @ p
        mu = p @ q
        alpha = rho / mu
        x += alpha * p
        r -= alpha * q
        z = diag_inv * r
        rho_new = r @ z
        beta = rho_new / rho
        rho = rho_new
        p = z + beta * p
        it += 1
        res = np.linalg.norm(r) / nb
    return x, it, np.linalg.norm(b - A @ x) / nb


def gmres(A, b, diag_inv=None, tol=1e-8, restart=30, maxit=10000, x0=None):
    b = np.asarray(b, dtype=np.complex128)
    n = b.shape[0]
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.complex128)
    if diag_inv is None:
        diag_inv = np.ones_like(b)
    nb = np.linalg.norm(b)
    if nb == 0.0:
        return x, 0, 0.0
    it = 0
    while True:
        r = b - A @ x
        beta = np.linalg.norm(r)
        if beta / nb <= tol or it >= maxit:
            break
        V = np.zeros((restart + 1, n), dtype=np.complex128)
        H = np.zeros((restart + 1, restart), dtype=np.complex128)
        cs = np.zeros(restart, dtype=np.complex128)
        sn = np.zeros(restart, dtype=np.complex128)
        g = np.zeros(restart + 1, dtype=np.complex128)
        g[0] = beta
        V[0] = r / beta
        k = 0
        for j in range(restart):
            w = A @ (diag_inv * V[j])
            for i in range(j + 1):
                H[i, j] = np.vdot(V[i], w)
                w = w - H[i, j] * V[i]
            H[j + 1, j] = np.linalg.norm(w)
            if H[j + 1, j] != 0:
                V[j + 1] = w / H[j + 1, j]
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -np.conj(sn[i]) * H[i, j] + np.conj(cs[i]) * H[i + 1, j]
                H[i, j] = t
            a, bb = H[j, j], H[j + 1, j]
            den = np.sqrt(abs(a) ** 2 + abs(bb) ** 2)
            cs[j] = a / den if den else 1.0
            sn[j] = bb / den if den else 0.0
            # rotation [c s; -conj(s) conj(c)] with c = a/den, s = bb/den maps (a,bb) -> (den,0)
            cs[j], sn[j] = np.conj(cs[j]), np.conj(sn[j])
            H[j, j] = cs[j] * a + sn[j] * bb
            H[j + 1, j] = 0.0
            g[j + 1] = -np.conj(sn[j]) * g[j]
            g[j] = cs[j] * g[j]
            k = j + 1
            it += 1
            if abs(g[j + 1]) / nb <= tol or it >= maxit:
                break
        y = np.linalg.solve(np.triu(H[:k, :k]), g[:k])
        x = x + diag_inv * (V[:k].T @ y)
    return x, it, np.linalg.norm(b - A @ x) / nb
