"""C + OpenMP CPU baseline of the forward solve -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper of ``oracle/libpgcpu.so`` (``cpu_baseline.c``, built by
``make -C oracle``): element matrices in the reference's quadrature form
(``petgem/hvfem.py:258-314``) with the reference basis tabulated by
``oracle.basis`` per local DOF and orientation value, CSR assembly of
A = K - i w mu0 M (``petgem/solver.py:191-235``), the PEC condition
(``petgem/solver.py:552-577``) and COCG + Jacobi to a true relative residual
(the KSP step, ``petgem/solver.py:584-590``), all on every host core.

Used as the timed CPU baseline of ``bench.py`` and pinned against the numpy
oracle (``tests/test_oracle_cpu_baseline.py``: assembled values, boundary
condition, SuperLU solution).
"""
import ctypes
import os
import time

import numpy as np

from . import basis as B
from . import fem
from . import quadrature as Q
from . import topology as T

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpgcpu.so")
_lib = None
_P = ctypes.c_void_p
_I64 = ctypes.c_int64


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.pgcpu_threads.restype = ctypes.c_int
        L.pgcpu_assemble.restype = ctypes.c_int
        L.pgcpu_assemble.argtypes = [_I64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P,
                                     _P, _P, _P, _P, _I64, _P, _P, ctypes.c_double, _P]
        L.pgcpu_dirichlet.restype = None
        L.pgcpu_dirichlet.argtypes = [_I64, _P, _P, _P, _P, _P]
        L.pgcpu_cocg.restype = ctypes.c_int
        L.pgcpu_cocg.argtypes = [_I64, _P, _P, _P, _P, _P, ctypes.c_double, ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def threads():
    return int(lib().pgcpu_threads())


def basis_tables(p):
    """S, C (n, no, nq, 3): reference basis and curls of local DOF i for
    orientation value o (edge flag 0/1, face code 0..5) at the order-2p
    quadrature points; w (nq)."""
    pts, w = Q.rule_for_order(p)
    n = B.ndof_element(p)
    no = 6 if p >= 2 else 2
    S = np.zeros((n, no, pts.shape[0], 3))
    C = np.zeros_like(S)
    for o in range(no):
        # edge functions depend on their edge's flag only, face functions on
        # their face's code only (hvfem.py:379,400)
        s, c = B.shape_functions(pts, [o % 2] * 6, [o] * 4, p)
        S[:, o] = s.transpose(1, 0, 2)
        C[:, o] = c.transpose(1, 0, 2)
    return np.ascontiguousarray(S), np.ascontiguousarray(C), np.ascontiguousarray(w)


def dof_orientation(edge_orient, face_orient, p):
    """(E, n) orientation value of every local DOF (hvfem.py:52-93 block order)."""
    de, df, dv = T.dof_counts(p)
    eo = np.asarray(edge_orient, dtype=np.int32)
    fo = np.asarray(face_orient, dtype=np.int32)
    parts = [np.repeat(eo, de, axis=1)]
    if df:
        parts.append(np.repeat(fo, df, axis=1))
    if dv:
        parts.append(np.zeros((eo.shape[0], dv), dtype=np.int32))
    return np.ascontiguousarray(np.concatenate(parts, axis=1), dtype=np.int32)


class System:
    """Inputs of one forward solve prepared for the C baseline (untimed set-up:
    basis tables, per-DOF orientations, the CSR pattern)."""

    def __init__(self, coords, sigma, edge_orient, face_orient, dofs, ndof, p, rowptr, cols):
        self.p = int(p)
        self.coords = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 12)
        self.E = self.coords.shape[0]
        self.sigma = np.ascontiguousarray(sigma, dtype=np.float64)
        self.orient = dof_orientation(edge_orient, face_orient, p)
        self.dofs = np.ascontiguousarray(dofs, dtype=np.int32)
        self.n = self.dofs.shape[1]
        self.ndof = int(ndof)
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        self.cols = np.ascontiguousarray(cols, dtype=np.int32)
        self.S, self.C, self.w = basis_tables(self.p)
        self.values = np.empty((self.cols.size, 2))

    def assemble(self, frequency):
        rc = lib().pgcpu_assemble(self.E, self.n, self.w.size, self.S.shape[1], _p(self.coords),
                                  _p(self.sigma), _p(self.orient), _p(self.S), _p(self.C),
                                  _p(self.w), _p(self.dofs), self.ndof, _p(self.rowptr),
                                  _p(self.cols), 2.0 * np.pi * float(frequency), _p(self.values))
        if rc != 0:
            raise RuntimeError("an element DOF pair is missing from the CSR pattern")
        return self.values

    def dirichlet(self, bdofs, b):
        mask = np.zeros(self.ndof, dtype=np.uint8)
        mask[np.asarray(bdofs, dtype=np.int64)] = 1
        bb = np.ascontiguousarray(np.stack([np.real(b), np.imag(b)], axis=1))
        lib().pgcpu_dirichlet(self.ndof, _p(self.rowptr), _p(self.cols), _p(mask),
                              _p(self.values), _p(bb))
        return bb

    def cocg(self, bb, tol=1e-8, maxit=100000):
        x = np.zeros_like(bb)
        rel = ctypes.c_double()
        it = lib().pgcpu_cocg(self.ndof, _p(self.rowptr), _p(self.cols), _p(self.values), _p(bb),
                              _p(x), float(tol), int(maxit), ctypes.byref(rel))
        return x[:, 0] + 1j * x[:, 1], int(it), rel.value

    def matrix(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.values[:, 0] + 1j * self.values[:, 1], self.cols, self.rowptr),
                             shape=(self.ndof, self.ndof))


def forward(system, frequency, b, bdofs, tol=1e-8, maxit=100000):
    """Assembly + boundary condition + COCG, each timed (seconds)."""
    t0 = time.perf_counter()
    system.assemble(frequency)
    t1 = time.perf_counter()
    bb = system.dirichlet(bdofs, b)
    t2 = time.perf_counter()
    x, it, rel = system.cocg(bb, tol=tol, maxit=maxit)
    t3 = time.perf_counter()
    return dict(x=x, iterations=it, relres=rel, assembly_s=t1 - t0, bc_s=t2 - t1,
                solve_s=t3 - t2, total_s=t3 - t0, threads=threads())


def source(coords, dofs, edge_orient, face_orient, p, ndof, src):
    """Right-hand side of the point dipole (the oracle's restatement)."""
    el = int(fem.locate(coords, np.asarray(src["position"])[None, :])[0])
    return fem.source_vector(coords, dofs, edge_orient, face_orient, p, ndof, src, el)
