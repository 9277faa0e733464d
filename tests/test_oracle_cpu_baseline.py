"""The C + OpenMP CPU baseline (oracle/cpu_baseline.c) against the numpy
oracle: assembled values, the PEC condition and the solution."""
import numpy as np
import pytest

from oracle import cpu, fem, topology as otopo
from petgem_amd.synthetic import Workload


def _case(n, p):
    w = Workload(n, p, noise=0.1, seed=4, n_rx=5)
    T = w.topo
    eo, fo = otopo.orientations(T.elems_nodes, T.elems_edges, T.edges_nodes, T.elem_faces_edges)
    indptr, cols = fem.csr_pattern(T.dofs, T.ndof)
    sysm = cpu.System(T.coords, w.sigma, eo, fo, T.dofs, T.ndof, p, indptr, cols)
    return w, T, eo, fo, sysm


@pytest.mark.parametrize("n,p", [(5, 1), (3, 2), (2, 3)])
def test_assembly_and_bc_match_oracle(n, p):
    w, T, eo, fo, sysm = _case(n, p)
    sysm.assemble(w.frequency)
    M, K = fem.all_element_matrices(T.coords, eo, fo, w.sigma, p)
    A = fem.assemble(T.dofs, M, K, w.frequency, T.ndof)
    Ac = sysm.matrix()
    assert np.array_equal(Ac.indptr, A.indptr) and np.array_equal(Ac.indices, A.indices)
    assert np.abs(Ac.data - A.data).max() <= 1e-12 * np.abs(A.data).max()
    b = cpu.source(T.coords, T.dofs, eo, fo, p, T.ndof, w.source)
    bb = sysm.dirichlet(T.boundary_dofs, b)
    Abc, bbc = fem.apply_dirichlet(A, b, T.boundary_dofs)
    assert np.abs(sysm.matrix().data - Abc.data).max() <= 1e-12 * np.abs(Abc.data).max()
    assert np.array_equal(bb[:, 0] + 1j * bb[:, 1], bbc)


@pytest.mark.parametrize("n,p", [(6, 1), (3, 2)])
def test_cocg_solution_matches_superlu(n, p):
    w, T, eo, fo, sysm = _case(n, p)
    b = cpu.source(T.coords, T.dofs, eo, fo, p, T.ndof, w.source)
    r = cpu.forward(sysm, w.frequency, b, T.boundary_dofs, tol=1e-10)
    ref = fem.forward(dict(edge_orient=eo, face_orient=fo, dofs=T.dofs, ndof=T.ndof,
                           boundary_dofs=T.boundary_dofs), T.coords, w.sigma, p, w.source,
                      w.receivers)
    assert r["relres"] <= 1e-10 and r["threads"] >= 1
    assert np.abs(r["x"] - ref["x"]).max() <= 1e-7 * np.abs(ref["x"]).max()
