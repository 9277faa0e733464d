"""Failure handling on the device path (every kernel exits normally; the errors
come from host-side checks or device flags read back by the entry points), and
run-to-run bitwise reproducibility of assembly and solve."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from petgem_amd import _native as nat, core  # noqa: E402
from petgem_amd.synthetic import Workload  # noqa: E402

DEV = "cuda"


def _mesh(w):
    mesh = core.DeviceMesh.from_topology(w.topo, DEV)
    return mesh, core.Assembler(mesh)


def test_degenerate_element_is_rejected():
    w = Workload(3, 2)
    coords = w.topo.coords.copy()
    coords[5, 1] = coords[5, 0]  # two coincident vertices: det J = 0
    T = w.topo
    mesh = core.DeviceMesh(T.p, coords, T.elems_nodes, T.elems_edges, T.elem_edges_nodes,
                           T.elem_faces_edges, T.dofs, T.n_edges, T.n_faces, T.elems_faces, DEV)
    asm = core.Assembler(mesh)
    with pytest.raises(nat.PgError, match="degenerate element"):
        asm.geometry(w.sigma)


@pytest.mark.parametrize("bad", [-1.0, np.nan, np.inf])
def test_invalid_conductivity_is_rejected(bad):
    w = Workload(3, 1)
    mesh, asm = _mesh(w)
    sigma = w.sigma.copy()
    sigma[7, 0] = bad
    with pytest.raises(nat.PgError, match="conductivity"):
        asm.geometry(sigma)


def test_negative_determinant_follows_the_reference_sign():
    """Swapping two vertices of one element flips the signed determinant; the
    reference keeps the sign (hvfem.py:265), so the element matrices change sign."""
    w = Workload(3, 1)
    T = w.topo
    mesh, asm = _mesh(w)
    asm.geometry(w.sigma)
    M0, K0 = asm.element_matrices()
    en = T.elems_nodes.copy()
    en[4, [1, 2]] = en[4, [2, 1]]
    from petgem_amd.mesh import Topology
    T2 = Topology(T.points, en, 1)
    m2 = core.DeviceMesh.from_topology(T2, DEV)
    a2 = core.Assembler(m2)
    a2.geometry(w.sigma)
    M2, K2 = a2.element_matrices()
    M0, M2 = M0.cpu().numpy(), M2.cpu().numpy()
    # same element, reversed orientation: M keeps its spectrum up to the sign of det
    e0 = np.sort(np.abs(np.linalg.eigvalsh(M0[4])))
    e2 = np.sort(np.abs(np.linalg.eigvalsh(M2[4])))
    assert np.allclose(e0, e2, rtol=1e-10, atol=0)
    assert (np.linalg.eigvalsh(M2[4]) < 0).all() and (np.linalg.eigvalsh(M0[4]) > 0).all()


def test_source_outside_the_domain_is_rejected():
    w = Workload(3, 2)
    mesh, asm = _mesh(w)
    el = core.locate_points(mesh, np.array([[0.0, 0.0, 9000.0]])).cpu().numpy()
    assert el[0] == -1
    with pytest.raises(nat.PgError, match="outside"):
        core.source_vector(mesh, int(el[0]), [0.0, 0.0, 9000.0], (1.0, 0.0, 0.0), 1.0)


def test_receivers_outside_the_mesh_give_zero_fields():
    w = Workload(3, 2)
    mesh, asm = _mesh(w)
    pts = np.array([[0.0, 0.0, -1000.0], [0.0, 0.0, 9000.0]])
    el = core.locate_points(mesh, pts)
    assert el.cpu().numpy()[1] == -1
    x = torch.ones((mesh.ndof, 2), dtype=torch.float64, device=DEV)
    f = core.as_complex(core.receiver_fields(mesh, pts, el, x, 1.0))
    assert np.abs(f[1]).max() == 0.0 and np.abs(f[0]).max() > 0.0


def test_zero_diagonal_is_rejected_by_the_preconditioner():
    w = Workload(3, 1)
    mesh, asm = _mesh(w)
    asm.geometry(w.sigma)
    vals = asm.assemble(core.omega_of(1.0))
    pl = asm.plan
    rp = pl.h_rowptr
    row = 3
    s0, s1 = rp[row], rp[row + 1]
    diag = s0 + int(np.flatnonzero(pl.h_cols[s0:s1] == row)[0])
    vals[diag] = 0.0
    with pytest.raises(nat.PgError, match="singular diagonal"):
        core.KrylovSolver(pl, vals, pc="jacobi")


def test_non_convergence_is_reported_not_raised():
    w = Workload(4, 2)
    mesh, asm = _mesh(w)
    asm.geometry(w.sigma)
    om = core.omega_of(1.0)
    ent_bc, dof_bc = mesh.entity_mask(w.topo.boundary_dofs)
    vals = asm.assemble(om, ent_bc=ent_bc)
    el = int(core.locate_points(mesh, np.array(w.source["position"])[None]).cpu()[0])
    b = core.source_vector(mesh, el, w.source["position"], (1.0, 0.0, 0.0), om)
    core.apply_dirichlet(asm.plan, None, dof_bc, b)
    x, info = core.KrylovSolver(asm.plan, vals, rtol=1e-10, maxit=5).solve(b)
    assert info["status"] == -1 and info["iterations"] <= 5 + 25 and info["relres"] > 1e-10


def test_assembly_and_solve_are_bitwise_reproducible():
    """Assembly, and a solve with one kernel configuration, repeat bitwise.  The
    SpMV configuration is picked by timing when a solver is created, and the
    configurations sum a row in different orders, so the second solver takes
    the first one's choice (like=); the first one also solves twice."""
    w = Workload(5, 2, noise=0.1, seed=9)
    mesh, asm = _mesh(w)
    om = core.omega_of(1.0)
    ent_bc, dof_bc = mesh.entity_mask(w.topo.boundary_dofs)
    el = int(core.locate_points(mesh, np.array(w.source["position"])[None]).cpu()[0])
    runs, first = [], None
    for _ in range(2):
        asm.geometry(w.sigma)
        vals = asm.assemble(om, ent_bc=ent_bc)
        b = core.source_vector(mesh, el, w.source["position"], (1.0, 0.0, 0.0), om)
        core.apply_dirichlet(asm.plan, None, dof_bc, b)
        S = core.KrylovSolver(asm.plan, vals, precision="mixed", pc="bjacobi", rtol=1e-9,
                              block_ptr=mesh.block_ptr(), entity_order=mesh.entity_order(),
                              spmv="entity_pairs", like=first)
        first = first or S
        x, info = S.solve(b)
        runs.append((vals.clone(), x.clone(), info["iterations"]))
    x2, info2 = first.solve(b)
    assert torch.equal(runs[0][0], runs[1][0])
    assert torch.equal(runs[0][1], runs[1][1]) and runs[0][2] == runs[1][2]
    assert torch.equal(x2, runs[0][1]) and info2["iterations"] == runs[0][2]
