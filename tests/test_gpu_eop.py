"""Element-by-element operator (y = A x without the assembled matrix) against
the oracle's assembled FP64 system and the device CSR SpMV.

Bars: 1e-12 relative (max norm) in complex128 against the oracle's CSR product,
whose summation order differs; 1e-5 in complex64 (single-precision products
of 6-45 terms per element and 1-8 elements per row).  At p = 3 the operator
has a complex64 product only."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from petgem_amd import _native as nat  # noqa: E402
from petgem_amd import core  # noqa: E402
from petgem_amd.synthetic import Workload  # noqa: E402
from oracle import fem, topology as otopo  # noqa: E402

DEV = "cuda"


def _setup(n, p, seed=7, frequency=1.0):
    w = Workload(n, p, noise=0.1, seed=seed, frequency=frequency)
    T = w.topo
    mesh = core.DeviceMesh.from_topology(T, DEV)
    asm = core.Assembler(mesh)
    asm.geometry(w.sigma)
    return w, T, mesh, asm


def _oracle_matrix(T, sigma, p, frequency):
    eo, fo = otopo.orientations(T.elems_nodes, T.elems_edges, T.edges_nodes, T.elem_faces_edges)
    M, K = fem.all_element_matrices(T.coords, eo, fo, sigma, p)
    A = fem.assemble(T.dofs, M, K, frequency, T.ndof)
    A, _ = fem.apply_dirichlet(A, np.zeros(T.ndof, complex), T.boundary_dofs)
    return A.tocsr()


def _rand(ndof, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(ndof) + 1j * rng.standard_normal(ndof)


@pytest.mark.parametrize("p,n", [(1, 8), (2, 5), (3, 3)])
@pytest.mark.parametrize("order", ["morton", "natural"])
def test_apply_matches_oracle_system(p, n, order):
    w, T, mesh, asm = _setup(n, p)
    omega = core.omega_of(w.frequency)
    ent_bc, _ = mesh.entity_mask(T.boundary_dofs)
    eop = core.ElementOperator(asm, omega, ent_bc,
                               mesh.entity_order() if order == "morton" else None)
    A = _oracle_matrix(T, w.sigma, p, w.frequency)
    for seed in (1, 2):
        x = _rand(T.ndof, seed)
        y_ref = A @ x
        x32 = core.from_complex(x, DEV).float()
        y32 = eop.apply(x32).double().cpu().numpy()
        y32 = y32[:, 0] + 1j * y32[:, 1]
        err32 = np.abs(y32 - y_ref).max() / np.abs(y_ref).max()
        y = core.as_complex(eop.apply(core.from_complex(x, DEV)))
        err = np.abs(y - y_ref).max() / np.abs(y_ref).max()
        print(f"p={p} n={n} {order}: complex128 err {err:.2e}, complex64 err {err32:.2e}",
              flush=True)
        assert err <= 1e-12
        assert err32 <= 1e-5
    # PEC rows are the identity
    bd = np.asarray(T.boundary_dofs)
    assert np.abs(y[bd] - x[bd]).max() == 0.0


@pytest.mark.parametrize("p,n", [(1, 12), (2, 8)])
def test_apply_matches_device_csr_without_pec(p, n):
    """No PEC entities: against the device CSR SpMV of pg_assemble's matrix."""
    w, T, mesh, asm = _setup(n, p, seed=3)
    omega = core.omega_of(0.5)
    vals = asm.assemble(omega)
    eop = core.ElementOperator(asm, omega, None, mesh.entity_order())
    x = core.from_complex(_rand(T.ndof, 5), DEV)
    y_csr = core.as_complex(core.spmv(asm.plan, vals, x))
    y = core.as_complex(eop.apply(x))
    err = np.abs(y - y_csr).max() / np.abs(y_csr).max()
    print(f"p={p} n={n} no PEC: err vs CSR {err:.2e}", flush=True)
    assert err <= 1e-12


def test_refresh_follows_new_conductivity_and_frequency():
    """After Assembler.geometry with another conductivity model and a refresh at
    another frequency, the operator is the reassembled system's."""
    w, T, mesh, asm = _setup(5, 2)
    ent_bc, _ = mesh.entity_mask(T.boundary_dofs)
    eop = core.ElementOperator(asm, core.omega_of(1.0), ent_bc, mesh.entity_order())
    sigma2 = w.sigma * np.exp(0.5 * np.random.default_rng(4).standard_normal(w.sigma.shape))
    asm.geometry(sigma2)
    eop.refresh(core.omega_of(0.25))
    vals = asm.assemble(core.omega_of(0.25), ent_bc=ent_bc)
    x = core.from_complex(_rand(T.ndof, 9), DEV)
    y_csr = core.as_complex(core.spmv(asm.plan, vals, x))
    y = core.as_complex(eop.apply(x))
    err = np.abs(y - y_csr).max() / np.abs(y_csr).max()
    print(f"refresh: err vs reassembled CSR {err:.2e}", flush=True)
    assert err <= 1e-12


def test_repeated_apply_is_bitwise_reproducible():
    w, T, mesh, asm = _setup(6, 2)
    ent_bc, _ = mesh.entity_mask(T.boundary_dofs)
    eop = core.ElementOperator(asm, core.omega_of(1.0), ent_bc, mesh.entity_order())
    x = core.from_complex(_rand(T.ndof, 3), DEV)
    y0 = eop.apply(x)
    for _ in range(3):
        assert torch.equal(eop.apply(x), y0)


@pytest.mark.parametrize("p,n,precision", [(2, 6, "fp64"), (2, 6, "mixed"), (3, 4, "mixed")])
def test_solver_iterations_on_operator_converge_on_csr_residual(p, n, precision):
    """spmv="element": the Krylov iterations run on the operator, the true
    residual (on the CSR matrix) reaches the tolerance, and the solution
    matches the CSR-kernel solve, in both precisions (at p = 3 the complex128
    products run on the matrix cores)."""
    w, T, mesh, asm = _setup(n, p, seed=11)
    omega = core.omega_of(w.frequency)
    ent_bc, dof_bc = mesh.entity_mask(T.boundary_dofs)
    vals = asm.assemble(omega, ent_bc=ent_bc)
    el = int(core.locate_points(mesh, np.array(w.source["position"])[None]).cpu()[0])
    b = core.source_vector(mesh, el, w.source["position"], core.dipole_moment(w.source), omega)
    core.apply_dirichlet(asm.plan, None, dof_bc, b)
    order = mesh.entity_order()
    eop = core.ElementOperator(asm, omega, ent_bc, order)
    kw = dict(precision=precision, pc="bjacobi", block_ptr=mesh.block_ptr(), entity_order=order,
              rtol=1e-8)
    S_e = core.KrylovSolver(asm.plan, vals, spmv="element", eop=eop, **kw)
    S_c = core.KrylovSolver(asm.plan, vals, spmv="entity", **kw)
    prof = S_e.profile(reps=3)
    assert prof["g64"] == 3000  # p = 3 included: the matrix-core complex128 product
    xe, ie = S_e.solve(b)
    xc, ic = S_c.solve(b)
    print(f"{precision}: element iters {ie['iterations']} relres {ie['relres']:.2e}; "
          f"CSR iters {ic['iterations']}", flush=True)
    assert ie["status"] == 0 and ie["relres"] <= 1e-8
    xe, xc = core.as_complex(xe), core.as_complex(xc)
    assert np.abs(xe - xc).max() <= 1e-6 * np.abs(xc).max()


@pytest.mark.parametrize("p", [1, 2, 3])
def test_apply_on_reference_unstructured_mesh(test_mesh, p):
    """The reference's own unstructured test mesh (9453 tets, four materials;
    many orientation combinations, so many partly filled wavefronts): the
    operator against the oracle's assembled system with the PEC condition."""
    from petgem_amd.mesh import Topology
    pts, en, phys = test_mesh
    T = Topology(pts, en, p)
    sig = np.stack([np.array([1.0, 0.01, 1.0, 3.3333])[phys - 1]] * 2, axis=1)
    mesh = core.DeviceMesh.from_topology(T, DEV)
    asm = core.Assembler(mesh)
    asm.geometry(sig)
    ent_bc, _ = mesh.entity_mask(T.boundary_dofs)
    eop = core.ElementOperator(asm, core.omega_of(2.0), ent_bc, mesh.entity_order())
    A = _oracle_matrix(T, sig, p, 2.0)
    x = _rand(T.ndof, 4)
    y_ref = A @ x
    y = core.as_complex(eop.apply(core.from_complex(x, DEV)))
    err = np.abs(y - y_ref).max() / np.abs(y_ref).max()
    y32 = eop.apply(core.from_complex(x, DEV).float()).double().cpu().numpy()
    err32 = np.abs(y32[:, 0] + 1j * y32[:, 1] - y_ref).max() / np.abs(y_ref).max()
    eo, fo = otopo.orientations(T.elems_nodes, T.elems_edges, T.edges_nodes, T.elem_faces_edges)
    combos = len({(tuple(a), tuple(b)) for a, b in zip(eo.tolist(), fo.tolist())})
    print(f"reference mesh p={p}: {combos} orientation combinations, complex128 err {err:.2e}, "
          f"complex64 err {err32:.2e}", flush=True)
    assert err <= 1e-12 and err32 <= 1e-5
