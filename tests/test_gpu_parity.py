"""GPU parity of the HIP hot path against the oracle (CPU restatement of the reference).

Bars (stated per test): bit-exact for integer work (orientations, CSR
pattern, point location), 1e-12 relative (max-norm per element / per
matrix) for element matrices and assembled values, 1e-6 relative for
receiver fields against an exact (SuperLU) FP64 solve, true residual
<= 1e-8 for every Krylov solve.
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from petgem_amd import _native as nat  # noqa: E402
from petgem_amd import core  # noqa: E402
from petgem_amd.mesh import Topology, LOCAL_EDGES, FACE_EDGE_TRIPLES  # noqa: E402
from petgem_amd.synthetic import Workload  # noqa: E402
from oracle import fem, topology as otopo, krylov  # noqa: E402

DEV = "cuda"
# face code -> (k1, k2) positions, inverse of hvfem.py:199-211
CODE_POS = {0: (1, 2), 1: (3, 1), 2: (2, 3), 3: (3, 2), 4: (1, 3), 5: (2, 1)}


def _t(a, dt):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(DEV)


def synth_connectivity(eo, fo):
    """Per-element scratch arrays that make computeElementOrientation yield (eo, fo)."""
    E = eo.shape[0]
    en = np.arange(4 * E).reshape(E, 4)
    ee = np.arange(6 * E).reshape(E, 6)
    edn = np.zeros((E, 6, 2), dtype=np.int64)
    for i, (a, b) in enumerate(LOCAL_EDGES):
        flag = eo[:, i].astype(bool)
        edn[:, i, 0] = np.where(flag, en[:, b], en[:, a])
        edn[:, i, 1] = np.where(flag, en[:, a], en[:, b])
    fe = np.zeros((E, 4, 3), dtype=np.int64)
    for e in range(E):
        for f in range(4):
            g = ee[e, FACE_EDGE_TRIPLES[f]]
            k1, k2 = CODE_POS[int(fo[e, f])]
            k3 = 6 - k1 - k2
            fe[e, f, k1 - 1], fe[e, f, k2 - 1], fe[e, f, k3 - 1] = g[0], g[1], g[2]
    return en, ee, edn, fe


def gpu_element_matrices(p, coords, en, ee, edn, fe, sigma):
    L = nat.lib()
    E = coords.shape[0]
    c, en_, ee_ = _t(coords.reshape(E, 12), np.float64), _t(en, np.int32), _t(ee, np.int32)
    edn_, fe_ = _t(edn.reshape(E, 12), np.int32), _t(fe.reshape(E, 12), np.int32)
    sg = _t(sigma, np.float64)
    ws = torch.empty(int(L.pg_element_workspace_bytes(E)), dtype=torch.uint8, device=DEV)
    tab = core.Tables(p)
    s = nat.stream_handle()
    nat.check(L.pg_element_geometry(p, E, nat.ptr(c), nat.ptr(en_), nat.ptr(ee_), nat.ptr(edn_),
                                    nat.ptr(fe_), nat.ptr(sg), nat.ptr(ws), s))
    n = p * (p + 2) * (p + 3) // 2
    M = torch.empty((E, n, n), dtype=torch.float64, device=DEV)
    K = torch.empty_like(M)
    nat.check(L.pg_element_matrices(tab.handle, E, nat.ptr(ws), nat.ptr(M), nat.ptr(K), s))
    torch.cuda.synchronize()
    return M.cpu().numpy(), K.cpu().numpy()


def rel_err_per_element(A, B):
    den = np.abs(B).reshape(B.shape[0], -1).max(axis=1)
    return (np.abs(A - B).reshape(A.shape[0], -1).max(axis=1) / den).max()


@pytest.mark.parametrize("p", [1, 2, 3])
def test_element_matrices_vs_reference_golden(golden, p):
    """Against computeElementalMatrices itself (golden vectors, arbitrary orientations)."""
    nodes = golden["elem_nodes"]
    eo, fo = golden["elem_edge_orient"], golden["elem_face_orient"]
    en, ee, edn, fe = synth_connectivity(eo, fo)
    M, K = gpu_element_matrices(p, nodes, en, ee, edn, fe, golden["elem_sigma"])
    print(f"p={p} golden M err {rel_err_per_element(M, golden[f'elem_M_p{p}']):.2e} "
          f"K err {rel_err_per_element(K, golden[f'elem_K_p{p}']):.2e}", flush=True)
    assert rel_err_per_element(M, golden[f"elem_M_p{p}"]) <= 1e-12
    assert rel_err_per_element(K, golden[f"elem_K_p{p}"]) <= 1e-12


def test_orientation_bit_exact(test_mesh, topology_golden):
    """Against the reference's own computeElementOrientation on its test mesh
    (tests/golden/topology_vectors.npz, hvfem.py:122-220) and the oracle."""
    pts, en, _ = test_mesh
    T = Topology(pts, en, 2)
    eo_ref, fo_ref = otopo.orientations(en, T.elems_edges, T.edges_nodes, T.elem_faces_edges)
    assert np.array_equal(eo_ref, topology_golden["mesh_edge_orient"])
    assert np.array_equal(fo_ref, topology_golden["mesh_face_orient"])
    E = T.n_elems
    eo = torch.empty((E, 6), dtype=torch.int32, device=DEV)
    fo = torch.empty((E, 4), dtype=torch.int32, device=DEV)
    # keep every input tensor alive until the kernel has run
    ins = [_t(en, np.int32), _t(T.elems_edges, np.int32),
           _t(T.elem_edges_nodes.reshape(E, 12), np.int32),
           _t(T.elem_faces_edges.reshape(E, 12), np.int32)]
    nat.check(nat.lib().pg_element_orientation(
        E, *[nat.ptr(t) for t in ins], nat.ptr(eo), nat.ptr(fo), nat.stream_handle()))
    torch.cuda.synchronize()
    assert np.array_equal(eo.cpu().numpy(), eo_ref)
    assert np.array_equal(fo.cpu().numpy(), fo_ref)
    assert len(np.unique(fo_ref)) == 6  # every face code is exercised by this mesh


@pytest.mark.parametrize("p", [1, 2, 3])
def test_element_matrices_vs_oracle_mesh(test_mesh, p):
    pts, en, phys = test_mesh
    T = Topology(pts, en, p)
    sel = np.arange(0, T.n_elems, 7)[:600]
    sig = np.stack([np.array([1.0, 0.01, 1.0, 3.3333])[phys - 1]] * 2, axis=1)
    sig[:, 1] *= 0.5  # anisotropic: sigma_v != sigma_h
    M, K = gpu_element_matrices(p, T.coords[sel], en[sel], T.elems_edges[sel],
                                T.elem_edges_nodes[sel], T.elem_faces_edges[sel], sig[sel])
    eo, fo = otopo.orientations(en[sel], T.elems_edges[sel], T.edges_nodes,
                                T.elem_faces_edges[sel])
    Mr, Kr = fem.all_element_matrices(T.coords[sel], eo, fo, sig[sel], p)
    em, ek = rel_err_per_element(M, Mr), rel_err_per_element(K, Kr)
    print(f"p={p} mesh M err {em:.2e} K err {ek:.2e}", flush=True)
    assert em <= 1e-12 and ek <= 1e-12


def _oracle_system(T, sigma, freq):
    eo, fo = otopo.orientations(T.elems_nodes, T.elems_edges, T.edges_nodes, T.elem_faces_edges)
    M, K = fem.all_element_matrices(T.coords, eo, fo, sigma, T.p)
    A = fem.assemble(T.dofs, M, K, freq, T.ndof)
    return A, eo, fo


@pytest.mark.parametrize("p,n", [(1, 6), (2, 5), (3, 3)])
def test_assembly_vs_oracle(p, n):
    w = Workload(n, p, noise=0.1, seed=3)
    T = w.topo
    mesh = core.DeviceMesh.from_topology(T, DEV)
    asm = core.Assembler(mesh)
    asm.geometry(w.sigma)
    omega = core.omega_of(w.frequency)
    vals = core.as_complex(asm.assemble(omega))
    A, _, _ = _oracle_system(T, w.sigma, w.frequency)
    assert np.array_equal(asm.plan.h_rowptr, A.indptr)
    assert np.array_equal(asm.plan.h_cols, A.indices)
    err = np.abs(vals - A.data).max() / np.abs(A.data).max()
    print(f"p={p} assembly rel err {err:.2e} nnz={A.nnz}", flush=True)
    assert err <= 1e-12
    # fixed-order, atomic-free: bitwise reproducible and equal to the ordered
    # sum of the device's own element matrices
    vals2 = core.as_complex(asm.assemble(omega))
    assert np.array_equal(vals, vals2)
    Mg, Kg = asm.element_matrices()
    Mg, Kg = Mg.cpu().numpy(), Kg.cpu().numpy()
    om = omega * core.MU0
    nn = Mg.shape[1]
    rows = np.repeat(T.dofs, nn, axis=1).astype(np.int64)
    cols = np.tile(T.dofs, (1, nn)).astype(np.int64)
    prow = np.repeat(np.arange(T.ndof), np.diff(A.indptr)).astype(np.int64)
    keys = prow * T.ndof + A.indices
    idx = np.searchsorted(keys, rows * T.ndof + cols)
    re = np.zeros(A.nnz)
    im = np.zeros(A.nnz)
    # element-major order = ascending element index per slot, as ADD_VALUES on one rank
    np.add.at(re, idx.ravel(), Kg.ravel())
    np.add.at(im, idx.ravel(), -(om * Mg.ravel()))
    assert np.array_equal(vals.real, re) and np.array_equal(vals.imag, im)


@pytest.mark.parametrize("p,n", [(1, 6), (2, 5), (2, 8), (3, 3)])
def test_row_major_assembly_is_bitwise_the_two_phase(p, n):
    """The row-major kernel (plans exported with entity_pair; its short-row
    variant, and the long-row one when entity_row_max is unset) and the
    two-phase (p <= 2) or one-pass (p = 3) kernels (entity_pair NULL) give
    bitwise the same CSR values, with and without the PEC fold, for A(omega)
    and for K, M."""
    w = Workload(n, p, noise=0.1, seed=5)
    T = w.topo
    mesh = core.DeviceMesh.from_topology(T, DEV)
    asm = core.Assembler(mesh)
    assert asm.plan.entity_pair is not None
    asm.geometry(w.sigma)
    omega = core.omega_of(w.frequency)
    ent_bc, _ = mesh.entity_mask(T.boundary_dofs)
    res = {}
    assert 0 < asm.plan.dev_struct.entity_row_max <= (96 if p < 3 else 192)
    for mode in ("rows", "rows_long", "two_phase"):
        saved = (asm.plan.dev_struct.entity_pair, asm.plan.dev_struct.entity_row_max)
        if mode == "rows_long":
            asm.plan.dev_struct.entity_row_max = 0
        if mode == "two_phase":
            asm.plan.dev_struct.entity_pair = None
        try:
            res[mode] = (asm.assemble(omega).clone(), asm.assemble(omega, ent_bc=ent_bc).clone(),
                         [t.clone() for t in core.assemble_km(asm, ent_bc)])
        finally:
            asm.plan.dev_struct.entity_pair, asm.plan.dev_struct.entity_row_max = saved
    for mode in ("rows", "rows_long"):
        (a0, a1, km_a), (b0, b1, km_b) = res[mode], res["two_phase"]
        assert torch.equal(a0, b0) and torch.equal(a1, b1), mode
        assert torch.equal(km_a[0], km_b[0]) and torch.equal(km_a[1], km_b[1]), mode


def test_dirichlet_fold_equals_separate():
    w = Workload(5, 2)
    T = w.topo
    mesh = core.DeviceMesh.from_topology(T, DEV)
    asm = core.Assembler(mesh)
    asm.geometry(w.sigma)
    omega = core.omega_of(w.frequency)
    ent_bc, dof_bc = mesh.entity_mask(T.boundary_dofs)
    v_fold = asm.assemble(omega, ent_bc=ent_bc)
    v_sep = asm.assemble(omega)
    b = torch.zeros((mesh.ndof, 2), dtype=torch.float64, device=DEV)
    b[:, 0] = 1.0
    core.apply_dirichlet(asm.plan, v_sep, dof_bc, b)
    assert torch.equal(v_fold, v_sep)
    A, _, _ = _oracle_system(T, w.sigma, w.frequency)
    Abc, bbc = fem.apply_dirichlet(A, np.ones(T.ndof, dtype=complex), T.boundary_dofs)
    got = core.as_complex(v_sep)
    assert np.abs(got - Abc.data).max() <= 1e-12 * np.abs(Abc.data).max()
    assert np.array_equal(core.as_complex(b), bbc)


def test_locate_and_source_vs_oracle(test_mesh):
    pts, en, phys = test_mesh
    for p in (1, 2, 3):
        T = Topology(pts, en, p)
        mesh = core.DeviceMesh.from_topology(T, DEV)
        rx = np.load(__import__("os").path.join(
            __import__("os").path.dirname(__file__), "golden", "receiver_pos.npy"))
        probe = np.vstack([rx, [[1750.0, 1750.0, -975.0], [1e6, 0, 0]]])
        got = core.locate_points(mesh, probe).cpu().numpy()
        ref = fem.locate(T.coords, probe)
        assert np.array_equal(got, ref)
        assert got[-1] == -1
        src = dict(position=[1750.0, 1750.0, -975.0], azimuth=30.0, dip=10.0, current=2.0,
                   length=1.5, frequency=2.0)
        el = int(got[-2])
        eo, fo = otopo.orientations(en, T.elems_edges, T.edges_nodes, T.elem_faces_edges)
        bref = fem.source_vector(T.coords, T.dofs, eo, fo, p, T.ndof, src, el)
        b = core.source_vector(mesh, el, src["position"], core.dipole_moment(src),
                               core.omega_of(2.0))
        bg = core.as_complex(b)
        assert np.abs(bg - bref).max() <= 1e-12 * np.abs(bref).max()
        # receiver interpolation of an arbitrary complex vector
        rng = np.random.default_rng(p)
        x = rng.standard_normal(T.ndof) + 1j * rng.standard_normal(T.ndof)
        f = core.as_complex(core.receiver_fields(mesh, rx, got[:-2], core.from_complex(x, DEV),
                                                 core.omega_of(2.0)))
        fref = fem.receiver_fields(x, T.coords, T.dofs, eo, fo, p, 2.0, rx, ref[:-2])
        assert np.abs(f - fref).max() <= 1e-12 * np.abs(fref).max()


def test_spmv_vs_scipy():
    w = Workload(6, 2)
    T = w.topo
    mesh = core.DeviceMesh.from_topology(T, DEV)
    asm = core.Assembler(mesh)
    asm.geometry(w.sigma)
    v = asm.assemble(core.omega_of(1.0))
    A = sp.csr_matrix((core.as_complex(v), asm.plan.h_cols, asm.plan.h_rowptr),
                      shape=(T.ndof, T.ndof))
    rng = np.random.default_rng(0)
    x = rng.standard_normal(T.ndof) + 1j * rng.standard_normal(T.ndof)
    y = core.as_complex(core.spmv(asm.plan, v, core.from_complex(x, DEV)))
    yr = A @ x
    assert np.abs(y - yr).max() <= 1e-13 * np.abs(yr).max()


def _forward_gpu(w, precision="fp64", rtol=1e-8, ksp="cocg", pc="jacobi", spmv="auto"):
    T = w.topo
    mesh = core.DeviceMesh.from_topology(T, DEV)
    asm = core.Assembler(mesh)
    asm.geometry(w.sigma)
    omega = core.omega_of(w.frequency)
    ent_bc, dof_bc = mesh.entity_mask(T.boundary_dofs)
    vals = asm.assemble(omega)
    el = int(core.locate_points(mesh, np.array(w.source["position"])[None]).cpu()[0])
    b = core.source_vector(mesh, el, w.source["position"], core.dipole_moment(w.source), omega)
    core.apply_dirichlet(asm.plan, vals, dof_bc, b)
    order = None if spmv == "entity_natural" else mesh.entity_order()
    eop = core.ElementOperator(asm, omega, ent_bc, order) if spmv == "element" else None
    solver = core.KrylovSolver(asm.plan, vals, ksp=ksp, precision=precision, rtol=rtol, pc=pc,
                               spmv="entity" if spmv == "entity_natural" else spmv,
                               block_ptr=mesh.block_ptr(), entity_order=order, eop=eop)
    x, info = solver.solve(b)
    rx_el = core.locate_points(mesh, w.receivers)
    f = core.as_complex(core.receiver_fields(mesh, w.receivers, rx_el, x, omega))
    A = sp.csr_matrix((core.as_complex(vals), asm.plan.h_cols, asm.plan.h_rowptr),
                      shape=(T.ndof, T.ndof))
    return dict(x=core.as_complex(x), info=info, fields=f, A=A, b=core.as_complex(b))


def _oracle_forward(w):
    T = w.topo
    eo, fo = otopo.orientations(T.elems_nodes, T.elems_edges, T.edges_nodes, T.elem_faces_edges)
    return fem.forward(dict(edge_orient=eo, face_orient=fo, dofs=T.dofs, ndof=T.ndof,
                            boundary_dofs=T.boundary_dofs), T.coords, w.sigma, T.p, w.source,
                       w.receivers)


@pytest.mark.parametrize("p,n", [(1, 10), (2, 6), (3, 4)])
@pytest.mark.parametrize("precision,pc,spmv", [("fp64", "jacobi", "row"),
                                               ("fp64", "bjacobi", "entity"),
                                               ("mixed", "jacobi", "entity"),
                                               ("mixed", "bjacobi", "row"),
                                               ("mixed", "bjacobi", "entity_natural"),
                                               ("fp64", "bjacobi", "element"),
                                               ("mixed", "bjacobi", "element")])
def test_forward_solve_vs_superlu(p, n, precision, pc, spmv):
    w = Workload(n, p, noise=0.1, seed=7, n_rx=21)
    g = _forward_gpu(w, precision=precision, pc=pc, spmv=spmv)
    r = _oracle_forward(w)
    info = g["info"]
    true_rel = np.linalg.norm(g["b"] - g["A"] @ g["x"]) / np.linalg.norm(g["b"])
    assert np.abs(g["b"] - r["bbc"]).max() <= 1e-12 * np.abs(r["bbc"]).max()
    assert np.abs(g["A"].data - r["Abc"].data).max() <= 1e-12 * np.abs(r["Abc"].data).max()
    ferr = np.abs(g["fields"] - r["fields"]).max() / np.abs(r["fields"]).max()
    print(f"p={p} n={n} {precision} {pc} {spmv}: iters={info['iterations']} refinements={info['refinements']} "
          f"relres={info['relres']:.2e} true={true_rel:.2e} field err={ferr:.2e}", flush=True)
    assert info["status"] == 0
    assert info["relres"] <= 1e-8 and true_rel <= 1e-8
    assert ferr <= 1e-6


@pytest.mark.parametrize("n,spmv", [(3, "entity"), (4, "entity"), (4, "element")])
def test_p3_mixed_at_quarter_hertz_vs_superlu(n, spmv):
    """C3's second frequency: p = 3 at 0.25 Hz in mixed precision (complex64
    iterations under FP64 refinement, with the stagnation watch free to hand
    the solve over to FP64) against SuperLU, for the CSR and element-product
    iterations."""
    w = Workload(n, 3, noise=0.1, seed=9, n_rx=21, frequency=0.25)
    g = _forward_gpu(w, precision="mixed", pc="bjacobi", spmv=spmv)
    r = _oracle_forward(w)
    info = g["info"]
    true_rel = np.linalg.norm(g["b"] - g["A"] @ g["x"]) / np.linalg.norm(g["b"])
    ferr = np.abs(g["fields"] - r["fields"]).max() / np.abs(r["fields"]).max()
    xerr = np.abs(g["x"] - r["x"]).max() / np.abs(r["x"]).max()
    print(f"p=3 n={n} 0.25 Hz mixed {spmv}: iters={info['iterations']} "
          f"refinements={info['refinements']} fp64_fallback={info['fp64_fallback']} "
          f"true relres={true_rel:.2e} field err={ferr:.2e} x err={xerr:.2e}", flush=True)
    assert info["status"] == 0 and info["relres"] <= 1e-8 and true_rel <= 1e-8
    assert ferr <= 1e-6


def test_c1_full_size_fp64_and_mixed():
    """C1 (20^3 Kuhn, p=1, 59,660 DOFs) at full size: true residual and FP64/mixed agreement."""
    w = Workload(20, 1, n_rx=41)
    g64 = _forward_gpu(w, "fp64")
    g32 = _forward_gpu(w, "mixed")
    for g in (g64, g32):
        rel = np.linalg.norm(g["b"] - g["A"] @ g["x"]) / np.linalg.norm(g["b"])
        assert rel <= 1e-8 and g["info"]["status"] == 0
    d = np.abs(g64["fields"] - g32["fields"]).max() / np.abs(g64["fields"]).max()
    print(f"C1 iters fp64={g64['info']['iterations']} mixed={g32['info']['iterations']} "
          f"(refinements {g32['info']['refinements']}) field diff {d:.2e}", flush=True)
    assert d <= 1e-6


def test_mixed_stagnation_hands_over_to_fp64():
    """A complex64 round declared stagnant (here after a single iteration without a
    10 % residual drop) hands the solve over to FP64, which still converges."""
    w = Workload(5, 2, noise=0.1, seed=3)
    T = w.topo
    mesh = core.DeviceMesh.from_topology(T, DEV)
    asm = core.Assembler(mesh)
    asm.geometry(w.sigma)
    omega = core.omega_of(w.frequency)
    ent_bc, dof_bc = mesh.entity_mask(T.boundary_dofs)
    vals = asm.assemble(omega, ent_bc=ent_bc)
    el = int(core.locate_points(mesh, np.array(w.source["position"])[None]).cpu()[0])
    b = core.source_vector(mesh, el, w.source["position"], core.dipole_moment(w.source), omega)
    core.apply_dirichlet(asm.plan, None, dof_bc, b)
    ref, _ = core.KrylovSolver(asm.plan, vals, rtol=1e-10).solve(b)
    S = core.KrylovSolver(asm.plan, vals, precision="mixed", rtol=1e-8, stall_iters=1,
                          pc="bjacobi", block_ptr=mesh.block_ptr(),
                          entity_order=mesh.entity_order())
    x, info = S.solve(b)
    print(f"stagnation hand-over: {info}", flush=True)
    assert info["fp64_fallback"] == 1 and info["status"] == 0 and info["relres"] <= 1e-8
    xr, xh = core.as_complex(ref), core.as_complex(x)
    assert np.abs(xh - xr).max() <= 1e-6 * np.abs(xr).max()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_column_pair_spmv_matches_entity_rows(precision):
    """p = 2 (two DOFs per entity): the column-pair entity-row SpMV gives the
    solution of the scalar entity-row SpMV and of SuperLU."""
    w = Workload(6, 2, noise=0.1, seed=11, n_rx=21)
    g2 = _forward_gpu(w, precision=precision, pc="bjacobi", spmv="entity_pairs")
    g1 = _forward_gpu(w, precision=precision, pc="bjacobi", spmv="entity")
    r = _oracle_forward(w)
    for g in (g1, g2):
        assert g["info"]["status"] == 0 and g["info"]["relres"] <= 1e-8
    ferr = np.abs(g2["fields"] - r["fields"]).max() / np.abs(r["fields"]).max()
    assert ferr <= 1e-6
    assert np.abs(g2["x"] - g1["x"]).max() <= 1e-6 * np.abs(g1["x"]).max()
