"""Multi-mesh batch (uneven meshes, one launch set per iteration, HIP-graph
replay) against per-mesh single solves, and graph replays against eager launches."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from petgem_amd import core  # noqa: E402
from petgem_amd.synthetic import Workload  # noqa: E402
from oracle import fem, topology as otopo  # noqa: E402


def _system(n, p, freq, seed):
    w = Workload(n, p, noise=0.2, seed=seed)
    T = w.topo
    mesh = core.DeviceMesh.from_topology(T, "cuda")
    asm = core.Assembler(mesh)
    asm.geometry(w.sigma)
    ent_bc, dof_bc = mesh.entity_mask(T.boundary_dofs)
    om = core.omega_of(freq)
    vals = asm.assemble(om, ent_bc=ent_bc)
    el = int(core.locate_points(mesh, np.array([[0.0, 0.0, -950.0]])).cpu()[0])
    b = core.source_vector(mesh, el, [0.0, 0.0, -950.0], (1.0, 0.0, 0.0), om)
    core.apply_dirichlet(asm.plan, None, dof_bc, b)
    return asm.plan, vals, b


CASES = [(3, 1, 0.5), (4, 2, 1.0), (5, 2, 0.25), (3, 3, 1.0), (6, 1, 2.0), (2, 2, 1.0)]


def test_multi_mesh_matches_single_solves():
    systems = [_system(n, p, f, seed=i) for i, (n, p, f) in enumerate(CASES)]
    ms = core.MultiMeshSolver([(pl, v) for pl, v, _ in systems], rtol=1e-10, check_every=10)
    xs, info = ms.solve([b for _, _, b in systems])
    print(f"multi-mesh iterations {info['iterations']} relres max {info['relres'].max():.2e} "
          f"graph launches {info['launches']}", flush=True)
    assert (info["status"] == 0).all() and info["relres"].max() <= 1e-10
    for (pl, v, b), x in zip(systems, xs):
        x1, inf1 = core.KrylovSolver(pl, v, rtol=1e-10).solve(b)
        xh, x1h = core.as_complex(x), core.as_complex(x1)
        assert np.abs(xh - x1h).max() <= 1e-7 * np.abs(x1h).max()


def test_multi_mesh_matches_oracle():
    """Three meshes of the batch (p = 1, 2, 3, different frequencies) against
    the oracle's SuperLU solution of each mesh's system: solution and receiver
    fields within 1e-6."""
    picks = [0, 2, 3]
    ws, systems = [], []
    for i in picks:
        n, p, f = CASES[i]
        w = Workload(n, p, noise=0.2, seed=i, n_rx=21, frequency=f)
        ws.append(w)
        systems.append(_system(n, p, f, seed=i))
    ms = core.MultiMeshSolver([(pl, v) for pl, v, _ in systems], rtol=1e-10, check_every=10)
    xs, info = ms.solve([b for _, _, b in systems])
    assert (info["status"] == 0).all() and info["relres"].max() <= 1e-10
    for w, x in zip(ws, xs):
        T = w.topo
        eo, fo = otopo.orientations(T.elems_nodes, T.elems_edges, T.edges_nodes,
                                    T.elem_faces_edges)
        ref = fem.forward(dict(edge_orient=eo, face_orient=fo, dofs=T.dofs, ndof=T.ndof,
                               boundary_dofs=T.boundary_dofs), T.coords, w.sigma, T.p,
                          dict(w.source, position=[0.0, 0.0, -950.0]), w.receivers)
        mesh = core.DeviceMesh.from_topology(T, "cuda")
        rx_el = core.locate_points(mesh, w.receivers)
        fields = core.as_complex(core.receiver_fields(mesh, w.receivers, rx_el, x,
                                                      core.omega_of(w.frequency)))
        xerr = np.abs(core.as_complex(x) - ref["x"]).max() / np.abs(ref["x"]).max()
        ferr = np.abs(fields - ref["fields"]).max() / np.abs(ref["fields"]).max()
        print(f"multi-mesh p={T.p} n={w.n} {w.frequency} Hz: x err {xerr:.2e}, field err "
              f"{ferr:.2e}", flush=True)
        assert xerr <= 1e-6 and ferr <= 1e-6


def test_graph_replays_match_eager_launches():
    systems = [_system(n, p, f, seed=10 + i) for i, (n, p, f) in enumerate(CASES[:4])]
    pairs = [(pl, v) for pl, v, _ in systems]
    eager = core.MultiMeshSolver(pairs, rtol=1e-9, check_every=7, use_graph=False)
    graph = core.MultiMeshSolver(pairs, rtol=1e-9, check_every=7, use_graph=True)
    for rep in range(3):
        bs = [(1.0 + rep) * b for _, _, b in systems]
        xe, ie = eager.solve(bs)
        torch.cuda.synchronize()
        xg, ig = graph.solve(bs)
        assert (ig["status"] == 0).all()
        assert np.array_equal(ie["iterations"], ig["iterations"])
        for a, c in zip(xe, xg):
            assert torch.equal(a, c), f"replay {rep}: graph and eager results differ"


def test_solve_each_on_streams_is_bitwise_per_mesh():
    """core.solve_each (several meshes at a time on several streams, each mesh
    with its own solver and element operator) gives bitwise each solver's own
    solution."""
    solvers, bs = [], []
    for i, (n, p, f) in enumerate(CASES):
        w = Workload(n, p, noise=0.2, seed=i)
        T = w.topo
        mesh = core.DeviceMesh.from_topology(T, "cuda")
        asm = core.Assembler(mesh)
        asm.geometry(w.sigma)
        ent_bc, dof_bc = mesh.entity_mask(T.boundary_dofs)
        om = core.omega_of(f)
        vals = asm.assemble(om, ent_bc=ent_bc)
        el = int(core.locate_points(mesh, np.array([[0.0, 0.0, -950.0]])).cpu()[0])
        b = core.source_vector(mesh, el, [0.0, 0.0, -950.0], (1.0, 0.0, 0.0), om)
        core.apply_dirichlet(asm.plan, None, dof_bc, b)
        order = mesh.entity_order()
        eop = core.ElementOperator(asm, om, ent_bc, order)
        solvers.append(core.KrylovSolver(asm.plan, vals, precision="mixed", pc="bjacobi",
                                         rtol=1e-8, block_ptr=mesh.block_ptr(),
                                         entity_order=order, eop=eop))
        bs.append(b)
    xs, infos = core.solve_each(solvers, bs, workers=3)
    for S, b, x, info in zip(solvers, bs, xs, infos):
        assert info["status"] == 0 and info["relres"] <= 1e-8
        x1, _ = S.solve(b)
        assert torch.equal(x, x1)


# ---------------------------------------------------------------------------
# element products + entity-block Jacobi (pg_eop_create_multi)
# ---------------------------------------------------------------------------
def _system_e(n, p, freq, seed, n_rx=0):
    w = Workload(n, p, noise=0.2, seed=seed, n_rx=max(n_rx, 1), frequency=freq)
    T = w.topo
    mesh = core.DeviceMesh.from_topology(T, "cuda")
    asm = core.Assembler(mesh)
    asm.geometry(w.sigma)
    ent_bc, dof_bc = mesh.entity_mask(T.boundary_dofs)
    om = core.omega_of(freq)
    vals = asm.assemble(om, ent_bc=ent_bc)
    el = int(core.locate_points(mesh, np.array([[0.0, 0.0, -950.0]])).cpu()[0])
    b = core.source_vector(mesh, el, [0.0, 0.0, -950.0], (1.0, 0.0, 0.0), om)
    core.apply_dirichlet(asm.plan, None, dof_bc, b)
    return dict(w=w, mesh=mesh, asm=asm, plan=asm.plan, vals=vals, b=b, om=om, ent_bc=ent_bc,
                order=mesh.entity_order())


E_CASES = {1: [(4, 0.5), (6, 2.0), (3, 1.0)], 2: [(4, 1.0), (2, 0.25), (5, 0.5)],
           3: [(3, 1.0), (2, 0.5)]}


def _element_batch(syss, rtol, check_every=10, use_graph=True, precision="fp64", **kw):
    return core.MultiMeshSolver([(s["plan"], s["vals"]) for s in syss], rtol=rtol,
                                check_every=check_every, use_graph=use_graph,
                                element=[(s["asm"], s["om"], s["ent_bc"], s["order"])
                                         for s in syss], precision=precision, **kw)


@pytest.mark.parametrize("p,precision", [(1, "fp64"), (2, "fp64"), (3, "fp64"), (1, "mixed"),
                                         (2, "mixed"), (3, "mixed")])
def test_multi_mesh_element_matches_oracle(p, precision):
    """Meshes of one order p (sizes and frequencies differ) in one batch on the
    multi-mesh element operator with entity-block Jacobi, in FP64 and in
    complex64 rounds under FP64 refreshes: every mesh's solution and receiver
    fields within 1e-6 of the oracle's SuperLU solution of its system."""
    syss = [_system_e(n, p, f, seed=20 + i, n_rx=21) for i, (n, f) in enumerate(E_CASES[p])]
    print(f"multi-mesh element p={p} {precision}: meshes {[s['mesh'].n_elems for s in syss]}",
          flush=True)
    ms = _element_batch(syss, rtol=1e-10, precision=precision)
    xs, info = ms.solve([s["b"] for s in syss])
    print(f"  iterations {info['iterations']} relres max {info['relres'].max():.2e}", flush=True)
    assert (info["status"] == 0).all() and info["relres"].max() <= 1e-10
    for s, x in zip(syss, xs):
        w = s["w"]
        T = w.topo
        eo, fo = otopo.orientations(T.elems_nodes, T.elems_edges, T.edges_nodes,
                                    T.elem_faces_edges)
        ref = fem.forward(dict(edge_orient=eo, face_orient=fo, dofs=T.dofs, ndof=T.ndof,
                               boundary_dofs=T.boundary_dofs), T.coords, w.sigma, T.p,
                          dict(w.source, position=[0.0, 0.0, -950.0]), w.receivers)
        rx_el = core.locate_points(s["mesh"], w.receivers)
        fields = core.as_complex(core.receiver_fields(s["mesh"], w.receivers, rx_el, x, s["om"]))
        xerr = np.abs(core.as_complex(x) - ref["x"]).max() / np.abs(ref["x"]).max()
        ferr = np.abs(fields - ref["fields"]).max() / np.abs(ref["fields"]).max()
        print(f"  n={w.n} {w.frequency} Hz: x err {xerr:.2e}, field err {ferr:.2e}", flush=True)
        assert xerr <= 1e-6 and ferr <= 1e-6


def test_multi_mesh_element_matches_single_solver_and_replays():
    """The element-operator batch agrees with each mesh's single-RHS solver on
    the same method (FP64, entity-block Jacobi, element products) to 1e-7, its
    graph replays are bitwise its eager launches, and a second solve with other
    right-hand sides starts afresh."""
    syss = [_system_e(n, 2, f, seed=40 + i) for i, (n, f) in
            enumerate([(5, 1.0), (3, 0.25), (6, 0.5), (4, 2.0)])]
    graph = _element_batch(syss, rtol=1e-9, check_every=7)
    eager = _element_batch(syss, rtol=1e-9, check_every=7, use_graph=False)
    for rep in range(2):
        bs = [(1.0 + rep) * s["b"] for s in syss]
        xg, ig = graph.solve(bs)
        torch.cuda.synchronize()
        xe, ie = eager.solve(bs)
        assert (ig["status"] == 0).all() and ig["relres"].max() <= 1e-9
        assert np.array_equal(ie["iterations"], ig["iterations"])
        for a, c in zip(xe, xg):
            assert torch.equal(a, c), f"solve {rep}: graph and eager results differ"
    for s, x in zip(syss, xg):
        eop = core.ElementOperator(s["asm"], s["om"], s["ent_bc"], s["order"])
        S = core.KrylovSolver(s["plan"], s["vals"], precision="fp64", pc="bjacobi", rtol=1e-9,
                              block_ptr=s["mesh"].block_ptr(), entity_order=s["order"], eop=eop,
                              spmv="element")
        x1, inf1 = S.solve(2.0 * s["b"])
        xh, x1h = core.as_complex(x), core.as_complex(x1)
        assert np.abs(xh - x1h).max() <= 1e-7 * np.abs(x1h).max()


def test_multi_mesh_mixed_rounds_and_fp64_hand_over():
    """Mixed precision on the multi-mesh element operator: the solve is
    repeatable bitwise (graph replays of the complex64 rounds included), a
    stall limit small enough to stop every complex64 round early still
    converges (meshes refreshed or handed to FP64), and both agree with the
    FP64 batch to the tolerance."""
    syss = [_system_e(n, 2, f, seed=60 + i) for i, (n, f) in
            enumerate([(5, 0.1), (3, 1.0), (6, 0.25), (4, 2.0)])]
    bs = [s["b"] for s in syss]
    ref, iref = _element_batch(syss, rtol=1e-10).solve(bs)
    mx = _element_batch(syss, rtol=1e-10, check_every=7, precision="mixed")
    x1, i1 = mx.solve(bs)
    torch.cuda.synchronize()
    x2, i2 = mx.solve(bs)
    assert (i1["status"] == 0).all() and i1["relres"].max() <= 1e-10
    assert np.array_equal(i1["iterations"], i2["iterations"])
    for a, c in zip(x1, x2):
        assert torch.equal(a, c)
    tight = _element_batch(syss, rtol=1e-10, check_every=5, precision="mixed", stall_iters=3,
                           use_graph=False)
    x3, i3 = tight.solve(bs)
    print(f"mixed iterations {i1['iterations']}, with stall limit 3 {i3['iterations']}, "
          f"FP64 {iref['iterations']}", flush=True)
    assert (i3["status"] == 0).all() and i3["relres"].max() <= 1e-10
    for xs in (x1, x3):
        for a, r in zip(xs, ref):
            ah, rh = core.as_complex(a), core.as_complex(r)
            assert np.abs(ah - rh).max() <= 1e-7 * np.abs(rh).max()


def test_multi_mesh_lanes_are_bitwise_the_single_batch():
    """Meshes dealt to several concurrent batches (lanes, own streams) give
    bitwise the results of one batch over all of them: element products in
    FP64 and mixed, and the CSR batch."""
    syss = [_system_e(n, 2, f, seed=80 + i) for i, (n, f) in
            enumerate([(5, 1.0), (3, 0.25), (6, 0.5), (4, 2.0), (2, 1.0)])]
    bs = [s["b"] for s in syss]
    for prec in ("fp64", "mixed"):
        x1, i1 = _element_batch(syss, rtol=1e-9, precision=prec).solve(bs)
        x3, i3 = _element_batch(syss, rtol=1e-9, precision=prec, lanes=3).solve(bs)
        assert (i3["status"] == 0).all() and np.array_equal(i1["iterations"], i3["iterations"])
        for a, c in zip(x1, x3):
            assert torch.equal(a, c), f"{prec}: lanes change a mesh's result"
    pairs = [(s["plan"], s["vals"]) for s in syss]
    x1, i1 = core.MultiMeshSolver(pairs, rtol=1e-9).solve(bs)
    x2, i2 = core.MultiMeshSolver(pairs, rtol=1e-9, lanes=2).solve(bs)
    assert np.array_equal(i1["iterations"], i2["iterations"])
    for a, c in zip(x1, x2):
        assert torch.equal(a, c)
