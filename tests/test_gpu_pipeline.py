"""GPU tests of the drop-in boundary: GMRES on the reference's own assembled
system, and the whole ``kernel`` pipeline (params file -> scratch files ->
Solver.setup/assembly/run -> HDF5 fields) against the oracle."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from petgem_amd import core, h5, petsc_io  # noqa: E402
from oracle import fem, topology as otopo  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DEV = "cuda"


class _CSR:
    """Minimal plan-like view (ndof, nnz, rowptr, cols) of an arbitrary CSR matrix."""

    def __init__(self, A):
        self.ndof, self.nnz = A.shape[0], A.nnz
        self.rowptr = torch.from_numpy(A.indptr.astype(np.int32)).to(DEV)
        self.cols = torch.from_numpy(A.indices.astype(np.int32)).to(DEV)


@pytest.mark.parametrize("ksp,restart", [("gmres", 30), ("gmres", 60)])
def test_gmres_on_reference_matrix_fixture(ksp, restart):
    """tests/data/matrix-A.dat + vector-b.dat of the reference (4184 DOFs, p=1, BC rows)."""
    ip, ix, d, sh = petsc_io.read_mat(os.path.join(GOLDEN, "matrix-A.dat"))
    b = petsc_io.read_vec(os.path.join(GOLDEN, "vector-b.dat"))
    A = sp.csr_matrix((d, ix, ip), shape=sh)
    A.sort_indices()
    xs = spla.spsolve(A.tocsc(), b)
    csr = _CSR(A)
    vals = core.from_complex(A.data, DEV)
    S = core.KrylovSolver(csr, vals, ksp=ksp, pc="jacobi", rtol=1e-10, restart=restart,
                          maxit=20000)
    x, info = S.solve(core.from_complex(b, DEV))
    xg = core.as_complex(x)
    rel = np.linalg.norm(b - A @ xg) / np.linalg.norm(b)
    print(f"GMRES({restart}) on matrix-A.dat: {info}, true relres {rel:.2e}", flush=True)
    assert info["status"] == 0 and rel <= 1e-10
    assert np.abs(xg - xs).max() <= 1e-7 * np.abs(xs).max()


def _write_case(tmp, p, ksp="gmres", pc="sor", precision=None, spmv=None):
    rx = np.load(os.path.join(GOLDEN, "receiver_pos.npy"))
    h5.write_data(os.path.join(tmp, "receiver_pos.h5"), rx)
    params = f"""
model:
  mode: csem
  csem:
    sigma:
      horizontal: [1., 0.01, 1., 3.3333]
      vertical: [1., 0.01, 1., 3.3333]
    source:
      frequency: 2.
      position: [1750., 1750., -975.]
      azimuth: 0.
      dip: 0.
      current: 1.
      length: 1.
  mesh: {os.path.join(GOLDEN, 'test_mesh.msh')}
  receivers: {os.path.join(tmp, 'receiver_pos.h5')}
run:
  nord: {p}
  cuda: True
output:
  vtk: False
  directory: {os.path.join(tmp, 'out')}
"""
    with open(os.path.join(tmp, "params.yaml"), "w") as f:
        f.write(params)
    with open(os.path.join(tmp, "petsc.opts"), "w") as f:
        f.write(f"# Solver options for PETSc\n-ksp_type {ksp}\n-pc_type {pc}\n-ksp_rtol 1e-10\n")
        if precision:
            f.write(f"-pg_precision {precision}\n")
        if spmv:
            f.write(f"-pg_spmv {spmv}\n")
    return os.path.join(tmp, "params.yaml"), os.path.join(tmp, "petsc.opts")


@pytest.mark.parametrize("p,ksp,precision,spmv", [(1, "gmres", None, None),
                                                  (1, "cocg", "mixed", None),
                                                  (2, "cocg", None, None),
                                                  (2, "cocg", "mixed", None),
                                                  (2, "cocg", "mixed", "element")])
def test_kernel_pipeline_vs_oracle(tmp_path, p, ksp, precision, spmv):
    from petgem_amd import kernel
    from petgem_amd.common import Timers
    from petgem_amd.mesh import read_gmsh_v2, Topology
    Timers().reset()
    params, opts = _write_case(str(tmp_path), p, ksp=ksp, precision=precision, spmv=spmv)
    path = kernel.main(["kernel.py", "-options_file", opts, params])
    assert os.path.exists(path)
    out = os.path.join(str(tmp_path), "out")
    for name in ("nodes.dat", "dofs.dat", "boundaries.dat", "source.dat", "x0.dat"):
        assert os.path.exists(os.path.join(out, name)), name
    # the reference's run-time key and the E/H groups exist; fields match the oracle
    assert h5.read_dataset(path, "model/run-time (s)") > 0
    pts, en, phys = read_gmsh_v2(os.path.join(GOLDEN, "test_mesh.msh"))
    T = Topology(pts, en, p)
    eo, fo = otopo.orientations(en, T.elems_edges, T.edges_nodes, T.elem_faces_edges)
    sig = np.stack([np.array([1.0, 0.01, 1.0, 3.3333])[phys - 1]] * 2, axis=1)
    src = dict(position=[1750.0, 1750.0, -975.0], azimuth=0.0, dip=0.0, current=1.0,
               length=1.0, frequency=2.0)
    rx = np.load(os.path.join(GOLDEN, "receiver_pos.npy"))
    x = petsc_io.read_vec(os.path.join(out, "x0.dat"))
    if p == 1:
        ref = fem.forward(dict(edge_orient=eo, face_orient=fo, dofs=T.dofs, ndof=T.ndof,
                               boundary_dofs=T.boundary_dofs), T.coords, sig, p, src, rx)
        assert np.abs(x - ref["x"]).max() <= 1e-6 * np.abs(ref["x"]).max()
        fref = ref["fields"]
    else:  # p = 2: SuperLU is slow on this mesh; check the residual of the written solution
        M, K = fem.all_element_matrices(T.coords, eo, fo, sig, p)
        A = fem.assemble(T.dofs, M, K, 2.0, T.ndof)
        el = int(fem.locate(T.coords, np.array([src["position"]]))[0])
        bb = fem.source_vector(T.coords, T.dofs, eo, fo, p, T.ndof, src, el)
        Abc, bbc = fem.apply_dirichlet(A, bb, T.boundary_dofs)
        assert np.linalg.norm(bbc - Abc @ x) / np.linalg.norm(bbc) <= 1e-9
        rx_el = fem.locate(T.coords, rx)
        fref = fem.receiver_fields(x, T.coords, T.dofs, eo, fo, p, 2.0, rx, rx_el)
    from petgem_amd.postprocessing import field_interpolator
    mesh = core.DeviceMesh.from_topology(T)
    f, keep = field_interpolator(mesh, core.from_complex(x, DEV), rx, 2.0)
    assert keep.size == rx.shape[0]
    assert np.abs(f - fref).max() <= 1e-6 * np.abs(fref).max()


class _RefTimer:
    """The reference's Timer (common.py:391-419): stop() takes no argument."""

    def __init__(self):
        self._start, self.elapsed = 0.0, 0.0

    def start(self):
        import time
        self._start = time.time()

    def stop(self):
        import time
        if self._start > 0:
            self.elapsed += time.time() - self._start
            self._start = 0.0


class _RefTimers:
    def __init__(self):
        self._elems = {}

    def __getitem__(self, key):
        return self._elems.setdefault(key, _RefTimer())


def test_reference_driver_with_only_the_solver_import_changed(tmp_path, monkeypatch):
    """The reference's kernel.py sequence (kernel.py:22-88) with one line
    changed, ``from petgem_amd.solver import Solver``: ``Solver()`` takes the
    options file from the command line as ``ksp.setFromOptions()`` does
    (solver.py:588), and the reference's own ``Timers()`` singleton -- a
    stand-in ``petgem.common`` module here, loaded first as the reference's
    kernel.py loads it -- receives the Assembly and Solver times that the
    reference's postprocessing adds up (postprocessing.py:381-383)."""
    import sys
    import types
    inst = _RefTimers()
    common = types.ModuleType("petgem.common")
    common.Timers = lambda opath=None: inst
    pkg = types.ModuleType("petgem")
    pkg.common = common
    monkeypatch.setitem(sys.modules, "petgem", pkg)
    monkeypatch.setitem(sys.modules, "petgem.common", common)
    params, opts = _write_case(str(tmp_path), 2, ksp="cocg", pc="jacobi", precision="mixed")
    monkeypatch.setattr(sys, "argv", ["kernel.py", "-options_file", opts, params])

    from petgem_amd.common import InputParameters
    from petgem_amd.postprocessing import Postprocessing
    from petgem_amd.preprocessing import Preprocessing
    from petgem_amd.solver import Solver  # the changed import
    input_setup = InputParameters(sys.argv[3])
    Preprocessing().run(input_setup)
    solver = Solver()
    solver.setup(input_setup)
    solver.assembly(input_setup)
    solver.run(input_setup)
    o = solver.options
    assert (o.ksp_type, o.pc_type, o.precision, o.rtol) == ("cocg", "jacobi", "mixed", 1e-10)
    assert solver.info["status"] == 0 and solver.info["relres"] <= 1e-10
    assert solver.info["refinements"] >= 1  # complex64 inner solves under FP64 refinement
    del solver
    path = Postprocessing().run(input_setup)
    assert inst["Assembly"].elapsed > 0 and inst["Solver"].elapsed > 0
    rt = h5.read_dataset(path, "model/run-time (s)")
    assert rt == pytest.approx(inst["Assembly"].elapsed + inst["Solver"].elapsed)
