"""Drop-in ``Solver``: the reference's setup / assembly / run on the MI355X path.

Same call sequence and timer names as ``petgem/solver.py:30-598``
(``kernel.py:64-76``: ``Solver(); setup(); assembly(); run(); del``), same
scratch-file inputs and ``x{i}.dat`` output, CSEM mode:

* ``setup`` -- reads the scratch files (``solver.py:39-128``) and moves the
  per-element arrays into HBM.
* ``assembly`` -- element geometry/orientation, fused atomic-free CSR assembly
  of A = K - i w mu0 M and the point-dipole source (``solver.py:131-316``),
  all in HIP kernels.  The integer assembly plan (CSR pattern + per-block
  contributor lists) is built on the host once per mesh, like the
  reference's preallocation + MatSetValues insertion.
* ``run`` -- PEC boundary condition (``solver.py:552-577``) and the Krylov solve
  (``solver.py:584-590``) with the options of the PETSc options file
  (``-ksp_type``, ``-pc_type``, ``-ksp_rtol``...), then ``x0.dat``.

There is no CPU fallback: without the HIP library every method raises.
"""
import os
import sys

import numpy as np

from . import _native as nat  # noqa: F401  (fails loudly when the library is missing)
from . import core, petsc_io
from .common import Print, SolverOptions, Timers, stop_timer
from .mesh import dof_counts


def _ints(a):
    return np.rint(np.real(a)).astype(np.int64)


class Solver:
    """Class for solver (CSEM mode, one polarization)."""

    def __init__(self, options=None):
        # without explicit options the PETSc-style flags come from the command
        # line, -options_file included, as the reference's ksp.setFromOptions()
        # reads them (solver.py:588)
        self.options = options if options is not None else SolverOptions(sys.argv[1:])
        self.info = None

    # ------------------------------------------------------------------ setup
    def setup(self, inputSetup):
        Timers()["Setup"].start()
        out = inputSetup.output["directory_scratch"]
        Print.master("     Importing files")
        rd = lambda name: petsc_io.read_dense(os.path.join(out, name))  # noqa: E731
        coords = np.real(rd("nodes.dat"))
        en = _ints(rd("meshConnectivity.dat"))
        ee = _ints(rd("edges.dat"))
        edn = _ints(rd("edgesNodes.dat"))
        ef = _ints(rd("faces.dat"))
        fe = _ints(rd("facesEdges.dat"))
        dofs = _ints(rd("dofs.dat"))
        self.sigma = np.ascontiguousarray(np.real(rd("conductivityModel.dat")))
        nnz = petsc_io.read_vec(os.path.join(out, "nnz.dat"))
        self.total_num_dofs = nnz.size
        self.boundaries = _ints(petsc_io.read_vec(os.path.join(out, "boundaries.dat")))
        self.source_data = np.real(petsc_io.read_vec(os.path.join(out, "source.dat")))
        p = int(inputSetup.run["nord"])
        E = en.shape[0]
        de, df, dv, n = dof_counts(p)
        if dofs.shape != (E, n):
            raise ValueError(f"dofs.dat has shape {dofs.shape}, expected ({E},{n}) for nord={p}")
        n_edges = int(ee.max()) + 1
        n_faces = int(ef.max()) + 1
        self.mesh = core.DeviceMesh(p, coords.reshape(E, 4, 3), en, ee, edn.reshape(E, 6, 2),
                                    fe.reshape(E, 4, 3), dofs, n_edges, n_faces, ef)
        if self.mesh.ndof != self.total_num_dofs:
            raise ValueError(f"nnz.dat length {self.total_num_dofs} != {self.mesh.ndof} DOFs")
        stop_timer("Setup")

    # --------------------------------------------------------------- assembly
    def assembly(self, inputSetup):
        import torch
        Timers()["Assembly"].start()
        Print.master("     Assembling linear system")
        src = inputSetup.model["csem"]["source"]
        self.omega = core.omega_of(src["frequency"])
        m = self.mesh
        self.assembler = core.Assembler(m)
        self.assembler.geometry(self.sigma)
        self.A = self.assembler.assemble(self.omega)
        # source: evaluated on the source-element record alone (solver.py:269-316)
        sd = self.source_data
        n = m.n_dof_elem
        rec = core.DeviceMesh(m.p, sd[4:16].reshape(1, 4, 3), _ints(sd[0:4])[None],
                              _ints(sd[32:38])[None], _ints(sd[38:50]).reshape(1, 6, 2),
                              _ints(sd[20:32]).reshape(1, 4, 3), _ints(sd[50:50 + n])[None],
                              m.n_edges, m.n_faces, _ints(sd[16:20])[None], m.device)
        rec.ndof = m.ndof
        self.b = torch.zeros((m.ndof, 2), dtype=torch.float64, device=m.device)
        core.source_vector(rec, 0, src["position"], core.dipole_moment(src), self.omega, b=self.b)
        self.x = torch.zeros_like(self.b)
        stop_timer("Assembly")

    # -------------------------------------------------------------------- run
    def run(self, inputSetup):
        out = inputSetup.output["directory_scratch"]
        Print.master("     Solving linear system")
        Timers()["SetBoundaries"].start()
        ent_bc, dof_bc = self.mesh.entity_mask(self.boundaries)
        core.apply_dirichlet(self.assembler.plan, self.A, dof_bc, self.b)
        stop_timer("SetBoundaries")
        Timers()["Solver"].start()
        o = self.options
        for note in o.notes:
            Print.master(f"     solver option mapped: {note}")
        order = self.mesh.entity_order()
        # element-by-element products (PEC on whole entities, as the reference's
        # boundary lists are; complex64 only at p = 3): used when the autotuner
        # finds them faster
        bp = self.mesh.block_ptr()
        whole = np.array_equal(np.repeat(ent_bc.cpu().numpy()[:bp.size - 1], np.diff(bp)),
                               dof_bc.cpu().numpy())
        eop = (core.ElementOperator(self.assembler, self.omega, ent_bc, order)
               if whole and o.spmv in ("auto", "element") else None)
        if o.spmv == "element" and eop is None:
            raise ValueError("-pg_spmv element needs the PEC condition on whole entities")
        solver = core.KrylovSolver(self.assembler.plan, self.A, ksp=o.ksp_type, pc=o.pc_type,
                                   precision=o.precision, rtol=o.rtol, maxit=o.max_it,
                                   restart=o.restart, inner_rtol=o.inner_rtol,
                                   block_ptr=bp, entity_order=order, eop=eop, spmv=o.spmv)
        self.x, self.info = solver.solve(self.b, self.x)
        stop_timer("Solver")
        Print.master(f"     {o.ksp_type}/{o.pc_type} ({o.precision}): {self.info['iterations']} "
                     f"iterations, relative residual {self.info['relres']:.3e}")
        if self.info["status"] != 0:
            Print.master("     WARNING: the Krylov solve did not converge to the requested tolerance")
        petsc_io.write_vec(os.path.join(out, "x0.dat"), core.as_complex(self.x))
