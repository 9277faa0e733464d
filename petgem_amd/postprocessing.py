"""Postprocessing: receiver fields and the HDF5 results file.

Restates ``Postprocessing.run`` + ``fieldInterpolator``
(``petgem/postprocessing.py:42-476,479-616``) for CSEM mode: the receivers
are located (lowest-index element within barycentric tolerance 1e-12;
receivers outside the mesh are dropped with a message), E = sum x_j N_j and
H = sum x_j curl N_j / (i w mu0) are evaluated on the device, and the
results file ``<directory>/csem_petgemV<version>_<YYYY-MM-DD>.h5`` holds the
reference's groups and keys (``machine/...``, ``model/...``,
``model/E-fields/{x,y,z}``, ``model/H-fields/{x,y,z}``,
``model/run-time (s)`` = Assembly + Solver wall time), plus the solver's
iteration count, true relative residual and precision mode.
VTK output and the MT impedance post-processing are out of scope.
"""
import datetime
import os
import platform

import numpy as np

from . import core, h5, petsc_io
from .common import Print, Timers
from .mesh import Topology, read_gmsh_v2

CODE_VERSION = "1.0"  # the reference version whose file layout is reproduced


def field_interpolator(mesh, x, receivers, frequency):
    """(nRx_found,6) complex fields and the kept receiver indices."""
    rx = np.atleast_2d(np.asarray(receivers, dtype=np.float64))
    el = core.locate_points(mesh, rx).cpu().numpy()
    keep = np.nonzero((el >= 0) & (el < mesh.n_elems))[0]
    if keep.size != rx.shape[0]:
        Print.master("        The following receivers were not located and will not be taken "
                     f"into account {np.setdiff1d(np.arange(rx.shape[0]), keep)}")
    if keep.size == 0:
        raise ValueError("No point has been found. Nothing to do. Aborting")
    f = core.receiver_fields(mesh, rx[keep], el[keep].astype(np.int32), x,
                             core.omega_of(frequency))
    return core.as_complex(f), keep


class Postprocessing:
    def run(self, inputSetup, solver=None):
        Timers()["Postprocessing"].start()
        model, run, output = inputSetup.model, inputSetup.run, inputSetup.output
        src = model["csem"]["source"]
        p = int(run["nord"])
        if solver is not None:
            mesh, x = solver.mesh, solver.x
        else:
            import torch
            pts, en, _ = read_gmsh_v2(model["mesh"])
            topo = Topology(pts, en, p)
            mesh = core.DeviceMesh.from_topology(topo)
            xs = petsc_io.read_vec(os.path.join(output["directory_scratch"], "x0.dat"))
            x = core.from_complex(xs, torch.device("cuda"))
        receivers = h5.read_dataset(model["receivers"], "data")
        fields, keep = field_interpolator(mesh, x, receivers, src["frequency"])
        date = datetime.date.today().strftime("%Y-%m-%d")
        path = os.path.join(output["directory"], f"csem_petgemV{CODE_VERSION}_{date}.h5")
        un = platform.uname()
        with h5.Writer(path) as w:
            w.dataset("machine/machine", f"{un.node}. {un.release}. {un.processor}")
            w.dataset("machine/num_processors", 1)
            w.dataset("machine/petgem_version", CODE_VERSION)
            w.dataset("machine/backend", "petgem_amd (HIP, gfx950)")
            w.dataset("model/date", datetime.datetime.today().isoformat())
            w.dataset("model/mesh_file", str(model["mesh"]))
            w.dataset("model/receivers_file", str(model["receivers"]))
            w.dataset("model/nord", p)
            w.dataset("model/dof", int(mesh.ndof))
            w.dataset("model/cuda", bool(run.get("cuda", False)))
            w.dataset("model/vtk", bool(output.get("vtk", False)))
            w.dataset("model/mode", "csem")
            w.dataset("model/num-polarizations", 1)
            if solver is not None:
                w.dataset("model/solver", solver.options.ksp_type)
                w.dataset("model/preconditioner", solver.options.pc_type)
                w.dataset("model/precision", solver.options.precision)
                w.dataset("model/iterations", int(solver.info["iterations"]))
                w.dataset("model/relative-residual", float(solver.info["relres"]))
            w.dataset("model/run-time (s)", Timers()["Assembly"].elapsed + Timers()["Solver"].elapsed)
            sig = model["csem"]["sigma"]
            if run.get("conductivity_from_file"):
                w.dataset("model/sigma-file", str(sig["file"]))
            else:
                w.dataset("model/sigma_horizontal (S/m)", np.asarray(sig["horizontal"], dtype=float))
                w.dataset("model/sigma_vertical (S/m)", np.asarray(sig["vertical"], dtype=float))
            w.dataset("model/frequency (Hz)", float(src["frequency"]))
            w.dataset("model/source_position (m)", np.asarray(src["position"], dtype=float))
            w.dataset("model/source_azimuth (deg)", float(src["azimuth"]))
            w.dataset("model/source_dip (deg)", float(src["dip"]))
            w.dataset("model/source_current (Am)", float(src["current"]))
            w.dataset("model/source_length (m)", float(src["length"]))
            w.dataset("model/receivers_kept", keep.astype(np.int64))
            for k, c in enumerate("xyz"):
                w.dataset(f"model/E-fields/{c}", fields[:, k])
                w.dataset(f"model/H-fields/{c}", fields[:, 3 + k])
        if output.get("remove_scratch"):
            d = output["directory_scratch"]
            for name in os.listdir(d):
                if name.endswith((".dat", ".info")):
                    os.remove(os.path.join(d, name))
        Timers()["Postprocessing"].stop()
        Print.master(f"     Results written to {path}")
        self.path, self.fields = path, fields
        return path
