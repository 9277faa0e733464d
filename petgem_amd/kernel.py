"""petgem driver on the MI355X path (the reference's ``kernel.py``).

    python -m petgem_amd.kernel -options_file petsc.opts params.yaml

Same pipeline as ``kernel.py:15-88`` of the reference: params -> preprocessing
(scratch files) -> Solver.setup / assembly / run -> postprocessing (HDF5).
The PETSc options file provides ``-ksp_type``, ``-pc_type``, ``-ksp_rtol``
(plus ``-pg_precision fp64|mixed``).
"""
import sys

from .common import InputParameters, Print, SolverOptions, Timers, params_path_from_argv
from .postprocessing import Postprocessing
from .preprocessing import Preprocessing
from .solver import Solver


def main(argv=None):
    argv = list(sys.argv if argv is None else argv)
    params = params_path_from_argv(argv)
    input_setup = InputParameters(params)
    options = SolverOptions(argv[1:])
    Timers(input_setup.output["directory"])
    Print.header()
    Print.master(" ")
    Print.master("  Data preprocessing")
    Preprocessing().run(input_setup)
    Print.master(" ")
    Print.master("  Run modelling")
    solver = Solver(options)
    solver.setup(input_setup)
    solver.assembly(input_setup)
    solver.run(input_setup)
    Print.master(" ")
    Print.master("  Data postprocessing")
    path = Postprocessing().run(input_setup, solver)
    del solver
    for k, t in Timers().items():
        Print.master(f"     {k:16s} {t.elapsed:10.3f} s")
    return path


if __name__ == "__main__":
    main()
