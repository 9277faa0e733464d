"""Parameters file, solver options and timers of the petgem driver.

``InputParameters`` accepts the reference's YAML params file
(``petgem/common.py:129-361``; schema ``doc/source/_downloads/params.yaml``)
with the same validation messages and the same derived keys
(``run.conductivity_from_file``, ``run.num_polarizations``,
``output.directory_scratch``, ``output.remove_scratch``).  Only the CSEM
mode and orders 1..3 are accepted here (the MT branch and p > 3 are out of
this build's scope and are rejected with a clear error).

``SolverOptions`` reads the PETSc-style flags the reference passes through
``-options_file`` (``kernel.py:15``, ``solver.py:588``): ``-ksp_type``,
``-pc_type``, ``-ksp_rtol``, ``-ksp_max_it``, ``-ksp_gmres_restart``, plus
``-pg_precision fp64|mixed``, ``-pg_inner_rtol`` and
``-pg_spmv auto|row|entity|entity_pairs|element`` (the product kernel family;
auto = the fastest by timing, element = the element-by-element operator).

``Timers`` keeps named wall-clock timers with the reference's names
(``Setup``, ``Assembly``, ``SetBoundaries``, ``Solver``, ...;
``petgem/common.py:367-436``); a timer that brackets device work is stopped
only after the device has drained.
"""
import os
import sys
import time

import yaml


class ParamsError(ValueError):
    pass


def _fail(msg):
    raise ParamsError(msg.strip())


class InputParameters:
    """The params file as object attributes ``model``, ``run``, ``output``."""

    def __init__(self, params, parEnv=None):
        with open(params, "r") as f:
            inputs = yaml.safe_load(f)
        for k in ("model", "run", "output"):
            if k not in inputs:
                _fail(f"'{k}' section not provided. Please, verify the parameter file consistency.")
        self.model = inputs["model"]
        self.run = inputs["run"]
        self.output = inputs["output"]
        if "mode" not in self.model:
            _fail("Modeling mode not provided. Please, verify the parameter file consistency.")
        if self.model["mode"] not in ("csem", "mt"):
            _fail("Modeling mode not supported.")
        if self.model["mode"] == "mt":
            _fail("Modeling mode 'mt' is not supported by petgem_amd (CSEM only).")
        if "csem" not in self.model:
            _fail("csem parameters not provided. Please, verify the parameter file consistency.")
        cff, npol = self._verify_csem(self.model["csem"])
        self.run["conductivity_from_file"] = cff
        self.run["num_polarizations"] = npol
        if "mesh" not in self.model:
            _fail("mesh parameter not provided. Please, verify the parameter file consistency.")
        if "receivers" not in self.model:
            _fail("receivers parameter not provided. Please, verify the parameter file consistency.")
        if "nord" not in self.run:
            _fail("nord parameter not provided. Please, verify the parameter file consistency.")
        if not (1 <= int(self.run["nord"]) <= 6):
            _fail("Vector finite element basis order not supported. Please, select a valid "
                  "order (1,2,3,4,5,6).")
        if int(self.run["nord"]) > 3:
            _fail("Vector finite element basis order > 3 is not supported by petgem_amd (1,2,3).")
        if "cuda" not in self.run:
            self.run["cuda"] = False
        elif self.run["cuda"] not in (True, False):
            _fail("cuda option not supported. Please, select a valid order (True/False).")
        self.output.setdefault("vtk", False)
        if "directory" not in self.output:
            _fail("output directory parameter not provided. Please, verify the parameter file "
                  "consistency.")
        os.makedirs(self.output["directory"], exist_ok=True)
        if "directory_scratch" not in self.output:
            self.output["directory_scratch"] = self.output["directory"]
            self.output["remove_scratch"] = False
        else:
            os.makedirs(self.output["directory_scratch"], exist_ok=True)
            self.output["remove_scratch"] = True

    @staticmethod
    def _verify_csem(p):
        if "sigma" not in p:
            _fail("csem parameters not provided. Please, verify the parameter file consistency.")
        s = p["sigma"]
        if "file" in s:
            if "horizontal" in s or "vertical" in s:
                _fail("sigma parameters invalid. Please, verify the parameter file consistency.")
            cff = True
        elif "horizontal" in s and "vertical" in s:
            cff = False
        else:
            _fail("sigma parameters invalid. Please, verify the parameter file consistency.")
        if "source" not in p:
            _fail("source parameters not provided. Please, verify the parameter file consistency.")
        src = p["source"]
        if len(src) != 6:
            _fail("number of source parameters is not consistent. Please, verify the parameter "
                  "file consistency.")
        for k in ("frequency", "position", "azimuth", "dip", "current", "length"):
            if k not in src:
                _fail(f"{k} parameter not provided. Please, verify the parameter file consistency.")
        return cff, 1


class SolverOptions:
    """PETSc-style solver flags (options file and/or command line)."""

    KSP = {"cocg": "cocg", "gmres": "gmres", "fgmres": "gmres", "lgmres": "gmres",
           "cg": "cocg", "bcgs": "cocg", "tfqmr": "cocg", "preonly": "cocg"}
    PC = {"jacobi": "jacobi", "none": "none", "bjacobi": "bjacobi", "pbjacobi": "bjacobi",
          "sor": "bjacobi", "asm": "bjacobi", "ilu": "bjacobi", "gamg": "bjacobi", "lu": "bjacobi"}

    def __init__(self, argv=None, options_file=None):
        self.ksp_type = "gmres"            # PETSc's default KSP
        self.pc_type = "jacobi"
        self.rtol = 1e-8
        self.max_it = 100000
        self.restart = 30
        self.precision = "fp64"
        self.inner_rtol = None  # complex64 round tolerance: by basis order (core.mixed_defaults)
        self.spmv = "auto"
        self.notes = []
        flags = []
        if options_file:
            flags += self._read_file(options_file)
        argv = list(argv or [])
        i = 0
        while i < len(argv):
            a = argv[i]
            if a == "-options_file" and i + 1 < len(argv):
                flags += self._read_file(argv[i + 1])
                i += 2
                continue
            if a.startswith("-") and not a.startswith("--"):
                val = argv[i + 1] if i + 1 < len(argv) and not argv[i + 1].startswith("-") else None
                flags.append((a, val))
                i += 2 if val is not None else 1
                continue
            i += 1
        for k, v in flags:
            self._apply(k, v)

    @staticmethod
    def _read_file(path):
        out = []
        with open(path) as f:
            for line in f:
                line = line.split("#", 1)[0].strip()
                if not line:
                    continue
                parts = line.split()
                out.append((parts[0], parts[1] if len(parts) > 1 else None))
        return out

    def _apply(self, k, v):
        if k == "-ksp_type" and v:
            self.ksp_type = self.KSP.get(v, "cocg")
            if self.KSP.get(v) is None or v not in ("cocg", "gmres"):
                self.notes.append(f"-ksp_type {v} -> {self.ksp_type}")
        elif k == "-pc_type" and v:
            self.pc_type = self.PC.get(v, "jacobi")
            if v not in ("jacobi", "none", "bjacobi"):
                self.notes.append(f"-pc_type {v} -> {self.pc_type}")
        elif k == "-ksp_rtol" and v:
            self.rtol = float(v)
        elif k == "-ksp_max_it" and v:
            self.max_it = int(v)
        elif k == "-ksp_gmres_restart" and v:
            self.restart = int(v)
        elif k == "-pg_precision" and v:
            if v not in ("fp64", "mixed"):
                raise ParamsError(f"-pg_precision must be fp64 or mixed, not {v}")
            self.precision = v
        elif k == "-pg_inner_rtol" and v:
            self.inner_rtol = float(v)
        elif k == "-pg_spmv" and v:
            if v not in ("auto", "row", "entity", "entity_pairs", "element"):
                raise ParamsError(f"-pg_spmv must be auto, row, entity, entity_pairs or "
                                  f"element, not {v}")
            self.spmv = v

    def ksp_name(self):
        return self.ksp_type


class Timer:
    def __init__(self, elapsed=0.0):
        self._start = 0.0
        self.elapsed = elapsed

    def start(self):
        self._start = time.time()

    def stop(self, sync=None):
        if sync is not None:
            sync()
        if self._start > 0:
            self.elapsed += time.time() - self._start
            self._start = 0.0

    def reset(self):
        self.elapsed = 0.0


class _Timers:
    def __init__(self):
        self._elems = {}

    def __getitem__(self, key):
        return self._elems.setdefault(key, Timer())

    def __setitem__(self, key, value):
        self._elems[key] = Timer(value)

    def items(self):
        return self._elems.items()

    def reset(self):
        self._elems = {}


_TIMERS = _Timers()


def Timers(opath=None):  # noqa: N802 - same name as the reference's singleton
    """The timers singleton.  When the reference's ``petgem.common`` is loaded
    (its ``kernel.py`` imports it before creating the solver), its own
    ``Timers()`` singleton is returned, so that the reference's postprocessing
    (``postprocessing.py:381-383``: ``Timers()["Assembly"].elapsed +
    Timers()["Solver"].elapsed``) reads the times this solver recorded."""
    ref = sys.modules.get("petgem.common")
    if ref is not None and callable(getattr(ref, "Timers", None)):
        return ref.Timers()
    return _TIMERS


def stop_timer(name, sync=True):
    """Stop timer `name` after the device has drained (the reference's Timer.stop
    takes no argument, so the drain happens here)."""
    if sync:
        device_sync()
    Timers()[name].stop()


class Print:
    """Rank-0 console output (single process here)."""

    @staticmethod
    def master(msg=""):
        print(msg, flush=True)

    @staticmethod
    def header():
        Print.master("%" * 75)
        Print.master("  petgem_amd -- CSEM forward modelling on MI355X (gfx950)")
        Print.master("%" * 75)


def device_sync():
    import torch
    if torch.cuda.is_available():
        torch.cuda.synchronize()


def params_path_from_argv(argv):
    """The params file is the last non-flag argument (``kernel.py:35`` uses argv[3])."""
    cands = [a for a in argv[1:] if a.endswith((".yaml", ".yml"))]
    if cands:
        return cands[-1]
    if len(argv) >= 4:
        return argv[3]
    raise ParamsError("usage: python -m petgem_amd.kernel -options_file petsc.opts params.yaml")


if __name__ == "__main__":  # pragma: no cover
    print(InputParameters(params_path_from_argv(sys.argv)).run)
