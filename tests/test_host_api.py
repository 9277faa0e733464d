"""Host-side logic of the drop-in boundary, on the CPU: solver options from the
command line, the timers singleton shared with a loaded reference
``petgem.common``, and the check that concurrently running solvers own their
device state.  No GPU work is launched."""
import sys
import types

import pytest

from petgem_amd import common, core


def test_solver_without_options_reads_the_command_line(tmp_path, monkeypatch):
    opts = tmp_path / "petsc.opts"
    opts.write_text("# PETSc options\n-ksp_type cocg\n-pc_type sor\n-ksp_rtol 1e-9\n"
                    "-pg_precision mixed\n")
    monkeypatch.setattr(sys, "argv", ["kernel.py", "-options_file", str(opts), "params.yaml",
                                      "-ksp_max_it", "77"])
    from petgem_amd.solver import Solver
    o = Solver().options
    assert (o.ksp_type, o.pc_type, o.rtol, o.precision, o.max_it) == (
        "cocg", "bjacobi", 1e-9, "mixed", 77)
    assert any("sor" in n for n in o.notes)  # the mapping is reported


def test_timers_are_the_reference_singleton_when_it_is_loaded(monkeypatch):
    own = common.Timers()
    assert common.Timers() is own
    ref = object()
    mod = types.ModuleType("petgem.common")
    mod.Timers = lambda opath=None: ref
    monkeypatch.setitem(sys.modules, "petgem.common", mod)
    assert common.Timers() is ref
    monkeypatch.delitem(sys.modules, "petgem.common")
    assert common.Timers() is own


class _Fake:
    def __init__(self, h, eop=None):
        import ctypes
        self.handle = ctypes.c_void_p(h)
        self.eop = eop


def test_concurrent_solvers_must_not_share_device_state():
    e1, e2 = _Fake(101), _Fake(102)
    core._check_independent([_Fake(1, e1), _Fake(2, e2)])
    s = _Fake(3, e1)
    with pytest.raises(ValueError, match="same solver"):
        core._check_independent([s, s])
    with pytest.raises(ValueError, match="ElementOperator"):
        core._check_independent([_Fake(1, e1), _Fake(2, e1)])
    with pytest.raises(ValueError, match="same solver"):
        core.solve_concurrently([s, s], [None, None])
    with pytest.raises(ValueError, match="one solver per right-hand side"):
        core.solve_each([s], [None, None])
