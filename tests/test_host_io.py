"""Host-side boundary pieces on the CPU: PETSc binary codec, HDF5 codec,
params-file validation and solver-option parsing (no GPU)."""
import os

import numpy as np
import pytest

from petgem_amd import h5, petsc_io
from petgem_amd.common import InputParameters, ParamsError, SolverOptions

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_petsc_codec_roundtrips_reference_files(tmp_path):
    for name in ("matrix-A.dat", "vector-b.dat"):
        src = os.path.join(GOLDEN, name)
        dst = str(tmp_path / name)
        if name.startswith("matrix"):
            ip, ix, d, sh = petsc_io.read_mat(src)
            assert sh == (4184, 4184) and ix.size == 49931
            petsc_io.write_aij(dst, ip, ix, d, sh)
        else:
            v = petsc_io.read_vec(src)
            assert v.size == 4184 and np.count_nonzero(v) == 3032
            petsc_io.write_vec(dst, v)
        assert open(src, "rb").read() == open(dst, "rb").read()


def test_petsc_dense_roundtrip(tmp_path):
    a = np.arange(24, dtype=float).reshape(6, 4) - 3.5
    petsc_io.write_dense(str(tmp_path / "m.dat"), a)
    assert np.array_equal(np.real(petsc_io.read_dense(str(tmp_path / "m.dat"))), a)


def test_hdf5_roundtrip(tmp_path):
    rx = np.load(os.path.join(GOLDEN, "receiver_pos.npy"))
    p = str(tmp_path / "rx.h5")
    h5.write_data(p, rx)
    assert np.array_equal(h5.read_dataset(p), rx)
    with h5.Writer(str(tmp_path / "o.h5")) as w:
        w.dataset("model/E-fields/x", np.array([1 + 2j, -3j]))
        w.dataset("model/run-time (s)", 1.25)
        w.dataset("model/mode", "csem")
    assert h5.read_dataset(str(tmp_path / "o.h5"), "model/run-time (s)") == 1.25


def _params(tmp_path, **over):
    d = dict(nord=2, mode="csem", sigma="horizontal: [1.]\n      vertical: [1.]")
    d.update(over)
    txt = f"""
model:
  mode: {d['mode']}
  csem:
    sigma:
      {d['sigma']}
    source:
      frequency: 1.
      position: [0., 0., 0.]
      azimuth: 0.
      dip: 0.
      current: 1.
      length: 1.
  mesh: m.msh
  receivers: r.h5
run:
  nord: {d['nord']}
output:
  directory: {tmp_path / 'out'}
"""
    p = tmp_path / "params.yaml"
    p.write_text(txt)
    return str(p)


def test_params_accepts_reference_schema(tmp_path):
    ip = InputParameters(_params(tmp_path))
    assert ip.run["num_polarizations"] == 1 and ip.run["conductivity_from_file"] is False
    assert ip.run["cuda"] is False and ip.output["remove_scratch"] is False
    assert ip.output["directory_scratch"] == ip.output["directory"]


@pytest.mark.parametrize("over", [dict(nord=4), dict(nord=0), dict(mode="mt"), dict(mode="xx"),
                                  dict(sigma="file: s.h5\n      horizontal: [1.]")])
def test_params_rejects(tmp_path, over):
    with pytest.raises(ParamsError):
        InputParameters(_params(tmp_path, **over))


def test_solver_options_from_reference_opts_file(tmp_path):
    f = tmp_path / "petsc.opts"
    f.write_text("# Solver options for PETSc\n-ksp_type gmres\n-pc_type sor\n-ksp_rtol 1e-8\n")
    o = SolverOptions(["kernel.py", "-options_file", str(f), "params.yaml"])
    assert (o.ksp_type, o.pc_type, o.rtol) == ("gmres", "bjacobi", 1e-8)
    assert any("sor" in n for n in o.notes)
    o = SolverOptions(["-ksp_type", "cocg", "-pg_precision", "mixed", "-ksp_max_it", "50"])
    assert (o.ksp_type, o.precision, o.max_it, o.spmv) == ("cocg", "mixed", 50, "auto")
    assert SolverOptions(["-pg_spmv", "element"]).spmv == "element"
    with pytest.raises(ParamsError):
        SolverOptions(["-pg_spmv", "magic"])
