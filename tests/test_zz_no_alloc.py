"""The adjoint-gradient entry point that allocates nothing
(pg_adjoint_gradient_noalloc) against the one that allocates
(pg_adjoint_gradient), both through the C ABI: two calls in a row, the second
on a smaller mesh with fewer right-hand sides, give bitwise the same gradient;
a call beyond the staging limit (4096 right-hand sides) is refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from petgem_amd import _native as nat, core  # noqa: E402
from petgem_amd.synthetic import Workload  # noqa: E402


def _case(n, p, nrhs, seed):
    w = Workload(n, p, block=False)
    T = w.topo
    mesh = core.DeviceMesh.from_topology(T, "cuda")
    asm = core.Assembler(mesh)
    asm.geometry(w.sigma)
    rng = np.random.default_rng(seed)
    x = torch.from_numpy(rng.standard_normal((nrhs, T.ndof, 2))).cuda()
    lam = torch.from_numpy(rng.standard_normal((nrhs, T.ndof, 2))).cuda()
    omega = np.ascontiguousarray(2.0 * np.pi * rng.uniform(0.1, 2.0, nrhs))
    return mesh, asm, x, lam, omega


def _gradient(fn, mesh, asm, x, lam, omega):
    g = torch.zeros(mesh.n_elems, dtype=torch.float64, device="cuda")
    nat.check(fn(asm.tables.handle, mesh.n_elems, nat.ptr(asm.workspace), nat.ptr(mesh.dofs),
                 int(omega.size), mesh.ndof, nat.ptr(x), nat.ptr(lam), nat.ptr(omega), nat.ptr(g),
                 nat.stream_handle()))
    torch.cuda.synchronize()
    return g.cpu().numpy()


def test_noalloc_gradient_equals_allocating_entry_point():
    L = nat.lib()
    # first call: p = 3, 9 right-hand sides (two chunks of the kernel's
    # right-hand-side loop); second call: smaller mesh, p = 2, 2 right-hand sides
    for n, p, nrhs, seed in [(4, 3, 9, 0), (3, 2, 2, 1)]:
        print(f"noalloc case n={n} p={p} nrhs={nrhs}", flush=True)
        mesh, asm, x, lam, omega = _case(n, p, nrhs, seed)
        g_ref = _gradient(L.pg_adjoint_gradient, mesh, asm, x, lam, omega)
        g_na = _gradient(L.pg_adjoint_gradient_noalloc, mesh, asm, x, lam, omega)
        assert np.abs(g_ref).max() > 0.0
        assert np.array_equal(g_na, g_ref), np.abs(g_na - g_ref).max()


def test_noalloc_gradient_rejects_more_than_the_staging_limit():
    mesh, asm, x, lam, _ = _case(2, 1, 1, 2)
    omega = np.full(4097, 2.0 * np.pi)
    g = torch.zeros(mesh.n_elems, dtype=torch.float64, device="cuda")
    rc = nat.lib().pg_adjoint_gradient_noalloc(
        asm.tables.handle, mesh.n_elems, nat.ptr(asm.workspace), nat.ptr(mesh.dofs), 4097,
        mesh.ndof, nat.ptr(x), nat.ptr(lam), nat.ptr(omega), nat.ptr(g), nat.stream_handle())
    assert rc == nat.PG_ERR_ARG
    assert "4096" in nat.lib().pg_last_error().decode()
