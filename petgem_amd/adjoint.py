"""Adjoint gradient of a receiver misfit w.r.t. per-element log-conductivity.

No counterpart in the reference (an extension of its forward path).  For
sources s and frequencies f, with A_f x_sf = b_sf the PEC-constrained system
of the reference's discretisation and E_sf = R x_sf the receiver fields
(``postprocessing.py:566-614``), the misfit is

    Phi = 1/2 sum_{s,f,r,c} w_{sfrc} |E_sfrc - E^obs_sfrc|^2,
    w = 1/(|E^obs|^2 + floor^2)

(relative data errors with an optional absolute noise floor; floor = 0 gives
purely relative weights, which make the misfit as sensitive to the solver
tolerance as the weakest observed field is small),

and, since A_f = A_f^T (complex symmetric), one adjoint solve per (s,f) with
the same operator gives

    lam_sf = A_f^-1 R^T (w conj(E_sf - E^obs_sf)),
    dPhi/dlog(sigma_e) = sum_{s,f} -omega_f mu0 Im(lam_e^T M_e(sigma) x_e),

because dA_e/dlog(sigma_e) = -i omega mu0 M_e(sigma) (M_e is linear in
sigma; horizontal and vertical conductivity scale together).  The element
mass matrices are recomputed on the device inside the contraction kernel,
never stored.  Boundary rows/columns are not differentiated (the PEC rows
are fixed), which the device code realises by zeroing the adjoint source on
boundary DOFs.

With ``batched=True`` every (source, frequency) pair becomes one column of the
batched multi-RHS solver (one pass over K and M per iteration serves all
columns), for the forward and for the adjoint solves, and the contraction
reads the solutions in the solver's (ndof, nrhs) layout.
"""
import math

import numpy as np
import torch

from . import _native as nat
from . import core


def batch_width(r):
    """Smallest width of the batched solver holding r columns (powers of two
    below 64, multiples of 64 up to 256); extra columns are zero."""
    if not 1 <= r <= 256:
        raise ValueError("a batch holds 1..256 columns")
    if r <= 64:
        w = 1
        while w < r:
            w *= 2
        return w
    return -(-r // 64) * 64


class InversionProblem:
    """Forward fields, misfit and adjoint gradient for many sources and frequencies."""

    def __init__(self, topo, source_positions, frequencies, receivers, observed=None,
                 rtol=1e-10, precision="fp64", moment=(1.0, 0.0, 0.0), device=None,
                 batched=False, max_batch=64, noise_floor=0.0, concurrency=2):
        self.topo = topo
        # per-RHS mode: right-hand sides solved `concurrency` at a time, each
        # worker with its own solver, element operator and stream
        self.concurrency = max(1, int(concurrency))
        self.noise_floor = float(noise_floor)
        self.mesh = core.DeviceMesh.from_topology(topo, device)
        self.asm = core.Assembler(self.mesh)
        self.ent_bc, self.dof_bc = self.mesh.entity_mask(topo.boundary_dofs)
        self.sources = np.atleast_2d(np.asarray(source_positions, dtype=np.float64))
        self.freqs = [float(f) for f in frequencies]
        self.rx = np.atleast_2d(np.asarray(receivers, dtype=np.float64))
        self.moment = np.asarray(moment, dtype=np.float64)
        self.rtol, self.precision = rtol, precision
        self.src_el = core.locate_points(self.mesh, self.sources).cpu().numpy()
        if (self.src_el < 0).any():
            raise ValueError("a source lies outside the mesh")
        self.rx_el = core.locate_points(self.mesh, self.rx)
        if (self.rx_el.cpu().numpy() < 0).any():
            raise ValueError("a receiver lies outside the mesh")
        self.rx_dev = torch.from_numpy(self.rx).to(self.mesh.device)
        nnz = self.asm.plan.nnz
        self.batched = bool(batched)
        if self.batched:
            self.K = torch.empty(nnz, dtype=torch.float64, device=self.mesh.device)
            self.M = torch.empty_like(self.K)
            cols = [(si, fi) for fi in range(len(self.freqs)) for si in range(len(self.sources))]
            mb = max(1, min(int(max_batch), 256))
            self.chunks = [cols[i:i + mb] for i in range(0, len(cols), mb)]
            self.batches = [None] * len(self.chunks)
            self.X = None
        else:
            self.values = [torch.empty((nnz, 2), dtype=torch.float64, device=self.mesh.device)
                           for _ in self.freqs]
            self.solvers = [None] * len(self.freqs)
            # element-by-element operators, one per frequency and worker: the
            # solvers' autotuners use them when they are the faster SpMV
            self.eops = [None] * len(self.freqs)
        self.observed = None if observed is None else np.asarray(observed)
        self.x = None
        self.stats = dict(solves=0, iterations=0)

    # ---------------------------------------------------------------- forward
    def _set_model(self, log_sigma):
        ls = log_sigma.detach() if torch.is_tensor(log_sigma) else torch.from_numpy(
            np.asarray(log_sigma, dtype=np.float64))
        sig = torch.exp(ls.to(self.mesh.device, torch.float64))
        self.asm.geometry(torch.stack([sig, sig], dim=1).contiguous())
        if self.batched:
            core.assemble_km(self.asm, self.ent_bc, out=(self.K, self.M))
            for ci, ch in enumerate(self.chunks):
                if self.batches[ci] is None:
                    w = batch_width(len(ch))
                    fr = [self.freqs[fi] for _, fi in ch] + [self.freqs[ch[0][1]]] * (w - len(ch))
                    self.batches[ci] = core.BatchSolver(self.asm.plan, self.K, self.M, fr,
                                                        rtol=self.rtol,
                                                        block_ptr=self.mesh.block_ptr(),
                                                        entity_order=self.mesh.entity_order(),
                                                        precision=self.precision, pc="bjacobi")
                else:
                    self.batches[ci].refresh()
            return
        for fi, f in enumerate(self.freqs):
            om = core.omega_of(f)
            self.asm.assemble(om, ent_bc=self.ent_bc, out=self.values[fi])
            if self.solvers[fi] is None:
                # COCG + entity-block Jacobi; the SpMV is the faster of the
                # entity-row CSR kernels (Morton order) and the element products
                order = self.mesh.entity_order()
                self.eops[fi] = [core.ElementOperator(self.asm, om, self.ent_bc, order)
                                 for _ in range(self.concurrency)]
                # the workers' solvers take the first one's kernel choices, so that
                # every right-hand side is solved with the same arithmetic
                self.solvers[fi] = []
                for e in self.eops[fi]:
                    self.solvers[fi].append(core.KrylovSolver(
                        self.asm.plan, self.values[fi], precision=self.precision, rtol=self.rtol,
                        pc="bjacobi", block_ptr=self.mesh.block_ptr(), entity_order=order, eop=e,
                        like=self.solvers[fi][0] if self.solvers[fi] else None))
            else:
                for e, S in zip(self.eops[fi], self.solvers[fi]):
                    e.refresh(om)
                    S.refresh()

    def _solve_all(self, fi, bs):
        """Solutions of A_f x = b for every b in bs (concurrently when configured)."""
        solvers = self.solvers[fi]
        if len(solvers) > 1 and len(bs) > 1:
            xs, infos = core.solve_concurrently(solvers, bs)
        else:
            xs, infos = zip(*[solvers[0].solve(b) for b in bs]) if bs else ((), ())
        for info in infos:
            if info["status"] != 0:
                raise RuntimeError(f"Krylov solve failed: {info}")
            self.stats["solves"] += 1
            self.stats["iterations"] += info["iterations"]
        return list(xs)

    def _solve_batch(self, ci, column):
        """Solve chunk ci with right-hand side column(si, fi) -> (ndof, 2) per pair."""
        bs = self.batches[ci]
        ch = self.chunks[ci]
        B = torch.zeros((self.mesh.ndof, bs.nrhs, 2), dtype=torch.float64, device=self.mesh.device)
        for c, (si, fi) in enumerate(ch):
            B[:, c, :] = column(si, fi)
        X, info = bs.solve(B)
        st = info["status"][:len(ch)]
        if (st != 0).any():
            raise RuntimeError(f"batched Krylov solve failed: status {st}, "
                               f"relres {info['relres'][:len(ch)]}")
        self.stats["solves"] += len(ch)
        self.stats["iterations"] += int(info["iterations"][:len(ch)].sum())
        return X

    def _source(self, si, fi):
        b = core.source_vector(self.mesh, int(self.src_el[si]), self.sources[si], self.moment,
                               core.omega_of(self.freqs[fi]))
        self._mask(b)
        return b

    def _adjoint_source(self, si, fi, coef):
        m = self.mesh
        cdev = core.from_complex(coef, m.device)
        rhs = torch.zeros((m.ndof, 2), dtype=torch.float64, device=m.device)
        nat.check(nat.lib().pg_adjoint_rhs(m.p, self.rx.shape[0], nat.ptr(self.rx_dev),
                                           nat.ptr(self.rx_el), nat.ptr(m.coords),
                                           nat.ptr(m.elems_nodes), nat.ptr(m.elems_edges),
                                           nat.ptr(m.edges_nodes), nat.ptr(m.faces_edges),
                                           nat.ptr(m.dofs), nat.ptr(cdev), nat.ptr(rhs),
                                           nat.stream_handle()))
        self._mask(rhs)
        return rhs

    def _mask(self, v):
        nat.check(nat.lib().pg_apply_dirichlet(self.mesh.ndof, nat.ptr(self.asm.plan.rowptr),
                                               nat.ptr(self.asm.plan.cols), nat.ptr(self.dof_bc),
                                               None, nat.ptr(v), nat.stream_handle()))

    def forward(self, log_sigma):
        """Receiver E fields (S, F, R, 3) complex128 and the solutions kept on the device."""
        self._set_model(log_sigma)
        S, F = self.sources.shape[0], len(self.freqs)
        out = np.zeros((S, F, self.rx.shape[0], 3), dtype=np.complex128)
        if self.batched:
            self.X = []
            for ci, ch in enumerate(self.chunks):
                X = self._solve_batch(ci, self._source)
                self.X.append(X)
                for c, (si, fi) in enumerate(ch):
                    fld = core.receiver_fields(self.mesh, self.rx, self.rx_el,
                                               X[:, c, :].contiguous(),
                                               core.omega_of(self.freqs[fi]))
                    out[si, fi] = core.as_complex(fld)[:, 0:3]
            return out
        self.x = [[None] * F for _ in range(S)]
        for fi, f in enumerate(self.freqs):
            om = core.omega_of(f)
            xs = self._solve_all(fi, [self._source(si, fi) for si in range(S)])
            for si, x in enumerate(xs):
                self.x[si][fi] = x
                fld = core.receiver_fields(self.mesh, self.rx, self.rx_el, x, om)
                out[si, fi] = core.as_complex(fld)[:, 0:3]
        return out

    def weights(self):
        """w = 1 / (|E_obs|^2 + floor^2), per observed component."""
        return 1.0 / np.maximum(np.abs(self.observed) ** 2 + self.noise_floor ** 2, 1e-300)

    def misfit(self, fields):
        d = fields - self.observed
        return 0.5 * float(np.sum(self.weights() * np.abs(d) ** 2))

    # --------------------------------------------------------------- gradient
    def misfit_and_gradient(self, log_sigma):
        if self.observed is None:
            raise ValueError("observed data not set")
        fields = self.forward(log_sigma)
        phi = self.misfit(fields)
        m = self.mesh
        g = torch.zeros(m.n_elems, dtype=torch.float64, device=m.device)
        w = self.weights()
        L = nat.lib()

        def adjoint_source(si, fi):
            coef = w[si, fi] * np.conj(fields[si, fi] - self.observed[si, fi])
            return self._adjoint_source(si, fi, coef)

        if self.batched:
            for ci, ch in enumerate(self.chunks):
                lam = self._solve_batch(ci, adjoint_source)
                width = lam.shape[1]
                omega = np.zeros(width, dtype=np.float64)
                omega[:len(ch)] = [core.omega_of(self.freqs[fi]) for _, fi in ch]
                nat.check(L.pg_adjoint_gradient_rows(self.asm.tables.handle, m.n_elems,
                                                     nat.ptr(self.asm.workspace), nat.ptr(m.dofs),
                                                     width, m.ndof, nat.ptr(self.X[ci]),
                                                     nat.ptr(lam), nat.ptr(omega), nat.ptr(g),
                                                     nat.stream_handle()))
            return phi, g
        for fi, f in enumerate(self.freqs):
            om = core.omega_of(f)
            S = self.sources.shape[0]
            lams = self._solve_all(fi, [adjoint_source(si, fi) for si in range(S)])
            for si, lam in enumerate(lams):
                omega = np.array([om], dtype=np.float64)
                nat.check(L.pg_adjoint_gradient(self.asm.tables.handle, m.n_elems,
                                                nat.ptr(self.asm.workspace), nat.ptr(m.dofs), 1,
                                                m.ndof, nat.ptr(self.x[si][fi]), nat.ptr(lam),
                                                nat.ptr(omega), nat.ptr(g), nat.stream_handle()))
        return phi, g


class LogConductivityMisfit(torch.autograd.Function):
    """Phi(log sigma) with the adjoint gradient as its backward."""

    @staticmethod
    def forward(ctx, log_sigma, problem):
        phi, g = problem.misfit_and_gradient(log_sigma)
        ctx.save_for_backward(g)
        return torch.tensor(phi, dtype=torch.float64, device=log_sigma.device)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g.to(grad_out.device), None


def misfit(log_sigma, problem):
    return LogConductivityMisfit.apply(log_sigma, problem)


def synthetic_observations(problem, log_sigma):
    """Fields of a model, to be used as observed data."""
    return problem.forward(log_sigma)


MU0 = 4.0 * math.pi * 1e-7
