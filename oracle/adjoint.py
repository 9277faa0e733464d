"""FP64 CPU reference gradient of the receiver misfit w.r.t. per-element log sigma.

No reference counterpart (the reference has no inversion); this is the
oracle for the device adjoint.  Built on the oracle's own forward pieces
(``fem``): SuperLU solves of the PEC-constrained system, receiver fields as
in ``postprocessing.py:566-614``, and

    g_e = sum_{s,f} -omega mu0 Im(lam_e^T M_e(sigma) x_e),
    lam = A^-1 R^T (w conj(E - E_obs)),  w = 1/(|E_obs|^2 + floor^2).

Pinned independently of that formula by central finite differences of the
misfit itself (``tests/test_oracle_adjoint.py``).
"""
import numpy as np
import scipy.sparse.linalg as spla

from . import fem, topology as otopo

MU0 = fem.MU0


class OracleInversion:
    def __init__(self, topo, source_positions, frequencies, receivers, observed=None,
                 moment=(1.0, 0.0, 0.0), noise_floor=0.0):
        self.T = topo
        self.noise_floor = float(noise_floor)
        T = topo
        self.eo, self.fo = otopo.orientations(T.elems_nodes, T.elems_edges, T.edges_nodes,
                                              T.elem_faces_edges)
        self.sources = np.atleast_2d(np.asarray(source_positions, dtype=np.float64))
        self.freqs = [float(f) for f in frequencies]
        self.rx = np.atleast_2d(np.asarray(receivers, dtype=np.float64))
        self.moment = np.asarray(moment, dtype=np.float64)
        self.src_el = fem.locate(T.coords, self.sources)
        self.rx_el = fem.locate(T.coords, self.rx)
        self.observed = observed
        self.basis = []
        for r, e in enumerate(self.rx_el):
            xi = fem.xyz_to_reference(T.coords[e], self.rx[r])
            N, _ = fem.physical_basis(T.coords[e], self.eo[e], self.fo[e], T.p, xi)
            self.basis.append(N[0])                     # (n,3)

    def _solve_all(self, log_sigma):
        T = self.T
        s = np.exp(np.asarray(log_sigma, dtype=np.float64))
        sig = np.stack([s, s], axis=1)
        M, K = fem.all_element_matrices(T.coords, self.eo, self.fo, sig, T.p)
        out = {}
        for fi, f in enumerate(self.freqs):
            A = fem.assemble(T.dofs, M, K, f, T.ndof)
            Abc, _ = fem.apply_dirichlet(A, np.zeros(T.ndof), T.boundary_dofs)
            lu = spla.splu(Abc.tocsc())
            for si in range(self.sources.shape[0]):
                src = dict(position=self.sources[si], azimuth=0.0, dip=0.0, current=1.0,
                           length=1.0, frequency=f)
                b = fem.source_vector(T.coords, T.dofs, self.eo, self.fo, T.p, T.ndof, src,
                                      int(self.src_el[si]))
                b[T.boundary_dofs] = 0.0
                out[(si, fi)] = (lu, lu.solve(b))
        return M, out

    def fields(self, x):
        T = self.T
        return np.array([x[T.dofs[e]] @ self.basis[r] for r, e in enumerate(self.rx_el)])

    def forward(self, log_sigma):
        _, sols = self._solve_all(log_sigma)
        S, F = self.sources.shape[0], len(self.freqs)
        out = np.zeros((S, F, self.rx.shape[0], 3), dtype=np.complex128)
        for (si, fi), (_, x) in sols.items():
            out[si, fi] = self.fields(x)
        return out

    def weights(self):
        # relative data errors with an absolute floor: 1 / (|E_obs|^2 + floor^2)
        return 1.0 / np.maximum(np.abs(self.observed) ** 2 + self.noise_floor ** 2, 1e-300)

    def misfit(self, fields):
        w = self.weights()
        return 0.5 * float(np.sum(w * np.abs(fields - self.observed) ** 2))

    def misfit_and_gradient(self, log_sigma):
        T = self.T
        M, sols = self._solve_all(log_sigma)
        w = self.weights()
        S, F = self.sources.shape[0], len(self.freqs)
        fields = np.zeros((S, F, self.rx.shape[0], 3), dtype=np.complex128)
        g = np.zeros(T.n_elems)
        for (si, fi), (lu, x) in sols.items():
            fields[si, fi] = self.fields(x)
            coef = w[si, fi] * np.conj(fields[si, fi] - self.observed[si, fi])   # (R,3)
            rhs = np.zeros(T.ndof, dtype=np.complex128)
            for r, e in enumerate(self.rx_el):
                np.add.at(rhs, T.dofs[e], self.basis[r] @ coef[r])
            rhs[T.boundary_dofs] = 0.0
            lam = lu.solve(rhs)
            om = 2.0 * np.pi * self.freqs[fi] * MU0
            xe = x[T.dofs]                                # (E,n)
            le = lam[T.dofs]
            g += -om * np.imag(np.einsum("ei,eij,ej->e", le, M, xe))
        return self.misfit(fields), g
