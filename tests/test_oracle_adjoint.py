"""The oracle's adjoint gradient against central finite differences of the
misfit (CPU).  This pins the gradient formula independently of its
derivation: a wrong sign, a missing conjugate or a dropped weight fails."""
import numpy as np
import pytest

from oracle.adjoint import OracleInversion
from petgem_amd.synthetic import Workload, receiver_line


@pytest.mark.parametrize("p,floor", [(1, 0.0), (2, 0.0), (2, 0.3)])
def test_oracle_gradient_matches_central_fd(p, floor):
    """floor: absolute noise floor of the weights, as a fraction of the largest
    observed field (0.3 changes most weights substantially)."""
    w = Workload(4, p, block=False)
    T = w.topo
    src = np.array([[0.0, 0.0, -950.0], [700.0, 300.0, -1200.0]])
    rx = receiver_line(9, z=-2000.0, x0=-3000.0, x1=3000.0, y=500.0)
    true = np.log(w.sigma[:, 0]) + 0.3 * np.random.default_rng(1).standard_normal(T.n_elems)
    inv = OracleInversion(T, src, [0.5, 1.0], rx)
    inv.observed = inv.forward(true)
    inv.noise_floor = floor * np.abs(inv.observed).max()
    m0 = np.log(w.sigma[:, 0])
    phi, g = inv.misfit_and_gradient(m0)
    assert phi > 0
    h = 1e-4
    for e in np.random.default_rng(2).choice(T.n_elems, 6, replace=False):
        mp, mm = m0.copy(), m0.copy()
        mp[e] += h
        mm[e] -= h
        fd = (inv.misfit(inv.forward(mp)) - inv.misfit(inv.forward(mm))) / (2 * h)
        assert abs(fd - g[e]) <= 1e-5 * np.abs(g).max() + 1e-4 * abs(fd), (e, fd, g[e])
