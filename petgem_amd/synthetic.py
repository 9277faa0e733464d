"""Synthetic CSEM workloads: structured Kuhn tetrahedral grids and layered models.

A cube [-L/2, L/2]^3 is divided into n^3 hexahedra, each split into the 6
Kuhn (Freudenthal) tetrahedra that share the hexahedron's main diagonal, so
the mesh is conforming.  Node order inside every tetrahedron is chosen so
that the signed volume is positive (the reference uses the signed Jacobian
determinant, ``petgem/hvfem.py:265,313-314``, so negatively oriented
elements would give negative-definite element matrices).

The conductivity model is the layered marine model of the benchmark
workloads: seawater 3.3 S/m above z = -1 km, sediments 1.0 S/m down to
z = -3 km, basement 0.1 S/m below, and an optional resistive block
(0.01 S/m, 2 x 2 x 0.2 km centred at z = -2 km), assigned by element
centroid, with optional seeded log-normal per-element noise.
"""
import itertools

import numpy as np

from .mesh import Topology

SEAWATER, SEDIMENTS, BASEMENT, RESISTOR = 3.3, 1.0, 0.1, 0.01


def kuhn_grid(n, length=10000.0, origin=None):
    """Nodes (nNodes,3) and tetrahedra (6 n^3, 4) of an n^3 Kuhn grid."""
    if origin is None:
        origin = -0.5 * length
    h = length / n
    ax = origin + h * np.arange(n + 1)
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)

    def nid(i, j, k):
        return i + (n + 1) * (j + (n + 1) * k)

    ii, jj, kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    # hexahedra in x-fastest order
    ii, jj, kk = (a.transpose(2, 1, 0).ravel() for a in (ii, jj, kk))
    tets = []
    for perm in itertools.permutations(range(3)):
        path = [np.zeros(3, dtype=np.int64)]
        for ax_ in perm:
            nxt = path[-1].copy()
            nxt[ax_] += 1
            path.append(nxt)
        vs = [nid(ii + d[0], jj + d[1], kk + d[2]) for d in path]
        tets.append(np.stack(vs, axis=1))
    tets = np.stack(tets, axis=1).reshape(-1, 4)                       # hex-major, 6 per hex
    c = pts[tets]
    det = np.linalg.det(c[:, 1:4] - c[:, 0:1])
    neg = det < 0
    tets[neg, 2], tets[neg, 3] = tets[neg, 3].copy(), tets[neg, 2].copy()
    return pts, tets


def layered_sigma(centroids, block=True, noise=0.0, seed=0):
    """(E,2) horizontal/vertical conductivity of the layered marine model."""
    z = centroids[:, 2]
    s = np.where(z > -1000.0, SEAWATER, np.where(z > -3000.0, SEDIMENTS, BASEMENT))
    if block:
        inb = ((np.abs(centroids[:, 0]) <= 1000.0) & (np.abs(centroids[:, 1]) <= 1000.0)
               & (np.abs(z + 2000.0) <= 100.0))
        s = np.where(inb, RESISTOR, s)
    if noise > 0.0:
        rng = np.random.default_rng(seed)
        s = s * np.exp(noise * rng.standard_normal(s.shape[0]))
    return np.stack([s, s], axis=1)


def receiver_line(n_rx, z=-1000.0, x0=-4000.0, x1=4000.0, y=0.0):
    x = np.linspace(x0, x1, n_rx)
    return np.stack([x, np.full_like(x, y), np.full_like(x, z)], axis=1)


class Workload:
    """One synthetic CSEM problem: mesh, order, model, source and receivers."""

    def __init__(self, n, p, block=True, noise=0.0, seed=0, n_rx=41, frequency=1.0,
                 source=(0.0, 0.0, -950.0), length=10000.0):
        self.n, self.p = n, p
        self.points, self.elems_nodes = kuhn_grid(n, length)
        self.topo = Topology(self.points, self.elems_nodes, p)
        cen = self.topo.coords.mean(axis=1)
        self.sigma = layered_sigma(cen, block=block, noise=noise, seed=seed)
        self.receivers = receiver_line(n_rx)
        self.frequency = float(frequency)
        self.source = dict(position=list(source), azimuth=0.0, dip=0.0, current=1.0,
                           length=1.0, frequency=self.frequency)
        self.seed = seed

    def describe(self):
        t = self.topo
        return dict(hexes=f"{self.n}^3", elements=t.n_elems, p=self.p, dofs=t.ndof,
                    boundary_dofs=int(t.boundary_dofs.size), frequency_hz=self.frequency,
                    receivers=int(self.receivers.shape[0]))


# The five workloads of the benchmark (sizes; see bench.py for which one is timed)
CONFIGS = {
    "c1": dict(n=20, p=1, block=True, noise=0.0, n_rx=41),
    "c2": dict(n=40, p=2, block=True, noise=0.1, n_rx=101),
    "c3": dict(n=24, p=3, block=True, noise=0.0, n_rx=101),
    "c4": dict(n=30, p=2, block=True, noise=0.0, n_rx=101),
}


def multimesh_systems(device, rtol=1e-8):
    """The multi-mesh workload: 32 meshes of 8^3 .. 24^3 Kuhn hexes x 6 (p = 2),
    seeded random 3-layer conductivities (0.01 - 3.3 S/m, layer tops at -1 and
    -3 km), an x-directed HED at (0, 0, -950) m in each, 0.1 / 0.25 / 0.5 / 1 Hz
    cyclic.  Returns per mesh a dict: mesh, assembler (after its geometry),
    plan, vals (assembled, PEC applied), b, omega, ent_bc, order."""
    import torch
    from . import core
    sizes = np.round(np.linspace(8, 24, 32)).astype(int)
    freqs = [0.1, 0.25, 0.5, 1.0]
    out = []
    for i, n in enumerate(sizes):
        w = Workload(int(n), 2, block=False, seed=i)
        rng = np.random.default_rng(1000 + i)
        layers = np.exp(rng.uniform(np.log(0.01), np.log(3.3), 3))
        z = w.topo.coords.mean(axis=1)[:, 2]
        s = np.where(z > -1000.0, layers[0], np.where(z > -3000.0, layers[1], layers[2]))
        mesh = core.DeviceMesh.from_topology(w.topo, torch.device(device))
        asm = core.Assembler(mesh)
        asm.geometry(np.stack([s, s], axis=1))
        ent_bc, dof_bc = mesh.entity_mask(w.topo.boundary_dofs)
        om = core.omega_of(freqs[i % 4])
        vals = asm.assemble(om, ent_bc=ent_bc)
        el = int(core.locate_points(mesh, np.array([[0.0, 0.0, -950.0]])).cpu()[0])
        b = core.source_vector(mesh, el, [0.0, 0.0, -950.0], (1.0, 0.0, 0.0), om)
        core.apply_dirichlet(asm.plan, None, dof_bc, b)
        out.append(dict(mesh=mesh, asm=asm, plan=asm.plan, vals=vals, b=b, omega=om,
                        ent_bc=ent_bc, order=mesh.entity_order()))
    return out
