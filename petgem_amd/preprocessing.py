"""Preprocessing: mesh -> scratch files, the reference's stage-1 contract.

Restates ``Preprocessing.run`` (``petgem/preprocessing.py:40-538``): reads
the gmsh v2 mesh, numbers edges/faces/DOFs, builds the conductivity model,
the PEC boundary DOFs and the source-element record, and writes them into
``output.directory_scratch`` as PETSc binary files with the reference's names
and layouts, so that a reference ``Solver`` could read them and ours reads
the reference's:

``nodes.dat`` (E,12) | ``meshConnectivity.dat`` (E,4) | ``edges.dat`` (E,6) |
``edgesNodes.dat`` (E,12) | ``faces.dat`` (E,4) | ``facesEdges.dat`` (E,12) |
``dofs.dat`` (E,n) | ``conductivityModel.dat`` (E,2) | ``boundaries.dat`` |
``source.dat`` (50+n) | ``nnz.dat`` (N).

Differences from the reference, by design: the mesh reader is our own (no
meshio), HDF5 goes through ``petgem_amd.h5`` (no h5py needed) and the source
is located with the device kernel ``pg_locate_points`` (same rule as
``Delaunay.find_simplex(bruteforce=True, tol=1e-12)``: lowest index).
``boundaries.dat`` lists the boundary edge DOFs then the boundary face DOFs,
each block ascending, which is the order ``computeBoundaries``
(``mesh.py:280-321``) produces.
"""
import os

import numpy as np

from . import h5, petsc_io
from .common import Print, Timers
from .mesh import Topology, read_gmsh_v2

VALENCE = np.array([50, 200, 400, 800, 1400, 2500])  # preprocessing.py:470


def conductivity_model(inputSetup, n_elems, phys):
    model = inputSetup.model["csem"]
    if inputSetup.run.get("conductivity_from_file"):
        sig = h5.read_dataset(model["sigma"]["file"], "data")
        sig = np.atleast_2d(sig)
        if sig.shape != (n_elems, 2):
            raise ValueError(f"conductivity file must hold a (nElems,2) 'data' dataset, got {sig.shape}")
        return sig
    hor = np.asarray(model["sigma"]["horizontal"], dtype=np.float64)
    ver = np.asarray(model["sigma"]["vertical"], dtype=np.float64)
    idx = np.asarray(phys, dtype=np.int64) - 1
    if idx.min() < 0 or idx.max() >= hor.size or idx.max() >= ver.size:
        raise ValueError("physical group without a conductivity value in the params file")
    return np.stack([hor[idx], ver[idx]], axis=1)


def source_record(topo, element):
    """Source-element record (``preprocessing.py:427-451``)."""
    e = int(element)
    n = topo.n_dof_elem
    rec = np.zeros(50 + n)
    rec[0:4] = topo.elems_nodes[e]
    rec[4:16] = topo.coords[e].ravel()
    rec[16:20] = topo.elems_faces[e]
    rec[20:32] = topo.faces_edges[topo.elems_faces[e]].ravel()
    rec[32:38] = topo.elems_edges[e]
    rec[38:50] = topo.edges_nodes[topo.elems_edges[e]].ravel()
    rec[50:] = topo.dofs[e]
    return rec


def locate(topo, points):
    """Lowest-index containing element of each point (device kernel)."""
    from . import core
    import torch
    dev = torch.device("cuda")
    c = torch.from_numpy(np.ascontiguousarray(topo.coords.reshape(-1, 12))).to(dev)

    class _M:  # minimal view for core.locate_points
        n_elems = topo.n_elems
        coords = c
        device = dev
    return core.locate_points(_M, np.atleast_2d(points)).cpu().numpy()


class Preprocessing:
    def run(self, inputSetup):
        Timers()["Preprocessing"].start()
        model, run, output = inputSetup.model, inputSetup.run, inputSetup.output
        out = output["directory_scratch"]
        os.makedirs(out, exist_ok=True)
        p = int(run["nord"])
        Print.master("     Reading mesh")
        pts, en, phys = read_gmsh_v2(model["mesh"])
        topo = Topology(pts, en, p)
        E = topo.n_elems
        Print.master("     Writing scratch files")
        petsc_io.write_dense(os.path.join(out, "nodes.dat"), topo.coords.reshape(E, 12))
        petsc_io.write_dense(os.path.join(out, "meshConnectivity.dat"), en)
        petsc_io.write_dense(os.path.join(out, "edges.dat"), topo.elems_edges)
        petsc_io.write_dense(os.path.join(out, "edgesNodes.dat"), topo.elem_edges_nodes.reshape(E, 12))
        petsc_io.write_dense(os.path.join(out, "faces.dat"), topo.elems_faces)
        petsc_io.write_dense(os.path.join(out, "facesEdges.dat"), topo.elem_faces_edges.reshape(E, 12))
        petsc_io.write_dense(os.path.join(out, "dofs.dat"), topo.dofs)
        sigma = conductivity_model(inputSetup, E, phys)
        petsc_io.write_dense(os.path.join(out, "conductivityModel.dat"), sigma)
        petsc_io.write_vec(os.path.join(out, "boundaries.dat"), topo.boundary_dofs.astype(np.float64))
        src = model["csem"]["source"]
        el = int(locate(topo, np.asarray(src["position"], dtype=np.float64)[None])[0])
        if el < 0:
            raise ValueError("Source no located in the computational domain. Please, verify source "
                             "position or improve the mesh quality.")
        petsc_io.write_vec(os.path.join(out, "source.dat"), source_record(topo, el))
        petsc_io.write_vec(os.path.join(out, "nnz.dat"),
                           np.full(topo.ndof, VALENCE[p - 1], dtype=np.float64))
        receivers = h5.read_dataset(model["receivers"], "data")
        nrx = 1 if receivers.ndim == 1 else receivers.shape[0]
        Timers()["Preprocessing"].stop()
        Print.master(" ")
        Print.master("  Mesh statistics")
        Print.master(f"     Mesh file:            {model['mesh']}")
        Print.master(f"     Number of elements:   {E}")
        Print.master(f"     Number of faces:      {topo.n_faces}")
        Print.master(f"     Number of edges:      {topo.n_edges}")
        Print.master(f"     Number of dofs:       {topo.ndof}")
        Print.master(f"     Number of boundaries: {topo.boundary_dofs.size}")
        Print.master(f"     Number of receivers:  {nrx}")
        Print.master(f"     Source element:       {el}")
        self.topology = topo
        return topo
