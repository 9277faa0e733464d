"""Device-resident CSEM system: mesh arrays, assembly, source, BC, solve, receivers.

Thin host orchestration over the C ABI; every array that the hot path touches
lives in HBM as a torch tensor on the current HIP device, and all arithmetic
happens in the HIP kernels of ``libpetgem_amd.so``.
"""
import ctypes
import math

import numpy as np
import torch

from . import _native as nat
from .mesh import dof_counts

MU0 = 4.0 * math.pi * 1e-7


def _dev(device):
    return torch.device(device if device is not None else "cuda")


def _i32(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).to(dev)


def _f64(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)


def omega_of(frequency):
    """Angular frequency exactly as the reference forms it (``solver.py:174``)."""
    return float(frequency) * 2.0 * math.pi


def dipole_moment(source):
    """I*L times the rotated unit vector (``solver.py:253-268``, ``hvfem.py:2303-2344``)."""
    az = math.radians(source.get("azimuth", 0.0))
    dip = math.radians(source.get("dip", 0.0))
    a = az
    b = dip
    rot = np.array([math.cos(a) * math.cos(b), math.sin(a) * math.cos(b), math.sin(b)])
    il = float(source.get("current", 1.0)) * float(source.get("length", 1.0))
    return il * rot


def _spread3(v):
    """Spread the low 10 bits of v so that two zero bits follow each one."""
    v = v.astype(np.uint64) & np.uint64(0x3FF)
    v = (v | (v << np.uint64(16))) & np.uint64(0x030000FF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x0300F00F)
    v = (v | (v << np.uint64(4))) & np.uint64(0x030C30C3)
    v = (v | (v << np.uint64(2))) & np.uint64(0x09249249)
    return v


def _morton3(x, y, z):
    return _spread3(x) | (_spread3(y) << np.uint64(1)) | (_spread3(z) << np.uint64(2))


class DeviceMesh:
    """Per-element arrays in the reference's scratch layout, resident on the device."""

    def __init__(self, p, coords, elems_nodes, elems_edges, edges_nodes, faces_edges, dofs,
                 n_edges, n_faces, elems_faces, device=None):
        dev = _dev(device)
        self.device = dev
        self.p = int(p)
        self.n_elems = int(np.asarray(elems_nodes).shape[0])
        self.n_edges = int(n_edges)
        self.n_faces = int(n_faces)
        de, df, dv, n = dof_counts(self.p)
        self.n_dof_elem = n
        self.ndof = self.n_edges * de + self.n_faces * df + self.n_elems * dv
        self.coords = _f64(np.asarray(coords).reshape(self.n_elems, 12), dev)
        self.elems_nodes = _i32(elems_nodes, dev)
        self.elems_edges = _i32(elems_edges, dev)
        self.edges_nodes = _i32(np.asarray(edges_nodes).reshape(self.n_elems, 12), dev)
        self.faces_edges = _i32(np.asarray(faces_edges).reshape(self.n_elems, 12), dev)
        self.dofs = _i32(dofs, dev)
        # host copies needed by the (host-side, integer) assembly plan and the
        # entity schedule
        self.h_elems_edges = np.ascontiguousarray(elems_edges, dtype=np.int32)
        self.h_elems_faces = np.ascontiguousarray(elems_faces, dtype=np.int32)
        self.h_centroids = np.asarray(coords, dtype=np.float64).reshape(self.n_elems, 4, 3).mean(1)

    @classmethod
    def from_topology(cls, topo, device=None):
        return cls(topo.p, topo.coords, topo.elems_nodes, topo.elems_edges,
                   topo.elem_edges_nodes, topo.elem_faces_edges, topo.dofs, topo.n_edges,
                   topo.n_faces, topo.elems_faces, device)

    def block_ptr(self):
        """DOF ranges of the mesh entities (edges, faces, volumes): entity-block Jacobi blocks."""
        de, df, dv, _ = dof_counts(self.p)
        sizes = [np.full(self.n_edges, de)]
        if df:
            sizes.append(np.full(self.n_faces, df))
        if dv:
            sizes.append(np.full(self.n_elems, dv))
        return np.concatenate([[0], np.cumsum(np.concatenate(sizes))]).astype(np.int32)

    def entity_order(self):
        """The DOF-carrying entities (block_ptr order: edges, faces, volumes) sorted
        by the 3-D Morton code of their position (the centroid of an element that
        contains them, on a 1024^3 grid over the mesh's bounding box).  Visiting
        the entity rows in this order keeps the SpMV's gathers spatially local."""
        de, df, dv, _ = dof_counts(self.p)
        c = self.h_centroids
        parts = []
        pe = np.empty((self.n_edges, 3))
        pe[self.h_elems_edges.ravel()] = np.repeat(c, 6, axis=0)
        parts.append(pe)
        if df:
            pf = np.empty((self.n_faces, 3))
            pf[self.h_elems_faces.ravel()] = np.repeat(c, 4, axis=0)
            parts.append(pf)
        if dv:
            parts.append(c)
        pos = np.concatenate(parts)
        lo, hi = pos.min(0), pos.max(0)
        q = np.clip(((pos - lo) / np.maximum(hi - lo, 1e-300) * 1023.0).astype(np.int64), 0, 1023)
        return np.argsort(_morton3(q[:, 0], q[:, 1], q[:, 2]), kind="stable").astype(np.int32)

    def entity_mask(self, boundary_dofs):
        """(per-entity, per-DOF) uint8 boundary masks from a list of boundary DOFs."""
        de, df, dv, _ = dof_counts(self.p)
        dof_bc = np.zeros(self.ndof, dtype=np.uint8)
        dof_bc[np.asarray(boundary_dofs, dtype=np.int64)] = 1
        ent = np.zeros(self.n_edges + self.n_faces + self.n_elems, dtype=np.uint8)
        ent[: self.n_edges] = dof_bc[0: self.n_edges * de: de]
        if df:
            f0 = self.n_edges * de
            ent[self.n_edges: self.n_edges + self.n_faces] = dof_bc[f0: f0 + self.n_faces * df: df]
        return (torch.from_numpy(ent).to(self.device), torch.from_numpy(dof_bc).to(self.device))


class Tables:
    """Device-resident reference-element tables for order p."""

    def __init__(self, p):
        L = nat.lib()
        h = ctypes.c_void_p()
        nat.check(L.pg_tables_create(int(p), ctypes.byref(h)))
        self.handle = h
        self.p = p

    def __del__(self):
        try:
            if self.handle:
                nat.lib().pg_tables_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class Plan:
    """Assembly plan (CSR pattern + per-block contribution lists) on the device."""

    def __init__(self, mesh):
        L = nat.lib()
        h = ctypes.c_void_p()
        nat.check(L.pg_plan_build(mesh.p, mesh.n_elems, mesh.n_edges, mesh.n_faces,
                                  nat.ptr(mesh.h_elems_edges), nat.ptr(mesh.h_elems_faces),
                                  ctypes.byref(h)))
        try:
            info = np.zeros(6, dtype=np.int64)
            nat.check(L.pg_plan_info(h, nat.ptr(info)))
            self.ndof, self.nnz, self.npairs, self.ncontrib, self.n_entities, self.maxd = (
                int(v) for v in info)
            rowptr = np.empty(self.ndof + 1, dtype=np.int32)
            cols = np.empty(self.nnz, dtype=np.int32)
            slot0 = np.empty(self.npairs, dtype=np.int32)
            rowlen = np.empty(self.npairs, dtype=np.int32)
            cbeg = np.empty(self.npairs + 1, dtype=np.int32)
            ab = np.empty(2 * self.npairs, dtype=np.int32)
            contrib = np.empty(self.ncontrib, dtype=np.uint32)
            nat.check(L.pg_plan_export(h, nat.ptr(rowptr), nat.ptr(cols), nat.ptr(slot0),
                                       nat.ptr(rowlen), nat.ptr(cbeg), nat.ptr(ab),
                                       nat.ptr(contrib)))
            # first pair per entity, when every entity fits the row-major
            # numeric assembly (else the two-phase kernels run)
            epair = np.empty(self.n_entities + 1, dtype=np.int32)
            row_max = ctypes.c_int32(0)
            rc = L.pg_plan_entity_pairs(h, nat.ptr(epair), ctypes.byref(row_max))
            if rc == nat.PG_ERR_UNSUPPORTED:
                epair = None
            else:
                nat.check(rc)
        finally:
            L.pg_plan_destroy(h)
        if self.ndof != mesh.ndof:
            raise RuntimeError(f"plan ndof {self.ndof} != mesh ndof {mesh.ndof}")
        dev = mesh.device
        self.h_rowptr, self.h_cols = rowptr, cols
        self.rowptr = torch.from_numpy(rowptr).to(dev)
        self.cols = torch.from_numpy(cols).to(dev)
        self.slot0 = torch.from_numpy(slot0).to(dev)
        self.rowlen = torch.from_numpy(rowlen).to(dev)
        self.cbeg = torch.from_numpy(cbeg).to(dev)
        self.ab = torch.from_numpy(ab).to(dev)
        self.contrib = torch.from_numpy(contrib.view(np.int32)).to(dev)
        self.entity_pair = None if epair is None else torch.from_numpy(epair).to(dev)
        self.dev_struct = nat.PlanDev(
            mesh.p, mesh.n_edges, mesh.n_faces, mesh.n_elems, self.ndof, self.nnz, self.npairs,
            self.rowptr.data_ptr(), self.cols.data_ptr(), self.slot0.data_ptr(),
            self.rowlen.data_ptr(), self.cbeg.data_ptr(), self.ab.data_ptr(),
            self.contrib.data_ptr(),
            None if self.entity_pair is None else self.entity_pair.data_ptr(),
            0 if self.entity_pair is None else row_max.value)


class Assembler:
    """Element geometry + fused atomic-free assembly for one mesh and order."""

    def __init__(self, mesh, tables=None, plan=None):
        self.mesh = mesh
        self.tables = tables if tables is not None else Tables(mesh.p)
        self.plan = plan if plan is not None else Plan(mesh)
        nbytes = int(nat.lib().pg_element_workspace_bytes(mesh.n_elems))
        self.workspace = torch.empty(nbytes, dtype=torch.uint8, device=mesh.device)
        self._sigma = None

    def geometry(self, sigma, stream=None):
        """Fill the per-element workspace (Jacobian coefficients, sigma, orientation)."""
        m = self.mesh
        sig = sigma if torch.is_tensor(sigma) else _f64(sigma, m.device)
        self._sigma = sig
        nat.check(nat.lib().pg_element_geometry(
            m.p, m.n_elems, nat.ptr(m.coords), nat.ptr(m.elems_nodes), nat.ptr(m.elems_edges),
            nat.ptr(m.edges_nodes), nat.ptr(m.faces_edges), nat.ptr(sig), nat.ptr(self.workspace),
            nat.stream_handle(stream)))

    def element_matrices(self, stream=None):
        m = self.mesh
        n = m.n_dof_elem
        M = torch.empty((m.n_elems, n, n), dtype=torch.float64, device=m.device)
        K = torch.empty_like(M)
        nat.check(nat.lib().pg_element_matrices(self.tables.handle, m.n_elems,
                                                nat.ptr(self.workspace), nat.ptr(M), nat.ptr(K),
                                                nat.stream_handle(stream)))
        return M, K

    def assemble(self, omega, ent_bc=None, out=None, stream=None):
        """CSR values (nnz,2) float64 = complex128 interleaved."""
        if out is None:
            out = torch.empty((self.plan.nnz, 2), dtype=torch.float64, device=self.mesh.device)
        nat.check(nat.lib().pg_assemble(self.tables.handle, ctypes.byref(self.plan.dev_struct),
                                        nat.ptr(self.workspace), float(omega), nat.ptr(ent_bc),
                                        nat.ptr(out), nat.stream_handle(stream)))
        return out


def assemble_km(assembler, ent_bc=None, out=None, stream=None):
    """Stiffness and mass parts of the assembled system, each (nnz,) float64
    (written into ``out = (K, M)`` when given)."""
    pl = assembler.plan
    if out is None:
        K = torch.empty(pl.nnz, dtype=torch.float64, device=assembler.mesh.device)
        M = torch.empty_like(K)
    else:
        K, M = out
    nat.check(nat.lib().pg_assemble_km(assembler.tables.handle, ctypes.byref(pl.dev_struct),
                                       nat.ptr(assembler.workspace), nat.ptr(ent_bc), nat.ptr(K),
                                       nat.ptr(M), nat.stream_handle(stream)))
    return K, M


class BatchSolver:
    """Many right-hand sides (each with its own frequency) in one batched COCG."""

    def __init__(self, plan, K, M, frequencies, rtol=1e-8, maxit=100000, block_ptr=None,
                 entity_order=None, precision="fp64", inner_rtol=1e-4, stall_iters=0,
                 pc="jacobi", eop=None, stream=None):
        """eop: an ElementOperator of the same system (p = 1, 2; 64, 128, 192 or
        256 right-hand sides): the iterations apply A_r element by element,
        forming each element's entries once for all columns."""
        self.plan, self.K, self.M = plan, K, M
        self.omega = np.array([omega_of(f) for f in frequencies], dtype=np.float64)
        self.nrhs = self.omega.size
        h = ctypes.c_void_p()
        nat.check(nat.lib().pg_batch_create(plan.ndof, plan.nnz, nat.ptr(plan.rowptr),
                                            nat.ptr(plan.cols), nat.ptr(K), nat.ptr(M),
                                            self.nrhs, nat.ptr(self.omega), float(rtol), int(maxit),
                                            ctypes.byref(h), nat.stream_handle(stream)))
        self.handle = h
        if block_ptr is not None:
            bp = np.ascontiguousarray(block_ptr, dtype=np.int32)
            od = None if entity_order is None else np.ascontiguousarray(entity_order, np.int32)
            if od is not None and od.size != bp.size - 1:
                raise ValueError("entity_order needs one entry per block_ptr entity")
            nat.check(nat.lib().pg_batch_set_entities(self.handle, bp.ctypes.data, bp.size - 1,
                                                      None if od is None else od.ctypes.data,
                                                      nat.stream_handle(stream)))
        if pc == "bjacobi":
            if block_ptr is None:
                raise ValueError("block Jacobi needs block_ptr (entity rows)")
            nat.check(nat.lib().pg_batch_set_block_jacobi(self.handle, nat.stream_handle(stream)))
        elif pc != "jacobi":
            raise ValueError(f"pc must be jacobi or bjacobi, not {pc}")
        if precision == "mixed":
            if block_ptr is None:
                raise ValueError("the mixed-precision batch needs block_ptr (entity rows)")
            nat.check(nat.lib().pg_batch_set_mixed(self.handle, float(inner_rtol),
                                                   int(stall_iters), nat.stream_handle(stream)))
        elif precision != "fp64":
            raise ValueError(f"precision must be fp64 or mixed, not {precision}")
        self.eop = eop  # kept alive as long as the batch
        if eop is not None:
            nat.check(nat.lib().pg_batch_set_element_operator(self.handle, eop.handle,
                                                              nat.ptr(eop._ws),
                                                              nat.stream_handle(stream)))

    def refresh(self, stream=None):
        """Re-derive the Jacobi diagonals after K and M were reassembled in place
        (with an element operator, refresh the operator first)."""
        nat.check(nat.lib().pg_batch_refresh(self.handle, nat.stream_handle(stream)))

    def solve(self, B, stream=None):
        """B: (ndof, nrhs, 2) float64 (complex interleaved, row-major); returns X, info."""
        if tuple(B.shape) != (self.plan.ndof, self.nrhs, 2) or not B.is_contiguous():
            raise ValueError("B must be a contiguous (ndof, nrhs, 2) float64 tensor")
        X = torch.empty_like(B)
        it = np.zeros(self.nrhs, dtype=np.int32)
        rel = np.zeros(self.nrhs, dtype=np.float64)
        st = np.zeros(self.nrhs, dtype=np.int32)
        nat.check(nat.lib().pg_batch_solve(self.handle, nat.ptr(B), nat.ptr(X), nat.ptr(it),
                                           nat.ptr(rel), nat.ptr(st), nat.stream_handle(stream)))
        return X, dict(iterations=it, relres=rel, status=st)

    def __del__(self):
        try:
            if self.handle:
                nat.lib().pg_batch_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class MultiMeshSolver:
    """Independent meshes of uneven size solved together (one COCG recurrence per
    mesh, all advancing in the same launches; the iterations between host
    checks replayed from a captured HIP graph).  Point Jacobi and CSR products,
    or with `element` entity-block Jacobi and the element products of all
    meshes in one launch.  With lanes > 1 the meshes are dealt to that many
    such batches (largest first, each to the lane with the fewest DOFs so
    far), which run concurrently on their own streams; a mesh's result does
    not depend on the other meshes of its batch, so it is bitwise the same."""

    def __init__(self, systems, rtol=1e-8, maxit=100000, check_every=25, use_graph=True,
                 element=None, precision="fp64", inner_rtol=1e-4, stall_iters=800, lanes=1,
                 stream=None):
        """systems: list of (plan, values) with values (nnz, 2) float64 on the device.
        element: optional list, one per mesh in the order of `systems`, of
        (assembler, omega, ent_bc, entity_order) of the same systems (the
        assembler after its geometry; ent_bc and entity_order may be None).
        precision "mixed" (with `element`): complex64 rounds under FP64
        refreshes per mesh (inner_rtol, stall_iters as in KrylovSolver at p <= 2)."""
        if precision not in ("fp64", "mixed"):
            raise ValueError(f"precision must be 'fp64' or 'mixed', not {precision!r}")
        if precision == "mixed" and element is None:
            raise ValueError("precision='mixed' needs the element operator (element=...)")
        self.handle = None
        self.eop = None
        self.lanes = None
        k = len(systems)
        lanes = max(1, min(int(lanes), k))
        if lanes > 1:
            nd = [pl.ndof for pl, _ in systems]
            load, groups = [0] * lanes, [[] for _ in range(lanes)]
            for i in sorted(range(k), key=lambda i: -nd[i]):
                j = int(np.argmin(load))
                groups[j].append(i)
                load[j] += nd[i]
            self.groups = [sorted(g) for g in groups]
            self.lanes = [MultiMeshSolver([systems[i] for i in g], rtol, maxit, check_every,
                                          use_graph, None if element is None else
                                          [element[i] for i in g], precision, inner_rtol,
                                          stall_iters, 1, stream) for g in self.groups]
            self.k = k
            self.n = np.array([pl.ndof for pl, _ in systems], dtype=np.int64)
            return
        self.n = np.array([pl.ndof for pl, _ in systems], dtype=np.int64)
        nnz = np.array([pl.nnz for pl, _ in systems], dtype=np.int64)
        arr = lambda ts: (ctypes.c_void_p * k)(*[nat.ptr(t) for t in ts])  # noqa: E731
        rp = arr([pl.rowptr for pl, _ in systems])
        ci = arr([pl.cols for pl, _ in systems])
        va = arr([v for _, v in systems])
        h = ctypes.c_void_p()
        nat.check(nat.lib().pg_multi_create(k, self.n.ctypes.data, nnz.ctypes.data, rp, ci, va,
                                            float(rtol), int(maxit), int(check_every),
                                            1 if use_graph else 0, ctypes.byref(h),
                                            nat.stream_handle(stream)))
        self.handle = h
        self.k = k
        if element is not None:
            self._set_element(element, stream)
        if precision == "mixed":
            nat.check(nat.lib().pg_multi_set_mixed(self.handle, float(inner_rtol),
                                                   int(stall_iters), nat.stream_handle(stream)))

    def _set_element(self, element, stream):
        if len(element) != self.k:
            raise ValueError(f"need one (assembler, omega, ent_bc, order) per mesh ({self.k})")
        meshes = [a.mesh for a, _, _, _ in element]
        p = meshes[0].p
        if any(m.p != p for m in meshes):
            raise ValueError("the meshes of a multi-mesh element operator share one order p")
        k = self.k
        bps = [np.ascontiguousarray(m.block_ptr(), np.int32) for m in meshes]
        ords = [None if o is None else np.ascontiguousarray(o, np.int32)
                for _, _, _, o in element]
        dev_arr = lambda ts: (ctypes.c_void_p * k)(*[nat.ptr(t) for t in ts])  # noqa: E731
        host_arr = lambda xs: (ctypes.c_void_p * k)(  # noqa: E731
            *[None if x is None else x.ctypes.data for x in xs])
        ne = np.array([m.n_elems for m in meshes], dtype=np.int64)
        nd = np.array([m.ndof for m in meshes], dtype=np.int64)
        nb = np.array([b.size - 1 for b in bps], dtype=np.int64)
        om = np.array([float(o) for _, o, _, _ in element], dtype=np.float64)
        if not np.array_equal(nd, self.n):
            raise ValueError("element operator meshes differ from the systems")
        h = ctypes.c_void_p()
        nat.check(nat.lib().pg_eop_create_multi(
            p, k, ne.ctypes.data, dev_arr([a.workspace for a, _, _, _ in element]),
            dev_arr([m.dofs for m in meshes]), nd.ctypes.data, host_arr(bps), nb.ctypes.data,
            host_arr(ords), dev_arr([e for _, _, e, _ in element]), om.ctypes.data,
            ctypes.byref(h), nat.stream_handle(stream)))
        self.eop = h
        nat.check(nat.lib().pg_multi_set_element_operator(self.handle, self.eop,
                                                          nat.stream_handle(stream)))

    def _solve_lanes(self, bs):
        import threading
        torch.cuda.synchronize()  # right-hand sides complete
        res, errors = [None] * len(self.lanes), []

        def work(j):
            try:
                res[j] = self.lanes[j].solve([bs[i] for i in self.groups[j]])
            except Exception as e:  # reported after every lane has stopped
                errors.append(e)

        threads = [threading.Thread(target=work, args=(j,)) for j in range(len(self.lanes))]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        xs = [None] * self.k
        it = np.zeros(self.k, dtype=np.int32)
        rel = np.zeros(self.k, dtype=np.float64)
        st = np.zeros(self.k, dtype=np.int32)
        launches = 0
        for g, (xg, ig) in zip(self.groups, res):
            for a, i in enumerate(g):
                xs[i] = xg[a]
                it[i], rel[i], st[i] = ig["iterations"][a], ig["relres"][a], ig["status"][a]
            launches = max(launches, ig["launches"])
        return xs, dict(iterations=it, relres=rel, status=st, launches=launches)

    def solve(self, bs, stream=None):
        """bs: list of (ndof_i, 2) float64 right-hand sides; returns xs, info."""
        if len(bs) != self.k:
            raise ValueError(f"need {self.k} right-hand sides")
        if self.lanes is not None:
            return self._solve_lanes(bs)
        for b, n in zip(bs, self.n):
            if tuple(b.shape) != (int(n), 2) or b.dtype != torch.float64 or not b.is_contiguous():
                raise ValueError("each b must be a contiguous (ndof, 2) float64 tensor")
        xs = [torch.empty_like(b) for b in bs]
        arr = lambda ts: (ctypes.c_void_p * self.k)(*[nat.ptr(t) for t in ts])  # noqa: E731
        it = np.zeros(self.k, dtype=np.int32)
        rel = np.zeros(self.k, dtype=np.float64)
        st = np.zeros(self.k, dtype=np.int32)
        launches = ctypes.c_int32()
        nat.check(nat.lib().pg_multi_solve(self.handle, arr(bs), arr(xs), nat.ptr(it),
                                           nat.ptr(rel), nat.ptr(st), ctypes.byref(launches),
                                           nat.stream_handle(stream)))
        return xs, dict(iterations=it, relres=rel, status=st, launches=launches.value)

    def __del__(self):
        try:
            if self.handle:
                nat.lib().pg_multi_destroy(self.handle)
                self.handle = None
            if self.eop:  # after the batch that uses it
                nat.lib().pg_eop_destroy(self.eop)
                self.eop = None
        except Exception:
            pass


def apply_dirichlet(plan, values, dof_bc, b=None, stream=None):
    nat.check(nat.lib().pg_apply_dirichlet(plan.ndof, nat.ptr(plan.rowptr), nat.ptr(plan.cols),
                                           nat.ptr(dof_bc), nat.ptr(values), nat.ptr(b),
                                           nat.stream_handle(stream)))


def locate_points(mesh, points, tol=1e-12, stream=None):
    pts = _f64(np.atleast_2d(points), mesh.device)
    out = torch.empty(pts.shape[0], dtype=torch.int32, device=mesh.device)
    nat.check(nat.lib().pg_locate_points(mesh.n_elems, nat.ptr(mesh.coords), pts.shape[0],
                                         nat.ptr(pts), float(tol), nat.ptr(out),
                                         nat.stream_handle(stream)))
    return out


def source_vector(mesh, element, position, moment, omega, b=None, stream=None):
    """Add the point-dipole contribution to b (complex128 as (N,2) float64)."""
    element = int(element)
    if not 0 <= element < mesh.n_elems:
        raise nat.PgError(nat.PG_ERR_NOT_FOUND,
                          f"source element {element} not in the mesh (source outside the domain?)")
    if b is None:
        b = torch.zeros((mesh.ndof, 2), dtype=torch.float64, device=mesh.device)
    elif tuple(b.shape) != (mesh.ndof, 2) or b.dtype != torch.float64:
        raise ValueError("b must be a (ndof,2) float64 tensor")
    pos = _f64(np.asarray(position, dtype=np.float64), mesh.device)
    mom = _f64(np.asarray(moment, dtype=np.float64), mesh.device)
    nat.check(nat.lib().pg_source_vector(mesh.p, int(element), nat.ptr(mesh.coords),
                                         nat.ptr(mesh.elems_nodes), nat.ptr(mesh.elems_edges),
                                         nat.ptr(mesh.edges_nodes), nat.ptr(mesh.faces_edges),
                                         nat.ptr(mesh.dofs), nat.ptr(pos), nat.ptr(mom),
                                         float(omega), nat.ptr(b), nat.stream_handle(stream)))
    return b


def receiver_fields(mesh, points, elements, x, omega, stream=None):
    pts = _f64(np.atleast_2d(points), mesh.device)
    el = elements if torch.is_tensor(elements) else _i32(elements, mesh.device)
    el_h = el.detach().cpu().numpy()
    if el_h.shape[0] != pts.shape[0] or (el_h >= mesh.n_elems).any() or el.dtype != torch.int32:
        raise ValueError("receiver elements must be int32, one per point, < n_elems (-1 = none)")
    if tuple(x.shape) != (mesh.ndof, 2) or x.dtype != torch.float64:
        raise ValueError("x must be a (ndof,2) float64 tensor")
    out = torch.empty((pts.shape[0], 6, 2), dtype=torch.float64, device=mesh.device)
    nat.check(nat.lib().pg_receiver_fields(mesh.p, pts.shape[0], nat.ptr(pts), nat.ptr(el),
                                           nat.ptr(mesh.coords), nat.ptr(mesh.elems_nodes),
                                           nat.ptr(mesh.elems_edges), nat.ptr(mesh.edges_nodes),
                                           nat.ptr(mesh.faces_edges), nat.ptr(mesh.dofs),
                                           nat.ptr(x), float(omega), nat.ptr(out),
                                           nat.stream_handle(stream)))
    return out


_KSP = {"cocg": nat.PG_KSP_COCG, "gmres": nat.PG_KSP_GMRES}
_PC = {"none": nat.PG_PC_NONE, "jacobi": nat.PG_PC_JACOBI, "bjacobi": nat.PG_PC_BJACOBI}
_PREC = {"fp64": nat.PG_PREC_FP64, "mixed": nat.PG_PREC_MIXED}


_SPMV = {"auto": 0, "row": 1, "entity": 2, "entity_pairs": 3, "element": 4}


class ElementOperator:
    """Element-by-element product y = A x of the system Assembler.assemble builds
    (A = sum_e K_e - i omega mu0 M_e, PEC entities as identity rows), without
    the assembled matrix: the element matrices are recomputed from the geometry
    workspace and the tables at every product.  Orders p = 1, 2 (complex128 and
    complex64) and 3 (complex64).  Passed to KrylovSolver(eop=...), whose
    iterations then use it when it is the faster kernel (or always, with
    spmv="element")."""

    def __init__(self, assembler, omega, ent_bc=None, entity_order=None, stream=None):
        m = assembler.mesh
        self.mesh = m
        self.ndof = m.ndof
        self._bptr = np.ascontiguousarray(m.block_ptr(), np.int32)
        self._order = None if entity_order is None else np.ascontiguousarray(entity_order,
                                                                            np.int32)
        self._ent_bc = ent_bc
        self._ws = assembler.workspace
        h = ctypes.c_void_p()
        nat.check(nat.lib().pg_eop_create(
            m.p, m.n_elems, nat.ptr(self._ws), nat.ptr(m.dofs), m.ndof,
            self._bptr.ctypes.data, self._bptr.size - 1,
            None if self._order is None else self._order.ctypes.data, nat.ptr(ent_bc),
            float(omega), ctypes.byref(h), nat.stream_handle(stream)))
        self.handle = h

    def refresh(self, omega, stream=None):
        """Re-read the element coefficients (after Assembler.geometry with a new
        conductivity, or for another frequency)."""
        nat.check(nat.lib().pg_eop_refresh(self.handle, nat.ptr(self._ws), float(omega),
                                           nat.stream_handle(stream)))

    def apply(self, x, stream=None):
        """y = A x; x (ndof, 2) float64 (complex128) or float32 (complex64)."""
        y = torch.empty_like(x)
        fn = nat.lib().pg_eop_apply_z if x.dtype == torch.float64 else nat.lib().pg_eop_apply_c
        nat.check(fn(self.handle, nat.ptr(x), nat.ptr(y), nat.stream_handle(stream)))
        return y

    def __del__(self):
        try:
            if self.handle:
                nat.lib().pg_eop_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def mixed_defaults(p):
    """(inner_rtol, stall_iters) of the single-RHS solver's complex64 rounds by
    basis order.  A round is watched 800 iterations for stagnation (at 400 the
    low-frequency rounds were cut before their tenfold gain and handed over to
    FP64: C4 at 0.1 Hz 0.505 -> 0.369 s per right-hand side, p = 3 at
    0.25 Hz 2.59 -> 1.69 s; no change at 1 Hz).  At p = 3 the 45-DOF complex64
    element products carry more rounding, so a round aims for 1e-3 (0.98 ->
    0.81 s at 1 Hz); p <= 2 keeps 1e-4 (best at p = 2)."""
    return (1e-3, 800) if int(p) >= 3 else (1e-4, 800)


class KrylovSolver:
    """Device Krylov solver for a complex symmetric CSR matrix (values complex128).
    inner_rtol / stall_iters (mixed precision) default by the plan's basis
    order (mixed_defaults)."""

    def __init__(self, plan, values, ksp="cocg", pc="jacobi", precision="fp64", rtol=1e-8,
                 maxit=100000, check_every=25, inner_rtol=None, restart=30, block_ptr=None,
                 spmv="auto", entity_order=None, stall_iters=None, eop=None, like=None,
                 stream=None):
        dev_struct = getattr(plan, "dev_struct", None)
        d_inner, d_stall = mixed_defaults(dev_struct.p if dev_struct is not None else 2)
        inner_rtol = d_inner if inner_rtol is None else inner_rtol
        stall_iters = d_stall if stall_iters is None else stall_iters
        self.plan = plan
        self.values = values
        self.eop = eop  # kept alive as long as the solver
        if pc == "bjacobi" and block_ptr is None:
            raise ValueError("pc='bjacobi' needs block_ptr (DeviceMesh.block_ptr())")
        # block_ptr (the mesh entities' DOF ranges) also enables the entity-row SpMV
        self._bptr = None if block_ptr is None else np.ascontiguousarray(block_ptr, np.int32)
        nb = 0 if self._bptr is None else self._bptr.size - 1
        # entity_order (DeviceMesh.entity_order()): the entity-row SpMV's visiting order
        self._order = None if entity_order is None else np.ascontiguousarray(entity_order,
                                                                            np.int32)
        if self._order is not None and self._order.size != nb:
            raise ValueError("entity_order needs one entry per block_ptr entity")
        self.opts = nat.SolveOpts(_KSP[ksp], _PC[pc], _PREC[precision], int(restart), int(maxit),
                                  int(check_every), float(rtol), float(inner_rtol),
                                  None if self._bptr is None else self._bptr.ctypes.data, nb,
                                  _SPMV[spmv], int(stall_iters),
                                  None if self._order is None else self._order.ctypes.data,
                                  None if eop is None else eop.handle,
                                  None if like is None else like.handle)
        h = ctypes.c_void_p()
        nat.check(nat.lib().pg_solver_create(plan.ndof, plan.nnz, nat.ptr(plan.rowptr),
                                             nat.ptr(plan.cols), nat.ptr(values),
                                             ctypes.byref(self.opts), ctypes.byref(h),
                                             nat.stream_handle(stream)))
        self.handle = h

    def refresh(self, stream=None):
        """Re-derive the preconditioner (and complex64 copy) after reassembly in place."""
        nat.check(nat.lib().pg_solver_refresh(self.handle, nat.stream_handle(stream)))

    def profile(self, reps=20, stream=None):
        """Average device time (ms) of this solver's SpMV kernel(s), HIP events on `stream`."""
        ms64, ms32 = ctypes.c_double(), ctypes.c_double()
        g64, g32 = ctypes.c_int32(), ctypes.c_int32()
        nat.check(nat.lib().pg_solver_profile(self.handle, int(reps), ctypes.byref(ms64),
                                              ctypes.byref(ms32), ctypes.byref(g64),
                                              ctypes.byref(g32), nat.stream_handle(stream)))
        return dict(ms64=ms64.value, ms32=ms32.value, g64=g64.value, g32=g32.value)

    def solve(self, b, x=None, stream=None):
        if x is None:
            x = torch.zeros_like(b)
        info = nat.SolveInfo()
        nat.check(nat.lib().pg_solver_solve(self.handle, nat.ptr(b), nat.ptr(x),
                                            ctypes.byref(info), nat.stream_handle(stream)))
        return x, dict(iterations=info.iterations, refinements=info.refinements,
                       status=info.status, relres=info.relres, spmv_ms=info.spmv_ms,
                       spmv_samples=info.spmv_samples, fp64_fallback=info.fp64_fallback,
                       spmv_first_ms=info.spmv_first_ms)

    def __del__(self):
        try:
            if self.handle:
                nat.lib().pg_solver_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def _check_independent(solvers):
    """Solvers that may run at the same time must not share device state: each
    KrylovSolver owns its vectors and scalars, and each ElementOperator its
    local-result buffer.  Raises ValueError when the list names one solver, or
    one element operator, twice (the two would race and give wrong solutions
    without an error)."""
    seen_s, seen_e = set(), set()
    for S in solvers:
        h = S.handle.value
        if h in seen_s:
            raise ValueError("the same solver object would run on two streams at once: give "
                             "every concurrent worker its own KrylovSolver")
        seen_s.add(h)
        if getattr(S, "eop", None) is not None:
            he = S.eop.handle.value
            if he in seen_e:
                raise ValueError("two concurrent solvers share one ElementOperator (its "
                                 "local-result buffer is per operator): give each its own")
            seen_e.add(he)


def _run_on_streams(nworkers, rhs, pick, order=None):
    """Solve rhs[i] with the solver pick(k, i) on worker k (K = nworkers host
    threads, one HIP stream each; the library call releases the interpreter
    lock); each worker takes the next unsolved index of `order` (default
    0, 1, ...) when it is free.  Returns (solutions, infos); raises the first
    failure once every worker has stopped."""
    import threading
    K = max(1, min(int(nworkers), len(rhs)))
    streams = [torch.cuda.Stream() for _ in range(K)]
    out, infos, errors = [None] * len(rhs), [None] * len(rhs), []
    queue = list(range(len(rhs))) if order is None else list(order)
    lock = threading.Lock()
    torch.cuda.synchronize()  # right-hand sides and operators complete

    def work(k):
        try:
            with torch.cuda.stream(streams[k]):
                while True:
                    with lock:
                        if not queue:
                            break
                        i = queue.pop(0)
                    x = torch.zeros_like(rhs[i])
                    out[i], infos[i] = pick(k, i).solve(rhs[i], x, stream=streams[k])
            streams[k].synchronize()
        except Exception as e:  # reported after every worker has stopped
            errors.append(e)

    threads = [threading.Thread(target=work, args=(k,)) for k in range(K)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    torch.cuda.synchronize()
    if errors:
        raise errors[0]
    cur = torch.cuda.current_stream()
    for x in out:
        x.record_stream(cur)  # allocated on a worker stream, used from here on
    return out, infos


def solve_concurrently(solvers, rhs):
    """Solve A x_i = rhs[i] for every right-hand side, len(solvers) at a time:
    worker k drives solvers[k] on its own HIP stream, taking the next right-hand
    side whenever it is free.  The solvers must be independent objects of the same system
    (each with its own element operator, whose local-result buffer is per
    operator) with the same kernel choices (``KrylovSolver(like=...)``).  A
    solve's arithmetic does not depend on what runs beside it, so the
    solutions are then bitwise those of one solve after another.  Returns
    (solutions, infos) in the order of `rhs`; raises the first failure."""
    _check_independent(solvers)
    return _run_on_streams(len(solvers), rhs, lambda k, i: solvers[k])


def solve_each(solvers, rhs, workers=4):
    """Solve solvers[i] x_i = rhs[i] (independent systems, e.g. many meshes),
    `workers` at a time on their own streams, largest system first, each
    worker taking the next one when it is free.  Each solve is bitwise what it
    is on its own.  Returns (solutions, infos); raises the first failure."""
    if len(solvers) != len(rhs):
        raise ValueError("solve_each needs one solver per right-hand side")
    if workers > 1:
        _check_independent(solvers)
    order = sorted(range(len(rhs)), key=lambda i: -rhs[i].shape[0])
    return _run_on_streams(workers, rhs, lambda k, i: solvers[i], order)


def spmv(plan, values, x, stream=None):
    y = torch.empty_like(x)
    nat.check(nat.lib().pg_spmv_z(plan.ndof, nat.ptr(plan.rowptr), nat.ptr(plan.cols),
                                  nat.ptr(values), nat.ptr(x), nat.ptr(y),
                                  nat.stream_handle(stream)))
    return y


def as_complex(t):
    """(…,2) float64 tensor -> numpy complex128."""
    a = t.detach().cpu().numpy()
    return a[..., 0] + 1j * a[..., 1]


def from_complex(z, device):
    z = np.asarray(z, dtype=np.complex128)
    return torch.from_numpy(np.stack([z.real, z.imag], axis=-1).copy()).to(device)
