"""The C-ABI library on the CPU: it loads, exports every declared entry point,
and its host-only parts (assembly plan, reference tables, argument errors)
match the oracle exactly.  No GPU work is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

from petgem_amd import _native as nat
from petgem_amd.mesh import Topology
from petgem_amd.synthetic import Workload
from oracle import basis as B, fem, quadrature as Q

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "petgem_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = nat.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
        assert s in nat.SIGNATURES, f"{s} has no ctypes signature"
    assert L.pg_abi_version() == 1


def _plan(T):
    L = nat.lib()
    h = ctypes.c_void_p()
    ee = np.ascontiguousarray(T.elems_edges, np.int32)
    ef = np.ascontiguousarray(T.elems_faces, np.int32)
    nat.check(L.pg_plan_build(T.p, T.n_elems, T.n_edges, T.n_faces, nat.ptr(ee), nat.ptr(ef),
                              ctypes.byref(h)))
    info = np.zeros(6, np.int64)
    nat.check(L.pg_plan_info(h, nat.ptr(info)))
    nd, nnz, npairs, nc = (int(v) for v in info[:4])
    out = dict(rowptr=np.empty(nd + 1, np.int32), cols=np.empty(nnz, np.int32),
               slot0=np.empty(npairs, np.int32), rowlen=np.empty(npairs, np.int32),
               cbeg=np.empty(npairs + 1, np.int32), ab=np.empty(2 * npairs, np.int32),
               contrib=np.empty(nc, np.uint32))
    nat.check(L.pg_plan_export(h, *[nat.ptr(out[k]) for k in
                                    ("rowptr", "cols", "slot0", "rowlen", "cbeg", "ab", "contrib")]))
    L.pg_plan_destroy(h)
    return info, out


@pytest.mark.parametrize("p", [1, 2, 3])
def test_plan_pattern_bit_exact(test_mesh, p):
    pts, en, _ = test_mesh
    T = Topology(pts, en, p)
    info, pl = _plan(T)
    indptr, cols = fem.csr_pattern(T.dofs, T.ndof)
    assert info[0] == T.ndof
    assert np.array_equal(pl["rowptr"], indptr) and np.array_equal(pl["cols"], cols)


def test_plan_contributions_cover_every_element_entry():
    """Every (element, i, j) lands in exactly one slot, listed in ascending element order."""
    w = Workload(3, 2)
    T = w.topo
    info, pl = _plan(T)
    n = T.n_dof_elem
    p = T.p
    ab = pl["ab"].reshape(-1, 2)
    count = np.zeros(pl["cols"].size, dtype=np.int64)
    loc_off = lambda la: la * p if la < 6 else 6 * p + (la - 6) * p * (p - 1)  # noqa: E731
    loc_n = lambda la: p if la < 6 else p * (p - 1)  # noqa: E731
    for k in range(ab.shape[0]):
        c0, c1 = pl["cbeg"][k], pl["cbeg"][k + 1]
        es = pl["contrib"][c0:c1] >> 8
        assert np.all(np.diff(es.astype(np.int64)) > 0)
        for u in pl["contrib"][c0:c1]:
            e, la, lb = int(u >> 8), int((u >> 4) & 15), int(u & 15)
            for a in range(loc_n(la)):
                for b in range(loc_n(lb)):
                    r, c = T.dofs[e, loc_off(la) + a], T.dofs[e, loc_off(lb) + b]
                    s = pl["slot0"][k] + a * pl["rowlen"][k] + b
                    assert pl["cols"][s] == c and pl["rowptr"][r] <= s < pl["rowptr"][r + 1]
                    count[s] += 1
    assert count.sum() == T.n_elems * n * n


@pytest.mark.parametrize("p", [1, 2, 3])
def test_plan_entity_pairs_index_each_entity_s_pairs(test_mesh, p):
    """entity_pair[A] .. entity_pair[A+1] are exactly the pairs (A, B).  At p <= 2
    the reference mesh fits the row-major assembly's bounds; at p = 3 its rows
    (up to 303 columns) do not, and the export says so."""
    pts, en, _ = test_mesh
    T = Topology(pts, en, p)
    ee = np.ascontiguousarray(T.elems_edges, np.int32)
    ef = np.ascontiguousarray(T.elems_faces, np.int32)
    L = nat.lib()
    h = ctypes.c_void_p()
    nat.check(L.pg_plan_build(T.p, T.n_elems, T.n_edges, T.n_faces, nat.ptr(ee), nat.ptr(ef),
                              ctypes.byref(h)))
    try:
        info = np.zeros(6, np.int64)
        nat.check(L.pg_plan_info(h, nat.ptr(info)))
        ep = np.empty(int(info[4]) + 1, np.int32)
        rmax = ctypes.c_int32(0)
        rc = L.pg_plan_entity_pairs(h, nat.ptr(ep), ctypes.byref(rmax))
    finally:
        L.pg_plan_destroy(h)
    if p == 3:
        assert rc == nat.PG_ERR_UNSUPPORTED and "row length 303" in L.pg_last_error().decode()
        return
    assert rc == 0
    assert rmax.value == np.diff(_plan(T)[1]["rowptr"]).max()
    _, pl = _plan(T)
    a = pl["ab"][0::2]
    assert ep[0] == 0 and ep[-1] == a.size
    assert np.array_equal(ep, np.searchsorted(a, np.arange(ep.size), side="left"))
    # every entity with DOFs has its diagonal pair among its pairs
    b = pl["ab"][1::2]
    for A in np.unique(a):
        assert (b[ep[A]:ep[A + 1]] == A).sum() == 1


@pytest.mark.parametrize("p,n", [(1, 4), (2, 4), (3, 3)])
def test_plan_invariants_of_the_row_major_assembly(p, n):
    """What the row-major assembly kernel indexes its LDS with: every entity
    with DOFs has exactly one diagonal pair, whose contributors are its
    elements in ascending order; its pairs' contributors number (elements) x
    (local entities), each (element, local entity) exactly once; every pair
    block lies inside the entity's row, at columns ascending with the pair."""
    T = Workload(n, p, seed=1).topo
    ee = np.ascontiguousarray(T.elems_edges, np.int32)
    ef = np.ascontiguousarray(T.elems_faces, np.int32)
    L = nat.lib()
    h = ctypes.c_void_p()
    nat.check(L.pg_plan_build(T.p, T.n_elems, T.n_edges, T.n_faces, nat.ptr(ee), nat.ptr(ef),
                              ctypes.byref(h)))
    try:
        info = np.zeros(6, np.int64)
        nat.check(L.pg_plan_info(h, nat.ptr(info)))
        ep = np.empty(int(info[4]) + 1, np.int32)
        rmax = ctypes.c_int32(0)
        nat.check(L.pg_plan_entity_pairs(h, nat.ptr(ep), ctypes.byref(rmax)))
    finally:
        L.pg_plan_destroy(h)
    _, pl = _plan(T)
    nE, nF = T.n_edges, T.n_faces
    de, df, dv = p, p * (p - 1), p * (p - 1) * (p - 2) // 2
    nl = 6 + (4 if df else 0) + (1 if dv else 0)
    nd = lambda A: de if A < nE else (df if A < nE + nF else dv)  # noqa: E731
    ab, slot0, cbeg, contrib = pl["ab"], pl["slot0"], pl["cbeg"], pl["contrib"]
    for A in range(ep.size - 1):
        if nd(A) == 0:
            assert ep[A] == ep[A + 1]
            continue
        r0 = A * de if A < nE else (nE * de + (A - nE) * df if A < nE + nF
                                    else nE * de + nF * df + (A - nE - nF) * dv)
        rowbase, rowlen = pl["rowptr"][r0], pl["rowptr"][r0 + 1] - pl["rowptr"][r0]
        assert 0 < rowlen <= rmax.value <= 192
        pis = np.arange(ep[A], ep[A + 1])
        assert 0 < pis.size <= 96 and (ab[2 * pis] == A).all()
        B = ab[2 * pis + 1]
        off = slot0[pis] - rowbase
        widths = np.array([nd(b) for b in B])
        assert off[0] == 0 and (np.diff(off) == widths[:-1]).all() and off[-1] + widths[-1] == rowlen
        (pd,) = np.nonzero(B == A)[0]
        inc = contrib[cbeg[pis[pd]]:cbeg[pis[pd] + 1]] >> 8
        assert 0 < inc.size <= 24 and (np.diff(inc.astype(np.int64)) > 0).all()
        c = contrib[cbeg[pis[0]]:cbeg[pis[-1] + 1]]
        assert c.size == inc.size * nl
        keys = (c >> 8).astype(np.int64) * 16 + (c & 15)
        assert np.unique(keys).size == c.size and np.isin(c >> 8, inc).all()


@pytest.mark.parametrize("p", [1, 2, 3])
def test_reference_tables_match_oracle_contraction(p):
    V = B.ndof_element(p) and (12 * p + 24 * p * (p - 1) + p * (p - 1) * (p - 2) // 2)
    T = np.zeros(V * V * 12)
    nat.check(nat.lib().pg_tables_host(p, nat.ptr(T), T.size))
    T = T.reshape(V, V, 12)
    pts, w = Q.rule_for_order(p)
    S = np.zeros((len(w), V, 3))
    C = np.zeros_like(S)
    for e in range(6):
        for o in range(2):
            v = (e * 2 + o) * p
            S[:, v:v + p], C[:, v:v + p] = B.edge_functions(pts, e, o, p)
    if p >= 2:
        for f in range(4):
            for o in range(6):
                v = 12 * p + (f * 6 + o) * p * (p - 1)
                S[:, v:v + p * (p - 1)], C[:, v:v + p * (p - 1)] = B.face_functions(pts, f, o, p)
    if p >= 3:
        S[:, 12 * p + 24 * p * (p - 1):], C[:, 12 * p + 24 * p * (p - 1):] = B.bubble_functions(pts, p)
    R = np.einsum("q,qia,qjb->ijab", w, S, S)
    RK = np.einsum("q,qia,qjb->ijab", w, C, C)
    combos = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    for c, (a, b) in enumerate(combos):
        refm = R[..., a, b] if a == b else R[..., a, b] + R[..., b, a]
        refk = RK[..., a, b] if a == b else RK[..., a, b] + RK[..., b, a]
        assert np.abs(T[..., c] - refm).max() <= 1e-14
        assert np.abs(T[..., 6 + c] - refk).max() <= 1e-14


def test_errors_are_reported():
    L = nat.lib()
    h = ctypes.c_void_p()
    ee = np.zeros((1, 6), np.int32)
    rc = L.pg_plan_build(4, 1, 6, 4, nat.ptr(ee), nat.ptr(ee), ctypes.byref(h))
    assert rc == nat.PG_ERR_UNSUPPORTED and b"order" in L.pg_last_error()
    with pytest.raises(nat.PgError):
        nat.check(L.pg_plan_build(2, 1, 6, 4, nat.ptr(np.full((1, 6), 9, np.int32)),
                                  nat.ptr(np.zeros((1, 4), np.int32)), ctypes.byref(h)))


def test_config_sizes_match_workload_table():
    """Exact sizes of the benchmark workloads (N, nnz, boundary DOFs)."""
    for n, p, ndof, nnz, nb in [(20, 1, 59660, 938060, 7200), (40, 2, 2480240, 106797280, 96000),
                                (30, 2, 1053180, 45133560, 54000)]:
        w = Workload(n, p)
        info, _ = (None, None)
        L = nat.lib()
        h = ctypes.c_void_p()
        T = w.topo
        ee = np.ascontiguousarray(T.elems_edges, np.int32)
        ef = np.ascontiguousarray(T.elems_faces, np.int32)
        nat.check(L.pg_plan_build(p, T.n_elems, T.n_edges, T.n_faces, nat.ptr(ee), nat.ptr(ef),
                                  ctypes.byref(h)))
        inf = np.zeros(6, np.int64)
        L.pg_plan_info(h, nat.ptr(inf))
        L.pg_plan_destroy(h)
        assert (T.ndof, int(inf[1]), T.boundary_dofs.size) == (ndof, nnz, nb)


def test_entity_order_is_morton_permutation():
    """DeviceMesh.entity_order: a permutation of the block_ptr entities whose
    positions follow the Morton curve (spatially local)."""
    from petgem_amd.core import DeviceMesh, _morton3
    for p in (1, 2, 3):
        w = Workload(4, p)
        m = DeviceMesh.from_topology(w.topo, "cpu")
        order = m.entity_order()
        nb = m.block_ptr().size - 1
        assert order.dtype == np.int32 and order.size == nb
        assert np.array_equal(np.sort(order), np.arange(nb))
    # the Morton key interleaves x, y, z bits (x lowest)
    assert list(_morton3(np.array([1, 0, 0, 3]), np.array([0, 1, 0, 0]),
                         np.array([0, 0, 1, 0]))) == [1, 2, 4, 9]


def test_eop_zero_mask_header_matches_tables():
    """The structural-zero masks compiled into the element-product kernel are the
    ones the current tables give (over every orientation combination)."""
    import re
    from petgem_amd import gen_eop_mask
    text = open(gen_eop_mask.OUT).read()
    for p in (1, 2, 3):
        body = re.search(r"kZero%d\[\d+\] = \{([^}]*)\}" % p, text).group(1)
        header = [int(x, 16) for x in re.findall(r"0x[0-9a-f]+", body)]
        assert header == [int(v) for v in gen_eop_mask.structural_zeros(p)], p


def test_element_operator_rejects_bad_arguments_before_device_work():
    """pg_eop_create validates the order and the entity ranges on the host before
    it touches any device pointer (the pointers here are never dereferenced)."""
    L = nat.lib()
    h = ctypes.c_void_p()
    fake = ctypes.c_void_p(256)  # stands for a device pointer; not read
    bp = np.array([0, 2, 4], np.int32)
    rc = L.pg_eop_create(4, 1, fake, fake, 4, nat.ptr(bp), 2, None, None, 1.0, ctypes.byref(h),
                         None)
    assert rc == nat.PG_ERR_UNSUPPORTED and b"p=4" in L.pg_last_error()
    bad_end = np.array([0, 2, 5], np.int32)
    rc = L.pg_eop_create(2, 1, fake, fake, 4, nat.ptr(bad_end), 2, None, None, 1.0,
                         ctypes.byref(h), None)
    assert rc == nat.PG_ERR_ARG and b"block_ptr" in L.pg_last_error()
    too_big = np.array([0, 3, 4], np.int32)
    rc = L.pg_eop_create(2, 1, fake, fake, 4, nat.ptr(too_big), 2, None, None, 1.0,
                         ctypes.byref(h), None)
    assert rc == nat.PG_ERR_ARG and b"entity sizes" in L.pg_last_error()
    rc = L.pg_eop_create(2, 0, fake, fake, 4, nat.ptr(bp), 2, None, None, 1.0, ctypes.byref(h),
                         None)
    assert rc == nat.PG_ERR_ARG


def test_multi_mesh_element_operator_rejects_bad_arguments_before_device_work():
    """pg_eop_create_multi checks the order and every mesh's entity ranges on the
    host before it touches a device pointer."""
    L = nat.lib()
    h = ctypes.c_void_p()
    fake = ctypes.c_void_p(256)  # stands for a device pointer; not read
    bp = np.array([0, 2, 4], np.int32)
    bad = np.array([0, 2, 5], np.int32)
    P2 = ctypes.c_void_p * 2
    ne = np.array([1, 1], np.int64)
    nd = np.array([4, 4], np.int64)
    nb = np.array([2, 2], np.int64)
    om = np.array([1.0, 2.0])
    args = lambda bps: (ne.ctypes.data, P2(fake, fake), P2(fake, fake), nd.ctypes.data,  # noqa
                        P2(*[b.ctypes.data for b in bps]), nb.ctypes.data, None, None,
                        om.ctypes.data, ctypes.byref(h), None)
    rc = L.pg_eop_create_multi(4, 2, *args([bp, bp]))
    assert rc == nat.PG_ERR_UNSUPPORTED and b"p=4" in L.pg_last_error()
    rc = L.pg_eop_create_multi(2, 2, *args([bp, bad]))
    assert rc == nat.PG_ERR_ARG and b"block_ptr" in L.pg_last_error()
    rc = L.pg_eop_create_multi(2, 0, *args([bp, bp]))
    assert rc == nat.PG_ERR_ARG
