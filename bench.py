#!/usr/bin/env python3
"""Benchmark of the CSEM forward hot path on one MI355X.

Workload (the largest single-GPU configuration the headline metric is quoted
on): p = 2 Nedelec elements on a 40^3 Kuhn grid (384,000 tetrahedra,
2,480,240 DOFs, 106,797,280 CSR non-zeros), layered marine model with seeded
log-normal noise (sigma_log = 0.1), x-directed HED at (0, 0, -950) m, 1 Hz.

One step = numeric assembly (element geometry + fused atomic-free CSR
assembly) + point-dipole source + PEC boundary condition + preconditioner
set-up + Krylov solve to a true relative residual <= 1e-8, with every input
already resident in HBM.  The Krylov iterations apply A either with a CSR
SpMV of the assembled matrix or element by element (the element matrices
recomputed from the tables at every product), whichever the solver's
autotuner measures faster; the true residual always uses the assembled CSR.
The headline value is the FP64 step time (complex128 throughout, the
workload as stated); the mixed-precision step (complex64 Krylov iterations
under FP64 residual replacement) is the metric's second leg, reported beside
it with its speed-up and its receiver-field difference from FP64.  Prints one
JSON line.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

METRIC = "p=2 CSEM solve time @1e-8 relres (FP64 vs mixed); batched RHS/s; adjoint grad/s"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 (vector and matrix)
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(nnz, n, bv):
    """Algorithmic bytes of one CSR SpMV: values + int32 columns + row pointers + x, y."""
    return nnz * (bv + 4) + 4 * (n + 1) + 2 * n * bv


def eop_flops(n_elems, p):
    """Algorithmic flops of one element-by-element product.  Per element and
    upper-triangle pair (i, j >= i): the mass part of the entry from its m
    nonzero table components (one MUL + m-1 FMA = 2m-1 flops), the curl-curl
    part likewise, then each nonzero part applied to x_j (2 FMA = 4 flops) and,
    off the diagonal, to x_i (4 more).  The components that vanish for every
    orientation (csrc/pg_eop_mask.h) are not computed and not counted."""
    import re
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "petgem_amd", "csrc",
                        "pg_eop_mask.h")
    body = re.search(r"kZero%d\[\d+\] = \{([^}]*)\}" % p, open(path).read()).group(1)
    masks = [int(v, 16) for v in re.findall(r"0x[0-9a-f]+", body)]
    n = p * (p + 2) * (p + 3) // 2
    total, k = 0, 0
    for i in range(n):
        for j in range(i, n):
            z = masks[k]
            k += 1
            for part in (z & 0x3F, (z >> 6) & 0x3F):
                m = 6 - bin(part).count("1")
                if m:
                    total += 2 * m - 1 + (4 if j == i else 8)
    return n_elems * total


def spmv_key(code, dmax, vt, p=2):
    """Kernel name (as the profile summaries print it) of the solve's SpMV from the
    solver's configuration code kind*1000 + G*10 + U."""
    kind, g, u = code // 1000, (code % 1000) // 10, code % 10
    if kind == 3:  # element products (p = 2: paired gathers on the solver's aligned vectors)
        if p == 3 and vt != "float2":
            return "k_eop_mfma3"  # p = 3 complex128: the matrix-core kernel
        return "k_eop_local<{},{},{}>".format(p, "float" if vt == "float2" else "double",
                                              "true" if p == 2 else "false")
    if kind == 2:
        return f"k_spmv_er2<{g},{u},1,{vt},{vt}>"
    if kind == 1:
        d = 1 if dmax <= 1 else (2 if dmax <= 2 else 6)
        return f"k_spmv_er<{g},{u},{d},1,{vt},{vt}>"
    return f"k_spmv<{g},{u},1,{vt},{vt}>"


def roofline_entry(code, res, standalone_ms, vt, T, plan, dmax, traffic):
    """Roofline of the first kernel of the solve's SpMV.

    CSR kinds: HBM-bound, algorithmic bytes = spmv_bytes.  Element-by-element
    operator (kind 3): its element-product kernel is bound by the vector ALU
    (FP32 or FP64 FMA issue), so it is priced in algorithmic flops (eop_flops)
    against the FP32 / FP64 peak; the summation kernel that follows it is
    HBM-bound and reported beside it with its algorithmic bytes."""
    bv = 8 if vt == "float2" else 16
    n, nnz = plan.ndof, plan.nnz
    key = spmv_key(code, dmax, vt, T.p)
    out = {"samples": res["spmv_samples"], "spmv_ms": round(res["spmv_ms"], 5),
           "spmv_ms_standalone": round(standalone_ms, 5)}
    if code // 1000 == 3:
        ms = res["spmv_first_ms"]
        fl = eop_flops(T.n_elems, T.p)
        peak = FP32_PEAK_TFLOPS if vt == "float2" else FP64_PEAK_TFLOPS
        ach = fl / (ms * 1e-3) / 1e12
        # summation kernel: 16-byte entity descriptors, one 4-byte slot per
        # (entity, element) incidence, the n local results of every element,
        # p read and q written
        nle = 6 if T.p == 1 else 10
        nent = T.n_edges + (T.n_faces if T.p >= 2 else 0)
        sum_bytes = 16 * nent + 4 * T.n_elems * nle + T.n_elems * T.n_dof_elem * bv + 2 * n * bv
        sum_ms = max(res["spmv_ms"] - ms, 1e-9)
        out.update({
            "bound": "valu", "kernel": f"{key} (element products y_e = A_e p_e)",
            "achieved": round(ach, 2), "peak": peak, "unit": "TFLOP/s",
            "frac": round(ach / peak, 4), "traffic": traffic.get(key),
            "algorithmic_flops": fl, "avg_ms": round(ms, 5),
            "note": ("vector-ALU FMA kernel (no matrix-core work); packed FP32 is disabled "
                     "in this build, so FP32 issues at most half the published peak"
                     if vt == "float2" else
                     "vector-ALU FP64 FMA kernel (no matrix-core work); peak = MI355X FP64 "
                     "vector rate"),
            "sum_kernel": {"bound": "hbm", "avg_ms": round(sum_ms, 5),
                           "algorithmic_bytes": sum_bytes,
                           "achieved": round(sum_bytes / (sum_ms * 1e-3) / 1e9, 1),
                           "unit": "GB/s", "peak": HBM_PEAK_GBS,
                           "frac": round(sum_bytes / (sum_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
            "csr_equivalent_bytes": spmv_bytes(nnz, n, bv),
        })
    else:
        ms = res["spmv_ms"]
        by = spmv_bytes(nnz, n, bv)
        ach = by / (ms * 1e-3) / 1e9
        out.update({"bound": "hbm", "kernel": f"{key} (q = A p + p^T q partials)",
                    "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic.get(key),
                    "algorithmic_bytes": by, "avg_ms": round(ms, 5)})
    return out


def measured_traffic(config):
    """HBM bytes per launch of each SpMV configuration, from the committed
    FETCH_SIZE / WRITE_SIZE profile of this workload ({} if there is none)."""
    from petgem_amd.profsum import parse_traffic_table
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"spmv_traffic_{config}.md")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return parse_traffic_table(f.read())


def cpu_baseline(w, plan_rowptr, plan_cols):
    """The CPU restatement in C + OpenMP (oracle/cpu_baseline.c) on every host
    core OpenMP is given: element matrices in the reference's quadrature form,
    CSR assembly, PEC condition and COCG + Jacobi to a true relative residual
    of 1e-8 on the full workload, timed end to end (no sampling, no
    extrapolation).  Beside it, the same on C1 (p = 1, 20^3) and SciPy's
    SuperLU on C1."""
    import scipy.sparse.linalg as spla
    from oracle import cpu, fem, topology as otopo
    from petgem_amd.synthetic import CONFIGS, Workload

    def run(wl, rowptr, cols):
        T = wl.topo
        eo, fo = otopo.orientations(T.elems_nodes, T.elems_edges, T.edges_nodes,
                                    T.elem_faces_edges)
        S = cpu.System(T.coords, wl.sigma, eo, fo, T.dofs, T.ndof, T.p, rowptr, cols)
        b = cpu.source(T.coords, T.dofs, eo, fo, T.p, T.ndof, wl.source)
        return S, b, cpu.forward(S, wl.frequency, b, T.boundary_dofs, tol=1e-8)

    S, b, r = run(w, plan_rowptr, plan_cols)
    if r["relres"] > 1e-8:
        raise RuntimeError(f"CPU baseline did not converge: {r['relres']}")
    # C1: the same restatement, and SuperLU on the same system
    w1 = Workload(**CONFIGS["c1"], seed=2026)
    T1 = w1.topo
    ip, ci = fem.csr_pattern(T1.dofs, T1.ndof)
    S1, b1, r1 = run(w1, ip, ci)
    bb = S1.dirichlet(T1.boundary_dofs, b1)
    t0 = time.perf_counter()
    x1 = spla.spsolve(S1.matrix().tocsc(), bb[:, 0] + 1j * bb[:, 1])
    t_lu = time.perf_counter() - t0
    lu_diff = float(np.abs(x1 - r1["x"]).max() / np.abs(x1).max())
    return dict(value=r["total_s"], unit="s", cores=r["threads"], kind="port",
                sample=(f"full workload on {r['threads']} OpenMP threads (oracle/cpu_baseline.c): "
                        f"assembly {r['assembly_s']:.2f} s, PEC {r['bc_s']:.2f} s, COCG + Jacobi "
                        f"{r['iterations']} iterations to true relres {r['relres']:.1e} in "
                        f"{r['solve_s']:.1f} s; no extrapolation"),
                assembly_s=r["assembly_s"], solve_s=r["solve_s"], iterations=r["iterations"],
                c1={"total_s": round(r1["total_s"], 3), "iterations": r1["iterations"],
                    "relres": r1["relres"], "superlu_s": round(t_lu, 3),
                    "superlu_vs_cocg_rel_diff": lu_diff})


def batched_leg(args, log):
    """C4 batching: 64 HED positions (100 m apart along x at z = -950 m) x 4
    frequencies = 256 right-hand sides on the 30^3 p=2 mesh.

    (a) The batched multi-RHS solver, one --batch-width-column batch per chunk
        of one frequency's positions: COCG per column (--batch-precision),
        entity-block Jacobi, the columns' products applied element by element
        with each element's entries formed once for all columns (the element
        operator in the batch).  Every right-hand side is solved and timed,
        K and M assembled once per batch.
    (b) The single-RHS solver (mixed precision, entity-block Jacobi, element
        products) two right-hand sides at a time on two streams
        (core.solve_concurrently), on --single-positions of the positions x 4
        frequencies (default: all 256 right-hand sides).
    `rhs_per_s` is the faster of the two over the same right-hand sides
    (`mode` names it); both rates are reported."""
    import torch
    from petgem_amd import core
    from petgem_amd.synthetic import CONFIGS, Workload

    w = Workload(**CONFIGS["c4"], seed=2026)
    T = w.topo
    dev = torch.device("cuda")
    mesh = core.DeviceMesh.from_topology(T, dev)
    asm = core.Assembler(mesh)
    ent_bc, dof_bc = mesh.entity_mask(T.boundary_dofs)
    sigma = torch.from_numpy(w.sigma).to(dev)
    K = torch.empty(asm.plan.nnz, dtype=torch.float64, device=dev)
    M = torch.empty_like(K)
    xs_all = -3150.0 + 100.0 * np.arange(64)
    xs = xs_all[np.linspace(0, 63, args.batch_positions).round().astype(int)]
    srcs = np.stack([xs, np.zeros_like(xs), np.full_like(xs, -950.0)], axis=1)
    src_el = core.locate_points(mesh, srcs).cpu().numpy()
    freqs = [0.1, 0.25, 0.5, 1.0]
    width = args.batch_width
    chunks = []
    for f in freqs:  # one frequency per batch: each batch stops when its columns converge
        cf = [(si, f) for si in range(len(xs))]
        chunks += [cf[i:i + width] for i in range(0, len(cf), width)]
    ncols = sum(len(ch) for ch in chunks)
    asm.geometry(sigma)
    core.assemble_km(asm, ent_bc, out=(K, M))
    bptr, order = mesh.block_ptr(), mesh.entity_order()
    eop = core.ElementOperator(asm, core.omega_of(1.0), ent_bc, order)
    use_eop = width in (64, 128, 192, 256)
    solvers = [core.BatchSolver(asm.plan, K, M, [f for _, f in ch], rtol=args.rtol,
                                block_ptr=bptr, entity_order=order, pc="bjacobi",
                                precision=args.batch_precision,
                                eop=eop if use_eop and len(ch) == width else None)
               for ch in chunks]
    B = torch.zeros((T.ndof, width, 2), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    asm.geometry(sigma)
    core.assemble_km(asm, ent_bc, out=(K, M))
    eop.refresh(core.omega_of(1.0))
    its, worst = [], 0.0
    for ch, bs in zip(chunks, solvers):
        bs.refresh()
        for c, (si, f) in enumerate(ch):
            b = core.source_vector(mesh, int(src_el[si]), srcs[si], (1.0, 0.0, 0.0),
                                   core.omega_of(f))
            core.apply_dirichlet(asm.plan, None, dof_bc, b)
            B[:, c, :] = b
        X, info = bs.solve(B[:, :len(ch)].contiguous() if len(ch) < width else B)
        if (info["status"] != 0).any():
            raise RuntimeError(f"batched solve did not converge: {info}")
        its.append(int(info["iterations"].max()))
        worst = max(worst, float(info["relres"].max()))
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    log(f"batched: {t:.2f} s for {ncols} RHS ({ncols / t:.2f} RHS/s), iterations per batch "
        f"{its}, worst relres {worst:.2e}")
    # the single-RHS solver, two at a time on two streams, on a subset of positions
    sel = np.linspace(0, len(xs) - 1, min(args.single_positions, len(xs))).round().astype(int)
    seq_t, nseq = 0.0, 0
    vals = torch.empty((asm.plan.nnz, 2), dtype=torch.float64, device=dev)
    for f in freqs:
        asm.assemble(core.omega_of(f), ent_bc=ent_bc, out=vals)
        eops = [core.ElementOperator(asm, core.omega_of(f), ent_bc, order) for _ in range(2)]
        Ss = []
        for e in eops:  # the second takes the first one's kernel choices
            Ss.append(core.KrylovSolver(asm.plan, vals, precision="mixed", pc="bjacobi",
                                        rtol=args.rtol, block_ptr=bptr, entity_order=order,
                                        eop=e, like=Ss[0] if Ss else None))
        rhs = []
        for si in sel:
            b = core.source_vector(mesh, int(src_el[si]), srcs[si], (1.0, 0.0, 0.0),
                                   core.omega_of(f))
            core.apply_dirichlet(asm.plan, None, dof_bc, b)
            rhs.append(b)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, infos = core.solve_concurrently(Ss, rhs)
        torch.cuda.synchronize()
        seq_t += time.perf_counter() - t0
        nseq += len(rhs)
        for inf in infos:
            if inf["status"] != 0:
                raise RuntimeError(f"single solve did not converge: {inf}")
        del Ss, eops
    log(f"single-RHS solves two at a time (mixed, block Jacobi, element products): {nseq} RHS in "
        f"{seq_t:.2f} s ({nseq / seq_t:.2f} RHS/s)")
    r_batch, r_single = ncols / t, nseq / seq_t
    best = "batched_solver" if r_batch >= r_single else "single_rhs_two_streams"
    return {"rhs_per_s": round(max(r_batch, r_single), 3), "mode": best,
            "batched_solver_rhs_per_s": round(r_batch, 3),
            "s_total": round(t, 3), "nrhs": ncols, "width": width,
            "precision": args.batch_precision, "element_products": use_eop,
            "batch_iterations": its, "worst_relres": worst,
            "single_rhs_two_streams_rhs_per_s": round(nseq / seq_t, 3),
            "single_rhs_sample": int(nseq),
            "batched_vs_single_rhs_two_streams": round((ncols / t) / (nseq / seq_t), 3),
            "workload": (f"c4: p=2, 30^3 Kuhn x 6 = {T.n_elems} tets, {T.ndof} DOFs, "
                         f"{asm.plan.nnz} nnz; {len(xs)} HED x-positions (100 m apart at "
                         f"z=-950 m) x {freqs} Hz = {ncols} RHS, rtol {args.rtol}; batches of "
                         f"{width} columns of one frequency, COCG + entity-block Jacobi per "
                         f"column ({args.batch_precision}), element products; K/M assembled "
                         "once; single_rhs_two_streams = the single-RHS solver (mixed, "
                         "entity-block Jacobi, element products) two at a time on two "
                         f"streams over {nseq} of the right-hand sides")}


def c3_leg(args, log):
    """Mixed precision at p = 3: 24^3 Kuhn hexes x 6 = 82,944 tets (p = 3), the
    layered model with the resistive block, HED at (0, 0, -950) m, 1 Hz and
    0.25 Hz.  At each frequency the FP64 solve (COCG + entity-block Jacobi,
    products on the FP64 matrix cores or the CSR, as the autotuner picks) and
    the mixed solve (complex64 iterations under FP64 refinement) of the same
    system, one warm-up and one timed solve each; receiver fields of the two
    compared."""
    import torch
    from petgem_amd import core
    from petgem_amd.synthetic import CONFIGS, Workload

    w = Workload(**CONFIGS["c3"], seed=2026)
    T = w.topo
    dev = torch.device("cuda")
    mesh = core.DeviceMesh.from_topology(T, dev)
    asm = core.Assembler(mesh)
    plan = asm.plan
    ent_bc, dof_bc = mesh.entity_mask(T.boundary_dofs)
    asm.geometry(torch.from_numpy(w.sigma).to(dev))
    bptr, order = mesh.block_ptr(), mesh.entity_order()
    src_el = int(core.locate_points(mesh, np.array(w.source["position"])[None]).cpu()[0])
    rx_el = core.locate_points(mesh, w.receivers)
    moment = core.dipole_moment(w.source)
    out = {"ndof": int(plan.ndof), "n_elems": int(T.n_elems)}
    for f in (1.0, 0.25):
        om = core.omega_of(f)
        vals = asm.assemble(om, ent_bc=ent_bc)
        b = torch.zeros((plan.ndof, 2), dtype=torch.float64, device=dev)
        core.source_vector(mesh, src_el, w.source["position"], moment, om, b=b)
        core.apply_dirichlet(plan, None, dof_bc, b)
        eop = core.ElementOperator(asm, om, ent_bc, order)
        res = {}
        for prec in ("fp64", "mixed"):
            S = core.KrylovSolver(plan, vals, precision=prec, rtol=args.rtol, pc=args.pc,
                                  block_ptr=bptr, entity_order=order, eop=eop)
            S.solve(b)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            x, info = S.solve(b)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            if info["status"] != 0 or info["relres"] > args.rtol:
                raise RuntimeError(f"C3 {prec} solve at {f} Hz did not converge: {info}")
            fields = core.as_complex(core.receiver_fields(mesh, w.receivers, rx_el, x, om))
            res[prec] = (dt, info, fields)
            log(f"c3 {f} Hz {prec}: {dt*1e3:.1f} ms, {info['iterations']} iterations, "
                f"relres {info['relres']:.2e}")
            del S
        f64, fmx = res["fp64"][2], res["mixed"][2]
        key = f"{f:g}Hz"
        out[key] = {"fp64_s": round(res["fp64"][0], 4), "mixed_s": round(res["mixed"][0], 4),
                    "mixed_speedup": round(res["fp64"][0] / res["mixed"][0], 3),
                    "fp64_iterations": int(res["fp64"][1]["iterations"]),
                    "mixed_iterations": int(res["mixed"][1]["iterations"]),
                    "mixed_refinements": int(res["mixed"][1]["refinements"]),
                    "fields_mixed_vs_fp64": float(np.abs(fmx - f64).max() / np.abs(f64).max())}
        del eop, vals
        torch.cuda.empty_cache()
    out["workload"] = (f"c3: p=3, 24^3 Kuhn x 6 = {T.n_elems} tets, {plan.ndof} DOFs, layered "
                       "model with the resistive block, HED at (0,0,-950) m, 1 and 0.25 Hz, rtol "
                       f"{args.rtol}; FP64 vs mixed solve of each, {args.pc}")
    return out


def multimesh_leg(args, log):
    """32 independent meshes, 8^3 .. 24^3 Kuhn hexes (p=2), seeded random layered
    conductivities (0.01 - 3.3 S/m), one HED each (synthetic.multimesh_systems),
    solved as multi-mesh batches (HIP-graph replay; one COCG per mesh) on the
    meshes' element products with entity-block Jacobi: in four concurrent lanes
    in FP64 and in complex64 rounds, and as one FP64 batch; beside them the
    batch on CSR products and point Jacobi, the latter's systems one after
    another, and single-RHS solvers (mixed, entity-block Jacobi, element
    products) four meshes at a time on four streams.  `rhs_per_s` is the
    faster of the fastest batch variant and the four-stream single solves."""
    import torch
    from petgem_amd import core
    from petgem_amd.synthetic import multimesh_systems

    systems, bs, fast, elem = [], [], [], []
    for m in multimesh_systems("cuda"):
        systems.append((m["plan"], m["vals"]))
        bs.append(m["b"])
        elem.append((m["asm"], m["omega"], m["ent_bc"], m["order"]))
        eop = core.ElementOperator(m["asm"], m["omega"], m["ent_bc"], m["order"])
        fast.append(core.KrylovSolver(m["plan"], m["vals"], precision="mixed", pc="bjacobi",
                                      rtol=args.rtol, block_ptr=m["mesh"].block_ptr(),
                                      entity_order=m["order"], eop=eop))
    ndof = int(sum(pl.ndof for pl, _ in systems))

    def timed(ms, name):
        times = []
        for it in range(1 + min(2, max(1, args.steps))):  # a warm-up and at most two timed runs
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            xs, info = ms.solve(bs)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            if (info["status"] != 0).any():
                raise RuntimeError(f"multi-mesh solve did not converge: {info}")
            log(f"multi-mesh ({name}) step {it}: {dt*1e3:.1f} ms, iterations "
                f"{info['iterations'].min()}..{info['iterations'].max()}, graph launches "
                f"{info['launches']}")
            if it >= 1:
                times.append(dt)
        return float(np.median(times)), info

    lanes = 4
    ms_l = core.MultiMeshSolver(systems, rtol=args.rtol, check_every=25, use_graph=True,
                                element=elem, lanes=lanes)
    t_l, info_l = timed(ms_l, f"element products, block Jacobi, FP64, {lanes} lanes")
    del ms_l
    ms_m = core.MultiMeshSolver(systems, rtol=args.rtol, check_every=25, use_graph=True,
                                element=elem, precision="mixed", lanes=lanes)
    t_m, info_m = timed(ms_m, f"element products, block Jacobi, mixed, {lanes} lanes")
    del ms_m
    ms_e = core.MultiMeshSolver(systems, rtol=args.rtol, check_every=25, use_graph=True,
                                element=elem)
    t_e, info_e = timed(ms_e, "element products, block Jacobi, FP64, one batch")
    del ms_e
    ms = core.MultiMeshSolver(systems, rtol=args.rtol, check_every=25, use_graph=True)
    t, info = timed(ms, "CSR, point Jacobi")
    del ms
    singles = [core.KrylovSolver(pl, v, rtol=args.rtol) for pl, v in systems]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for S, b in zip(singles, bs):
        x, inf = S.solve(b)
        if inf["status"] != 0:
            raise RuntimeError(f"single solve did not converge: {inf}")
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, infs = core.solve_each(fast, bs, workers=4)
    torch.cuda.synchronize()
    t_fast = time.perf_counter() - t0
    for inf in infs:
        if inf["status"] != 0:
            raise RuntimeError(f"single solve did not converge: {inf}")
    log(f"multi-mesh: batch on element products FP64 in {lanes} lanes {t_l*1e3:.1f} ms, mixed "
        f"in {lanes} lanes {t_m*1e3:.1f} ms, FP64 one batch {t_e*1e3:.1f} ms, on CSR "
        f"{t*1e3:.1f} ms vs one after another (CSR) {t_seq*1e3:.1f} ms; single-RHS on element "
        f"products, four at a time: {t_fast*1e3:.1f} ms")
    # the fastest batch variant
    t_b, info_b, var_b = min((t_l, info_l, f"fp64, {lanes} lanes"),
                             (t_m, info_m, f"mixed, {lanes} lanes"),
                             (t_e, info_e, "fp64, one batch"), key=lambda v: v[0])
    best = "multi_mesh_batch" if t_b <= t_fast else "single_rhs_four_streams"
    return {"rhs_per_s": round(len(systems) / min(t_b, t_fast), 3), "mode": best,
            "batch_rhs_per_s": round(len(systems) / t_b, 3), "s_per_batch": round(t_b, 4),
            "batch_variant": var_b,
            "batch_iterations_max": int(info_b["iterations"].max()),
            "batch_graph_launches": info_b["launches"],
            "speedup_vs_single_rhs_four_streams": round(t_fast / t_b, 3),
            "fp64_lanes_s": round(t_l, 4), "mixed_lanes_s": round(t_m, 4),
            "mixed_iterations_max": int(info_m["iterations"].max()),
            "fp64_one_batch_s": round(t_e, 4),
            "fp64_iterations_max": int(info_e["iterations"].max()),
            "single_rhs_four_streams_s": round(t_fast, 4),
            "csr_batch_s": round(t, 4), "csr_batch_iterations_max": int(info["iterations"].max()),
            "meshes": len(systems), "total_dofs": ndof,
            "sequential_csr_s": round(t_seq, 4),
            "workload": ("32 meshes, 8^3..24^3 Kuhn hexes x 6 (p=2), seeded random 3-layer "
                         "conductivities 0.01-3.3 S/m, HED at (0,0,-950) m, 0.1/0.25/0.5/1 Hz "
                         f"cyclic, rtol {args.rtol}; batch = one COCG per mesh with its "
                         "batch's element products in one launch and entity-block Jacobi, in "
                         "FP64 or in complex64 rounds under FP64 refreshes per mesh, as one "
                         "batch or dealt to 4 concurrent batches (lanes) on their own streams, "
                         "the fastest variant reported as the batch; HIP-graph replay of 25 "
                         "iterations between host checks, the launches cut to the meshes still "
                         "iterating; csr_batch = "
                         "the same batch on CSR products and point Jacobi; sequential_csr = its "
                         "systems one after another through the single-RHS solver; "
                         "single_rhs_four_streams = single-RHS solvers (mixed, entity-block "
                         "Jacobi, element products), four meshes at a time")}


def adjoint_leg(args, log):
    """Adjoint gradient: 8 HED sources x 2 frequencies, 101 receivers, 30^3 p=2 mesh;
    observations from the model with the resistive block, gradient at the model
    without it, w.r.t. every element's log-conductivity.  Timed two ways: forward
    and adjoint as two batched 16-column solves (--adjoint-precision), and as 32
    single-RHS solves (mixed precision, entity-block Jacobi, two at a time on two
    streams); `grad_per_s` is the faster of the two."""
    import torch
    from petgem_amd.adjoint import InversionProblem
    from petgem_amd.synthetic import Workload, receiver_line

    w_true = Workload(30, 2, block=True, n_rx=101, seed=2026)
    w0 = Workload(30, 2, block=False, n_rx=101, seed=2026)
    T = w_true.topo
    xs = np.linspace(-1750.0, 1750.0, 8)
    src = np.stack([xs, np.zeros_like(xs), np.full_like(xs, -950.0)], axis=1)
    rx = receiver_line(101, z=-1000.0)
    m0 = torch.tensor(np.log(w0.sigma[:, 0]), dtype=torch.float64, device="cuda")
    res, grads = {}, {}
    observed = floor = None
    for mode, kw in (("batched", dict(batched=True, precision=args.adjoint_precision)),
                     ("per_rhs", dict(batched=False, precision="mixed"))):
        prob = InversionProblem(T, src, [0.25, 1.0], rx, rtol=args.rtol,
                                device=torch.device("cuda"), **kw)
        if observed is None:
            # one set of observations for both modes; data errors relative with
            # an absolute floor 1e-6 of the largest field (without a floor the
            # 1/|E_obs|^2 weights of the weakest fields magnify the solver
            # tolerance into the misfit)
            observed = prob.forward(np.log(w_true.sigma[:, 0]))
            floor = 1e-6 * float(np.abs(observed).max())
        prob.observed = observed
        prob.noise_floor = floor
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        phi, g = prob.misfit_and_gradient(m0)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        log(f"adjoint ({mode}): {dt*1e3:.1f} ms, misfit {phi:.6e}, |g|max {float(g.abs().max()):.4e}")
        res[mode] = dt
        grads[mode] = g
        del prob
    gdiff = float((grads["batched"] - grads["per_rhs"]).abs().max() / grads["per_rhs"].abs().max())
    best = min(res, key=res.get)
    return {"grad_per_s": round(1.0 / res[best], 4), "s_per_gradient": round(res[best], 4),
            "mode": best, "batched_s": round(res["batched"], 4),
            "per_rhs_s": round(res["per_rhs"], 4), "grad_rel_diff_between_modes": gdiff,
            "n_params": int(T.n_elems), "solves_per_gradient": 32,
            "workload": (f"c5: p=2, 30^3 Kuhn x 6 = {T.n_elems} tets, {T.ndof} DOFs; 8 HED x "
                         "2 frequencies (0.25, 1 Hz), 101 receivers, rtol "
                         f"{args.rtol}; misfit weights 1/(|E_obs|^2 + (1e-6 max|E_obs|)^2); "
                         "forward + adjoint solves + element contraction; "
                         f"batched = two 16-column batched solves ({args.adjoint_precision}, "
                         "entity-block Jacobi), per_rhs = 32 single solves (mixed, entity-block "
                         "Jacobi, element products, two at a time on two streams)")}


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline (a full solve on the host cores, about a minute)")
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--frequency", type=float, default=None,
                    help="source frequency of the solve leg (default: the workload's, 1 Hz)")
    ap.add_argument("--pc", default="bjacobi", choices=["jacobi", "bjacobi"])
    ap.add_argument("--legs", default="solve,dropin,c3,batch,multimesh,adjoint",
                    help="comma list of: solve (headline), dropin (the reference's petsc.opts "
                         "solver), c3 (p = 3 FP64 vs mixed), batch (RHS/s), multimesh "
                         "(RHS/s), adjoint (grad/s)")
    ap.add_argument("--batch-width", type=int, default=64,
                    help="columns per batched solve (chunks of one frequency's positions)")
    ap.add_argument("--batch-precision", default="fp64", choices=["fp64", "mixed"],
                    help="batched leg: FP64, or complex64 under FP64 refinement (at C4 the "
                         "complex64 rounds stagnate at 0.1 Hz and the FP64 batch is faster)")
    ap.add_argument("--adjoint-precision", default="mixed", choices=["fp64", "mixed"],
                    help="the adjoint leg's batched solves")
    ap.add_argument("--batch-positions", type=int, default=64,
                    help="HED positions of the batched leg (of 64; x 4 frequencies)")
    ap.add_argument("--single-positions", type=int, default=64,
                    help="positions (x 4 frequencies) of the single-RHS solves")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: the multi-GPU sharding of the "
                          "path is designed but not implemented; run --gpus 1"}))
        sys.exit(2)

    import torch
    from petgem_amd import core
    from petgem_amd.synthetic import CONFIGS, Workload

    cfg = dict(CONFIGS[args.config])
    if args.frequency:
        cfg["frequency"] = args.frequency
    t0 = time.perf_counter()
    w = Workload(**cfg, seed=2026)
    T = w.topo
    log(f"workload {args.config}: {w.describe()} built in {time.perf_counter()-t0:.1f}s")
    dev = torch.device("cuda")
    t0 = time.perf_counter()
    mesh = core.DeviceMesh.from_topology(T, dev)
    asm = core.Assembler(mesh)
    plan = asm.plan
    sigma = torch.from_numpy(w.sigma).to(dev)
    ent_bc, dof_bc = mesh.entity_mask(T.boundary_dofs)
    omega = core.omega_of(w.frequency)
    moment = core.dipole_moment(w.source)
    src_el = int(core.locate_points(mesh, np.array(w.source["position"])[None]).cpu()[0])
    values = torch.empty((plan.nnz, 2), dtype=torch.float64, device=dev)
    b = torch.zeros((plan.ndof, 2), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    log(f"setup (upload + assembly plan) {time.perf_counter()-t0:.1f}s; nnz={plan.nnz}")

    def assemble():
        # PEC rows/columns folded into the assembly kernel; b masked separately
        asm.geometry(sigma)
        asm.assemble(omega, ent_bc=ent_bc, out=values)
        b.zero_()
        core.source_vector(mesh, src_el, w.source["position"], moment, omega, b=b)
        core.apply_dirichlet(plan, None, dof_bc, b)

    assemble()
    bptr = mesh.block_ptr()
    order = mesh.entity_order()
    # element-by-element operator: a candidate of the solvers' autotuners (at
    # p = 3 for the complex64 iterations only)
    eop = core.ElementOperator(asm, omega, ent_bc, order)
    solvers = {prec: core.KrylovSolver(plan, values, precision=prec, rtol=args.rtol, pc=args.pc,
                                       block_ptr=bptr, entity_order=order, eop=eop)
               for prec in ("fp64", "mixed")}
    results = {}
    for prec in ("fp64", "mixed"):
        S = solvers[prec]
        times, infos = [], []
        for it in range(args.warmup + args.steps):
            x = torch.zeros_like(b)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            assemble()
            if eop is not None:
                eop.refresh(omega)
            S.refresh()
            x, info = S.solve(b, x)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            if info["status"] != 0 or info["relres"] > args.rtol:
                raise RuntimeError(f"{prec} solve did not converge: {info}")
            if it >= args.warmup:
                times.append(dt)
                infos.append(info)
            log(f"{prec} step {it}: {dt*1e3:.1f} ms, {info}")
        results[prec] = dict(t=float(np.median(times)), times=times, info=infos[-1], x=x,
                             spmv_ms=float(np.mean([i["spmv_ms"] for i in infos])),
                             spmv_first_ms=float(np.mean([i["spmv_first_ms"] for i in infos])),
                             spmv_samples=int(np.sum([i["spmv_samples"] for i in infos])))

    # assembly alone (HIP events on the current stream)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(5):
        assemble()
    ev1.record()
    torch.cuda.synchronize()
    t_asm_ms = ev0.elapsed_time(ev1) / 5

    # receivers (mixed solution) and agreement with FP64
    rx_el = core.locate_points(mesh, w.receivers)
    f64 = core.as_complex(core.receiver_fields(mesh, w.receivers, rx_el, results["fp64"]["x"], omega))
    f32 = core.as_complex(core.receiver_fields(mesh, w.receivers, rx_el, results["mixed"]["x"], omega))
    field_diff = float(np.abs(f64 - f32).max() / np.abs(f64).max())

    # roofline of the dominant kernel: the first kernel of the mixed solve's
    # SpMV.  Its average duration is sampled with HIP events on the solver's
    # stream inside the timed solves; the standalone profile() timing of the
    # whole SpMV is kept as a cross-check.
    prof = solvers["mixed"].profile(reps=20)
    prof64 = solvers["fp64"].profile(reps=20)
    n, nnz = plan.ndof, plan.nnz
    dmax = int(np.diff(bptr).max())
    traffic = measured_traffic(args.config)
    roof32 = roofline_entry(prof["g32"], results["mixed"], prof["ms32"], "float2", T, plan,
                            dmax, traffic)
    roof64 = roofline_entry(prof64["g64"], results["fp64"], prof64["ms64"], "double2", T, plan,
                            dmax, traffic)
    k32_name = spmv_key(prof["g32"], dmax, "float2", T.p)
    k64_name = spmv_key(prof64["g64"], dmax, "double2", T.p)
    log(f"SpMV mixed: {roof32}")
    log(f"SpMV fp64: {roof64}")
    log(f"assembly {t_asm_ms:.2f} ms")

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (full solve on the host cores) ...")
        cpu = cpu_baseline(w, plan.h_rowptr, plan.h_cols)
        log(f"cpu baseline {cpu['value']:.1f} s: {cpu['sample']}; C1 {cpu['c1']}")

    legs = set(args.legs.split(","))
    dropin = None
    if "dropin" in legs:
        # the reference's own example options (petsc.opts: -ksp_type gmres -pc_type
        # sor -ksp_rtol 1e-8) as the drop-in Solver maps them: GMRES(30), FP64,
        # entity-block Jacobi for SOR, products autotuned (element operator)
        Sg = core.KrylovSolver(plan, values, ksp="gmres", pc="bjacobi", precision="fp64",
                               rtol=args.rtol, restart=30, maxit=30000, block_ptr=bptr,
                               entity_order=order, eop=eop)
        xg = torch.zeros_like(b)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        assemble()
        eop.refresh(omega)
        Sg.refresh()
        xg, ig = Sg.solve(b, xg)
        torch.cuda.synchronize()
        tg = time.perf_counter() - t0
        fg = core.as_complex(core.receiver_fields(mesh, w.receivers, rx_el, xg, omega))
        dropin = {"s": round(tg, 4), "iterations": ig["iterations"], "relres": ig["relres"],
                  "status": ig["status"], "spmv": spmv_key(Sg.profile(reps=3)["g64"], dmax,
                                                           "double2", T.p),
                  "receiver_field_diff_vs_cocg_fp64": float(np.abs(fg - f64).max()
                                                            / np.abs(f64).max()),
                  "options": "-ksp_type gmres -pc_type sor -ksp_rtol 1e-8 (petsc.opts:2-4) -> "
                             "GMRES(30) FP64, entity-block Jacobi, true-residual stop"}
        log(f"drop-in petsc.opts (GMRES(30) + bjacobi): {dropin}")
        del Sg
    del solvers, values
    torch.cuda.empty_cache()
    c3 = c3_leg(args, log) if "c3" in legs else None
    batched = batched_leg(args, log) if "batch" in legs else None
    multimesh = multimesh_leg(args, log) if "multimesh" in legs else None
    adjoint = adjoint_leg(args, log) if "adjoint" in legs else None

    t_mixed, t_fp64 = results["mixed"]["t"], results["fp64"]["t"]
    line = {
        "metric": METRIC,
        "value": round(t_fp64, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_fp64 * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128",
        "data": "synthetic",
        "config": {
            "workload": (f"{args.config}: p={T.p} Nedelec, {w.n}^3 Kuhn hexes x 6 = {T.n_elems} "
                         f"tets, {T.ndof} DOFs, {nnz} nnz, layered marine model"
                         f"{' + log-normal noise 0.1' if cfg.get('noise') else ''}, "
                         f"HED at (0,0,-950) m, {w.frequency} Hz, rtol {args.rtol}"),
            "solver": f"COCG + {args.pc}, true-residual stop",
            "spmv": {"mixed": k32_name, "fp64": k64_name,
                     "choice": "autotuned per precision among CSR kernels and the "
                               "element-by-element operator"},
        },
        "fp64_s": round(t_fp64, 6),
        "mixed": {"s": round(t_mixed, 6), "speedup_vs_fp64": round(t_fp64 / t_mixed, 3),
                  "dtype": "complex64 Krylov iterations and SpMV; complex128 assembly, "
                           "residual b - Ax, solution update",
                  "roofline": roof32},
        "mixed_s": round(t_mixed, 6),
        "mixed_speedup": round(t_fp64 / t_mixed, 3),
        "iterations": {"fp64": results["fp64"]["info"]["iterations"],
                       "mixed": results["mixed"]["info"]["iterations"],
                       "mixed_refinements": results["mixed"]["info"]["refinements"]},
        "relres": {"fp64": results["fp64"]["info"]["relres"],
                   "mixed": results["mixed"]["info"]["relres"]},
        "receiver_field_diff_mixed_vs_fp64": field_diff,
        "assembly_ms": round(t_asm_ms, 3),
        "roofline": roof64,
        "dropin_petsc_opts": dropin,
        "c3_mixed_precision": c3,
        "batched": batched,
        "multimesh": multimesh,
        "adjoint": adjoint,
        "cpu_baseline": None if cpu is None else {
            k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample", "assembly_s",
                                "solve_s", "iterations", "c1")},
        "host": platform.processor() or platform.machine(),
    }
    if cpu is not None:
        line["cpu_baseline"]["speedup_vs_gpu_fp64"] = round(cpu["value"] / t_fp64, 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
