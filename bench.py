#!/usr/bin/env python3
"""Benchmark of the MI355X matrix-free Laplacian apply (one GPU).

Default (the contract line): 3D unit cube, 64^3 hexahedra with interior
vertices perturbed by U[-0.2h, 0.2h] (seed 1), Q4 (5 nodes/dir, 16,581,375
free dofs), 6 Gauss points/dir, FP64 apply y = A u with u ~ U[-1,1] (seed 0)
resident in HBM.  A step is one y = A u over all cells: the zero fill of y
(its shared dofs receive atomic adds) and one dispatch of the apply kernel.
Prints one JSON line with
GDOF/s, the roofline of the apply (HIP events on its stream around each
apply) and the CPU oracle timed on the host cores.

Other modes (one JSON line each): --mode sweep (Q2..Q7 at ~16M dofs, FP64 and
FP32), --mode cg (Q7 32^3: FP64 CG vs mixed-precision refinement to 1e-10),
--mode adjoint (Q5 48^3 kappa-gradient kernel), --mode q9 (24^3 Q9 solves and
adjoint), --mode ref2d (the reference's own 2D 1024^2 Q4-Q7 benchmark).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 vector
FP32_PEAK_TFLOPS = 157.3    # MI355X FP32 vector
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "apply3d_f64_fusedgeom_q4_64cube_traffic.txt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def alg_bytes(n_free, n_cells, n, q, D, s_vec=8, s_g=8):
    """Algorithmic HBM bytes of one apply: read x + write y, the int32 cell
    dofs, the stored geometric factor."""
    return ((s_vec + s_vec) * n_free + 4 * n ** D * n_cells
            + s_g * (D * (D + 1) // 2) * q ** D * n_cells)


def alg_bytes_fused(n_free, n_cells, n, D, s_vec=8):
    """Algorithmic HBM bytes of one apply that recomputes the geometric
    factor from each cell's multilinear map: x, y, cell dofs, the map's
    coefficient vectors (7 x 3 values per cell in 3D, 3 x 2 in 2D)."""
    ncoef = 21 if D == 3 else 6
    return (s_vec + s_vec) * n_free + 4 * n ** D * n_cells + s_vec * ncoef * n_cells


STORED_G_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "apply3d_f64_q4_64cube_traffic.txt")


def read_traffic(path=TRAFFIC_FILE):
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def build_problem(pkg, D, m, p, q, perturb=0.2, kappa=None):
    model = pkg.CartesianModel(D, m, perturb=perturb, seed=1)
    space = pkg.FESpace(model, p)
    mp = pkg.SumFactorizationMap.with_points(D, p, q)
    A = pkg.LazyMatrix(mp, space, kappa=kappa)
    return model, space, mp, A


def stats_ms(ts):
    ts = sorted(ts)
    return {"mean": sum(ts) / len(ts), "min": ts[0], "median": ts[len(ts) // 2],
            "max": ts[-1]}


def time_apply(torch, A, x, y, steps, warmup, reduced=None, detail=False):
    """Wall time per step, and the apply kernel's mean duration from the
    library's HIP events recorded on the launch stream around the kernel
    alone (not the fill of y).  detail: also the min / median / max of the
    kernel and of each step (fill + kernel, HIP events on the stream around
    each y = A x), as the reference's benchmark reports minima and maxima
    (profiling/benchmark.jl:7-14)."""
    stream = torch.cuda.current_stream()

    def step():
        if reduced:
            A.mul_reduced(y, x, reduced)
        else:
            A.mul(y, x)

    for _ in range(warmup):
        step()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    torch.cuda.synchronize()
    A.timing(True)
    t0 = time.perf_counter()
    for i in range(steps):
        ev[i][0].record(stream)
        step()
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    A.timing(False)
    kern = A.read_timing()
    assert len(kern) == steps, (len(kern), steps)
    kstat = stats_ms(kern)
    if not detail:
        return wall, kstat["mean"] / 1e3
    step_ms = [a.elapsed_time(b) for a, b in ev]
    return wall, kstat["mean"] / 1e3, {"kernel_ms": kstat, "step_ms": stats_ms(step_ms)}


def cpu_info():
    """Usable host threads (this process's CPU affinity, capped by
    OMP_NUM_THREADS where set: the box's CPU share) and the CPU model."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return max(1, n), model


def mode_apply(args, torch, pkg):
    import numpy as np
    D, m, p, q = 3, 64, 4, 6
    if args.order != 4:  # another order of the sweep (~16M dofs, q = p + 2)
        p, q, m = args.order, args.order + 2, int(round(256 / args.order))
    log("setting up 3D %d^3 Q%d q=%d perturbed ..." % (m, p, q))
    model, space, mp, A = build_problem(pkg, D, m, p, q)
    n = p + 1
    xh = pkg.uniform(0, space.n_free, -1.0, 1.0)
    x = torch.from_numpy(xh).cuda()
    y = torch.empty_like(x)
    log("timing %d steps after %d warmup ..." % (args.steps, args.warmup))
    wall, kern, det = time_apply(torch, A, x, y, args.steps, args.warmup, detail=True)
    # the default apply recomputes G from the vertices: bytes are x, y, ids
    # and vertices, and the bound is FP64 arithmetic (the reference's FLOP
    # model, which counts the metric work at every quadrature point)
    Balg = alg_bytes_fused(space.n_free, space.n_cells, n, D)
    Falg = pkg.count_manual_flops_poisson_matrix_free(space.n_cells, D, (q,) * D, (n,) * D,
                                                      count_only_fma_ops=False,
                                                      structured_mesh_opt=False)
    t_hbm, t_flop = Balg / (HBM_PEAK_GBS * 1e9), Falg / (FP64_PEAK_TFLOPS * 1e12)
    achieved = Falg / kern / 1e12
    out = {
        "metric": "Laplacian apply GDOF/s (3D Q4-Q7, FP64 vs mixed CG@1e-10), % of roofline",
        "value": space.n_free / wall / 1e9,
        "unit": "GDOF/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded perturbed mesh, u ~ U[-1,1])",
        "config": {"workload": "3D unit cube %d^3 hexes perturbed +-0.2h, Q%d (%d nodes/dir), "
                               "%d Gauss pts/dir, FP64 apply y = A u" % (m, p, p + 1, q),
                   "n_free": space.n_free, "n_cells": space.n_cells, "order": p, "nquad": q},
        "roofline": {"bound": "fp64" if t_flop >= t_hbm else "hbm",
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": read_traffic() if args.order == 4 else None,
                     "traffic_source": "rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this "
                                       "kernel on this workload, scaled by a 1 GiB copy "
                                       "(profiles/apply3d_f64_fusedgeom_q4_64cube_traffic.txt); "
                                       "not measured in this run",
                     "peak_note": "FP64 peak, vector ALU = matrix cores on MI355X",
                     "kernel": "k_apply3d<%d,%d,double,double,fused geometry>" % (n, q),
                     "kernel_ms": kern * 1e3,
                     "kernel_ms_min": det["kernel_ms"]["min"],
                     "kernel_ms_median": det["kernel_ms"]["median"],
                     "kernel_timing": "HIP events on the launch stream around the apply kernel "
                                      "alone, every timed step (mean above)",
                     "step_ms": det["step_ms"],
                     "dispatches_per_apply": A.info()["launches_per_apply"],
                     "alg_bytes": Balg, "alg_flops": Falg,
                     "achieved_GBs": Balg / kern / 1e9},
    }
    # host vectors (sf_apply_host: x staged in, A x, y staged out over PCIe);
    # a side figure, never `value`
    yh = np.empty_like(xh)
    A.mul(yh, xh)
    t0 = time.perf_counter()
    for _ in range(5):
        A.mul(yh, xh)
    th = (time.perf_counter() - t0) / 5
    out["host_vector_apply"] = {"gdofs": space.n_free / th / 1e9, "ms": th * 1e3,
                                "note": "sf_apply_host on pageable numpy vectors: "
                                        "PCIe-inclusive"}
    # the stored-G form of the same apply (G read from HBM, HBM-bound)
    os.environ["SFHIP_FUSED_GEOMETRY"] = "0"
    model_s, space_s, mp_s, A_s = build_problem(pkg, D, m, p, q)
    del os.environ["SFHIP_FUSED_GEOMETRY"]
    y_s = torch.empty_like(x)
    wall_s, kern_s = time_apply(torch, A_s, x, y_s, args.steps, args.warmup)
    Bs = alg_bytes(space.n_free, space.n_cells, n, q, D)
    out["stored_g_variant"] = {"gdofs": space.n_free / wall_s / 1e9, "kernel_ms": kern_s * 1e3,
                               "alg_bytes": Bs, "hbm_GBs": Bs / kern_s / 1e9,
                               "hbm_frac": Bs / kern_s / 1e9 / HBM_PEAK_GBS,
                               "traffic": (read_traffic(STORED_G_TRAFFIC_FILE)
                                           if args.order == 4 else None),
                               "max_rel_diff_vs_fused": float(torch.linalg.norm(y_s - y) /
                                                              torch.linalg.norm(y))}
    A_s.free()
    # parity + CPU baseline from the same oracle run
    yd = y.cpu().numpy()
    if not args.no_cpu:
        import oracle
        threads, model_name = cpu_info()
        X = model.cell_vertices()
        log("CPU oracle (%d threads, %s) ..." % (threads, model_name))
        y_ref = oracle.apply(D, n, q, space.cell_dofs, X, xh, mp.tables, nthreads=threads)
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            y_ref = oracle.apply(D, n, q, space.cell_dofs, X, xh, mp.tables, nthreads=threads)
            ts.append(time.perf_counter() - t0)
        log("CPU oracle, 1 thread ...")
        t0 = time.perf_counter()
        oracle.apply(D, n, q, space.cell_dofs, X, xh, mp.tables, nthreads=1)
        t1c = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": space.n_free / min(ts) / 1e9, "unit": "GDOF/s",
                               "cores": threads, "kind": "port",
                               "sample": "full 64^3 Q4 q=6 apply (the same workload), OpenMP C "
                                         "oracle; value = best of 3 after 1 warm-up on the "
                                         "usable host threads",
                               "mean_gdofs": space.n_free / (sum(ts) / len(ts)) / 1e9,
                               "one_core_gdofs": space.n_free / t1c / 1e9,
                               "one_core_s": t1c, "cpu_model": model_name}
        out["parity_rel_l2"] = float(np.linalg.norm(yd - y_ref) / np.linalg.norm(y_ref))
        A.free()
        out["assembled_comparator"] = assembled_comparator(torch, pkg, threads)
        return out
    A.free()
    return out


def assembled_comparator(torch, pkg, threads):
    """Matrix-free vs assembled at a reduced 3D size (full-size CSR would
    not fit): the oracle's brute-force element matrices assembled into CSR,
    SpMV on the host (scipy, one thread) and on the GPU (torch CSR), the
    OpenMP oracle on 1 and on `threads` cores, and the GPU matrix-free
    apply, all on the same seeded vector."""
    import numpy as np
    import oracle
    from oracle import brute
    D, m, p, q = 3, 16, 4, 6
    n = p + 1
    log("assembled comparator: 3D %d^3 Q%d q=%d ..." % (m, p, q))
    model, space, mp, A = build_problem(pkg, D, m, p, q)
    X = model.cell_vertices()
    B, Dm, pts, w = brute.tables(p, q)
    G, _ = brute.geometry_numpy(D, X, pts, w)
    Acsr = brute.assemble_csr(space.cell_dofs, brute.element_matrices(D, B, Dm, G), space.n_free)
    xh = pkg.uniform(0, space.n_free, -1.0, 1.0)

    def best(f, reps):
        f()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            f()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    y_sp = Acsr @ xh
    t_cpu_csr = best(lambda: Acsr @ xh, 5)
    t_or1 = best(lambda: oracle.apply(D, n, q, space.cell_dofs, X, xh, mp.tables, nthreads=1), 2)
    t_orN = best(lambda: oracle.apply(D, n, q, space.cell_dofs, X, xh, mp.tables,
                                      nthreads=threads), 3)
    x = torch.from_numpy(xh).cuda()
    y = torch.empty_like(x)
    wall, kern = time_apply(torch, A, x, y, 50, 5)
    err_mf = float(np.linalg.norm(y.cpu().numpy() - y_sp) / np.linalg.norm(y_sp))
    Ag = torch.sparse_csr_tensor(torch.from_numpy(Acsr.indptr.astype(np.int64)),
                                 torch.from_numpy(Acsr.indices.astype(np.int64)),
                                 torch.from_numpy(Acsr.data), size=Acsr.shape).cuda()
    yg = Ag @ x
    torch.cuda.synchronize()
    err_csr = float(np.linalg.norm(yg.cpu().numpy() - y_sp) / np.linalg.norm(y_sp))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for _ in range(5):
        yg = Ag @ x
    torch.cuda.synchronize()
    e0.record()
    for _ in range(50):
        yg = Ag @ x
    e1.record()
    torch.cuda.synchronize()
    t_gpu_csr = e0.elapsed_time(e1) / 50 / 1e3
    nf = space.n_free
    res = {"workload": "3D %d^3 Q%d q=%d perturbed, FP64 (full-size CSR would not fit)" % (m, p, q),
           "n_free": nf, "csr_nnz": int(Acsr.nnz),
           "gpu_matrix_free_gdofs": nf / wall / 1e9,
           "gpu_csr_gdofs": nf / t_gpu_csr / 1e9,
           "cpu_csr_1core_gdofs": nf / t_cpu_csr / 1e9,
           "cpu_oracle_1core_gdofs": nf / t_or1 / 1e9,
           "cpu_oracle_gdofs": nf / t_orN / 1e9, "cpu_oracle_cores": threads,
           "matrix_free_vs_csr_rel_l2": err_mf, "gpu_csr_vs_cpu_csr_rel_l2": err_csr}
    log("assembled comparator", res)
    A.free()
    return res


def mode_sweep(args, torch, pkg):
    rows = []
    for p in [int(t) for t in args.orders.split(",") if t]:
        q = p + 2
        m = int(round(256 / p))
        model, space, mp, A = build_problem(pkg, 3, m, p, q)
        n = p + 1
        x = torch.from_numpy(pkg.uniform(0, space.n_free, -1, 1)).cuda()
        y = torch.empty_like(x)
        wall, kern, det = time_apply(torch, A, x, y, args.steps, args.warmup, detail=True)
        # fused-geometry FP64 apply: bound by the larger of its bytes and the
        # reference's FLOP model
        B64 = alg_bytes_fused(space.n_free, space.n_cells, n, 3)
        F = pkg.count_manual_flops_poisson_matrix_free(space.n_cells, 3, (q,) * 3, (n,) * 3,
                                                       count_only_fma_ops=False,
                                                       structured_mesh_opt=False)
        t_hbm, t_flop = B64 / (HBM_PEAK_GBS * 1e9), F / (FP64_PEAK_TFLOPS * 1e12)
        row = {"order": p, "nquad": q, "m": m, "n_free": space.n_free,
               "f64_gdofs": space.n_free / wall / 1e9, "f64_kernel_ms": kern * 1e3,
               "f64_kernel_ms_median": det["kernel_ms"]["median"],
               "f64_kernel_ms_min": det["kernel_ms"]["min"],
               "f64_roofline_frac": max(t_hbm, t_flop) / kern,
               "f64_roofline_frac_median": max(t_hbm, t_flop) / (det["kernel_ms"]["median"] * 1e-3),
               "f64_bound": "fp64" if t_flop >= t_hbm else "hbm",
               "f64_tflops": F / kern / 1e12, "f64_GBs": B64 / kern / 1e9}
        for prec in ("f32", "f32_gbf16", "mfma_f32", "mfma_bf16x3", "mfma_bf16", "mfma_f16"):
            if prec == "mfma_f16" and n != 8:  # the FP16 GEMM apply is built for n = 8
                continue
            x32 = x.float()
            y32 = torch.empty_like(x32)
            A.prepare(prec)
            w32, k32 = time_apply(torch, A, x32, y32, args.steps, args.warmup, reduced=prec)
            sg = 4 if prec in ("f32", "mfma_f32") else 2
            B32 = alg_bytes(space.n_free, space.n_cells, n, q, 3, s_vec=4, s_g=sg)
            row[prec + "_gdofs"] = space.n_free / w32 / 1e9
            row[prec + "_kernel_ms"] = k32 * 1e3
            row[prec + "_GBs"] = B32 / k32 / 1e9
        log(json.dumps(row))
        rows.append(row)
        A.free()
        del x, y
        torch.cuda.empty_cache()
    return {"metric": "Q2-Q7 apply sweep at ~16M dofs (3D perturbed, q = p + 2)", "rows": rows,
            "unit": "GDOF/s", "n_gpus": 1}


def mode_cg(args, torch, pkg):
    import numpy as np
    D, m, p, q = 3, 32, 7, 9
    log("setting up 3D 32^3 Q7 q=9 perturbed ...")
    model, space, mp, A = build_problem(pkg, D, m, p, q)
    b = torch.from_numpy(pkg.normal(2, space.n_free)).cuda()
    res = {"metric": "CG to 1e-10: FP64 vs mixed-precision refinement, 3D 32^3 Q7 q=9",
           "n_free": space.n_free, "n_gpus": 1}
    # the reference's CG benchmark (profiling/benchmark.jl:16-25): min and
    # max over 10 runs of cg!(x, A, b, maxiter = 10) after one warm-up
    times = []
    for i in range(11):
        x10 = torch.zeros_like(b)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pkg.cg_(x10, A, b, rtol=1e-10, maxiter=10)
        torch.cuda.synchronize()
        if i:
            times.append(time.perf_counter() - t0)
    res["cg_maxiter10_ms"] = {"min": min(times) * 1e3, "max": max(times) * 1e3}
    log("cg maxiter=10", res["cg_maxiter10_ms"])
    x64 = torch.zeros_like(b)
    st = pkg.cg_(x64, A, b, rtol=1e-10, maxiter=args.maxit, Pl=args.pl)
    res["f64"] = st
    res["preconditioner"] = args.pl
    log("f64", st)
    for spec in [t for t in args.mixed.split(",") if t]:
        parts = spec.split(":")
        inner, delta = parts[0], parts[1]
        rec = parts[2] if len(parts) > 2 else "f32"
        x = torch.zeros_like(b)
        st = pkg.cg_mixed_(x, A, b, rtol=1e-10, inner=inner, delta=float(delta),
                           maxiter=args.maxit, recurrence=rec, Pl=args.pl)
        st["err_vs_f64"] = float(torch.linalg.norm(x - x64) / torch.linalg.norm(x64))
        st["speedup_vs_f64"] = res["f64"]["seconds"] / st["seconds"]
        res["mixed_%s_delta%s_rec%s" % (inner, delta, rec)] = st
        log("mixed", inner, delta, rec, st)
    for spec in [t for t in args.inner.split(",") if t]:
        inner, irt = spec.split(":")
        irt = float(irt)
        x = torch.zeros_like(b)
        try:
            st = pkg.cg_ir_(x, A, b, rtol=1e-10, inner=inner, inner_rtol=irt,
                            inner_maxiter=args.maxit, max_outer=args.max_outer, Pl=args.pl)
        except pkg.SFError as e:  # e.g. refinement stalls with bf16 operands
            res["%s_inner%g" % (inner, irt)] = {"error": str(e)}
            log(inner, irt, "failed:", e)
            continue
        st["err_vs_f64"] = float(torch.linalg.norm(x - x64) / torch.linalg.norm(x64))
        st["speedup_vs_f64"] = res["f64"]["seconds"] / st["seconds"]
        res["%s_inner%g" % (inner, irt)] = st
        log(inner, irt, st)
    _ = np
    A.free()
    return res


def fd_check(torch, A, kappa, u, d, kb, ub, ndir, h=1e-6):
    """Central finite differences of J = 1/2 |A(kappa) u - d|^2 along ndir
    random unit directions in kappa and in u (h = 1e-6, FP64), against the
    adjoint's <dJ/dkappa, dk> and <dJ/du, du>.  J is quadratic in each, so
    the only error is rounding; J(+) - J(-) is formed as
    1/2 (r+ - r-) . (r+ + r-) to keep the subtraction at the residual level.
    Error measure per direction: |fd - an| / (|an| + 1e-3 |grad|)."""
    gen = torch.Generator(device="cuda")
    kd = torch.from_numpy(kappa).cuda()
    y = torch.empty_like(u)

    def resid(v):
        A.mul(y, v)
        return y - d

    def central(rp, rm):
        return 0.5 * float(((rp - rm) * (rp + rm)).sum()) / (2 * h)

    errs_k, errs_u = [], []
    for s in range(ndir):
        gen.manual_seed(1000 + s)
        dk = torch.randn(kd.shape, generator=gen, device="cuda", dtype=torch.float64)
        dk /= torch.linalg.norm(dk)
        A.set_kappa((kd + h * dk).cpu().numpy())
        rp = resid(u)
        A.set_kappa((kd - h * dk).cpu().numpy())
        rm = resid(u)
        an = float((kb * dk).sum())
        errs_k.append(abs(central(rp, rm) - an) / (abs(an) + 1e-3 * float(torch.linalg.norm(kb))))
    A.set_kappa(kappa)
    for s in range(ndir):
        gen.manual_seed(2000 + s)
        du = torch.randn(u.shape, generator=gen, device="cuda", dtype=torch.float64)
        du /= torch.linalg.norm(du)
        rp = resid(u + h * du)
        rm = resid(u - h * du)
        an = float((ub * du).sum())
        errs_u.append(abs(central(rp, rm) - an) / (abs(an) + 1e-3 * float(torch.linalg.norm(ub))))
    return {"directions": ndir, "h": h, "max_err_dkappa": max(errs_k),
            "max_err_du": max(errs_u), "median_err_dkappa": sorted(errs_k)[ndir // 2],
            "median_err_du": sorted(errs_u)[ndir // 2],
            "directions_from": "torch.randn on the device, seeds 1000+s (kappa), 2000+s (u)"}


def mode_adjoint(args, torch, pkg):
    """Differentiation workload: 3D 48^3 perturbed hexes, Q5, 7 Gauss
    pts/dir, log-normal kappa per quadrature point.  The pullback of
    J = 1/2 |A(kappa) u - d|^2 (dJ/dkappa and dJ/du = A(kappa) r) by the
    block-assembled fused-geometry kernel (the default) and by the per-cell
    stored-G kernel (a plan built with SFHIP_FUSED_GEOMETRY=0), and the
    kappa apply in both forms."""
    import ctypes
    import numpy as np
    D, m, p, q = 3, 48, 5, 7
    n = p + 1
    log("setting up 3D 48^3 Q5 q=7 perturbed with kappa ...")
    model = pkg.CartesianModel(D, m, perturb=0.2, seed=1)
    space = pkg.FESpace(model, p)
    mp = pkg.SumFactorizationMap.with_points(D, p, q)
    nq = q ** D
    nf, E = space.n_free, space.n_cells
    kappa = np.exp(pkg.normal(3, E * nq))
    u = torch.from_numpy(pkg.uniform(0, nf, -1, 1)).cuda()
    d = torch.from_numpy(pkg.uniform(4, nf, -1, 1)).cuda()
    res = {"metric": "adjoint (dJ/dkappa + dJ/du) 3D 48^3 Q5 q=7, log-normal kappa",
           "n_free": nf, "n_cells": E, "n_gpus": 1}
    outs = {}
    for form in ("fused", "stored_g"):
        if form == "stored_g":
            os.environ["SFHIP_FUSED_GEOMETRY"] = "0"
        A = pkg.LazyMatrix(mp, space, kappa=kappa)
        os.environ.pop("SFHIP_FUSED_GEOMETRY", None)
        y = torch.empty_like(u)
        wall, kern = time_apply(torch, A, u, y, args.steps, args.warmup)
        r = y - d
        kb = torch.empty(E * nq, dtype=torch.float64, device="cuda")
        ub = torch.empty_like(u)
        call = lambda: pkg._lib.check(A._L.sf_adjoint(
            A.plan, ctypes.c_void_p(u.data_ptr()), ctypes.c_void_p(r.data_ptr()),
            ctypes.c_void_p(kb.data_ptr()), ctypes.c_void_p(ub.data_ptr()), None))
        for _ in range(args.warmup):
            call()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            call()
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / args.steps
        if form == "fused":
            # u, r read and dJ/du written; cell dofs; the trilinear map's 21
            # coefficients per cell; kappa read and dJ/dkappa written per
            # quadrature point
            B = 8 * 3 * nf + 4 * n ** 3 * E + 8 * 21 * E + 8 * 2 * nq * E
        else:
            # the same with the six G entries per quadrature point
            B = 8 * 3 * nf + 4 * n ** 3 * E + 8 * (6 + 1 + 1) * nq * E
        res[form] = {"adjoint_ms": t * 1e3, "alg_bytes": B, "GB/s": B / t / 1e9,
                     "frac_hbm": B / t / 1e9 / HBM_PEAK_GBS,
                     "kappa_apply_kernel_ms": kern * 1e3,
                     "adjoint_over_apply": t / kern}
        outs[form] = (y.clone(), kb.clone(), ub.clone())
        log(form, res[form])
        if form == "fused" and args.fd_dirs > 0:
            log("central finite differences along %d directions ..." % args.fd_dirs)
            res["fd_check"] = fd_check(torch, A, kappa, u, d, kb, ub, args.fd_dirs)
            log("fd", res["fd_check"])
        A.free()
    for i, name in enumerate(("Au", "dJdkappa", "dJdu")):
        a, b = outs["fused"][i], outs["stored_g"][i]
        res["rel_l2_fused_vs_stored_" + name] = float(torch.linalg.norm(a - b)
                                                      / torch.linalg.norm(b))
    return res


def mode_q9(args, torch, pkg):
    """Highest order at scale: 3D 24^3 perturbed hexes, Q9 (10 nodes/dir),
    11 Gauss pts/dir, log-normal kappa = exp(g), g ~ N(0,1) per quadrature
    point.  FP64 apply (roofline), the reduced-precision and MFMA applies
    (matrix-core utilisation, error against the FP64 apply), FP64 CG and the
    mixed-precision CG to 1e-10, and the adjoint gradient."""
    import ctypes
    import numpy as np
    D, m, p, q = 3, 24, 9, 11
    n = p + 1
    log("setting up 3D 24^3 Q9 q=11 perturbed, log-normal kappa ...")
    model = pkg.CartesianModel(D, m, perturb=0.2, seed=1)
    space = pkg.FESpace(model, p)
    mp = pkg.SumFactorizationMap.with_points(D, p, q)
    nq = q ** D
    kappa = np.exp(pkg.normal(3, space.n_cells * nq))
    A = pkg.LazyMatrix(mp, space, kappa=kappa)
    nf, E = space.n_free, space.n_cells
    res = {"metric": "3D 24^3 Q9 q=11 perturbed, log-normal kappa", "n_free": nf,
           "n_cells": E, "n_gpus": 1}
    xh = pkg.uniform(0, nf, -1.0, 1.0)
    x = torch.from_numpy(xh).cuda()
    y = torch.empty_like(x)
    wall, kern = time_apply(torch, A, x, y, args.steps, args.warmup)
    # fused-geometry apply reading kappa per quadrature point: bytes of the
    # vectors, cell dofs, map coefficients and kappa; FLOPs of the reference's
    # model
    B64 = alg_bytes_fused(nf, E, n, D) + 8 * nq * E
    F = pkg.count_manual_flops_poisson_matrix_free(E, D, (q,) * D, (n,) * D,
                                                   count_only_fma_ops=False,
                                                   structured_mesh_opt=False)
    res["f64_apply"] = {"gdofs": nf / wall / 1e9, "kernel_ms": kern * 1e3,
                        "alg_bytes": B64, "alg_flops": F,
                        "hbm_frac": B64 / kern / 1e9 / HBM_PEAK_GBS,
                        "fp64_frac": F / kern / 1e12 / FP64_PEAK_TFLOPS}
    # the FP64 apply at this (n, q) with kappa is checked against the oracle
    # by the GPU tests; the reduced forms are measured against it here
    y64 = y.clone()
    # tensor-contraction FMAs of the MFMA schedule (useful work, no padding)
    fma_cell = 2 * (2 * n ** 3 * q + 3 * n ** 2 * q ** 2 + 3 * n * q ** 3)
    x32 = x.float()
    for prec in ("f32", "f32_gbf16", "mfma_f32", "mfma_bf16x3", "mfma_bf16"):
        A.prepare(prec)
        y32 = torch.empty_like(x32)
        w32, k32 = time_apply(torch, A, x32, y32, args.steps, args.warmup, reduced=prec)
        err = float(torch.linalg.norm(y32.double() - y64) / torch.linalg.norm(y64))
        r = {"gdofs": nf / w32 / 1e9, "kernel_ms": k32 * 1e3, "rel_l2_vs_f64": err}
        if prec == "mfma_f32":
            useful = 2.0 * fma_cell * E
            r["useful_tflops"] = useful / k32 / 1e12
            r["mfma_util_vs_fp32_matrix_peak"] = useful / k32 / 157.3e12
        elif prec.startswith("mfma"):
            terms = 6 if prec.endswith("x3") else 1
            useful = 2.0 * fma_cell * E
            r["useful_tflops"] = useful / k32 / 1e12
            r["mfma_util_vs_bf16_dense_peak"] = useful * terms / k32 / 2.5e15
        res[prec] = r
        log(prec, r)
    b = torch.from_numpy(pkg.normal(2, nf)).cuda()
    x64 = torch.zeros_like(b)
    # log-normal kappa multiplies the condition number by about
    # max(kappa) / min(kappa): both solves run under the same iteration cap
    # and report the residual they reach
    st = pkg.cg_(x64, A, b, rtol=1e-10, maxiter=args.q9_maxit, Pl=args.pl)
    res["cg_f64"] = st
    res["preconditioner"] = args.pl
    log("cg f64", st)
    for rec in [t for t in args.q9_mixed.split(",") if t]:
        xm = torch.zeros_like(b)
        st = pkg.cg_mixed_(xm, A, b, rtol=1e-10, inner="f32", delta=0.1, maxiter=args.q9_maxit,
                           recurrence=rec, Pl=args.pl)
        st["err_vs_f64"] = float(torch.linalg.norm(xm - x64) / torch.linalg.norm(x64))
        st["speedup_vs_f64"] = res["cg_f64"]["seconds"] / st["seconds"]
        res["cg_mixed_f32_rec" + rec] = st
        log("cg mixed f32", rec, st)
    xi = torch.zeros_like(b)
    st = pkg.cg_ir_(xi, A, b, rtol=1e-10, inner="f32", inner_rtol=1e-2,
                    inner_maxiter=args.q9_maxit, max_outer=args.max_outer, Pl=args.pl)
    st["err_vs_f64"] = float(torch.linalg.norm(xi - x64) / torch.linalg.norm(x64))
    st["speedup_vs_f64"] = res["cg_f64"]["seconds"] / st["seconds"]
    res["cg_ir_f32"] = st
    log("cg refinement f32", st)
    # adjoint of J = 1/2 |A(kappa) u - d|^2
    d = torch.from_numpy(pkg.uniform(4, nf, -1.0, 1.0)).cuda()
    A.mul(y, x)
    rr = y - d
    kb = torch.empty(E * nq, dtype=torch.float64, device="cuda")
    ub = torch.empty_like(x)
    call = lambda: pkg._lib.check(A._L.sf_adjoint(A.plan, ctypes.c_void_p(x.data_ptr()),
                                                  ctypes.c_void_p(rr.data_ptr()),
                                                  ctypes.c_void_p(kb.data_ptr()),
                                                  ctypes.c_void_p(ub.data_ptr()), None))
    for _ in range(3):
        call()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        call()
    torch.cuda.synchronize()
    ta = (time.perf_counter() - t0) / 10
    # fused-geometry adjoint: vectors, cell dofs, map coefficients, kappa,
    # dJ/dkappa
    Badj = 8 * 3 * nf + 4 * n ** 3 * E + 8 * 21 * E + 8 * 2 * nq * E
    res["adjoint"] = {"ms": ta * 1e3, "hbm_frac": Badj / ta / 1e9 / HBM_PEAK_GBS,
                      "over_f64_apply": ta / kern}
    log("adjoint", res["adjoint"])
    A.free()
    return res


# The reference's own GPU figures (test/my_plot.jl:4-7): seconds per apply of
# its CUDA.jl kernel on the 2D unit square, 1024^2 Cartesian cells, orders 4-7
# with Gauss degree 2(p - 1) (test/benchmark_v3.jl:37-49), on the V100 of its
# job scripts, and its one-core CPU sum-factorized mul!.
REF2D_GPU_S = {4: 0.003, 5: 0.0051, 6: 0.00805, 7: 0.014}
REF2D_CPU_S = {4: 0.5, 5: 0.74, 6: 0.91, 7: 1.09}


def mode_ref2d(args, torch, pkg):
    """The reference's benchmark workload: 2D unit square, 1024^2 Cartesian
    cells, Q4-Q7, Gauss degree 2(p - 1) -> q = p points per direction
    (Gridap's Gauss rule takes ceil((degree + 1) / 2) points), homogeneous
    Dirichlet, FP64 y = A u.  The reference's kernel applied only the
    quadrature weights (a Cartesian mesh needs no Jacobian); this apply reads
    the geometric factor recomputed from the cell vertices, so it does
    more work per cell.  Timed like
    test/benchmark_v3.jl:20-33: a warm-up, then the minimum (and maximum)
    over the timed applies, next to the mean and median."""
    rows = []
    for p in [int(t) for t in args.orders.split(",") if t]:
        if p not in REF2D_GPU_S:
            continue
        q = p
        model, space, mp, A = build_problem(pkg, 2, 1024, p, q, perturb=0.0)
        x = torch.from_numpy(pkg.uniform(0, space.n_free, -1, 1)).cuda()
        y = torch.empty_like(x)
        wall, kern, det = time_apply(torch, A, x, y, args.steps, args.warmup, detail=True)
        step = det["step_ms"]
        nf = space.n_free
        # roofline of the fused-geometry kernel: its bytes, and the
        # reference's FLOP model (Utils.jl:33-72, general geometry)
        B = alg_bytes_fused(nf, space.n_cells, p + 1, 2)
        F = pkg.count_manual_flops_poisson_matrix_free(space.n_cells, 2, (q,) * 2, (p + 1,) * 2,
                                                       count_only_fma_ops=False,
                                                       structured_mesh_opt=False)
        km = det["kernel_ms"]["median"] * 1e-3
        t_hbm, t_flop = B / (HBM_PEAK_GBS * 1e9), F / (FP64_PEAK_TFLOPS * 1e12)
        row = {"order": p, "nquad": q, "cells": space.n_cells, "n_free": nf,
               "kernel_ms": det["kernel_ms"], "step_ms": step,
               "gdofs_min_step": nf / (step["min"] * 1e-3) / 1e9,
               "gdofs_median_step": nf / (step["median"] * 1e-3) / 1e9,
               "reference_v100_s": REF2D_GPU_S[p],
               "reference_v100_gdofs": nf / REF2D_GPU_S[p] / 1e9,
               "speedup_vs_reference_gpu": REF2D_GPU_S[p] / (step["min"] * 1e-3),
               "reference_cpu_1core_s": REF2D_CPU_S[p],
               "alg_bytes": B, "alg_flops": F,
               "hbm_GBs_median": B / km / 1e9, "tflops_median": F / km / 1e12,
               "roofline_bound": "fp64" if t_flop >= t_hbm else "hbm",
               "roofline_frac_median": max(t_hbm, t_flop) / km}
        log(json.dumps(row))
        rows.append(row)
        A.free()
        del x, y
        torch.cuda.empty_cache()
    return {"metric": "the reference's 2D benchmark: apply of 1024^2 Q4-Q7 (q = p), FP64",
            "rows": rows, "unit": "GDOF/s", "n_gpus": 1,
            "note": "step = zero fill of y + the apply kernel; the reference's times are "
                    "test/my_plot.jl's hand-entered figures (V100, CUDA.jl gpu_mul_v3!, "
                    "weights only)"}


def mode_copycal(args, torch, pkg):
    """A plain device copy of 1 GiB, three times: the reference dispatch that
    calibrates the FETCH_SIZE / WRITE_SIZE counters (bytes moved are known)."""
    n = (1 << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    return {"metric": "device copy of 1 GiB (counter calibration)", "bytes_read": n * 8,
            "bytes_written": n * 8, "n_gpus": 1}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--mode", default="apply",
                    choices=["apply", "sweep", "cg", "adjoint", "q9", "ref2d", "copycal"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--order", type=int, default=4,
                    help="apply mode: polynomial order (4: the 64^3 Q4 contract workload; "
                         "others: the sweep's mesh of about 16M dofs with q = p + 2)")
    ap.add_argument("--orders", default="2,3,4,5,6,7",
                    help="sweep mode: polynomial orders (comma list; q = p + 2); "
                         "ref2d mode: orders among 4-7")
    ap.add_argument("--maxit", type=int, default=300000)
    ap.add_argument("--max-outer", type=int, default=50)
    ap.add_argument("--fd-dirs", type=int, default=64,
                    help="adjoint mode: central-difference directions in kappa and in u")
    ap.add_argument("--q9-maxit", type=int, default=40000,
                    help="q9 mode: iteration cap of the CG solves")
    ap.add_argument("--q9-mixed", default="f64p32,f64,f32",
                    help="q9 mode: mixed-CG recurrences to run (comma list)")
    ap.add_argument("--pl", default=None, choices=[None, "jacobi", "gll"],
                    help="cg / q9 modes: preconditioner of every solve (None: the "
                         "reference's unpreconditioned CG)")
    ap.add_argument("--inner", default="f32:1e-2,f32_gbf16:1e-2,mfma_f16:1e-2,mfma_bf16x3:1e-2,mfma_bf16:1e-2",
                    help="cg mode: restarted refinement runs, comma list of "
                         "inner_precision:inner_rtol")
    ap.add_argument("--mixed", default="f32:0.1:f64p32,f32:0.1:f64,f32:0.1,mfma_f16:0.1:f64,mfma_f16:0.1,f32_gbf16:0.1:f64,mfma_bf16x3:0.1:f64",
                    help="cg mode: reliable-update mixed CG runs, comma list of "
                         "inner_precision:delta[:recurrence f32|f64|f64p32]")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (--gpus %d requested); the apply "
              "would shard by cell slabs with a halo sum of interface dofs, which is not "
              "built here." % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    import __graft_entry__
    pkg = __graft_entry__.import_package()
    torch.cuda.set_device(0)
    fn = {"apply": mode_apply, "sweep": mode_sweep, "cg": mode_cg, "adjoint": mode_adjoint,
          "q9": mode_q9, "ref2d": mode_ref2d, "copycal": mode_copycal}
    out = fn[args.mode](args, torch, pkg)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
