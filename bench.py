#!/usr/bin/env python3
"""Benchmark: 3D steady-state thermal FEM solves/s to 1e-8 relative residual on one MI355X.

Workload: 128 x 128 x 256 nodes (4,194,304 DOF), 8-node trilinear hexahedra, 40 seeded
anisotropic layers (kxy, kz log-uniform in [1,150] W/m/K), cylindrical 1e16 W/m^3
source in the 0.1 um active layer, Dirichlet 300 K at the bottom (plask_amd.workloads.thermal_3d).
A step is one complete solve from the reference start field (inittemp = 300 K, made on
the device) to ||B - A T|| <= 1e-8 ||B - A T_init||, with the element data, source
and Dirichlet set already resident in HBM.  The solution stays on the device.

Printed JSON line: value = solves/s of FP64 PCG preconditioned by the multigrid V-cycle
(in-plane semicoarsening, exact line smoother along the layer axis).  Extra fields give
the single-level forms of the same solve (exact line preconditioner, Jacobi), the
FP32-operator forms (FP64 refinement or reliable updates), the learned-warm-start form
when trained weights are present, per-kernel-class timings of profiled solves (HIP events
around each launch on the context stream), the roofline of the kernel function with
the largest total time in a profiled headline solve (all multigrid levels), and the CPU path
timed on this host: the reference's NSPCG (CG + IC(0)) compiled from its sources, and this
build's serial IC(0)-PCG restatement, both on the reference's 14-diagonal storage.
"""
import argparse
import csv
import glob
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from plask_amd import workloads  # noqa: E402
from plask_amd.fem import (KCAT_APPLY, KCAT_MG_COARSE, KCAT_MG_RESIDUAL, KCAT_MG_SMOOTH,  # noqa: E402
                           KCAT_MG_TRANSFER, KCAT_PRECOND, PF_FP32_REFINE, PF_FP32_RELIABLE, PF_FP64, PF_FP64_PREC32,
                           PRECOND_JACOBI, PRECOND_LINE, PRECOND_MG, KERNEL_NAMES, FemContext)


HBM_PEAK_GBS = 8000.0
# Full-solve iteration count of the CPU path (serial IC(0)-CG on the 14-diagonal
# storage, same inputs, same stopping rule) for 128x128x256, measured once by a
# complete CPU solve; used to scale the bounded per-iteration sample to a whole
# solve.  Iteration counts depend on the algorithm and the inputs, not the host.
CPU_FULL_ITERS = {(128, 128, 256): 161}
# The same for the reference's own NSPCG (CG + IC(0), zeta = maxerr = 1e-8; a complete run
# on 128x128x256 took 161 iterations, 77.6 s on one core, and reached ||B-AT||/||B-AT_init||
# = 9.5e-9).
CPU_REF_ITERS = {(128, 128, 256): 161}
WEIGHTS = os.path.join(ROOT, "plask_amd", "data", "warmstart_64x64x128.npz")
TRAFFIC = os.path.join(ROOT, "profiles", "traffic.json")
# relative reduction asked of each FP32 run: restarted refinement needs deep inner solves,
# the reliable-update form replaces the residual early and keeps the search direction
INNER = {PF_FP64: 0.0, PF_FP32_REFINE: 1e-5, PF_FP32_RELIABLE: 1e-2, PF_FP64_PREC32: 0.0}
# with the multigrid preconditioner each iteration reduces the residual ~10x, so a kept
# direction pays off only over a longer run
INNER_MG_RELIABLE = 1e-3


def make_ctx(p, profile=False):
    c = FemContext(0, profile=profile)
    c.set_grid(p["axes"])
    c.set_coeffs(p["k_h"], p["k_v"])
    c.set_source(p["q"])
    c.set_dirichlet(p["dir_nodes"], p["dir_vals"])
    return c


def kernel_bytes(shape, m=8):
    """Algorithmic HBM bytes per launch (FP64) of the kernels of one PCG iteration.
    apply_fused: read z, p_old, x; write p_new, q, x (48 B/node); k_xy, k_z
      (16 B/element); Dirichlet bitmask (N/8 B).
    line (partitioned, segments of m planes): read r, q, a, w, c'; write r, z
      (56 B/node); separator factors 3 x 8 B per separator per column.
    apply_plain: read x, write y (16 B/node), 16 B/element, bitmask.
    mg_residual (V-cycle, solve grid): r = b - A x: read x, b; write r (24 B/node),
      16 B/element, bitmask.
    mg_smooth (V-cycle, solve grid): damped line sweep, first (read r, q, a, w, c';
      write r, x) or last (read r_s, x, r, a, w, c'; write x): 56 B/node + separators.
    mg_transfer: restriction (read fine r, write coarse b) and prolongation (read coarse x,
      read and write fine x), averaged: 12 B per fine node + 8 B per coarse node.
    FP32 forms: half the bytes of vectors and element data (line: 16-plane segments)."""
    nx, ny, nz = shape
    N, E = nx * ny * nz, (nx - 1) * (ny - 1) * (nz - 1)
    mask = 8 * ((N + 63) // 64)
    P = (nz + m - 1) // m
    P32 = (nz + 15) // 16
    Nc = (nx // 2 + 1 if nx > 3 else nx) * (ny // 2 + 1 if ny > 3 else ny) * nz
    return {"apply_fused": 48 * N + 16 * E + mask,
            "line": 56 * N + 24 * (P - 1) * nx * ny,
            "apply_plain": 16 * N + 16 * E + mask,
            "apply_fused_fp32": 24 * N + 8 * E + mask,
            "line_fp32": 28 * N + 12 * (P32 - 1) * nx * ny,
            "mg_residual": 24 * N + 16 * E + mask,
            "mg_smooth": 56 * N + 24 * (P - 1) * nx * ny,
            "mg_transfer": 12 * N + 8 * Nc}


def per_kernel(kstats):
    """Records of pf_kernel_stats grouped by kernel function: total time, launches and
    algorithmic bytes over all levels, the coarse-level share, and the per-level rows."""
    out = {}
    total = sum(r["seconds"] for r in kstats)
    for r in kstats:
        k = out.setdefault(r["name"], dict(seconds=0.0, launches=0, bytes=0.0, levels=[]))
        k["seconds"] += r["seconds"]
        k["launches"] += r["launches"]
        k["bytes"] += r["bytes"]
        k["levels"].append(dict(level=r["level"], launches=r["launches"],
                                avg_us=round(1e6 * r["seconds"] / max(r["launches"], 1), 2),
                                bytes_per_launch=int(r["bytes"] / max(r["launches"], 1)),
                                gbs=round(r["bytes"] / max(r["seconds"], 1e-30) / 1e9, 1)))
    for k in out.values():
        k["levels"].sort(key=lambda r: r["level"])
        k["share"] = k["seconds"] / max(total, 1e-30)
    coarse = sum(r["seconds"] for r in kstats if r["level"] > 0)
    return dict(kernels=out, total_s=total, coarse_share=coarse / max(total, 1e-30))


def _kernel_class(name):
    """Kernel class of a dispatch (the names pf_kernel_stats reports), or None for kernels
    outside the PCG iterations (set-up, initial residual, norms, copies)."""
    for key, k in (("k_mg_smooth", 2), ("k_mg_restrict", 3), ("k_mg_prolong_add", 4), ("k_cg_part", 5),
                   ("k_cg_update", 6)):
        if key in name:
            return KERNEL_NAMES[k]
    m = re.search(r"k_apply(?:3d|2d)<[^,]+,\s*(\d)", name)
    if m:
        return {0: KERNEL_NAMES[0], 2: KERNEL_NAMES[1], 3: KERNEL_NAMES[1]}.get(int(m.group(1)))
    return None


def _pmc_dispatches(folder, counter):
    """(kernel name, grid size, value) of every dispatch in a rocprofv3 --pmc csv output
    folder, in dispatch order."""
    rows = []
    for f in sorted(glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True)):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") == counter:
                    rows.append((int(r.get("Dispatch_Id", len(rows))), r["Kernel_Name"], int(r["Grid_Size"]),
                                 float(r["Counter_Value"])))
    rows.sort()
    return [r[1:] for r in rows]


def traffic_from(fetch_dir, write_dir, size, copy_bytes=1 << 30):
    """HBM bytes per launch of each kernel class from two rocprofv3 --pmc runs of
    `bench.py --counters` (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each counter
    scaled by what it reads for the k_stream_copy launches of known size.  Dispatches
    that exit at once after the stop (below 1 % of the largest value of the same kernel
    and grid) are left out; a class's figure is the mean over its launches on every
    multigrid level, as the roofline's achieved bytes are."""
    out = {"workload": "thermal3d_%dx%dx%d" % size,
           "unit": "HBM bytes per launch, mean over the launches of all multigrid levels",
           "source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE -- python bench.py --counters "
                     "(one FP64 multigrid solve); python bench.py --traffic-from FETCH_DIR WRITE_DIR",
           "calibration": {"kernel": "k_stream_copy, %d bytes read + %d written" % (copy_bytes, copy_bytes)},
           "kernels": {}, "per_grid": {}}
    per = {}
    for counter, folder in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        rows = _pmc_dispatches(folder, counter)
        copy = [v for n, g, v in rows if "k_stream_copy" in n]
        if not copy:
            raise SystemExit("traffic_from: no k_stream_copy dispatch in %s" % folder)
        factor = copy_bytes / 1024.0 / float(np.median(copy))
        out["calibration"][counter + "_KiB"] = float(np.median(copy))
        out["calibration"][counter.split("_")[0].lower() + "_factor"] = round(factor, 4)
        groups = {}
        for n, g, v in rows:
            c = _kernel_class(n)
            if c is not None:
                groups.setdefault((c, n, g), []).append(v)
        for (c, n, g), vals in groups.items():
            top = max(vals)
            kept = [v for v in vals if v >= 0.01 * top] if top > 0 else vals
            rec = per.setdefault((c, n, g), {"launches": len(kept)})
            rec[counter] = 1024.0 * factor * float(np.mean(kept))
    for (c, n, g), rec in sorted(per.items(), key=lambda kv: (kv[0][0], -kv[0][2])):
        b = rec.get("FETCH_SIZE", 0.0) + rec.get("WRITE_SIZE", 0.0)
        out["per_grid"].setdefault(c, []).append({"kernel": n, "grid": g, "launches": rec["launches"],
                                                  "read_bytes": int(rec.get("FETCH_SIZE", 0.0)),
                                                  "written_bytes": int(rec.get("WRITE_SIZE", 0.0)),
                                                  "bytes": int(b)})
    for c, lst in out["per_grid"].items():
        n = sum(r["launches"] for r in lst)
        out["kernels"][c] = int(sum(r["launches"] * r["bytes"] for r in lst) / max(n, 1))
    return out


def roofline_line(by_kernel, size, copy_gbs):
    """Roofline of the kernel function with the largest total time of the profiled FP64
    multigrid solve: achieved = its algorithmic bytes summed over all its launches (every
    level) / its summed duration = bytes per launch / average launch duration.  traffic:
    the calibrated hardware counters of that kernel, from the rocprofv3 --pmc summary
    committed in profiles/traffic.json (measured in a separate counter run, not here)."""
    ks = by_kernel["kernels"]
    top = max(ks, key=lambda k: ks[k]["seconds"])
    k = ks[top]
    avg_us = 1e6 * k["seconds"] / k["launches"]
    bpl = k["bytes"] / k["launches"]
    achieved = bpl / (1e-6 * avg_us) / 1e9
    traffic = None
    src = None
    if os.path.exists(TRAFFIC):
        with open(TRAFFIC) as fh:
            tr = json.load(fh)
        if tr.get("workload") == "thermal3d_%dx%dx%d" % size and top in tr.get("kernels", {}):
            traffic = tr["kernels"][top]
            src = "profiles/traffic.json (%s)" % tr.get("source", "rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE")
    return {"bound": "hbm", "kernel": top, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": src,
            "bytes_per_launch": int(bpl), "avg_launch_us": round(avg_us, 2), "launches": k["launches"],
            "time_share_of_kernel_time": round(k["share"], 3),
            "coarse_level_share_of_kernel_time": round(by_kernel["coarse_share"], 3),
            "per_level": k["levels"],
            "kernel_shares": {n: round(v["share"], 3) for n, v in sorted(ks.items(), key=lambda kv: -kv[1]["seconds"])},
            "stream_copy_gbs": round(copy_gbs, 1)}


def cpu_reference(A, seconds, full_iters):
    """The reference's NSPCG (CG + IC(0), its default; compiled from the reference sources
    into oracle/_ref) on the assembled 14-diagonal system, one core.  NSPCG factorises on
    every call, so two calls stopped after k1 and k2 iterations give the set-up time and
    the time per iteration; the complete solve is run instead when its estimate fits in
    `seconds`.  None when the library was not built."""
    import oracle as O
    if O.ref_lib() is None:
        return None
    x0 = np.full(A.rank, 300.0)
    k1, k2 = 2, 12
    t0 = time.perf_counter()
    O.nspcg_solve(A, x0, "ic", 1e-30, k1)
    t1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.nspcg_solve(A, x0, "ic", 1e-30, k2)
    t2 = time.perf_counter() - t0
    t_it = max(t2 - t1, 1e-9) / (k2 - k1)
    setup = max(t1 - k1 * t_it, 0.0)
    res = {"setup_s": setup, "s_per_iter": t_it}
    if full_iters and setup + full_iters * t_it > seconds:
        res.update(iters=full_iters, full_solve=False, solve_s=setup + full_iters * t_it)
    else:
        t0 = time.perf_counter()
        u, it, zo, ier, _, _ = O.nspcg_solve(A, x0, "ic", 1e-8, 5000)
        t = time.perf_counter() - t0
        r0 = np.linalg.norm(A.B - A.mult(x0))
        res.update(iters=it, full_solve=True, solve_s=t, ier=ier,
                   rel_resid=float(np.linalg.norm(A.B - A.mult(u)) / r0))
    return res


def cpu_baseline(p, seconds, full_iters):
    """CPU path on this host, one core: assembly of the 14-diagonal system (the
    reference's SparseBandMatrix storage), the IC(0) factor, and IC(0)-PCG to the same
    stopping rule as the GPU (oracle restatement, not NSPCG itself).  The whole solve runs
    when its estimate fits in `seconds`; otherwise a few iterations are timed and scaled
    to the iteration count of a complete solve (full_iters)."""
    import oracle as O
    g = O.Grid(p["axes"])
    t0 = time.perf_counter()
    A = O.assemble(g, 0, p["k_h"], p["k_v"], p["q"], O.ST_DIA, p["dir_nodes"], p["dir_vals"])
    t_asm = time.perf_counter() - t0
    x0 = np.full(g.nnodes, 300.0)
    t0 = time.perf_counter()
    O.ic0_factor(A)
    t_fact = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, st, it, rel = O.dia_pcg(A, x0, precond=2, rtol=1e-30, maxit=3)
    t_iter = max(time.perf_counter() - t0 - t_fact, 1e-9) / max(it, 1)
    est = t_fact + (full_iters or 0) * t_iter
    res = {"assemble_s": t_asm, "ic0_factor_s": t_fact, "s_per_iter": t_iter}
    if full_iters and est > seconds:
        res.update(iters=full_iters, full_solve=False, solve_s=t_asm + t_fact + full_iters * t_iter)
    else:
        t0 = time.perf_counter()
        _, st, it, rel = O.dia_pcg(A, x0, precond=2, rtol=1e-8, maxit=100000)
        t_run = time.perf_counter() - t0
        res.update(iters=it, full_solve=True, rel_resid=rel, solve_s=t_asm + t_run,
                   s_per_iter=(t_run - t_fact) / max(it, 1))
    res["value"] = 1.0 / res["solve_s"]
    return res, A


def timed(fn, steps, warmup):
    for _ in range(warmup):
        fn()
    t0 = time.perf_counter()
    for _ in range(steps):
        st = fn()
    return (time.perf_counter() - t0) / steps, st


def learned(ctx, p, size, steps, warmup, T64):
    """FP32 + FP64-refinement multigrid PCG from the warm-start network guess (prediction on the
    64x64x128 grid of the same stack, trilinear onto the solve grid); the network
    forward pass and the interpolation are inside the timed step."""
    from plask_amd import warmstart
    from plask_amd.train_warmstart import load_weights
    params, nshape = load_weights(WEIGHTS)
    net = warmstart.WarmNet(nshape)
    net.set_params(params)
    ws = warmstart.WarmStart(net, nshape)
    pc = workloads.thermal_3d(seed=2, n=nshape)
    feat = warmstart.features(pc)          # inputs of the network (host, not timed)

    def step():
        ws.start_on_device(ctx, pc, feat)
        return ctx.solve_warm(precision=PF_FP32_REFINE, rtol=1e-8, inner_rtol=INNER[PF_FP32_REFINE],
                              init_value=300.0, checked=True, precond=PRECOND_MG)

    t, st = timed(step, steps, warmup)
    T = ctx.solution()
    t0 = time.perf_counter()
    ws.start_on_device(ctx, pc, feat)
    t_net = time.perf_counter() - t0
    t0 = time.perf_counter()
    feat = warmstart.features(pc)
    t_feat = time.perf_counter() - t0
    return dict(solves_per_s=1.0 / t, ms=1e3 * t, inner_iters=st.iters, outer=st.outer, rel_resid=st.rel_resid,
                fell_back_to_uniform_start=bool(st.start),
                network_and_prolong_ms=1e3 * t_net, host_features_ms_untimed=1e3 * t_feat, max_rel_diff_vs_fp64_mg=float(np.abs(T - T64).max() / np.abs(T64).max()))


def config1(steps, with_cpu):
    """2D cylindrical stack, 200 (r) x 400 (z) nodes (plask_amd.workloads.thermal_2d_cyl, seed 1):
    GPU FP64 multigrid-PCG solve to 1e-8 (and the line-PCG form) vs the reference method on
    the CPU (assembly + band Cholesky, kd = 201, serial restatement), with the max relative
    difference."""
    p = workloads.thermal_2d_cyl(seed=1)
    c = FemContext(0)
    c.set_grid(p["axes"], cylindrical=True)
    c.set_coeffs(p["k_h"], p["k_v"])
    c.set_source(p["q"])
    c.set_dirichlet(p["dir_nodes"], p["dir_vals"])
    t_line, st_line = timed(lambda: c.solve_uniform(300.0, rtol=1e-8, precond=PRECOND_LINE), steps, 1)
    t, st = timed(lambda: c.solve_uniform(300.0, rtol=1e-8, precond=PRECOND_MG), steps, 1)
    T = c.solution()
    out = dict(workload="thermal2d_cyl_200x400", gpu_ms=1e3 * t, method="fp64 pcg, multigrid V-cycle",
               iters=st.iters, rel_resid=st.rel_resid, gpu_ms_line_pcg=1e3 * t_line, iters_line_pcg=st_line.iters)
    if with_cpu:
        import oracle as O
        g = O.Grid(p["axes"], cyl=True)
        t0 = time.perf_counter()
        A = O.assemble(g, 0, p["k_h"], p["k_v"], p["q"], O.ST_BAND, p["dir_nodes"], p["dir_vals"])
        t_asm = time.perf_counter() - t0
        # every LAPACK this host offers (MKL when present, SciPy's OpenBLAS), each with its
        # thread pool warmed up, median of 3; the fastest is the baseline
        lap = {}
        vendors = [O.lapack_vendor()]
        if vendors[0][0] != "openblas":
            vendors.append(("openblas", None))
        for vendor in vendors:
            O.lapack_band_cholesky_solve(A, vendor=vendor)
            ts = []
            for _ in range(3):
                t0 = time.perf_counter()
                Tl, vname, nthreads = O.lapack_band_cholesky_solve(A, vendor=vendor)
                ts.append(time.perf_counter() - t0)
            lap[vname] = (float(np.median(ts)), nthreads, Tl)
        vname = min(lap, key=lambda k: lap[k][0])
        t_lapack, nthreads, Tl = lap[vname]
        t0 = time.perf_counter()
        Tr = O.band_cholesky_solve(A)
        t_text = time.perf_counter() - t0
        tc = t_asm + t_lapack
        out.update(cpu_ms=1e3 * tc, cpu_assemble_ms=1e3 * t_asm, cpu_lapack_ms=1e3 * t_lapack,
                   cpu_lapack="%s dpbtrf/dpbtrs, %d threads (median of 3; fastest of %s)" % (
                       vname, nthreads, ", ".join("%s %.1f ms" % (k, 1e3 * v[0]) for k, v in lap.items())),
                   cpu_assembly_threads=1, kd=int(A.kd),
                   cpu_textbook_band_cholesky_ms_1core=1e3 * t_text,
                   max_rel_diff_vs_lapack=float(np.abs(T - Tl).max() / np.abs(Tl).max()),
                   max_rel_diff_vs_band_cholesky=float(np.abs(T - Tr).max() / np.abs(Tr).max()),
                   speedup_vs_cpu=tc / t)
    c.close()
    return out


def config3(cpu_seconds):
    """256 uneven 2D cylindrical Shockley instances (plask_amd.workloads.shockley_cyl_batch,
    seeds 3000+i; 50x100 to 400x800 nodes): one persistent batched launch per solve
    (multigrid-preconditioned CG inside each workgroup, Picard loop <= 30 on the device),
    autotuned workgroup size / count.  CPU reference path (assembly + band Cholesky per
    Picard loop, serial oracle) on a bounded sample: one loop timed on a few instances, a
    cost model t = c N kd^2 (band Cholesky, kd = n_r + 1) fitted to them and summed over
    the loops every instance ran on the GPU."""
    import oracle as O
    from plask_amd.batch import ShockleyBatch
    insts = workloads.shockley_cyl_batch(n_inst=256)
    nr = np.array([len(p["axes"][0]) for p in insts])
    nz = np.array([len(p["axes"][1]) for p in insts])
    b = ShockleyBatch(insts)
    blk, nwg, t_tune = b.autotune(maxloops=30)
    _, stats, t = b.solve(want_v=False, maxloops=30)
    loops = np.array([s["loops"] for s in stats])
    its = np.array([s["iters"] for s in stats])
    status = np.array([s["status"] for s in stats])
    b.close()
    # CPU sample: one Picard loop (assembly + band Cholesky) on instances of increasing size
    order = np.argsort(nr * nz)
    cpu = []
    spent = 0.0
    for i in order[::32]:
        if spent > cpu_seconds:
            break
        p = insts[i]
        r, z = p["axes"]
        g = O.Grid([r, z], cyl=True)
        E0 = len(r) - 1
        elems = np.arange(p["jlo"] * E0, p["jhi"] * E0)
        bot = [[e % E0 + p["jlo"] * len(r), e % E0 + 1 + p["jlo"] * len(r)] for e in elems]
        top = [[e % E0 + p["jhi"] * len(r), e % E0 + 1 + p["jhi"] * len(r)] for e in elems]
        J = dict(elems=elems, bottom=bot, top=top, height=z[p["jhi"]] - z[p["jlo"]], beta=p["beta"], js=p["js"])
        t0 = time.perf_counter()
        O.electrical_compute(g, p["sh"].copy(), p["sv"].copy(), p["dir_nodes"], p["dir_vals"], [J], 1)
        dt = time.perf_counter() - t0
        spent += dt
        cpu.append((int(i), float(dt)))
        print("config 3 CPU sample: instance %d (%d x %d nodes): %.3f s per loop" % (i, nr[i], nz[i], dt),
              file=sys.stderr, flush=True)
    work = lambda i: float(nr[i] * nz[i]) * float(nr[i] + 1) ** 2
    c = sum(dt for _, dt in cpu) / sum(work(i) for i, _ in cpu)
    cpu_total = c * sum(loops[i] * work(i) for i in range(len(insts)))
    # algorithmic bytes (FP64, each value read or written once per pass): per CG iteration
    # the marched operator apply (p, two conductivities, q: 32 B/node), the x / r update
    # (48), p = z + beta p and (r, z) (40) and the V-cycle, 232 B per node of every level
    # (two line sweeps 56 + 64, two marched residuals 2 x 40, restriction 12, prolongation
    # 20) over levels of 2 N nodes in all (axis 0 halved per level); per Picard loop the line
    # factors and coarse conductivities of all levels, the junction update and the currents
    # (~150 B/node)
    nodes = (nr * nz).astype(np.float64)
    bytes_alg = float(np.sum(nodes * (its * (32 + 48 + 40 + 232 * 2) + loops * 150)))
    roofline = {"bound": "hbm", "kernel": "k_batch_shockley2d (whole batch)", "achieved": round(bytes_alg / t / 1e9, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(bytes_alg / t / 1e9 / HBM_PEAK_GBS, 4),
                "traffic": None, "bytes": bytes_alg,
                "note": "one workgroup per instance: each instance's passes are chains of dependent loads on one CU; "
                        "the time is the slowest instances' (largest, most CG iterations), not the total bytes"}
    return dict(workload="shockley_cyl_batch_256", instances=len(insts), nodes_total=int((nr * nz).sum()),
                roofline=roofline,
                gpu_s=t, block=blk, workgroups=nwg, autotune_s=t_tune, picard_loops_mean=float(loops.mean()),
                picard_loops_max=int(loops.max()), cg_iters_total=int(its.sum()),
                converged=int((status == 0).sum()), stopped_at_maxloops=int((status == 2).sum()),
                cpu_sample=[dict(instance=i, seconds_per_loop=dt) for i, dt in cpu],
                cpu_estimate_s=cpu_total, cpu_cores=1, speedup_vs_cpu=cpu_total / t)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", default="128,128,256")
    ap.add_argument("--cpu-seconds", type=float, default=45.0,
                    help="CPU baseline: run the complete CPU solve when its estimate fits in this many seconds")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-graphs", action="store_true", help="plain kernel launches instead of hipGraph replay")
    ap.add_argument("--config3", action="store_true",
                    help="the batched 2D Shockley workload (256 instances) instead of the 3D benchmark")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the timed headline solves (for kernel-trace runs): no comparison forms, "
                         "profiled solves, CPU baseline or 2D case")
    ap.add_argument("--counters", action="store_true",
                    help="short run for hardware-counter profiles: stream-copy calibration, one FP64 "
                         "multigrid solve")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="summarise two rocprofv3 --pmc output folders of --counters runs (FETCH_SIZE, "
                         "WRITE_SIZE) into HBM bytes per launch per kernel class (JSON on stdout)")
    args = ap.parse_args()
    if args.traffic_from:
        size = tuple(int(v) for v in args.size.split(","))
        print(json.dumps(traffic_from(args.traffic_from[0], args.traffic_from[1], size), indent=1))
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the solve path has no multi-GPU "
              "decomposition); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.config3:
        print(json.dumps(config3(args.cpu_seconds)))
        return
    size = tuple(int(v) for v in args.size.split(","))
    p = workloads.thermal_3d(seed=2, n=size)
    N = int(np.prod(size))
    E = int(np.prod([s - 1 for s in size]))
    ctx = make_ctx(p)
    if args.no_graphs:
        ctx.set_graphs(False)
    rtol = 1e-8
    if args.counters:
        cb = 1 << 30
        t = ctx.stream_copy(cb, reps=3)
        print("stream copy %d bytes: %.1f us, %.0f GB/s" % (cb, 1e6 * t, 2 * cb / t / 1e9), file=sys.stderr)
        st = ctx.solve_uniform(300.0, rtol=rtol, precond=PRECOND_MG)
        print(json.dumps({"counters_run": True, "copy_bytes": cb, "iters": st.iters}))
        return

    def step(prec, pc, inner=None):
        return ctx.solve_uniform(300.0, precision=prec, rtol=rtol, inner_rtol=inner, precond=pc)

    def form(name, prec, pc, inner=None, steps=args.steps, warmup=args.warmup):
        t, st = timed(lambda: step(prec, pc, inner), steps, warmup)
        T = ctx.solution()
        f = dict(method=name, solves_per_s=1.0 / t, ms=1e3 * t, iters=st.iters, outer=st.outer,
                 rel_resid=st.rel_resid)
        if prec in (PF_FP32_REFINE, PF_FP32_RELIABLE):
            f["inner_rtol"] = inner
        if T64 is not None:
            f["max_rel_diff_vs_fp64_mg"] = float(np.abs(T - T64).max() / np.abs(T64).max())
            f["speedup_vs_fp64_mg"] = fp64["ms"] / f["ms"]
        print("%s: %.2f solves/s, %.2f ms/solve, %d iterations (%d outer), rel resid %.2e" %
              (name, f["solves_per_s"], f["ms"], st.iters, st.outer, st.rel_resid), file=sys.stderr, flush=True)
        return f, T

    # headline: FP64 PCG preconditioned by the multigrid V-cycle
    T64 = None
    fp64, T64 = form("fp64 pcg, multigrid V-cycle (in-plane semicoarsening, exact line smoother)", PF_FP64,
                     PRECOND_MG)
    fp64.pop("inner_rtol", None)
    if args.headline_only:
        print(json.dumps({"metric": "3D thermal solves/s to 1e-8 rel. residual (4.2M DOF)", "value":
                          round(fp64["solves_per_s"], 4), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
                          "warmup": args.warmup, "ms_per_step": round(fp64["ms"], 3), "headline_only": True,
                          "config": {"workload": "thermal3d_%dx%dx%d" % size, "method": fp64["method"],
                                     "iters": fp64["iters"], "rel_resid": fp64["rel_resid"]}}))
        return
    # the single-level forms of the same solve, for comparison
    line64, _ = form("fp64 pcg, exact line preconditioner", PF_FP64, PRECOND_LINE)
    jac, _ = form("fp64 pcg, Jacobi preconditioner", PF_FP64, PRECOND_JACOBI, steps=1, warmup=1)
    # the pair of the benchmark description: FP64 Jacobi-PCG vs FP32 matrix-free apply + FP64 refinement
    jac32, _ = form("fp32 operator + fp64 refinement, Jacobi preconditioner", PF_FP32_REFINE, PRECOND_JACOBI,
                    1e-4, steps=1, warmup=1)
    # compute(1) equivalent: a coefficient change (device-evaluated conductivity law, the
    # setMatrix part) followed by the solve -- the set-up after the change (load vector,
    # diagonal, line factors and coarse operators of all levels in both precisions) is timed
    ctx.set_powerlaw(p["k_h"], p["k_v"])

    def compute1():
        ctx.coeffs_update()
        return ctx.solve_uniform(300.0, rtol=rtol, precond=PRECOND_MG)
    tc1, stc1 = timed(compute1, args.steps, args.warmup)
    comp1 = dict(method="coefficient update + set-up + fp64 multigrid pcg solve", ms=1e3 * tc1,
                 solves_per_s=1.0 / tc1, iters=stc1.iters, setup_ms=1e3 * tc1 - fp64["ms"])
    print("compute(1) equivalent: %.2f ms" % comp1["ms"], file=sys.stderr, flush=True)
    mixed = {}
    warm = None
    if not args.no_mixed:
        mixed["mg_fp32_refine"], _ = form("fp32 operator + fp64 refinement, multigrid", PF_FP32_REFINE, PRECOND_MG,
                                          INNER[PF_FP32_REFINE])
        mixed["mg_fp32_reliable"], _ = form("fp32 operator, reliable updates, multigrid", PF_FP32_RELIABLE,
                                            PRECOND_MG, INNER_MG_RELIABLE)
        mixed["mg_fp64_cg_fp32_precond"], _ = form("fp64 cg, multigrid preconditioner in fp32", PF_FP64_PREC32,
                                                   PRECOND_MG)
        mixed["line_fp32_reliable"], _ = form("fp32 operator, reliable updates, line preconditioner",
                                              PF_FP32_RELIABLE, PRECOND_LINE, INNER[PF_FP32_RELIABLE])
        if os.path.exists(WEIGHTS):
            warm = learned(ctx, p, size, args.steps, args.warmup, T64)
            warm["speedup_vs_fp64_mg"] = fp64["ms"] / warm["ms"]
            print("learned start + fp32 refinement, multigrid: %.2f solves/s, %.2f ms/solve, %d inner its" %
                  (warm["solves_per_s"], warm["ms"], warm["inner_iters"]), file=sys.stderr, flush=True)
    best = min(mixed.values(), key=lambda f: f["ms"]) if mixed else None

    # roofline: a profiled FP64 multigrid solve -- HIP events around every launch on the
    # context stream (no graph replay), one record per (kernel, multigrid level); launches
    # after the stop are not counted.  The kernel of the roofline is the kernel function with
    # the largest total time over all levels (how a kernel trace ranks them).
    def profiled(prec, pc, inner=None):
        ctx.set_profile(True)
        st = ctx.solve_uniform(300.0, precision=prec, rtol=rtol, inner_rtol=inner, precond=pc)
        ctx.set_profile(False)
        t, n = ctx.kernel_times(6)
        return t, n, st.iters, ctx.kernel_stats()
    kb = kernel_bytes(size)
    kern = {}
    t, n, its, kstats = profiled(PF_FP64, PRECOND_MG)
    names = {KCAT_APPLY: "apply_fused", KCAT_MG_RESIDUAL: "mg_residual", KCAT_MG_SMOOTH: "mg_smooth",
             KCAT_MG_TRANSFER: "mg_transfer"}
    for cat, name in names.items():
        kern[name] = dict(avg_us=1e6 * t[cat], launches=int(n[cat]), bytes=kb[name])
    kern["mg_vcycle"] = dict(avg_us=1e6 * t[KCAT_PRECOND], launches=int(n[KCAT_PRECOND]))
    kern["mg_coarse_levels"] = dict(avg_us=1e6 * t[KCAT_MG_COARSE], launches=int(n[KCAT_MG_COARSE]))
    by_kernel = per_kernel(kstats)
    t, n, _, _ = profiled(PF_FP32_REFINE, PRECOND_MG, INNER[PF_FP32_REFINE])
    kern["apply_fused_fp32"] = dict(avg_us=1e6 * t[KCAT_APPLY], launches=int(n[KCAT_APPLY]),
                                    bytes=kb["apply_fused_fp32"])
    kern["mg_vcycle_fp32"] = dict(avg_us=1e6 * t[KCAT_PRECOND], launches=int(n[KCAT_PRECOND]))
    t, n, _, _ = profiled(PF_FP64, PRECOND_LINE)
    kern["line"] = dict(avg_us=1e6 * t[KCAT_PRECOND], launches=int(n[KCAT_PRECOND]), bytes=kb["line"])
    kern["apply_plain"] = dict(avg_us=1e6 * ctx.time_apply(PF_FP64, reps=50), bytes=kb["apply_plain"])
    for k in kern.values():
        if "bytes" in k and k["avg_us"] > 0:
            k["gbs"] = round(k["bytes"] / (1e-6 * k["avg_us"]) / 1e9, 1)
        k["avg_us"] = round(k["avg_us"], 2)
    t_copy = ctx.stream_copy(1 << 30, reps=5)
    copy_gbs = 2 * (1 << 30) / t_copy / 1e9
    roofline = roofline_line(by_kernel, size, copy_gbs)

    cpu = cpu_port = None
    if not args.no_cpu:
        full = CPU_FULL_ITERS.get(size)
        c, A_dia = cpu_baseline(p, args.cpu_seconds, full)
        how = ("the complete solve timed (%d iterations)" % c["iters"] if c["full_solve"] else
               "3 iterations timed (%.3f s/iter) and scaled to the %d iterations of a complete solve" %
               (c["s_per_iter"], c["iters"]))
        cpu_port = {"value": c["value"], "unit": "solves/s", "cores": 1, "kind": "port",
               "sample": "serial IC(0)-PCG (this build's restatement of the reference's default incomplete-Cholesky "
                         "preconditioner family, not NSPCG) on the 14-diagonal SparseBandMatrix storage of the same "
                         "system, same stopping rule: assembly %.3f s + IC(0) factor %.3f s + iterations; %s"
                         % (c["assemble_s"], c["ic0_factor_s"], how),
               "assemble_s": round(c["assemble_s"], 3), "ic0_factor_s": round(c["ic0_factor_s"], 3),
               "s_per_iter": c["s_per_iter"], "iters": c["iters"], "full_solve": c["full_solve"]}
        cpu = cpu_port
        r = cpu_reference(A_dia, args.cpu_seconds, CPU_REF_ITERS.get(size))
        del A_dia
        if r is not None:
            solve = c["assemble_s"] + r["solve_s"]
            how = ("the complete solve timed (%d iterations, rel. residual %.2e)" % (r["iters"], r["rel_resid"])
                   if r["full_solve"] else
                   "runs stopped after 2 and 12 iterations timed (set-up %.3f s, %.3f s/iter) and scaled to the "
                   "%d iterations of a complete run at zeta = 1e-8" % (r["setup_s"], r["s_per_iter"], r["iters"]))
            cpu = {"value": 1.0 / solve, "unit": "solves/s", "cores": 1, "kind": "reference",
                   "sample": "the reference's NSPCG (CG + IC(0), its default solver; nspcg.c compiled from the "
                             "reference sources) on the 14-diagonal SparseBandMatrix system of the same workload: "
                             "assembly %.3f s (this build's restatement of setMatrix) + %s" % (c["assemble_s"], how),
                   "assemble_s": round(c["assemble_s"], 3), "setup_s": round(r["setup_s"], 3),
                   "s_per_iter": r["s_per_iter"], "iters": r["iters"], "full_solve": r["full_solve"]}

    c1 = config1(args.steps, not args.no_cpu)
    print("config 1 (2D cyl): %.2f ms GPU" % c1["gpu_ms"], file=sys.stderr, flush=True)

    line = {
        "metric": "3D thermal solves/s to 1e-8 rel. residual (4.2M DOF); operator-apply HBM GB/s",
        "value": round(fp64["solves_per_s"], 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fp64["ms"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded layered stack, SplitMix64)",
        "config": {"workload": "thermal3d_%dx%dx%d" % size, "method": fp64["method"],
                   "dof": N, "elements": E, "rtol": rtol,
                   "residual": "||B-AT||/||B-AT_init||, T_init = 300 K",
                   "iters": fp64["iters"], "rel_resid": fp64["rel_resid"]},
        "fp64_line_pcg": line64,
        "fp64_jacobi_pcg": jac,
        "config2_pair": {"fp64_jacobi_pcg": jac, "fp32_apply_fp64_refinement_jacobi": jac32,
                         "speedup_fp32_refine_vs_fp64": jac["ms"] / jac32["ms"]},
        "compute1_equivalent": comp1,
        "mixed_precision": best,
        "mixed_precision_forms": mixed,
        "learned_warm_start": warm,
        "config1_2d_cyl": c1,
        "kernels": kern,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_port": cpu_port if cpu is not cpu_port else None,
    }
    if cpu and cpu.get("value"):
        line["speedup_vs_cpu"] = round(fp64["solves_per_s"] / cpu["value"], 1)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
