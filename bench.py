#!/usr/bin/env python3
"""Benchmark: 3D steady-state thermal FEM solves/s to 1e-8 relative residual on one MI355X.

Workload: 128 x 128 x 256 nodes (4,194,304 DOF), 8-node trilinear hexahedra, 40 seeded
anisotropic layers (kxy, kz log-uniform in [1,150] W/m/K), cylindrical 1e16 W/m^3
source in the 0.1 um active layer, Dirichlet 300 K at the bottom (plask_amd.workloads.thermal_3d).
A step is one complete solve from the reference start field (inittemp = 300 K, made on
the device) to ||B - A T|| <= 1e-8 ||B - A T_init||, with the element data, source
and Dirichlet set already resident in HBM.  The solution stays on the device.

Printed JSON line: value = solves/s of the plain FP64 Jacobi-PCG; extra fields give the
FP32-operator + FP64-refinement form, the roofline of the dominant kernel (HIP events
around each launch of an extra profiled solve) and the CPU reference path timed on
this host (serial IC(0)-CG on the reference's 14-diagonal storage, bounded sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from plask_amd import workloads  # noqa: E402
from plask_amd.fem import PF_FP32_REFINE, PF_FP64, PRECOND_JACOBI, PRECOND_LINE, FemContext  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Full-solve iteration counts of the CPU reference path (serial IC(0)-CG and Jacobi-CG
# on the 14-diagonal storage, same inputs, same stopping rule) for 128x128x256, used to
# extrapolate the bounded per-iteration sample to a whole solve.  Iteration counts are a
# property of the algorithm and the inputs, not of the host.
CPU_FULL_ITERS = {"ic0": 161, "jacobi": None}


def make_ctx(p, profile=False):
    c = FemContext(0, profile=profile)
    c.set_grid(p["axes"])
    c.set_coeffs(p["k_h"], p["k_v"])
    c.set_source(p["q"])
    c.set_dirichlet(p["dir_nodes"], p["dir_vals"])
    return c


def cpu_baseline(p, seconds, full_iters):
    """Serial oracle: assemble the 14-diagonal system, time IC(0)-CG iterations."""
    import oracle as O
    g = O.Grid(p["axes"])
    t0 = time.perf_counter()
    A = O.assemble(g, 0, p["k_h"], p["k_v"], p["q"], O.ST_DIA, p["dir_nodes"], p["dir_vals"])
    t_asm = time.perf_counter() - t0
    x0 = np.full(g.nnodes, 300.0)
    # probe one iteration to size the sample
    t0 = time.perf_counter()
    O.dia_pcg(A, x0, precond=2, rtol=1e-30, maxit=1)
    t_one = time.perf_counter() - t0
    n = max(2, int(seconds / max(t_one, 1e-3)))
    t0 = time.perf_counter()
    _, st, it, rel = O.dia_pcg(A, x0, precond=2, rtol=1e-30, maxit=n)
    t_run = time.perf_counter() - t0
    t_iter = t_run / max(it, 1)
    res = {"assemble_s": t_asm, "iters_sampled": it, "s_per_iter": t_iter}
    if full_iters:
        res["solve_s"] = t_asm + full_iters * t_iter
        res["value"] = 1.0 / res["solve_s"]
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", default="128,128,256")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the solve path has no multi-GPU "
              "decomposition); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    size = tuple(int(v) for v in args.size.split(","))
    p = workloads.thermal_3d(seed=2, n=size)
    N = int(np.prod(size))
    E = int(np.prod([s - 1 for s in size]))
    ctx = make_ctx(p)
    rtol = 1e-8

    def step(prec, pc=PRECOND_LINE):
        return ctx.solve_uniform(300.0, precision=prec, rtol=rtol, inner_rtol=1e-4 if prec else 0.0, precond=pc)

    for _ in range(args.warmup):
        st = step(PF_FP64)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = step(PF_FP64)
    t = time.perf_counter() - t0
    fp64 = dict(solves_per_s=args.steps / t, ms=1e3 * t / args.steps, iters=st.iters, rel_resid=st.rel_resid)
    print("fp64: %.2f solves/s, %.1f ms/solve, %d iterations, rel resid %.2e" %
          (fp64["solves_per_s"], fp64["ms"], st.iters, st.rel_resid), file=sys.stderr, flush=True)
    T64 = ctx.solution()

    # plain Jacobi-PCG for reference (one untimed-warm solve + timed)
    step(PF_FP64, PRECOND_JACOBI)
    t0 = time.perf_counter()
    stj = step(PF_FP64, PRECOND_JACOBI)
    tj = time.perf_counter() - t0
    jac = dict(solves_per_s=1.0 / tj, ms=1e3 * tj, iters=stj.iters, rel_resid=stj.rel_resid)
    print("fp64 jacobi: %.2f solves/s, %d iterations" % (jac["solves_per_s"], stj.iters), file=sys.stderr, flush=True)

    mixed = None
    if not args.no_mixed:
        for _ in range(args.warmup):
            st = step(PF_FP32_REFINE)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st = step(PF_FP32_REFINE)
        t = time.perf_counter() - t0
        T32 = ctx.solution()
        mixed = dict(solves_per_s=args.steps / t, ms=1e3 * t / args.steps, inner_iters=st.iters, outer=st.outer,
                     rel_resid=st.rel_resid, max_rel_diff_vs_fp64=float(np.abs(T32 - T64).max() / np.abs(T64).max()))
        print("fp32+refine: %.2f solves/s, %.1f ms/solve, %d inner its, %d outer, rel resid %.2e" %
              (mixed["solves_per_s"], mixed["ms"], st.iters, st.outer, st.rel_resid), file=sys.stderr, flush=True)

    # roofline: one profiled solve (HIP events around each kernel launch on the context stream)
    ctx.set_profile(True)
    ctx.solve_uniform(300.0, precision=PF_FP64, rtol=rtol)
    ctx.set_profile(False)
    (t_apply, t_upd), (n_apply, n_upd) = ctx.kernel_times()
    mask_bytes = 8 * ((N + 63) // 64)
    bytes_apply = 32 * N + 16 * E + mask_bytes      # z, p_old read; p_new, q write; kxy, kz; mask
    bytes_upd = 104 * N  # fwd: x,p,r,q,a,w read, x,r,z write; bwd: z,c',r read, z write
    t_plain = ctx.time_apply(PF_FP64, reps=50)
    bytes_plain = 16 * N + 16 * E + mask_bytes      # x read, y write, kxy, kz, mask
    kern = {
        "apply_fused": dict(avg_s=t_apply, launches=int(n_apply), bytes=bytes_apply, gbs=bytes_apply / t_apply / 1e9),
        "cg_line": dict(avg_s=t_upd, launches=int(n_upd), bytes=bytes_upd, gbs=bytes_upd / t_upd / 1e9),
        "apply_plain": dict(avg_s=t_plain, bytes=bytes_plain, gbs=bytes_plain / t_plain / 1e9),
    }
    dom = "apply_fused" if t_apply >= t_upd else "cg_line"
    k = kern[dom]
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(k["gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(k["gbs"] / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_launch": k["bytes"], "avg_launch_us": round(1e6 * k["avg_s"], 2)}

    cpu = None
    if not args.no_cpu:
        full = CPU_FULL_ITERS["ic0"] if size == (128, 128, 256) else None
        c = cpu_baseline(p, args.cpu_seconds, full)
        cpu = {"value": c.get("value"), "unit": "solves/s", "cores": 1, "kind": "port",
               "sample": "serial IC(0)-CG (reference default preconditioner, NSPCG ic2 semantics) on the "
                         "14-diagonal SparseBandMatrix storage of the same system: assembly timed whole, "
                         "%d iterations timed (%.3f s/iter) and scaled to the %s iterations of the full "
                         "solve to the same stopping rule" % (c["iters_sampled"], c["s_per_iter"], full),
               "assemble_s": round(c["assemble_s"], 3), "s_per_iter": c["s_per_iter"]}

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
        "config": {"workload": "thermal3d_%dx%dx%d_fp64_zline_pcg" % size, "dof": N, "elements": E,
                   "rtol": rtol, "residual": "||B-AT||/||B-AT_init||, T_init = 300 K",
                   "iters": fp64["iters"], "rel_resid": fp64["rel_resid"]},
        "mixed_precision": mixed,
        "fp64_jacobi_pcg": jac,
        "kernels": kern,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
