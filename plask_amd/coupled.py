"""Self-consistent thermo-electrical loop on a seeded 3D layered device.

    python -m plask_amd.coupled --shape 96,96,192 --outer 20

Device (SplitMix64 seed 5): 12 layers bottom -> top, thickness U[0.05, 2] um,
electrical conductivity log-uniform in [1e-2, 1e5] S/m, thermal conductivity
k(T) = k0 (300/T)^alpha with k0 log-uniform in [1, 150] W/m/K and alpha uniform
in [0.5, 1.5]; layer 6 is the 0.02 um junction (role 'active', js log-uniform in
[1e-2, 1e2] A/m^2, beta uniform in [10, 25] 1/V).  100 x 100 um, voltage
U[1.4, 2.2] V on the top nodes with x^2 + y^2 <= 10^2 um^2, 0 V and 300 K at the
bottom.  The loop is ``outer`` iterations of (electrical.compute(6),
thermal.compute(1)) as ThermoElectric.compute (thermoelectric.py:209-211), run once
with FP64 PCG and once with FP32 PCG + FP64 refinement; the final temperatures are
compared (and a third run with FP64 CG and the preconditioner in FP32).  The coupled fields
stay on the device (ThermoElectric.link); a fourth run passes them through the providers'
host arrays, for comparison.  The 3D solvers
precondition with the multigrid V-cycle by default.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import geometry, material, mesh, workloads
from .meta.shockley import ThermoElectric3D


def build(seed=5, shape=(96, 96, 192), nlayers=12, junction=6, width=100.0, contact_r=10.0):
    rng = workloads.SplitMix64(seed)
    t = [rng.uniform(0.05, 2.0) for _ in range(nlayers)]
    sig = [rng.loguniform(1e-2, 1e5) for _ in range(nlayers)]
    k0 = [rng.loguniform(1.0, 150.0) for _ in range(nlayers)]
    alpha = [rng.uniform(0.5, 1.5) for _ in range(nlayers)]
    js = rng.loguniform(1e-2, 1e2)
    beta = rng.uniform(10.0, 25.0)
    V = rng.uniform(1.4, 2.2)
    t[junction] = 0.02
    stack = geometry.Stack3D(xcenter=0.0, ycenter=0.0)
    for i in range(nlayers):
        m = material.PowerLaw((k0[i], k0[i]), (alpha[i], alpha[i]), cond=(sig[i], sig[i]), name="layer%d" % i)
        stack.append(geometry.Cuboid(width, width, t[i], m, role="active" if i == junction else None))
    space = geometry.Cartesian3D(stack)
    nx, ny, nz = shape
    x = np.linspace(-width / 2, width / 2, nx)
    y = np.linspace(-width / 2, width / 2, ny)
    z, _ = workloads.layered_axis(t, nz - 1, minimum=2)
    msh = mesh.Rectangular3D(x, y, z)
    te = ThermoElectric3D("te")
    te.geometry = space
    te.mesh = msh
    te.thermal.temperature_boundary.append(mesh.Rectangular3D.Bottom(), 300.0)
    ztop = z[-1]
    contact = mesh.Rectangular3D.Predicate(
        lambda p: (np.abs(p[:, 2] - ztop) < 1e-9) & (p[:, 0] ** 2 + p[:, 1] ** 2 <= contact_r ** 2 + 1e-9),
        vectorized=True)
    te.electrical.voltage_boundary.append(contact, V)
    te.electrical.voltage_boundary.append(mesh.Rectangular3D.Bottom(), 0.0)
    te.electrical.beta = beta
    te.electrical.js = js
    return te, dict(seed=seed, shape=shape, layers=t, sigma=sig, k0=k0, alpha=alpha, js=js, beta=beta, V=V)


def run(te, outer=20, tfreq=6, verbose=False, link=True):
    """outer x (electrical.compute(tfreq), thermal.compute(1)); returns timings.  link: the
    coupled fields stay on the device (ThermoElectric.link), else they pass through the
    providers' host arrays."""
    te.invalidate()
    te.initialize()
    linked = te.link() if link else False
    t_el = t_th = 0.0
    hist = []
    try:
        for _ in range(outer):
            t0 = time.perf_counter()
            verr = te.electrical.compute(tfreq)
            t1 = time.perf_counter()
            terr = te.thermal.compute(1)
            t2 = time.perf_counter()
            t_el += t1 - t0
            t_th += t2 - t1
            hist.append((float(verr), float(terr)))
            if verbose:
                print("outer %d: electrical err %.3g %% (%.2f s), thermal err %.3g K (%.2f s)" %
                      (len(hist), verr, t1 - t0, terr, t2 - t1), file=sys.stderr, flush=True)
    finally:
        te.unlink()
    return dict(electrical_s=t_el, thermal_s=t_th, total_s=t_el + t_th, history=hist, linked=linked,
                current_mA=te.get_total_current(), Tmax=float(te.thermal._temps().max()))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--shape", default="96,96,192")
    ap.add_argument("--outer", type=int, default=20)
    # the Joule heating raises T by ~0.07 K over 300 K: the FP64 relative residual of the
    # thermal system bottoms out near 4e-8 (cancellation in B - A T), so ask for 1e-7
    ap.add_argument("--rtol", type=float, default=1e-7, help="thermal PCG tolerance")
    a = ap.parse_args(argv)
    shape = tuple(int(v) for v in a.shape.split(","))
    out = {}
    T = {}
    for name, prec, link in (("fp64", 0, True), ("fp32_refine", 1, True), ("fp64_cg_fp32_precond", 3, True),
                             ("fp64_host_exchange", 0, False)):
        te, info = build(shape=shape)
        te.thermal.iterative.precision = prec
        te.electrical.iterative.precision = prec
        te.thermal.iterative.maxerr = a.rtol
        t0 = time.perf_counter()
        te.initialize()
        t_init = time.perf_counter() - t0
        r = run(te, a.outer, verbose=True, link=link)
        r["init_s"] = t_init
        T[name] = te.thermal._temps().copy()
        out[name] = r
        print("%s: %.2f s (electrical %.2f s, thermal %.2f s), I = %.6g mA, Tmax = %.6f K" %
              (name, r["total_s"], r["electrical_s"], r["thermal_s"], r["current_mA"], r["Tmax"]),
              file=sys.stderr, flush=True)
    d = np.abs(T["fp32_refine"] - T["fp64"]).max() / np.abs(T["fp64"]).max()
    d3 = np.abs(T["fp64_cg_fp32_precond"] - T["fp64"]).max() / np.abs(T["fp64"]).max()
    dh = np.abs(T["fp64_host_exchange"] - T["fp64"]).max() / np.abs(T["fp64"]).max()
    summ = dict(config=dict(info, layers=None), outer=a.outer, thermal_rtol=a.rtol,
                preconditioner="multigrid V-cycle (solver default)",
                fp64={k: v for k, v in out["fp64"].items() if k != "history"},
                fp32_refine={k: v for k, v in out["fp32_refine"].items() if k != "history"},
                fp64_cg_fp32_precond={k: v for k, v in out["fp64_cg_fp32_precond"].items() if k != "history"},
                fp64_host_exchange={k: v for k, v in out["fp64_host_exchange"].items() if k != "history"},
                speedup=out["fp64"]["total_s"] / out["fp32_refine"]["total_s"],
                speedup_device_exchange=out["fp64_host_exchange"]["total_s"] / out["fp64"]["total_s"],
                max_rel_T_diff=float(d), max_rel_T_diff_fp32_precond=float(d3),
                max_rel_T_diff_host_exchange=float(dh),
                history_fp64=out["fp64"]["history"])
    print(json.dumps(summ))
    return summ


if __name__ == "__main__":
    main()
