"""Train the warm-start network on solver-generated stacks and evaluate it.

    python -m plask_amd.train_warmstart --shape 64,64,128 --train 2000 --held 200 \
        --out plask_amd/data/warmstart_64x64x128.npz

Training data: thermal_3d stacks (seeds 4000 + i; 40 layers, kxy and kz
log-uniform in [1, 150] W/m/K, cylindrical source in the 0.1 um active layer,
300 K at the bottom) solved by the GPU solver to 1e-8; the target is
(T - 300) / target_scale.  Held-out stacks (seeds 9000 + i) are solved from the
uniform inittemp start and from the network guess, both to 1e-8 relative
residual (||B - A T|| / ||B - A T_init||, T_init = 300 K); iterations and times
are reported.  Input features are computed by a pool of host processes while
the GPU solves.  The weights and a JSON summary are written.
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

from . import fem, warmstart, workloads


def _feat_worker(args):
    seed, shape = args
    p = workloads.thermal_3d(seed=seed, n=shape)
    return seed, warmstart.features(p), warmstart.z_edge_conductance(p)


def _setup(ctx, p):
    ctx.set_grid(p["axes"])
    ctx.set_coeffs(p["k_h"], p["k_v"])
    ctx.set_source(p["q"])
    ctx.set_dirichlet(p["dir_nodes"], p["dir_vals"])


def load_weights(path):
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32), tuple(int(v) for v in z["shape"])


def evaluate(weights, shape, held, precond=fem.PRECOND_MG, log=print):
    """Held-out stacks (seeds 9000 + i) solved by PCG (multigrid by default) to 1e-8 of the
    uniform start's residual: from the uniform inittemp start, from the network guess as it
    is (unchecked), and from the guess checked by the residual norm; with, per stack, the
    ratio of the guess's error energy to the uniform start's, estimated as (r, M r) with the
    preconditioner M (the quantity CG's progress depends on; M approximates A^-1)."""
    params, wshape = load_weights(weights)
    if tuple(wshape) != tuple(shape):
        raise ValueError("weights are for %s, not %s" % (wshape, shape))
    ctx = fem.FemContext()
    net = warmstart.WarmNet(shape, max_samples=1)
    net.set_params(params)
    ws = warmstart.WarmStart(net, shape)
    rows = []
    for j in range(held):
        p = workloads.thermal_3d(seed=9000 + j, n=shape)
        _setup(ctx, p)
        if j == 0:
            net.set_mask(warmstart.dirichlet_mask(p))
        t0 = time.perf_counter()
        st_u = ctx.solve_uniform(300.0, rtol=1e-8, precond=precond)
        t_u = time.perf_counter() - t0
        T = ctx.solution()
        f = warmstart.features(p)
        t0 = time.perf_counter()
        ws.start_on_device(ctx, p, f)
        t_g = time.perf_counter() - t0
        xg = ctx.solution()
        xg[p["dir_nodes"]] = p["dir_vals"]   # the solve scatters the Dirichlet values into its start
        t0 = time.perf_counter()
        st_f = ctx.solve_warm(rtol=1e-8, init_value=300.0, precond=precond)
        t_f = time.perf_counter() - t0
        ws.start_on_device(ctx, p, f)
        st_c = ctx.solve_warm(rtol=1e-8, init_value=300.0, precond=precond, checked=True)
        xu = np.full(ctx.nnodes, 300.0)
        xu[p["dir_nodes"]] = p["dir_vals"]
        B = ctx.rhs()
        ru, rg = B - ctx.apply(xu), B - ctx.apply(xg)
        eu = float(ru @ ctx.precond_apply(ru, fem.PF_FP64, precond))
        eg = float(rg @ ctx.precond_apply(rg, fem.PF_FP64, precond))
        rise = np.abs(T - 300.0).max()
        rows.append(dict(seed=9000 + j, it_uniform=st_u.iters, it_net=st_f.iters, it_checked=st_c.iters,
                         fallback=int(st_c.start), energy_ratio=eg / eu,
                         resid_ratio=float(np.linalg.norm(rg) / np.linalg.norm(ru)),
                         guess_err=float(np.abs(xg - T).max() / rise), ms_uniform=1e3 * t_u, ms_guess=1e3 * t_g,
                         ms_net_solve=1e3 * t_f))
        if (j + 1) % 50 == 0:
            log("held-out %d/%d" % (j + 1, held))
    ctx.close()
    a = lambda k: np.array([r[k] for r in rows], dtype=np.float64)
    summ = dict(shape=list(shape), held=held, precond=int(precond),
                iters_uniform_mean=float(a("it_uniform").mean()), iters_net_mean=float(a("it_net").mean()),
                iters_checked_mean=float(a("it_checked").mean()),
                net_fewer_iters=int((a("it_net") < a("it_uniform")).sum()),
                net_more_iters=int((a("it_net") > a("it_uniform")).sum()),
                energy_ratio_median=float(np.median(a("energy_ratio"))),
                energy_ratio_below_1=int((a("energy_ratio") < 1.0).sum()),
                resid_ratio_median=float(np.median(a("resid_ratio"))),
                guess_err_median=float(np.median(a("guess_err"))),
                ms_uniform_mean=float(a("ms_uniform").mean()), ms_guess_mean=float(a("ms_guess").mean()),
                ms_net_solve_mean=float(a("ms_net_solve").mean()), fallbacks=int(a("fallback").sum()))
    return summ, rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--shape", default="64,64,128")
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--held", type=int, default=200)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--lam", type=float, default=1.0, help="weight of the x/y neighbour-difference loss term")
    ap.add_argument("--mu", type=float, default=1.0,
                    help="weight of the vertical-conductance energy term (relative to the target's own); 0: off")
    ap.add_argument("--out", default="")
    ap.add_argument("--summary", default="")
    ap.add_argument("--eval", default="", metavar="WEIGHTS",
                    help="only evaluate these weights on the held-out stacks (multigrid PCG; see evaluate)")
    a = ap.parse_args(argv)
    shape = tuple(int(v) for v in a.shape.split(","))
    log = lambda *s: print(*s, file=sys.stderr, flush=True)
    if a.eval:
        summ, rows = evaluate(a.eval, shape, a.held, log=log)
        if a.summary:
            with open(a.summary, "w") as fh:
                json.dump(dict(summary=summ, rows=rows), fh, indent=1)
        print(json.dumps(summ))
        return summ
    # the feature workers are forked before this process touches the GPU
    pool = mp.get_context("fork").Pool(a.workers)
    ctx = fem.FemContext()
    net = warmstart.WarmNet(shape, max_samples=a.train)
    net.set_params(warmstart.init_params(a.seed))
    net.set_loss(a.lam)
    t0 = time.perf_counter()
    iters_train = []
    try:
        seeds = list(range(4000, 4000 + a.train))
        for i, (seed, f, cond) in enumerate(pool.imap(_feat_worker, [(s, shape) for s in seeds], chunksize=4)):
            p = workloads.thermal_3d(seed=seed, n=shape)
            _setup(ctx, p)
            st = ctx.solve_uniform(300.0, rtol=1e-8)
            T = ctx.solution()
            if i == 0:
                net.set_mask(warmstart.dirichlet_mask(p))
            t = (T - 300.0) / warmstart.target_scale(p)
            net.set_sample(i, f, t)
            if a.mu > 0:
                net.set_sample_weights(i, warmstart.z_edge_weights(p, t, a.mu, conductance=cond))
            iters_train.append(st.iters)
            if (i + 1) % 200 == 0:
                log("data %d/%d (%.0f s)" % (i + 1, a.train, time.perf_counter() - t0))
    finally:
        pool.close()
        pool.join()
    t_data = time.perf_counter() - t0
    rng = np.random.default_rng(a.seed)
    t0 = time.perf_counter()
    losses = []
    for s in range(a.steps):
        lr = a.lr * 0.5 * (1 + np.cos(np.pi * s / a.steps))
        losses.append(net.train_batch(rng.choice(a.train, a.batch, replace=False), lr))
        if (s + 1) % 250 == 0:
            log("step %d loss %.4g (%.0f s)" % (s + 1, np.mean(losses[-50:]), time.perf_counter() - t0))
    t_train = time.perf_counter() - t0
    params = net.params()
    if a.out:
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        np.savez(a.out, params=params, shape=np.array(shape))
    # held-out evaluation
    ws = warmstart.WarmStart(net, shape)
    rows = []
    for j in range(a.held):
        p = workloads.thermal_3d(seed=9000 + j, n=shape)
        _setup(ctx, p)
        st0 = ctx.solve_uniform(300.0, rtol=1e-8)
        T0 = ctx.solution()
        st3 = ctx.solve_uniform(300.0, rtol=1e-8, precision=fem.PF_FP32_REFINE)
        f = warmstart.features(p)
        x0 = ws.guess(p, p)
        tg = time.perf_counter()
        ws.start_on_device(ctx, p, f)
        tg = time.perf_counter() - tg
        st1 = ctx.solve_warm(rtol=1e-8, init_value=300.0, checked=True)
        T1 = ctx.solution()
        ws.start_on_device(ctx, p, f)
        st2 = ctx.solve_warm(rtol=1e-8, init_value=300.0, checked=True, precision=fem.PF_FP32_REFINE)
        rise = np.abs(T0 - 300.0).max()
        if (j + 1) % 50 == 0:
            log("held-out %d/%d" % (j + 1, a.held))
        rows.append(dict(seed=9000 + j, it_uniform=st0.iters, it_net=st1.iters, it_net_fp32r=st2.iters,
                         it_uniform_fp32r=st3.iters, fallback=int(st1.start), guess_err=float(np.abs(x0 - T0).max() / rise),
                         sol_diff=float(np.abs(T1 - T0).max() / rise), rel_resid=st1.rel_resid, guess_s=tg))
    it0 = np.array([r["it_uniform"] for r in rows])
    it1 = np.array([r["it_net"] for r in rows])
    summ = dict(shape=shape, train=a.train, held=a.held, steps=a.steps, batch=a.batch, lr=a.lr,
                data_s=t_data, train_s=t_train, train_ms_per_step=1e3 * t_train / max(a.steps, 1),
                loss_first=float(np.mean(losses[:20])), loss_last=float(np.mean(losses[-20:])),
                train_iters_mean=float(np.mean(iters_train)),
                held_iters_uniform_mean=float(it0.mean()), held_iters_net_mean=float(it1.mean()),
                held_iters_ratio=float(it1.sum() / it0.sum()),
                held_iters_fp32r_uniform_mean=float(np.mean([r["it_uniform_fp32r"] for r in rows])),
                held_iters_fp32r_net_mean=float(np.mean([r["it_net_fp32r"] for r in rows])),
                held_guess_err_median=float(np.median([r["guess_err"] for r in rows])),
                held_max_sol_diff=float(max(r["sol_diff"] for r in rows)),
                held_fallbacks=int(sum(r["fallback"] for r in rows)))
    log(json.dumps(summ))
    if a.summary:
        with open(a.summary, "w") as fh:
            json.dump(dict(summary=summ, rows=rows), fh, indent=1)
    print(json.dumps(summ))
    return summ


if __name__ == "__main__":
    main()
