"""GPU multigrid preconditioner: one V-cycle against the CPU statement
(oracle/multigrid.py), and multigrid-preconditioned solves against band Cholesky."""
import numpy as np
import pytest

import oracle as O
from oracle import multigrid as MG
from plask_amd.fem import (PF_FP32_REFINE, PF_FP32_RELIABLE, PF_FP64, PF_FP64_PREC32, PRECOND_JACOBI, PRECOND_LINE,
                           PRECOND_MG, FemContext, FemError)

pytestmark = pytest.mark.gpu


def _rand_problem(shape, seed, extra=5):
    rng = np.random.default_rng(seed)
    axes = [np.cumsum(np.r_[0.0, rng.uniform(0.05, 2.0, n - 1)]) for n in shape]
    g = O.Grid(axes)
    E = g.nelems
    kh = rng.uniform(0.5, 150.0, E)
    kv = rng.uniform(0.5, 150.0, E)
    q = rng.uniform(0, 1e16, E)
    plane = shape[0] * shape[1]
    dn = np.r_[np.arange(plane), rng.choice(np.arange(plane, g.nnodes), extra, replace=False)]
    dv = rng.uniform(280, 320, len(dn))
    return g, axes, kh, kv, q, dn.astype(np.int64), dv


def _ctx(axes, kh, kv, q, dn, dv):
    c = FemContext()
    c.set_grid(axes)
    c.set_coeffs(kh, kv)
    c.set_source(q)
    c.set_dirichlet(dn, dv)
    return c


def _levels(n0, n1):
    lv = 1
    while n0 > 3 or n1 > 3:
        n0 = n0 // 2 + 1 if n0 > 3 else n0
        n1 = n1 // 2 + 1 if n1 > 3 else n1
        lv += 1
    return lv


@pytest.mark.parametrize("fuse", [False, True])
@pytest.mark.parametrize("shape", [(9, 7, 11), (17, 9, 13), (33, 40, 30), (10, 6, 257), (6, 5, 3)])
def test_vcycle_matches_oracle(shape, fuse):
    """z = V r on the GPU (FP64 and FP32) vs the CPU statement of the same V-cycle, with the
    grid transfers as kernels of their own and fused into the sweeps / residuals."""
    g, axes, kh, kv, q, dn, dv = _rand_problem(shape, seed=21 + shape[0])
    print("vcycle", shape, "fuse", fuse, flush=True)
    c = _ctx(axes, kh, kv, q, dn, dv)
    c.set_multigrid_fusion(fuse)
    vc = MG.VCycle(axes, kh, kv, dn)
    r = np.random.default_rng(4).uniform(-1, 1, g.nnodes)
    r[dn] = 0.0   # PCG residuals vanish at Dirichlet nodes
    z_ref = vc.apply(r)
    z = c.precond_apply(r, PF_FP64, PRECOND_MG)
    assert c.multigrid_levels() == len(vc.levels) == _levels(shape[0], shape[1])
    scale = np.abs(z_ref).max()
    assert np.abs(z - z_ref).max() <= 1e-10 * scale, np.abs(z - z_ref).max() / scale
    z32 = c.precond_apply(r, PF_FP32_RELIABLE, PRECOND_MG)
    # FP32 line solves on these random tall grids: norm-wise 1e-3
    assert np.linalg.norm(z32 - z_ref) <= 1e-3 * np.linalg.norm(z_ref), np.linalg.norm(z32 - z_ref) / np.linalg.norm(z_ref)
    c.close()


@pytest.mark.parametrize("shape", [(9, 7, 11), (24, 20, 30), (33, 40, 70), (10, 6, 257), (128, 20, 12), (5, 4, 6)])
@pytest.mark.parametrize("precision", [PF_FP64, PF_FP32_REFINE, PF_FP32_RELIABLE, PF_FP64_PREC32])
def test_mg_solve_matches_band_cholesky(shape, precision):
    g, axes, kh, kv, q, dn, dv = _rand_problem(shape, seed=11)
    print("mg solve", shape, precision, flush=True)
    A = O.assemble(g, 0, kh, kv, q, O.ST_BAND, dn, dv)
    T_ref = O.band_cholesky_solve(A)
    c = _ctx(axes, kh, kv, q, dn, dv)
    T, st = c.solve(x0=np.full(g.nnodes, 300.0), precision=precision, rtol=1e-10, precond=PRECOND_MG)
    assert st.rel_resid <= 1e-10
    rel = np.abs(T - T_ref).max() / np.abs(T_ref).max()
    assert rel <= 1e-7, (precision, rel)
    c.close()


@pytest.mark.parametrize("precision", [PF_FP64, PF_FP32_REFINE, PF_FP64_PREC32])
def test_fused_transfers_give_the_same_solve(precision):
    """The fused-transfer V-cycle is the same preconditioner: same iterations, same
    solution to rounding."""
    g, axes, kh, kv, q, dn, dv = _rand_problem((33, 40, 70), seed=12)
    print("mg fused solve", precision, flush=True)
    out = []
    for fuse in (False, True):
        c = _ctx(axes, kh, kv, q, dn, dv)
        c.set_multigrid_fusion(fuse)
        out.append(c.solve(x0=np.full(g.nnodes, 300.0), precision=precision, rtol=1e-10, precond=PRECOND_MG))
        c.close()
    (T0, s0), (T1, s1) = out
    assert s0.iters == s1.iters
    assert np.abs(T1 - T0).max() <= 1e-9 * np.abs(T0).max()


def test_mg_iterations_on_layered_stack():
    """Seeded layered stack 48 x 48 x 96: the V-cycle preconditioner needs a handful of
    PCG iterations where the line preconditioner needs over a hundred; same solution."""
    from plask_amd import workloads
    p = workloads.thermal_3d(seed=2, n=(48, 48, 96))
    print("mg layered 48x48x96", flush=True)
    c = FemContext()
    c.set_grid(p["axes"])
    c.set_coeffs(p["k_h"], p["k_v"])
    c.set_source(p["q"])
    c.set_dirichlet(p["dir_nodes"], p["dir_vals"])
    st_line = c.solve_uniform(300.0, rtol=1e-8, precond=PRECOND_LINE)
    T_line = c.solution()
    st_mg = c.solve_uniform(300.0, rtol=1e-8, precond=PRECOND_MG)
    T_mg = c.solution()
    print("line", st_line.iters, "mg", st_mg.iters, flush=True)
    assert st_mg.iters <= 16
    assert st_line.iters >= 5 * st_mg.iters
    assert np.abs(T_mg - T_line).max() / np.abs(T_line).max() <= 1e-9
    # mixed precision with the V-cycle in FP32
    st32 = c.solve_uniform(300.0, precision=PF_FP32_RELIABLE, rtol=1e-8, precond=PRECOND_MG)
    T32 = c.solution()
    assert st32.rel_resid <= 1e-8
    assert np.abs(T32 - T_line).max() / np.abs(T_line).max() <= 1e-9
    c.close()


def test_mg_needs_partitioned_line_solve():
    g, axes, kh, kv, q, dn, dv = _rand_problem((9, 7, 11), seed=2)
    c = _ctx(axes, kh, kv, q, dn, dv)
    c.set_line_segment(-1)
    with pytest.raises(FemError):
        c.solve(x0=np.full(g.nnodes, 300.0), precond=PRECOND_MG)
    c.set_line_segment(0)
    T, st = c.solve(x0=np.full(g.nnodes, 300.0), rtol=1e-10, precond=PRECOND_MG)
    assert st.rel_resid <= 1e-10
    c.close()


def _rand_problem_2d(shape, cyl, seed):
    rng = np.random.default_rng(seed)
    axes = [np.cumsum(np.r_[0.0, rng.uniform(0.05, 2.0, n - 1)]) for n in shape]
    g = O.Grid(axes, cyl=cyl)
    E = g.nelems
    kh = rng.uniform(0.5, 150.0, E)
    kv = rng.uniform(0.5, 150.0, E)
    q = rng.uniform(0, 1e16, E)
    dn = np.r_[np.arange(shape[0]), rng.choice(np.arange(shape[0], g.nnodes), 3, replace=False)]
    dv = rng.uniform(280, 320, len(dn))
    return g, axes, kh, kv, q, dn.astype(np.int64), dv


def _ctx2(axes, cyl, kh, kv, q, dn, dv):
    c = FemContext()
    c.set_grid(axes, cylindrical=cyl)
    c.set_coeffs(kh, kv)
    c.set_source(q)
    c.set_dirichlet(dn, dv)
    return c


@pytest.mark.parametrize("shape,cyl", [((40, 60), True), ((31, 45), False), ((5, 300), False), ((200, 41), True)])
def test_vcycle_2d_matches_oracle(shape, cyl):
    """2D (axis 0 coarsened, axis 1 the layer axis): one V-cycle vs the CPU statement."""
    g, axes, kh, kv, q, dn, dv = _rand_problem_2d(shape, cyl, seed=shape[0])
    print("vcycle 2d", shape, cyl, flush=True)
    c = _ctx2(axes, cyl, kh, kv, q, dn, dv)
    vc = MG.VCycle(axes, kh, kv, dn, cyl=cyl)
    r = np.random.default_rng(6).uniform(-1, 1, g.nnodes)
    r[dn] = 0.0
    z_ref = vc.apply(r)
    z = c.precond_apply(r, PF_FP64, PRECOND_MG)
    assert c.multigrid_levels() == len(vc.levels)
    scale = np.abs(z_ref).max()
    assert np.abs(z - z_ref).max() <= 1e-10 * scale, np.abs(z - z_ref).max() / scale
    c.close()


@pytest.mark.parametrize("shape,cyl", [((40, 60), True), ((31, 45), False), ((5, 1100), False), ((200, 41), True)])
@pytest.mark.parametrize("precision", [PF_FP64, PF_FP32_RELIABLE])
def test_mg_solve_2d_matches_band_cholesky(shape, cyl, precision):
    g, axes, kh, kv, q, dn, dv = _rand_problem_2d(shape, cyl, seed=7)
    print("mg solve 2d", shape, cyl, precision, flush=True)
    A = O.assemble(g, 0, kh, kv, q, O.ST_BAND, dn, dv)
    T_ref = O.band_cholesky_solve(A)
    c = _ctx2(axes, cyl, kh, kv, q, dn, dv)
    T, st = c.solve(x0=np.full(g.nnodes, 300.0), precision=precision, rtol=1e-10, precond=PRECOND_MG)
    assert st.rel_resid <= 1e-10
    rel = np.abs(T - T_ref).max() / np.abs(T_ref).max()
    assert rel <= 1e-7, (precision, rel)
    c.close()


@pytest.mark.parametrize("shape", [(9, 7, 11), (24, 20, 30), (10, 6, 257)])
def test_fp32_preconditioner_in_fp64_cg(shape):
    """PF_FP64_PREC32: FP64 CG, line or multigrid preconditioner applied in FP32 -- same
    FP64 solution; the Jacobi preconditioner is refused for this mode."""
    g, axes, kh, kv, q, dn, dv = _rand_problem(shape, seed=13)
    print("fp64 cg, fp32 preconditioner", shape, flush=True)
    A = O.assemble(g, 0, kh, kv, q, O.ST_BAND, dn, dv)
    T_ref = O.band_cholesky_solve(A)
    c = _ctx(axes, kh, kv, q, dn, dv)
    for pc in (PRECOND_LINE, PRECOND_MG):
        T, st = c.solve(x0=np.full(g.nnodes, 300.0), precision=PF_FP64_PREC32, rtol=1e-10, precond=pc)
        assert st.rel_resid <= 1e-10
        assert np.abs(T - T_ref).max() / np.abs(T_ref).max() <= 1e-7
    with pytest.raises(FemError):
        c.solve(x0=np.full(g.nnodes, 300.0), precision=PF_FP64_PREC32, precond=PRECOND_JACOBI)
    c.close()


def test_graph_cache_follows_the_preconditioner():
    """Cached iteration graphs are kept per precision and preconditioner: alternating
    FP64 / FP32 solves with different preconditioners give each form's own result."""
    g, axes, kh, kv, q, dn, dv = _rand_problem((24, 20, 30), seed=17)
    print("graph cache", flush=True)
    A = O.assemble(g, 0, kh, kv, q, O.ST_BAND, dn, dv)
    T_ref = O.band_cholesky_solve(A)
    c = _ctx(axes, kh, kv, q, dn, dv)
    seq = [(PF_FP64, PRECOND_LINE), (PF_FP32_REFINE, PRECOND_MG), (PF_FP64, PRECOND_MG), (PF_FP64, PRECOND_JACOBI),
           (PF_FP32_REFINE, PRECOND_LINE), (PF_FP64, PRECOND_MG), (PF_FP64_PREC32, PRECOND_MG), (PF_FP64, PRECOND_LINE)]
    iters = {}
    for prec, pc in seq:
        T, st = c.solve(x0=np.full(g.nnodes, 300.0), precision=prec, rtol=1e-10, precond=pc)
        assert np.abs(T - T_ref).max() / np.abs(T_ref).max() <= 1e-7, (prec, pc)
        iters.setdefault((prec, pc), []).append(st.iters)
    for k, v in iters.items():
        assert max(v) - min(v) <= 1, (k, v)
    c.close()
