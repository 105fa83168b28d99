"""GPU parity of the matrix-free operator, load vector and PCG solve against the
CPU oracle (FP64 tolerances stated per test)."""
import numpy as np
import pytest

import oracle as O
from plask_amd.fem import (PF_FP32_REFINE, PF_FP32_RELIABLE, PF_FP64, PF_THERMAL, PF_ELECTRICAL, PRECOND_JACOBI,
                           PRECOND_LINE, FemContext)

pytestmark = pytest.mark.gpu


def _rand_problem(shape, cyl=False, seed=0, ndir_extra=5):
    rng = np.random.default_rng(seed)
    axes = [np.cumsum(np.r_[0.0, rng.uniform(0.05, 2.0, n - 1)]) for n in shape]
    if cyl:
        axes[0] = axes[0] + 0.0
    g = O.Grid(axes, cyl=cyl)
    E = g.nelems
    kh = rng.uniform(0.5, 150.0, E)
    kv = rng.uniform(0.5, 150.0, E)
    q = rng.uniform(0, 1e16, E)
    plane = shape[0] * (shape[1] if len(shape) == 3 else 1)
    dn = np.r_[np.arange(plane), rng.choice(np.arange(plane, g.nnodes), ndir_extra, replace=False)]
    dv = rng.uniform(280, 320, len(dn))
    return g, axes, kh, kv, q, dn.astype(np.int64), dv


def _ctx(axes, kh, kv, q, dn, dv, cyl=False, eq=PF_THERMAL):
    c = FemContext()
    c.set_grid(axes, cylindrical=cyl)
    c.set_coeffs(kh, kv, equation=eq)
    c.set_source(q)
    c.set_dirichlet(dn, dv)
    return c


SHAPES3 = [(5, 4, 6), (17, 9, 13), (64, 7, 9), (96, 11, 21), (128, 20, 12), (33, 40, 70)]
SHAPES2 = [(5, 4), (17, 33), (200, 41), (64, 300)]


@pytest.mark.parametrize("shape", SHAPES3 + SHAPES2)
def test_apply_matches_oracle(shape):
    cyl = len(shape) == 2 and shape[0] % 2 == 0
    g, axes, kh, kv, q, dn, dv = _rand_problem(shape, cyl=cyl, seed=len(shape) * 100 + shape[0])
    print("apply", shape, "cyl", cyl, flush=True)
    c = _ctx(axes, kh, kv, q, dn, dv, cyl=cyl)
    mask = np.zeros(g.nnodes, np.uint8)
    mask[dn] = 1
    x = np.random.default_rng(1).uniform(-1, 1, g.nnodes)
    y_ref = O.apply(g, 0, kh, kv, mask, x)
    y = c.apply(x)
    scale = np.abs(y_ref).max()
    assert np.abs(y - y_ref).max() <= 1e-13 * scale, np.abs(y - y_ref).max() / scale


@pytest.mark.parametrize("shape", SHAPES3)
@pytest.mark.parametrize("kind,planes", [(0, 16), (1, 4), (1, 5), (1, 64)])
def test_apply_kernels_match_oracle(shape, kind, planes):
    """Both 3D operator kernels (tile blocks, row blocks with several plane
    chunks) give the oracle's operator and residual."""
    g, axes, kh, kv, q, dn, dv = _rand_problem(shape, seed=7 + shape[0])
    print("apply kernel", kind, planes, shape, flush=True)
    c = _ctx(axes, kh, kv, q, dn, dv)
    c.set_apply(kind, planes)
    mask = np.zeros(g.nnodes, np.uint8)
    mask[dn] = 1
    x = np.random.default_rng(2).uniform(-1, 1, g.nnodes)
    y_ref = O.apply(g, 0, kh, kv, mask, x)
    y = c.apply(x)
    assert np.abs(y - y_ref).max() <= 1e-13 * np.abs(y_ref).max()
    A = O.assemble(g, 0, kh, kv, q, O.ST_BAND, dn, dv)
    b = c.rhs()
    assert np.abs(b - A.B).max() <= 1e-12 * np.abs(A.B).max()
    T, st = c.solve(x0=np.full(g.nnodes, 300.0), rtol=1e-10)
    assert st.rel_resid <= 1e-10


@pytest.mark.parametrize("shape", [(5, 4, 6), (17, 9, 13), (96, 11, 21), (5, 4), (64, 33)])
def test_rhs_matches_oracle(shape):
    cyl = len(shape) == 2
    g, axes, kh, kv, q, dn, dv = _rand_problem(shape, cyl=cyl, seed=7)
    print("rhs", shape, flush=True)
    c = _ctx(axes, kh, kv, q, dn, dv, cyl=cyl)
    A = O.assemble(g, 0, kh, kv, q, O.ST_BAND, dn, dv)
    b = c.rhs()
    assert np.abs(b - A.B).max() <= 1e-12 * np.abs(A.B).max()


@pytest.mark.parametrize("shape,cyl", [((9, 7, 11), False), ((24, 20, 30), False), ((40, 60), True), ((31, 45), False)])
@pytest.mark.parametrize("precision", [PF_FP64, PF_FP32_REFINE, PF_FP32_RELIABLE])
@pytest.mark.parametrize("precond", [PRECOND_LINE, PRECOND_JACOBI])
def test_solve_matches_band_cholesky(shape, cyl, precision, precond):
    g, axes, kh, kv, q, dn, dv = _rand_problem(shape, cyl=cyl, seed=3)
    print("solve", shape, cyl, precision, precond, flush=True)
    A = O.assemble(g, 0, kh, kv, q, O.ST_BAND, dn, dv)
    T_ref = O.band_cholesky_solve(A)
    c = _ctx(axes, kh, kv, q, dn, dv, cyl=cyl)
    T, st = c.solve(x0=np.full(g.nnodes, 300.0), precision=precision, rtol=1e-12, precond=precond)
    assert st.rel_resid <= 1e-12
    rel = np.abs(T - T_ref).max() / np.abs(T_ref).max()
    assert rel <= 1e-8, rel
    # the residual reported is the true one of the eliminated system
    r = A.B - A.mult(T)
    assert np.linalg.norm(r) / np.linalg.norm(A.B) <= 2e-12


@pytest.mark.parametrize("shape,cyl", [((9, 7, 11), False), ((6, 5, 3), False), ((10, 6, 257), False),
                                       ((40, 60), True), ((5, 1100), False), ((33, 2), False)])
@pytest.mark.parametrize("seg", [-1, 0, 4, 16, 32])
def test_line_schedules_same_preconditioner(shape, cyl, seg):
    """The partitioned line solve (segments coupled through the separator Schur
    complement) is the same exact line preconditioner as the sequential one: same
    solution and the same PCG iteration count (+-1 from rounding).  rtol 1e-10: the
    tall random grids here are too ill-conditioned for 1e-12 in FP64."""
    g, axes, kh, kv, q, dn, dv = _rand_problem(shape, cyl=cyl, seed=5)
    print("line schedule", shape, cyl, seg, flush=True)
    A = O.assemble(g, 0, kh, kv, q, O.ST_BAND, dn, dv)
    T_ref = O.band_cholesky_solve(A)
    c = _ctx(axes, kh, kv, q, dn, dv, cyl=cyl)
    c.set_line_segment(-1)
    T_seq, st_seq = c.solve(x0=np.full(g.nnodes, 300.0), precision=PF_FP64, rtol=1e-10, precond=PRECOND_LINE)
    c.set_line_segment(seg)
    for precision in (PF_FP64, PF_FP32_REFINE, PF_FP32_RELIABLE):
        print("  precision", precision, flush=True)
        T, st = c.solve(x0=np.full(g.nnodes, 300.0), precision=precision, rtol=1e-10, precond=PRECOND_LINE)
        assert st.rel_resid <= 1e-10
        rel = np.abs(T - T_ref).max() / np.abs(T_ref).max()
        assert rel <= 1e-7, (precision, rel)
        if precision == PF_FP64:
            assert abs(st.iters - st_seq.iters) <= 1, (st.iters, st_seq.iters)
            assert np.abs(T - T_seq).max() / np.abs(T_seq).max() <= 1e-9
