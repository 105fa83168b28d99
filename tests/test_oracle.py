"""Pins of the CPU oracle (oracle/fem_oracle.c) by checks that do not retype its formulas:

* element matrices vs 2x2(x2) Gauss quadrature of the Q1 bilinear/trilinear forms
* assembly + Dirichlet elimination vs an independent numpy COO assembly
* band storage vs symmetric-diagonal storage of the same system
* band Cholesky vs LAPACK dpbtrf/dpbtrs (scipy) and numpy dense solves
* layered stacks vs the closed-form 1D temperature profile (2D, cylindrical, 3D)
* PCG (Jacobi, IC(0)) vs the direct solution
"""
import itertools

import numpy as np
import pytest
import scipy.linalg.lapack as lapack

import oracle as O
from plask_amd import workloads


def gauss_K(d, sizes, kdiag):
    """Q1 stiffness of a box by 2-point Gauss quadrature: K_ab = int k_d dphi_a/dx_d dphi_b/dx_d."""
    gp = [0.5 - 0.5 / np.sqrt(3), 0.5 + 0.5 / np.sqrt(3)]
    n = 2 ** d
    K = np.zeros((n, n))
    for pt in itertools.product(gp, repeat=d):
        w = np.prod(sizes) / 2 ** d
        grads = np.zeros((n, d))
        for a in range(n):
            bits = [(a >> t) & 1 for t in range(d)]
            for dd in range(d):
                g = 1.0
                for t in range(d):
                    xi = pt[t]
                    if t == dd:
                        g *= (1 if bits[t] else -1) / sizes[t]
                    else:
                        g *= xi if bits[t] else 1 - xi
                grads[a, dd] = g
        K += w * (grads * np.asarray(kdiag)) @ grads.T
    return K


def numpy_assemble(grid, kh, kv, q, dn, dv, scale3d=1e-6):
    """Independent assembly (dense) with the reference's unit scalings and elimination."""
    N = grid.nnodes
    A = np.zeros((N, N))
    F = np.zeros(N)
    n = grid.n
    d = grid.dim
    ranges = [range(m - 1) for m in n]
    for ii in itertools.product(*ranges[::-1]):
        ii = ii[::-1]
        e = ii[0] + (n[0] - 1) * (ii[1] + ((n[1] - 1) * ii[2] if d == 3 else 0))
        sizes = [grid.axes[a][ii[a] + 1] - grid.axes[a][ii[a]] for a in range(d)]
        if d == 2:
            k = [kh[e], kv[e]]
            r = 0.5 * (grid.axes[0][ii[0]] + grid.axes[0][ii[0] + 1]) if grid.cyl else 1.0
            Ke = r * gauss_K(2, sizes, k)
            f = 1e-12 * r * np.prod(sizes) * q[e] / 4
        else:
            k = [kh[e] * scale3d, kh[e] * scale3d, kv[e] * scale3d]
            Ke = gauss_K(3, sizes, k)
            f = 1e-18 * np.prod(sizes) * q[e] / 8
        nodes = []
        for a in range(2 ** d):
            bits = [(a >> t) & 1 for t in range(d)]
            nodes.append(sum((ii[t] + bits[t]) * grid.s[t] for t in range(d)))
        for a in range(2 ** d):
            F[nodes[a]] += f
            for b in range(2 ** d):
                A[nodes[a], nodes[b]] += Ke[a, b]
    B = F - A[:, dn] @ dv
    B[dn] = dv
    A[dn, :] = 0
    A[:, dn] = 0
    A[dn, dn] = 1
    return A, B


def rand_case(shape, cyl=False, seed=0, order=None):
    rng = np.random.default_rng(seed)
    axes = [np.cumsum(np.r_[0.1 if cyl and a == 0 else 0.0, rng.uniform(0.1, 2.0, n - 1)]) for a, n in enumerate(shape)]
    g = O.Grid(axes, cyl=cyl, order=order)
    E = g.nelems
    kh = rng.uniform(1, 150, E)
    kv = rng.uniform(1, 150, E)
    q = rng.uniform(0, 1e15, E)
    plane = [g.node_index(*ii) for ii in itertools.product(*[range(m) for m in shape[:-1]])]
    dn = np.array([p + 0 for p in plane], dtype=np.int64)  # bottom face (last axis index 0)
    dv = rng.uniform(290, 310, len(dn))
    return g, kh, kv, q, dn, dv


def test_element_matrices_vs_gauss_quadrature():
    rng = np.random.default_rng(5)
    for _ in range(5):
        s = rng.uniform(0.1, 3, 3)
        k = rng.uniform(1, 100, 2)
        g = O.Grid([np.array([0, s[0]]), np.array([0, s[1]]), np.array([0, s[2]])])
        A = O.assemble(g, 0, [k[0]], [k[1]], None, O.ST_BAND)
        K = A.dense()
        Kg = gauss_K(3, s, [k[0] * 1e-6, k[0] * 1e-6, k[1] * 1e-6])
        assert np.allclose(K, Kg, rtol=1e-12, atol=1e-20)
        g2 = O.Grid([np.array([0, s[0]]), np.array([0, s[1]])])
        K2 = O.assemble(g2, 0, [k[0]], [k[1]], None, O.ST_BAND).dense()
        assert np.allclose(K2, gauss_K(2, s[:2], k), rtol=1e-12)


@pytest.mark.parametrize("shape,cyl,order", [((4, 5), False, [1, 0]), ((5, 4), True, [0, 1]),
                                             ((3, 4, 5), False, None), ((4, 3, 3), False, [0, 2, 1])])
def test_assembly_vs_independent_numpy(shape, cyl, order):
    g, kh, kv, q, dn, dv = rand_case(shape, cyl=cyl, seed=11, order=order)
    A_np, B_np = numpy_assemble(g, kh, kv, q, dn, dv)
    for kind in (O.ST_BAND, O.ST_DIA):
        A = O.assemble(g, 0, kh, kv, q, kind, dn, dv)
        assert np.allclose(A.dense(), A_np, rtol=1e-12, atol=1e-14 * np.abs(A_np).max())
        assert np.allclose(A.B, B_np, rtol=1e-12, atol=1e-12 * np.abs(B_np).max())


def test_band_cholesky_vs_lapack_and_dense():
    g, kh, kv, q, dn, dv = rand_case((9, 13), seed=2, order=[1, 0])
    A = O.assemble(g, 0, kh, kv, q, O.ST_BAND, dn, dv)
    x = O.band_cholesky_solve(A)
    assert np.allclose(x, np.linalg.solve(A.dense(), A.B), rtol=1e-11)
    # LAPACK 'L' band: ab[i-j, j]; oracle stores (r,c) at ld*c + r
    n, kd, ld = A.rank, A.kd, A.ld
    ab = np.zeros((kd + 1, n))
    for c in range(n):
        for r in range(c, min(n, c + kd + 1)):
            ab[r - c, c] = A.data[ld * c + r]
    c_, info = lapack.dpbtrf(ab, lower=1)
    assert info == 0
    xl, info = lapack.dpbtrs(c_, A.B, lower=1)
    assert np.allclose(x, xl, rtol=1e-12)


def test_not_positive_definite_reports_minor():
    g, kh, kv, q, dn, dv = rand_case((5, 5), seed=3)
    A = O.assemble(g, 0, -kh, -kv, q, O.ST_BAND, dn, dv)
    with pytest.raises(ArithmeticError, match="not positive-definite"):
        O.band_cholesky_solve(A)


def layered_exact(z, layers, active, q, T0):
    bounds = np.cumsum([0] + [t for t, k in layers])
    za, ta, ka = bounds[active], layers[active][0], layers[active][1]

    def T(zz):
        if zz <= za + 1e-12:
            s = sum((min(zz, hi) - lo) / k for (t, k), lo, hi in zip(layers, bounds[:-1], bounds[1:]) if zz > lo)
            return T0 + q * ta * 1e-12 * s
        d = min(zz, za + ta) - za
        return T(za) + q * 1e-12 * (ta * d - d * d / 2) / ka

    return np.array([T(v) for v in z])


@pytest.mark.parametrize("dim,cyl", [(2, False), (2, True), (3, False)])
def test_layered_stack_closed_form(dim, cyl):
    layers = [(1.0, 10.0), (0.5, 3.0), (0.1, 50.0), (2.0, 100.0)]
    active, q, T0 = 2, 1e15, 300.0
    z = [0.0]
    for t, k in layers:
        z += list(np.linspace(z[-1], z[-1] + t, 4)[1:])
    z = np.array(z)
    lat = [np.linspace(0, 8, 5)] * (dim - 1)
    if cyl:
        lat = [np.linspace(0.0, 8, 6)]
    axes = lat + [z]
    g = O.Grid(axes, cyl=cyl)
    bounds = np.cumsum([0] + [t for t, k in layers])
    lid = np.searchsorted(bounds, 0.5 * (z[1:] + z[:-1])) - 1
    per = int(np.prod([len(a) - 1 for a in lat]))
    kel = np.repeat(np.array([layers[i][1] for i in lid]), per)
    heat = np.repeat(np.where(lid == active, q, 0.0), per)
    plane = int(np.prod([len(a) for a in lat]))
    dn = np.arange(plane)
    # 3D conductivities are W/m/K like 2D; the 1e-6 um scaling cancels against the load's 1e-18
    A = O.assemble(g, 0, kel, kel, heat, O.ST_BAND, dn, np.full(plane, T0))
    T = O.band_cholesky_solve(A).reshape(len(z), plane)
    Tex = layered_exact(z, layers, active, q, T0)
    assert np.abs(T - Tex[:, None]).max() <= 1e-10 * Tex.max()


@pytest.mark.parametrize("precond", [0, 1, 2])
def test_pcg_matches_direct(precond):
    g, kh, kv, q, dn, dv = rand_case((6, 5, 7), seed=4)
    Ab = O.assemble(g, 0, kh, kv, q, O.ST_BAND, dn, dv)
    Ad = O.assemble(g, 0, kh, kv, q, O.ST_DIA, dn, dv)
    x_ref = O.band_cholesky_solve(Ab)
    x, st, it, rel = O.dia_pcg(Ad, np.full(g.nnodes, 300.0), precond=precond, rtol=1e-13)
    assert st >= 0 and rel <= 1e-13
    assert np.abs(x - x_ref).max() <= 1e-10 * np.abs(x_ref).max()


def test_matrix_free_apply_matches_assembled():
    for shape in [(4, 5), (3, 4, 5)]:
        g, kh, kv, q, dn, dv = rand_case(shape, seed=9)
        A = O.assemble(g, 0, kh, kv, q, O.ST_BAND, dn, dv)
        m = np.zeros(g.nnodes, np.uint8)
        m[dn] = 1
        x = np.random.default_rng(0).normal(size=g.nnodes)
        assert np.allclose(O.apply(g, 0, kh, kv, m, x), A.mult(x), rtol=1e-12, atol=1e-14)


@pytest.mark.skipif(O.ref_lib() is None, reason="oracle/_ref/libnspcg.so not built (reference sources absent)")
def test_cg_restatement_pinned_by_reference_nspcg():
    """The oracle's PCG (NSPCG itcg order, oracle/fem_oracle.c) against the reference's own
    NSPCG built from its sources (oracle/_ref): Jacobi-CG takes the same number of
    iterations to the same accuracy, IC(0)-CG within one iteration (the two stop on
    different tests: NSPCG's error estimate vs the true residual); both agree with band
    Cholesky."""
    p = workloads.thermal_3d(seed=2, n=(12, 10, 24), nlayers=6, active=3, width=20.0)
    g = O.Grid(p["axes"])
    T_ref = O.band_cholesky_solve(O.assemble(g, 0, p["k_h"], p["k_v"], p["q"], O.ST_BAND, p["dir_nodes"],
                                             p["dir_vals"]))
    A = O.assemble(g, 0, p["k_h"], p["k_v"], p["q"], O.ST_DIA, p["dir_nodes"], p["dir_vals"])
    x0 = np.full(g.nnodes, 300.0)
    for pc, ours in (("jac", 1), ("ic", 2)):
        u, it, _, ier, _, _ = O.nspcg_solve(A, x0, pc, 1e-10, 5000)
        v, st, it2, rel = O.dia_pcg(A, x0, precond=ours, rtol=1e-10)
        assert ier == 0 and st >= 0
        assert np.abs(u - T_ref).max() <= 1e-9 * np.abs(T_ref).max()
        assert np.abs(v - T_ref).max() <= 1e-9 * np.abs(T_ref).max()
        assert abs(it - it2) <= (0 if pc == "jac" else 1), (pc, it, it2)


@pytest.mark.skipif(O.ref_lib() is None, reason="oracle/_ref/libnspcg.so not built (reference sources absent)")
def test_reference_nspcg_default_tolerance_accuracy():
    """PLaSK's default iterative settings (maxerr = zeta = 1e-6, IC(0)) give a solution
    about 1e-6 relative from the direct one -- the accuracy the reference's 3D path
    delivers by default (the GPU solves are run to a 1e-8 residual, well inside it)."""
    p = workloads.thermal_3d(seed=2, n=(12, 10, 24), nlayers=6, active=3, width=20.0)
    g = O.Grid(p["axes"])
    T_ref = O.band_cholesky_solve(O.assemble(g, 0, p["k_h"], p["k_v"], p["q"], O.ST_BAND, p["dir_nodes"],
                                             p["dir_vals"]))
    A = O.assemble(g, 0, p["k_h"], p["k_v"], p["q"], O.ST_DIA, p["dir_nodes"], p["dir_vals"])
    u, it, _, ier, _, _ = O.nspcg_solve(A, np.full(g.nnodes, 300.0), "ic", 1e-6, 1000)
    err = np.abs(u - T_ref).max() / np.abs(T_ref).max()
    assert ier == 0 and 1e-9 < err < 1e-5, err
