"""Pins of the CPU statement of the multigrid preconditioner (oracle/multigrid.py):
the nested-space identity of the coarse operator and the symmetry / positivity of
the V-cycle."""
import numpy as np
import pytest
import scipy.sparse.linalg as spl

import oracle as O
from oracle import multigrid as MG


def _layered(shape, seed, extra_dirichlet=0):
    rng = np.random.default_rng(seed)
    axes = [np.cumsum(np.r_[0.0, rng.uniform(0.1, 2.0, n - 1)]) for n in shape]
    n0, n1, n2 = shape
    per_layer_h = rng.uniform(0.5, 150.0, n2 - 1)
    per_layer_v = rng.uniform(0.5, 150.0, n2 - 1)
    e = (n0 - 1) * (n1 - 1)
    kh = np.repeat(per_layer_h, e)
    kv = np.repeat(per_layer_v, e)
    dn = np.arange(n0 * n1)
    if extra_dirichlet:
        dn = np.r_[dn, rng.choice(np.arange(n0 * n1, n0 * n1 * n2), extra_dirichlet, replace=False)]
    return axes, kh, kv, dn


@pytest.mark.parametrize("shape", [(9, 7, 5), (8, 11, 4), (4, 13, 3), (17, 6, 4)])
def test_coarse_operator_is_galerkin_for_layered_coefficients(shape):
    """Coarse Q1 functions are fine Q1 functions (interpolation P is exact on them), so
    with conductivities constant in each element layer the rediscretized coarse operator
    equals R A P on the free nodes.  A wrong coarse axis, weight, averaging or mask
    breaks the identity."""
    axes, kh, kv, dn = _layered(shape, seed=sum(shape))
    vc = MG.VCycle(axes, kh, kv, dn)
    assert len(vc.levels) >= 2
    for l, P in enumerate(vc.P):
        F, Cl = vc.levels[l], vc.levels[l + 1]
        G = (P.T @ F.A @ P).toarray()
        Ac = Cl.A.toarray()
        free = ~Cl.dmask
        scale = np.abs(Ac[np.ix_(free, free)]).max()
        assert np.abs(G[np.ix_(free, free)] - Ac[np.ix_(free, free)]).max() <= 1e-12 * scale


def test_coarse_conductivity_is_area_mean():
    """Laterally varying conductivity: each coarse element holds the area-weighted mean of
    the fine elements it covers (checked on one hand-computed element)."""
    axes = [np.array([0.0, 1.0, 3.0, 4.0, 6.0]), np.array([0.0, 2.0, 3.0]), np.array([0.0, 1.0, 2.0])]
    kh = np.arange(1.0, 1.0 + 4 * 2 * 2)
    kv = 2.0 * kh
    vc = MG.VCycle(axes, kh, kv, np.arange(10))
    C = vc.levels[1]
    assert C.n == [3, 3, 3]   # x: nodes 0, 2, 4; y: 3 nodes stay (not coarsened)
    # coarse element (0, 0, 0) covers fine x elements 0, 1 (widths 1, 2) of row 0 (height 2)
    want = (1.0 * 2 * kh[0] + 2.0 * 2 * kh[1]) / (1.0 * 2 + 2.0 * 2)
    assert C.kh[0] == pytest.approx(want, rel=1e-14)
    assert C.kv[0] == pytest.approx(2.0 * want, rel=1e-14)


@pytest.mark.parametrize("shape,extra", [((7, 6, 5), 0), ((9, 5, 6), 3)])
def test_vcycle_is_symmetric_positive_definite(shape, extra):
    """The V-cycle with R = P^T, the same damped line sweep before and after, and
    symmetric coarse sweeps is an SPD operator on the free nodes (what PCG needs)."""
    axes, kh, kv, dn = _layered(shape, seed=3, extra_dirichlet=extra)
    rng = np.random.default_rng(0)
    kh = kh * rng.uniform(0.5, 2.0, kh.size)   # lateral variation too
    vc = MG.VCycle(axes, kh, kv, dn)
    free = ~vc.levels[0].dmask
    idx = np.nonzero(free)[0]
    Z = np.zeros((len(idx), len(idx)))
    for c, j in enumerate(idx):
        e = np.zeros(vc.levels[0].N)
        e[j] = 1.0
        Z[:, c] = vc.apply(e)[idx]
    assert np.abs(Z - Z.T).max() <= 1e-12 * np.abs(Z).max()
    assert np.linalg.eigvalsh(0.5 * (Z + Z.T)).min() > 0.0


def test_mg_pcg_iterations_on_layered_stack():
    """PCG with the V-cycle on a seeded layered stack: a handful of iterations where the
    line preconditioner alone needs many (the reason for the preconditioner)."""
    from plask_amd import workloads
    p = workloads.thermal_3d(seed=2, n=(24, 24, 48))
    vc = MG.VCycle(p["axes"], p["k_h"], p["k_v"], p["dir_nodes"])
    g = O.Grid(p["axes"])
    A = O.assemble(g, 0, p["k_h"], p["k_v"], p["q"], O.ST_DIA, p["dir_nodes"], p["dir_vals"])
    S = MG.csr(A)
    L0 = vc.levels[0]

    def pcg(prec):
        x = np.full(A.rank, 300.0)
        x[p["dir_nodes"]] = p["dir_vals"]
        r = A.B - S @ x
        b0 = np.linalg.norm(r)
        z = prec(r)
        pv = z.copy()
        rz = r @ z
        for it in range(1, 2000):
            q = S @ pv
            a = rz / (pv @ q)
            x += a * pv
            r -= a * q
            if np.linalg.norm(r) <= 1e-8 * b0:
                return it
            z = prec(r)
            rzn = r @ z
            pv = z + (rzn / rz) * pv
            rz = rzn
        return -1

    it_mg = pcg(vc.apply)
    it_line = pcg(L0.Mlu.solve)
    assert 0 < it_mg <= 15
    assert it_line >= 4 * it_mg


def _random_anisotropic(shape, seed):
    rng = np.random.default_rng(seed)
    axes = [np.cumsum(np.r_[0.0, rng.uniform(0.05, 2.0, n - 1)]) for n in shape]
    E = int(np.prod([n - 1 for n in shape]))
    kh = rng.uniform(0.5, 150.0, E)
    kv = rng.uniform(0.5, 150.0, E)
    plane = shape[0] * shape[1]
    dn = np.r_[np.arange(plane), rng.choice(np.arange(plane, int(np.prod(shape))), 5, replace=False)]
    return axes, kh, kv, dn


@pytest.mark.parametrize("shape,seed", [((9, 7, 11), 11), ((12, 10, 9), 3)])
def test_line_preconditioned_spectrum_bounded_by_three(shape, seed):
    """lambda_max(M_line^-1 A) <= 3 (the element-by-element bound that fixes the smoother
    damping); random element sizes and conductivities come close to it."""
    axes, kh, kv, dn = _random_anisotropic(shape, seed)
    vc = MG.VCycle(axes, kh, kv, dn)
    L = vc.levels[0]
    op = spl.LinearOperator((L.N, L.N), matvec=lambda v: L.Mlu.solve(L.A @ v))
    lam = spl.eigs(op, k=1, which="LM", return_eigenvectors=False, maxiter=5000, tol=1e-8)[0].real
    assert 2.25 < lam <= 3.0 + 1e-9


def test_vcycle_spd_on_random_anisotropic_grid():
    """With the default damping the V-cycle stays positive definite where omega = 0.75
    (omega * lambda_max > 2) does not."""
    axes, kh, kv, dn = _random_anisotropic((9, 7, 11), 11)

    def min_eig(omega):
        vc = MG.VCycle(axes, kh, kv, dn, omega=omega)
        idx = np.nonzero(~vc.levels[0].dmask)[0]
        Z = np.zeros((len(idx), len(idx)))
        for c, j in enumerate(idx):
            e = np.zeros(vc.levels[0].N)
            e[j] = 1.0
            Z[:, c] = vc.apply(e)[idx]
        return np.linalg.eigvalsh(0.5 * (Z + Z.T)).min()

    assert min_eig(0.65) > 0.0
    assert min_eig(0.75) < 0.0


@pytest.mark.parametrize("shape", [(9, 5), (12, 7), (4, 6)])
def test_coarse_operator_is_galerkin_2d(shape):
    """2D Cartesian: same nested-space identity with axis 0 coarsened."""
    rng = np.random.default_rng(sum(shape))
    axes = [np.cumsum(np.r_[0.0, rng.uniform(0.1, 2.0, n - 1)]) for n in shape]
    kh = np.repeat(rng.uniform(0.5, 150.0, shape[1] - 1), shape[0] - 1)
    kv = np.repeat(rng.uniform(0.5, 150.0, shape[1] - 1), shape[0] - 1)
    vc = MG.VCycle(axes, kh, kv, np.arange(shape[0]))
    assert len(vc.levels) >= 2
    for l, P in enumerate(vc.P):
        F, Cl = vc.levels[l], vc.levels[l + 1]
        G = (P.T @ F.A @ P).toarray()
        Ac = Cl.A.toarray()
        free = ~Cl.dmask
        scale = np.abs(Ac[np.ix_(free, free)]).max()
        assert np.abs(G[np.ix_(free, free)] - Ac[np.ix_(free, free)]).max() <= 1e-12 * scale


def test_cylindrical_coarse_conductivity_keeps_radial_conductance():
    """Cylindrical: the coarse conductivity times the coarse mid radius and width equals the
    sum of k r h over the fine elements it covers (the weight the element matrix carries)."""
    axes = [np.array([0.0, 1.0, 3.0, 4.0, 6.0]), np.array([0.0, 1.0, 2.0])]
    kh = np.arange(1.0, 9.0)
    vc = MG.VCycle(axes, kh, 2.0 * kh, np.arange(5), cyl=True)
    C = vc.levels[1]
    assert C.n == [3, 3]
    r = 0.5 * (axes[0][1:] + axes[0][:-1])
    h = np.diff(axes[0])
    want = (kh[0] * r[0] * h[0] + kh[1] * r[1] * h[1]) / (1.5 * 3.0)   # coarse element [0, 3]: r 1.5, h 3
    assert C.kh[0] == pytest.approx(want, rel=1e-14)


def test_vcycle_spd_2d_cylindrical():
    rng = np.random.default_rng(5)
    shape = (11, 9)
    axes = [np.cumsum(np.r_[0.0, rng.uniform(0.05, 2.0, n - 1)]) for n in shape]
    E = (shape[0] - 1) * (shape[1] - 1)
    kh = rng.uniform(0.5, 150.0, E)
    kv = rng.uniform(0.5, 150.0, E)
    vc = MG.VCycle(axes, kh, kv, np.r_[np.arange(shape[0]), [30, 47]], cyl=True)
    idx = np.nonzero(~vc.levels[0].dmask)[0]
    Z = np.zeros((len(idx), len(idx)))
    for c, j in enumerate(idx):
        e = np.zeros(vc.levels[0].N)
        e[j] = 1.0
        Z[:, c] = vc.apply(e)[idx]
    assert np.abs(Z - Z.T).max() <= 1e-12 * np.abs(Z).max()
    assert np.linalg.eigvalsh(0.5 * (Z + Z.T)).min() > 0.0
