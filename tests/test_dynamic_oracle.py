"""Pins of the oracle's transient thermal loop (oracle.thermal_dynamic) on the CPU.

The oracle restates DynamicThermalFem{2D,3D}Solver::compute and setMatrix
(solvers/thermal/dynamic/femT2d.cpp:146-243, :413-440; femT3d.cpp:140-234) with lumped
capacities.  What fixes it independently of its own code:

* energy balance: with no held nodes, uniform material and uniform heat, K 1 = 0 and every
  node warms by q dt / (cp rho) per step, whatever theta (2D, cylindrical, 3D);
* a cosine mode along a uniform axis is an exact eigenvector of the lumped pencil
  (K v = mu C v with mu = kx (1 - cos(pi / (n - 1))) / (2 c) in 2D, kx / (4 c) in 3D), so
  its amplitude follows g = (1 - (1 - theta) mu) / (1 + theta mu) per step;
* theta = 0 is the explicit update T' = T + (F - K T) / C, with K T from the oracle's
  matrix-free element loop (orc_apply), a second implementation of K;
* theta = 1 makes B = C, so a held node takes val + C_r T_r (the reference eliminates the
  condition from A and F only) and the free rows satisfy (K + C) T' = C T + F with the
  held values in T';
* theta = 1 with vanishing capacity is the steady solve (band Cholesky of the eliminated
  static system).
"""
import numpy as np
import pytest

import oracle as O

CPRHO = 1.7e6  # J / (m^3 K)


def _cap_2d(x, y, dt, cyl=False, cprho=CPRHO):
    """Element capacities as setMatrix forms them (femT2d.cpp:167, :298)."""
    w, h = np.diff(x), np.diff(y)
    c = cprho * 0.25e-12 * (h[:, None] * w[None, :]) / dt / 1e-9
    if cyl:
        c = c * (0.5 * (x[1:] + x[:-1]))[None, :]
    return c.ravel()


def _cap_3d(x, y, z, dt, cprho=CPRHO):
    """femT3d.cpp:195: cp rho 0.125e-9 dx dy dz / timestep."""
    v = np.diff(z)[:, None, None] * np.diff(y)[None, :, None] * np.diff(x)[None, None, :]
    return (cprho * 0.125e-9 * v / dt).ravel()


@pytest.mark.parametrize("theta", [0.0, 0.5, 1.0])
@pytest.mark.parametrize("kind", ["2d", "cyl", "3d"])
def test_uniform_heating_energy_balance(kind, theta):
    dt, q, k = 0.1, 3e15, 40.0
    if kind == "3d":
        x, y, z = np.linspace(0, 3, 5), np.array([0, 0.4, 1.5, 2.0]), np.array([0, 0.3, 0.5, 1.4])
        g = O.Grid([x, y, z])
        cap = _cap_3d(x, y, z, dt)
    else:
        x, y = np.array([0.0, 0.5, 1.2, 3.0, 3.1]), np.array([0.0, 0.2, 0.9, 1.0])
        g = O.Grid([x, y], cyl=(kind == "cyl"))
        cap = _cap_2d(x, y, dt, cyl=(kind == "cyl"))
    E = g.nelems
    T = O.thermal_dynamic(g, np.full(E, k), np.full(E, k * 0.7), np.full(E, q), cap, theta,
                          np.full(g.nnodes, 300.0), 4)
    rise = q * dt * 1e-9 / CPRHO
    for n, Tn in enumerate(T, 1):
        assert np.abs(Tn - (300.0 + n * rise)).max() <= 1e-8 * n * rise


def _cosine_case(dim, nx=9):
    x = np.linspace(0.0, 4.0, nx)
    y = np.array([0.0, 1.0, 2.2, 3.0, 4.5])       # any spacing across the mode
    z = np.array([0.0, 0.9, 2.0, 3.1])
    L = x[-1]
    v = np.cos(np.pi * x / L)
    return x, y, z, v


@pytest.mark.parametrize("theta", [0.0, 0.5, 1.0])
@pytest.mark.parametrize("dim", [2, 3])
def test_cosine_mode_decay(dim, theta):
    # mu about 0.43; the explicit step needs mu_max < 2 over all modes (the Nyquist mode
    # along x has 2 / (1 - cos(pi / 8)) = 26 times the cosine's mu)
    dt, k = (6.0 if theta == 0.0 else 40.0), 30.0
    x, y, z, v = _cosine_case(dim)
    w = x[1] - x[0]
    cosd = np.cos(np.pi / (len(x) - 1))
    if dim == 2:
        g = O.Grid([x, y])
        cap = _cap_2d(x, y, dt)
        # per element: kx = k h / w, c_e = cap; mu = kx (1 - cos) / (2 c_e) (same for every row)
        h = np.diff(y)[0]
        c_e = cap[0]
        mu = (k * h / w) * (1 - cosd) / (2 * c_e)
        T0 = 300.0 + 5.0 * np.tile(v, len(y))
    else:
        g = O.Grid([x, y, z])
        cap = _cap_3d(x, y, z, dt)
        dy, dz = np.diff(y)[0], np.diff(z)[0]
        c_e = cap[0]
        mu = (k * 1e-6 * dy * dz / w) * (1 - cosd) / (4 * c_e)
        T0 = 300.0 + 5.0 * np.tile(v, len(y) * len(z))
    E = g.nelems
    T = O.thermal_dynamic(g, np.full(E, k), np.full(E, 0.3 * k), None, cap, theta, T0, 5)
    gain = (1 - (1 - theta) * mu) / (1 + theta * mu)
    assert 0.05 < abs(gain) < 0.999
    for n, Tn in enumerate(T, 1):
        expect = 300.0 + 5.0 * gain ** n * (T0 - 300.0) / 5.0
        assert np.abs(Tn - expect).max() <= 1e-10 * 5.0


def _held_case():
    x = np.array([0.0, 0.4, 1.0, 1.9, 2.5, 3.0])
    y = np.array([0.0, 0.5, 0.8, 1.6, 2.0])
    g = O.Grid([x, y])
    E = g.nelems
    rng = np.random.default_rng(7)
    kh = rng.uniform(5.0, 50.0, E)
    kv = kh * rng.uniform(0.5, 1.0, E)
    q = np.where(rng.random(E) < 0.4, 2e15, 0.0)
    cap = _cap_2d(x, y, 0.2) * rng.uniform(0.5, 2.0, E) * 1e-5
    dn = np.arange(len(x))                          # bottom row held
    dv = np.full(len(dn), 300.0)
    T0 = 300.0 + rng.uniform(0.0, 3.0, g.nnodes)
    return g, kh, kv, q, cap, dn, dv, T0


def test_explicit_step_is_the_textbook_update():
    g, kh, kv, q, cap, dn, dv, T0 = _held_case()
    T1 = O.thermal_dynamic(g, kh, kv, q, cap, 0.0, T0, 1, dn, dv)[0]
    KT = O.apply(g, 0, kh, kv, np.zeros(g.nnodes, np.uint8), T0)   # matrix-free K T
    F = O.assemble(g, 0, kh, kv, q, O.ST_BAND).B
    Cn = O.node_capacity(g, cap)
    free = np.ones(g.nnodes, bool)
    free[dn] = False
    expect = T0 + (F - KT) / Cn
    assert np.abs(T1[free] - expect[free]).max() <= 1e-11 * np.abs(expect).max()
    held = dv + Cn[dn] * T0[dn] - KT[dn]
    assert np.abs(T1[dn] - held).max() <= 1e-11 * np.abs(held).max()


def test_implicit_step_held_nodes_and_free_rows():
    g, kh, kv, q, cap, dn, dv, T0 = _held_case()
    T1 = O.thermal_dynamic(g, kh, kv, q, cap, 1.0, T0, 1, dn, dv)[0]
    Cn = O.node_capacity(g, cap)
    assert np.allclose(T1[dn], dv + Cn[dn] * T0[dn], rtol=1e-13, atol=0)
    Th = T1.copy()
    Th[dn] = dv
    F = O.assemble(g, 0, kh, kv, q, O.ST_BAND).B
    lhs = O.apply(g, 0, kh, kv, np.zeros(g.nnodes, np.uint8), Th) + Cn * T1
    free = np.ones(g.nnodes, bool)
    free[dn] = False
    rhs = Cn * T0 + F
    assert np.abs(lhs - rhs)[free].max() <= 1e-11 * np.abs(rhs[free]).max()


def test_vanishing_capacity_gives_the_steady_solution():
    g, kh, kv, q, cap, dn, dv, T0 = _held_case()
    T1 = O.thermal_dynamic(g, kh, kv, q, cap * 1e-12, 1.0, T0, 1, dn, dv)[0]
    A = O.assemble(g, 0, kh, kv, q, O.ST_BAND, dn, dv)
    Ts = O.band_cholesky_solve(A)
    rise = np.abs(Ts - 300.0).max()
    assert rise > 1.0
    assert np.abs(T1 - Ts).max() <= 1e-6 * rise


@pytest.mark.parametrize("theta", [0.0, 0.5, 1.0])
def test_fixed_held_nodes_keep_the_steady_state(theta):
    """With held nodes kept at their values the steady solution is a fixed point of every
    theta step (K T* = F on the free rows), and the held values never move."""
    g, kh, kv, q, cap, dn, dv, T0 = _held_case()
    Ts = O.band_cholesky_solve(O.assemble(g, 0, kh, kv, q, O.ST_BAND, dn, dv))
    T = O.thermal_dynamic(g, kh, kv, q, cap * 1e5, theta, Ts, 3, dn, dv, held="fixed")
    for Tn in T:
        assert np.abs(Tn - Ts).max() <= 1e-9 * np.abs(Ts).max()
        assert np.array_equal(Tn[dn], dv)
    # from a cold start the implicit steps approach it
    T = O.thermal_dynamic(g, kh, kv, q, cap, 1.0, np.full(g.nnodes, 300.0), 60, dn, dv, held="fixed")
    assert np.abs(T[-1] - Ts).max() <= 1e-6 * np.abs(Ts - 300.0).max()


def test_node_capacity_sums_the_elements():
    x, y, z = np.linspace(0, 1, 4), np.linspace(0, 1, 3), np.linspace(0, 1, 5)
    g = O.Grid([x, y, z])
    cap = np.arange(1, g.nelems + 1, dtype=float)
    Cn = O.node_capacity(g, cap)
    assert np.isclose(Cn.sum(), 8 * cap.sum())
    # corner node 0 touches element 0 only; the far corner the last element only
    assert Cn[0] == cap[0] and Cn[-1] == cap[-1]
    # an interior node touches 8 elements
    i, j, k = 1, 1, 2
    n = g.node_index(i, j, k)
    ex = [(k - 1 + c2) * 6 + (j - 1 + c1) * 3 + (i - 1 + c0) for c0 in (0, 1) for c1 in (0, 1) for c2 in (0, 1)]
    assert np.isclose(Cn[n], cap[ex].sum())
