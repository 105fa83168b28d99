"""Pins of the CPU statement of the thermal boundary loads (oracle/bc.py): closed-form
1D solutions, the consistent load of the cylindrical weights, and the reference's 3D
placement of wall loads."""
import numpy as np
import pytest
from scipy.optimize import brentq

import oracle as O
from oracle import bc as BC


def _grid_axes(shape, seed):
    rng = np.random.default_rng(seed)
    return [np.cumsum(np.r_[0.0, rng.uniform(0.2, 2.0, n - 1)]) for n in shape]


def _solve(axes, k_h, k_v, dn, dv, F, cyl=False):
    g = O.Grid(axes, cyl=cyl)
    E = g.nelems
    A = O.assemble(g, 0, np.full(E, k_h), np.full(E, k_v), np.zeros(E), O.ST_BAND, dn, dv)
    return O.band_cholesky_solve(A, BC.with_load(A.B, F, dn))


def test_flux_2d_matches_the_1d_solution():
    """Flux v (W/m^2, outward) on the top of a 2D slab held at T0 below: with elements
    aligned to the slab the Q1 solution is the exact T0 - 1e-6 v z / k_v (z in um)."""
    axes = _grid_axes((9, 12), seed=1)
    n0, n1 = 9, 12
    top = np.arange((n1 - 1) * n0, n1 * n0)
    v, kv, T0 = 2.5e5, 7.0, 320.0
    F = BC.load_2d(axes, flux=BC.first_values([(top, v)]))
    dn = np.arange(n0)
    T = _solve(axes, 2.0 * kv, kv, dn, np.full(n0, T0), F)
    z = np.repeat(axes[1], n0)
    assert np.allclose(T, T0 - 1e-6 * v * z / kv, rtol=0, atol=1e-10 * T0)


def test_flux_3d_goes_to_the_elements_lower_face():
    """3D: the reference adds a wall's four loads to the element's local nodes 0-3, so a
    top flux enters one plane below the top: T is linear up to that plane, constant
    above (the 1D solution with the load moved down by one element layer)."""
    shape = (5, 4, 7)
    axes = _grid_axes(shape, seed=2)
    n0, n1, n2 = shape
    plane = n0 * n1
    top = np.arange((n2 - 1) * plane, n2 * plane)
    v, kv, T0 = -4.0e5, 3.0, 300.0
    F = BC.load_3d(axes, flux=BC.first_values([(top, v)]))
    assert np.all(F[(n2 - 1) * plane:] == 0.0)
    dn = np.arange(plane)
    T = _solve(axes, 0.5 * kv, kv, dn, np.full(plane, T0), F)
    zc = np.minimum(axes[2], axes[2][n2 - 2])
    want = np.repeat(T0 - 1e-6 * v * zc / kv, plane)
    assert np.allclose(T, want, rtol=0, atol=1e-10 * T0)


def test_cylindrical_flux_is_the_consistent_load():
    """Cylindrical top flux: each node's load is -1e-6 v times the integral of its hat
    function times r over the top edges (two-point Gauss, exact here); a flux on the
    outer side r = R loads each node with -0.5e-6 v R times the adjacent edge lengths."""
    axes = _grid_axes((8, 6), seed=3)
    n0, n1 = 8, 6
    r = axes[0]
    v = 1.7e4
    top = np.arange((n1 - 1) * n0, n1 * n0)
    F = BC.load_2d(axes, flux=BC.first_values([(top, v)]), cyl=True)
    want = np.zeros(n0)
    gp = np.array([-1.0, 1.0]) / np.sqrt(3.0)
    for i in range(n0 - 1):
        a, b = r[i], r[i + 1]
        for s in gp:
            x = 0.5 * (a + b) + 0.5 * (b - a) * s
            w = 0.5 * (b - a)
            want[i] += w * (b - x) / (b - a) * x
            want[i + 1] += w * (x - a) / (b - a) * x
    assert np.allclose(F[top], -1e-6 * v * want, rtol=1e-13, atol=0)
    assert F[top].sum() == pytest.approx(-1e-6 * v * 0.5 * (r[-1] ** 2 - r[0] ** 2), rel=1e-13)
    right = np.arange(n0 - 1, n0 * n1, n0)
    F = BC.load_2d(axes, flux=BC.first_values([(right, v)]), cyl=True)
    h = np.diff(axes[1])
    lens = np.r_[h, 0.0] + np.r_[0.0, h]
    assert np.allclose(F[right], -0.5e-6 * v * r[-1] * lens, rtol=1e-13, atol=0)
    assert np.count_nonzero(F) == n1


def test_radiation_2d_picard_fixed_point():
    """Radiation on the top of a slab held at T0: the Picard loop (load at the previous
    temperatures) converges to the balance k (T0 - T_top) / L = eps SB (T_top^4 - Ta^4)."""
    axes = _grid_axes((6, 9), seed=4)
    n0, n1 = 6, 9
    top = np.arange((n1 - 1) * n0, n1 * n0)
    eps, Ta, T0, kv = 0.8, 300.0, 600.0, 0.05
    rad = BC.first_values([(top, (eps, Ta))])
    dn = np.arange(n0)
    T = np.full(n0 * n1, T0)
    for _ in range(200):
        F = BC.load_2d(axes, radiation=rad, T=T)
        Tn = _solve(axes, kv, kv, dn, np.full(n0, T0), F)
        done = np.abs(Tn - T).max() < 1e-12 * T0
        T = Tn
        if done:
            break
    L = 1e-6 * axes[1][-1]
    want = brentq(lambda t: kv * (T0 - t) / L - eps * BC.SB * (t ** 4 - Ta ** 4), Ta, T0, xtol=1e-13)
    assert T[top] == pytest.approx(np.full(n0, want), rel=1e-11)


def test_first_place_wins():
    a, b = np.array([1, 2, 3]), np.array([3, 4])
    d = BC.first_values([(a, 5.0), (b, 7.0)])
    assert d == {1: 5.0, 2: 5.0, 3: 5.0, 4: 7.0}


def test_3d_radiation_reads_the_wall_node_number():
    """The reference's 3D radiation term takes the temperature of mesh node wall[i] (a
    local number 0-7), not of the wall's own node: with the top plane at 500 K and
    every other node at 400 K, a top radiation condition radiates at 400 K."""
    axes = [np.arange(3.0)] * 3
    N = 27
    top = np.arange(18, 27)
    T = np.full(N, 400.0)
    T[top] = 500.0
    F = BC.load_3d(axes, radiation=BC.first_values([(top, (1.0, 0.0))]), T=T)
    per = -0.25e-12 * BC.SB * 400.0 ** 4   # per top wall and local node 0-3
    counts = np.zeros(N)
    for j in range(2):
        for i in range(2):
            for l in range(4):
                counts[(i + (l & 1)) + 3 * (j + (l >> 1)) + 9 * 1] += 1
    assert np.allclose(F, per * counts, rtol=1e-14, atol=0)


# --- convection (the matrix terms take part in the elimination) ---------------------------

def _conv_solve(axes, k, dn, dv, conv_entries, Fc, cyl=False):
    g = O.Grid(axes, cyl=cyl)
    E = g.nelems
    A0 = O.assemble(g, 0, np.full(E, k), np.full(E, k), np.zeros(E), O.ST_BAND)
    return BC.solve_with_matrix(A0.dense(), A0.B, conv_entries, Fc, dn, dv)


def test_convection_2d_slab_closed_form():
    """2D slab held at T0 below, convection (h, T_amb) on top.  The reference's equations
    at a top node (per unit of the node's width): k (T_top - T_below) / dz + h T_top =
    1e-6 h T_amb -- its matrix lacks the 1e-6 um -> m of its load (therm2d.cpp:246-252).
    The profile stays linear, T = T0 + g z with g = (1e-6 h T_amb - h T0) / (k + h L);
    for k = 1, L = 10, h = 10, T0 = T_amb = 300 the top sits at 2.97 K."""
    axes = [np.array([0.0, 0.7, 2.0, 2.2, 4.0]), np.array([0.0, 1.5, 4.0, 7.0, 10.0])]
    n0, n1 = 5, 5
    top = np.arange((n1 - 1) * n0, n1 * n0)
    k, h, T0, Tamb = 1.0, 10.0, 300.0, 300.0
    Fc, Kc = BC.convection_2d(axes, BC.first_values([(top, (h, Tamb))]))
    dn = np.arange(n0)
    T = _conv_solve(axes, k, dn, np.full(n0, T0), Kc, Fc)
    L = axes[1][-1]
    g = (1e-6 * h * Tamb - h * T0) / (k + h * L)
    z = np.repeat(axes[1], n0)
    assert np.allclose(T, T0 + g * z, rtol=0, atol=1e-9 * T0)
    assert abs(T[-1] - 2.9703) < 1e-3


def test_convection_3d_bottom_closed_form():
    """3D slab with convection on its bottom wall (whose nodes are the element's local
    nodes 0-3, so the reference's placement is the geometric one) and T_top held.  The
    reference's wall matrix 0.125e-12 area (c_i + c_j) / (9 | 18 | 36) sums to a quarter
    of its load's 0.25e-12 area c per node (therm3d.cpp:252-257), so per unit area
    k 1e-6 (T_b - T_a) / dz + 0.25e-12 h T_b = 1e-12 h T_amb and the profile is linear:
    T_b = (1e-12 h T_amb + 1e-6 k T_top / L) / (1e-6 k / L + 0.25e-12 h)."""
    axes = [np.array([0.0, 1.0, 2.5, 3.0]), np.array([0.0, 0.5, 2.0]), np.array([0.0, 0.4, 1.6, 3.0, 5.0])]
    n0, n1, n2 = 4, 3, 5
    plane = n0 * n1
    bottom = np.arange(plane)
    top = np.arange((n2 - 1) * plane, n2 * plane)
    k, h, Ttop, Tamb = 2.0, 3e5, 310.0, 290.0
    Fc, Kc = BC.convection_3d(axes, BC.first_values([(bottom, (h, Tamb))]))
    T = _conv_solve(axes, k, top, np.full(plane, Ttop), Kc, Fc)
    L = axes[2][-1]
    Tb = (1e-12 * h * Tamb + 1e-6 * k * Ttop / L) / (1e-6 * k / L + 0.25e-12 * h)
    z = np.repeat(axes[2], plane)
    assert np.allclose(T, Tb + (Ttop - Tb) * z / L, rtol=0, atol=1e-9 * Ttop)
    assert abs(Tb - Ttop) > 1.0


def test_convection_3d_top_wall_goes_to_the_lower_face():
    """A top convective wall puts its terms on the element's local nodes 0-3: the top
    plane then carries no equation but conduction and stays at the plane below's value."""
    axes = [np.array([0.0, 1.0, 2.0]), np.array([0.0, 1.0]), np.array([0.0, 1.0, 2.0, 3.0])]
    n0, n1, n2 = 3, 2, 4
    plane = n0 * n1
    top = np.arange((n2 - 1) * plane, n2 * plane)
    Fc, Kc = BC.convection_3d(axes, BC.first_values([(top, (1e5, 350.0))]))
    assert not Fc[top].any() and Fc[(n2 - 2) * plane:(n2 - 1) * plane].all()
    T = _conv_solve(axes, 1.0, np.arange(plane), np.full(plane, 300.0), Kc, Fc)
    assert np.allclose(T[top], T[(n2 - 2) * plane:(n2 - 1) * plane], rtol=0, atol=1e-9)


def test_convection_cylindrical_far_from_the_axis():
    """An annulus far from the axis (r in [1000, 1004]) is the 2D slab times r: the
    reference's cylindrical terms reduce to the Cartesian closed form to O(len / r)."""
    axes = [np.array([1000.0, 1001.0, 1002.5, 1004.0]), np.array([0.0, 2.0, 5.0, 10.0])]
    n0, n1 = 4, 4
    top = np.arange((n1 - 1) * n0, n1 * n0)
    k, h, T0, Tamb = 1.0, 10.0, 300.0, 300.0
    Fc, Kc = BC.convection_2d(axes, BC.first_values([(top, (h, Tamb))]), cyl=True)
    T = _conv_solve(axes, k, np.arange(n0), np.full(n0, T0), Kc, Fc, cyl=True)
    g = (1e-6 * h * Tamb - h * T0) / (k + h * axes[1][-1])
    z = np.repeat(axes[1], n0)
    assert np.allclose(T, T0 + g * z, rtol=0, atol=2e-3 * T0)


def test_convection_matrix_is_positive_semidefinite():
    axes = [np.array([0.0, 1.0, 2.0, 3.0]), np.array([0.0, 1.0, 3.0]), np.array([0.0, 1.0, 2.0])]
    conv = BC.first_values([(np.arange(36), (50.0, 300.0))])   # every node
    _, Kc = BC.convection_3d(axes, conv)
    M = np.zeros((36, 36))
    for r, c, v in Kc:
        M[r, c] += v
        if r != c:
            M[c, r] += v
    assert np.linalg.eigvalsh(M).min() > -1e-12 * np.abs(M).max()
