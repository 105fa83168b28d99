"""Pins of the oracle's electrical functions on the CPU (orc_currents, orc_junction_update,
electrical_compute), independent of their own code:

* a linear potential gives j = -sigma grad V exactly: 1 V/um = 1e6 V/m, 1 A/m^2 = 1e-7
  kA/cm^2, so j = -0.1 sigma dV/dx (kA/cm^2) and the Joule heat sigma (1e6 dV/dx)^2 (W/m^3)
  (electr2d.cpp:405-409, :457-461; electr3d.cpp:399-411, :459-475), in 2D and 3D;
* the junction law (beta.hpp:43-46) iterated at a fixed junction voltage U converges to the
  Shockley diode current j = 1e-7 js (exp(beta U) - 1) kA/cm^2 (js in A/m^2): its fixed
  point is ln(1e7 j / js + 1) = beta U;
* the Picard loop (electr2d.cpp:363-440) on a junction between two good conductors with
  1 V across carries the reference's analytic current 1e-3 js (e^beta - 1) mA through
  1000 x 1000 um^2 (solvers/electrical/shockley/tests/shockley2d.py:54-70).
"""
import math

import numpy as np
import pytest

import oracle as O


@pytest.mark.parametrize("dim", [2, 3])
def test_currents_of_a_linear_potential(dim):
    rng = np.random.default_rng(dim)
    axes = [np.cumsum(np.r_[0.0, rng.uniform(0.2, 1.5, n - 1)]) for n in ((5, 4) if dim == 2 else (4, 3, 5))]
    g = O.Grid(axes)
    grad = np.array([0.3, -1.2, 0.7][:dim])      # V/um
    X = np.stack(np.meshgrid(*axes, indexing="ij"), -1)
    X = X.transpose(tuple(range(dim))[::-1] + (dim,)).reshape(-1, dim)
    V = 0.25 + X @ grad
    E = g.nelems
    sh = rng.uniform(1.0, 100.0, E)
    sv = rng.uniform(1.0, 100.0, E)
    cur, heat = O.currents(g, sh, sv, V)
    sig = np.column_stack([sh] * (dim - 1) + [sv])
    assert np.allclose(cur, -0.1 * sig * grad, rtol=1e-12, atol=0)
    assert np.allclose(heat, (sig * (1e6 * grad) ** 2).sum(axis=1), rtol=1e-12, atol=0)


@pytest.mark.parametrize("beta,U", [(10.0, 1.0), (18.0, 0.6), (5.0, 0.2)])
def test_junction_law_fixed_point_is_the_diode_current(beta, U):
    js, h = 2.5, 0.05
    V = np.array([0.0, 0.0, U, U])               # bottom corners 0, 1; top corners 2, 3
    corners = np.array([0, 1, 2, 3], dtype=np.int64)
    sig = np.array([5.0])
    for _ in range(400):
        O.junction_update([0], corners, 2, V, h, beta, js, False, sig)
    j = 0.1 * sig[0] * U / h
    want = 1e-7 * js * math.expm1(beta * U)
    assert abs(j - want) <= 1e-9 * want


def test_stable_update_is_the_mean_of_the_fast_one():
    V = np.array([0.0, 0.0, 0.8, 0.8])
    corners = np.array([0, 1, 2, 3], dtype=np.int64)
    a, b = np.array([3.0]), np.array([3.0])
    O.junction_update([0], corners, 2, V, 0.02, 12.0, 1.0, False, a)
    O.junction_update([0], corners, 2, V, 0.02, 12.0, 1.0, True, b)
    assert abs(b[0] - 0.5 * (3.0 + a[0])) <= 1e-15 * b[0]


def test_picard_loop_gives_the_analytic_current():
    beta, js, h = 10.0, 1.0, 0.02
    x = np.array([0.0, 500.0, 1000.0])
    y = np.r_[np.linspace(0.0, 300.0, 4), 300.0 + h, 300.0 + h + np.linspace(100.0, 300.0, 3)]
    g = O.Grid([x, y])
    n0, n1 = len(x), len(y)
    E = g.nelems
    sh = np.full(E, 1e12)   # the conductors' drop stays below 1e-9 of the current
    sv = np.full(E, 1e12)
    kb = 3                                      # the junction is element row 3
    elems = np.arange(kb * (n0 - 1), (kb + 1) * (n0 - 1))
    bottom = [[i + kb * n0, i + 1 + kb * n0] for i in range(n0 - 1)]
    top = [[i + (kb + 1) * n0, i + 1 + (kb + 1) * n0] for i in range(n0 - 1)]
    J = dict(elems=elems, bottom=bottom, top=top, height=h, beta=beta, js=js)
    dn = np.r_[np.arange(n0), np.arange((n1 - 1) * n0, n1 * n0)]
    dv = np.r_[np.full(n0, 1.0), np.zeros(n0)]
    V, cur, heat, sig, loops, err = O.electrical_compute(g, sh, sv, dn, dv, [J], 0, maxerr=1e-9)
    assert err <= 1e-9
    jy = np.abs(cur[elems, 1])                  # kA/cm^2
    current_mA = float(np.sum(jy * np.diff(x)) * 1000.0 * 1e-8) * 1e6
    assert abs(current_mA - 1e-3 * js * math.expm1(beta)) <= 1e-6 * 1e-3 * js * math.expm1(beta)
