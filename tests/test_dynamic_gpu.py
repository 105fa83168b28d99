"""Transient thermal solver on the GPU: the theta-method step of pf_dynamic_step against
the oracle's restatement of DynamicThermalFem2DSolver::compute (solvers/thermal/dynamic/
femT2d.cpp:413-440, femT3d.cpp:140-234), the closed forms at full size, and the solver
classes of thermal.dynamic.

Tolerances: FP64 PCG to 1e-12 of the start residual per step; node temperatures within
1e-9 of the largest |T| (oracle parity) and 1e-7 K (closed forms).
"""
import numpy as np
import pytest

import oracle as O
from plask_amd import geometry, material, mesh
from plask_amd.fem import PF_HELD_FIXED, PF_HELD_REFERENCE, PF_THERMAL, PRECOND_LINE, PRECOND_MG, FemContext
from plask_amd.thermal.dynamic import Dynamic2D, Dynamic3D, DynamicCyl

pytestmark = pytest.mark.gpu

CPRHO = 1.7e6


def _problem(shape, cyl, seed, held):
    rng = np.random.default_rng(seed)
    axes = [np.cumsum(np.r_[0.0, rng.uniform(0.2, 1.5, n - 1)]) for n in shape]
    g = O.Grid(axes, cyl=cyl)
    E = g.nelems
    kh = rng.uniform(2.0, 80.0, E)
    kv = rng.uniform(2.0, 80.0, E)
    q = np.where(rng.random(E) < 0.3, rng.uniform(1e14, 3e15, E), 0.0)
    # capacities of the order of the element stiffness
    d = [np.diff(a) for a in axes]
    if len(shape) == 3:
        vol = (d[2][:, None, None] * d[1][None, :, None] * d[0][None, None, :]).ravel()
    else:
        vol = (d[1][:, None] * d[0][None, :]).ravel()
    scale = 1e-6 if len(shape) == 3 else 1.0
    cap = scale * rng.uniform(5.0, 40.0, E) * vol
    if cyl:
        cap = cap * np.tile(0.5 * (axes[0][1:] + axes[0][:-1]), E // (len(axes[0]) - 1))
    plane = shape[0] * (shape[1] if len(shape) == 3 else 1)
    if held:
        dn = np.r_[np.arange(plane), rng.choice(np.arange(plane, g.nnodes), 3, replace=False)].astype(np.int64)
        dv = rng.uniform(290.0, 310.0, len(dn))
    else:
        dn, dv = np.zeros(0, np.int64), np.zeros(0)
    T0 = 300.0 + rng.uniform(-2.0, 2.0, g.nnodes)
    return g, axes, kh, kv, q, cap, dn, dv, T0


CASES = [((17, 13), False), ((12, 9), True), ((9, 8, 7), False), ((21, 6, 11), False)]


@pytest.mark.parametrize("theta", [0.0, 0.5, 1.0])
@pytest.mark.parametrize("held", ["none", "fixed", "reference"])
@pytest.mark.parametrize("shape,cyl", CASES)
def test_steps_match_oracle(shape, cyl, held, theta):
    g, axes, kh, kv, q, cap, dn, dv, T0 = _problem(shape, cyl, seed=len(shape) * 10 + shape[0], held=held != "none")
    print("dynamic", shape, "cyl", cyl, "held", held, "theta", theta, flush=True)
    mode = "fixed" if held == "none" else held
    ref = O.thermal_dynamic(g, kh, kv, q, cap, theta, T0, 4, dn, dv, held=mode)
    c = FemContext()
    c.set_grid(axes, cylindrical=cyl)
    c.set_coeffs(kh, kv, equation=PF_THERMAL)
    c.set_source(q)
    c.set_dirichlet(dn, dv)
    c.set_solution(T0)
    c.set_capacity(cap, theta)
    for n in range(4):
        st = c.dynamic_step(rtol=1e-12, precond=PRECOND_MG if n % 2 == 0 else PRECOND_LINE,
                            held=PF_HELD_REFERENCE if held == "reference" else PF_HELD_FIXED)
        T = c.solution()
        scale = np.abs(ref[n]).max()
        err = np.abs(T - ref[n]).max()
        assert err <= 1e-9 * scale, (n, err / scale, st.iters)


def test_capacity_off_restores_the_steady_operator():
    g, axes, kh, kv, q, cap, dn, dv, T0 = _problem((9, 8, 7), False, seed=5, held=True)
    c = FemContext()
    c.set_grid(axes)
    c.set_coeffs(kh, kv)
    c.set_source(q)
    c.set_dirichlet(dn, dv)
    x_ref, _ = c.solve(x0=T0, rtol=1e-12)
    c.set_solution(T0)
    c.set_capacity(cap, 0.3)
    c.dynamic_step(rtol=1e-12)
    c.set_capacity(None)
    x, _ = c.solve(x0=T0, rtol=1e-12)
    assert np.abs(x - x_ref).max() <= 1e-10 * np.abs(x_ref).max()
    # a coefficient update while transient keeps theta K in the operator
    c.set_capacity(cap, 1.0)
    c.set_coeffs(kh * 2.0, kv * 2.0)
    c.set_solution(T0)
    c.dynamic_step(rtol=1e-12, held=PF_HELD_REFERENCE)
    ref = O.thermal_dynamic(g, kh * 2.0, kv * 2.0, q, cap, 1.0, T0, 1, dn, dv)[0]
    assert np.abs(c.solution() - ref).max() <= 1e-9 * np.abs(ref).max()


def test_bad_capacity_input_is_rejected():
    c = FemContext()
    c.set_grid([np.linspace(0, 1, 5), np.linspace(0, 1, 4)])
    c.set_coeffs(np.ones(12), np.ones(12))
    with pytest.raises(RuntimeError):
        c.set_capacity(np.ones(11), 0.5)
    with pytest.raises(RuntimeError):
        c.set_capacity(np.ones(12), 1.5)
    with pytest.raises(RuntimeError):
        c.set_capacity(-np.ones(12), 0.5)
    with pytest.raises(RuntimeError):
        c.dynamic_step()


@pytest.mark.parametrize("theta", [0.5, 1.0])
def test_cosine_mode_decay_full_size(theta):
    """A 3D grid of 2.3 M nodes: the cosine mode along the uniform x axis decays by
    g = (1 - (1 - theta) mu) / (1 + theta mu) per step, exactly for the lumped pencil.
    Each step to 1e-9 of max(start residual, 1e-4 ||B||)."""
    nx, ny, nz = 193, 120, 100
    x = np.linspace(0.0, 96.0, nx)
    rng = np.random.default_rng(3)
    y = np.cumsum(np.r_[0.0, rng.uniform(0.3, 1.2, ny - 1)])
    z = np.cumsum(np.r_[0.0, rng.uniform(0.1, 0.8, nz - 1)])
    E = (nx - 1) * (ny - 1) * (nz - 1)
    k, dt = 30.0, 16000.0   # mu about 0.3
    v = np.diff(z)[:, None, None] * np.diff(y)[None, :, None] * np.diff(x)[None, None, :]
    cap = (CPRHO * 0.125e-9 * v / dt).ravel()
    w = x[1] - x[0]
    beta = CPRHO * 0.125e-9 / dt
    mu = k * 1e-6 * (1 - np.cos(np.pi / (nx - 1))) / (4 * beta * w * w)
    gain = (1 - (1 - theta) * mu) / (1 + theta * mu)
    mode = np.cos(np.pi * x / x[-1])
    T0 = 300.0 + 5.0 * np.tile(mode, ny * nz)
    print("cosine decay", (nx, ny, nz), "theta", theta, "mu", mu, flush=True)
    c = FemContext()
    c.set_grid([x, y, z])
    c.set_coeffs(np.full(E, k), np.full(E, 0.4 * k))
    c.set_source(None)
    c.set_dirichlet(np.zeros(0, np.int64), np.zeros(0))
    c.set_solution(T0)
    c.set_capacity(cap, theta)
    assert 0.2 < mu < 0.5
    for n in range(1, 4):
        c.dynamic_step(rtol=1e-9)
    T = c.solution()
    expect = 300.0 + gain ** 3 * (T0 - 300.0)
    assert np.abs(T - expect).max() <= 1e-7


def _slab(cls, held):
    cp, rho = 330.0, 5300.0
    m = material.Const(thermk=(45.0, 40.0), cp=cp, dens=rho, name="Sl")
    if cls is Dynamic3D:
        stack = geometry.Stack3D()
        stack.append(geometry.Cuboid(6., 5., 2.0, m))
        s = cls("dyn3")
        s.geometry = geometry.Cartesian3D(stack)
        s.mesh = mesh.Rectangular3D(np.linspace(0, 6, 7), np.linspace(0, 5, 6), np.linspace(0, 2, 5))
    else:
        stack = geometry.Stack2D()
        stack.append(geometry.Rectangle(6., 2.0, m))
        s = cls("dyn2")
        s.geometry = geometry.Cylindrical(stack) if cls is DynamicCyl else geometry.Cartesian2D(stack)
        s.mesh = mesh.Rectangular2D(np.linspace(0, 6, 13), np.linspace(0, 2, 6))
    if held:
        s.temperature_boundary.append(s.mesh.Bottom(), 300.)
    s.inHeat = 2e15
    s.iterative.maxerr = 1e-13
    return s, cp * rho


@pytest.mark.parametrize("cls", [Dynamic2D, DynamicCyl, Dynamic3D])
def test_solver_uniform_heating(cls):
    """compute(time) runs the reference loop t = 0, dt, ... while t < time + dt/2
    (femT2d.cpp:418-419): compute(1.0) at dt = 0.1 takes 11 steps and reports time 1.0;
    each step warms every node by q dt / (cp rho)."""
    s, cprho = _slab(cls, held=False)
    s.timestep = 0.1
    s.compute(1.0)
    assert abs(s.time - 1.0) < 1e-12
    T = s.outTemperature(s.mesh)
    rise = 2e15 * 0.1e-9 / cprho
    assert np.abs(T - (300.0 + 11 * rise)).max() <= 1e-7
    s.compute(0.5)
    assert abs(s.time - 1.5) < 1e-12
    T = s.outTemperature(s.mesh)
    assert np.abs(T - (300.0 + 17 * rise)).max() <= 1e-7


@pytest.mark.parametrize("held", ["fixed", "reference"])
@pytest.mark.parametrize("theta", [0.0, 0.5, 1.0])
@pytest.mark.parametrize("cls", [Dynamic2D, DynamicCyl, Dynamic3D])
def test_solver_matches_oracle(cls, theta, held):
    """A held bottom, kept at 300 K or with the reference's update (conditions eliminated
    from A and F only, femT2d.cpp:243), and the capacities of setMatrix, against the
    oracle loop."""
    s, cprho = _slab(cls, held=True)
    s.timestep = 0.05
    s.methodparam = theta
    s.held_nodes = held
    s.compute(0.2)   # 5 steps
    T = s._core.solution()
    ax = [np.asarray(a) for a in s.mesh.axes]
    g = O.Grid(ax, cyl=(cls is DynamicCyl))
    E = g.nelems
    if cls is Dynamic3D:
        v = np.diff(ax[2])[:, None, None] * np.diff(ax[1])[None, :, None] * np.diff(ax[0])[None, None, :]
        cap = cprho * 0.125e-9 * v.ravel() / 0.05
    else:
        a = (np.diff(ax[1])[:, None] * np.diff(ax[0])[None, :]).ravel()
        cap = cprho * 0.25e-12 * a / 0.05 / 1e-9
        if cls is DynamicCyl:
            cap = cap * np.tile(0.5 * (ax[0][1:] + ax[0][:-1]), E // (len(ax[0]) - 1))
    dn = np.arange(len(ax[0]) * (len(ax[1]) if cls is Dynamic3D else 1))
    ref = O.thermal_dynamic(g, np.full(E, 45.0), np.full(E, 40.0), np.full(E, 2e15), cap, theta,
                            np.full(g.nnodes, 300.0), 5, dn, np.full(len(dn), 300.0), held=held)[-1]
    assert np.abs(T - ref).max() <= 1e-9 * np.abs(ref).max()
