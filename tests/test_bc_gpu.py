"""Heat-flux, radiation and convection boundary conditions through thermal.static on the
GPU, against the CPU statement of the reference's boundary terms (oracle/bc.py) and band
Cholesky / dense Cholesky."""
import numpy as np
import pytest

import oracle as O
from oracle import bc as BC
from plask_amd import geometry, material, mesh
from plask_amd.thermal.static import Convection, Radiation, Static2D, Static3D, StaticCyl

pytestmark = pytest.mark.gpu


def _stack_2d(cls, cyl, seed=7, nx=17, nlayers=5, kscale=1.0):
    rng = np.random.default_rng(seed)
    stack = geometry.Stack2D()
    for i in range(nlayers):
        t = float(rng.uniform(0.3, 2.0))
        k = kscale * float(np.exp(rng.uniform(0, 4)))
        stack.append(geometry.Rectangle(15., t, material.Const(thermk=(k, 0.5 * k), name="B%d" % i)))
    s = cls("therm")
    s.geometry = geometry.Cylindrical(stack) if cyl else geometry.Cartesian2D(stack)
    s.mesh = mesh.Rectangular2D.DivideGenerator(prediv=(nx - 1, 3))
    s.temperature_boundary.append(s.mesh.Bottom(), 300.)
    s.iterative.maxerr = 1e-13
    return s


def _stack_3d(seed=8, shape=(7, 6, 9), kscale=1.0):
    rng = np.random.default_rng(seed)
    stack = geometry.Stack3D(xcenter=0.0, ycenter=0.0)
    ts = []
    for i in range(3):
        t = float(rng.uniform(0.5, 2.0))
        k = kscale * float(np.exp(rng.uniform(0, 4)))
        stack.append(geometry.Cuboid(12., 10., t, material.Const(thermk=(k, 0.5 * k), name="C%d" % i)))
        ts.append(t)
    z = np.concatenate([[0.0], np.cumsum(ts)])
    zz = np.unique(np.concatenate([np.linspace(z[i], z[i + 1], 4) for i in range(3)]))
    s = Static3D("therm3")
    s.geometry = geometry.Cartesian3D(stack)
    s.mesh = mesh.Rectangular3D(np.linspace(-6., 6., shape[0]), np.linspace(-5., 5., shape[1]), zz)
    s.temperature_boundary.append(s.mesh.Bottom(), 300.)
    s.iterative.maxerr = 1e-13
    return s


def _oracle_system(s, cyl=False):
    msh = s.mesh
    g = O.Grid(msh.axes, cyl=cyl)
    k = np.array([m.thermk(300.) for m in s._mat])
    k = k if k.ndim == 2 else np.column_stack([k, k])
    dn, dv = s.temperature_boundary.resolve(msh, s.geometry)
    A = O.assemble(g, 0, k[:, 0], k[:, 1], np.zeros(g.nelems), O.ST_BAND, dn, dv)
    return A, dn


def _nodes(s, place):
    return place.nodes(s.mesh, s.geometry)


@pytest.mark.parametrize("cls,cyl", [(Static2D, False), (StaticCyl, True)])
def test_heat_flux_2d_vs_oracle(cls, cyl):
    s = _stack_2d(cls, cyl)
    print("heat flux 2d", cls.__name__, flush=True)
    top, right = s.mesh.Top(), s.mesh.Right()
    s.heatflux_boundary.append(top, 3.0e5)
    s.heatflux_boundary.append(right, -1.0e5)   # the shared corner keeps the first value
    s.initialize()
    s.compute(1)
    T = s._core.solution()
    A, dn = _oracle_system(s, cyl)
    F = BC.load_2d(s.mesh.axes, flux=BC.first_values([(_nodes(s, top), 3.0e5), (_nodes(s, right), -1.0e5)]),
                   cyl=cyl)
    T_ref = O.band_cholesky_solve(A, BC.with_load(A.B, F, dn))
    assert np.abs(T - T_ref).max() <= 1e-9 * np.abs(T_ref).max()
    assert np.abs(T_ref - 300.).max() > 0.1   # the load matters


def test_heat_flux_3d_vs_oracle():
    s = _stack_3d()
    print("heat flux 3d", flush=True)
    top, left = s.mesh.Top(), s.mesh.Left()
    s.heatflux_boundary.append(top, -2.0e6)
    s.heatflux_boundary.append(left, 5.0e5)
    s.initialize()
    s.compute(1)
    T = s._core.solution()
    A, dn = _oracle_system(s)
    F = BC.load_3d(s.mesh.axes, flux=BC.first_values([(_nodes(s, top), -2.0e6), (_nodes(s, left), 5.0e5)]))
    T_ref = O.band_cholesky_solve(A, BC.with_load(A.B, F, dn))
    assert np.abs(T - T_ref).max() <= 1e-9 * np.abs(T_ref).max()
    assert np.abs(T_ref - 300.).max() > 0.1


def _picard(A, dn, load, T0, loops):
    T = T0.copy()
    for _ in range(loops):
        T = O.band_cholesky_solve(A, BC.with_load(A.B, load(T), dn))
    return T


@pytest.mark.parametrize("cls,cyl", [(Static2D, False), (StaticCyl, True)])
def test_radiation_2d_picard_vs_oracle(cls, cyl):
    s = _stack_2d(cls, cyl, kscale=0.02)
    print("radiation 2d", cls.__name__, flush=True)
    top = s.mesh.Top()
    s.radiation_boundary.append(top, Radiation(0.9, 900.))
    s.maxerr = 0.0   # run every loop
    s.initialize()
    s.compute(3)
    T = s._core.solution()
    A, dn = _oracle_system(s, cyl)
    rad = BC.first_values([(_nodes(s, top), (0.9, 900.))])
    T_ref = _picard(A, dn, lambda t: BC.load_2d(s.mesh.axes, radiation=rad, T=t, cyl=cyl),
                    np.full(A.rank, 300.), 3)
    assert np.abs(T - T_ref).max() <= 1e-9 * np.abs(T_ref).max()
    assert np.abs(T_ref - 300.).max() > 1.0


def test_radiation_3d_picard_vs_oracle():
    s = _stack_3d(kscale=0.02)
    print("radiation 3d", flush=True)
    top = s.mesh.Top()
    s.radiation_boundary.append(top, {"emissivity": 0.7, "ambient": 1200.})
    s.maxerr = 0.0
    s.initialize()
    s.compute(2)
    T = s._core.solution()
    A, dn = _oracle_system(s)
    rad = BC.first_values([(_nodes(s, top), (0.7, 1200.))])
    T_ref = _picard(A, dn, lambda t: BC.load_3d(s.mesh.axes, radiation=rad, T=t), np.full(A.rank, 300.), 2)
    assert np.abs(T - T_ref).max() <= 1e-9 * np.abs(T_ref).max()
    assert np.abs(T_ref - 300.).max() > 1.0


def test_boundary_loads_clear():
    s = _stack_2d(Static2D, False)
    s.initialize()
    s.compute(1)
    T0 = s._core.solution()
    s.heatflux_boundary.append(s.mesh.Top(), 1.0e6)
    s.compute(1)
    assert np.abs(s._core.solution() - T0).max() > 0.1
    s.heatflux_boundary.clear()
    s.compute(1)
    assert np.abs(s._core.solution() - T0).max() <= 1e-9 * 300.


def _oracle_parts(s, cyl=False):
    """Unconstrained stiffness and load of a solver's mesh (element conductivities at 300 K)."""
    msh = s.mesh
    g = O.Grid(msh.axes, cyl=cyl)
    k = np.array([m.thermk(300.) for m in s._mat])
    k = k if k.ndim == 2 else np.column_stack([k, k])
    A0 = O.assemble(g, 0, k[:, 0], k[:, 1], np.zeros(g.nelems), O.ST_BAND)
    dn, dv = s.temperature_boundary.resolve(msh, s.geometry)
    return A0, dn, dv


@pytest.mark.parametrize("cls,cyl", [(Static2D, False), (StaticCyl, True)])
def test_convection_2d_vs_oracle(cls, cyl):
    """Convection on the top and the right side (first place wins at the corner), a flux
    on the left: the operator's boundary matrix and the loads against the oracle's
    statement of therm2d.cpp:242-254 / :370-392 (dense, eliminated)."""
    s = _stack_2d(cls, cyl)
    print("convection 2d", cls.__name__, flush=True)
    top, right, left = s.mesh.Top(), s.mesh.Right(), s.mesh.Left()
    s.convection_boundary.append(top, Convection(40., 350.))
    s.convection_boundary.append(right, {"coeff": 5., "ambient": 280.})
    s.heatflux_boundary.append(left, 2.0e5)
    s.initialize()
    s.compute(1)
    T = s._core.solution()
    A0, dn, dv = _oracle_parts(s, cyl)
    conv = BC.first_values([(_nodes(s, top), (40., 350.)), (_nodes(s, right), (5., 280.))])
    Fc, Kc = BC.convection_2d(s.mesh.axes, conv, cyl=cyl)
    Ff = BC.load_2d(s.mesh.axes, flux=BC.first_values([(_nodes(s, left), 2.0e5)]), cyl=cyl)
    T_ref = BC.solve_with_matrix(A0.dense(), A0.B, Kc, Fc + Ff, dn, dv)
    assert np.abs(T - T_ref).max() <= 1e-9 * np.abs(T_ref).max()
    assert np.abs(T_ref - 300.).max() > 1.0


def test_convection_3d_vs_oracle():
    """Convection on the top (whose terms the reference places on the plane below) and on
    the y-low side wall, against the oracle (therm3d.cpp:250-259)."""
    s = _stack_3d()
    print("convection 3d", flush=True)
    top, front = s.mesh.Top(), s.mesh.Front()
    s.convection_boundary.append(top, Convection(3e5, 400.))
    s.convection_boundary.append(front, Convection(1e5, 250.))
    s.initialize()
    s.compute(1)
    T = s._core.solution()
    A0, dn, dv = _oracle_parts(s)
    conv = BC.first_values([(_nodes(s, top), (3e5, 400.)), (_nodes(s, front), (1e5, 250.))])
    Fc, Kc = BC.convection_3d(s.mesh.axes, conv)
    T_ref = BC.solve_with_matrix(A0.dense(), A0.B, Kc, Fc, dn, dv)
    assert np.abs(T - T_ref).max() <= 1e-9 * np.abs(T_ref).max()
    assert np.abs(T_ref - 300.).max() > 1.0


def test_convection_2d_closed_form_through_the_solver():
    """The 2D slab of the reference's equations (matrix without the 1e-6 of the load):
    T = T0 + g z, g = (1e-6 h T_amb - h T0) / (k + h L); 2.97 K on top for k = 1, L = 10,
    h = 10, T0 = T_amb = 300."""
    stack = geometry.Stack2D()
    stack.append(geometry.Rectangle(4., 10., material.Const(thermk=(1.0, 1.0), name="K1")))
    s = Static2D("slab")
    s.geometry = geometry.Cartesian2D(stack)
    s.mesh = mesh.Rectangular2D(np.array([0.0, 0.7, 2.0, 4.0]), np.array([0.0, 1.5, 4.0, 7.0, 10.0]))
    s.temperature_boundary.append(s.mesh.Bottom(), 300.)
    s.convection_boundary.append(s.mesh.Top(), Convection(10., 300.))
    s.iterative.maxerr = 1e-13
    s.compute(1)
    T = s._core.solution()
    g = (1e-6 * 10. * 300. - 10. * 300.) / (1. + 10. * 10.)
    z = np.repeat(s.mesh.axes[1], 4)
    assert np.allclose(T, 300. + g * z, rtol=0, atol=1e-8 * 300.)


def test_apply_with_boundary_matrix_vs_oracle():
    """pf_boundary_matrix_set through the C ABI: pf_apply equals the dense eliminated
    operator K + Kc, and a solve matches the oracle, with both preconditioners."""
    from plask_amd.fem import PRECOND_LINE, PRECOND_MG, FemContext
    axes = _grid_axes_3d()
    g = O.Grid(axes)
    E = g.nelems
    rng = np.random.default_rng(4)
    kh, kv = rng.uniform(1, 50, E), rng.uniform(1, 50, E)
    A0 = O.assemble(g, 0, kh, kv, np.zeros(E), O.ST_BAND)
    n0, n1, n2 = (len(a) for a in axes)
    plane = n0 * n1
    dn = np.arange(plane)
    dv = np.full(plane, 300.)
    side = np.array([i + n0 * (j + n1 * k) for k in range(n2) for j in range(n1) for i in (n0 - 1,)])
    conv = BC.first_values([(np.arange((n2 - 1) * plane, n2 * plane), (2e5, 350.)), (side, (5e4, 280.))])
    Fc, Kc = BC.convection_3d(axes, conv)
    c = FemContext()
    c.set_grid(axes)
    c.set_coeffs(kh, kv)
    c.set_source(None)
    c.set_dirichlet(dn, dv)
    rows, cols, vals = (np.array([e[i] for e in Kc]) for i in range(3))
    c.set_boundary_matrix(rows, cols, vals)
    nz = np.nonzero(Fc)[0]
    c.set_load(nz, Fc[nz])
    K = A0.dense()
    for r, cc, v in Kc:
        K[r, cc] += v
        if r != cc:
            K[cc, r] += v
    free = np.ones(g.nnodes, bool)
    free[dn] = False
    x = rng.uniform(-1, 1, g.nnodes)
    y = c.apply(x)
    xm = np.where(free, x, 0.0)
    want = np.where(free, K @ xm, x)
    assert np.abs(y - want).max() <= 1e-12 * np.abs(want).max()
    T_ref = BC.solve_with_matrix(A0.dense(), A0.B, Kc, Fc, dn, dv)
    for pc in (PRECOND_LINE, PRECOND_MG):
        T, st = c.solve(x0=np.full(g.nnodes, 300.), rtol=1e-12, precond=pc)
        assert np.abs(T - T_ref).max() <= 1e-9 * np.abs(T_ref).max(), (pc, st.iters)
    # clearing the matrix restores the plain operator
    c.set_boundary_matrix([], [], [])
    y0 = c.apply(x)
    assert np.abs(y0 - np.where(free, A0.dense() @ xm, x)).max() <= 1e-12 * np.abs(y0).max()
    c.close()


def _grid_axes_3d():
    rng = np.random.default_rng(11)
    return [np.cumsum(np.r_[0.0, rng.uniform(0.3, 1.5, n - 1)]) for n in (9, 7, 8)]


def test_condition_values():
    r = Radiation.of({"emissivity": 0.5})
    assert r["ambient"] == 300.0 and r.emissivity == 0.5
    c = Convection.of((4.0, 350.0))
    c["coeff"] = 5.0
    assert c.coeff == 5.0 and c["ambient"] == 350.0
    assert repr(c) == "{'coeff': 5.0, 'ambient': 350.0}"
    with pytest.raises(KeyError):
        c["emissivity"]


def test_load_set_through_the_c_abi():
    """pf_load_set: repeated nodes are summed into the right-hand side, Dirichlet rows
    keep their values, and an empty list removes the load."""
    from plask_amd.fem import FemContext
    axes = [np.linspace(0, 4, 5), np.linspace(0, 3, 4), np.linspace(0, 2, 6)]
    E = 4 * 3 * 5
    c = FemContext()
    c.set_grid(axes)
    c.set_coeffs(np.full(E, 3.0), np.full(E, 1.5))
    c.set_source(np.full(E, 1e15))
    dn = np.arange(20)
    c.set_dirichlet(dn, np.full(20, 300.0))
    b0 = c.rhs()
    c.set_load([45, 61, 45, 3], [1.0, 2.0, 0.5, 7.0])
    d = c.rhs() - b0
    want = np.zeros_like(d)
    want[45], want[61] = 1.5, 2.0   # node 3 is a Dirichlet node
    assert np.allclose(d, want, rtol=0, atol=1e-12)
    assert d[3] == 0.0
    c.set_load([], [])
    assert np.array_equal(c.rhs(), b0)
    c.close()
