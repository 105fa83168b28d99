"""Device post-processing and providers (pf_postprocess, pf_interp_elements) on the GPU.

* heat flux (saveHeatFluxes, solvers/thermal/static/therm2d.cpp:504-540,
  therm3d.cpp:342-384): in a layered slab with a flux through one face and the opposite
  face held the Q1 solution is piecewise linear, so every element's flux equals the
  applied one whatever the layers' conductivities -- for device laws and for
  conductivities handed over (thickness-dependent materials, evaluated with the leaf
  height as the reference does);
* element fields at points: linear fields are reproduced exactly between midpoints and
  clamped to the outermost midpoint up to the mesh boundary (against SciPy's
  RegularGridInterpolator on clamped points), fill outside;
* outCurrentDensity / outHeat of the Shockley solver against the oracle's element
  currents and Joule heat (electr2d.cpp:395-464) of the same potential.
"""
import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator

import oracle as O
from plask_amd import geometry, material, mesh
from plask_amd.fem import PF_POST_CURRENT, PF_POST_FLUX, PF_POST_JOULE, FemContext
from plask_amd.thermal.static import Static2D, Static3D, StaticCyl

pytestmark = pytest.mark.gpu


@material.simple()
class ThickDependent(material.Material):
    """k depends on the thickness argument: the flux must use the leaf height."""
    def thermk(self, T, h=float("inf")):
        return (3.0 + 0.5 * min(h, 10.0), 2.0 + 0.25 * min(h, 10.0))


def _slab(cls, law=True):
    ts = [0.7, 1.9, 0.4, 1.2]
    ks = [(30., 20.), (3., 1.5), (60., 45.), (8., 2.)]
    if cls is Static3D:
        stack = geometry.Stack3D()
        for t, k in zip(ts, ks):
            m = material.Const(thermk=k, name="S%g" % t) if law else ThickDependent()
            stack.append(geometry.Cuboid(5., 4., t, m))
        s = cls("slab3")
        s.geometry = geometry.Cartesian3D(stack)
        s.mesh = mesh.Rectangular3D.DivideGenerator(prediv=(4, 3, 3))
    else:
        stack = geometry.Stack2D()
        for t, k in zip(ts, ks):
            m = material.Const(thermk=k, name="S%g" % t) if law else ThickDependent()
            stack.append(geometry.Rectangle(6., t, m))
        s = cls("slab2")
        s.geometry = geometry.Cylindrical(stack) if cls is StaticCyl else geometry.Cartesian2D(stack)
        s.mesh = mesh.Rectangular2D.DivideGenerator(prediv=(5, 3))
    s.temperature_boundary.append(s.mesh.Bottom(), 300.)
    s.iterative.maxerr = 1e-13
    return s


@pytest.mark.parametrize("law", [True, False])
@pytest.mark.parametrize("cls", [Static2D, Static3D])
def test_heat_flux_is_the_applied_flux(cls, law):
    """2D: a flux v (W/m^2, outward positive) on the top, the bottom held: -k dT/dz = v in
    every element.  3D: the reference puts a wall's flux on the element's local nodes 0-3
    (therm3d.cpp:160-163), which is the geometric face only for a bottom wall, so there the
    flux enters at the bottom and leaves through the held top: -k dT/dz = -v."""
    v = -2.0e5   # heat into the slab
    s = _slab(cls, law)
    print("flux", cls.__name__, "law", law, flush=True)
    if cls is Static3D:
        s.temperature_boundary.clear()
        s.temperature_boundary.append(s.mesh.Top(), 300.)
        s.heatflux_boundary.append(s.mesh.Bottom(), v)
        want = -v
    else:
        s.heatflux_boundary.append(s.mesh.Top(), v)
        want = v
    s.compute(1)
    em = s.mesh.element_mesh()
    q = s.outHeatFlux(em)
    assert q.shape == (len(em), s.DIM)
    assert np.abs(q[:, -1] - want).max() <= 1e-6 * abs(v)
    assert np.abs(q[:, :-1]).max() <= 1e-6 * abs(v)
    # the same at arbitrary points (device interpolation of the element field)
    rng = np.random.default_rng(1)
    pts = rng.uniform(0.0, 1.0, (200, s.DIM)) * np.array([a[-1] for a in s.mesh.axes])
    qp = s.outHeatFlux(pts)
    assert np.abs(qp[:, -1] - want).max() <= 1e-6 * abs(v)


def test_heat_flux_matches_the_gradient_through_the_c_abi():
    """pf_postprocess(FLUX) with given conductivities: -k times the corner-difference
    gradient (x 1e6 um -> m), checked on a linear temperature field, where it is exact."""
    axes = [np.array([0.0, 0.5, 1.7, 2.0]), np.array([0.0, 1.0, 1.5]), np.array([0.0, 0.3, 1.0, 2.2])]
    g = [1.5, -0.7, 2.5]   # K/um
    X = np.stack(np.meshgrid(*axes, indexing="ij"), -1).transpose(2, 1, 0, 3).reshape(-1, 3)
    T = 300.0 + X @ np.array(g)
    c = FemContext()
    c.set_grid(axes)
    E = 3 * 2 * 3
    kh = np.linspace(1.0, 5.0, E)
    kv = np.linspace(2.0, 3.0, E)
    c.set_coeffs(kh, kv)
    c.set_solution(T)
    q = c.heat_flux(kh, kv)
    want = -1e6 * np.column_stack([kh * g[0], kh * g[1], kv * g[2]])
    assert np.allclose(q, want, rtol=1e-12, atol=0)
    # no arrays: the conductivities as set
    assert np.allclose(c.heat_flux(), want, rtol=1e-12, atol=0)
    c.close()


def test_element_interpolation_against_scipy():
    axes = [np.array([0.0, 0.4, 1.0, 2.5, 3.0]), np.array([0.0, 0.9, 1.2, 3.0])]
    c = FemContext()
    c.set_grid(axes)
    E = 4 * 3
    rng = np.random.default_rng(2)
    c.set_coeffs(np.ones(E), np.ones(E))
    T = rng.uniform(290.0, 310.0, 20)
    c.set_solution(T)
    q = c.heat_flux(np.linspace(1, 2, E), np.linspace(3, 4, E))
    pts = np.column_stack([rng.uniform(-0.5, 3.5, 400), rng.uniform(-0.5, 3.5, 400)])
    got = c.interp_elements(PF_POST_FLUX, pts, fill=-7.0)
    mids = [0.5 * (a[1:] + a[:-1]) for a in axes]
    clipped = np.column_stack([np.clip(pts[:, a], mids[a][0], mids[a][-1]) for a in range(2)])
    inside = np.all([(pts[:, a] >= axes[a][0]) & (pts[:, a] <= axes[a][-1]) for a in range(2)], axis=0)
    for comp in range(2):
        f = RegularGridInterpolator(mids, q[:, comp].reshape(3, 4).T, method="linear")
        want = np.where(inside, f(clipped), -7.0)
        assert np.allclose(got[:, comp], want, rtol=1e-13, atol=1e-9)
    assert (~inside).any() and inside.any()
    c.close()


def test_shockley_fields_against_the_oracle():
    from test_solvers_gpu import _shockley2d
    from plask_amd.electrical.shockley import Shockley2D
    s = _shockley2d(Shockley2D, False)
    s.iterative.maxerr = 1e-13
    s.compute(3)
    V = s._core.solution()
    g = O.Grid(s.mesh.axes)
    sh = s._core.conductivity_h()
    sv = s._core.conductivity_v()
    cur, heat = O.currents(g, sh, sv, V)
    em = s.mesh.element_mesh()
    j = s.outCurrentDensity(em)
    h = s.outHeat(em)
    assert np.abs(j - cur.reshape(-1, 2)).max() <= 1e-9 * np.abs(cur).max()
    assert np.abs(h - heat).max() <= 1e-9 * np.abs(heat).max()
    pts = em.points()[::7] + 1e-3
    jp = s.outCurrentDensity(pts)
    mids = [np.asarray(a) for a in em.axes]
    f = RegularGridInterpolator(mids, cur.reshape(len(mids[1]), len(mids[0]), 2).transpose(1, 0, 2),
                                method="linear")
    clipped = np.column_stack([np.clip(pts[:, a], mids[a][0], mids[a][-1]) for a in range(2)])
    assert np.abs(jp - f(clipped)).max() <= 1e-9 * np.abs(cur).max()
    hp = s.outHeat(pts)
    fh = RegularGridInterpolator(mids, heat.reshape(len(mids[1]), len(mids[0])).T, method="linear")
    assert np.abs(hp - fh(clipped)).max() <= 1e-9 * np.abs(heat).max()
    # the same fields through the C ABI's post-processing entry
    got = s._core.postprocess(PF_POST_CURRENT, None, None, True).reshape(-1, 2)
    assert np.abs(got - cur.reshape(-1, 2)).max() <= 1e-9 * np.abs(cur).max()
    assert np.abs(s._core.postprocess(PF_POST_JOULE, None, None, True) - heat).max() <= 1e-9 * np.abs(heat).max()
