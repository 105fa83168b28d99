"""Device interpolation of the solution at arbitrary points (pf_interp_nodes, used by
outTemperature / outVoltage for general destination meshes) against the host
interpolation (SciPy RegularGridInterpolator) of the same node values."""
import numpy as np
import pytest

from plask_amd import geometry, material, mesh
from plask_amd import solverbase as sb
from plask_amd.fem import FemContext
from plask_amd.thermal.static import Static2D, Static3D, StaticCyl

pytestmark = pytest.mark.gpu


def _points(axes, n, rng):
    """Random points inside, outside, exactly on nodes and on the upper boundary."""
    lo = np.array([a[0] for a in axes])
    hi = np.array([a[-1] for a in axes])
    span = hi - lo
    inside = lo + rng.uniform(0, 1, (n, len(axes))) * span
    outside = lo + rng.uniform(-0.3, 1.3, (n // 4, len(axes))) * span
    nodes = np.stack([rng.choice(a, n // 4) for a in axes], axis=1)
    upper = np.tile(hi, (3, 1))
    return np.concatenate([inside, outside, nodes, upper])


@pytest.mark.parametrize("shape", [(9, 7), (5, 11, 6)])
def test_interp_matches_host(shape):
    rng = np.random.default_rng(sum(shape))
    axes = [np.cumsum(np.r_[0.0, rng.uniform(0.1, 2.0, n - 1)]) for n in shape]
    N = int(np.prod(shape))
    E = int(np.prod([n - 1 for n in shape]))
    c = FemContext()
    c.set_grid(axes)
    c.set_coeffs(np.full(E, 2.0), np.full(E, 1.0))
    c.set_source(np.zeros(E))
    c.set_dirichlet(np.arange(shape[0]), np.zeros(shape[0]))
    vals = rng.uniform(-5.0, 5.0, N)
    c.set_solution(vals)
    pts = _points(axes, 400, rng)
    got = c.interp(pts, fill=-7.0)
    want = sb.interp_nodes(axes, vals, pts.tolist(), fill=-7.0)
    assert np.allclose(got, want, rtol=0, atol=1e-12 * np.abs(vals).max())
    assert (got == -7.0).sum() == (want == -7.0).sum() > 0
    c.close()


@pytest.mark.parametrize("cls", [Static2D, StaticCyl, Static3D])
def test_provider_uses_device_interpolation(cls):
    """outTemperature on a general mesh equals the host interpolation of the solution."""
    rng = np.random.default_rng(5)
    if cls is Static3D:
        stack = geometry.Stack3D(xcenter=0.0, ycenter=0.0)
        stack.append(geometry.Cuboid(8., 6., 1.5, material.Const(thermk=(3.0, 2.0), name="A")))
        stack.append(geometry.Cuboid(8., 6., 0.7, material.Const(thermk=(30.0, 20.0), name="B")))
        s = cls("t3")
        s.geometry = geometry.Cartesian3D(stack)
        s.mesh = mesh.Rectangular3D(np.linspace(-4, 4, 7), np.linspace(-3, 3, 6), np.linspace(0, 2.2, 9))
        dest = mesh.Rectangular3D(np.linspace(-5, 5, 13), np.linspace(-2.5, 2.5, 4), np.linspace(0.1, 2.4, 7))
    else:
        stack = geometry.Stack2D()
        stack.append(geometry.Rectangle(8., 1.5, material.Const(thermk=(3.0, 2.0), name="A")))
        stack.append(geometry.Rectangle(8., 0.7, material.Const(thermk=(30.0, 20.0), name="B")))
        s = cls("t2")
        s.geometry = geometry.Cylindrical(stack) if cls is StaticCyl else geometry.Cartesian2D(stack)
        s.mesh = mesh.Rectangular2D(np.linspace(0, 8, 9), np.linspace(0, 2.2, 8))
        dest = mesh.Rectangular2D(np.linspace(-1, 9, 23), np.linspace(0.05, 2.5, 11))
    s.temperature_boundary.append(s.mesh.Bottom(), 300.)
    em = s.mesh.element_mesh().points()
    s.inHeat = rng.uniform(0, 5e15, len(em))
    s.iterative.maxerr = 1e-12
    s.compute(1)
    got = s.outTemperature(dest)
    want = sb.interp_nodes(s._axes, s._core.solution(), dest.points().tolist(), fill=300.0)
    assert np.abs(got - want).max() <= 1e-12 * 400.0
    assert np.abs(got - 300.0).max() > 1e-3
