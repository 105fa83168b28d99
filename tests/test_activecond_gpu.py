"""ActiveCond2D / ActiveCondCyl / ActiveCond3D (PythonCondSolver, solvers/electrical/
shockley/python/electr_python.cpp:113-190) on the GPU.

The reference's own test, solvers/electrical/shockley/tests/cond2d.py:33-56: a 0.2 um
junction between two 300 um conductors, 1 V, cond = 0.05 + |j| -> sigma = 0.1 S/m,
j = 0.05 kA/cm^2, 500 mA through 1000 x 1000 um^2 (to 3 decimals, as there).  The
cylindrical and 3D forms carry the same current density.  A cond callable that states
Shockley's law reproduces Shockley2D (beta.hpp:43-46) loop by loop.
"""
import math

import numpy as np
import pytest

from plask_amd import geometry, material, mesh
from plask_amd.electrical.shockley import ActiveCond2D, ActiveCond3D, ActiveCondCyl, Shockley2D

pytestmark = pytest.mark.gpu


@material.simple()
class Conductor12(material.Material):
    def cond(self, T):
        return (1e+12, 1e+12)

    def eps(self, T):
        return 1.


def _cond_solver(cls, jh=0.2):
    rect_args = (1000., 300.)
    if cls is ActiveCond3D:
        rect = geometry.Cuboid(1000., 1000., 300., Conductor12())
        junc = geometry.Cuboid(1000., 1000., jh, "GaAs")
        junc.role = "active"
        stack = geometry.Stack3D()
        space_cls = geometry.Cartesian3D
    else:
        rect = geometry.Rectangle(*rect_args, Conductor12())
        junc = geometry.Rectangle(1000., jh, "GaAs")
        junc.role = "active"
        stack = geometry.Stack2D()
        space_cls = None
    stack.append(rect)
    stack.append(junc)
    stack.append(rect)
    if cls is ActiveCond3D:
        space = space_cls(stack)
        gen = mesh.Rectangular3D.DivideGenerator()
        gen.prediv = 2, 2, 1
    elif cls is ActiveCondCyl:
        space = geometry.Cylindrical(stack)
        gen = mesh.Rectangular2D.DivideGenerator()
        gen.prediv = 2, 1
    else:
        space = geometry.Cartesian2D(stack, length=1000.)
        gen = mesh.Rectangular2D.DivideGenerator()
        gen.prediv = 2, 1
    s = cls("electrical")
    s.geometry = space
    s.mesh = gen
    s.maxerr = 1e-5
    s.voltage_boundary.append(s.mesh.Top(), 0.)
    s.voltage_boundary.append(s.mesh.Bottom(), 1.)
    return s


@pytest.mark.parametrize("cls,area_cm2", [(ActiveCond2D, 1e-2), (ActiveCondCyl, math.pi * 1e-2),
                                          (ActiveCond3D, 1e-2)])
def test_reference_cond2d(cls, area_cm2):
    s = _cond_solver(cls)
    s.cond = lambda U, j, T: 0.05 + abs(j)
    s.compute()
    correct_current = 50.0 * area_cm2 * 1e3   # 0.05 kA/cm^2 -> mA
    assert abs(s.get_total_current() - correct_current) < 0.5e-3 * max(1.0, correct_current / 500.)


def test_scalar_only_callable_and_pair_result():
    """A callable that only takes numbers (math, max) is called per element; an
    (in-plane, vertical) pair sets both components (electr_python.cpp:174-177)."""
    s1 = _cond_solver(ActiveCond2D)
    s1.cond = lambda U, j, T: max(0.05 + math.fabs(j), 1e-3)
    s1.compute()
    assert abs(s1.get_total_current() - 500.) < 0.5e-3
    s2 = _cond_solver(ActiveCond2D)
    s2.cond = lambda U, j, T: (2.0, 0.05 + abs(j))
    s2.compute()
    assert abs(s2.get_total_current() - 500.) < 0.5e-3
    msh = s2.mesh.elements.mesh
    geo = s2.geometry
    cond = np.asarray(s2.outConductivity(msh))
    act = np.array([geo.has_role("active", p) for p in msh])
    assert act.any()
    assert np.allclose(cond[act, 0], 2.0) and np.allclose(cond[act, 1], 0.1, rtol=1e-4)


def test_cond_attribute_rules():
    s = _cond_solver(ActiveCond2D)
    with pytest.raises(TypeError):
        s.cond = 5.0
    assert s.cond is None
    f = lambda U, j, T: 1.0   # noqa: E731
    s.cond1 = f
    assert s.cond1 is f
    with pytest.raises(IndexError):   # junction 0 has no law
        s.compute()
    with pytest.raises(AttributeError):
        s.beta


def test_shockley_law_as_callable_matches_shockley2d():
    """cond = 10 |j| beta h / ln(1e7 |j| / js + 1) through the host callback equals the
    device law of Shockley2D, loop by loop (same loop count, same potentials)."""
    def build(cls):
        s = _cond_solver(cls, jh=0.02)
        s.maxerr = 1e-5
        return s
    ref = build(Shockley2D)
    ref.beta, ref.js = 10.0, 1.0
    got = build(ActiveCond2D)
    got.cond = lambda U, j, T: 10. * np.abs(j) * 10.0 * 0.02 / np.log(1e7 * np.abs(j) / 1.0 + 1.)
    for loops in (1, 3, 0):
        e_ref = ref.compute(loops)
        e_got = got.compute(loops)
        assert ref._core.loopno == got._core.loopno
        assert abs(e_ref - e_got) <= 1e-6 * max(abs(e_ref), 1e-12) + 1e-9
    V_ref = ref._core.solution()
    V = got._core.solution()
    assert np.abs(V - V_ref).max() <= 1e-9
    assert abs(got.get_total_current() - ref.get_total_current()) <= 1e-9 * abs(ref.get_total_current())
    assert abs(ref.get_total_current() - 1e-3 * (math.exp(10.) - 1)) < 0.5e-3
