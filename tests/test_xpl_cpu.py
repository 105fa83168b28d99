"""XPL solver configuration (plask_amd.xpl) on the CPU: solver blocks written in the
reference's schema (solvers/thermal/static/solvers.yml, solvers/thermal/dynamic/solvers.yml,
solvers/electrical/shockley/solvers.yml, solvers/meta/shockley/solvers.yml) give the same
solver state as the same configuration written in Python."""
import numpy as np
import pytest

from plask_amd import geometry, material, mesh, xpl
from plask_amd.electrical.shockley import ActiveCond2D, ShockleyCyl
from plask_amd.meta.shockley import ThermoElectric2D
from plask_amd.thermal.dynamic import DynamicCyl
from plask_amd.thermal.static import Convection, Radiation, Static2D, Static3D


def _manager():
    cont = geometry.Rectangle(4., 1., material.Const(thermk=(50., 50.), cond=(1e6, 1e6), name="M"))
    junc = geometry.Rectangle(10., 0.1, "GaAs")
    junc.role = "active"
    body = geometry.Rectangle(10., 5., material.Const(thermk=(40., 30.), cond=(1e4, 1e4), name="B"))
    shelf = geometry.Shelf()
    shelf.append(cont)
    stack = geometry.Stack2D()
    stack.prepend(shelf)
    stack.prepend(junc)
    stack.prepend(body)
    g2 = geometry.Cylindrical(stack)
    g2c = geometry.Cartesian2D(stack)
    msh = mesh.Rectangular2D(np.linspace(0, 10, 11), np.r_[0.0, 2.5, 5.0, 5.1, 5.6, 6.1])
    return xpl.Manager(geometry={"main": g2, "flat": g2c}, mesh={"default": msh}, objects={"contact": cont},
                       defines={"T0": 290.0})


def _same_bc(a, b, msh, geo):
    """Same places (as node sets) with the same values, in the same order."""
    assert len(a) == len(b)
    for (pa, va), (pb, vb) in zip(a, b):
        assert np.array_equal(pa.nodes(msh, geo), pb.nodes(msh, geo))
        assert repr(va) == repr(vb)


THERMAL = """
<thermal name="THERM" solver="Static2D" lib="static">
  <geometry ref="flat"/>
  <mesh ref="default" empty-elements="exclude"/>
  <temperature>
    <condition place="bottom" value="{T0 + 10}"/>
  </temperature>
  <heatflux>
    <condition value="2e5"><place side="top" object="contact"/></condition>
  </heatflux>
  <convection>
    <condition place="right" coeff="10" ambient="300"/>
  </convection>
  <radiation>
    <condition emissivity="0.8" ambient="310" place="left"/>
  </radiation>
  <loop inittemp="305" maxerr="0.01"/>
  <matrix algorithm="iterative">
    <iterative accelerator="cg" preconditioner="ljac" noconv="error" maxit="500" maxerr="1e-9" nfact="4"/>
  </matrix>
</thermal>
"""


def test_thermal_block_matches_python():
    m = _manager()
    s = xpl.configure(Static2D("THERM"), THERMAL, m)
    ref = Static2D("THERM")
    ref.geometry = m.geo["flat"]
    ref.mesh = m.msh["default"]
    ref.empty_elements = "exclude"
    ref.temperature_boundary.append(ref.mesh.Bottom(), 300.0)
    ref.heatflux_boundary.append(ref.mesh.TopOf(m.obj["contact"]), 2e5)
    ref.convection_boundary.append(ref.mesh.Right(), Convection(10., 300.))
    ref.radiation_boundary.append(ref.mesh.Left(), Radiation(0.8, 310.))
    assert s.geometry is ref.geometry and s.mesh is ref.mesh and s.empty_elements == "exclude"
    assert s.inittemp == 305.0 and s.maxerr == 0.01 and s.algorithm == "iterative"
    it = s.iterative
    assert (it.accelerator, it.preconditioner, it.noconv, it.maxit, it.maxerr, it.nfact) == \
        ("cg", "ljac", "error", 500, 1e-9, 4)
    assert it.precond == 2    # ljac: the line preconditioner
    for kind in ("temperature", "heatflux", "convection", "radiation"):
        _same_bc(getattr(s, kind + "_boundary"), getattr(ref, kind + "_boundary"), ref.mesh, ref.geometry)
    top = s.heatflux_boundary[0].place.nodes(ref.mesh, ref.geometry)
    assert len(top) == 5     # the contact's top edge, x = 0..4


ELECTRICAL = """
<solvers>
  <electrical name="ELEC" solver="ShockleyCyl" lib="shockley">
    <geometry ref="main"/>
    <mesh ref="default"/>
    <voltage>
      <condition value="1.5" placename="ctop"><place side="top" object="contact"/></condition>
      <condition placeref="ctop" value="2.0"/>
      <condition value="0"><place side="bottom"/></condition>
    </voltage>
    <junction beta0="18" js0="0.5" js1="2" beta1="{T0 / 29}"/>
    <loop start-cond="0.1,3" convergence="stable" maxerr="0.02"/>
    <contacts pcond="7" ncond="70"/>
    <matrix><iterative maxerr="1e-11"/></matrix>
  </electrical>
  <thermal name="DYN" solver="DynamicCyl" lib="dynamic">
    <geometry ref="main"/>
    <mesh ref="default"/>
    <temperature><condition place="bottom" value="300"/></temperature>
    <loop inittemp="301" timestep="0.05" rebuildfreq="3" logfreq="20"/>
    <matrix methodparam="1" lumping="yes" algorithm="cholesky"/>
  </thermal>
</solvers>
"""


def test_electrical_and_dynamic_blocks():
    m = _manager()
    got = xpl.load_solvers(ELECTRICAL, m)
    e, d = got["ELEC"], got["DYN"]
    assert isinstance(e, ShockleyCyl) and isinstance(d, DynamicCyl)
    assert (e.beta0, e.js0, e.js1, e.beta1) == (18.0, 0.5, 2.0, 10.0)
    assert e.start_cond == (0.1, 3.0) and e.convergence == "stable" and e.maxerr == 0.02
    assert (e.pcond, e.ncond) == (7.0, 70.0) and e.iterative.maxerr == 1e-11
    ref = ShockleyCyl("ELEC")
    ref.geometry, ref.mesh = m.geo["main"], m.msh["default"]
    top = ref.mesh.TopOf(m.obj["contact"])
    ref.voltage_boundary.append(top, 1.5)
    ref.voltage_boundary.append(top, 2.0)
    ref.voltage_boundary.append(ref.mesh.Bottom(), 0.0)
    _same_bc(e.voltage_boundary, ref.voltage_boundary, ref.mesh, ref.geometry)
    assert m.boundaries["ctop"] is e.voltage_boundary[0].place
    assert (d.inittemp, d.timestep, d.rebuildfreq, d.logfreq, d.methodparam, d.lumping, d.algorithm) == \
        (301.0, 0.05, 3, 20, 1.0, True, "cholesky")


META = """
<meta name="TE" solver="ThermoElectric2D" lib="shockley">
  <geometry thermal="flat" electrical="flat"/>
  <mesh thermal="default" electrical="default" empty-elements="include"/>
  <junction beta="12" js="1.5"/>
  <contacts pcond="9"/>
  <loop tfreq="4" inittemp="295" maxterr="0.2" maxcerr="0.3"/>
  <tmatrix algorithm="gauss"/>
  <ematrix><iterative maxit="999"/></ematrix>
  <temperature><condition place="bottom" value="300"/></temperature>
  <voltage><condition place="top" value="1"/><condition place="bottom" value="0"/></voltage>
</meta>
"""


def test_meta_block():
    m = _manager()
    te = xpl.configure(ThermoElectric2D("TE"), META, m)
    th, el = te.thermal, te.electrical
    assert th.geometry is m.geo["flat"] and el.mesh is m.msh["default"] and el.empty_elements == "include"
    assert (el.beta, el.js, el.pcond, te.tfreq) == (12.0, 1.5, 9.0, 4)
    assert (th.inittemp, th.maxerr, el.maxerr) == (295.0, 0.2, 0.3)
    assert th.algorithm == "gauss" and el.iterative.maxit == 999
    assert len(th.temperature_boundary) == 1 and len(el.voltage_boundary) == 2


def test_plask_document_with_defines():
    doc = """<plask><defines><define name="U" value="0.5"/><define name="V" value="2 * U"/></defines>
    <solvers><electrical name="A" solver="ActiveCond2D" lib="shockley"><geometry ref="flat"/><mesh ref="default"/>
    <voltage><condition place="top" value="{V}"/></voltage><loop maxerr="1e-5"/></electrical></solvers></plask>"""
    m = _manager()
    m.defines = {}
    got = xpl.load_solvers(doc, m)
    a = got["A"]
    assert isinstance(a, ActiveCond2D) and a.maxerr == 1e-5
    assert a.voltage_boundary[0].value == 1.0


@pytest.mark.parametrize("block,msg", [
    ('<electrical name="x" solver="Shockley2D"><potential/></electrical>', "renamed to <voltage>"),
    ('<thermal name="x" solver="Static3D"><loop inittemp="abc"/></thermal>', "illegal value"),
    ('<thermal name="x" solver="Static3D"><bogus/></thermal>', "unexpected tag"),
    ('<electrical name="x" solver="Shockley2D"><loop start-cond="1,2" start-cond-inplane="3"/></electrical>',
     "two values"),
    ('<thermal name="x" solver="Static2D"><temperature><condition value="1"/></temperature></thermal>',
     "<place>"),
    ('<thermal name="x" solver="Static2D"><matrix><iterative preconditioner="zzz"/></matrix></thermal>',
     "unknown preconditioner"),
    ('<electrical name="x" solver="ActiveCond2D"><junction beta="1"/></electrical>', "unexpected tag"),
])
def test_errors(block, msg):
    el = xpl._element(block)
    cls = {"Shockley2D": xpl.Shockley2D, "Static3D": Static3D, "Static2D": Static2D,
           "ActiveCond2D": ActiveCond2D}[el.get("solver")]
    with pytest.raises(xpl.XMLError, match=msg):
        xpl.configure(cls(el.get("name")), el, xpl.Manager())


def test_line_and_side_of_places():
    """<place line=.../> takes the mesh line nearest 'at' from start to stop (widened by a
    node within 1% of an interval, rectangular_common.hpp:86-115); side-of places take
    the object's face line over its extent."""
    msh = mesh.Rectangular2D(np.array([0.0, 1.0, 2.0, 3.0]), np.array([0.0, 1.0, 2.0, 3.0, 4.0]))
    b = mesh.Rectangular.VerticalLineNear(1.9, 0.995, 3.0)
    assert list(b.nodes(msh)) == [2 + 4 * k for k in (1, 2, 3)]
    b = mesh.Rectangular.HorizontalLineNear(2.2, 0.5, 2.0)
    assert list(b.nodes(msh)) == [1 + 4 * 2, 2 + 4 * 2]
    r = geometry.Rectangle(2.0, 2.0, material.Const(thermk=1.0, name="R"))
    st = geometry.Stack2D()
    st.prepend(r)
    g = geometry.Cartesian2D(st)
    assert list(mesh.Rectangular.RightOf(r).nodes(msh, g)) == [2, 6, 10]
    assert list(mesh.Rectangular.LeftOf(r).nodes(msh, g)) == [0, 4, 8]
    assert list(mesh.Rectangular.TopOf(r).nodes(msh, g)) == [8, 9, 10]
