"""XPL solver configuration: the ``<solvers>`` section of a PLaSK .xpl file.

``load_solvers(source, manager)`` reads a ``<solvers>`` element (or a whole ``<plask>``
document, whose ``<defines>`` are evaluated first) and returns {name: solver} for the
solver classes of this package; ``configure(solver, element, manager)`` applies one
solver element to an existing solver.  The tags and attributes are the reference's:

* thermal.static (ThermalFem2DSolver / ThermalFem3DSolver::loadConfiguration,
  solvers/thermal/static/therm2d.cpp:40-68, therm3d.cpp:40-68): ``<temperature>``,
  ``<heatflux>``, ``<convection>``, ``<radiation>``, ``<loop inittemp maxerr>``;
* thermal.dynamic (femT2d.cpp:47-75): ``<temperature>``, ``<loop inittemp timestep
  rebuildfreq logfreq>``, ``<matrix methodparam lumping ...>``;
* electrical.shockley (ElectricalFem2DSolver::parseConfiguration, electr2d.cpp:41-86;
  BetaSolver::loadConfiguration, beta.cpp:31-63): ``<voltage>``, ``<loop start-cond
  start-cond-inplane convergence maxerr>``, ``<contacts pcond ncond>``, ``<junction beta
  js beta# js# pnjcond>`` (not for ActiveCond*);
* meta.shockley ThermoElectric* (thermoelectric.py:131-172): ``<geometry thermal
  electrical>``, ``<mesh thermal electrical empty-elements>``, ``<junction>``,
  ``<contacts>``, ``<loop tfreq inittemp maxterr maxcerr>``, ``<tmatrix>``, ``<ematrix>``
  and the five boundary-condition tags;
* every FEM solver (parseFemConfiguration, plask/common/fem/fem_solver.hpp:38-111):
  ``<matrix algorithm><iterative accelerator preconditioner noconv maxit maxerr nfact
  omega ndeg lvfill ltrunc nsave nrestart/></matrix>``, and ``<geometry ref>``,
  ``<mesh ref empty-elements>``.

Boundary conditions (Manager::readBoundaryConditions, plask/manager.hpp:705-737):
``<condition place="top" value="300"/>``, ``<condition placeref="name" .../>`` or
``<condition value="1"><place side="top" object="contact" path="p"/></condition>``, with
``placename`` storing the place; 2D also ``<place line="vertical|horizontal" at start
stop/>``.  Values: ``value``, or ``coeff`` + ``ambient`` (convection), ``emissivity`` +
``ambient`` (radiation) (solvers/thermal/static/common.hpp:50-58).

Attribute values may hold ``{expression}`` parts, evaluated with the defines (Python
expressions over earlier defines and the math module).  Geometries, meshes, geometry
objects and path hints come from the ``Manager`` (this package has no XPL geometry
reader): ``Manager(geometry={...}, mesh={...}, objects={...}, paths={...})``.
"""
import math
import re
import xml.etree.ElementTree as ET

from . import mesh as _mesh
from .electrical.shockley import (ActiveCond2D, ActiveCond3D, ActiveCondCyl, Shockley2D, Shockley3D,
                                  ShockleyCyl, _ActiveCondBase, _ShockleyBase)
from .meta.shockley import ThermoElectric, ThermoElectric2D, ThermoElectric3D, ThermoElectricCyl
from .thermal.dynamic import Dynamic2D, Dynamic3D, DynamicCyl, _DynamicBase
from .thermal.static import Convection, Radiation, Static2D, Static3D, StaticCyl, _StaticBase


class XMLError(ValueError):
    """Malformed or unsupported solver configuration."""


SOLVERS = {
    ("thermal", "static"): {c.__name__: c for c in (Static2D, StaticCyl, Static3D)},
    ("thermal", "dynamic"): {c.__name__: c for c in (Dynamic2D, DynamicCyl, Dynamic3D)},
    ("electrical", "shockley"): {c.__name__: c for c in (Shockley2D, ShockleyCyl, Shockley3D, ActiveCond2D,
                                                         ActiveCondCyl, ActiveCond3D)},
    ("meta", "shockley"): {c.__name__: c for c in (ThermoElectric2D, ThermoElectricCyl, ThermoElectric3D)},
}


class Manager:
    """What solver configurations refer to: geometries, meshes (or mesh generators),
    geometry objects and path hints by name, the defines and the named places."""

    def __init__(self, geometry=None, mesh=None, objects=None, paths=None, defines=None):
        self.geo = dict(geometry or {})
        self.msh = dict(mesh or {})
        self.obj = dict(objects or {})
        self.pth = dict(paths or {})
        self.defines = dict(defines or {})
        self.boundaries = {}


# -- attribute values ---------------------------------------------------------------
_EXPR = re.compile(r"\{([^{}]*)\}")


def _text(el, attr, manager):
    v = el.get(attr)
    if v is None:
        return None
    if "{" not in v:
        return v
    ns = {k: getattr(math, k) for k in dir(math) if not k.startswith("_")}
    ns.update(manager.defines)

    def sub(m):
        try:
            return repr(eval(m.group(1), {"__builtins__": {}}, ns))  # noqa: S307 -- user's own configuration
        except Exception as e:
            raise XMLError("<%s>: cannot evaluate '{%s}' in attribute '%s': %s" % (el.tag, m.group(1), attr, e))
    out = _EXPR.sub(sub, v)
    return out[1:-1] if (out.startswith("'") and out.endswith("'")) else out


def _float(el, attr, manager, required=False):
    v = _text(el, attr, manager)
    if v is None:
        if required:
            raise XMLError("<%s>: attribute '%s' is required" % (el.tag, attr))
        return None
    try:
        return float(v)
    except ValueError:
        raise XMLError("<%s>: attribute '%s' has illegal value '%s'" % (el.tag, attr, v))


def _int(el, attr, manager):
    v = _text(el, attr, manager)
    if v is None:
        return None
    try:
        return int(v)
    except ValueError:
        raise XMLError("<%s>: attribute '%s' has illegal value '%s'" % (el.tag, attr, v))


def _bool(el, attr, manager):
    v = _text(el, attr, manager)
    if v is None:
        return None
    t = v.strip().lower()
    if t in ("yes", "true", "1", "on"):
        return True
    if t in ("no", "false", "0", "off"):
        return False
    raise XMLError("<%s>: attribute '%s' has illegal value '%s'" % (el.tag, attr, v))


def _choice(el, attr, manager, choices):
    v = _text(el, attr, manager)
    if v is None:
        return None
    t = v.strip().lower()
    if t not in choices:
        raise XMLError("<%s>: attribute '%s' has illegal value '%s'" % (el.tag, attr, v))
    return t


def _check_attrs(el, allowed):
    for a in el.attrib:
        if a not in allowed:
            raise XMLError("<%s>: unexpected attribute '%s'" % (el.tag, a))


def _no_children(el):
    for ch in el:
        raise XMLError("<%s>: unexpected tag <%s>" % (el.tag, ch.tag))


# -- boundary conditions --------------------------------------------------------------
_SIDES = {"bottom": "Bottom", "top": "Top", "left": "Left", "right": "Right", "front": "Front", "back": "Back"}


def _place_from_tag(el, manager, dim):
    side = el.get("side")
    line = el.get("line")
    if side and line:
        raise XMLError("<place>: conflicting attributes 'side' and 'line'")
    if side:
        side = side.strip().lower()
        if side not in _SIDES or (dim == 2 and side in ("front", "back")):
            raise XMLError("<place>: attribute 'side' has illegal value '%s'" % side)
        obj = el.get("object")
        if obj is None:
            return getattr(_mesh.Rectangular, _SIDES[side])()
        if obj not in manager.obj:
            raise XMLError("<place>: no geometry object '%s'" % obj)
        path = el.get("path")
        if path is not None and path not in manager.pth:
            raise XMLError("<place>: no path '%s'" % path)
        return getattr(_mesh.Rectangular, _SIDES[side] + "Of")(manager.obj[obj],
                                                            manager.pth[path] if path is not None else None)
    if line:
        if dim != 2:
            raise XMLError("<place>: 'line' places are 2D")
        at = _float(el, "at", manager, True)
        start = _float(el, "start", manager, True)
        stop = _float(el, "stop", manager, True)
        if line == "vertical":
            return _mesh.Rectangular.VerticalLineNear(at, start, stop)
        if line == "horizontal":
            return _mesh.Rectangular.HorizontalLineNear(at, start, stop)
        raise XMLError("<place>: attribute 'line' has illegal value '%s'" % line)
    raise XMLError("<place>: 'side' or 'line' is required")


def _value(el, kind, manager):
    if kind == "convection":
        return Convection(_float(el, "coeff", manager, True), _float(el, "ambient", manager, True))
    if kind == "radiation":
        return Radiation(_float(el, "emissivity", manager, True), _float(el, "ambient", manager, True))
    return _float(el, "value", manager, True)


def read_conditions(el, dest, kind, manager, dim):
    """Append the <condition> children of el to the BoundaryConditions dest."""
    for cond in el:
        if cond.tag != "condition":
            raise XMLError("<%s>: unexpected tag <%s>" % (el.tag, cond.tag))
        value = _value(cond, kind, manager)
        place = cond.get("place")
        if place is not None:
            p = place.strip().lower()
            if p not in _SIDES or (dim == 2 and p in ("front", "back")):
                raise XMLError("<condition>: cannot parse boundary place from string \"%s\"" % place)
            boundary = getattr(_mesh.Rectangular, _SIDES[p])()
        elif cond.get("placeref") is not None:
            ref = cond.get("placeref")
            if ref not in manager.boundaries:
                raise XMLError("<condition>: no place named '%s'" % ref)
            boundary = manager.boundaries[ref]
        else:
            kids = list(cond)
            if len(kids) != 1 or kids[0].tag != "place":
                raise XMLError("<condition>: one <place> tag (or a 'place' / 'placeref' attribute) is required")
            boundary = _place_from_tag(kids[0], manager, dim)
        if cond.get("placename") is not None:
            manager.boundaries[cond.get("placename")] = boundary
        dest.append(boundary, value)


# -- matrix ---------------------------------------------------------------------------
_ITER_ATTRS = {"accelerator": str, "preconditioner": str, "noconv": str, "maxit": int, "maxerr": float,
               "nfact": int, "omega": float, "ndeg": int, "lvfill": int, "ltrunc": int, "nsave": int,
               "nrestart": int}


def _iterative(el, params, manager):
    _check_attrs(el, _ITER_ATTRS)
    for attr, typ in _ITER_ATTRS.items():
        if el.get(attr) is None:
            continue
        v = {str: _text, int: _int, float: _float}[typ](el, attr, manager)
        try:
            setattr(params, attr, v.lower() if typ is str else v)
        except ValueError as e:
            raise XMLError("<iterative>: %s" % e)
    _no_children(el)


def _matrix(el, solver, manager, extra=()):
    _check_attrs(el, ("algorithm",) + tuple(extra))
    alg = _choice(el, "algorithm", manager, ("cholesky", "gauss", "iterative"))
    if alg is not None:
        solver.algorithm = alg
    for ch in el:
        if ch.tag != "iterative":
            raise XMLError("<%s>: unexpected tag <%s>" % (el.tag, ch.tag))
        _iterative(ch, solver.iterative, manager)


# -- solvers --------------------------------------------------------------------------
def _geometry_mesh(el, solver, manager):
    if el.tag == "geometry":
        _check_attrs(el, ("ref",))
        ref = el.get("ref")
        if ref not in manager.geo:
            raise XMLError("<geometry>: no geometry '%s'" % ref)
        solver.geometry = manager.geo[ref]
        return True
    if el.tag == "mesh":
        _check_attrs(el, ("ref", "empty-elements"))
        ref = el.get("ref")
        if ref not in manager.msh:
            raise XMLError("<mesh>: no mesh '%s'" % ref)
        solver.mesh = manager.msh[ref]
        ee = _choice(el, "empty-elements", manager, ("default", "include", "exclude", "included", "excluded"))
        if ee is not None:
            solver.empty_elements = ee
        return True
    return False


def _configure_static(solver, el, manager):
    for t in el:
        if _geometry_mesh(t, solver, manager):
            continue
        if t.tag in ("temperature", "heatflux", "convection", "radiation"):
            read_conditions(t, getattr(solver, t.tag + "_boundary"), t.tag, manager, solver.DIM)
        elif t.tag == "loop":
            _check_attrs(t, ("inittemp", "maxerr"))
            v = _float(t, "inittemp", manager)
            if v is not None:
                solver.inittemp = v
            v = _float(t, "maxerr", manager)
            if v is not None:
                solver.maxerr = v
        elif t.tag == "matrix":
            _matrix(t, solver, manager)
        else:
            raise XMLError("<%s>: unexpected tag <%s>" % (el.tag, t.tag))


def _configure_dynamic(solver, el, manager):
    for t in el:
        if _geometry_mesh(t, solver, manager):
            continue
        if t.tag == "temperature":
            read_conditions(t, solver.temperature_boundary, "temperature", manager, solver.DIM)
        elif t.tag == "loop":
            _check_attrs(t, ("inittemp", "timestep", "rebuildfreq", "logfreq"))
            for attr, conv in (("inittemp", _float), ("timestep", _float), ("rebuildfreq", _int), ("logfreq", _int)):
                v = conv(t, attr, manager)
                if v is not None:
                    setattr(solver, attr, v)
        elif t.tag == "matrix":
            v = _float(t, "methodparam", manager)
            if v is not None:
                solver.methodparam = v
            v = _bool(t, "lumping", manager)
            if v is not None:
                solver.lumping = v
            _matrix(t, solver, manager, extra=("methodparam", "lumping"))
        else:
            raise XMLError("<%s>: unexpected tag <%s>" % (el.tag, t.tag))


def _junction(t, solver, manager):
    for attr in t.attrib:
        if attr in ("wavelength", "heat"):
            continue
        if attr == "pnjcond":
            solver.start_cond = _float(t, attr, manager)
            continue
        m = re.fullmatch(r"(beta|js)(\d*)", attr)
        if not m:
            raise XMLError("<junction>: unexpected attribute '%s'" % attr)
        setattr(solver, attr, _float(t, attr, manager))
    _no_children(t)


def _electrical_tag(t, solver, manager):
    """One tag of ElectricalFem2DSolver::parseConfiguration; False when not one of them."""
    if t.tag == "potential":
        raise XMLError("<potential> boundary conditions have been permanently renamed to <voltage>")
    if t.tag == "voltage":
        read_conditions(t, solver.voltage_boundary, "voltage", manager, solver.DIM)
    elif t.tag == "loop":
        _check_attrs(t, ("start-cond", "start-cond-inplane", "convergence", "maxerr"))
        if t.get("start-cond") is not None or t.get("start-cond-inplane") is not None:
            c0, c1 = solver.start_cond
            vert = _text(t, "start-cond", manager)
            if vert is not None:
                if "," not in vert:
                    c1 = float(vert)
                else:
                    if t.get("start-cond-inplane") is not None:
                        raise XMLError("<loop>: tag attribute 'start-cond' has two values, but attribute "
                                       "'start-cond-inplane' is also provided")
                    a, b = vert.split(",", 1)
                    c0, c1 = float(a), float(b)
            v = _float(t, "start-cond-inplane", manager)
            if v is not None:
                c0 = v
            solver.start_cond = (c0, c1)
        v = _choice(t, "convergence", manager, ("fast", "stable"))
        if v is not None:
            solver.convergence = v
        v = _float(t, "maxerr", manager)
        if v is not None:
            solver.maxerr = v
    elif t.tag == "contacts":
        _check_attrs(t, ("pcond", "ncond"))
        for attr in ("pcond", "ncond"):
            v = _float(t, attr, manager)
            if v is not None:
                setattr(solver, attr, v)
    elif t.tag == "matrix":
        _matrix(t, solver, manager)
    else:
        return False
    return True


def _configure_electrical(solver, el, manager):
    beta = not isinstance(solver, _ActiveCondBase)
    for t in el:
        if _geometry_mesh(t, solver, manager):
            continue
        if beta and t.tag == "junction":
            _junction(t, solver, manager)
        elif not _electrical_tag(t, solver, manager):
            raise XMLError("<%s>: unexpected tag <%s>" % (el.tag, t.tag))


def _configure_meta(solver, el, manager):
    th, el_ = solver.thermal, solver.electrical
    for t in el:
        if t.tag == "geometry":
            for attr, s in (("thermal", th), ("electrical", el_)):
                ref = t.get(attr)
                if ref is not None:
                    if ref not in manager.geo:
                        raise XMLError("<geometry>: no geometry '%s'" % ref)
                    s.geometry = manager.geo[ref]
        elif t.tag == "mesh":
            for attr, s in (("thermal", th), ("electrical", el_)):
                ref = t.get(attr)
                if ref is not None:
                    if ref not in manager.msh:
                        raise XMLError("<mesh>: no mesh '%s'" % ref)
                    s.mesh = manager.msh[ref]
            ee = _choice(t, "empty-elements", manager, ("default", "include", "exclude", "included", "excluded"))
            if ee is not None:
                el_.empty_elements = ee
        elif t.tag == "junction":
            _junction(t, el_, manager)
        elif t.tag == "contacts":
            _electrical_tag(t, el_, manager)
        elif t.tag == "loop":
            _check_attrs(t, ("tfreq", "inittemp", "maxterr", "maxcerr"))
            v = _int(t, "tfreq", manager)
            if v is not None:
                solver.tfreq = v
            for attr, s, pyattr in (("inittemp", th, "inittemp"), ("maxterr", th, "maxerr"), ("maxcerr", el_, "maxerr")):
                v = _float(t, attr, manager)
                if v is not None:
                    setattr(s, pyattr, v)
        elif t.tag in ("tmatrix", "ematrix"):
            _matrix(t, th if t.tag == "tmatrix" else el_, manager)
        elif t.tag in ("temperature", "heatflux", "convection", "radiation"):
            read_conditions(t, getattr(th, t.tag + "_boundary"), t.tag, manager, th.DIM)
        elif t.tag == "voltage":
            read_conditions(t, el_.voltage_boundary, "voltage", manager, el_.DIM)
        else:
            raise XMLError("<%s>: unexpected tag <%s>" % (el.tag, t.tag))


def configure(solver, element, manager=None):
    """Apply one solver element (its children) to solver."""
    manager = manager or Manager()
    el = _element(element)
    if isinstance(solver, _StaticBase):
        _configure_static(solver, el, manager)
    elif isinstance(solver, _DynamicBase):
        _configure_dynamic(solver, el, manager)
    elif isinstance(solver, _ShockleyBase):
        _configure_electrical(solver, el, manager)
    elif isinstance(solver, ThermoElectric):
        _configure_meta(solver, el, manager)
    else:
        raise TypeError("no XPL configuration for %s" % type(solver).__name__)
    return solver


def _element(source):
    if isinstance(source, ET.Element):
        return source
    text = str(source)
    if text.lstrip().startswith("<"):
        return ET.fromstring(text)
    return ET.parse(text).getroot()


def read_defines(el, manager):
    """<defines><define name value/></defines>: each value evaluated in order."""
    for d in el:
        if d.tag != "define":
            raise XMLError("<defines>: unexpected tag <%s>" % d.tag)
        name, value = d.get("name"), d.get("value")
        if name is None or value is None:
            raise XMLError("<define>: 'name' and 'value' are required")
        if name in manager.defines:   # values given by the caller take precedence (command line)
            continue
        ns = {k: getattr(math, k) for k in dir(math) if not k.startswith("_")}
        ns.update(manager.defines)
        try:
            manager.defines[name] = eval(value, {"__builtins__": {}}, ns)  # noqa: S307 -- user's own configuration
        except Exception:
            manager.defines[name] = value


def load_solvers(source, manager=None):
    """{name: solver} for every solver element of a <solvers> section (or a <plask> file)."""
    manager = manager or Manager()
    root = _element(source)
    if root.tag == "plask":
        d = root.find("defines")
        if d is not None:
            read_defines(d, manager)
        sec = root.find("solvers")
        if sec is None:
            return {}
    elif root.tag == "solvers":
        sec = root
    else:
        raise XMLError("expected <plask> or <solvers>, got <%s>" % root.tag)
    out = {}
    for el in sec:
        name, cls_name, lib = el.get("name"), el.get("solver"), el.get("lib")
        if not name or not cls_name:
            raise XMLError("<%s>: attributes 'name' and 'solver' are required" % el.tag)
        table = {}
        for (cat, lb), classes in SOLVERS.items():
            if cat == el.tag and (lib is None or lb == lib):
                table.update(classes)
        if cls_name not in table:
            raise XMLError("<%s>: no solver '%s'%s in this package" % (el.tag, cls_name,
                                                                    " (lib '%s')" % lib if lib else ""))
        solver = table[cls_name](name)
        configure(solver, el, manager)
        out[name] = solver
    return out
