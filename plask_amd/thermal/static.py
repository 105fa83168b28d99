"""``thermal.static``: steady-state FEM thermal solvers Static2D, StaticCyl, Static3D.

Same classes and attributes as the reference module (solvers/thermal/static/python/
therm_python.cpp:97-152): compute(loops=0) -> err, inHeat, outTemperature,
outHeatFlux, outThermalConductivity, temperature_boundary, heatflux_boundary,
convection_boundary, radiation_boundary, inittemp, maxerr, algorithm, iterative,
empty_elements.  The linear solve runs on the GPU (matrix-free operator, PCG).
"""
import numpy as np

from .. import mesh as _mesh
from .. import solverbase as _sb
from .._core import ThermalCore

_POST_FLUX = 0   # PF_POST_FLUX


class _BcValue:
    """Value of a convection / radiation boundary condition (therm_python.cpp:25-86):
    ``Cls(first, ambient)``; also built from a dict {first: ..., 'ambient': ...}
    (ambient defaults to 300 K) or a (first, ambient) pair; item access by key."""
    FIRST = ""

    def __init__(self, first, ambient=300.0):
        setattr(self, self.FIRST, float(first))
        self.ambient = float(ambient)

    @classmethod
    def of(cls, v):
        if isinstance(v, cls):
            return v
        if isinstance(v, dict):
            return cls(v[cls.FIRST], v.get("ambient", 300.0))
        if isinstance(v, (tuple, list)) and len(v) == 2:
            return cls(*v)
        raise TypeError("%s boundary value must be %s(%s, ambient), a dict or a pair" %
                        (cls.__name__, cls.__name__, cls.FIRST))

    def __getitem__(self, key):
        if key in (self.FIRST, "ambient"):
            return getattr(self, key)
        raise KeyError(key)

    def __setitem__(self, key, value):
        if key not in (self.FIRST, "ambient"):
            raise KeyError(key)
        setattr(self, key, float(value))

    def __repr__(self):
        return "{'%s': %r, 'ambient': %r}" % (self.FIRST, getattr(self, self.FIRST), self.ambient)

    def __str__(self):
        return "%s (%sK)" % (getattr(self, self.FIRST), self.ambient)


class Convection(_BcValue):
    """Convective boundary condition value: coeff (W/m^2/K), ambient (K)."""
    FIRST = "coeff"


class Radiation(_BcValue):
    """Radiative boundary condition value: emissivity (-), ambient (K)."""
    FIRST = "emissivity"


def _bc_lists(bcs, mesh, geometry, convert):
    """Nodes and converted values of a condition list, in list order (the first place
    holding a node gives its value, BoundaryConditions::getValue)."""
    nodes, vals = [], []
    for place, value in bcs:
        nd = np.asarray(place.nodes(mesh, geometry), dtype=np.int64)
        nodes.append(nd)
        vals.extend([convert(value)] * len(nd))
    return (np.concatenate(nodes) if nodes else np.zeros(0, np.int64)), vals


def _setup_conductivity(solver):
    """Element conductivities of a thermal solver's core: one device law k0 (300/T)^alpha
    per leaf when every material has one, else a host callback evaluating thermk(T, h)
    per element at the element temperatures (setMatrix, therm3d.cpp:209-216)."""
    c = solver._core
    leaf_mats = [lf[0].material for lf in solver._geometry._leaves] + [solver._mat_bg]
    laws = [m.thermk_law() if m is not None else None for m in leaf_mats]
    solver._laws_on_device = all(l is not None for l in laws)
    if solver._laws_on_device:
        L = np.array(laws)[solver._leaf]
        L[solver._excl, 0], L[solver._excl, 1] = _sb.MASKED_K   # masked-out (empty) elements
        c.set_conductivity(L[:, 0].copy(), L[:, 1].copy(), L[:, 2].copy(), L[:, 3].copy())
        return
    mats = solver._mat
    h = solver._h
    groups = {}
    for e, m in enumerate(mats):
        groups.setdefault(id(m), (m, []))[1].append(e)
    groups = [(m, np.array(ix)) for m, ix in groups.values()]
    excl = solver._excl

    def kfun(tavg):
        kh = np.full(len(tavg), _sb.MASKED_K[0])
        kv = np.full(len(tavg), _sb.MASKED_K[1])
        for m, ix in groups:
            ix = ix[~excl[ix]]
            if len(ix) and m is None:
                raise ValueError("%s: element without material" % solver.id)
            for e in ix:
                k = m.thermk(float(tavg[e]), float(h[e]))
                k = (k, k) if np.isscalar(k) else k
                kh[e], kv[e] = float(k[0]), float(k[1])
        return kh, kv

    c.set_conductivity_callback(kfun)
    # a placeholder law so that the device holds valid coefficients before the first loop
    E = len(mats)
    c.set_conductivity(np.ones(E), np.ones(E), np.zeros(E), np.zeros(E))
    c.set_conductivity_callback(kfun)


class _StaticBase(_sb.FemSolver):
    _core_class = ThermalCore
    inHeat = _sb.ReceiverAttr(0.0)

    def __init__(self, name=""):
        super().__init__(name)
        self.inittemp = 300.0
        self.maxerr = 0.05
        self.temperature_boundary = _mesh.BoundaryConditions()
        self.heatflux_boundary = _mesh.BoundaryConditions()
        self.convection_boundary = _mesh.BoundaryConditions()
        self.radiation_boundary = _mesh.BoundaryConditions()
        self._T = None

    @property
    def inHeatDensity(self):
        return self.inHeat

    @inHeatDensity.setter
    def inHeatDensity(self, v):
        self.inHeat = v

    def _on_initialize(self):
        c = self._core
        c.inittemp = float(self.inittemp)
        self._h = self._thickness()
        _setup_conductivity(self)
        c.initialize()
        self._T = None

    def compute(self, loops=0):
        """Run thermal calculations; returns the maximum temperature update (K)."""
        self.initialize()
        c = self._core
        c.inittemp = float(self.inittemp)
        c.maxerr = float(self.maxerr)
        if not getattr(self, "_heat_linked", False):
            q = np.array(self.inHeat(self._element_mesh()), dtype=np.float64)
            q[self._excl] = 0.0
            c.set_heat(q)
        # else: linked to an electrical solver on the same mesh -- the core copies its Joule
        # heat into the source on the device
        nodes, vals = self.temperature_boundary.resolve(self.mesh, self.geometry)
        if len(self._isolated):  # nodes of the masked-out region hold the initial temperature
            nodes = np.concatenate([nodes, self._isolated])
            vals = np.concatenate([vals, np.full(len(self._isolated), float(self.inittemp))])
        c.set_dirichlet(nodes.astype(np.int64), vals)
        fn, fv = _bc_lists(self.heatflux_boundary, self.mesh, self.geometry, float)
        rn, rv = _bc_lists(self.radiation_boundary, self.mesh, self.geometry, Radiation.of)
        c.set_boundary_loads(fn, np.asarray(fv, dtype=np.float64), rn,
                             np.array([v.emissivity for v in rv], dtype=np.float64),
                             np.array([v.ambient for v in rv], dtype=np.float64))
        cn, cv = _bc_lists(self.convection_boundary, self.mesh, self.geometry, Convection.of)
        c.set_convection(cn, np.array([v.coeff for v in cv], dtype=np.float64),
                         np.array([v.ambient for v in cv], dtype=np.float64))
        err = self._run_core(c.compute, int(loops))
        self._T = None
        return err

    @property
    def err(self):
        return self._core.toterr if self._core is not None else 0.0

    def _temps(self):
        if self._T is None:
            self._T = self._core.solution()
        return self._T

    # -- providers ---------------------------------------------------------------
    def outTemperature(self, mesh, interpolation="linear"):
        """Temperatures (K); inittemp everywhere before the first compute
        (ThermalFem2DSolver::getTemperatures, therm2d.cpp:543-552)."""
        if not self._initialized or self._core is None or self._core.loopno == 0:
            return np.full(len(_sb.points_of(mesh)), float(self.inittemp))
        return _sb.interp_nodes(self._axes, self._temps, mesh, fill=300.0, device=self._core)

    def _element_T(self):
        T = self._temps().reshape([len(a) for a in self._axes][::-1])
        if self.DIM == 2:
            return 0.25 * (T[:-1, :-1] + T[:-1, 1:] + T[1:, :-1] + T[1:, 1:]).ravel()
        return 0.125 * (T[:-1, :-1, :-1] + T[:-1, :-1, 1:] + T[:-1, 1:, :-1] + T[:-1, 1:, 1:] +
                        T[1:, :-1, :-1] + T[1:, :-1, 1:] + T[1:, 1:, :-1] + T[1:, 1:, 1:]).ravel()

    def _element_k(self, leaf_height=False):
        tav = self._element_T() if (self._core is not None and self._core.loopno > 0) else \
            np.full(len(self._mat), float(self.inittemp))
        h = self._h
        leaves = self._geometry._leaves
        if leaf_height:
            h = np.array([lf[2][-1] for lf in leaves] + [np.inf])[self._leaf]
        # per leaf (the last entry: the background): a closed-form law k0 (300/T)^alpha
        # for the whole group at once, else the material's thermk per element
        mats = [lf[0].material for lf in leaves] + [self._mat_bg]
        li = np.where(self._leaf < 0, len(mats) - 1, self._leaf)
        order = np.argsort(li, kind="stable")
        bounds = np.searchsorted(li[order], np.arange(len(mats) + 1))
        k = np.empty((len(self._mat), 2))
        for g, m in enumerate(mats):
            ix = order[bounds[g]:bounds[g + 1]]
            if not len(ix):
                continue
            law = m.thermk_law() if m is not None else None
            if m is None:
                k[ix] = np.nan
            elif law is not None:
                t = 300.0 / tav[ix]
                k[ix, 0] = law[0] * t ** law[2]
                k[ix, 1] = law[1] * t ** law[3]
            else:
                for e in ix:
                    v = m.thermk(float(tav[e]), float(h[e]))
                    k[e] = (v, v) if np.isscalar(v) else v
        return k

    def outThermalConductivity(self, mesh, interpolation="nearest"):
        """Element conductivities (in-plane, vertical) at the element temperatures."""
        self.initialize()
        k = self._element_k()
        k[self._excl] = np.nan   # not in the masked mesh
        return _sb.nearest_elements(self._axes, k, mesh)

    def outHeatFlux(self, mesh, interpolation="linear"):
        """Heat flux -k grad T (W/m^2) on the device (saveHeatFluxes, therm2d.cpp:504-540,
        therm3d.cpp:342-384).  k is thermk at the final element temperatures with the
        leaf's height (therm2d.cpp:525-531): the device law when every material has one
        (no height dependence), else evaluated here and handed over.  Interpolated on the
        element mesh on the device (getHeatFluxes, therm2d.cpp:556-569)."""
        if self._core is None or self._core.loopno == 0:
            return np.zeros((len(_sb.points_of(mesh)), self.DIM))
        c = self._core
        if getattr(self, "_laws_on_device", False):
            fetch_k = (None, None)
        else:
            k = self._element_k(leaf_height=True)
            fetch_k = (k[:, 0].copy(), k[:, 1].copy())
        mids = [0.5 * (a[1:] + a[:-1]) for a in self._axes]
        if isinstance(mesh, _mesh.Rectangular) and mesh.dim == self.DIM and _sb._same_axes(mesh.axes, mids):
            return c.postprocess(_POST_FLUX, fetch_k[0], fetch_k[1], True).reshape(-1, self.DIM)
        c.postprocess(_POST_FLUX, fetch_k[0], fetch_k[1], False)
        pts = np.ascontiguousarray(_sb.points_of(mesh), dtype=np.float64).reshape(-1, self.DIM)
        return c.interp_elements(_POST_FLUX, pts, 0.0).reshape(-1, self.DIM)


class Static2D(_StaticBase):
    """Finite element thermal solver for 2D Cartesian geometry."""
    DIM, CYL = 2, False


class StaticCyl(_StaticBase):
    """Finite element thermal solver for 2D cylindrical geometry."""
    DIM, CYL = 2, True


class Static3D(_StaticBase):
    """Finite element thermal solver for 3D geometry."""
    DIM, CYL = 3, False
