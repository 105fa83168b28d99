"""``thermal.dynamic``: transient FEM thermal solvers Dynamic2D, DynamicCyl, Dynamic3D.

Same classes and attributes as the reference module (solvers/thermal/dynamic/python/
dynamic_python.cpp): compute(time), inHeat, outTemperature, outHeatFlux,
outThermalConductivity, temperature_boundary, inittemp, timestep (ns), methodparam
(0 explicit, 0.5 Crank-Nicolson, 1 implicit), lumping, rebuildfreq, logfreq, time /
elapsed_time, algorithm, iterative, empty_elements.  held_nodes (not in the reference):
'fixed' keeps the temperature-boundary nodes at their values; 'reference' gives them the
reference's value + (B T)_r (femT2d.cpp:243 eliminates the conditions from A and F only,
a term that grows step by step with realistic 2D capacities).  Each time step solves
(theta K + C) T^{n+1} = (C - (1 - theta) K) T^n + F on the GPU (the same matrix-free
operator as the steady solvers with the lumped heat-capacity diagonal added, PCG from
the previous step).
"""
import numpy as np

from .. import mesh as _mesh
from .. import solverbase as _sb
from .._core import DynamicCore
from .static import _StaticBase, _setup_conductivity


class _DynamicBase(_sb.FemSolver):
    _core_class = DynamicCore
    inHeat = _sb.ReceiverAttr(0.0)

    def __init__(self, name=""):
        super().__init__(name)
        # femT3d.cpp:22-38 / femT2d.cpp:22-38
        self.inittemp = 300.0
        self.timestep = 0.1
        self.methodparam = 0.5
        self.lumping = True
        self.rebuildfreq = 0
        self.logfreq = 500
        self.held_nodes = "fixed"
        self.temperature_boundary = _mesh.BoundaryConditions()
        self._T = None

    @property
    def time(self):
        """Time of calculations performed so far since the last invalidation (ns)."""
        return self._core.elapstime if (self._core is not None and self._initialized) else 0.0

    @property
    def elapsed_time(self):
        """Alias for time (obsolete)."""
        return self.time

    def _capacity(self, tavg):
        """Element capacities over the time step in the units of the stiffness, as setMatrix
        forms them: 3D cp rho 0.125e-9 dx dy dz / timestep (femT3d.cpp:230); 2D cp rho
        0.25e-12 w h / timestep / 1e-9 (femT2d.cpp:167), times the mid radius in cylindrical
        geometry (femT2d.cpp:298).  Masked-out elements: none."""
        vol = self._element_volume()
        if self.DIM == 3:
            f = 0.125e-9 * vol / float(self.timestep)
        else:
            f = 0.25e-12 * vol / float(self.timestep) / 1e-9
            if self.CYL:
                f = f * np.tile(0.5 * (self._axes[0][1:] + self._axes[0][:-1]), len(self._axes[1]) - 1)
        c = np.zeros(len(tavg))
        for m, ix in self._groups:
            ix = ix[~self._excl[ix]]
            if not len(ix):
                continue
            if m is None:
                raise ValueError("%s: element without material" % self.id)
            cd = m.heat_capacity_const()
            if cd is not None:
                c[ix] = cd * f[ix]
            else:
                for e in ix:
                    t = float(tavg[e])
                    c[e] = m.cp(t) * m.dens(t) * f[e]
        return c

    def _on_initialize(self):
        if not self.lumping:
            raise NotImplementedError("%s: the consistent (non-lumped) heat-capacity matrix is not implemented on "
                                      "the GPU path; lumping = True (the reference's default) is" % self.id)
        c = self._core
        c.inittemp = float(self.inittemp)
        self._h = self._thickness()
        _setup_conductivity(self)
        groups = {}
        for e, m in enumerate(self._mat):
            groups.setdefault(id(m), (m, []))[1].append(e)
        self._groups = [(m, np.array(ix)) for m, ix in groups.values()]
        c.set_capacity_callback(self._capacity)
        c.initialize()
        self._T = None

    def compute(self, time):
        """Run thermal calculations for `time` ns (femT3d.cpp:282-331)."""
        self.initialize()
        c = self._core
        c.inittemp = float(self.inittemp)
        c.timestep = float(self.timestep)
        c.methodparam = float(self.methodparam)
        c.rebuildfreq = int(self.rebuildfreq)
        c.logfreq = int(self.logfreq)
        if self.held_nodes not in ("fixed", "reference"):
            raise ValueError("%s: held_nodes must be 'fixed' or 'reference'" % self.id)
        c.held = 0 if self.held_nodes == "fixed" else 1
        q = np.array(self.inHeat(self._element_mesh()), dtype=np.float64)
        q[self._excl] = 0.0
        c.set_heat(q)
        nodes, vals = self.temperature_boundary.resolve(self.mesh, self.geometry)
        if len(self._isolated):  # nodes of the masked-out region hold the initial temperature
            nodes = np.concatenate([nodes, self._isolated])
            vals = np.concatenate([vals, np.full(len(self._isolated), float(self.inittemp))])
        c.set_dirichlet(nodes.astype(np.int64), vals)
        r = self._run_core(c.compute, float(time))
        self._T = None
        return r

    def _temps(self):
        if self._T is None:
            self._T = self._core.solution()
        return self._T

    # the providers read the same node temperatures as the steady solvers'
    outTemperature = _StaticBase.outTemperature
    _element_T = _StaticBase._element_T
    _element_k = _StaticBase._element_k
    outThermalConductivity = _StaticBase.outThermalConductivity
    outHeatFlux = _StaticBase.outHeatFlux


class Dynamic2D(_DynamicBase):
    """Finite element transient thermal solver for 2D Cartesian geometry."""
    DIM, CYL = 2, False


class DynamicCyl(_DynamicBase):
    """Finite element transient thermal solver for 2D cylindrical geometry."""
    DIM, CYL = 2, True


class Dynamic3D(_DynamicBase):
    """Finite element transient thermal solver for 3D geometry."""
    DIM, CYL = 3, False
