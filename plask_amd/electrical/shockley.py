"""``electrical.shockley``: Shockley-junction FEM solvers Shockley2D, ShockleyCyl, Shockley3D.

Same classes and attributes as the reference module (solvers/electrical/shockley/
python/electr_python.cpp:195-297): compute(loops=0), get_total_current(nact=0), err,
inTemperature, outVoltage, outCurrentDensity, outHeat, outConductivity,
voltage_boundary, maxerr, convergence, pcond, ncond, start_cond (pnjcond), beta/beta#,
js/js# (float or callable of T), get_electrostatic_energy, get_capacitance,
get_total_heat.  outPotential raises TypeError as in the reference.

ActiveCond2D, ActiveCondCyl, ActiveCond3D (PythonCondSolver, electr_python.cpp:113-190)
take the junction conductivity from a Python callable cond/cond# (U in V, j in kA/cm^2,
T in K -> S/m, or an (in-plane, vertical) pair) instead of the Shockley law.
"""
import math
import re

import numpy as np

from .. import mesh as _mesh
from .. import solverbase as _sb
from .._core import ElectricalCore

EPSILON0 = 8.854187817e-12  # F/m (plask/phys/constants.hpp)
_POST_CURRENT, _POST_JOULE = 1, 2   # PF_POST_CURRENT, PF_POST_JOULE


def _active_number(roles):
    """Junction number + 1 from 'active'/'active#'/'junction'/'junction#' roles, else 0
    (ElectricalFem2DSolver::isActive, electr2d.hpp:132-155)."""
    no = 0
    for r in roles:
        m = re.fullmatch(r"(active|junction)(\d*)", r)
        if not m:
            continue
        if no:
            raise ValueError("multiple 'active'/'junction' roles specified")
        no = int(m.group(2)) + 1 if m.group(2) else 1
    return no


class _ShockleyBase(_sb.FemSolver):
    _core_class = ElectricalCore
    inTemperature = _sb.ReceiverAttr(300.0)
    _PARAMS = re.compile(r"(beta|js)(\d*)")

    def __init__(self, name=""):
        object.__setattr__(self, "_beta", {})
        object.__setattr__(self, "_js", {})
        object.__setattr__(self, "_cond", {})
        super().__init__(name)
        self.voltage_boundary = _mesh.BoundaryConditions()
        self.maxerr = 0.05
        self.convergence = "fast"
        self.pcond = 5.0
        self.ncond = 50.0
        self._start_cond = (0.0, 5.0)
        # contact rows (sigma up to 1e9 S/m) dominate the residual norm; the junction rows
        # (sigma ~ 1e-4 S/m) need a much smaller relative residual to be resolved
        self._iter.maxerr = 1e-12

    # beta#, js# (BetaSolver, beta.hpp:49-77; Python callables electr_python.cpp:96-111)
    def __getattr__(self, attr):
        m = type(self)._PARAMS.fullmatch(attr)
        if m and m.group(1) == "cond":
            n = int(m.group(2)) if m.group(2) else 0
            return object.__getattribute__(self, "_cond").get(n, None)
        if m:
            d = object.__getattribute__(self, "_" + m.group(1))
            n = int(m.group(2)) if m.group(2) else 0
            if n not in d:
                raise AttributeError("%s: no %s given for junction %d" % (self.id, m.group(1), n))
            return d[n]
        raise AttributeError("'%s' object has no attribute '%s'" % (type(self).__name__, attr))

    def __setattr__(self, attr, value):
        m = type(self)._PARAMS.fullmatch(attr)
        if m and m.group(1) == "cond":
            n = int(m.group(2)) if m.group(2) else 0
            if not callable(value):
                raise TypeError("%s: cond%d must be a a callable" % (self.id, n))
            self._cond[n] = value
            self.invalidate()
            return
        if m:
            if not (callable(value) or isinstance(value, (int, float))):
                raise TypeError("%s: %s must be a float or a callable" % (self.id, attr))
            n = int(m.group(2)) if m.group(2) else 0
            getattr(self, "_" + m.group(1))[n] = value
            self.invalidate()
            return
        object.__setattr__(self, attr, value)

    @property
    def start_cond(self):
        return self._start_cond

    @start_cond.setter
    def start_cond(self, v):
        self._start_cond = (0.0, float(v)) if np.isscalar(v) else (float(v[0]), float(v[1]))
        self.invalidate()

    pnjcond = start_cond

    @property
    def outPotential(self):
        raise TypeError("%s: 'outPotential' is reserved for drift-diffusion model; use 'outVoltage' instead" % self.id)

    # -- initialization -----------------------------------------------------------
    def _on_initialize(self):
        c = self._core
        self._cond_groups = None
        m = [len(a) - 1 for a in self._axes]
        E = int(np.prod(m))
        leaf_act = np.array([_active_number(r) for r in self._leaf_roles], dtype=np.int64)
        act = leaf_act[self._leaf]                       # index -1: background (no junction)
        self._act = act
        shape = m[::-1]  # [k][j][i] or [j][i]
        actg = act.reshape(shape)
        nj = int(act.max()) if E else 0
        z = self._axes[-1]
        n0 = len(self._axes[0])
        n1 = len(self._axes[1]) if self.DIM == 3 else 1
        plane = n0 * n1
        elems, bot, top, jid, heights, self._junctions = [], [], [], [], [], []
        for n in range(1, nj + 1):
            where = np.argwhere(actg == n)
            if len(where) == 0:
                heights.append(1.0)
                self._junctions.append(None)
                continue
            kb, kt = int(where[:, 0].min()), int(where[:, 0].max()) + 1
            h = float(z[kt] - z[kb])
            heights.append(h)
            self._junctions.append((kb, kt, h))
            for w in where:
                if self.DIM == 2:
                    k, i = int(w[0]), int(w[1])
                    e = i + m[0] * k
                    cols = [i, i + 1]
                else:
                    k, j, i = int(w[0]), int(w[1]), int(w[2])
                    e = i + m[0] * (j + m[1] * k)
                    cols = [i + n0 * j, i + 1 + n0 * j, i + n0 * (j + 1), i + 1 + n0 * (j + 1)]
                elems.append(e)
                bot.extend([cc + kb * plane for cc in cols])
                top.extend([cc + kt * plane for cc in cols])
                jid.append(n - 1)
        self._jelems = np.array(elems, dtype=np.int64)
        self._jid = np.array(jid, dtype=np.int32)
        c.set_start_cond(float(self._start_cond[1]), float(self._start_cond[0]))
        c.set_junctions(self._jelems, 2 if self.DIM == 2 else 4, np.array(bot, dtype=np.int64),
                        np.array(top, dtype=np.int64), list(self._jid), heights)
        leaves = self._geometry._leaves
        air_noheat = True
        leaf_noheat = np.array([(lf[0].material is None or getattr(lf[0].material, "kind", "") == "empty" or
                                 "noheat" in r) for lf, r in zip(leaves, self._leaf_roles)] + [air_noheat],
                               dtype=np.uint8)
        noheat = leaf_noheat[self._leaf]
        sel = (act > 0).astype(np.uint8) if len(elems) else np.zeros(0, np.uint8)
        c.set_flags(sel, noheat)
        c.initialize()

    @staticmethod
    def _mat_air():
        from .. import material
        return material.Air()

    def _load_conductivities(self, T):
        """Element conductivities (loadConductivities, electr2d.cpp:331-353).  Elements
        are grouped by (material, role class); materials with a closed-form law
        (cond_law: s0 (300/T)^alpha) are evaluated for a whole group at once."""
        E = len(self._mat)
        self._build_cond_groups()
        sh = np.empty(E)
        sv = np.empty(E)
        T = np.asarray(T, dtype=np.float64)
        for kind, mm, ix in self._cond_groups:
            if kind in ("act", "none"):
                sh[ix] = sv[ix] = 0.0
            elif kind == "p":
                sh[ix] = sv[ix] = self.pcond
            elif kind == "n":
                sh[ix] = sv[ix] = self.ncond
            else:
                law = mm.cond_law()
                if law is not None:
                    t = 300.0 / T[ix]
                    sh[ix] = law[0] * t ** law[2]
                    sv[ix] = law[1] * t ** law[3]
                else:
                    for e in ix:
                        v = mm.cond(float(T[e]))
                        sh[e], sv[e] = (v, v) if np.isscalar(v) else (float(v[0]), float(v[1]))
        return sh, sv

    def _cond_law_arrays(self):
        """The element conductivities as the device law s0 (300 / T)^alpha: (s0_h, s0_v,
        alpha_h, alpha_v) per element, NaN s0_h at the junction elements (their conductivity
        is the junction update's), when every material conductivity is such a law or
        constant and beta and js do not depend on temperature; else None."""
        if isinstance(self, _ActiveCondBase):   # the junction law takes the temperature
            return None
        if any(callable(v) for v in list(self._beta.values()) + list(self._js.values())):
            return None
        E = len(self._mat)
        self._build_cond_groups()
        s0h, s0v, ah, av = np.zeros(E), np.zeros(E), np.zeros(E), np.zeros(E)
        for kind, mm, ix in self._cond_groups:
            if kind == "p":
                s0h[ix] = s0v[ix] = self.pcond
            elif kind == "n":
                s0h[ix] = s0v[ix] = self.ncond
            elif kind == "mat":
                law = mm.cond_law()
                if law is None:
                    return None
                s0h[ix], s0v[ix], ah[ix], av[ix] = law
        s0h[self._excl], s0v[self._excl] = _sb.MASKED_K
        ah[self._excl] = av[self._excl] = 0.0
        s0h[np.asarray(self._jelems, dtype=np.int64)] = np.nan
        return s0h, s0v, ah, av

    def _build_cond_groups(self):
        """Elements grouped by (material, role class), once per initialization."""
        if getattr(self, "_cond_groups", None) is None:
            leaves = self._geometry._leaves
            nl = len(leaves)
            leaf = np.where(self._leaf < 0, nl, self._leaf)   # background air -> entry nl
            groups = {}
            for k in range(nl + 1):
                mm = leaves[k][0].material if k < nl else self._mat_air()
                r = self._leaf_roles[k] if k < nl else frozenset()
                if _active_number(r):
                    key = ("act", None)
                elif "p-contact" in r:
                    key = ("p", None)
                elif "n-contact" in r:
                    key = ("n", None)
                elif mm is None:
                    key = ("none", None)
                else:
                    key = ("mat", id(mm))
                groups.setdefault(key, (mm, []))[1].append(k)
            order = np.argsort(leaf, kind="stable")
            bounds = np.searchsorted(leaf[order], np.arange(nl + 2))
            self._cond_groups = []
            for key, (mm, ks) in groups.items():
                ix = np.concatenate([order[bounds[k]:bounds[k + 1]] for k in ks])
                if len(ix):
                    self._cond_groups.append((key[0], mm, np.sort(ix)))

    def compute(self, loops=0):
        """Run electrical calculations; returns the maximum current-density change (%)."""
        self.initialize()
        c = self._core
        c.maxerr = float(self.maxerr)
        c.stable = (str(self.convergence).lower() == "stable")
        if not (getattr(self, "_temp_linked", False) and c.device_coeffs):
            T = self.inTemperature(self._element_mesh())
            sh, sv = self._load_conductivities(T)
            sh[self._excl], sv[self._excl] = _sb.MASKED_K   # masked-out (empty) elements
            c.set_conductivity(sh, sv)
            self._junction_setup(np.asarray(T, dtype=np.float64))
        # else: linked to a thermal solver on the same mesh -- the core evaluates the
        # conductivity law at its temperatures on the device (junction parameters constant)
        nodes, vals = self.voltage_boundary.resolve(self.mesh, self.geometry)
        if len(self._isolated):  # nodes of the masked-out region: held at 0 V, coupled to nothing
            nodes = np.concatenate([nodes, self._isolated])
            vals = np.concatenate([vals, np.zeros(len(self._isolated))])
        c.set_dirichlet(nodes.astype(np.int64), vals)
        err = self._run_core(c.compute, int(loops))
        self._V = None
        return err

    def _junction_temps(self, T):
        """Per junction element the temperature of its column's element in the junction's
        middle row (electr2d.cpp:255)."""
        m = [len(a) - 1 for a in self._axes]
        plane = m[0] if self.DIM == 2 else m[0] * m[1]
        e = np.asarray(self._jelems, dtype=np.int64)
        kmid = np.array([0 if j is None else (j[0] + j[1]) // 2 for j in self._junctions], dtype=np.int64)
        return np.asarray(T)[e % plane + plane * kmid[np.asarray(self._jid, dtype=np.int64)]]

    def _junction_setup(self, T):
        """beta, js per junction element at its temperature (BetaSolver getBeta/getJs and the
        Python callables of electr_python.cpp:102-108)."""
        Tj = self._junction_temps(T)
        beta = np.empty(len(self._jelems))
        js = np.empty(len(self._jelems))
        for q, n in enumerate(self._jid):
            b = self._beta.get(int(n), None)
            j = self._js.get(int(n), 1.0)
            if b is None:
                raise ValueError("%s: no beta given for junction %d" % (self.id, n))
            beta[q] = b(float(Tj[q])) if callable(b) else float(b)
            js[q] = j(float(Tj[q])) if callable(j) else float(j)
        self._core.set_junction_params(beta, js)

    @property
    def err(self):
        return self._core.toterr if self._core is not None else 0.0

    def _volt(self):
        if getattr(self, "_V", None) is None:
            self._V = self._core.solution()
        return self._V

    # -- providers -----------------------------------------------------------------
    def outVoltage(self, mesh, interpolation="linear"):
        if self._core is None or self._core.loopno == 0:
            raise ValueError("%s: voltage not computed" % self.id)
        return _sb.interp_nodes(self._axes, self._volt, mesh, fill=np.nan, device=self._core)

    def _element_field(self, what, mesh):
        """A device element field at the points of mesh: the values themselves on the
        element mesh, else linear interpolation on the element mesh on the device
        (getCurrentDensities / getHeatDensities, electr2d.cpp:518-562)."""
        c = self._core
        mids = [0.5 * (a[1:] + a[:-1]) for a in self._axes]
        if isinstance(mesh, _mesh.Rectangular) and mesh.dim == self.DIM and _sb._same_axes(mesh.axes, mids):
            return c.currents().reshape(-1, self.DIM) if what == _POST_CURRENT else c.heats()
        pts = np.ascontiguousarray(_sb.points_of(mesh), dtype=np.float64).reshape(-1, self.DIM)
        out = c.interp_elements(what, pts, 0.0)
        return out.reshape(-1, self.DIM) if what == _POST_CURRENT else out

    def outCurrentDensity(self, mesh, interpolation="linear"):
        """Current density (kA/cm^2)."""
        if self._core is None or self._core.loopno == 0:
            raise ValueError("%s: current density not computed" % self.id)
        return self._element_field(_POST_CURRENT, mesh)

    def outHeat(self, mesh, interpolation="linear"):
        """Joule heat density (W/m^3) (getHeatDensities, electr2d.cpp:541-562)."""
        if self._core is None or self._core.loopno == 0:
            raise ValueError("%s: heat density not computed" % self.id)
        return self._element_field(_POST_JOULE, mesh)

    def outConductivity(self, mesh, interpolation="nearest"):
        self.initialize()
        if self._core.loopno == 0:
            T = self.inTemperature(self._element_mesh())
            sh, sv = self._load_conductivities(T)
            sv = sv.copy()
            sv[self._act > 0] = self._start_cond[1]
            sh = sh.copy()
            sh[self._act > 0] = self._start_cond[0]
        else:
            sh, sv = self._core.conductivity_h(), self._core.conductivity_v()
        k = np.column_stack([sh, sv])
        k[self._excl] = np.nan   # not in the masked mesh
        return _sb.nearest_elements(self._axes, k, mesh)

    # -- integrals -------------------------------------------------------------------
    def get_total_current(self, nact=0):
        """Current through the middle row of junction nact (mA) (getTotalCurrent,
        electr2d.cpp:466-505, electr3d.cpp:480-510)."""
        if nact >= len(self._junctions) or self._junctions[nact] is None:
            raise ValueError("%s: wrong active region number" % self.id)
        kb, kt, _ = self._junctions[nact]
        level = (kb + kt) // 2
        j = self._core.currents().reshape(-1, self.DIM)
        m = [len(a) - 1 for a in self._axes]
        s = self._esize
        if self.DIM == 2:
            row = slice(level * m[0], (level + 1) * m[0])
            jr = j[row, 1]
            act = self._act[row] > 0
            if self.CYL:
                r = self._axes[0]
                return float(np.sum(np.where(act, jr * (r[1:] ** 2 - r[:-1] ** 2), 0.0)) * math.pi * 0.01)
            return float(np.sum(np.where(act, jr * s[0], 0.0)) * self.geometry.length * 0.01)
        n = m[0] * m[1]
        row = slice(level * n, (level + 1) * n)
        area = (s[0][None, :] * s[1][:, None]).ravel()
        act = self._act[row] > 0
        return float(np.sum(np.where(act, j[row, 2] * area, 0.0)) * 0.01)

    def get_total_heat(self):
        """Total Joule heat (mW) (getTotalHeat, electr2d.cpp:627-649)."""
        h = self._core.heats()
        vol = self._element_volume()
        if self.DIM == 2 and self.CYL:
            r = 0.5 * (self._axes[0][1:] + self._axes[0][:-1])
            rr = np.tile(r, len(self._axes[1]) - 1)
            return float(2e-15 * math.pi * np.sum(vol * rr * h))
        if self.DIM == 2:
            return float(self.geometry.length * 1e-15 * np.sum(vol * h))
        return float(1e-15 * np.sum(vol * h))

    def _grad2(self):
        """|grad V|^2 per element in (V/m)^2 with the reference's element gradients."""
        V = self._volt().reshape([len(a) for a in self._axes][::-1])
        s = self._esize
        if self.DIM == 2:
            gx = 0.5e6 * (-V[:-1, :-1] + V[:-1, 1:] - V[1:, :-1] + V[1:, 1:]) / s[0][None, :]
            gy = 0.5e6 * (-V[:-1, :-1] - V[:-1, 1:] + V[1:, :-1] + V[1:, 1:]) / s[1][:, None]
            return (gx ** 2 + gy ** 2).ravel()
        tot = []
        for ax in range(3):
            acc = 0
            for c in range(8):
                sl = tuple(slice(1, None) if (c >> (2 - a)) & 1 else slice(None, -1) for a in range(3))
                acc = acc + (1.0 if (c >> (2 - ax)) & 1 else -1.0) * V[sl]
            tot.append(acc)
        gz = 0.25e6 * tot[0] / s[2][:, None, None]
        gy = 0.25e6 * tot[1] / s[1][None, :, None]
        gx = 0.25e6 * tot[2] / s[0][None, None, :]
        return (gx ** 2 + gy ** 2 + gz ** 2).ravel()

    def get_electrostatic_energy(self):
        """Energy of the electrostatic field (J) (getTotalEnergy, electr2d.cpp:574-615)."""
        T = np.asarray(self.inTemperature(self._element_mesh()), dtype=np.float64)
        # permittivity per leaf (last entry: the background), at each distinct temperature
        mats = [lf[0].material for lf in self._geometry._leaves] + [self._mat_bg]
        li = np.where(self._leaf < 0, len(mats) - 1, self._leaf)
        eps = np.empty(len(T))
        for g, mm in enumerate(mats):
            ix = np.nonzero(li == g)[0]
            if not len(ix):
                continue
            if mm is None:
                eps[ix] = 1.0
                continue
            tu, inv = np.unique(T[ix], return_inverse=True)
            eps[ix] = np.array([float(mm.eps(float(t))) for t in tu])[inv]
        w = np.where(self._excl, 0.0, eps * self._grad2() * self._element_volume())
        if self.DIM == 2 and self.CYL:
            r = np.tile(0.5 * (self._axes[0][1:] + self._axes[0][:-1]), len(self._axes[1]) - 1)
            return float(2.0 * math.pi * 0.5e-18 * EPSILON0 * np.sum(w * r))
        if self.DIM == 2:
            return float(self.geometry.length * 0.5e-18 * EPSILON0 * np.sum(w))
        return float(0.5e-18 * EPSILON0 * np.sum(w))

    def get_capacitance(self):
        """Capacitance (pF) = 2 W / U^2 with exactly two voltage conditions (electr2d.cpp:617-625)."""
        if len(self.voltage_boundary) != 2:
            raise ValueError("%s: cannot estimate applied voltage (exactly 2 voltage boundary conditions required)"
                             % self.id)
        U = self.voltage_boundary[0].value - self.voltage_boundary[1].value
        return 2e12 * self.get_electrostatic_energy() / (U * U)


def _cond_values(f, U, j, T):
    """cond(U, j, T) of a junction for arrays of its elements: one vectorised call when the
    callable takes arrays, else one call per element as the reference makes them.
    Returns (in-plane, vertical) arrays; a plain number is (0, value) (electr_python.cpp:174-177)."""
    n = len(U)

    def split(r):
        if isinstance(r, (tuple, list)) and len(r) == 2:
            return np.broadcast_to(np.asarray(r[0], float), (n,)), np.broadcast_to(np.asarray(r[1], float), (n,))
        a = np.asarray(r, dtype=float)
        if a.ndim == 2 and a.shape == (n, 2):
            return a[:, 0], a[:, 1]
        return np.zeros(n), np.broadcast_to(a, (n,))
    try:
        with np.errstate(all="ignore"):
            r = f(U, j, T)
        h, v = split(r)
        return np.array(h, dtype=float), np.array(v, dtype=float)
    except (TypeError, ValueError):
        pass
    h, v = np.zeros(n), np.zeros(n)
    for q in range(n):
        r = f(float(U[q]), float(j[q]), float(T[q]))
        if isinstance(r, (tuple, list)) or np.ndim(r) == 1:
            h[q], v[q] = float(r[0]), float(r[1])
        else:
            v[q] = float(r)
    return h, v


class _ActiveCondBase(_ShockleyBase):
    """Junction conductivity from cond# callables (PythonCondSolver::activeCond,
    electr_python.cpp:171-178), evaluated in each loop's junction update (electr2d.cpp:247-256)."""
    _PARAMS = re.compile(r"(cond)(\d*)")

    def _junction_setup(self, T):
        Tj = self._junction_temps(T)
        jid = np.asarray(self._jid)
        groups = [(n, np.nonzero(jid == n)[0]) for n in np.unique(jid)]
        for n, _ in groups:
            if self._cond.get(int(n)) is None:
                raise IndexError("no conductivity for active region %d" % n)

        def law(U, jy):
            ch, cv = np.zeros(len(U)), np.zeros(len(U))
            for n, ix in groups:
                ch[ix], cv[ix] = _cond_values(self._cond[int(n)], U[ix], jy[ix], Tj[ix])
            return ch, cv
        self._core.set_junction_law(law)


class Shockley2D(_ShockleyBase):
    """Finite element electrical solver for 2D Cartesian geometry."""
    DIM, CYL = 2, False


class ShockleyCyl(_ShockleyBase):
    """Finite element electrical solver for 2D cylindrical geometry."""
    DIM, CYL = 2, True


class Shockley3D(_ShockleyBase):
    """Finite element electrical solver for 3D Cartesian geometry."""
    DIM, CYL = 3, False


class ActiveCond2D(_ActiveCondBase):
    """Finite element electrical solver for 2D Cartesian geometry with a junction
    conductivity function."""
    DIM, CYL = 2, False


class ActiveCondCyl(_ActiveCondBase):
    """Finite element electrical solver for 2D cylindrical geometry with a junction
    conductivity function."""
    DIM, CYL = 2, True


class ActiveCond3D(_ActiveCondBase):
    """Finite element electrical solver for 3D Cartesian geometry with a junction
    conductivity function."""
    DIM, CYL = 3, False
