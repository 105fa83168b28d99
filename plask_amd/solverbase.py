"""Common parts of the FEM solver classes: geometry/mesh handling, element tables,
receivers and providers (the subset of plask Solver / FemSolverWithMaskedMesh /
ReceiverFor / ProviderFor the thermal and electrical solvers use).
"""
import logging

import numpy as np
from scipy.interpolate import RegularGridInterpolator

from . import material as _material
from . import mesh as _mesh

try:
    from . import _core
except ImportError as e:  # the compute path is the compiled core only
    raise ImportError("plask_amd._core not built (%s); run __graft_entry__.build()" % e)

log = logging.getLogger("plask_amd")

# Conductivity of masked-out (empty) elements: zero in-plane and 1e-30 vertical (the device
# operator takes vertical > 0); the held nodes of the masked region couple at 1e-30.
MASKED_K = (0.0, 1e-30)

# plask/log/log.hpp levels used by the solvers -> Python logging
_LEVELS = {0: logging.CRITICAL, 1: logging.ERROR, 2: logging.ERROR, 3: logging.WARNING, 4: logging.WARNING,
           5: logging.INFO, 6: logging.INFO, 7: logging.DEBUG, 8: logging.DEBUG, 9: logging.DEBUG}


def _logger(level, msg):
    log.log(_LEVELS.get(level, logging.INFO), msg)


class Receiver:
    """Receiver of a field: a constant, an array (values at the requested points), or a
    provider (any callable taking a mesh) -- the Python receiver semantics of
    python/plask/python/provider.hpp:114-305 (assignValue / connect)."""

    def __init__(self, owner, default):
        self._owner = owner
        self._value = default
        self.changed = True

    def set(self, value):
        self._value = value
        self.changed = True

    def __call__(self, mesh):
        v = self._value
        n = len(mesh)
        if callable(v):
            return np.asarray(v(mesh), dtype=np.float64).reshape(n, -1).squeeze(-1) if n else np.zeros(0)
        if np.isscalar(v):
            return np.full(n, float(v))
        a = np.asarray(v, dtype=np.float64)
        if a.shape[0] != n:
            raise ValueError("receiver data has %d values, mesh has %d points" % (a.shape[0], n))
        return a


class ReceiverAttr:
    def __init__(self, default):
        self.default = default

    def __set_name__(self, owner, name):
        self.name = "_recv_" + name

    def __get__(self, obj, objtype=None):
        if obj is None:
            return self
        r = obj.__dict__.get(self.name)
        if r is None:
            r = Receiver(obj, self.default)
            obj.__dict__[self.name] = r
        return r

    def __set__(self, obj, value):
        self.__get__(obj).set(value)
        obj._receiver_changed()


def points_of(dest):
    """(N, dim) points of a destination mesh or point list."""
    if isinstance(dest, _mesh.Rectangular):
        return dest.points()
    return np.atleast_2d(np.asarray(list(dest), dtype=np.float64))


def _same_axes(a, b):
    return len(a) == len(b) and all(len(x) == len(y) and np.array_equal(x, y) for x, y in zip(a, b))


def _corner_mean(axes, values):
    """Mean of the 2^d corner values of every element (natural order): linear
    interpolation of node values at the element midpoints."""
    v = np.asarray(values).reshape([len(a) for a in axes][::-1])
    d = len(axes)
    acc = 0.0
    for c in range(1 << d):
        acc = acc + v[tuple(slice(1, None) if (c >> a) & 1 else slice(None, -1) for a in range(d))]
    return (acc / (1 << d)).ravel()


def interp_nodes(axes, values, dest, fill, device=None):
    """Linear interpolation of node values (natural order) on a rectilinear mesh.
    Destinations that are the mesh itself or its element-midpoint mesh are
    evaluated directly (same values, no search).  values may be a callable returning
    them.  device: a solver core whose device solution holds these values; other
    destinations are then interpolated on the GPU (pf_interp_nodes)."""
    d = len(axes)
    if isinstance(dest, _mesh.Rectangular) and dest.dim == d:
        if _same_axes(dest.axes, axes):
            return np.array(values() if callable(values) else values, dtype=np.float64)
        if _same_axes(dest.axes, [0.5 * (a[1:] + a[:-1]) for a in axes]):
            return _corner_mean(axes, values() if callable(values) else values)
    if device is not None:
        pts = np.ascontiguousarray(points_of(dest), dtype=np.float64).reshape(-1, d)
        return device.interp(pts, np.nan if fill is None else float(fill))
    if callable(values):
        values = values()
    vals = np.asarray(values).reshape([len(a) for a in axes][::-1]).transpose(*range(d)[::-1])
    f = RegularGridInterpolator(axes, vals, method="linear", bounds_error=False, fill_value=np.nan)
    out = f(points_of(dest))
    if fill is not None:
        out = np.where(np.isnan(out), fill, out)
    return out


def interp_elements(axes, values, dest, inside=None, fill=0.0):
    """Linear interpolation of element (midpoint) values; points inside the mesh but
    outside the midpoint range take the edge value; points outside the mesh take fill."""
    mids = [0.5 * (a[1:] + a[:-1]) for a in axes]
    d = len(axes)
    if inside is None and isinstance(dest, _mesh.Rectangular) and dest.dim == d and _same_axes(dest.axes, mids):
        return np.array(values, dtype=np.float64)
    vals = np.asarray(values)
    comp = vals.shape[1:] if vals.ndim > 1 else ()
    grid = vals.reshape([len(m) for m in mids][::-1] + list(comp)).transpose(*(list(range(d)[::-1]) +
                                                                                list(range(d, d + len(comp)))))
    pts = points_of(dest)
    clipped = np.column_stack([np.clip(pts[:, a], mids[a][0], mids[a][-1]) for a in range(d)])
    if any(len(m) == 1 for m in mids):
        idx = tuple(np.clip(np.searchsorted(m, clipped[:, a]), 0, len(m) - 1) for a, m in enumerate(mids))
        out = grid[idx]
    else:
        f = RegularGridInterpolator(mids, grid, method="linear", bounds_error=False, fill_value=None)
        out = f(clipped)
    ok = np.all([(pts[:, a] >= axes[a][0]) & (pts[:, a] <= axes[a][-1]) for a in range(d)], axis=0)
    if inside is not None:
        ok &= inside(pts)
    out = np.where(ok.reshape((-1,) + (1,) * len(comp)), out, fill)
    return out


def nearest_elements(axes, values, dest):
    """Nearest-element lookup (INTERPOLATION_NEAREST on the element mesh)."""
    pts = points_of(dest)
    d = len(axes)
    idx = [np.clip(np.searchsorted(axes[a], pts[:, a], side="right") - 1, 0, len(axes[a]) - 2) for a in range(d)]
    m = [len(a) - 1 for a in axes]
    flat = idx[0] + m[0] * (idx[1] + (m[1] * idx[2] if d == 3 else 0))
    return np.asarray(values)[flat]


class IterativeParams(_core.IterParams):
    """``solver.iterative`` (IterativeMatrixParams, plask/common/fem/iterative_matrix.hpp:25-94).

    maxit, maxerr and noconv act as in the reference; the GPU always runs CG, so
    ``accelerator`` is recorded (any NSPCG name is accepted), and ``preconditioner`` picks
    the GPU preconditioner: 'jac' Jacobi, 'ljac'/'ljacx' the exact line solve along the
    layer axis, any other (the default 'ic' included) the multigrid V-cycle.  nfact, omega,
    ndeg, lvfill, ltrunc, nsave and nrestart are NSPCG settings, recorded only.  After a
    solve, converged / iters / err hold its outcome."""
    ACCELERATORS = ("cg", "si", "sor", "srcg", "srsi", "basic", "me", "cgnr", "lsqr", "odir", "omin", "ores",
                    "iom", "gmres", "usymlq", "usymqr", "landir", "lanmin", "lanres", "cgcr", "bcgs")
    PRECONDITIONERS = ("rich", "jac", "ljac", "ljacx", "sor", "ssor", "ic", "mic", "lsp", "neu", "lsor", "lssor",
                       "llsp", "lneu", "bic", "bicx", "mbic", "mbicx")
    NOCONV = ("error", "warning", "continue")

    def __init__(self):
        super().__init__()
        self._accel = "cg"
        self._prec = "ic"
        self.nfact, self.omega, self.ndeg, self.lvfill, self.ltrunc = 10, 1.0, 1, 0, 0
        self.nsave, self.nrestart = 5, 100000
        self._results = (True, 0, 0.0, 0.0)

    @property
    def accelerator(self):
        return self._accel

    @accelerator.setter
    def accelerator(self, v):
        v = str(v).lower()
        if v not in self.ACCELERATORS:
            raise ValueError("unknown accelerator '%s'" % v)
        self._accel = v

    @property
    def preconditioner(self):
        return self._prec

    @preconditioner.setter
    def preconditioner(self, v):
        v = str(v).lower()
        if v not in self.PRECONDITIONERS:
            raise ValueError("unknown preconditioner '%s'" % v)
        self._prec = v
        self.precond = 1 if v == "jac" else 2 if v in ("ljac", "ljacx") else 0

    @property
    def noconv(self):
        return self.NOCONV[_core.IterParams.noconv.__get__(self)]

    @noconv.setter
    def noconv(self, v):
        if isinstance(v, str):
            v = v.lower()
            if v not in self.NOCONV:
                raise ValueError("noconv must be one of %s" % ", ".join(self.NOCONV))
            v = self.NOCONV.index(v)
        _core.IterParams.noconv.__set__(self, int(v))

    def _pull(self, raw):
        self._results = (bool(raw.converged), int(raw.iters), float(raw.err), float(raw.seconds))

    converged = property(lambda self: self._results[0])
    iters = property(lambda self: self._results[1])
    err = property(lambda self: self._results[2])
    seconds = property(lambda self: self._results[3])


class FemSolver:
    """Shared geometry/mesh/boundary handling (FemSolverWithMaskedMesh, fem_solver.hpp:145-241)."""

    DIM = 2
    CYL = False
    _core_class = None

    def __init__(self, name=""):
        self.id = name
        self._geometry = None
        self._mesh = None
        self._msh = None
        self._core = None
        # the reference's defaults: Cholesky in 2D (fem_solver.hpp:32), iterative in 3D
        # (therm3d.cpp:32, electr3d.cpp:36); the GPU always iterates, the choice only sets
        # what empty_elements='default' means (fem_solver.hpp:182-188)
        self._algorithm = "cholesky" if self.DIM == 2 else "iterative"
        self._empty = "default"
        self._iter = IterativeParams()
        self._initialized = False

    @property
    def algorithm(self):
        return self._algorithm

    @algorithm.setter
    def algorithm(self, v):
        v = str(v).lower()
        if v not in ("cholesky", "gauss", "iterative"):
            raise ValueError("%s: algorithm must be 'cholesky', 'gauss' or 'iterative'" % self.id)
        self._algorithm = v
        self.invalidate()

    @property
    def empty_elements(self):
        """'default', 'include' or 'exclude' (EmptyElementsHandling, fem_solver.hpp:139-176)."""
        return self._empty

    @empty_elements.setter
    def empty_elements(self, v):
        if isinstance(v, bool):
            v = "include" if v else "exclude"
        v = {"included": "include", "excluded": "exclude"}.get(str(v).lower(), str(v).lower())
        if v not in ("default", "include", "exclude"):
            raise ValueError("%s: empty_elements must be 'default', 'include' or 'exclude'" % self.id)
        self._empty = v
        self.invalidate()

    def _excludes_empty(self):
        """Whether elements of empty materials (air) leave the mesh: 'exclude', or
        'default' with a direct algorithm (setupMaskedMesh, fem_solver.hpp:182-188)."""
        return self._empty == "exclude" or (self._empty == "default" and self._algorithm != "iterative")

    # -- geometry / mesh ------------------------------------------------------
    @property
    def geometry(self):
        return self._geometry

    @geometry.setter
    def geometry(self, g):
        if g is not None and g.dim != self.DIM:
            raise TypeError("%s: geometry dimension does not match" % self.id)
        if g is not None and bool(getattr(g, "cylindrical", False)) != self.CYL:
            raise TypeError("%s: wrong geometry type" % self.id)
        self._geometry = g
        self.invalidate()

    @property
    def mesh(self):
        if self._msh is None and self._mesh is not None and self._geometry is not None:
            m = self._mesh
            self._msh = m if isinstance(m, _mesh.Rectangular) else m(self._geometry)
        return self._msh

    @mesh.setter
    def mesh(self, m):
        self._mesh = m
        self._msh = None
        self.invalidate()

    @property
    def iterative(self):
        return self._iter

    def invalidate(self):
        self._initialized = False
        if self._core is not None:
            self._core.invalidate()

    def _run_core(self, fn, *args):
        """Call a core compute with the current solver.iterative settings and copy the
        solve's outcome back into solver.iterative."""
        self._core.iter = self._iter
        try:
            return fn(*args)
        finally:
            self._iter._pull(self._core.iter)

    def initialize(self):
        if self._initialized:
            return
        if self._geometry is None:
            raise ValueError("%s: no geometry" % self.id)
        if self.mesh is None:
            raise ValueError("%s: no mesh" % self.id)
        msh = self.mesh
        if self._core is None:
            self._core = self._core_class(self.DIM, self.CYL, 0, self.id)
            self._core.logger = _logger
        self._core.iter = self._iter
        self._core.set_mesh([np.asarray(a) for a in msh.axes])
        self._tables()
        self._on_initialize()
        self._initialized = True

    def _receiver_changed(self):
        pass

    # -- element tables --------------------------------------------------------
    def _tables(self):
        msh = self.mesh
        g = self._geometry
        self._leaf = g.leaf_table_grid(msh.element_mesh().axes)
        leaves = g._leaves
        air = _material.Air()
        self._mat_bg = air
        leaf_mat = [lf[0].material for lf in leaves] + [air]
        none = frozenset()
        leaf_roles = [frozenset(lf[0].roles) for lf in leaves] + [none]
        self._mat = [leaf_mat[k] for k in self._leaf]          # index -1: the air entry
        self._roles = [leaf_roles[k] for k in self._leaf]
        self._leaf_roles = leaf_roles
        self._axes = [np.asarray(a) for a in msh.axes]
        self._esize = [np.diff(a) for a in self._axes]
        # Masked mesh (EMPTY_ELEMENTS_EXCLUDED): elements of empty materials carry no
        # conductivity, heat or field energy, and nodes that only such elements touch are
        # held at a fixed value; the solution on the other nodes is the masked mesh's.
        n_el = len(self._mat)
        if self._excludes_empty():
            leaf_empty = np.array([lf[0].material is None or getattr(lf[0].material, "kind", "") == "empty"
                                   for lf in leaves] + [True])
            self._excl = leaf_empty[self._leaf]
            if self._excl.all():
                raise ValueError("%s: every element is empty" % self.id)
            m = [len(a) - 1 for a in self._axes][::-1]
            kept = np.zeros([k + 2 for k in m], bool)
            kept[tuple(slice(1, -1) for _ in m)] = ~self._excl.reshape(m)
            touched = np.zeros([k + 1 for k in m], bool)
            for c in range(1 << len(m)):
                off = [(c >> a) & 1 for a in range(len(m))]
                touched |= kept[tuple(slice(o, o + k + 1) for o, k in zip(off, m))]
            self._isolated = np.nonzero(~touched.ravel())[0].astype(np.int64)
        else:
            self._excl = np.zeros(n_el, bool)
            self._isolated = np.zeros(0, np.int64)

    def _thickness(self):
        """Per element the thickness of the same-material run in its column
        (ThermalFem2DSolver::onInitialize thickness scan, therm2d.cpp:81-113)."""
        m = [len(a) - 1 for a in self._axes]
        ncol = int(np.prod(m[:-1]))
        nz = m[-1]
        # material of each element as an index (same object = same material)
        mats = [lf[0].material for lf in self._geometry._leaves] + [None]
        first = {}
        leaf_mid = np.array([first.setdefault(id(x), len(first)) for x in mats], dtype=np.int64)
        lay = leaf_mid[self._leaf].reshape(nz, ncol)
        z = self._axes[-1]
        k = np.arange(nz)[:, None]
        # a run of equal materials starts where the material changes going up and ends
        # below the next change; every element takes its run's height
        starts = np.ones((nz, ncol), bool)
        starts[1:] = lay[1:] != lay[:-1]
        ends = np.ones((nz, ncol), bool)
        ends[:-1] = lay[1:] != lay[:-1]
        s = np.maximum.accumulate(np.where(starts, k, 0), axis=0)
        e = np.minimum.accumulate(np.where(ends, k, nz)[::-1], axis=0)[::-1]
        return (z[e + 1] - z[s]).ravel()

    def _element_volume(self):
        s = self._esize
        if self.DIM == 2:
            return (s[0][None, :] * s[1][:, None]).ravel()
        return (s[0][None, None, :] * s[1][None, :, None] * s[2][:, None, None]).ravel()

    def _element_mesh(self):
        return self.mesh.element_mesh()
