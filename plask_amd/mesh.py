"""Rectilinear meshes and boundary places (the subset of plask.mesh the path uses).

``Rectangular2D(axis0, axis1)`` / ``Rectangular3D(axis0, axis1, axis2)``: node
coordinates per axis (um).  ``mesh.elements.mesh`` is the element-midpoint mesh.
Boundary places (``Bottom()``, ``Top()``, ``Left()``, ``Right()``, ``TopOf(obj)``,
``BottomOf(obj)`` ...) resolve to node sets of a mesh in a geometry
(plask/mesh/boundary.hpp; the reference's rectangular boundaries).
``DivideGenerator(prediv=...)`` builds a mesh from the object edges of a geometry,
each interval split into ``prediv`` parts (a simplified divide generator; the reference
generator plask/mesh/generator_rectangular.cpp also refines gradually).
"""
import numpy as np


class _Elements:
    def __init__(self, mesh):
        self._mesh = mesh

    @property
    def mesh(self):
        return self._mesh.element_mesh()

    def __len__(self):
        return int(np.prod([len(a) - 1 for a in self._mesh.axes]))


class Rectangular:
    def __init__(self, *axes, ordering=None):
        self.axes = [np.asarray(sorted(set(float(v) for v in a)), dtype=np.float64) for a in axes]
        self.dim = len(axes)
        self.ordering = ordering

    def __len__(self):
        return int(np.prod([len(a) for a in self.axes]))

    @property
    def axis0(self):
        return self.axes[0]

    @property
    def axis1(self):
        return self.axes[1]

    @property
    def axis2(self):
        return self.axes[2]

    @property
    def elements(self):
        return _Elements(self)

    def element_mesh(self):
        return type(self)(*[0.5 * (a[1:] + a[:-1]) for a in self.axes])

    def points(self):
        """All mesh points, natural order (axis 0 fastest), shape (N, dim), read-only.
        Kept with the mesh (the axis arrays are not modified in place; assigning a new
        axis array recomputes them)."""
        key = tuple((id(a), len(a)) for a in self.axes)
        cached = getattr(self, "_points", None)
        if cached is not None and cached[0] == key:
            return cached[1]
        n = [len(a) for a in self.axes]
        pts = np.empty((int(np.prod(n)), self.dim))
        for a in range(self.dim):
            shape = [1] * self.dim
            shape[self.dim - 1 - a] = n[a]   # axis 0 varies fastest
            pts[:, a] = np.broadcast_to(self.axes[a].reshape(shape), n[::-1]).ravel()
        pts.flags.writeable = False
        self._points = (key, pts)
        return pts

    def __iter__(self):
        return iter(self.points())

    # boundary places -------------------------------------------------------
    @staticmethod
    def Bottom():
        return Boundary("bottom")

    @staticmethod
    def Top():
        return Boundary("top")

    @staticmethod
    def Left():
        return Boundary("left")

    @staticmethod
    def Right():
        return Boundary("right")

    @staticmethod
    def Front():
        return Boundary("front")

    @staticmethod
    def Back():
        return Boundary("back")

    @staticmethod
    def TopOf(obj, path=None):
        return Boundary("topof", (obj, path))

    @staticmethod
    def BottomOf(obj, path=None):
        return Boundary("bottomof", (obj, path))

    @staticmethod
    def LeftOf(obj, path=None):
        return Boundary("leftof", (obj, path))

    @staticmethod
    def RightOf(obj, path=None):
        return Boundary("rightof", (obj, path))

    @staticmethod
    def FrontOf(obj, path=None):
        return Boundary("frontof", (obj, path))

    @staticmethod
    def BackOf(obj, path=None):
        return Boundary("backof", (obj, path))

    @staticmethod
    def VerticalLineNear(at, start, stop):
        """2D: the mesh line of axis 0 nearest `at`, nodes from `start` to `stop` along axis 1
        (createVerticalBoundaryNear, plask/mesh/rectangular2d.cpp:140-145)."""
        return Boundary("vline", (float(at), float(start), float(stop)))

    @staticmethod
    def HorizontalLineNear(at, start, stop):
        """2D: the mesh line of axis 1 nearest `at`, nodes from `start` to `stop` along axis 0."""
        return Boundary("hline", (float(at), float(start), float(stop)))

    @staticmethod
    def Predicate(fn, vectorized=False):
        """Nodes where fn(point) is true; vectorized=True: fn takes the (N, dim)
        array of all points and returns N booleans."""
        return Boundary("predicate", (fn, bool(vectorized)))


class Rectangular2D(Rectangular):
    def __init__(self, axis0, axis1, ordering="01"):
        super().__init__(axis0, axis1, ordering=ordering)

    DivideGenerator = None  # set below


class Rectangular3D(Rectangular):
    def __init__(self, axis0, axis1, axis2, ordering=None):
        super().__init__(axis0, axis1, axis2, ordering=ordering)

    DivideGenerator = None


class Boundary:
    def __init__(self, kind, arg=None):
        self.kind = kind
        self.arg = arg

    def nodes(self, mesh, geometry=None):
        """Node indices (natural order) of this place."""
        n = [len(a) for a in mesh.axes]
        d = mesh.dim
        vert = d - 1
        if self.kind in ("bottom", "top", "left", "right", "back", "front"):
            # the nodes with index 0 (bottom, left, back) or n - 1 (top, right, front) along
            # one axis, in natural order (axis 0 fastest)
            if self.kind in ("bottom", "top"):
                ax = vert
            elif d == 3:
                ax = {"back": 0, "front": 0, "left": 1, "right": 1}[self.kind]
            else:
                ax = 0
            pos = 0 if self.kind in ("bottom", "left", "back") else n[ax] - 1
            strides = np.cumprod([1] + n[:-1])
            out = np.array([pos * strides[ax]], dtype=np.int64)
            for a in range(d - 1, -1, -1):   # outer axes first
                if a != ax:
                    out = (out[:, None] + strides[a] * np.arange(n[a], dtype=np.int64)[None, :]).ravel()
            return out
        if self.kind == "predicate":
            fn, vec = self.arg if isinstance(self.arg, tuple) else (self.arg, False)
            pts = mesh.points()
            if vec:
                ok = np.asarray(fn(pts), dtype=bool)
                if ok.shape != (len(pts),):
                    raise ValueError("vectorized predicate must return one boolean per point")
                return np.nonzero(ok)[0].astype(np.int64)
            return np.nonzero([bool(fn(p)) for p in pts])[0].astype(np.int64)
        if self.kind in ("leftof", "rightof", "frontof", "backof"):
            if geometry is None:
                raise ValueError("%s needs a geometry" % self.kind)
            if d == 3:
                ax = {"backof": 0, "frontof": 0, "leftof": 1, "rightof": 1}[self.kind]
            else:
                if self.kind in ("frontof", "backof"):
                    raise ValueError("%s is a 3D place" % self.kind)
                ax = 0
            hi = self.kind in ("rightof", "frontof")
            out = []
            for o, s in geometry.leaves_of(*self.arg):
                got = _box_face(mesh.axes, ax, hi, o, o + s)
                if got is not None:
                    out.append(got)
            return np.unique(np.concatenate(out)).astype(np.int64) if out else np.zeros(0, np.int64)
        if self.kind in ("vline", "hline"):
            if d != 2:
                raise ValueError("line places are 2D")
            at, start, stop = self.arg
            ax = 0 if self.kind == "vline" else 1   # the axis the line is at
            other = 1 - ax
            line = int(np.argmin(np.abs(mesh.axes[ax] - at)))
            lo, hi = _bounds_ext(mesh.axes[other], start, stop)
            if lo >= hi:
                return np.zeros(0, np.int64)
            idx = np.arange(lo, hi)
            return (line + n[0] * idx if ax == 0 else idx + n[0] * line).astype(np.int64)
        if self.kind in ("topof", "bottomof"):
            if geometry is None:
                raise ValueError("TopOf/BottomOf need a geometry")
            out = []
            pts = mesh.points()
            for o, s in geometry.leaves_of(*self.arg):
                zlev = o[-1] + (s[-1] if self.kind == "topof" else 0.0)
                ok = np.isclose(pts[:, vert], zlev, rtol=0, atol=1e-9 * max(1.0, abs(zlev)))
                for a in range(d - 1):
                    ok &= (pts[:, a] >= o[a] - 1e-9) & (pts[:, a] <= o[a] + s[a] + 1e-9)
                out.append(np.nonzero(ok)[0])
            return np.unique(np.concatenate(out)).astype(np.int64) if out else np.zeros(0, np.int64)
        raise ValueError("unknown boundary %s" % self.kind)


def _find_index(axis, x):
    """Index of the first coordinate >= x (MeshAxis::findIndex)."""
    return int(np.searchsorted(axis, x, side="left"))


def _bounds(axis, lo, hi):
    """[begin, end) of the coordinates in [lo, hi] (getIndexesInBounds,
    plask/mesh/rectangular_common.hpp:72-78)."""
    if lo > hi:
        return 0, 0
    b, e = _find_index(axis, lo), _find_index(axis, hi)
    if e != len(axis) and axis[e] == hi:
        e += 1
    return b, e


def _bounds_ext(axis, lo, hi):
    """_bounds widened by one coordinate at either end that lies within 1/100 of its
    interval of the bound (getIndexesInBoundsExt with tryMakeLower / tryMakeHigher,
    rectangular_common.hpp:86-115)."""
    b, e = _bounds(axis, lo, hi)
    if 0 < b < len(axis) and (lo - axis[b - 1]) * 100.0 < (axis[b] - axis[b - 1]):
        b -= 1
    if 0 < e < len(axis) and (axis[e] - hi) * 100.0 < (axis[e] - axis[e - 1]):
        e += 1
    return b, e


def _box_face(axes, ax, hi, lo_pt, hi_pt):
    """Nodes of the mesh line along `ax` at a box's low (getLineLo) or high (getLineHi)
    face, over the box's extent on the other axes; None when the box holds no line."""
    a = axes[ax]
    if hi:
        line = _find_index(a, hi_pt[ax])
        if not (line != len(a) and a[line] == hi_pt[ax]):
            if line == 0:
                return None
            line -= 1
            if a[line] < lo_pt[ax]:
                return None
    else:
        line = _find_index(a, lo_pt[ax])
        if line == len(a) or a[line] > hi_pt[ax]:
            return None
    ranges = []
    for o in range(len(axes)):
        if o == ax:
            ranges.append(np.array([line]))
        else:
            b, e = _bounds(axes[o], lo_pt[o], hi_pt[o])
            if b >= e:
                return None
            ranges.append(np.arange(b, e))
    n = [len(x) for x in axes]
    strides = np.cumprod([1] + n[:-1])
    out = np.zeros(1, np.int64)
    for o in range(len(axes) - 1, -1, -1):
        out = (out[:, None] + strides[o] * ranges[o][None, :]).ravel()
    return out


class BoundaryConditions:
    """Ordered list of (place, value) (plask/mesh/boundary_conditions.hpp:95-186)."""

    def __init__(self):
        self._items = []

    def append(self, place, value):
        self._items.append((place, value))

    def insert(self, i, place, value):
        self._items.insert(i, (place, value))

    def clear(self):
        self._items = []

    def __len__(self):
        return len(self._items)

    def __iter__(self):
        return iter(self._items)

    def __getitem__(self, i):
        place, value = self._items[i]

        class _C:
            pass

        c = _C()
        c.place, c.value = place, value
        return c

    def resolve(self, mesh, geometry):
        """(nodes, values) in list order; a node in several places takes the last value
        (setBC applied in order, plask/common/fem/matrix.hpp:111-118)."""
        nodes, vals = [], []
        for place, value in self._items:
            nd = place.nodes(mesh, geometry)
            nodes.append(nd)
            vals.append(np.full(len(nd), float(value)))
        if not nodes:
            return np.zeros(0, np.int64), np.zeros(0)
        return np.concatenate(nodes), np.concatenate(vals)


class DivideGenerator:
    def __init__(self, prediv=None, postdiv=None, gradual=True, **kw):
        self.prediv = prediv
        self.postdiv = postdiv

    def __call__(self, geometry):
        return self.generate(geometry)

    def generate(self, geometry):
        d = geometry.dim
        edges = [set() for _ in range(d)]
        for _, o, s in geometry._leaves:
            for a in range(d):
                edges[a].add(round(float(o[a]), 12))
                edges[a].add(round(float(o[a] + s[a]), 12))
        if getattr(geometry, "cylindrical", False):
            edges[0].add(0.0)
        div = self.prediv if self.prediv is not None else 1
        if isinstance(div, int):
            div = (div,) * d
        axes = []
        for a in range(d):
            e = np.array(sorted(edges[a]))
            pts = [e[0]]
            for lo, hi in zip(e[:-1], e[1:]):
                m = max(1, int(div[a]))
                pts.extend(list(lo + (hi - lo) * np.arange(1, m + 1) / m))
            axes.append(np.array(pts))
        return Rectangular2D(*axes) if d == 2 else Rectangular3D(*axes)


Rectangular2D.DivideGenerator = DivideGenerator
Rectangular3D.DivideGenerator = DivideGenerator
