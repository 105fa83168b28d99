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
