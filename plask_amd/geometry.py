"""Layer-stack geometry: the subset of PLaSK's geometry the thermal/electrical path reads.

The solvers ask a geometry two things per element midpoint: the material
(``getMaterial``) and the roles (``getRolesAt``).  This module provides blocks
(``Rectangle``/``Cuboid``), vertical stacks (``Stack2D``/``Stack3D``; ``append`` adds
on top, ``prepend`` at the bottom), horizontal shelves (``Shelf``) and the spaces
``Cartesian2D``, ``Cylindrical``, ``Cartesian3D``.  Points outside every object have
the background material air (kind 'empty'), as in the reference spaces.
"""
import numpy as np

from . import material as _mat


class GeometryObject:
    role = None

    @property
    def roles(self):
        r = self.role
        if r is None:
            return set()
        if isinstance(r, str):
            return {x.strip() for x in r.split(",") if x.strip()}
        return set(r)


class Block(GeometryObject):
    def __init__(self, size, material, role=None):
        self.size = tuple(float(s) for s in size)
        self.material = _mat.get(material) if material is not None else None
        self.role = role

    def bbox_size(self):
        return self.size

    def leaves(self, origin, acc, path=()):
        acc.append((self, tuple(origin), path))


def Rectangle(width, height, material, role=None):
    return Block((width, height), material, role)


def Cuboid(dx, dy, dz, material, role=None):
    return Block((dx, dy, dz), material, role)


class Placement:
    """Handle of one placement of an object in a container (returned by append/prepend);
    pass it as the path of TopOf/BottomOf to select that placement."""

    def __init__(self, obj):
        self.obj = obj

    def __getattr__(self, a):
        return getattr(self.obj, a)


class Stack(GeometryObject):
    """Vertical stack; items are aligned at the lower lateral edge (2D) or centred
    at xcenter/ycenter when given (3D, Stack3D(xcenter=0, ycenter=0))."""

    def __init__(self, dim, xcenter=None, ycenter=None, role=None):
        self.dim = dim
        self.items = []  # bottom -> top
        self.center = (xcenter, ycenter)
        self.role = role

    def append(self, obj):
        h = Placement(obj)
        self.items.append(h)
        return h

    def prepend(self, obj):
        h = Placement(obj)
        self.items.insert(0, h)
        return h

    def bbox_size(self):
        sizes = [it.obj.bbox_size() for it in self.items]
        lat = [max((s[a] for s in sizes), default=0.0) for a in range(self.dim - 1)]
        return tuple(lat) + (sum(s[-1] for s in sizes),)

    def _offset(self, size, a):
        c = self.center[a] if a < len(self.center) else None
        if c is None:
            return 0.0
        full = self.bbox_size()[a]
        return (full - size[a]) / 2.0  # centred

    def leaves(self, origin, acc, path=()):
        z = origin[-1]
        full = self.bbox_size()
        shift = [-(full[a] / 2.0) if (a < len(self.center) and self.center[a] is not None) else 0.0
                 for a in range(self.dim - 1)]
        for it in self.items:
            s = it.obj.bbox_size()
            o = [origin[a] + shift[a] + self._offset(s, a) for a in range(self.dim - 1)] + [z]
            it.obj.leaves(o, acc, path + (it,))
            z += s[-1]


def Stack2D(role=None):
    return Stack(2, role=role)


def Stack3D(xcenter=None, ycenter=None, role=None):
    return Stack(3, xcenter, ycenter, role=role)


class Shelf(GeometryObject):
    """Horizontal arrangement along axis 0 (with gaps), aligned at the bottom."""

    def __init__(self, role=None):
        self.items = []
        self.role = role

    def append(self, obj):
        self.items.append(obj)
        return obj

    def append_gap(self, width):
        self.items.append(float(width))

    def bbox_size(self):
        w = sum(it if isinstance(it, float) else it.bbox_size()[0] for it in self.items)
        h = max((it.bbox_size()[1] for it in self.items if not isinstance(it, float)), default=0.0)
        return (w, h)

    def leaves(self, origin, acc, path=()):
        x = origin[0]
        for it in self.items:
            if isinstance(it, float):
                x += it
                continue
            it.leaves((x, origin[1]), acc, path)
            x += it.bbox_size()[0]


class Space:
    """Calculation space over a root object; the leaf table is built once."""

    cylindrical = False

    def __init__(self, root, dim, length=None):
        self.root = root
        self.dim = dim
        self.length = float(length) if length is not None else 1.0
        acc = []
        root.leaves(tuple([0.0] * dim), acc)
        self._leaves = [(obj, np.array(org), np.array(obj.size)) for obj, org, _ in acc]
        self._paths = [path for _, _, path in acc]
        size = root.bbox_size()
        lo = np.min([o for _, o, _ in self._leaves], axis=0) if self._leaves else np.zeros(dim)
        self.bbox = (lo, lo + np.array(size))

    def leaf_at(self, pt):
        pt = np.asarray(pt, dtype=float)
        for obj, o, s in self._leaves:
            if np.all(pt >= o) and np.all(pt < o + s):
                return obj
        return None

    def get_material(self, *pt):
        pt = pt[0] if len(pt) == 1 else pt
        leaf = self.leaf_at(pt)
        return leaf.material if leaf is not None else _mat.Air()

    def get_roles(self, *pt):
        pt = pt[0] if len(pt) == 1 else pt
        leaf = self.leaf_at(pt)
        roles = set()
        if leaf is not None:
            roles |= leaf.roles
        return roles

    def has_role(self, role, *pt):
        return role in self.get_roles(*pt)

    def leaves_of(self, obj, path=None):
        """(origin, size) of the bounding box of every placement of obj (restricted to
        placements under `path`, a Placement handle, when given)."""
        if isinstance(obj, Placement):
            path, obj = obj, obj.obj
        acc = []
        obj.leaves(tuple([0.0] * self.dim), acc)
        kids = {id(o) for o, _, _ in acc}
        boxes = {}
        for (leaf, o, s), pth in zip(self._leaves, self._paths):
            if id(leaf) not in kids and leaf is not obj:
                continue
            if path is not None and not any(p is path for p in pth):
                continue
            # group the leaves of one placement of obj by the path prefix that places obj
            key = tuple(id(p) for p in pth if p.obj is obj) or (id(leaf), tuple(o))
            lo, hi = boxes.get(key, (o, o + s))
            boxes[key] = (np.minimum(lo, o), np.maximum(hi, o + s))
        return [(lo, hi - lo) for lo, hi in boxes.values()]

    def leaf_table(self, points):
        """Vectorized lookup: index into self._leaves per point (-1 outside)."""
        pts = np.asarray(points, dtype=float)
        idx = np.full(len(pts), -1, dtype=np.int64)
        for k, (obj, o, s) in enumerate(self._leaves):
            inside = np.all((pts >= o) & (pts < o + s), axis=1) & (idx < 0)
            idx[inside] = k
        return idx

    def leaf_table_grid(self, axes):
        """leaf_table of the points of a rectilinear grid (natural order, axis 0
        fastest), from per-axis interval tests: a leaf is a box."""
        axes = [np.asarray(a, dtype=float) for a in axes]
        shape = [len(a) for a in axes][::-1]
        idx = np.full(shape, -1, dtype=np.int64)
        d = len(axes)
        for k, (obj, o, s) in enumerate(self._leaves):
            inside = np.ones(shape, bool)
            for a in range(d):
                m = (axes[a] >= o[a]) & (axes[a] < o[a] + s[a])
                view = [1] * d
                view[d - 1 - a] = len(axes[a])
                inside &= m.reshape(view)
            inside &= idx < 0
            idx[inside] = k
        return idx.ravel()


class Cartesian2D(Space):
    def __init__(self, root, length=None):
        super().__init__(root, 2, length)


class Cylindrical(Space):
    cylindrical = True

    def __init__(self, root):
        super().__init__(root, 2)


Cylindrical2D = Cylindrical


class Cartesian3D(Space):
    def __init__(self, root):
        super().__init__(root, 3)
