"""Materials: the interface of PLaSK's Material the thermal/electrical path calls.

``thermk(T, h)`` -> conductivity (W/m/K, scalar or (in-plane, vertical)),
``cond(T)`` -> electrical conductivity (S/m, scalar or pair), ``eps(T)``,
``cp(T)`` -> specific heat (J/kg/K) and ``dens(T)`` -> density (kg/m^3), which the
transient solver reads.  (plask/material/material.hpp:649,691,698,705.)

Laws the GPU evaluates itself (no host round trip inside a Picard loop):
``Const`` and power laws k = k0 (300/T)^alpha (``PowerLaw``).  Any other Python
material is evaluated on the host per element at the element temperatures.
"""
import math


class Material:
    name = "material"
    kind = "material"  # 'empty' for air-like materials

    def thermk(self, T=300.0, h=math.inf):
        raise NotImplementedError("%s: thermk not defined" % self.name)

    def cond(self, T=300.0):
        raise NotImplementedError("%s: cond not defined" % self.name)

    def eps(self, T=300.0):
        return 1.0

    def cp(self, T=300.0):
        raise NotImplementedError("%s: cp not defined" % self.name)

    def dens(self, T=300.0):
        raise NotImplementedError("%s: dens not defined" % self.name)

    # device law: (k0_inplane, k0_vertical, alpha_inplane, alpha_vertical) or None
    def thermk_law(self):
        return None

    # cp * dens when it does not depend on T (evaluated once per material), else None
    def heat_capacity_const(self):
        return None

    def cond_law(self):
        return None

    def __repr__(self):
        return self.name


def simple(*args, **kwargs):
    """Decorator registering a Python material class (``@material.simple()``)."""

    def deco(cls):
        inst = cls()
        inst.name = cls.__name__
        _registry[cls.__name__] = inst
        return cls

    return deco


def _pair(v):
    if isinstance(v, (tuple, list)):
        return float(v[0]), float(v[1])
    return float(v), float(v)


class Const(Material):
    """Constant (possibly anisotropic) conductivities."""

    def __init__(self, thermk=None, cond=None, eps=1.0, name="const", cp=None, dens=None):
        self._k = _pair(thermk) if thermk is not None else None
        self._c = _pair(cond) if cond is not None else None
        self._eps = eps
        self._cp, self._dens = cp, dens
        self.name = name

    def cp(self, T=300.0):
        if self._cp is None:
            raise NotImplementedError("%s: cp not defined" % self.name)
        return float(self._cp)

    def dens(self, T=300.0):
        if self._dens is None:
            raise NotImplementedError("%s: dens not defined" % self.name)
        return float(self._dens)

    def thermk(self, T=300.0, h=math.inf):
        if self._k is None:
            raise NotImplementedError("%s: thermk not defined" % self.name)
        return self._k

    def cond(self, T=300.0):
        if self._c is None:
            raise NotImplementedError("%s: cond not defined" % self.name)
        return self._c

    def eps(self, T=300.0):
        return self._eps

    def thermk_law(self):
        return None if self._k is None else (self._k[0], self._k[1], 0.0, 0.0)

    def heat_capacity_const(self):
        return None if self._cp is None or self._dens is None else float(self._cp) * float(self._dens)

    def cond_law(self):
        return None if self._c is None else (self._c[0], self._c[1], 0.0, 0.0)


class PowerLaw(Material):
    """k(T) = k0 (300/T)^alpha for thermal conductivity (and optional constant cond)."""

    def __init__(self, k0, alpha, cond=None, eps=1.0, name="powerlaw", cp=None, dens=None):
        self.k0 = _pair(k0)
        self.alpha = _pair(alpha)
        self._c = _pair(cond) if cond is not None else None
        self._eps = eps
        self._cp, self._dens = cp, dens
        self.name = name

    def cp(self, T=300.0):
        if self._cp is None:
            raise NotImplementedError("%s: cp not defined" % self.name)
        return float(self._cp)

    def dens(self, T=300.0):
        if self._dens is None:
            raise NotImplementedError("%s: dens not defined" % self.name)
        return float(self._dens)

    def thermk(self, T=300.0, h=math.inf):
        t = 300.0 / T
        return (self.k0[0] * t ** self.alpha[0], self.k0[1] * t ** self.alpha[1])

    def cond(self, T=300.0):
        if self._c is None:
            raise NotImplementedError("%s: cond not defined" % self.name)
        return self._c

    def eps(self, T=300.0):
        return self._eps

    def thermk_law(self):
        return (self.k0[0], self.k0[1], self.alpha[0], self.alpha[1])

    def heat_capacity_const(self):
        return None if self._cp is None or self._dens is None else float(self._cp) * float(self._dens)

    def cond_law(self):
        return None if self._c is None else (self._c[0], self._c[1], 0.0, 0.0)


class _GaAs(PowerLaw):
    """GaAs with the commonly used k = 45 (300/T)^1.28 W/m/K and eps = 12.9, cp = 327 J/kg/K
    and density 5317.49 kg/m^3 (a stand-in for the reference material database, which is
    out of scope)."""

    def __init__(self):
        super().__init__((45.0, 45.0), (1.28, 1.28), cond=(1e-3, 1e-3), eps=12.9, name="GaAs", cp=0.327e3,
                         dens=5.31749e3)


class _Air(Material):
    """The background material of every space (plask/material/air.cpp): kind EMPTY,
    cond 0.55e-14 S/m, thermk 0.0258 (T/300)^0.84 W/m/K, cp 1007 J/kg/K, density
    1.161 kg/m^3 (air.cpp:52-94)."""
    name = "air"
    kind = "empty"

    def thermk(self, T=300.0, h=math.inf):
        k = 0.0258 * (T / 300.0) ** 0.84
        return (k, k)

    def cond(self, T=300.0):
        return (0.55e-14, 0.55e-14)

    def eps(self, T=300.0):
        return 1.0

    def cp(self, T=300.0):
        return 1.007e3

    def dens(self, T=300.0):
        return 1.161

    def thermk_law(self):
        return (0.0258, 0.0258, -0.84, -0.84)

    def heat_capacity_const(self):
        return 1.007e3 * 1.161

    def cond_law(self):
        return (0.55e-14, 0.55e-14, 0.0, 0.0)


_registry = {"GaAs": _GaAs(), "air": _Air(), "Air": _Air()}


def get(m):
    if isinstance(m, Material):
        return m
    if isinstance(m, type) and issubclass(m, Material):
        return m()
    if isinstance(m, str):
        if m in _registry:
            return _registry[m]
        raise KeyError("unknown material '%s'" % m)
    raise TypeError("not a material: %r" % (m,))


def GaAs():
    return _registry["GaAs"]


def Air():
    return _registry["air"]
