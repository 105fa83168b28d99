"""CPU statement of the thermal boundary terms (heat flux, radiation, convection).

TEST INFRASTRUCTURE ONLY (imported by ``tests/``).  Restates the load-vector part of
``setBoundaries`` as the thermal ``setMatrix`` calls it:

* 2D (solvers/thermal/static/therm2d.cpp:139-173): every element side (bottom, right,
  top, left) whose two nodes both hold a condition adds ``-0.5e-6 * len * value`` to
  each of them (:232-240 flux, :256-265 radiation); cylindrical (:358-368, :394-406):
  times the side radius on r = const sides and ``r_mid -/+ len/6`` along r (the node
  with the smaller index, i.e. the smaller r, takes the minus sign).
* 3D (therm3d.cpp:141-168, :242-249, :261-268): every element wall whose four nodes
  hold a condition adds ``-0.25e-12 * area * value`` of its i-th node to the element's
  local node i (0-3) -- the reference's indexing, kept; the radiation term reads the
  temperature of mesh node ``wall[i]`` (the local number), also kept.
* radiation value ``emissivity * SB * (T^4 - T_ambient^4)``, SB from
  plask/phys/constants.hpp:41; a node in several places takes the first one's value
  (boundary_conditions.hpp:182-186).

Loads are added to the right-hand side before the Dirichlet elimination, which then
overwrites the Dirichlet rows: ``with_load`` adds them to an eliminated system at the
free nodes, the same thing.

Convection (``convection_2d`` / ``convection_3d``) also adds matrix terms, which take part
in the elimination: ``solve_with_matrix`` assembles K + Kc, F + Fc densely and eliminates.
* 2D (therm2d.cpp:242-254): load ``0.5e-6 len coeff_n ambient_n``, matrix
  ``(coeff_a + coeff_b) len / 6`` on both diagonals and ``/ 12`` off it -- without the
  1e-6 of the load (the reference's unit mismatch, kept);
* cylindrical (:370-392): load ``0.125e-6 len (coeff_a + coeff_b)(ambient_a + ambient_b)``
  and the same matrix terms, times the side radius on r = const sides, and along r
  ``r_mid -/+ len/6`` (load, diagonal) or ``r_mid`` (off-diagonal);
* 3D (therm3d.cpp:250-259): load ``0.25e-12 area coeff ambient``, matrix
  ``0.125e-12 area (coeff_i + coeff_j) / (9 | 18 | 36)`` (same node, wall edge, wall
  diagonal), placed on the element's local nodes 0-3 like the flux.
"""
import numpy as np

SB = 5.670373e-8


def first_values(conditions):
    """{node: value} from [(nodes, value), ...] in list order, first place wins."""
    out = {}
    for nodes, value in conditions:
        for n in np.asarray(nodes, dtype=np.int64).tolist():
            out.setdefault(n, value)
    return out


def _radiated(v, t):
    emissivity, ambient = v
    return emissivity * SB * (t ** 4 - ambient ** 4)


def load_2d(axes, flux=None, radiation=None, T=None, cyl=False):
    """Boundary load vector of a 2D grid (natural order, axis 0 fastest)."""
    r_ax, z_ax = np.asarray(axes[0], float), np.asarray(axes[1], float)
    n0, n1 = len(r_ax), len(z_ax)
    F = np.zeros(n0 * n1)
    kinds = []
    if flux:
        kinds.append((flux, lambda v, node: v))
    if radiation:
        kinds.append((radiation, lambda v, node: _radiated(v, T[node])))
    for j in range(n1 - 1):
        for i in range(n0 - 1):
            lolo = j * n0 + i
            uplo = lolo + 1
            loup = lolo + n0
            upup = loup + 1
            width = r_ax[i + 1] - r_ax[i]
            height = z_ax[j + 1] - z_ax[j]
            rmid = 0.5 * (r_ax[i] + r_ax[i + 1])
            sides = (("bottom", lolo, uplo, width), ("right", uplo, upup, height),
                     ("top", upup, loup, width), ("left", loup, lolo, height))
            for cond, val in kinds:
                for side, a, b, L in sides:
                    if a not in cond or b not in cond:
                        continue
                    for n, other in ((a, b), (b, a)):
                        w = 1.0
                        if cyl:
                            if side == "left":
                                w = r_ax[i]
                            elif side == "right":
                                w = r_ax[i + 1]
                            else:
                                w = rmid - L / 6.0 if n < other else rmid + L / 6.0
                        F[n] += -0.5e-6 * L * val(cond[n], n) * w
    return F


_WALLS = ((0, 1, 2, 3), (4, 5, 6, 7), (0, 2, 4, 6), (1, 3, 5, 7), (0, 1, 4, 5), (2, 3, 6, 7))


def load_3d(axes, flux=None, radiation=None, T=None):
    """Boundary load vector of a 3D grid (natural order, axis 0 fastest)."""
    ax = [np.asarray(a, float) for a in axes]
    n0, n1, n2 = (len(a) for a in ax)
    F = np.zeros(n0 * n1 * n2)
    kinds = []
    if flux:
        kinds.append((flux, lambda v, local: v))
    if radiation:
        kinds.append((radiation, lambda v, local: _radiated(v, T[local])))
    for k in range(n2 - 1):
        for j in range(n1 - 1):
            for i in range(n0 - 1):
                idx = [(i + (l & 1)) + n0 * ((j + ((l >> 1) & 1)) + n1 * (k + (l >> 2))) for l in range(8)]
                dx = ax[0][i + 1] - ax[0][i]
                dy = ax[1][j + 1] - ax[1][j]
                dz = ax[2][k + 1] - ax[2][k]
                areas = (dx * dy, dy * dz, dz * dx)
                for cond, val in kinds:
                    for side, wall in enumerate(_WALLS):
                        if not all(idx[w] in cond for w in wall):
                            continue
                        for q in range(4):
                            F[idx[q]] += -0.25e-12 * areas[side // 2] * val(cond[idx[wall[q]]], wall[q])
    return F


def with_load(B, F, dirichlet_nodes):
    """Eliminated right-hand side plus a boundary load at the free nodes."""
    out = np.array(B, dtype=np.float64)
    free = np.ones(len(out), bool)
    free[np.asarray(dirichlet_nodes, dtype=np.int64)] = False
    out[free] += F[free]
    return out


def convection_2d(axes, conv, cyl=False):
    """Convection load vector and lower-triangle matrix entries [(row, col, value)] of a 2D
    grid; conv = {node: (coeff, ambient)}."""
    r_ax, z_ax = np.asarray(axes[0], float), np.asarray(axes[1], float)
    n0, n1 = len(r_ax), len(z_ax)
    F = np.zeros(n0 * n1)
    entries = []
    for j in range(n1 - 1):
        for i in range(n0 - 1):
            lolo = j * n0 + i
            uplo, loup = lolo + 1, lolo + n0
            upup = loup + 1
            width = r_ax[i + 1] - r_ax[i]
            height = z_ax[j + 1] - z_ax[j]
            rmid = 0.5 * (r_ax[i] + r_ax[i + 1])
            sides = (("bottom", lolo, uplo, width), ("right", uplo, upup, height),
                     ("top", upup, loup, width), ("left", loup, lolo, height))
            for side, a, b, L in sides:
                if a not in conv or b not in conv:
                    continue
                (ca, ta), (cb, tb) = conv[a], conv[b]
                if side == "left":
                    wa = wb = woff = r_ax[i]
                elif side == "right":
                    wa = wb = woff = r_ax[i + 1]
                else:
                    wa = rmid - L / 6.0 if a < b else rmid + L / 6.0
                    wb = rmid - L / 6.0 if b < a else rmid + L / 6.0
                    woff = rmid
                if not cyl:
                    wa = wb = woff = 1.0
                    F[a] += 0.5e-6 * L * ca * ta
                    F[b] += 0.5e-6 * L * cb * tb
                else:
                    f = 0.125e-6 * L * (ca + cb) * (ta + tb)
                    F[a] += f * wa
                    F[b] += f * wb
                entries.append((a, a, (ca + cb) * L / 6.0 * wa))
                entries.append((b, b, (cb + ca) * L / 6.0 * wb))
                entries.append((max(a, b), min(a, b), (ca + cb) * L / 12.0 * woff))
    return F, entries


def convection_3d(axes, conv):
    """Convection load vector and lower-triangle matrix entries of a 3D grid."""
    ax = [np.asarray(a, float) for a in axes]
    n0, n1, n2 = (len(a) for a in ax)
    F = np.zeros(n0 * n1 * n2)
    entries = []
    for k in range(n2 - 1):
        for j in range(n1 - 1):
            for i in range(n0 - 1):
                idx = [(i + (l & 1)) + n0 * ((j + ((l >> 1) & 1)) + n1 * (k + (l >> 2))) for l in range(8)]
                dx = ax[0][i + 1] - ax[0][i]
                dy = ax[1][j + 1] - ax[1][j]
                dz = ax[2][k + 1] - ax[2][k]
                areas = (dx * dy, dy * dz, dz * dx)
                for side, wall in enumerate(_WALLS):
                    if not all(idx[w] in conv for w in wall):
                        continue
                    area = areas[side // 2]
                    for p in range(4):
                        cp_, tp = conv[idx[wall[p]]]
                        F[idx[p]] += 0.25e-12 * area * cp_ * tp
                        for q in range(p + 1):
                            cq = conv[idx[wall[q]]][0]
                            d = 9.0 if p == q else (18.0 if (p ^ q) in (1, 2) else 36.0)
                            entries.append((max(idx[p], idx[q]), min(idx[p], idx[q]),
                                            0.125e-12 * area * (cp_ + cq) / d))
    return F, entries


def solve_with_matrix(K, F, entries, Fb, dnodes, dvals):
    """Dense K + boundary entries (symmetric), F + Fb, band-style Dirichlet elimination
    (matrix.hpp:132-145), solved by Cholesky."""
    import scipy.linalg as sl
    A = np.array(K, dtype=np.float64)
    B = np.asarray(F, dtype=np.float64) + np.asarray(Fb, dtype=np.float64)
    for r, c, v in entries:
        A[r, c] += v
        if r != c:
            A[c, r] += v
    for r, v in zip(np.asarray(dnodes, dtype=np.int64), np.asarray(dvals, dtype=np.float64)):
        B -= A[:, r] * v
        A[r, :] = 0.0
        A[:, r] = 0.0
        A[r, r] = 1.0
    for r, v in zip(np.asarray(dnodes, dtype=np.int64), np.asarray(dvals, dtype=np.float64)):
        B[r] = v
    return sl.cho_solve(sl.cho_factor(A, lower=True), B)
