"""CPU statement of the multigrid V-cycle preconditioner of the GPU solve path.

Test infrastructure only (imported by tests/, never by the product).  The
preconditioner has no counterpart in the reference (NSPCG offers Jacobi, IC and line
preconditioners, `iterative_matrix.hpp:436-459`); the reference's own arithmetic is
the operator it preconditions, restated in ``fem_oracle.c`` (`therm3d.cpp:209-240`,
Dirichlet elimination `matrix.hpp:111-145`).  This file states, with SciPy sparse
matrices, what the GPU kernels compute so that one V-cycle can be compared entry by
entry:

* levels: each in-plane axis (3D: axes 0 and 1; 2D: axis 0) keeps nodes 0, 2, 4, ...
  and its last node while it has more than 3 nodes; the layer axis keeps all planes;
* coarse operator: the oracle's eliminated Q1 operator on the coarse grid with
  area-weighted mean conductivities (2D: width-weighted, times the mid radius in
  cylindrical geometry) (``orc_assemble``); a coarse node is Dirichlet when the fine node
  at its position is;
* transfer: P = I_z (x) P_y (x) P_x (linear interpolation along each in-plane axis),
  masked to zero at Dirichlet nodes on both sides; restriction R = P^T;
* smoother: x <- x + omega M^-1 (b - A x), M the tridiagonal part of A along the layer
  axis (exact line solve); one sweep from zero before the coarse correction, one after,
  ``coarse_sweeps`` sweeps from zero on the coarsest level.

Damping: element by element, the column-block part of a Q1 box matrix with diagonal
conductivity is (kx' + ky')/3 I (x) I (x) M_z + kz'/9 I (x) I (x) S_z; in the
sum/difference basis its eigenvalue ratio to the full element matrix is at most 3
(mode delta-sigma-sigma with kx' >> ky', kz'), so lambda_max(M^-1 A) <= 3 and any
omega < 2/3 keeps the smoother convergent and the V-cycle positive definite.

Pinned in tests/test_multigrid_oracle.py by the nested-space identity (for
coefficients that vary only along the layer axis the coarse operator equals the
Galerkin product R A P: coarse Q1 functions are fine Q1 functions), by the
eigenvalue bound on random anisotropic grids, and by the symmetry and positivity of
the V-cycle operator.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from . import ST_DIA, Grid, assemble


def csr(A):
    """SciPy CSR matrix of an oracle DIA-storage matrix (symmetric, upper diagonals stored)."""
    N = A.rank
    rows, cols, vals = [], [], []
    for d, off in enumerate(A.offsets):
        dd = A.data[N * d:N * (d + 1)]
        r = np.arange(N - off)
        rows.append(r)
        cols.append(r + off)
        vals.append(dd[:N - off])
        if off:
            rows.append(r + off)
            cols.append(r)
            vals.append(dd[:N - off])
    M = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    M.eliminate_zeros()
    return M


def coarsen_axis(f, on):
    """Coarse node indices, per-fine-node interval and right weight, coarse coordinates."""
    n = len(f)
    ci = list(range(0, n, 2 if on else 1))
    if ci[-1] != n - 1:
        ci.append(n - 1)
    ci = np.array(ci)
    m = np.zeros(n, np.int64)
    w = np.zeros(n)
    for I in range(len(ci) - 1):
        a, b = ci[I], ci[I + 1]
        for i in range(a, b):
            m[i] = I
            w[i] = (f[i] - f[a]) / (f[b] - f[a])
    m[-1] = len(ci) - 1
    return ci, m, w, f[ci]


def interp_1d(m, w, nc):
    n = len(m)
    P = sp.lil_matrix((n, nc))
    for i in range(n):
        P[i, m[i]] += 1.0 - w[i]
        if w[i] != 0.0:
            P[i, m[i] + 1] += w[i]
    return P.tocsr()


class Level:
    def __init__(self, axes, kh, kv, dmask, cyl=False):
        self.axes = [np.asarray(a, dtype=np.float64) for a in axes]
        self.cyl = bool(cyl)
        self.n = [len(a) for a in self.axes]
        self.N = int(np.prod(self.n))
        self.plane = self.n[0] * self.n[1] if len(self.n) == 3 else self.n[0]
        self.kh = np.asarray(kh, dtype=np.float64)
        self.kv = np.asarray(kv, dtype=np.float64)
        self.dmask = np.asarray(dmask, dtype=bool)
        dn = np.nonzero(self.dmask)[0].astype(np.int64)
        A = assemble(Grid(self.axes, cyl=self.cyl), 0, self.kh, self.kv, np.zeros(self.kh.size), ST_DIA, dn,
                     np.zeros(len(dn)))
        self.A = csr(A)
        coo = self.A.tocoo()
        keep = (coo.row == coo.col) | (np.abs(coo.row - coo.col) == self.plane)
        self.M = sp.csc_matrix((coo.data[keep], (coo.row[keep], coo.col[keep])), shape=(self.N, self.N))
        self.Mlu = spl.splu(self.M)


def coarse_level_2d(L, max_n=3):
    """2D: axis 0 coarsened, axis 1 (the layer axis) kept; conductivity means weighted by
    element width, times the mid radius in cylindrical geometry."""
    if L.n[0] <= max_n:
        return None, None
    cx, mx, wx, ax = coarsen_axis(L.axes[0], True)
    n1 = L.n[1]
    e0f, e0c = L.n[0] - 1, len(cx) - 1
    w = np.diff(L.axes[0])
    if L.cyl:
        w = w * 0.5 * (L.axes[0][1:] + L.axes[0][:-1])
    khf = L.kh.reshape(n1 - 1, e0f)
    kvf = L.kv.reshape(n1 - 1, e0f)
    khc = np.zeros((n1 - 1, e0c))
    kvc = np.zeros((n1 - 1, e0c))
    for I in range(e0c):
        a = w[cx[I]:cx[I + 1]]
        khc[:, I] = (khf[:, cx[I]:cx[I + 1]] * a).sum(axis=1) / a.sum()
        kvc[:, I] = (kvf[:, cx[I]:cx[I + 1]] * a).sum(axis=1) / a.sum()
    cd = L.dmask.reshape(n1, L.n[0])[:, cx].ravel()
    C = Level([ax, L.axes[1]], khc.ravel(), kvc.ravel(), cd, cyl=L.cyl)
    P = sp.kron(sp.identity(n1), interp_1d(mx, wx, len(cx)))
    P = (sp.diags((~L.dmask).astype(float)) @ P @ sp.diags((~C.dmask).astype(float))).tocsr()
    return C, P


def coarse_level(L, max_n=3):
    """Next coarser level of L and the masked interpolation P (fine x coarse); None when
    no in-plane axis has more than ``max_n`` nodes."""
    if len(L.n) == 2:
        return coarse_level_2d(L, max_n)
    onx, ony = L.n[0] > max_n, L.n[1] > max_n
    if not (onx or ony):
        return None, None
    cx, mx, wx, ax = coarsen_axis(L.axes[0], onx)
    cy, my, wy, ay = coarsen_axis(L.axes[1], ony)
    n2 = L.n[2]
    e0f, e1f = L.n[0] - 1, L.n[1] - 1
    e0c, e1c = len(cx) - 1, len(cy) - 1
    hx, hy = np.diff(L.axes[0]), np.diff(L.axes[1])
    khf = L.kh.reshape(n2 - 1, e1f, e0f)
    kvf = L.kv.reshape(n2 - 1, e1f, e0f)
    khc = np.zeros((n2 - 1, e1c, e0c))
    kvc = np.zeros((n2 - 1, e1c, e0c))
    for J in range(e1c):
        for I in range(e0c):
            a = hy[cy[J]:cy[J + 1], None] * hx[None, cx[I]:cx[I + 1]]
            sl = (slice(None), slice(cy[J], cy[J + 1]), slice(cx[I], cx[I + 1]))
            khc[:, J, I] = (khf[sl] * a).sum(axis=(1, 2)) / a.sum()
            kvc[:, J, I] = (kvf[sl] * a).sum(axis=(1, 2)) / a.sum()
    fd = L.dmask.reshape(n2, L.n[1], L.n[0])
    cd = fd[:, cy][:, :, cx].ravel()
    C = Level([ax, ay, L.axes[2]], khc.ravel(), kvc.ravel(), cd)
    P = sp.kron(sp.identity(n2), sp.kron(interp_1d(my, wy, len(cy)), interp_1d(mx, wx, len(cx))))
    P = (sp.diags((~L.dmask).astype(float)) @ P @ sp.diags((~C.dmask).astype(float))).tocsr()
    return C, P


class VCycle:
    """Levels of a 2D or 3D grid (natural node order, axis 0 fastest; the last axis is the
    layer axis) and the V-cycle z = V r."""

    def __init__(self, axes, kh, kv, dnodes, omega=0.65, coarse_sweeps=1, max_levels=0, cyl=False):
        N = int(np.prod([len(a) for a in axes]))
        dmask = np.zeros(N, bool)
        dmask[np.asarray(dnodes, dtype=np.int64)] = True
        self.levels = [Level(axes, kh, kv, dmask, cyl=cyl)]
        self.P = []
        while max_levels == 0 or len(self.levels) < max_levels:
            C, P = coarse_level(self.levels[-1])
            if C is None:
                break
            self.levels.append(C)
            self.P.append(P)
        self.omega = omega
        self.coarse_sweeps = coarse_sweeps

    def _smooth(self, L, x, b):
        return x + self.omega * L.Mlu.solve(b - L.A @ x)

    def apply(self, r, level=0):
        L = self.levels[level]
        if level == len(self.levels) - 1:
            x = self.omega * L.Mlu.solve(r)
            for _ in range(self.coarse_sweeps - 1):
                x = self._smooth(L, x, r)
            return x
        x = self.omega * L.Mlu.solve(r)
        P = self.P[level]
        x = x + P @ self.apply(P.T @ (r - L.A @ x), level + 1)
        return self._smooth(L, x, r)
