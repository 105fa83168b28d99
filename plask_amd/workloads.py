"""Seeded synthetic layered-stack problems (the benchmark and test workloads).

Random numbers: SplitMix64 (Steele, Lea & Flood, "Fast splittable pseudorandom
number generators", OOPSLA 2014), u = (next >> 11) * 2^-53.  Uniform a + (b-a) u,
log-uniform exp(ln a + (ln b - ln a) u).

Each generator returns a plain dict of numpy arrays in natural order (axis 0
fastest): ``axes``, ``k_h``/``k_v`` per element (W/m/K or S/m), ``q`` per element
(W/m^3), ``dir_nodes``/``dir_vals``, and the layer table.
"""
import math

import numpy as np

_M64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & _M64

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & _M64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self, a=0.0, b=1.0):
        return a + (b - a) * ((self.next() >> 11) * (1.0 / 9007199254740992.0))

    def loguniform(self, a, b):
        return math.exp(math.log(a) + (math.log(b) - math.log(a)) * self.uniform())

    def randint(self, a, b):
        """uniform integer in [a, b]"""
        return a + int(self.uniform() * (b - a + 1))


def split_elements(thick, total, minimum=2):
    """Elements per layer proportional to thickness, >= minimum each, summing to total."""
    thick = np.asarray(thick, dtype=np.float64)
    n = np.maximum(minimum, np.round(total * thick / thick.sum()).astype(np.int64))
    while n.sum() > total:
        i = int(np.argmax(np.where(n > minimum, thick / n, -1)))
        n[i] -= 1
    while n.sum() < total:
        i = int(np.argmax(thick / n))
        n[i] += 1
    return n


def layered_axis(thick, nelem, minimum=2):
    """Vertical node axis: layers uniformly subdivided, element edges on every interface."""
    n = split_elements(thick, nelem, minimum)
    z = [0.0]
    for t, m in zip(thick, n):
        z0 = z[-1]
        z.extend(list(z0 + t * (np.arange(1, m + 1) / m)))
    return np.array(z), n


def element_layers(n_per_layer):
    return np.repeat(np.arange(len(n_per_layer)), n_per_layer)


def thermal_2d_cyl(seed=1, nr=200, nz=400, nlayers=12, active=6, rmax=50.0, peak=1e16, waist=4.0):
    """2D cylindrical steady-state thermal stack (Gaussian source in the 0.1 um active layer)."""
    rng = SplitMix64(seed)
    t = [rng.uniform(0.05, 2.0) for _ in range(nlayers)]
    k = [rng.loguniform(1.0, 150.0) for _ in range(nlayers)]
    t[active] = 0.1
    r = np.linspace(0.0, rmax, nr)
    z, npl = layered_axis(t, nz - 1)
    lay = element_layers(npl)                                # per element row
    kel = np.array(k)[lay]
    rm = 0.5 * (r[1:] + r[:-1])
    q_row = peak * np.exp(-2.0 * rm ** 2 / waist ** 2)
    E0 = nr - 1
    k_h = np.repeat(kel, E0)                                  # row-major: element (i, j) -> j*E0 + i
    q = np.zeros((nz - 1, E0))
    q[lay == active, :] = q_row
    dn = np.arange(nr, dtype=np.int64)                       # z = 0 row
    return dict(axes=[r, z], cyl=True, k_h=k_h, k_v=k_h.copy(), q=q.ravel(), dir_nodes=dn,
                dir_vals=np.full(nr, 300.0), thick=np.array(t), k_layers=np.array(k), layer_of_row=lay,
                active=active)


def thermal_3d(seed=2, n=(128, 128, 256), nlayers=40, active=20, width=100.0, peak=1e16, radius=4.0):
    """3D steady-state thermal stack: anisotropic layers, cylindrical source in the active layer."""
    rng = SplitMix64(seed)
    t = [rng.uniform(0.05, 2.0) for _ in range(nlayers)]
    kxy = [rng.loguniform(1.0, 150.0) for _ in range(nlayers)]
    kz = [rng.loguniform(1.0, 150.0) for _ in range(nlayers)]
    t[active] = 0.1
    nx, ny, nz = n
    x = np.linspace(0.0, width, nx)
    y = np.linspace(0.0, width, ny)
    z, npl = layered_axis(t, nz - 1, minimum=min(2, max(1, (nz - 1) // nlayers)))
    lay = element_layers(npl)
    E0, E1 = nx - 1, ny - 1
    k_h = np.repeat(np.array(kxy)[lay], E0 * E1)
    k_v = np.repeat(np.array(kz)[lay], E0 * E1)
    xm = 0.5 * (x[1:] + x[:-1])
    ym = 0.5 * (y[1:] + y[:-1])
    disk = ((xm[None, :] - width / 2) ** 2 + (ym[:, None] - width / 2) ** 2) <= radius ** 2   # [j, i]
    q = np.zeros((nz - 1, E1, E0))
    q[lay == active] = np.where(disk, peak, 0.0)[None]
    dn = np.arange(nx * ny, dtype=np.int64)
    return dict(axes=[x, y, z], cyl=False, k_h=k_h, k_v=k_v, q=q.ravel(), dir_nodes=dn,
                dir_vals=np.full(nx * ny, 300.0), thick=np.array(t), kxy_layers=np.array(kxy),
                kz_layers=np.array(kz), layer_of_row=lay, active=active)


def shockley_cyl_batch(n_inst=256, seed0=3000, nlayers=12, junction=6, rmax=50.0, contact_r=10.0,
                       nr_range=(50, 400), nz_range=(100, 800)):
    """Uneven 2D cylindrical Shockley instances (voltage / device sweep).

    Per instance i (SplitMix64 seed seed0 + i): node counts n_r ~ U_int[50,400],
    n_z ~ U_int[100,800]; 12 layers t ~ U[0.05,2] um with the junction layer (role
    active) 0.02 um; sigma ~ LU[1e-2,1e5] S/m per layer; js ~ LU[1e-2,1e2] A/m^2;
    beta ~ U[10,25] 1/V; V ~ U[1.4,2.2] V on top nodes with r <= 10 um, 0 V at z = 0.
    Junction elements start at (0, 5) S/m (electr2d.cpp:24).
    """
    out = []
    for i in range(n_inst):
        rng = SplitMix64(seed0 + i)
        nr = rng.randint(*nr_range)
        nz = rng.randint(*nz_range)
        t = [rng.uniform(0.05, 2.0) for _ in range(nlayers)]
        sig = [rng.loguniform(1e-2, 1e5) for _ in range(nlayers)]
        js = rng.loguniform(1e-2, 1e2)
        beta = rng.uniform(10.0, 25.0)
        V = rng.uniform(1.4, 2.2)
        t[junction] = 0.02
        r = np.linspace(0.0, rmax, nr)
        z, npl = layered_axis(t, nz - 1, minimum=1)
        lay = element_layers(npl)
        E0 = nr - 1
        sv_row = np.array(sig)[lay]
        sh = np.repeat(sv_row, E0)
        sv = sh.copy()
        jrows = np.nonzero(lay == junction)[0]
        jlo, jhi = int(jrows.min()), int(jrows.max()) + 1
        for k in range(jlo, jhi):
            sh[k * E0:(k + 1) * E0] = 0.0
            sv[k * E0:(k + 1) * E0] = 5.0
        top = (nz - 1) * nr + np.nonzero(r <= contact_r + 1e-12)[0]
        bot = np.arange(nr)
        dn = np.r_[bot, top].astype(np.int64)
        dv = np.r_[np.zeros(len(bot)), np.full(len(top), V)]
        out.append(dict(axes=[r, z], sh=sh, sv=sv, dir_nodes=dn, dir_vals=dv, jlo=jlo, jhi=jhi, beta=beta, js=js,
                        V=V, layers=np.array(t), sigma=np.array(sig)))
    return out
