"""Learned warm start for the 3D thermal solve (no counterpart in the reference,
which starts PCG from the previous solution or ``inittemp``, therm3d.cpp:79 /
iterative_matrix.hpp:205-209).

``WarmNet`` binds the ``pf_net_*`` entry points of ``include/plask_warmnet.h``: a
3-level 3D convolutional encoder-decoder (32/64/128 channels, 3x3x3 kernels, bf16
on the matrix cores) trained on the GPU by hand-written forward and backward
kernels.  ``features``/``target_scale`` turn a layered-stack problem into the
network's 8 input channels and its output normalisation; ``generate`` solves
seeded stacks with the GPU solver to make training data; ``WarmStart.guess``
turns a trained network into a start vector for ``FemContext.solve`` on any grid
of the same stack (the prediction is made on the network grid and interpolated).
A guess whose residual is not below that of the uniform start is rejected by the
caller in favour of the uniform (``inittemp``) start.
"""
import ctypes as C

import numpy as np

from . import fem, workloads

NIN = 8
LAYERS = [(8, 32), (32, 64), (64, 128), (128, 64), (64, 32)]


def _lib():
    L = fem.load_library()
    if getattr(L, "_net_declared", False):
        return L
    P, I, i64 = C.c_void_p, C.c_int, C.c_int64
    L.pf_net_create.argtypes = [C.POINTER(P), I, I, I, I, I]
    L.pf_net_destroy.argtypes = [P]
    L.pf_net_destroy.restype = None
    L.pf_net_last_error.argtypes = [P]
    L.pf_net_last_error.restype = C.c_char_p
    L.pf_net_num_params.argtypes = [P]
    L.pf_net_num_params.restype = i64
    L.pf_net_set_params.argtypes = [P, P, i64]
    L.pf_net_get_params.argtypes = [P, P, i64]
    L.pf_net_set_mask.argtypes = [P, P]
    L.pf_net_sample_set.argtypes = [P, I, P, P]
    L.pf_net_set_loss.argtypes = [P, C.c_float]
    L.pf_net_sample_weights.argtypes = [P, I, P]
    L.pf_net_predict.argtypes = [P, P, P]
    L.pf_net_predict_sample.argtypes = [P, I, P]
    L.pf_net_output_device.argtypes = [P]
    L.pf_net_output_device.restype = P
    L.pf_net_gradient.argtypes = [P, I, P, C.POINTER(C.c_double)]
    L.pf_net_train_batch.argtypes = [P, P, I, C.c_float, C.POINTER(C.c_double)]
    L._net_declared = True
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def num_params():
    return sum(27 * ci * co + co for ci, co in LAYERS) + 33


def init_params(seed=0):
    """He-normal convolution weights, zero biases, small head."""
    rng = np.random.default_rng(seed)
    parts = []
    for ci, co in LAYERS:
        parts.append(rng.normal(0.0, np.sqrt(2.0 / (27 * ci)), 27 * co * ci))
        parts.append(np.zeros(co))
    parts.append(rng.normal(0.0, np.sqrt(1.0 / 32), 32))
    parts.append(np.zeros(1))
    return np.concatenate(parts).astype(np.float32)


class WarmNet:
    """The network on one grid of (nx, ny, nz) nodes with room for max_samples
    training samples on the device."""

    def __init__(self, shape, max_samples=0, device=0):
        self.L = _lib()
        self.shape = tuple(int(s) for s in shape)
        self.nv = int(np.prod(self.shape))
        h = C.c_void_p()
        rc = self.L.pf_net_create(C.byref(h), int(device), *self.shape, int(max_samples))
        if rc != 0:
            raise fem.FemError(rc, "pf_net_create failed for grid %s (nx, ny, nz multiples of 4, nx*ny*nz/64 a "
                                   "multiple of 32, GPU %d)" % (self.shape, device))
        self.h = h
        self.max_samples = int(max_samples)
        self.n = int(self.L.pf_net_num_params(h))

    def close(self):
        if getattr(self, "h", None):
            self.L.pf_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc, what):
        if rc != 0:
            raise fem.FemError(rc, "%s: %s" % (what, self.L.pf_net_last_error(self.h).decode(errors="replace")))

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.n
        self._ck(self.L.pf_net_set_params(self.h, _ptr(p), self.n), "pf_net_set_params")

    def params(self):
        p = np.zeros(self.n, dtype=np.float32)
        self._ck(self.L.pf_net_get_params(self.h, _ptr(p), self.n), "pf_net_get_params")
        return p

    def set_mask(self, m):
        m = _f32(m)
        assert m.size == self.nv
        self._ck(self.L.pf_net_set_mask(self.h, _ptr(m)), "pf_net_set_mask")

    def set_loss(self, lam):
        """Weight of the neighbour-difference (gradient) term of the loss."""
        self._ck(self.L.pf_net_set_loss(self.h, float(lam)), "pf_net_set_loss")

    def set_sample(self, idx, feat, target):
        feat, target = _f32(feat), _f32(target)
        assert feat.size == self.nv * NIN and target.size == self.nv
        self._ck(self.L.pf_net_sample_set(self.h, int(idx), _ptr(feat), _ptr(target)), "pf_net_sample_set")

    def set_sample_weights(self, idx, wz):
        wz = _f32(wz)
        assert wz.size == self.nv
        self._ck(self.L.pf_net_sample_weights(self.h, int(idx), _ptr(wz)), "pf_net_sample_weights")

    def predict(self, feat):
        feat = _f32(feat)
        assert feat.size == self.nv * NIN
        out = np.zeros(self.nv, dtype=np.float32)
        self._ck(self.L.pf_net_predict(self.h, _ptr(feat), _ptr(out)), "pf_net_predict")
        return out

    def predict_device(self, feat):
        """Forward pass leaving the prediction on the device; returns its device address
        (valid until the next forward pass)."""
        feat = _f32(feat)
        assert feat.size == self.nv * NIN
        self._ck(self.L.pf_net_predict(self.h, _ptr(feat), None), "pf_net_predict")
        return int(self.L.pf_net_output_device(self.h))

    def predict_sample(self, idx):
        out = np.zeros(self.nv, dtype=np.float32)
        self._ck(self.L.pf_net_predict_sample(self.h, int(idx), _ptr(out)), "pf_net_predict_sample")
        return out

    def gradient(self, idx):
        g = np.zeros(self.n, dtype=np.float32)
        loss = C.c_double(0.0)
        self._ck(self.L.pf_net_gradient(self.h, int(idx), _ptr(g), C.byref(loss)), "pf_net_gradient")
        return loss.value, g

    def train_batch(self, idx, lr):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        loss = C.c_double(0.0)
        self._ck(self.L.pf_net_train_batch(self.h, _ptr(idx), len(idx), float(lr), C.byref(loss)),
                 "pf_net_train_batch")
        return loss.value


# ---------------------------------------------------------------------------
# problem -> network inputs / target
# ---------------------------------------------------------------------------
def _node_mean(ev, shape):
    """Mean of an element field over the (up to 8) elements touching each node."""
    nx, ny, nz = shape
    e = ev.reshape(nz - 1, ny - 1, nx - 1)
    acc = np.zeros((nz, ny, nx))
    cnt = np.zeros((nz, ny, nx))
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                acc[dz:dz + nz - 1, dy:dy + ny - 1, dx:dx + nx - 1] += e
                cnt[dz:dz + nz - 1, dy:dy + ny - 1, dx:dx + nx - 1] += 1
    return acc / cnt


def features(p):
    """(nodes, 8) float32 inputs in node order: log k_xy, log k_z (node means of
    the element values), heat source / 1e16, Dirichlet mask, z / height, the
    normalised vertical thermal resistance from the bottom, x and y relative to
    the centre."""
    x, y, z = p["axes"]
    shape = (len(x), len(y), len(z))
    nx, ny, nz = shape
    f = np.zeros((nz, ny, nx, NIN))
    f[..., 0] = _node_mean(np.log(p["k_h"]), shape) / 5.0
    f[..., 1] = _node_mean(np.log(p["k_v"]), shape) / 5.0
    f[..., 2] = _node_mean(p["q"] / 1e16, shape)
    m = np.zeros(nx * ny * nz)
    m[p["dir_nodes"]] = 1.0
    f[..., 3] = m.reshape(nz, ny, nx)
    f[..., 4] = (z / z[-1])[:, None, None]
    # vertical thermal resistance from the heat sink, sum dz / k_z along each
    # element column, normalised by the column total (node: mean of its columns)
    dz = np.diff(z)
    r = np.cumsum(dz[:, None, None] / p["k_v"].reshape(nz - 1, ny - 1, nx - 1), axis=0)
    r = np.concatenate([np.zeros((1, ny - 1, nx - 1)), r], axis=0) / r[-1][None]
    rn = np.zeros((nz, ny, nx))
    cnt = np.zeros((1, ny, nx))
    for dy in (0, 1):
        for dx in (0, 1):
            rn[:, dy:dy + ny - 1, dx:dx + nx - 1] += r
            cnt[:, dy:dy + ny - 1, dx:dx + nx - 1] += 1
    f[..., 5] = rn / cnt
    f[..., 6] = (x / x[-1] - 0.5)[None, None, :]
    f[..., 7] = (y / y[-1] - 0.5)[None, :, None]
    return f.reshape(-1, NIN).astype(np.float32)


def target_scale(p):
    """Temperature-rise scale of a stack (K): the rise at the active layer of the
    laterally averaged 1D problem, (total power / area) * sum_{layers below} t / k_z.
    Lengths in um, the source in W/m^3."""
    x, y, z = p["axes"]
    vol = np.outer(np.outer(np.diff(z), np.diff(y)), np.diff(x)).ravel()
    power = float((p["q"] * vol).sum()) * 1e-18            # W
    area = (x[-1] - x[0]) * (y[-1] - y[0]) * 1e-12         # m^2
    a = p["active"]
    r = float(np.sum(p["thick"][:a] / p["kz_layers"][:a])) * 1e-6   # m^2 K / W
    return max(power / area * r, 1e-6)


def z_edge_conductance(p):
    """Vertical conductance k_z dx dy / dz of the z edge above each node (mean of its up
    to 4 element columns; 0 on the top plane), node order."""
    x, y, z = p["axes"]
    nx, ny, nz = len(x), len(y), len(z)
    kz = p["k_v"].reshape(nz - 1, ny - 1, nx - 1)
    area = np.outer(np.diff(y), np.diff(x))[None]
    g = kz * area / np.diff(z)[:, None, None]
    acc = np.zeros((nz - 1, ny, nx))
    cnt = np.zeros((1, ny, nx))
    for dy in (0, 1):
        for dx in (0, 1):
            acc[:, dy:dy + ny - 1, dx:dx + nx - 1] += g
            cnt[:, dy:dy + ny - 1, dx:dx + nx - 1] += 1
    w = np.zeros((nz, ny, nx))
    w[:-1] = acc / cnt
    return w.ravel()


def z_edge_weights(p, target, mu=1.0, conductance=None):
    """Per-node weights of the vertical neighbour differences of the loss: the z-edge
    conductances scaled so that the target's own weighted energy is mu."""
    x, y, z = p["axes"]
    shape = (len(z), len(y), len(x))
    w = (z_edge_conductance(p) if conductance is None else conductance).reshape(shape)
    t = np.asarray(target, dtype=np.float64).reshape(shape)
    energy = float((w[:-1] * np.diff(t, axis=0) ** 2).sum())
    return (mu * w / max(energy, 1e-300)).ravel().astype(np.float32)


def dirichlet_mask(p):
    m = np.ones(int(np.prod([len(a) for a in p["axes"]])), dtype=np.float32)
    m[p["dir_nodes"]] = 0.0
    return m


def generate(seeds, shape, rtol=1e-8, ctx=None):
    """Solve the seeded stacks (thermal_3d distributions) on the GPU; yields
    (problem, features, normalised target, scale, iterations)."""
    ctx = ctx or fem.FemContext()
    for s in seeds:
        p = workloads.thermal_3d(seed=int(s), n=shape)
        ctx.set_grid(p["axes"])
        ctx.set_coeffs(p["k_h"], p["k_v"])
        ctx.set_source(p["q"])
        ctx.set_dirichlet(p["dir_nodes"], p["dir_vals"])
        st = ctx.solve_uniform(300.0, rtol=rtol)
        T = ctx.solution()
        S = target_scale(p)
        yield p, features(p), ((T - 300.0) / S).astype(np.float32), S, st.iters


class WarmStart:
    """A trained network applied to a problem on any grid of the same stack."""

    def __init__(self, net, net_shape):
        self.net = net
        self.net_shape = tuple(net_shape)

    def guess_coarse(self, p_net):
        """Prediction on the network grid (T in K)."""
        return 300.0 + target_scale(p_net) * self.net.predict(features(p_net)).astype(np.float64)

    def start_on_device(self, ctx, p_net, feat=None):
        """Predict on the network grid and prolong onto ctx's grid on the device
        (the next ctx.solve_warm starts from it)."""
        y = self.net.predict(features(p_net) if feat is None else feat)
        ctx.prolong(y, p_net["axes"], offset=300.0, scale=target_scale(p_net))

    def guess(self, p_net, p):
        """Start vector for problem p from the prediction on p_net (the same stack on
        the network grid), trilinear in physical coordinates; Dirichlet values kept."""
        return self.prolong(self.guess_coarse(p_net), p_net, p)

    @staticmethod
    def prolong(Tc, p_net, p):
        """Trilinear interpolation of node values on the grid of p_net to the nodes of
        p (separable: one axis at a time), Dirichlet values of p imposed."""
        out = np.asarray(Tc, dtype=np.float64).reshape([len(a) for a in p_net["axes"]][::-1])
        for ax in range(3):
            src, dst = p_net["axes"][ax], p["axes"][ax]
            i = np.clip(np.searchsorted(src, dst, side="right") - 1, 0, len(src) - 2)
            w = np.clip((dst - src[i]) / (src[i + 1] - src[i]), 0.0, 1.0)
            axis = 2 - ax                      # array axes are (z, y, x)
            lo = np.take(out, i, axis=axis)
            hi = np.take(out, i + 1, axis=axis)
            shape = [1, 1, 1]
            shape[axis] = len(dst)
            out = lo + (hi - lo) * w.reshape(shape)
        T = out.ravel()
        T[p["dir_nodes"]] = p["dir_vals"]
        return T
