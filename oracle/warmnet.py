"""CPU oracle of the warm-start network (TEST INFRASTRUCTURE ONLY: imported by
``tests/`` as the checker of the GPU kernels; the product never imports it).

NumPy restatement of the 3-level 3D encoder-decoder of ``plask_amd/csrc/warmnet.hip``
(no counterpart in the reference): forward pass and hand-derived backward pass.
``bf16=True`` rounds to bfloat16 (round to nearest even) at the same points as the
GPU kernels (weights, stored activations, stored activation gradients), so the GPU
can be compared closely; ``bf16=False`` is plain float64, whose backward pass is
pinned by central finite differences in the tests.

Layout: channels-last arrays ``[nz, ny, nx, c]``; parameters one flat vector, per
convolution ``w[27][cout][cin]`` then ``b[cout]``, then ``h[32]`` and ``hb``.
Tap t = (dz+1)*9 + (dy+1)*3 + (dx+1).
"""
import numpy as np

LAYERS = [(8, 32), (32, 64), (64, 128), (128, 64), (64, 32)]


def bf16(a):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def num_params():
    return sum(27 * ci * co + co for ci, co in LAYERS) + 33


def unpack(p):
    out, off = [], 0
    for ci, co in LAYERS:
        w = p[off:off + 27 * co * ci].reshape(27, co, ci)
        off += 27 * co * ci
        b = p[off:off + co]
        off += co
        out.append((w, b))
    h = p[off:off + 32]
    hb = p[off + 32]
    return out, h, hb


def _taps():
    for t in range(27):
        yield t, t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1


def conv3(x, w, b=None):
    """y[v, co] = sum_{tap, ci} x[v + off(tap), ci] w[tap, co, ci] (+ b), zero padding."""
    nz, ny, nx, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((nz, ny, nx, w.shape[1]))
    for t, dz, dy, dx in _taps():
        xs = xp[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]
        y += xs @ w[t].T
    if b is not None:
        y += b
    return y


def conv3_bwd(x, w, gy):
    """Gradients of conv3 with respect to x and w."""
    nz, ny, nx, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    gxp = np.zeros_like(xp)
    gw = np.zeros_like(w)
    g2 = gy.reshape(-1, gy.shape[-1])
    for t, dz, dy, dx in _taps():
        sl = (slice(1 + dz, 1 + dz + nz), slice(1 + dy, 1 + dy + ny), slice(1 + dx, 1 + dx + nx))
        gxp[sl] += gy @ w[t]
        gw[t] = g2.T @ xp[sl].reshape(-1, x.shape[-1])
    return gxp[1:-1, 1:-1, 1:-1], gw


def pool(x):
    nz, ny, nx, c = x.shape
    return x.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).mean(axis=(1, 3, 5))


def pool_bwd(g):
    return up(g) / 8.0


def up(x):
    return x.repeat(2, axis=0).repeat(2, axis=1).repeat(2, axis=2)


def up_bwd(g):
    nz, ny, nx, c = g.shape
    return g.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).sum(axis=(1, 3, 5))


def forward(p, x, rb=False):
    """Returns (y, cache).  rb: round to bf16 where the GPU stores bf16."""
    R = bf16 if rb else (lambda a: a)
    (W1, b1), (W2, b2), (W3, b3), (W4, b4), (W5, b5) = [(R(w), b) for w, b in unpack(p)[0]]
    _, h, hb = unpack(p)
    x = R(x)
    e1 = R(np.maximum(conv3(x, W1, b1), 0))
    p1 = R(pool(e1))
    e2 = R(np.maximum(conv3(p1, W2, b2), 0))
    p2 = R(pool(e2))
    e3 = R(np.maximum(conv3(p2, W3, b3), 0))
    u2 = up(e3)
    d2 = R(np.maximum(conv3(u2, W4, b4) + e2, 0))
    u1 = up(d2)
    d1 = R(np.maximum(conv3(u1, W5, b5) + e1, 0))
    y = d1 @ h + hb
    return y[..., None][..., 0], dict(x=x, e1=e1, p1=p1, e2=e2, p2=p2, e3=e3, u2=u2, d2=d2, u1=u1, d1=d1,
                                      W=[W1, W2, W3, W4, W5], h=h)


def loss_grad(p, x, t, mask=None, rb=False, lam=0.0, wz=None):
    """Loss [sum m e^2 + lam sum_{x,y edges} (e_a - e_b)^2] / sum m + sum_{z edges}
    w (e_a - e_b)^2 (e = y - t; w = wz[node below], or lam / sum m without wz) and its
    gradient."""
    R = bf16 if rb else (lambda a: a)
    y, c = forward(p, x, rb)
    m = np.ones_like(t) if mask is None else mask
    inv = 1.0 / m.sum()
    e = y - t
    loss = float((m * e * e).sum() * inv)
    gy = 2.0 * m * e * inv
    for ax in range(3):
        d = np.diff(e, axis=ax)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, -1)
        hi[ax] = slice(1, None)
        if ax == 0 and wz is not None:
            w = wz[tuple(lo)]
        else:
            w = lam * inv
        loss += float((w * d * d).sum())
        g = np.zeros_like(e)
        g[tuple(lo)] -= w * d
        g[tuple(hi)] += w * d
        gy += 2.0 * g
    W1, W2, W3, W4, W5 = c["W"]
    gh = (gy[..., None] * c["d1"]).reshape(-1, 32).sum(0)
    ghb = gy.sum()
    g5 = R(gy[..., None] * c["h"] * (c["d1"] > 0))
    gu1, gW5 = conv3_bwd(c["u1"], W5, g5)
    gb5 = g5.reshape(-1, 32).sum(0)
    g4 = R(R(up_bwd(R(gu1))) * (c["d2"] > 0))
    gu2, gW4 = conv3_bwd(c["u2"], W4, g4)
    gb4 = g4.reshape(-1, 64).sum(0)
    g3 = R(R(up_bwd(R(gu2))) * (c["e3"] > 0))
    gp2, gW3 = conv3_bwd(c["p2"], W3, g3)
    gb3 = g3.reshape(-1, 128).sum(0)
    g2 = R((g4 + pool_bwd(R(gp2))) * (c["e2"] > 0))
    gp1, gW2 = conv3_bwd(c["p1"], W2, g2)
    gb2 = g2.reshape(-1, 64).sum(0)
    g1 = R((g5 + pool_bwd(R(gp1))) * (c["e1"] > 0))
    _, gW1 = conv3_bwd(c["x"], W1, g1)
    gb1 = g1.reshape(-1, 32).sum(0)
    parts = []
    for gw, gb in ((gW1, gb1), (gW2, gb2), (gW3, gb3), (gW4, gb4), (gW5, gb5)):
        parts += [gw.ravel(), gb.ravel()]
    parts += [gh, np.array([ghb])]
    return loss, np.concatenate(parts), y
