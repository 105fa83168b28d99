"""CPU checks of the warm-start network oracle (oracle/warmnet.py): the backward
pass against central finite differences, the convolution against a direct loop,
the bf16 rounding against known values, and the input/scale helpers."""
import numpy as np
import pytest

from oracle import warmnet as W


def test_bf16_rounding_known_values():
    v = np.array([1.0, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, -2.5, 3.0e38, 1e-40])
    r = W.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                     # tie -> even
    assert r[2] == 1 + 2.0 ** -6           # tie -> even (upwards)
    assert r[3] == -2.5
    assert np.isfinite(r[4]) and abs(r[4] / 3.0e38 - 1) < 2 ** -8
    assert abs(r[5]) < 1e-39


def test_conv3_matches_direct_loop():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((4, 3, 5, 2))
    w = rng.standard_normal((27, 3, 2))
    b = rng.standard_normal(3)
    y = W.conv3(x, w, b)
    nz, ny, nx, _ = x.shape
    ref = np.zeros((nz, ny, nx, 3))
    for z in range(nz):
        for yy in range(ny):
            for xx in range(nx):
                s = b.copy()
                for t in range(27):
                    dz, dy, dx = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
                    zz, y2, x2 = z + dz, yy + dy, xx + dx
                    if 0 <= zz < nz and 0 <= y2 < ny and 0 <= x2 < nx:
                        s += w[t] @ x[zz, y2, x2]
                ref[z, yy, xx] = s
    assert np.abs(y - ref).max() < 1e-12


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = np.concatenate([np.r_[rng.normal(0, np.sqrt(2 / (27 * ci)), 27 * ci * co), rng.normal(0, 0.1, co)]
                        for ci, co in W.LAYERS] + [rng.normal(0, 0.2, 32), [0.1]])
    x = rng.standard_normal((8, 8, 8, 8))
    t = rng.standard_normal((8, 8, 8))
    mask = (rng.random((8, 8, 8)) > 0.2).astype(float)
    lam = 0.7
    wz = rng.random((8, 8, 8))
    loss, g, _ = W.loss_grad(p, x, t, mask, lam=lam, wz=wz)
    idx = np.r_[rng.choice(p.size - 33, 60, replace=False), np.arange(p.size - 33, p.size)]
    eps = 1e-6
    for i in idx:
        pp = p.copy()
        pp[i] += eps
        lp = W.loss_grad(pp, x, t, mask, lam=lam, wz=wz)[0]
        pp[i] -= 2 * eps
        lm = W.loss_grad(pp, x, t, mask, lam=lam, wz=wz)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_param_layout_matches_header():
    from plask_amd import warmstart
    assert W.num_params() == warmstart.num_params() == 560225
    p = warmstart.init_params(3)
    convs, h, hb = W.unpack(p.astype(np.float64))
    assert [w.shape for w, _ in convs] == [(27, co, ci) for ci, co in W.LAYERS]
    assert h.shape == (32,)


def test_features_and_scale():
    from plask_amd import warmstart, workloads
    p = workloads.thermal_3d(seed=4000, n=(16, 16, 64))
    f = warmstart.features(p)
    assert f.shape == (16 * 16 * 64, 8) and np.all(np.isfinite(f))
    assert np.all(f[:256, 3] == 1.0) and np.all(f[256:, 3] == 0.0)     # Dirichlet bottom plane
    assert 0.0 < f[:, 2].max() <= 1.0
    S = warmstart.target_scale(p)
    # 1D check: a laterally uniform source over the whole area gives exactly q t_a sum t/k
    p2 = dict(p)
    p2["q"] = np.where(p["q"] > 0, 1e16, 0.0)
    a = p["active"]
    lay = p["layer_of_row"]
    q_uniform = np.repeat(np.where(lay == a, 1e16, 0.0), 15 * 15)
    p2["q"] = q_uniform
    S2 = warmstart.target_scale(p2)
    expect = 1e16 * p["thick"][a] * 1e-6 * np.sum(p["thick"][:a] / p["kz_layers"][:a]) * 1e-6
    assert abs(S2 / expect - 1) < 1e-9
    assert 0 < S < S2
