"""Pinning the oracle's HIT forcing (hit.f90) on the CPU.

 - the forced modes are the integer wave vectors with 0 < |k| < kf
   (hit.f90:36, KF/KMIN = 2.5 gives 5 wave numbers per direction);
 - the synthesised fields are the trigonometric polynomial Re sum Fhat e^{ik.x}
   sampled at each component's staggered points: an FFT of the three fields,
   with each component's sampling offset undone, recovers coefficients that are
   divergence-free (k . G(k) = 0), vanish for |k| >= kf, and equal the
   solenoidal projection of the reported bhat (hit.f90:51-54);
 - bhat is an Ornstein-Uhlenbeck process: stationary E|b|^2 = (2 eps dt / TL^2)
   E|xi|^2 / (1 - a^2) with a = 1 - dt / TL and E|xi|^2 = 2 (Box-Muller pairs),
   lag-one correlation a (hit.f90:46-47);
 - with EPSSTAR = 0 one mode decays by a per step and gives the closed-form
   cosine field at the staggered points;
 - the ABC field (hit.f90:234-265) has each component independent of its own
   coordinate (discretely divergence-free), zero mean and amplitude 2C.
"""
import numpy as np
import pytest

import oracle
from melt_ibm_amd import inputs
from tests.helpers import make_cfg


@pytest.fixture(scope="module")
def lib():
    return oracle.load()


def _ctx(n=16, n3=None, **kw):
    cfg = make_cfg(n, n3, **kw)
    return cfg, oracle.context(cfg)


def test_forced_modes(lib):
    cfg, o = _ctx(16)
    o.hit_init(0.5, 0.05, 2.5, seed=3)
    nm, nforced = o.hit_nmodes()
    k = np.arange(-2, 3)
    K = np.stack(np.meshgrid(k, k, k, indexing="ij"), -1).reshape(-1, 3)
    m2 = np.sum(K * K, axis=1)
    assert nm == 5 and nforced == int(np.sum((m2 > 0) & (m2 < 6.25)))
    # HITForcing.in's KF/KMIN = 2.3: the same 5 wave numbers, |k|^2 <= 5 forced
    o.hit_init(0.5, 0.05, 2.3, seed=3)
    assert o.hit_nmodes() == (5, int(np.sum((m2 > 0) & (m2 < 2.3 ** 2))))
    # KF/KMIN = 1: wave numbers -1..1 but no mode with 0 < |k| < kf
    o.hit_init(0.5, 0.05, 1.0, seed=3)
    assert o.hit_nmodes() == (3, 0)


def _coefficients(cfg, f, offsets):
    """G(k) = F(k) + conj(F(-k)) of a field Re sum F e^{ik.x} sampled at
    x = offset + h * index (fields indexed [k, j, i])."""
    n3, n2, n1 = cfg.shape
    C = np.fft.fftn(f) / f.size                       # C[k3, k2, k1]
    k1 = np.fft.fftfreq(n1, 1.0 / n1) * 2 * np.pi / cfg.xlen
    k2 = np.fft.fftfreq(n2, 1.0 / n2) * 2 * np.pi / cfg.ylen
    k3 = np.fft.fftfreq(n3, 1.0 / n3) * 2 * np.pi / cfg.zlen
    ph = np.exp(-1j * (k1[None, None, :] * offsets[0] + k2[None, :, None] * offsets[1] + k3[:, None, None] * offsets[2]))
    return 2.0 * C * ph, (k1, k2, k3)


def test_synthesis_is_solenoidal_and_band_limited(lib):
    cfg, o = _ctx(16, 12, xlen=6.0, ylen=5.0, zlen=4.0)
    o.set_dt(0.02)
    o.hit_init(0.5, 0.05, 2.5, seed=11)
    for _ in range(3):
        o.hit_force()
    h = (cfg.xlen / cfg.n1m, cfg.ylen / cfg.n2m, cfg.zlen / cfg.n3m)
    offs = [(0.0, h[1] / 2, h[2] / 2), (h[0] / 2, 0.0, h[2] / 2), (h[0] / 2, h[1] / 2, 0.0)]   # u, v, w points
    G = []
    for c, name in enumerate(("forcx", "forcy", "forcz")):
        g, (k1, k2, k3) = _coefficients(cfg, o.download(name), offs[c])
        G.append(g)
    K1, K2, K3 = np.meshgrid(k3, k2, k1, indexing="ij")[::-1]
    div = K1 * G[0] + K2 * G[1] + K3 * G[2]
    scale = max(np.max(np.abs(g)) for g in G)
    assert scale > 0.0
    assert np.max(np.abs(div)) < 1e-12 * scale * 10
    kmag = np.sqrt(K1 ** 2 + K2 ** 2 + K3 ** 2)
    kf = 2.5 * 2 * np.pi / cfg.xlen
    assert max(np.max(np.abs(g[kmag >= kf])) for g in G) < 1e-13 * scale
    # the coefficients are the projection of the reported bhat
    bhat, steps = o.hit_state()                        # [n, m, l, dir]
    assert steps == 3
    wn = np.arange(-2, 3)
    for n_ in range(5):
        for m_ in range(5):
            for l_ in range(5):
                kv = np.array([wn[l_] * 2 * np.pi / cfg.xlen, wn[m_] * 2 * np.pi / cfg.ylen, wn[n_] * 2 * np.pi / cfg.zlen])
                km2 = kv @ kv
                if km2 == 0 or km2 >= kf * kf:
                    assert np.all(bhat[n_, m_, l_] == 0)
                    continue
                b = bhat[n_, m_, l_]
                F = b - kv * (kv @ b) / km2
                bm = bhat[4 - n_, 4 - m_, 4 - l_]
                Fm = bm - (-kv) * ((-kv) @ bm) / km2
                idx = (wn[n_] % cfg.n3m, wn[m_] % cfg.n2m, wn[l_] % cfg.n1m)
                for c in range(3):
                    assert abs(G[c][idx] - (F[c] + np.conj(Fm[c]))) < 1e-12 * scale, (l_, m_, n_, c)


def test_ou_statistics(lib):
    cfg, o = _ctx(8)
    tl, eps, dt = 0.5, 0.05, 0.05
    o.set_dt(dt)
    o.hit_init(tl, eps, 2.5, seed=5)
    a = 1.0 - dt / tl
    burn, steps = 60, 2000
    prev, num, den, var = None, 0.0, 0.0, []
    mask = None
    for s in range(burn + steps):
        o.hit_force()
        b, _ = o.hit_state()
        if mask is None:
            mask = np.abs(b) > 0        # forced modes and directions
        if s >= burn:
            v = b[mask]
            var.append(np.mean(np.abs(v) ** 2))
            if prev is not None:
                num += np.sum((v * np.conj(prev)).real)
                den += np.sum(np.abs(prev) ** 2)
            prev = v
    expect = (2.0 * eps * dt / tl ** 2) * 2.0 / (1.0 - a * a)
    assert abs(np.mean(var) / expect - 1.0) < 0.05
    assert abs(num / den - a) < 0.01


def test_single_mode_decay_closed_form(lib):
    cfg, o = _ctx(16, 12)
    tl, dt = 0.5, 0.02
    o.set_dt(dt)
    o.hit_init(tl, 0.0, 2.5, seed=0)
    bhat, _ = o.hit_state()
    b = np.zeros_like(bhat)
    bx = np.array([0.3 - 0.2j, 0.7 + 0.1j, -0.4 + 0.5j])    # k = (1, 0, 0) (x is l, the last index)
    bz = np.array([0.2 + 0.6j, -0.5 - 0.3j, 0.9 + 0.0j])    # k = (0, 0, 1)
    b[2, 2, 3] = bx
    b[3, 2, 2] = bz
    o.set_hit_state(b, 0)
    o.hit_force()
    a = 1.0 - dt / tl
    (xm, ym, zm), (xc, yc, zc) = inputs.grid_coords(cfg)
    z = np.zeros(cfg.shape)
    # Fhat = a b with the component along k removed
    fx = (a * bz[0] * np.exp(1j * zm))[:, None, None].real + z
    fy = (a * bx[1] * np.exp(1j * xm))[None, None, :].real + (a * bz[1] * np.exp(1j * zm))[:, None, None].real + z
    fz = (a * bx[2] * np.exp(1j * xm))[None, None, :].real + z
    for name, e in (("forcx", fx), ("forcy", fy), ("forcz", fz)):
        assert np.max(np.abs(o.download(name) - e)) < 1e-15 * 8, name
    b2, steps = o.hit_state()
    assert steps == 1 and np.max(np.abs(b2 - a * b)) < 1e-16


def test_abc_forcing(lib):
    cfg, o = _ctx(16, 12, xlen=6.0, ylen=5.0, zlen=4.0)
    c_hit = 0.1
    o.abc_force(c_hit)
    C = c_hit * (2 * np.pi / cfg.xlen) ** 2
    fx, fy, fz = (o.download(q) for q in ("forcx", "forcy", "forcz"))
    # each component is independent of its own coordinate: the staggered
    # divergence vanishes exactly
    assert np.all(fx == fx[:, :, :1]) and np.all(fy == fy[:, :1, :]) and np.all(fz == fz[:1, :, :])
    for f in (fx, fy, fz):
        assert abs(f.mean()) < 1e-15 and np.max(np.abs(f)) <= 2 * C + 1e-15
    (xm, ym, zm), _ = inputs.grid_coords(cfg)
    assert abs(fx[3, 5, 0] - C * (np.sin(2 * np.pi * zm[3] / cfg.zlen) + np.cos(2 * np.pi * ym[5] / cfg.ylen))) < 1e-15
