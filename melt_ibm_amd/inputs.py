"""Seeded synthetic inputs for the substep (no datasets exist for this path).

- `solenoidal_velocity`: staggered-grid velocity that is divergence-free to
  round-off under the reference's discrete divergence (divg.f90:17-20): the
  discrete curl of a random-phase vector potential with |k| <= kmax, scaled to a
  target rms.
- `low_mode_forcing`: HIT-like forcing fields (forcx/forcy/forcz) from random
  solenoidal Fourier modes |k| <= kf; the stochastic forcing synthesis itself
  (hit.f90) runs once per step outside the substep and is not on this path.
- `inqpr_temperature`: the reference's initial temperature (inqpr.f90:17-40).
- `cfl_dt`: the time step get_dt (cfl.f90:33-77) picks for a target CFL.
"""
from __future__ import annotations

import numpy as np

from .api import MeltConfig


def _random_potential(shape, kmax: float, rng: np.random.Generator) -> np.ndarray:
    n3, n2, n1 = shape
    k3 = np.fft.fftfreq(n3, 1.0 / n3)[:, None, None]
    k2 = np.fft.fftfreq(n2, 1.0 / n2)[None, :, None]
    k1 = np.fft.rfftfreq(n1, 1.0 / n1)[None, None, :]
    kk = np.sqrt(k1 ** 2 + k2 ** 2 + k3 ** 2)
    mask = (kk <= kmax) & (kk > 0)
    spec = np.zeros((n3, n2, n1 // 2 + 1), dtype=np.complex128)
    amp = np.where(mask, 1.0 / np.maximum(kk, 1.0) ** 2, 0.0)
    phase = rng.uniform(0.0, 2.0 * np.pi, size=spec.shape)
    spec[...] = amp * np.exp(1j * phase)
    return np.fft.irfftn(spec, s=shape, axes=(0, 1, 2))


def _curl(ax, ay, az, cfg: MeltConfig):
    dx1, dx2, dx3 = cfg.n1m / cfg.xlen, cfg.n2m / cfg.ylen, cfg.n3m / cfg.zlen

    def fwd(a, axis, d):  # a(+1) - a along axis (periodic), arrays indexed [k, j, i]
        return (np.roll(a, -1, axis=axis) - a) * d

    u = fwd(az, 1, dx2) - fwd(ay, 0, dx3)
    v = fwd(ax, 0, dx3) - fwd(az, 2, dx1)
    w = fwd(ay, 2, dx1) - fwd(ax, 1, dx2)
    return u, v, w


def solenoidal_velocity(cfg: MeltConfig, seed: int = 0, kmax: float = 4.0, rms: float = 0.1):
    rng = np.random.default_rng(seed)
    a = [_random_potential(cfg.shape, kmax, rng) for _ in range(3)]
    u, v, w = _curl(*a, cfg)
    s = rms / np.sqrt((u ** 2 + v ** 2 + w ** 2).mean() / 3.0)
    return u * s, v * s, w * s


def low_mode_forcing(cfg: MeltConfig, seed: int = 1, kf: float = 2.5, amplitude: float = 0.1):
    rng = np.random.default_rng(seed)
    a = [_random_potential(cfg.shape, kf, rng) for _ in range(3)]
    fx, fy, fz = _curl(*a, cfg)
    s = amplitude / np.sqrt((fx ** 2 + fy ** 2 + fz ** 2).mean() / 3.0)
    return fx * s, fy * s, fz * s


def grid_coords(cfg: MeltConfig):
    """xm, ym, zm (cell centres) and xc, yc, zc (faces), 0-based arrays, cordin.f90."""
    def c(n, L):
        i = np.arange(1, n + 2)
        xc = L * ((i - 1).astype(np.float64) / float(n))
        return xc[:-1], (xc[:-1] + xc[1:]) * 0.5

    xc, xm = c(cfg.n1m, cfg.xlen)
    yc, ym = c(cfg.n2m, cfg.ylen)
    zc, zm = c(cfg.n3m, cfg.zlen)
    return (xm, ym, zm), (xc, yc, zc)


def inqpr_temperature(cfg: MeltConfig, pos_cm, rad_p: float) -> np.ndarray:
    (xm, ym, zm), _ = grid_coords(cfg)
    dx1 = cfg.n1m / cfg.xlen
    r = np.sqrt((xm[None, None, :] - pos_cm[0]) ** 2 + (ym[None, :, None] - pos_cm[1]) ** 2 + (zm[:, None, None] - pos_cm[2]) ** 2)
    return cfg.tliq - (cfg.tliq - cfg.tsol) * 0.5 * (1.0 - np.tanh((r - rad_p) * dx1 / 2.0))


def cfl_dt(cfg: MeltConfig, u, v, w, cfl: float = 0.5, dt0: float = 1.0) -> float:
    dx1, dx2, dx3 = cfg.n1m / cfg.xlen, cfg.n2m / cfg.ylen, cfg.n3m / cfg.zlen
    c = np.max(np.abs(u) * dt0 * dx1 + np.abs(v) * dt0 * dx2 + np.abs(w) * dt0 * dx3)
    c = max(c, 6.0 / cfg.ren * dt0 * dx1 ** 2)
    return dt0 * cfl / c


def solenoidal_velocity_device(cfg: MeltConfig, seed: int = 0, kmax: float = 4.0, rms: float = 0.1):
    """The construction of `solenoidal_velocity` on the GPU (torch, float64):
    a random-phase vector potential with |k| <= kmax and amplitude 1/|k|^2,
    its forward-difference curl (divergence-free under divg.f90:17-20),
    scaled to the target rms.  Phases come from torch's generator seeded with
    `seed`, so the field differs from the host construction's but is
    reproducible.  Returns device tensors [k, j, i] and the dt of CFL 0.5
    (cfl.f90:33-77 with dt0 = 1)."""
    import torch
    dev = torch.device("cuda")
    n3, n2, n1 = cfg.shape
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))
    k3 = torch.fft.fftfreq(n3, 1.0 / n3, dtype=torch.float64, device=dev)[:, None, None]
    k2 = torch.fft.fftfreq(n2, 1.0 / n2, dtype=torch.float64, device=dev)[None, :, None]
    k1 = torch.fft.rfftfreq(n1, 1.0 / n1, dtype=torch.float64, device=dev)[None, None, :]
    kk = torch.sqrt(k1 ** 2 + k2 ** 2 + k3 ** 2)
    amp = torch.where((kk <= kmax) & (kk > 0), 1.0 / torch.clamp(kk, min=1.0) ** 2, torch.zeros_like(kk))
    del kk
    pot = []
    for _ in range(3):
        phase = torch.rand((n3, n2, n1 // 2 + 1), generator=gen, dtype=torch.float64, device=dev) * (2.0 * np.pi)
        pot.append(torch.fft.irfftn(torch.polar(amp.expand_as(phase).contiguous(), phase), s=(n3, n2, n1), dim=(0, 1, 2)))
        del phase
    d1, d2, d3 = cfg.n1m / cfg.xlen, cfg.n2m / cfg.ylen, cfg.n3m / cfg.zlen

    def fwd(a, axis, d):
        return (torch.roll(a, -1, dims=axis) - a) * d

    ax, ay, az = pot
    u = fwd(az, 1, d2) - fwd(ay, 0, d3)
    v = fwd(ax, 0, d3) - fwd(az, 2, d1)
    w = fwd(ay, 2, d1) - fwd(ax, 1, d2)
    del pot, ax, ay, az
    s = rms / torch.sqrt((u * u + v * v + w * w).mean() / 3.0)
    u, v, w = u * s, v * s, w * s
    c = torch.max(torch.abs(u) * d1 + torch.abs(v) * d2 + torch.abs(w) * d3).item()
    c = max(c, 6.0 / cfg.ren * d1 ** 2)
    torch.cuda.current_stream().synchronize()
    return u.contiguous(), v.contiguous(), w.contiguous(), 0.5 / c
