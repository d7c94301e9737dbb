"""Shared case builders for the parity tests: the same seeded state is loaded
into the HIP library and into the oracle."""
from __future__ import annotations

import numpy as np

from melt_ibm_amd import geometry, inputs
from melt_ibm_amd.api import MeltConfig, MeltContext


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    den = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / den)


def make_cfg(n=32, n3=None, **kw) -> MeltConfig:
    base = dict(n1m=n, n2m=n, n3m=n3 or n, ren=500.0, prandtl=1.0, betagz=0.5, latheat=1.0, dens_ratio=1.2,
                imlsfor=1, imlsstr=1, imelt=1, forcing=1)
    base.update(kw)
    return MeltConfig(**base)


def random_state(cfg: MeltConfig, seed: int = 0) -> dict:
    rng = np.random.default_rng(seed)
    u, v, w = inputs.solenoidal_velocity(cfg, seed=seed, rms=0.3)
    sh = cfg.shape
    st = {"vx": u, "vy": v, "vz": w,
          "temp": rng.uniform(0.0, 1.0, sh), "pr": rng.normal(0.0, 0.05, sh),
          "ru1": rng.normal(0, 0.1, sh), "ru2": rng.normal(0, 0.1, sh), "ru3": rng.normal(0, 0.1, sh),
          "rut": rng.normal(0, 0.1, sh), "VOFp": np.clip(rng.uniform(-0.3, 1.3, sh), 0, 1)}
    fx, fy, fz = inputs.low_mode_forcing(cfg, seed=seed + 1)
    st.update(forcx=fx, forcy=fy, forcz=fz)
    return st


def load_state(ctx: MeltContext, st: dict) -> None:
    for k, a in st.items():
        ctx.upload(k, a)


def body_state(ctx: MeltContext, seed: int = 3):
    """Give the body a non-trivial rigid motion (hdnlte's solid-velocity terms)."""
    rng = np.random.default_rng(seed)
    b = ctx.body()
    for i in range(3):
        b.pos_cm[i] = b.pos_cm[i] + rng.uniform(-0.3, 0.3)
        b.vel_cm[i] = rng.uniform(-0.1, 0.1)
        b.omega_c[i] = rng.uniform(-0.2, 0.2)
    ctx.set_body(b)


def with_body(ctx: MeltContext, level: int, rad: float) -> geometry.Mesh:
    m = geometry.sphere_level(level)
    ctx.set_geometry(m.xyz, m.vert_of_edge, m.edge_of_face, rad)
    return m


def melt_case(cfg: MeltConfig, level: int, rad: float, seed: int = 0):
    """Physical initial state of a melting sphere (inqpr temperature, solenoidal u)."""
    u, v, w = inputs.solenoidal_velocity(cfg, seed=seed, rms=0.1)
    fx, fy, fz = inputs.low_mode_forcing(cfg, seed=seed + 7)
    st = {"vx": u, "vy": v, "vz": w, "forcx": fx, "forcy": fy, "forcz": fz}
    dt = inputs.cfl_dt(cfg, u, v, w, cfl=0.5)
    return st, dt
