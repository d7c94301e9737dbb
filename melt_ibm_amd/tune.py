"""Autotuner for the launch shapes of the fused explicit/x-solve pass (z-march
planes per block, or line-tile shapes) and of the y/z line solves.

For a grid size and precision it times every candidate shape on a fluid-only
context (seeded solenoidal state, the kernels are the ones a melting run
launches) with the library's per-kernel HIP-event timers, and keeps the
fastest per pass.  Shapes never change results (every element is owned by
exactly one thread whatever the shape), only speed.  Results are cached per
(grid, precision) for the process, and in a JSON file when MELT_TUNE_CACHE
names one.
"""
from __future__ import annotations

import json
import os
from pathlib import Path

from . import inputs
from .api import MeltConfig, MeltContext, MeltError

_CACHE: dict[tuple[int, int], dict] = {}

XF_W = (1, 2, 4, 8)
XF_NT = (128, 256)
XF_Z = (4, 8, 16, 32)
SOLVE_W = (4, 8, 16, 32)
SOLVE_NT = (128, 256, 512, 1024)


def _load_file() -> dict:
    p = os.environ.get("MELT_TUNE_CACHE")
    if p and Path(p).exists():
        try:
            return json.loads(Path(p).read_text())
        except (OSError, ValueError):
            return {}
    return {}


def _save_file(key: str, val: dict) -> None:
    p = os.environ.get("MELT_TUNE_CACHE")
    if not p:
        return
    d = _load_file()
    d[key] = val
    Path(p).write_text(json.dumps(d, indent=1, sort_keys=True))


def _time(ctx: MeltContext, slots: tuple[str, ...], reps: int = 2) -> float:
    ctx.tschem()  # warm-up (and first use of this shape)
    ctx.sync()
    ctx.timing(True)
    ctx.timing_reset()
    for _ in range(reps):
        ctx.tschem()
    ctx.sync()
    tm = ctx.timing_read()
    ctx.timing(False)
    return sum(tm[s][0] / max(tm[s][1], 1) for s in slots)


def autotune(n: int, precision: int = 0, lib=None) -> dict:
    key = (n, precision)
    if key in _CACHE:
        return dict(_CACHE[key])
    fkey = f"{n}:{precision}"
    stored = _load_file().get(fkey)
    if stored is not None:
        _CACHE[key] = stored
        return dict(stored)
    cfg = MeltConfig(n1m=n, n2m=n, n3m=n, ren=500.0, imlsfor=0, imlsstr=0, imelt=0, forcing=0, precision=precision)
    ctx = MeltContext(cfg, lib=lib) if lib is not None else MeltContext(cfg)
    try:
        u, v, w = inputs.solenoidal_velocity(cfg, seed=0, rms=0.1)
        for name, a in (("vx", u), ("vy", v), ("vz", w)):
            ctx.upload(name, a)
        ctx.set_dt(inputs.cfl_dt(cfg, u, v, w, cfl=0.5))
        best = {}
        zero = {k: 0 for k in ("xf_w", "xf_nt", "solve_w", "solve_nt", "xf_z", "yz_jb")}
        # the explicit pass: z-march with D planes per block, or line tiles of W lines x nt threads
        xf = [{"xf_z": d, "xf_w": 0, "xf_nt": 0} for d in XF_Z if n % d == 0] + \
             [{"xf_z": 0, "xf_w": w_, "xf_nt": t} for w_ in XF_W if n % w_ == 0 for t in XF_NT]
        solve = [{"solve_w": w_, "solve_nt": t} for w_ in SOLVE_W if n % w_ == 0 for t in SOLVE_NT]
        # the y/z solves: one fused pass with 16 or 8 rows per block, or two passes
        yz = [{"yz_jb": 0}, {"yz_jb": 16}, {"yz_jb": 1}]
        for slots, grid in ((("explicit_xsolve",), xf), (("solve_y", "solve_z_update"), yz),
                            (("solve_y", "solve_z_update"), solve)):
            scores = []
            for cand in grid:
                try:
                    ctx.set_tuning(**cand)
                    scores.append((_time(ctx, slots), cand))
                except MeltError:
                    continue
            if scores:
                best.update({k: v for k, v in min(scores, key=lambda s: s[0])[1].items() if v})
            ctx.set_tuning(**zero)
    finally:
        ctx.close()
    _CACHE[key] = best
    _save_file(fkey, best)
    return dict(best)
