"""Batched form: many independent melting simulations of uneven grid sizes
advanced together on one GPU.

Every simulation owns a context (its fields, mesh and stream).  Its RK3 step is
captured once into a HIP graph (melt_graph_enable) and replayed; one ensemble
step enqueues every simulation's graph on its own stream, so the device runs
them concurrently, and then waits for all.  Launch shapes come from the
autotuner (melt_ibm_amd.tune), per grid size and precision.

The ensemble workload (48 simulations): grids drawn uniformly from
{48, 64, 96, 128}^3, a sphere of radius 0.3-0.7 meshed as a level-3 (grids up
to 64) or level-4 icosahedron, St uniform in [0.5, 4], Re uniform in
[200, 1000], seeded random-phase initial velocity (|k| <= 4, rms 0.1) from the
member's own seed, T = 1 in the fluid and 0 in the solid (inqpr), no HIT
forcing, dt from CFL 0.5, FP32 bulk with FP64 Lagrangian updates, remeshing
on (E_thresh = 0.7 h as in the reference's deck, or 0.95 of the shortest
initial edge where that is shorter).
"""
from __future__ import annotations

import sys
import time
from dataclasses import dataclass, field

import numpy as np

from . import geometry, inputs
from .api import MeltConfig, MeltContext, MeltError

GRIDS = (48, 64, 96, 128)

# The reference's fatal conditions that end one simulation (it stops or calls
# MPI_Abort): a body volume below V_thresh (particle.f90:218-224), remeshing
# that breaks the topology or leaves a volume residual above 1e-14
# (remesh_coarsen.f90:77-253, particle.f90:205-210), and, here, a surface that
# outgrew the device window sized at set-up.  In a batch they end only the
# simulation that raised them.
TERMINAL_CODES = (4, 6, 9)   # MELT_ERR_SMALL_VOLUME, MELT_ERR_REMESH, MELT_ERR_WINDOW


def terminal_status(ctx: MeltContext, e: MeltError) -> str:
    """"melted" for a positive volume below V_thresh; "diverged" for a tangled
    surface (a negative or non-finite volume, or one that left its window);
    "remesh failed" for the remeshing aborts; "error N" for anything else."""
    if e.code == 6:
        return "remesh failed"
    if e.code == 9:
        return "diverged"
    if e.code != 4:
        return f"error {e.code}"
    try:
        vol = ctx.body().volume
    except MeltError:
        vol = float("nan")
    return "melted" if np.isfinite(vol) and vol > 0.0 else "diverged"


@dataclass
class Member:
    n: int
    rad: float
    st: float
    re: float
    seed: int
    level: int = 3
    aspect: float = 1.0     # spheroid c/a (1 = sphere), volume-equivalent radius rad
    freq: int = 0           # icosphere frequency (faces = 20 freq^2); 0: 2**level
    steps: int = 0          # RK3 steps to run (serving)
    remesh: bool = True     # IREMESH = 1, E_thresh = min(0.7 h, 0.95 x shortest initial edge)

    @property
    def cells(self) -> int:
        return self.n ** 3

    def config(self, precision: int) -> MeltConfig:
        # latHeat = cp (Tliq - Tmelt) / St with cp = 1, Tliq = 1, Tmelt = 0.  With
        # remeshing, E_thresh is the reference deck's 0.7 h (PERCT_ETHRESH,
        # part.in), lowered to 0.95 of the shortest initial edge on a finer mesh:
        # the deck's edges start above E_thresh, and edges that the melting
        # shrinks are collapsed from there on.
        perc = 0.7
        if self.remesh:
            m = self.mesh()
            d = m.xyz[m.vert_of_edge[:, 0] - 1] - m.xyz[m.vert_of_edge[:, 1] - 1]
            perc = min(0.7, 0.95 * float(np.sqrt((d * d).sum(axis=1)).min()) * self.rad / (2.0 * np.pi / self.n))
        return MeltConfig(n1m=self.n, n2m=self.n, n3m=self.n, ren=self.re, prandtl=1.0, betagz=0.0, tmelt=0.0, tliq=1.0, tsol=0.0,
                          cpliquid=1.0, latheat=1.0 / self.st, dens_ratio=1.0, imlsfor=1, imlsstr=1, imelt=1,
                          iremesh=1 if self.remesh else 0, forcing=0, precision=precision, perc_ethresh=perc)

    def mesh(self) -> geometry.Mesh:
        f = self.freq or 2 ** self.level
        if self.aspect == 1.0:
            return geometry.icosphere(f)
        a = self.aspect ** (-1.0 / 3.0)
        return geometry.spheroid(a, self.aspect * a, f)


def ensemble_members(count: int = 48, seed: int = 0, remesh: bool = True) -> list[Member]:
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.choice(GRIDS))
        out.append(Member(n=n, rad=float(rng.uniform(0.3, 0.7)), st=float(rng.uniform(0.5, 4.0)), re=float(rng.uniform(200.0, 1000.0)),
                          seed=i, level=3 if n <= 64 else 4, remesh=remesh))
    return out


@dataclass
class Simulation:
    member: Member
    ctx: MeltContext
    dt: float
    step_ms: list = field(default_factory=list)

    @classmethod
    def build(cls, m: Member, precision: int = 1, lib=None, tuning: dict | None = None, graph: bool = True,
              device_inputs: bool = False) -> "Simulation":
        """Set the member up.  A mesh finer than the reference's ray tagging
        allows (more than n1m faces in one pencil, pencil_rayTag.f90:126-130)
        is replaced by the next coarser icosahedron level (or a frequency 3/4
        as fine).  device_inputs: synthesise the initial velocity on the GPU
        (inputs.solenoidal_velocity_device) instead of on the host."""
        while True:
            try:
                return cls._build(m, precision, lib, tuning, graph, device_inputs)
            except MeltError as e:
                f0 = m.freq or 2 ** m.level
                if e.code != 5 or f0 <= 2:
                    raise
                if m.freq:
                    m.freq = max(2, (3 * m.freq) // 4)
                else:
                    m.level -= 1
                f1 = m.freq or 2 ** m.level
                print(f"# member seed {m.seed} ({m.n}^3, R {m.rad:.2f}): pencil overflow, {20 * f0 * f0} -> {20 * f1 * f1} faces",
                      file=sys.stderr, flush=True)

    @classmethod
    def _build(cls, m: Member, precision: int, lib, tuning: dict | None, graph: bool, device_inputs: bool = False) -> "Simulation":
        cfg = m.config(precision)
        ctx = MeltContext(cfg, lib=lib) if lib is not None else MeltContext(cfg)
        try:
            mesh = m.mesh()
            ctx.set_geometry(mesh.xyz, mesh.vert_of_edge, mesh.edge_of_face, m.rad)
            ctx.inqpr()
            if device_inputs:
                import torch
                u, v, w, dt = inputs.solenoidal_velocity_device(cfg, seed=m.seed, kmax=4.0, rms=0.1)
                for name, a in (("vx", u), ("vy", v), ("vz", w)):
                    ctx.upload_device(name, a.data_ptr())
                del u, v, w
                torch.cuda.empty_cache()
            else:
                u, v, w = inputs.solenoidal_velocity(cfg, seed=m.seed, kmax=4.0, rms=0.1)
                for name, a in (("vx", u), ("vy", v), ("vz", w)):
                    ctx.upload(name, a)
                dt = inputs.cfl_dt(cfg, u, v, w, cfl=0.5)
            ctx.set_dt(dt)
            # many contexts share the GPU: each keeps its work on one stream
            # (their streams then spread over the device's hardware queues);
            # the ray-tagging pencils hold up to 4 n1m faces (the reference's
            # BUFF_SIZE, pencil_rayTag.f90:84, is n1m), so that members keep
            # their level-3/4 meshes on the coarse grids
            ctx.set_tuning(streams=1, pbuf=4 * m.n, **(tuning or {}))
            ctx.tagCells()
            ctx.sync()
            if graph:
                ctx.graph(True)
        except Exception:
            ctx.close()
            raise
        return cls(m, ctx, dt)

    def close(self) -> None:
        self.ctx.close()


class Ensemble:
    """With remeshing the host reads each member's short-edge flag every
    substep inside tschem (and edits the mesh when it is set), so the members
    are stepped from a pool of host threads: while one waits for its flag the
    others keep submitting work.  Without remeshing each step is one captured
    graph per member."""

    def __init__(self, members: list[Member], precision: int = 1, lib=None, tune: bool = True, graph: bool = True):
        self.tuning = {}
        if tune:
            from .tune import autotune
            for n in sorted({m.n for m in members}):
                self.tuning[n] = autotune(n, precision, lib=lib)
        self.sims = [Simulation.build(m, precision, lib, self.tuning.get(m.n), graph) for m in members]
        self.active = list(range(len(self.sims)))
        self.ended: dict[int, tuple[int, str]] = {}   # simulation index -> (ensemble step, terminal_status)
        self.ended_detail: dict[int, str] = {}        # simulation index -> the library's message
        self.steps_done = 0
        self.step_ms: list[float] = []
        self.step_cells: list[int] = []
        from concurrent.futures import ThreadPoolExecutor
        self.pool = ThreadPoolExecutor(max_workers=min(16, max(1, len(self.sims))))

    @property
    def cells(self) -> int:
        return sum(s.member.cells for s in self.sims)

    def step(self) -> float:
        """One RK3 step of every simulation still running.  The reference's
        terminal conditions (TERMINAL_CODES) end only their simulation, which
        is recorded in `ended` with terminal_status; the others continue.  Any
        other device error ends the ensemble.  Remeshing errors are raised by
        tschem itself (the host edits the mesh inside it), the others by the
        sync that reads the device's error word."""
        run = list(self.active)
        t0 = time.perf_counter()
        futs = [self.pool.submit(self.sims[i].ctx.tschem) for i in run]
        for i, f in zip(run, futs):
            try:
                f.result()
                self.sims[i].ctx.sync()
            except MeltError as e:
                if e.code not in TERMINAL_CODES:
                    raise
                self.ended[i] = (self.steps_done, terminal_status(self.sims[i].ctx, e))
                self.ended_detail[i] = str(e)
                self.active.remove(i)
        ms = (time.perf_counter() - t0) * 1e3
        self.steps_done += 1
        self.step_ms.append(ms)
        self.step_cells.append(sum(self.sims[i].member.cells for i in run))
        return ms

    def run(self, steps: int, warmup: int = 1) -> dict:
        for _ in range(warmup):
            self.step()
        self.step_ms, self.step_cells = [], []
        for _ in range(steps):
            self.step()
        a = np.array(self.step_ms)
        return {"sims": len(self.sims), "cells": self.cells, "steps": steps, "warmup": warmup,
                "step_ms_p50": float(np.percentile(a, 50)), "step_ms_p99": float(np.percentile(a, 99)),
                "step_ms_mean": float(a.mean()), "mcell_steps_per_s": sum(self.step_cells) / (a.sum() / 1e3) / 1e6,
                "ended": {str(i): {"step": k, "status": st} for i, (k, st) in sorted(self.ended.items())},
                "ended_detail": {str(i): d for i, d in sorted(self.ended_detail.items())},
                "faces": sorted({20 * (s.member.freq or 2 ** s.member.level) ** 2 for s in self.sims}),
                "tuning": {str(k): v for k, v in self.tuning.items()}}

    def close(self) -> None:
        self.pool.shutdown()
        for s in self.sims:
            s.close()
        self.sims = []
