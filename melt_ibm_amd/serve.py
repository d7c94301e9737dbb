"""Serving loop: melting-solid simulation requests arrive over time, are
admitted against an HBM budget, keep their state resident on the device and
advance concurrently (continuous batching).

Admission is first come first served: every request that has arrived and
fits in the free budget has its bytes reserved and goes to a builder thread
(a request that does not fit waits; later ones wait behind it).  The builder
sets the simulation up off the scheduler's path (context and mesh, initial
state synthesised on the GPU, initial ray tagging, and the first RK3 step)
and hands it to the scheduler.  The scheduler keeps one step in flight per
resident request: it polls each request's stream without blocking, and as
soon as a request's step has finished (its device error word read) it hands
that request's next step to a pool of host threads, which queue it on the
request's own stream (with remeshing, the default, tschem reads the device's
short-edge flag every substep and edits the mesh on the host when it is set,
so it returns only after the last substep is queued; without it, a step is
the replay of one captured graph).  Requests therefore advance at their own
pace, a small grid does not wait for a large one, and the steps of all
resident requests overlap on the device.  A request that has run its steps,
or hit a terminal condition, is retired and its bytes return to the budget.

Latency per request: the time between consecutive step completions of that
request (inter-step time, what a client streaming the request's steps sees,
including any wait behind other requests' work), beside the device latency of
each step (launch to completion), the queue wait (arrival to admission) and
the set-up time (admission to hand-over).

The request stream (python -m melt_ibm_amd.serve): Poisson arrivals, mean 4
requests/s, seed 7; grids from {64, 128, 256, 384}^3; spheres or spheroids of
aspect 1-3 meshed for a mean edge of about 0.9 h (2k to 60k triangles, as the
reference's deck relates its mesh to its grid); St in [0.5, 4]; 50-500 RK3
steps; 200 GB budget.
"""
from __future__ import annotations

import argparse
import json
import queue
import sys
import threading
import time
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import geometry
from .api import MeltError
from .ensemble import Member, Simulation, terminal_status

SERVE_GRIDS = (64, 128, 256, 384)
FREQ_RANGE = (10, 55)   # icosphere frequencies of the stream's meshes: 2000 to 60500 faces
EDGE_ON_H = 0.9         # mean edge / grid spacing the mesh is sized for


def freq_for(n: int, rad: float, aspect: float) -> int:
    """Icosphere frequency whose mesh has a mean edge of about 0.9 h on an n^3
    grid of the 2 pi cube (the reference's deck, part.in: a level-5 sphere of
    radius 0.2 on a 128^3 unit cube, edges near 0.9 h with E_thresh = 0.7 h),
    within the stream's 2000-60500 faces.  Edges scale as 1/frequency."""
    a = aspect ** (-1.0 / 3.0)
    m = geometry.icosphere(10) if aspect == 1.0 else geometry.spheroid(a, aspect * a, 10)
    d = m.xyz[m.vert_of_edge[:, 0] - 1] - m.xyz[m.vert_of_edge[:, 1] - 1]
    mean10 = float(np.sqrt((d * d).sum(axis=1)).mean()) * rad
    f = int(round(10.0 * mean10 / (EDGE_ON_H * 2.0 * np.pi / n)))
    return min(max(f, FREQ_RANGE[0]), FREQ_RANGE[1])


def request_bytes(m: Member, precision: int) -> int:
    """Device bytes of one simulation (what the library allocates, with a
    margin): 26 fields + 7 solve buffers, the half spectrum, the Lagrangian
    arrays (with the 125 tagging-band distances and the 486 gradient-probe
    coefficients per face) and the immersed-boundary windows, plus the
    transient initial-state synthesis (about 10 field-sized float64 buffers)."""
    esz = 4 if precision else 8
    n = m.n
    n1hp = ((n // 2 + 1) + 7) & ~7
    f = m.freq or 2 ** m.level
    nf = 20 * f * f
    nv = nf // 2 + 2
    win = min(n, int(np.ceil(2.0 * m.rad * max(1.0, m.aspect) * n / (2 * np.pi))) + 10) ** 3
    b = esz * n ** 3 * 33 + 2 * esz * n1hp * n * n + 8 * (nf * (4 * 27 + 40 + 125 + 486) + nv * 24) + win * (4 * 16 + 12)
    return int(b * 1.10) + 10 * 8 * n ** 3 + (64 << 20)


@dataclass
class Request:
    rid: int
    arrival: float
    member: Member
    sim: Simulation | None = None
    admitted: float = -1.0
    ready: float = -1.0
    finished: float = -1.0
    done_steps: int = 0
    step_ms: list = field(default_factory=list)       # inter-step times (completion to completion)
    device_ms: list = field(default_factory=list)     # launch to completion of each step
    bytes: int = 0
    status: str = "ok"
    inflight: bool = False
    future: Future | None = None                      # the host call that queued the current step
    launched: float = 0.0
    last_done: float = 0.0


def request_stream(duration_s: float, rate: float = 4.0, seed: int = 7, steps_scale: float = 1.0, grids=SERVE_GRIDS,
                   remesh: bool = True) -> list[Request]:
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration_s:
            break
        n = int(rng.choice(grids))
        aspect = 1.0 if rng.random() < 0.5 else float(rng.uniform(1.0, 3.0))
        rad = float(rng.uniform(0.3, 0.7))
        m = Member(n=n, rad=rad, st=float(rng.uniform(0.5, 4.0)), re=float(rng.uniform(200.0, 1000.0)),
                   seed=len(out), aspect=aspect, freq=freq_for(n, rad, aspect),
                   steps=max(1, int(round(rng.integers(50, 501) * steps_scale))), remesh=remesh)
        out.append(Request(len(out), t, m))
    return out


class Server:
    def __init__(self, budget_bytes: float = 200e9, precision: int = 1, lib=None, tune: bool = False,
                 device_inputs: bool = True):
        self.budget = int(budget_bytes)
        self.precision = precision
        self.lib = lib
        self.tune = tune
        self.device_inputs = device_inputs
        self.used = 0
        self.active: list[Request] = []
        self.done: list[Request] = []
        self.peak = 0
        self.tuning: dict[int, dict] = {}
        self._t0 = 0.0

    def _clock(self) -> float:
        return time.perf_counter() - self._t0

    # ------------------------------------------------------------ builder
    def _setup(self, r: Request) -> None:
        tuning = None
        if self.tune:
            from .tune import autotune
            if r.member.n not in self.tuning:
                self.tuning[r.member.n] = autotune(r.member.n, self.precision, lib=self.lib)
            tuning = self.tuning[r.member.n]
        r.sim = Simulation.build(r.member, self.precision, self.lib, tuning, graph=True, device_inputs=self.device_inputs)
        # the first step captures the graph (host work) here, off the scheduler
        t = self._clock()
        try:
            r.sim.ctx.tschem()
            r.sim.ctx.sync()
        except MeltError as e:
            r.status = terminal_status(r.sim.ctx, e)
        r.done_steps = 1
        r.last_done = self._clock()
        r.device_ms.append((r.last_done - t) * 1e3)

    def _builder(self, todo: queue.Queue, built: queue.Queue) -> None:
        while True:
            r = todo.get()
            if r is None:
                return
            try:
                self._setup(r)
            except MeltError as e:
                r.status = f"failed {e.code}"
            except Exception as e:  # noqa: BLE001 - any set-up failure ends only this request
                r.status = f"failed {type(e).__name__}"
            built.put(r)

    # ---------------------------------------------------------- scheduler
    def _retire(self, r: Request, keep_results: bool, results: dict, log) -> None:
        r.finished = self._clock()
        if keep_results and r.status == "ok" and r.sim is not None:
            results[r.rid] = {q: r.sim.ctx.download(q) for q in ("vx", "vy", "vz", "temp", "pr")}
            results[r.rid]["volume"] = r.sim.ctx.body().volume
        if r.sim is not None:
            r.sim.close()
            r.sim = None
        self.used -= r.bytes
        self.done.append(r)
        if log:
            print(f"# request {r.rid} ({r.member.n}^3, {r.done_steps}/{r.member.steps} steps, {r.status}) done at "
                  f"{r.finished:.2f} s", file=log, flush=True)

    def run(self, requests: list[Request], keep_results: bool = False, log=None) -> dict:
        pending = sorted(requests, key=lambda r: r.arrival)
        for r in pending:
            r.bytes = request_bytes(r.member, self.precision)
            if r.bytes > self.budget:
                raise ValueError(f"request {r.rid} ({r.member.n}^3) exceeds the HBM budget on its own")
        todo: queue.Queue = queue.Queue()
        built: queue.Queue = queue.Queue()
        worker = threading.Thread(target=self._builder, args=(todo, built), daemon=True)
        # steps are submitted from a pool of host threads: with remeshing,
        # tschem waits inside for each substep's short-edge flag
        pool = ThreadPoolExecutor(max_workers=16)
        self._t0 = time.perf_counter()
        worker.start()
        building = 0
        cell_steps = 0
        results: dict = {}
        try:
            while pending or self.active or building:
                now = self._clock()
                while pending and pending[0].arrival <= now and self.used + pending[0].bytes <= self.budget:
                    r = pending.pop(0)
                    r.admitted = now
                    self.used += r.bytes
                    self.peak = max(self.peak, self.used)
                    todo.put(r)
                    building += 1
                while True:
                    try:
                        r = built.get_nowait()
                    except queue.Empty:
                        break
                    building -= 1
                    r.ready = self._clock()
                    if r.status.startswith("failed"):
                        self._retire(r, keep_results, results, log)
                        continue
                    cell_steps += r.member.cells
                    self.active.append(r)
                progressed = False
                still = []
                for r in self.active:
                    if r.inflight:
                        if not (r.future.done() and r.sim.ctx.poll()):
                            still.append(r)
                            continue
                        t = self._clock()
                        r.inflight = False
                        progressed = True
                        try:
                            r.future.result()
                            r.sim.ctx.sync()
                        except MeltError as e:
                            r.status = terminal_status(r.sim.ctx, e)
                        r.device_ms.append((t - r.launched) * 1e3)
                        r.step_ms.append((t - r.last_done) * 1e3)
                        r.last_done = t
                        r.done_steps += 1
                        cell_steps += r.member.cells
                    if r.status != "ok" or r.done_steps >= r.member.steps:
                        self._retire(r, keep_results, results, log)
                        continue
                    r.launched = self._clock()
                    r.future = pool.submit(r.sim.ctx.tschem)
                    r.inflight = True
                    still.append(r)
                self.active = still
                if not progressed:
                    if not self.active and not building and pending:
                        time.sleep(max(0.0, min(pending[0].arrival - self._clock(), 0.005)))
                    else:
                        time.sleep(0.0002)
        finally:
            todo.put(None)
            worker.join()
            pool.shutdown()
        wall = self._clock()
        inter = np.concatenate([np.array(r.step_ms) for r in self.done if r.step_ms] or [np.zeros(1)])
        dev = np.concatenate([np.array(r.device_ms) for r in self.done if r.device_ms] or [np.zeros(1)])
        wait = np.array([r.admitted - r.arrival for r in self.done]) if self.done else np.zeros(1)
        setup = np.array([r.ready - r.admitted for r in self.done]) if self.done else np.zeros(1)
        status: dict = {}
        for r in self.done:
            status[r.status] = status.get(r.status, 0) + 1
        by_grid = {}
        for n in sorted({r.member.n for r in self.done}):
            a = np.concatenate([np.array(r.step_ms) for r in self.done if r.member.n == n and r.step_ms] or [np.zeros(1)])
            by_grid[str(n)] = {"requests": sum(1 for r in self.done if r.member.n == n),
                               "steps": int(sum(r.done_steps for r in self.done if r.member.n == n)),
                               "step_ms_p50": float(np.percentile(a, 50)), "step_ms_p99": float(np.percentile(a, 99))}
        rep = {"requests": len(self.done), "status": status, "wall_s": wall, "mcell_steps_per_s": cell_steps / wall / 1e6,
               "steps": int(sum(r.done_steps for r in self.done)),
               "step_ms_p50": float(np.percentile(inter, 50)), "step_ms_p99": float(np.percentile(inter, 99)),
               "device_step_ms_p50": float(np.percentile(dev, 50)), "device_step_ms_p99": float(np.percentile(dev, 99)),
               "queue_wait_s_p50": float(np.percentile(wait, 50)), "queue_wait_s_p99": float(np.percentile(wait, 99)),
               "setup_s_p50": float(np.percentile(setup, 50)), "setup_s_p99": float(np.percentile(setup, 99)),
               "peak_budget_gb": self.peak / 1e9, "budget_gb": self.budget / 1e9,
               "by_grid": by_grid,
               "grids": sorted({r.member.n for r in self.done}),
               "faces_range": [min(20 * (r.member.freq or 2 ** r.member.level) ** 2 for r in self.done),
                               max(20 * (r.member.freq or 2 ** r.member.level) ** 2 for r in self.done)] if self.done else []}
        if keep_results:
            rep["results"] = results
        return rep


def main() -> None:
    ap = argparse.ArgumentParser(description="serving loop over a Poisson request stream")
    ap.add_argument("--duration", type=float, default=4.0, help="arrival window (s)")
    ap.add_argument("--rate", type=float, default=4.0)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--steps-scale", type=float, default=1.0, help="scale the 50-500 step counts (shorter runs)")
    ap.add_argument("--precision", type=int, default=1)
    args = ap.parse_args()
    reqs = request_stream(args.duration, args.rate, args.seed, args.steps_scale)
    srv = Server(args.budget_gb * 1e9, args.precision)
    rep = srv.run(reqs, log=sys.stderr)
    print(json.dumps(rep))


if __name__ == "__main__":
    main()
