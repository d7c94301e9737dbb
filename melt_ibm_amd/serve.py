"""Serving loop: melting-solid simulation requests arrive over time, are
admitted against an HBM budget, keep their state resident on the device and
advance together (continuous batching).

Each scheduler tick admits, in arrival order, every request that has arrived
and fits in the free budget (a request that does not fit waits; later ones
wait behind it, so admission is first-come first-served), then advances every
resident simulation by one RK3 step: each step is a replay of that
simulation's captured graph on its own stream, so the steps of all resident
requests overlap on the device.  A request that has run its steps is retired
and its memory returns to the budget.  Per-request step latency is the time
from the tick's launch to that request's completion as seen by the host.

The request stream (python -m melt_ibm_amd.serve): Poisson arrivals, mean 4
requests/s, seed 7; grids from {64, 128, 256, 384}^3; spheres or spheroids of
aspect 1-3 whose triangle count grows with the grid (2k to 60k); St in
[0.5, 4]; 50-500 RK3 steps; 200 GB budget.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from dataclasses import dataclass, field

import numpy as np

from .api import MeltError
from .ensemble import Member, Simulation

SERVE_GRIDS = (64, 128, 256, 384)
FREQ_FOR_GRID = {64: 10, 128: 20, 256: 35, 384: 55}   # 2000, 8000, 24500, 60500 faces


def request_bytes(m: Member, precision: int) -> int:
    """Device bytes of one simulation (what the library allocates, with a
    margin): 26 fields + 7 solve buffers, the half spectrum, the Lagrangian
    arrays (with the 125 tagging-band distances and the 486 gradient-probe
    coefficients per face) and the immersed-boundary windows."""
    esz = 4 if precision else 8
    n = m.n
    n1hp = ((n // 2 + 1) + 7) & ~7
    f = m.freq or 2 ** m.level
    nf = 20 * f * f
    nv = nf // 2 + 2
    win = min(n, int(np.ceil(2.0 * m.rad * max(1.0, m.aspect) * n / (2 * np.pi))) + 10) ** 3
    b = esz * n ** 3 * 33 + 2 * esz * n1hp * n * n + 8 * (nf * (4 * 27 + 40 + 125 + 486) + nv * 24) + win * (4 * 16 + 12)
    return int(b * 1.10) + (64 << 20)


@dataclass
class Request:
    rid: int
    arrival: float
    member: Member
    sim: Simulation | None = None
    admitted: float = -1.0
    finished: float = -1.0
    done_steps: int = 0
    step_ms: list = field(default_factory=list)
    bytes: int = 0
    status: str = "ok"


def request_stream(duration_s: float, rate: float = 4.0, seed: int = 7, steps_scale: float = 1.0, grids=SERVE_GRIDS) -> list[Request]:
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration_s:
            break
        n = int(rng.choice(grids))
        aspect = 1.0 if rng.random() < 0.5 else float(rng.uniform(1.0, 3.0))
        m = Member(n=n, rad=float(rng.uniform(0.3, 0.7)), st=float(rng.uniform(0.5, 4.0)), re=float(rng.uniform(200.0, 1000.0)),
                   seed=len(out), aspect=aspect, freq=FREQ_FOR_GRID.get(n, 10),
                   steps=max(1, int(round(rng.integers(50, 501) * steps_scale))))
        out.append(Request(len(out), t, m))
    return out


class Server:
    def __init__(self, budget_bytes: float = 200e9, precision: int = 1, lib=None, tune: bool = False):
        self.budget = int(budget_bytes)
        self.precision = precision
        self.lib = lib
        self.tune = tune
        self.used = 0
        self.active: list[Request] = []
        self.done: list[Request] = []
        self.peak = 0
        self.tuning: dict[int, dict] = {}

    def _admit(self, r: Request, now: float) -> None:
        tuning = None
        if self.tune:
            from .tune import autotune
            if r.member.n not in self.tuning:
                self.tuning[r.member.n] = autotune(r.member.n, self.precision, lib=self.lib)
            tuning = self.tuning[r.member.n]
        r.sim = Simulation.build(r.member, self.precision, self.lib, tuning, graph=True)
        r.admitted = now
        self.used += r.bytes
        self.peak = max(self.peak, self.used)
        self.active.append(r)

    def run(self, requests: list[Request], keep_results: bool = False, log=None) -> dict:
        pending = sorted(requests, key=lambda r: r.arrival)
        for r in pending:
            r.bytes = request_bytes(r.member, self.precision)
            if r.bytes > self.budget:
                raise ValueError(f"request {r.rid} ({r.member.n}^3) exceeds the HBM budget on its own")
        t0 = time.perf_counter()
        cell_steps = 0
        results = {}
        while pending or self.active:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now and self.used + pending[0].bytes <= self.budget:
                self._admit(pending.pop(0), now)
            if not self.active:
                time.sleep(max(0.0, min(pending[0].arrival - now, 0.05)))
                continue
            ts = time.perf_counter()
            for r in self.active:
                r.sim.ctx.tschem()
            for r in self.active:
                try:
                    r.sim.ctx.sync()
                except MeltError as e:
                    # the reference's terminal conditions end the request, not the server:
                    # code 4 is a volume below V_thresh (particle.f90:218-224), which a
                    # tangled surface (melting front faster than a cell per step, no
                    # remeshing) also trips, with a negative or non-finite volume
                    r.status = f"error {e.code}"
                    if e.code == 4:
                        try:
                            vol = r.sim.ctx.body().volume
                        except MeltError:
                            vol = float("nan")
                        r.status = "melted" if np.isfinite(vol) and vol > 0.0 else "diverged"
                r.step_ms.append((time.perf_counter() - ts) * 1e3)
                r.done_steps += 1
                cell_steps += r.member.cells
            still = []
            for r in self.active:
                if r.done_steps >= r.member.steps or r.status != "ok":
                    r.finished = time.perf_counter() - t0
                    if keep_results and r.status == "ok":
                        results[r.rid] = {q: r.sim.ctx.download(q) for q in ("vx", "vy", "vz", "temp", "pr")}
                        results[r.rid]["volume"] = r.sim.ctx.body().volume
                    r.sim.close()
                    r.sim = None
                    self.used -= r.bytes
                    self.done.append(r)
                    if log:
                        print(f"# request {r.rid} ({r.member.n}^3, {r.done_steps}/{r.member.steps} steps, {r.status}) done at "
                              f"{r.finished:.2f} s", file=log, flush=True)
                else:
                    still.append(r)
            self.active = still
        wall = time.perf_counter() - t0
        lat = np.concatenate([np.array(r.step_ms) for r in self.done]) if self.done else np.zeros(1)
        wait = np.array([r.admitted - r.arrival for r in self.done]) if self.done else np.zeros(1)
        status = {}
        for r in self.done:
            status[r.status] = status.get(r.status, 0) + 1
        rep = {"requests": len(self.done), "status": status, "wall_s": wall, "mcell_steps_per_s": cell_steps / wall / 1e6,
               "step_ms_p50": float(np.percentile(lat, 50)), "step_ms_p99": float(np.percentile(lat, 99)),
               "queue_wait_s_p50": float(np.percentile(wait, 50)), "queue_wait_s_p99": float(np.percentile(wait, 99)),
               "peak_budget_gb": self.peak / 1e9, "budget_gb": self.budget / 1e9,
               "grids": sorted({r.member.n for r in self.done}),
               "faces_range": [min(20 * (r.member.freq or 2 ** r.member.level) ** 2 for r in self.done),
                               max(20 * (r.member.freq or 2 ** r.member.level) ** 2 for r in self.done)] if self.done else []}
        if keep_results:
            rep["results"] = results
        return rep


def main() -> None:
    ap = argparse.ArgumentParser(description="serving loop over a Poisson request stream")
    ap.add_argument("--duration", type=float, default=10.0, help="arrival window (s)")
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
