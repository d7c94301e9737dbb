"""Host logic of the serving loop and the ensemble, on the CPU: first-come first-served
admission under the byte budget, and how a terminal device error ends one
request without ending the server.  The simulations are stand-ins that count
steps and raise the device's small-volume error (particle.f90:218-224) at a
chosen step, with the volume the body then has, or the window error, or the
remeshing residual error (particle.f90:205-210), which tschem raises itself
because the host edits the mesh inside it."""
import math

import pytest

from melt_ibm_amd import serve
from melt_ibm_amd.api import MeltError
from melt_ibm_amd.ensemble import Member


class _Body:
    def __init__(self, volume):
        self.volume = volume


class _Ctx:
    def __init__(self, fail_at, volume_after):
        self.steps, self.fail_at, self.volume_after = 0, fail_at, volume_after

    def tschem(self):
        self.steps += 1
        if self.fail_at and self.steps == self.fail_at and self.volume_after == "remesh":
            raise MeltError(6, "remesh_smooth: non-zero remeshing residual")

    def poll(self):
        return True

    def sync(self):
        if self.fail_at and self.steps == self.fail_at:
            if self.volume_after == "window":
                raise MeltError(9, "device: immersed body outgrew the device window")
            raise MeltError(4, "device: Geometry V/VE is smaller than threshold")

    def body(self):
        failed = self.fail_at and self.steps >= self.fail_at
        if failed and self.volume_after is None:
            raise MeltError(1, "hipMemcpy: unspecified launch failure")
        return _Body(self.volume_after if failed else 0.5)


class _Sim:
    behaviour = {}     # member seed -> (fail_at step or 0, volume after the failure)
    built = []

    def __init__(self, m):
        self.member = m
        self.ctx = _Ctx(*self.behaviour.get(m.seed, (0, 0.0)))
        self.closed = False

    @classmethod
    def build(cls, m, precision, lib, tuning, graph=True, device_inputs=False):
        s = cls(m)
        cls.built.append((m.seed, s))
        return s

    def close(self):
        self.closed = True


@pytest.fixture
def fake_sims(monkeypatch):
    _Sim.behaviour, _Sim.built = {}, []
    monkeypatch.setattr(serve, "Simulation", _Sim)
    return _Sim


def _member(seed, steps, n=64):
    return Member(n=n, rad=0.5, st=1.0, re=500.0, seed=seed, level=3, steps=steps)


def test_terminal_errors_end_only_their_request(fake_sims):
    # 4: the body cannot even be read back after the error
    fake_sims.behaviour = {1: (2, 1e-7), 2: (3, -7.4e33), 3: (2, math.nan), 4: (1, None), 5: (3, "remesh"), 6: (1, "remesh")}
    reqs = [serve.Request(i, 0.0, _member(i, 5)) for i in range(7)]
    srv = serve.Server(budget_bytes=1e12, precision=1)
    rep = srv.run(reqs)
    status = {r.rid: r.status for r in srv.done}
    assert status == {0: "ok", 1: "melted", 2: "diverged", 3: "diverged", 4: "diverged", 5: "remesh failed", 6: "remesh failed"}
    steps = {r.rid: r.done_steps for r in srv.done}
    assert steps == {0: 5, 1: 2, 2: 3, 3: 2, 4: 1, 5: 3, 6: 1}
    assert rep["status"] == {"ok": 1, "melted": 1, "diverged": 3, "remesh failed": 2}
    assert srv.used == 0 and not srv.active
    assert all(s.closed for _, s in fake_sims.built)


def test_admission_is_first_come_first_served_within_the_budget(fake_sims):
    big, small = _member(0, 3, n=128), _member(1, 1, n=64)
    later = _member(2, 1, n=64)
    b_big, b_small = serve.request_bytes(big, 1), serve.request_bytes(small, 1)
    budget = b_big + b_small - 1          # the big request and a small one never fit together
    reqs = [serve.Request(0, 0.0, big), serve.Request(1, 0.0, small), serve.Request(2, 0.0, later)]
    srv = serve.Server(budget_bytes=budget, precision=1)
    rep = srv.run(reqs)
    assert [seed for seed, _ in fake_sims.built] == [0, 1, 2]
    admitted = {r.rid: r.admitted for r in srv.done}
    finished = {r.rid: r.finished for r in srv.done}
    # the small requests wait behind the big one, although they arrived at the same time
    assert admitted[1] >= finished[0] and admitted[2] >= finished[0]
    assert rep["peak_budget_gb"] * 1e9 <= budget
    assert rep["requests"] == 3 and rep["status"] == {"ok": 3}


def test_request_larger_than_the_budget_is_refused(fake_sims):
    m = _member(0, 1, n=256)
    with pytest.raises(ValueError, match="exceeds the HBM budget"):
        serve.Server(budget_bytes=serve.request_bytes(m, 1) - 1, precision=1).run([serve.Request(0, 0.0, m)])


def test_pencil_overflow_steps_the_mesh_down(monkeypatch):
    """A member whose mesh puts more than n1m faces in one ray-tagging pencil
    (pencil_rayTag.f90:126-130, device error 5) is rebuilt one icosahedron
    level coarser (or at 3/4 of its frequency); other errors propagate."""
    from melt_ibm_amd import ensemble
    tried = []

    def fake_build(cls, m, precision, lib, tuning, graph, device_inputs=False):
        f = m.freq or 2 ** m.level
        tried.append(f)
        if f > 8:
            raise MeltError(5, "device: pencil buffer overflow")
        return ("built", f)

    monkeypatch.setattr(ensemble.Simulation, "_build", classmethod(fake_build))
    assert ensemble.Simulation.build(Member(n=32, rad=1.0, st=1.0, re=500.0, seed=0, level=5)) == ("built", 8)
    assert tried == [32, 16, 8]
    tried.clear()
    assert ensemble.Simulation.build(Member(n=32, rad=1.0, st=1.0, re=500.0, seed=0, freq=20)) == ("built", 8)
    assert tried == [20, 15, 11, 8]

    def other_error(cls, m, precision, lib, tuning, graph, device_inputs=False):
        raise MeltError(3, "device: window overflow")

    monkeypatch.setattr(ensemble.Simulation, "_build", classmethod(other_error))
    with pytest.raises(MeltError) as ei:
        ensemble.Simulation.build(Member(n=32, rad=1.0, st=1.0, re=500.0, seed=0, level=4))
    assert ei.value.code == 3


def test_setup_failure_ends_only_its_request(fake_sims, monkeypatch):
    # any error while a request is set up (context, mesh, tagging) marks that
    # request failed and returns its bytes; the others run to completion
    real_build = fake_sims.build.__func__

    def build(cls, m, precision, lib, tuning, graph=True, device_inputs=False):
        if m.seed == 1:
            raise MeltError(2, "hipMalloc: out of memory")
        if m.seed == 2:
            raise RuntimeError("mesh file unreadable")
        return real_build(cls, m, precision, lib, tuning, graph, device_inputs)

    monkeypatch.setattr(fake_sims, "build", classmethod(build))
    reqs = [serve.Request(i, 0.0, _member(i, 3)) for i in range(4)]
    srv = serve.Server(budget_bytes=1e12, precision=1)
    rep = srv.run(reqs)
    status = {r.rid: r.status for r in srv.done}
    assert status == {0: "ok", 1: "failed 2", 2: "failed RuntimeError", 3: "ok"}
    assert {r.rid: r.done_steps for r in srv.done} == {0: 3, 1: 0, 2: 0, 3: 3}
    assert srv.used == 0 and rep["requests"] == 4


class _SlowCtx(_Ctx):
    """A step that stays in flight for `polls` polls after its launch."""

    def __init__(self, polls, log, rid):
        super().__init__(0, 0.0)
        self.polls, self.left, self.log, self.rid = polls, 0, log, rid

    def tschem(self):
        super().tschem()
        self.left = self.polls

    def poll(self):
        if self.left:
            self.left -= 1
            return False
        return True

    def sync(self):
        self.log.append(self.rid)


def test_requests_advance_at_their_own_pace(fake_sims):
    # one step of request 0 stays in flight for 200 polls, request 1's steps
    # for none: request 1 runs all its steps while request 0's second step is
    # still running (no lock-step tick), and each step's inter-step time is
    # recorded per request
    log = []
    reqs = [serve.Request(0, 0.0, _member(0, 2)), serve.Request(1, 0.0, _member(1, 20))]
    srv = serve.Server(budget_bytes=1e12, precision=1)
    orig = fake_sims.__init__

    def init(self, m):
        orig(self, m)
        self.ctx = _SlowCtx(200 if m.seed == 0 else 0, log, m.seed)

    fake_sims.__init__ = init
    try:
        rep = srv.run(reqs)
    finally:
        fake_sims.__init__ = orig
    done = {r.rid: r for r in srv.done}
    assert done[0].done_steps == 2 and done[1].done_steps == 20
    # every step of request 1 completed before request 0's second step did
    assert log.index(0, log.index(0) + 1) > max(i for i, rid in enumerate(log) if rid == 1)
    assert len(done[1].step_ms) == 19 and len(done[0].step_ms) == 1
    assert rep["steps"] == 22


def test_ensemble_retires_melted_members(monkeypatch):
    """A member whose body melts below V_thresh ends there (particle.f90:218-224),
    as "diverged" when its volume is negative, and one whose remeshing leaves a
    residual ends there too (particle.f90:205-210), while the others keep
    stepping; any other device error ends the ensemble."""
    from melt_ibm_amd import ensemble
    monkeypatch.setattr(ensemble, "Simulation", _Sim)
    _Sim.behaviour, _Sim.built = {1: (3, 1e-7), 2: (5, -3.0), 3: (4, "window"), 4: (2, "remesh")}, []
    ens = ensemble.Ensemble([_member(i, 0) for i in range(5)], precision=1, tune=False)
    st = ens.run(steps=6, warmup=0)
    assert st["ended"] == {"1": {"step": 2, "status": "melted"}, "2": {"step": 4, "status": "diverged"},
                           "3": {"step": 3, "status": "diverged"}, "4": {"step": 1, "status": "remesh failed"}}
    assert [s.ctx.steps for s in ens.sims] == [6, 3, 5, 4, 2]
    assert sum(ens.step_cells) == _member(0, 0).cells * (6 + 3 + 5 + 4 + 2)
    _Sim.behaviour, _Sim.built = {0: (2, None)}, []

    class _Broken(_Ctx):
        def sync(self):
            if self.steps == 2:
                raise MeltError(2, "hipStreamSynchronize: unspecified launch failure")

    ens = ensemble.Ensemble([_member(0, 0)], precision=1, tune=False)
    ens.sims[0].ctx = _Broken(0, 0.0)
    with pytest.raises(MeltError):
        ens.run(steps=3, warmup=0)
