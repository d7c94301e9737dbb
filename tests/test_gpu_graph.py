"""Graph replay of the RK3 step (melt_graph_enable) gives the eager result bit
for bit, replay after replay (with host work and synchronisation between
the replays), in both precisions, and across a dt change (re-capture)."""
import numpy as np
import pytest

from melt_ibm_amd.api import MeltContext
from melt_ibm_amd.inputs import inqpr_temperature
from tests.helpers import load_state, make_cfg, melt_case, with_body

pytestmark = pytest.mark.gpu

FIELDS = ("vx", "vy", "vz", "temp", "pr", "VOFp")


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_replay_matches_eager(gpu_lib, precision):
    n, level, rad = 32, 2, 1.2
    cfg = make_cfg(n, precision=precision)
    st, dt = melt_case(make_cfg(n), level, rad, seed=3)
    eager, graph = MeltContext(cfg, lib=gpu_lib), MeltContext(cfg, lib=gpu_lib)
    for c in (eager, graph):
        load_state(c, st)
        with_body(c, level, rad)
        c.upload("temp", inqpr_temperature(cfg, c.body().pos_cm[:], rad))
        c.set_dt(dt)
        c.tagCells()
    graph.graph(True)
    for step in range(4):
        if step == 2:  # new dt: the captured graph must be replaced
            for c in (eager, graph):
                c.set_dt(0.9 * dt)
        eager.tschem()
        graph.tschem()
        eager.sync()
        graph.sync()
        for q in FIELDS:
            assert np.array_equal(eager.download(q), graph.download(q)), f"step {step}: {q}"
        assert np.array_equal(eager.lag("xyzv"), graph.lag("xyzv")), f"step {step}: xyzv"
        be, bg = eager.body(), graph.body()
        assert be.volume == bg.volume and list(be.pos_cm) == list(bg.pos_cm)


@pytest.mark.parametrize("use_graph", [False, True])
def test_queued_probe_geometry_matches_recomputed(gpu_lib, use_graph):
    """The last substep of tschem queues the gradient probes' geometry stage
    for the next call; a change of the body in between cancels it (and drops a
    captured step), and the next call recomputes it.  Re-setting the same body
    state before every call must therefore give the same bits as not touching
    it: the queued coefficients equal freshly computed ones."""
    n, level, rad = 32, 2, 1.2
    cfg = make_cfg(n)
    st, dt = melt_case(cfg, level, rad, seed=5)
    plain, cancelled = MeltContext(cfg, lib=gpu_lib), MeltContext(cfg, lib=gpu_lib)
    for c in (plain, cancelled):
        load_state(c, st)
        with_body(c, level, rad)
        c.upload("temp", inqpr_temperature(cfg, c.body().pos_cm[:], rad))
        c.set_dt(dt)
        c.tagCells()
        c.graph(use_graph)
    for step in range(3):
        plain.tschem()
        cancelled.set_body(cancelled.body())
        cancelled.tschem()
        plain.sync()
        cancelled.sync()
        for q in FIELDS:
            assert np.array_equal(plain.download(q), cancelled.download(q)), f"step {step}: {q}"
        assert np.array_equal(plain.lag("xyzv"), cancelled.lag("xyzv")), f"step {step}: xyzv"
    plain.close()
    cancelled.close()


def test_substeps_between_graph_replays_match_eager(gpu_lib):
    """A captured step whose first substep starts from the probe coefficients
    queued by the previous call must not be replayed after melt_substep calls
    have moved the faces: graph tschem, substeps 1-3, graph tschem again (and
    a set_body in between) give the eager context's bits."""
    n, level, rad = 32, 2, 1.2
    cfg = make_cfg(n)
    st, dt = melt_case(cfg, level, rad, seed=11)
    eager, graph = MeltContext(cfg, lib=gpu_lib), MeltContext(cfg, lib=gpu_lib)
    for c in (eager, graph):
        load_state(c, st)
        with_body(c, level, rad)
        c.upload("temp", inqpr_temperature(cfg, c.body().pos_cm[:], rad))
        c.set_dt(dt)
        c.tagCells()
    graph.graph(True)
    script = ["tschem", "tschem", "sub", "tschem", "tschem", "sub", "body", "tschem", "tschem"]
    for i, op in enumerate(script):
        for c in (eager, graph):
            if op == "tschem":
                c.tschem()
            elif op == "sub":
                for ns in (1, 2, 3):
                    c.substep(ns)
            else:
                c.set_body(c.body())
            c.sync()
        for q in FIELDS:
            assert np.array_equal(eager.download(q), graph.download(q)), f"op {i} ({op}): {q}"
        assert np.array_equal(eager.lag("xyzv"), graph.lag("xyzv")), f"op {i} ({op}): xyzv"
    eager.close()
    graph.close()
