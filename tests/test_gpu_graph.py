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
