"""Snapshot and resume (the reference's continuation files): a context that
resumes from a snapshot continues bit for bit like the one that took it, with
HIT forcing on (bhat and the generator position carried over), in both
precisions and with graph replay; a snapshot is refused by a context whose
mesh or configuration differs."""
import numpy as np
import pytest

from melt_ibm_amd import geometry
from melt_ibm_amd.api import MeltContext, MeltError
from melt_ibm_amd.inputs import inqpr_temperature
from tests.helpers import load_state, make_cfg, melt_case

pytestmark = pytest.mark.gpu

FIELDS = ("vx", "vy", "vz", "temp", "pr", "VOFp", "forcx")


def _setup(cfg, lib, st, dt, level=2, rad=1.2):
    c = MeltContext(cfg, lib=lib)
    m = geometry.sphere_level(level)
    c.set_geometry(m.xyz, m.vert_of_edge, m.edge_of_face, rad)
    return c


def _step(c):
    c.hit_force()
    c.tschem()


@pytest.mark.parametrize("precision,graph", [(0, False), (1, False), (0, True)])
def test_resume_continues_bit_for_bit(gpu_lib, tmp_path, precision, graph):
    cfg = make_cfg(32, precision=precision)
    st, dt = melt_case(make_cfg(32), 2, 1.2, seed=9)
    a = _setup(cfg, gpu_lib, st, dt)
    load_state(a, st)
    a.upload("temp", inqpr_temperature(cfg, a.body().pos_cm[:], 1.2))
    a.set_dt(dt)
    a.hit_init(0.5, 0.05, 2.5, seed=7)
    a.tagCells()
    a.graph(graph)
    for _ in range(3):
        _step(a)
    a.sync()
    blob = a.snapshot()
    a.save(tmp_path / "state.snap")
    for _ in range(2):
        _step(a)
    a.sync()
    for source in ("blob", "file"):
        b = _setup(cfg, gpu_lib, st, dt)
        b.hit_init(0.9, 0.01, 1.5, seed=1)       # replaced by the snapshot's forcing state
        if source == "blob":
            b.restore(blob)
        else:
            b.load(tmp_path / "state.snap")
        b.graph(graph)
        for _ in range(2):
            _step(b)
        b.sync()
        for q in FIELDS:
            assert np.array_equal(a.download(q), b.download(q)), (source, q)
        assert np.array_equal(a.lag("xyzv"), b.lag("xyzv")), source
        ba, bb = a.body(), b.body()
        assert ba.volume == bb.volume and list(ba.pos_cm) == list(bb.pos_cm) and list(ba.omega_c) == list(bb.omega_c)
        assert np.array_equal(a.hit_state()[0], b.hit_state()[0]) and a.hit_state()[1] == b.hit_state()[1]
        b.close()
    a.close()


def test_snapshot_refused_by_a_different_context(gpu_lib):
    cfg = make_cfg(32)
    st, dt = melt_case(cfg, 2, 1.2, seed=9)
    a = _setup(cfg, gpu_lib, st, dt)
    a.set_dt(dt)
    blob = a.snapshot()
    other_mesh = _setup(cfg, gpu_lib, st, dt, level=3)
    with pytest.raises(MeltError) as e:
        other_mesh.restore(blob)
    assert e.value.code == 1
    other_cfg = _setup(make_cfg(32, ren=123.0), gpu_lib, st, dt)
    with pytest.raises(MeltError):
        other_cfg.restore(blob)
    with pytest.raises(MeltError):
        a.restore(blob[:100])
    for c in (a, other_mesh, other_cfg):
        c.close()
