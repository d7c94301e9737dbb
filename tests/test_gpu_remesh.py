"""Remeshing inside the fused RK3 step (IREMESH = 1, particle.f90:121-211)
against the oracle: a melting sphere (level-3 icosahedron, R 0.6, 64^3) with
E_thresh between the two shortest classes of edges, so the first substep
collapses every edge of the shortest class; the library measures the edges on
the device, remeshes on the host and refreshes the geometry on the device.
Topology (connectivity and ghost flags) and the event counters must equal the
oracle's; positions, fields and body state agree to the substep's parity bar
on the live entities.  With one stream and with the side stream; a snapshot
taken after remeshing resumes in a fresh context bit for bit."""
import numpy as np
import pytest

import oracle
from melt_ibm_amd import geometry
from melt_ibm_amd.api import MeltContext
from melt_ibm_amd.inputs import inqpr_temperature
from melt_ibm_amd.remesh import TriMesh
from tests.helpers import load_state, make_cfg, melt_case, rel_err, with_body

pytestmark = pytest.mark.gpu

N, LEVEL, RAD = 64, 3, 0.6


def remesh_cfg():
    el = np.unique(TriMesh.from_mesh(geometry.sphere_level(LEVEL), scale=RAD).e_lengths)
    h = 2 * np.pi / N
    return make_cfg(N, iremesh=1, perc_ethresh=0.5 * (el[0] + el[1]) / h)


def start(c, cfg, st, dt):
    load_state(c, st)
    with_body(c, LEVEL, RAD)
    c.upload("temp", inqpr_temperature(cfg, c.body().pos_cm[:], RAD))
    c.set_dt(dt)
    c.tagCells()


@pytest.mark.parametrize("streams", [0, 1])
def test_tschem_remeshes_like_the_oracle(gpu_lib, streams):
    cfg = remesh_cfg()
    st, dt = melt_case(cfg, LEVEL, RAD, seed=5)
    g, o = MeltContext(cfg, lib=gpu_lib), oracle.context(cfg)
    for c in (g, o):
        start(c, cfg, st, dt)
    if streams:
        g.set_tuning(streams=1)
    for step in range(2):
        print(f"# remesh parity step {step} streams={streams}", flush=True)
        g.tschem()
        o.tschem()
    g.sync()
    sg, so = g.remesh_stats(), o.remesh_stats()
    assert sg["events"] >= 1 and sg["collapses"] > 0, sg
    for k in ("events", "collapses", "deletions", "relaxations", "refreshes", "total_collapses"):
        assert sg[k] == so[k], (k, sg, so)
    assert abs(sg["residual"]) <= 1e-14 and abs(so["residual"]) <= 1e-14
    tg, to = g.topology(), o.topology()
    for k in tg:
        assert np.array_equal(tg[k], to[k]), k
    lv, lf = tg["ghost_v"] == 0, tg["ghost_f"] == 0
    assert lv.sum() < g.nv and lf.sum() < g.nf
    # vertices agree to about 1e-11 of their coordinates after two steps (the
    # same without remeshing); face and vertex areas are products of edge
    # vectors about 0.08 long, so the same absolute difference is a larger
    # fraction of them (measured 1.4e-10 of the largest face)
    for q, live, bar in (("xyzv", lv, 1e-10), ("tri_bar", lf, 1e-10), ("sur", lf, 1e-9), ("Avert", lv, 1e-9)):
        e = rel_err(g.lag(q)[live], o.lag(q)[live])
        assert e < bar, (q, e)
    # face velocities U + omega x r: the rigid-body velocities come from surface
    # integrals that nearly cancel, so they are measured against the flow's
    # velocity scale (rms 0.1), as in test_tschem_with_melting_body
    dvt = np.abs(g.lag("vel_tri")[lf] - o.lag("vel_tri")[lf]).max()
    assert dvt < 1e-9 * max(np.abs(o.lag("vel_tri")[lf]).max(), 0.1), dvt
    # The heat fluxes come from MLS probe fits whose moment inversions amplify
    # round-off of the fields and vertices: on this case without remeshing the
    # same two steps leave qw_o 1.4e-7, qw_oVert 4.2e-8 and vmelt 3.4e-8 apart
    # (positions 1e-11, T 4e-11), and remeshing does not change that level.
    for q, live in (("qw_oVert", lv), ("qw_iVert", lv), ("vmelt", lv), ("qw_o", lf), ("qw_i", lf)):
        e = rel_err(g.lag(q)[live], o.lag(q)[live])
        print(f"{q} rel err {e:.2e}", flush=True)
        assert e < 1e-6, (q, e)
    # fields: the oracle against itself, perturbed by one ulp, drifts to
    # u 5e-11..8e-11 and p 1.2e-9 by the third step on the reference case, so
    # two steps are held to 1e-9 (measured: u, v, w up to 1.2e-10)
    for q in ("vx", "vy", "vz", "temp", "pr", "VOFp"):
        e = rel_err(g.download(q), o.download(q))
        print(f"{q} rel err {e:.2e}", flush=True)
        assert e < 1e-9, (q, e)
    bg, bo = g.body(), o.body()
    assert abs(bg.volume - bo.volume) < 1e-10 * bo.volume      # measured 5e-12
    assert g.divgck() < 1e-10
    g.close()


def test_snapshot_after_remeshing_resumes(gpu_lib):
    cfg = remesh_cfg()
    st, dt = melt_case(cfg, LEVEL, RAD, seed=6)
    a = MeltContext(cfg, lib=gpu_lib)
    start(a, cfg, st, dt)
    a.tschem()
    assert a.remesh_stats()["events"] >= 1
    blob = a.snapshot()
    b = MeltContext(cfg, lib=gpu_lib)
    m = geometry.sphere_level(LEVEL)
    b.set_geometry(m.xyz, m.vert_of_edge, m.edge_of_face, RAD)
    b.restore(blob)
    for c in (a, b):
        c.tschem()
        c.sync()
    ta, tb = a.topology(), b.topology()
    for k in ta:
        assert np.array_equal(ta[k], tb[k]), k
    for q in ("vx", "vy", "vz", "temp", "pr"):
        assert np.array_equal(a.download(q), b.download(q)), q
    assert np.array_equal(a.lag("xyzv"), b.lag("xyzv"))
    a.close()
    b.close()
