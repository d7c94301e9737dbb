"""Batched ensemble, autotuner and serving loop.

Every simulation owns its context, so running it inside an ensemble or a
serving mix must give exactly what it gives alone (same kernels, same
order, deterministic reductions): the bar is bit-for-bit equality.  The
autotuner only changes launch shapes, so tuned and default runs are
bit-identical too.
"""
import numpy as np
import pytest

from melt_ibm_amd.ensemble import Ensemble, Member, Simulation
from melt_ibm_amd.serve import Request, Server, request_bytes
from melt_ibm_amd.tune import autotune

pytestmark = pytest.mark.gpu

FIELDS = ("vx", "vy", "vz", "temp", "pr")


def small_members():
    return [Member(n=32, rad=0.9, st=1.0, re=300.0, seed=0, level=2),
            Member(n=48, rad=1.1, st=2.5, re=800.0, seed=1, level=3),
            Member(n=32, rad=0.7, st=0.6, re=500.0, seed=2, level=2, aspect=2.0, freq=4),
            Member(n=48, rad=0.8, st=3.5, re=250.0, seed=3, level=3)]


def alone(m, steps, precision, lib, tuning=None, graph=False, device_inputs=False):
    s = Simulation.build(m, precision, lib, tuning, graph=graph, device_inputs=device_inputs)
    for _ in range(steps):
        s.ctx.tschem()
    s.ctx.sync()
    out = {q: s.ctx.download(q) for q in FIELDS}
    out["volume"] = s.ctx.body().volume
    s.close()
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_ensemble_members_match_alone(gpu_lib, precision):
    members = small_members()
    ens = Ensemble(members, precision=precision, lib=gpu_lib, tune=False)
    stats = ens.run(steps=2, warmup=1)
    assert stats["sims"] == 4 and stats["step_ms_p99"] >= stats["step_ms_p50"] > 0
    got = [{q: s.ctx.download(q) for q in FIELDS} | {"volume": s.ctx.body().volume} for s in ens.sims]
    ens.close()
    for m, g in zip(members, got):
        ref = alone(m, 3, precision, gpu_lib)
        for q in FIELDS:
            assert np.array_equal(g[q], ref[q]), (m, q)
        assert g["volume"] == ref["volume"]


def test_autotuned_shapes_do_not_change_results(gpu_lib):
    tuning = autotune(48, 0, lib=gpu_lib)
    assert set(tuning) <= {"xf_w", "xf_nt", "solve_w", "solve_nt", "xf_z", "yz_jb"} and tuning
    m = small_members()[1]
    a = alone(m, 2, 0, gpu_lib)
    b = alone(m, 2, 0, gpu_lib, tuning=tuning, graph=True)
    for q in FIELDS:
        assert np.array_equal(a[q], b[q]), q


def test_serving_results_match_alone_and_budget_holds(gpu_lib):
    members = small_members()
    steps = (3, 2, 4, 2)
    arrivals = (0.0, 0.0, 0.05, 0.1)
    reqs = []
    for i, (m, st, t) in enumerate(zip(members, steps, arrivals)):
        m.steps = st
        reqs.append(Request(i, t, m))
    # room for two 48^3 or three 32^3 simulations at a time: later ones must wait
    budget = request_bytes(members[1], 1) + request_bytes(members[3], 1) + 1
    srv = Server(budget, precision=1, lib=gpu_lib)
    rep = srv.run(reqs, keep_results=True)
    assert rep["requests"] == 4
    assert rep["peak_budget_gb"] * 1e9 <= budget
    assert rep["step_ms_p99"] >= rep["step_ms_p50"] > 0
    for i, m in enumerate(members):
        ref = alone(m, steps[i], 1, gpu_lib, device_inputs=True)    # the server synthesises inputs on the GPU
        for q in FIELDS:
            assert np.array_equal(rep["results"][i][q], ref[q]), (i, q)
        assert rep["results"][i]["volume"] == ref["volume"]


def test_diverged_request_ends_cheaply(gpu_lib):
    """A 256^3 spheroid (aspect 2.7, 24500 faces) at St 3.1 and Re 273 with
    remeshing off: its melting front moves about a cell per step, the surface
    tangles at step 29 and the volume check (particle.f90:218-224) ends the
    request, with a negative volume, so the server reports it as diverged.
    The blown-up coordinates put every thread of the tagging and forcing
    passes on the error word; that step costs as much as any other (measured
    4.2 ms, against 272 ms when each thread queued a compare-and-swap on it)."""
    m = Member(n=256, rad=0.5558868667770105, st=3.0961983157665, re=273.19648405043654, seed=6, aspect=2.6943004927317387,
               freq=35, steps=45, remesh=False)
    srv = Server(200e9, precision=1, lib=gpu_lib, device_inputs=False)     # the seeded host state that tangles
    rep = srv.run([Request(0, 0.0, m)])
    r = srv.done[0]
    assert rep["status"] == {"diverged": 1}
    assert r.done_steps < m.steps
    assert max(r.step_ms[1:]) < 60.0, r.step_ms


@pytest.mark.parametrize("precision", [0, 1])
def test_admission_estimate_covers_device_memory(gpu_lib, precision):
    torch = pytest.importorskip("torch")
    m = Member(n=64, rad=0.6, st=1.0, re=500.0, seed=0, level=3)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    s = Simulation.build(m, precision, gpu_lib)
    s.ctx.tschem()
    s.ctx.sync()
    used = free0 - torch.cuda.mem_get_info()[0]
    s.close()
    assert 0 < used <= request_bytes(m, precision), (used, request_bytes(m, precision))
