"""The diagnostics of the reference's time loop on the GPU against the oracle:
get_dt (cfl.f90:33-77) and CalcDissipation (diss.f90:11-201), from the same
state with a melting sphere (the dissipation re-tags the body's cells first).
get_dt is a max reduction of per-cell values computed in the same order:
bit-exact.  The dissipation sums differ from the oracle's serial sums only by
summation order (about 1e-13 relative)."""
import numpy as np
import pytest

import oracle
from melt_ibm_amd.api import MeltContext
from melt_ibm_amd.inputs import inqpr_temperature
from tests.helpers import load_state, make_cfg, random_state, with_body

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,n3", [(32, 32), (48, 24)])
def test_get_dt_and_dissipation_match_oracle(gpu_lib, n, n3):
    cfg = make_cfg(n, n3, prandtl=0.7)
    st = random_state(cfg, 6)
    st.pop("VOFp")
    g, o = MeltContext(cfg, lib=gpu_lib), oracle.context(cfg)
    for c in (g, o):
        load_state(c, st)
        with_body(c, 2, 1.1)
        c.upload("temp", inqpr_temperature(cfg, c.body().pos_cm[:], 1.1) + 0.1 * st["temp"])
        c.set_dt(0.004)
        c.tagCells()
    for cflfix in (0.5, 1.3):
        a, b = g.get_dt(cflfix), o.get_dt(cflfix)
        assert a == b, (cflfix, a, b)
    dg, do = g.dissipation(), o.dissipation()
    assert np.array_equal(g.download("VOFp"), o.download("VOFp"))
    assert dg["sumvof"] == do["sumvof"] and 0 < do["sumvof"] < n * n * n3
    for k in do:
        assert abs(dg[k] - do[k]) <= 1e-12 * abs(do[k]), (k, dg[k], do[k])
    g.close()
    o.close()


def test_dissipation_precision_form(gpu_lib):
    """The precision form's statistics (FP32 fields, FP64 sums) stay within
    1e-6 of the FP64 path's on the same state."""
    out = []
    for precision in (0, 1):
        cfg = make_cfg(32, precision=precision)
        st = random_state(make_cfg(32), 2)
        c = MeltContext(cfg, lib=gpu_lib)
        load_state(c, st)
        out.append(c.dissipation())
        c.close()
    for k in out[0]:
        if k != "sumvof":
            assert abs(out[1][k] - out[0][k]) <= 1e-6 * abs(out[0][k]), k


def test_timing_of_selected_slots(gpu_lib):
    """melt_timing_select times only the slots in its mask (the bench times
    the dominant kernel alone over its timed loop); melt_timing_enable times
    every slot; timing changes no result."""
    from tests.helpers import melt_case
    cfg = make_cfg(32)
    st, dt = melt_case(cfg, 2, 1.2, seed=3)
    a, b = MeltContext(cfg, lib=gpu_lib), MeltContext(cfg, lib=gpu_lib)
    for c in (a, b):
        load_state(c, st)
        with_body(c, 2, 1.2)
        c.set_dt(dt)
        c.tagCells()
    a.timing(True, slots=("explicit_xsolve",))
    a.timing_reset()
    a.tschem()
    b.tschem()
    tm = a.timing_read()
    assert tm["explicit_xsolve"][1] == 3 and tm["explicit_xsolve"][0] > 0.0
    assert all(n == 0 for k, (_, n) in tm.items() if k != "explicit_xsolve"), tm
    a.timing(True)
    a.timing_reset()
    a.tschem()
    b.tschem()
    tm = a.timing_read()
    a.timing(False)
    assert tm["substep"][1] == 3 and tm["project"][1] == 3 and tm["explicit_xsolve"][1] == 3
    a.sync(); b.sync()
    for q in ("vx", "vy", "vz", "temp", "pr"):
        assert np.array_equal(a.download(q), b.download(q)), q
    with pytest.raises(Exception):
        a._chk(a._f["timing_select"](a._ctx, 1 << 31))
    a.close(); b.close()
