"""HIT forcing (hit.f90) on the GPU against the oracle.

The OU update uses the same counter-based generator on both sides; the device's
log / cos / sin round within an ulp of the host's, so bhat agrees to about
1e-16 relative.  The synthesis sums the same trigonometric polynomial in a
separable order (the oracle follows the reference's direct nn / m / l loop),
so the fields agree to round-off of the field scale."""
import numpy as np
import pytest

import oracle
from melt_ibm_amd.api import MeltContext
from tests.helpers import make_cfg, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,n3,kf", [(32, 32, 2.5), (48, 24, 2.3), (16, 16, 4.5)])
def test_hit_forcing_matches_oracle(gpu_lib, n, n3, kf):
    cfg = make_cfg(n, n3, xlen=6.0, ylen=6.0, zlen=5.0)
    g, o = MeltContext(cfg, lib=gpu_lib), oracle.context(cfg)
    for c in (g, o):
        c.set_dt(0.013)
        c.hit_init(0.5, 0.05, kf, seed=42)
    assert g.hit_nmodes() == o.hit_nmodes()
    for step in range(4):
        g.hit_force()
        o.hit_force()
        bg, sg = g.hit_state()
        bo, so = o.hit_state()
        assert sg == so == step + 1
        assert np.max(np.abs(bg - bo)) <= 1e-14 * np.max(np.abs(bo)), step
        for q in ("forcx", "forcy", "forcz"):
            e = rel_err(g.download(q), o.download(q))
            assert e < 1e-13, (step, q, e)
    # continuation: a context resumed from the saved bhat and position goes on identically
    b, s = g.hit_state()
    r = MeltContext(cfg, lib=gpu_lib)
    r.set_dt(0.013)
    r.hit_init(0.5, 0.05, kf, seed=42)
    r.set_hit_state(b, s)
    g.hit_force()
    r.hit_force()
    for q in ("forcx", "forcy", "forcz"):
        assert np.array_equal(g.download(q), r.download(q)), q
    for c in (g, o, r):
        c.close()


def test_abc_forcing_matches_oracle(gpu_lib):
    cfg = make_cfg(32, 24, xlen=6.0, ylen=5.0, zlen=4.0)
    g, o = MeltContext(cfg, lib=gpu_lib), oracle.context(cfg)
    for c in (g, o):
        c.abc_force(0.1)
    for q in ("forcx", "forcy", "forcz"):
        assert rel_err(g.download(q), o.download(q)) < 1e-15, q
    g.close()
    o.close()


def test_hit_forcing_precision_form(gpu_lib):
    out = []
    for precision in (0, 1):
        c = MeltContext(make_cfg(32, precision=precision), lib=gpu_lib)
        c.set_dt(0.01)
        c.hit_init(0.5, 0.05, 2.5, seed=1)
        c.hit_force()
        out.append([c.download(q) for q in ("forcx", "forcy", "forcz")])
        c.close()
    for a, b in zip(*out):
        assert rel_err(b, a) < 1e-7
