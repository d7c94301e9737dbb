"""melt_divgck_noalloc (no context, no allocation: its scratch is a file-scope
device array) against melt_divgck on a context holding the same fields, to the
bit: two calls in a row, the second on a smaller grid, for FP64 and FP32
fields, plus the argument checks that refuse unsupported sizes."""
import ctypes as C

import numpy as np
import pytest

from melt_ibm_amd.api import MeltContext
from tests.helpers import make_cfg, random_state

pytestmark = pytest.mark.gpu


def _bind(lib):
    f = lib.melt_divgck_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p] * 3 + [C.c_int32] * 3 + [C.c_double] * 3 + [C.c_int32, C.c_void_p, C.c_void_p]
    return f


def _fields(cfg, seed):
    st = random_state(cfg, seed)
    rng = np.random.default_rng(seed + 100)
    return [st[q] + rng.normal(0.0, 0.05, cfg.shape) for q in ("vx", "vy", "vz")]   # not divergence-free


def _noalloc(f, fields, cfg, elem):
    import torch
    dt = torch.float64 if elem == 8 else torch.float32
    t = [torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dt) for a in fields]
    out = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    print(f"# divgck_noalloc {cfg.shape} elem {elem}", flush=True)
    rc = f(t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(), cfg.n1m, cfg.n2m, cfg.n3m, cfg.xlen, cfg.ylen, cfg.zlen, elem, None,
           out.data_ptr())
    assert rc == 0
    torch.cuda.synchronize()
    return out.item()


@pytest.mark.parametrize("precision", [0, 1])
def test_divgck_noalloc_matches_context_entry(gpu_lib, precision):
    f = _bind(gpu_lib)
    elem = 8 if precision == 0 else 4
    for n, n3, seed in ((48, 48, 1), (32, 24, 2)):     # the second call on a smaller grid
        cfg = make_cfg(n, n3, precision=precision, xlen=6.0, ylen=6.0, zlen=5.0)
        fields = _fields(cfg, seed)
        c = MeltContext(cfg, lib=gpu_lib)
        for q, a in zip(("vx", "vy", "vz"), fields):
            c.upload(q, a)
        ref = c.divgck()
        c.close()
        got = _noalloc(f, fields, cfg, elem)
        assert got == ref and ref > 0.0, (n, got, ref)


def test_divgck_noalloc_refuses_unsupported_sizes(gpu_lib):
    f = _bind(gpu_lib)
    buf = C.c_double(0.0)
    p = C.addressof(buf)      # never dereferenced: every call below is refused on the host
    assert f(p, p, p, 15, 16, 16, 1.0, 1.0, 1.0, 8, None, p) == 1          # n1 not a multiple of 2
    assert f(p, p, p, 18, 16, 16, 1.0, 1.0, 1.0, 4, None, p) == 1          # n1 not a multiple of 4
    assert f(p, p, p, 2048, 1024, 1024, 1.0, 1.0, 1.0, 8, None, p) == 1    # 2^31 cells
    assert f(p, p, p, 16, 16, 16, 1.0, 1.0, 1.0, 2, None, p) == 1          # element size
    assert f(None, p, p, 16, 16, 16, 1.0, 1.0, 1.0, 8, None, p) == 1
