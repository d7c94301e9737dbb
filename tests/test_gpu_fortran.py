"""The Fortran binding (melt_ibm_amd/fortran/melt_ibm_f.f90) driven by a Fortran
program the way the reference's time loop drives its own routines, against the
ctypes path on the same inputs, bit for bit.

The program (tests/fortran/drive_tschem.f90, built by build()) pushes
reference-shaped arrays q(1:n1,1:n2,0:n3m+1) with n1 = n1m+1 (padding column
and row, ghost planes: the library's ld1 > n1m, kghost = 1 layout), runs the
initial tagCells and two tschem steps of a melting sphere, pulls the fields
back with their ghost planes, and runs the reference's solxi / solxj /
solxk(q(1:n1,1:n2,1:n3m), betadx) sequence (invtr1.f90:66-68) on a given rhs.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from melt_ibm_amd.api import MeltContext
from melt_ibm_amd.inputs import inqpr_temperature
from tests.helpers import make_cfg, melt_case, with_body

pytestmark = pytest.mark.gpu

DRIVER = Path(__file__).resolve().parent / "fortran" / "drive_tschem"
PAD = 1.0e30


def _ref_shape(a: np.ndarray) -> np.ndarray:
    """[k, j, i] interior -> the reference's (1:n1m+1, 1:n2m+1, 0:n3m+1) array
    (Fortran order, so [k, j, i] in C order) with periodic ghost planes and a
    padding column / row the library must ignore."""
    n3, n2, n1 = a.shape
    out = np.full((n3 + 2, n2 + 1, n1 + 1), PAD)
    out[1:n3 + 1, :n2, :n1] = a
    out[0, :n2, :n1] = a[-1]
    out[n3 + 1, :n2, :n1] = a[0]
    return out


def test_fortran_driver_matches_ctypes(gpu_lib, tmp_path):
    if not DRIVER.exists():
        pytest.skip("amdflang was not available to build the Fortran driver")
    n, level, rad, nsteps = 32, 2, 1.2, 2
    cfg = make_cfg(n, 24)
    st, dt = melt_case(cfg, level, rad, seed=4)
    c = MeltContext(cfg, lib=gpu_lib)
    m = with_body(c, level, rad)
    temp = inqpr_temperature(cfg, c.body().pos_cm[:], rad)
    fields = {"vx": st["vx"], "vy": st["vy"], "vz": st["vz"], "temp": temp, "forcx": st["forcx"]}
    rng = np.random.default_rng(8)
    rhs, q = rng.normal(size=cfg.shape), rng.normal(size=cfg.shape)
    bx, by, bz = 0.31, 0.27, 0.45

    with open(tmp_path / "case.bin", "wb") as f:
        f.write(bytes(cfg.to_c()))
        np.array([m.xyz.shape[0], m.vert_of_edge.shape[0], m.edge_of_face.shape[0], nsteps], dtype=np.int32).tofile(f)
        np.array([rad, dt, bx, by, bz], dtype=np.float64).tofile(f)
        np.ascontiguousarray(m.xyz, dtype=np.float64).tofile(f)
        np.ascontiguousarray(m.vert_of_edge, dtype=np.int32).tofile(f)
        np.ascontiguousarray(m.edge_of_face, dtype=np.int32).tofile(f)
        for name in ("vx", "vy", "vz", "temp", "forcx"):
            _ref_shape(fields[name]).tofile(f)
        _ref_shape(rhs).tofile(f)
        _ref_shape(q).tofile(f)
    print(f"# Fortran driver {n}x{n}x24, {nsteps} steps", flush=True)
    res = subprocess.run([str(DRIVER), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr

    for name, a in fields.items():
        c.upload(name, a)
    c.set_dt(dt)
    c.tagCells()
    for _ in range(nsteps):
        c.tschem()
    c.sync()
    out = np.fromfile(tmp_path / "out.bin", dtype=np.float64)
    shp = (cfg.n3m + 2, cfg.n2m + 1, cfg.n1m + 1)
    per = int(np.prod(shp))
    assert out.size == 7 * per
    for k, name in enumerate(("vx", "vy", "vz", "temp", "pr", "VOFp")):
        got = out[k * per:(k + 1) * per].reshape(shp)
        assert np.array_equal(got[:, :cfg.n2m, :cfg.n1m], c.download_with_ghosts(name)), name
        assert np.all(got[:, cfg.n2m, :] == -7.0) and np.all(got[:, :, cfg.n1m] == -7.0), f"{name}: padding written"
        assert np.array_equal(got[0, :cfg.n2m, :cfg.n1m], got[cfg.n3m, :cfg.n2m, :cfg.n1m])

    # the reference's line-solve sequence: the array form of solxk on the same
    # rhs equals the ctypes array form and the device-resident field form
    qf = out[6 * per:].reshape(shp)
    c.upload("rhs", rhs)
    c.set_stage(1)
    c.solxi(bx)
    c.solxj(by)
    qa = np.ascontiguousarray(_ref_shape(q)[1:cfg.n3m + 1])
    c.solxk_array(qa, bz)
    assert np.array_equal(qf[1:cfg.n3m + 1], qa)
    assert np.all(qf[0] == _ref_shape(q)[0]) and np.all(qf[:, cfg.n2m, :] == PAD)
    c.upload("rhs", rhs)
    c.upload("vx", q)
    c.solxi(bx)
    c.solxj(by)
    c.solxk("vx", bz)
    assert np.array_equal(c.download("vx"), qf[1:cfg.n3m + 1, :cfg.n2m, :cfg.n1m])
    c.close()
