"""Pinning the oracle's Eulerian routines against closed forms (CPU, no GPU).

Every check below is stated independently of the oracle's own formula, so that
a plausible slip in the restatement fails it:
 - invtr1/2/3/invtrte: single Fourier modes are eigenfunctions of the discrete
   Laplacian, of the backward pressure difference (by cos a - cos b =
   -2 sin((a+b)/2) sin((a-b)/2)) and of the three Crank-Nicolson factors, so
   q_new = q + dt (gamma H + rho H_old + (alpha/Re) lambda q - alpha dp) / prod f
   holds in closed form with non-zero H, H_old and p (invtr1.f90:40-68,
   invtr2.f90, invtr3.f90, invtrte.f90:50-63);
 - prcalc: p gains phi (1 - alpha beta lambda) for a single mode (prcalc.f90:27-30);
 - hdnl1/2/3/hdnlte: the discrete flux divergence converges at second order to
   the analytic -div(u q) of smooth periodic fields on an anisotropic box
   (a scale, sign, spacing or neighbour slip leaves an O(1) or first-order
   error); the forcing terms are forc/len with each direction's own length;
 - hdnl3 buoyancy: betagz (T(zc) cos(h/2) - Tliq) for T = cos(z)
   (hdnl3.f90:66-67);
 - interp_vof: a step in VOFp puts 1/2 on the face between the last solid and
   the first fluid cell (tagCells.f90:105-107);
 - MLS supports: each component's 3x3x3 support is centred on the lattice
   point of that component's staggered grid nearest the face centroid, in
   (k, j, i) order (mlsWeight.f90:73-79, 164-170, 250-254, 345-352), shown by
   exact linear reproduction at the support coordinates implied by that rule;
 - cfl: a single non-zero face value s gives |s| dx / 2 along its own axis
   (cfl.f90:9-25);
 - inqpr: the tanh profile is radial, tends to Tsol inside and Tliq outside,
   crosses the mean at r = R with slope (Tliq - Tsol) dx1 / 4 (inqpr.f90:17-40).
"""
import numpy as np
import pytest

import oracle
from melt_ibm_amd import inputs
from tests.helpers import load_state, make_cfg, with_body

GAM = (8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0)
ROM = (0.0, -17.0 / 60.0, -5.0 / 12.0)


@pytest.fixture(scope="module")
def lib():
    return oracle.load()


def _grid(cfg):
    (xm, ym, zm), (xc, yc, zc) = inputs.grid_coords(cfg)
    return {"xm": xm, "ym": ym, "zm": zm, "xc": xc, "yc": yc, "zc": zc}


def _phase(cfg, g, kvec, at):
    """k.x on the [k, j, i] grid at the staggered location `at` = (x, y, z) names."""
    kx, ky, kz = (2 * np.pi * kvec[0] / cfg.xlen, 2 * np.pi * kvec[1] / cfg.ylen, 2 * np.pi * kvec[2] / cfg.zlen)
    return kx * g[at[0]][None, None, :] + ky * g[at[1]][None, :, None] + kz * g[at[2]][:, None, None]


def _lam(cfg, kvec):
    """Eigenvalue of the 7-point Laplacian for the mode kvec."""
    n = (cfg.n1m, cfg.n2m, cfg.n3m)
    d = (cfg.n1m / cfg.xlen, cfg.n2m / cfg.ylen, cfg.n3m / cfg.zlen)
    return -sum(2.0 * (1.0 - np.cos(2 * np.pi * kvec[a] / n[a])) * d[a] ** 2 for a in range(3))


def _adi(cfg, kvec, g_per_dx2):
    """prod_d (1 + 2 G_d (1 - cos(k_d h_d))) with G_d = g_per_dx2 * dx_d^2."""
    n = (cfg.n1m, cfg.n2m, cfg.n3m)
    d = (cfg.n1m / cfg.xlen, cfg.n2m / cfg.ylen, cfg.n3m / cfg.zlen)
    f = 1.0
    for a in range(3):
        f *= 1.0 + 2.0 * g_per_dx2 * d[a] ** 2 * (1.0 - np.cos(2 * np.pi * kvec[a] / n[a]))
    return f


# ------------------------------------------------------------------ invtr*
INVTR = {
    # routine: (field, location, explicit term H, its history, pressure axis or None)
    "invtr1": ("vx", ("xc", "ym", "zm"), "dq", "ru1", 0),
    "invtr2": ("vy", ("xm", "yc", "zm"), "dph", "ru2", 1),
    "invtr3": ("vz", ("xm", "ym", "zc"), "qcap", "ru3", 2),
    "invtrte": ("temp", ("xm", "ym", "zm"), "htemp", "rut", None),
}


@pytest.mark.parametrize("routine", sorted(INVTR))
@pytest.mark.parametrize("ns", [1, 2, 3])
def test_invtr_single_mode_closed_form(lib, routine, ns):
    cfg = make_cfg(16, 24, ren=20.0, prandtl=0.7, xlen=6.0, ylen=6.0, zlen=4.5)
    g = _grid(cfg)
    o = oracle.context(cfg)
    q_name, at, h_name, hist, pax = INVTR[routine]
    kvec = (2, 3, 4)
    th = _phase(cfg, g, kvec, at)
    thp = _phase(cfg, g, kvec, ("xm", "ym", "zm"))
    A, B, D, Cp = 0.8, 1.3, -0.6, 0.9
    q0 = A * np.cos(th)
    H = B * np.cos(th + 0.3)
    Hold = D * np.sin(th - 0.7)
    p = Cp * np.cos(thp + 0.5)
    dt = 0.05
    z = np.zeros(cfg.shape)
    load_state(o, {q: z for q in ("vx", "vy", "vz", "temp", "dq", "dph", "qcap", "htemp", "ru1", "ru2", "ru3", "rut")})
    load_state(o, {q_name: q0, h_name: H, hist: Hold, "pr": p})
    o.set_dt(dt)
    o.set_stage(ns)
    getattr(o, routine)()
    gam, rom = GAM[ns - 1], ROM[ns - 1]
    al = gam + rom
    if pax is None:
        nu, dp = 1.0 / (cfg.ren * cfg.prandtl), 0.0
    else:
        nu = 1.0 / cfg.ren
        n = (cfg.n1m, cfg.n2m, cfg.n3m)[pax]
        dxa = (cfg.n1m / cfg.xlen, cfg.n2m / cfg.ylen, cfg.n3m / cfg.zlen)[pax]
        # backward difference of p at the face midway between the two centres
        dp = -2.0 * Cp * np.sin(np.pi * kvec[pax] / n) * dxa * np.sin(th + 0.5)
    rhs = (gam * H + rom * Hold + al * nu * _lam(cfg, kvec) * q0 - al * dp) * dt
    expect = q0 + rhs / _adi(cfg, kvec, dt * nu * 0.5 * al)
    got = o.download(q_name)
    assert np.max(np.abs(got - expect)) < 1e-13 * np.max(np.abs(expect)), routine
    assert np.array_equal(o.download(hist), H), f"{routine}: history takes H"


def test_prcalc_single_mode(lib):
    cfg = make_cfg(16, 24, ren=2.0, xlen=6.0, ylen=6.0, zlen=4.5)
    g = _grid(cfg)
    o = oracle.context(cfg)
    kvec = (3, 1, 5)
    phi = 0.7 * np.cos(_phase(cfg, g, kvec, ("xm", "ym", "zm")) - 0.4)
    p0 = np.random.default_rng(2).normal(size=cfg.shape)
    dt, ns = 0.5, 2
    load_state(o, {"dph": phi, "pr": p0})
    o.set_dt(dt)
    o.set_stage(ns)
    o.prcalc()
    al, beta = GAM[ns - 1] + ROM[ns - 1], dt / cfg.ren * 0.5
    expect = p0 + phi * (1.0 - al * beta * _lam(cfg, kvec))
    assert al * beta * abs(_lam(cfg, kvec)) > 0.1     # the Laplacian term is not negligible here
    assert np.max(np.abs(o.download("pr") - expect)) < 1e-14 * np.max(np.abs(expect))


# ------------------------------------------------------- explicit terms
class Smooth:
    """Sum of products a * F1(kx x + p1) F2(ky y + p2) F3(kz z + p3) + c on a
    periodic box, with exact partial derivatives."""

    def __init__(self, cfg, terms, const=0.0):
        self.L = (cfg.xlen, cfg.ylen, cfg.zlen)
        self.terms = terms
        self.const = const

    def _f(self, kind, k, p, x, L, deriv):
        w = 2 * np.pi * k / L
        a = w * x + p
        if not deriv:
            return np.sin(a) if kind == "s" else np.cos(a)
        return w * np.cos(a) if kind == "s" else -w * np.sin(a)

    def __call__(self, x, y, z, d=None):
        out = 0.0 if d is not None else self.const
        for a, fx, fy, fz in self.terms:
            v = a
            for ax, (kind, k, p), c in zip(range(3), (fx, fy, fz), (x, y, z)):
                v = v * self._f(kind, k, p, c, self.L[ax], d == ax)
            out = out + v
        return out


def _fields(cfg):
    U = Smooth(cfg, [(1.0, ("s", 1, 0.3), ("c", 2, -0.1), ("c", 1, 0.2)), (0.5, ("c", 0, 0.0), ("c", 1, 0.0), ("c", 0, 0.0))])
    V = Smooth(cfg, [(1.0, ("c", 1, -0.4), ("s", 1, 0.7), ("c", 2, 0.0))], 0.3)
    W = Smooth(cfg, [(0.7, ("s", 2, 0.0), ("c", 1, -0.5), ("s", 1, 1.1))], -0.2)
    T = Smooth(cfg, [(1.0, ("c", 1, 0.1), ("s", 2, 0.0), ("c", 1, 0.0))], 0.5)
    return U, V, W, T


def _pts(g, at):
    return (g[at[0]][None, None, :], g[at[1]][None, :, None], g[at[2]][:, None, None])


def _flux_div(U, V, W, Q, P):
    """Analytic -(d(uq)/dx + d(vq)/dy + d(wq)/dz) at the points P."""
    x, y, z = P
    q = Q(x, y, z)
    return -(U(x, y, z, 0) * q + U(x, y, z) * Q(x, y, z, 0) + V(x, y, z, 1) * q + V(x, y, z) * Q(x, y, z, 1) +
             W(x, y, z, 2) * q + W(x, y, z) * Q(x, y, z, 2))


def _explicit_errors(n):
    cfg = make_cfg(n, 3 * n // 4, betagz=0.0, xlen=6.0, ylen=5.0, zlen=4.0)
    g = _grid(cfg)
    o = oracle.context(cfg)
    U, V, W, T = _fields(cfg)
    z = np.zeros(cfg.shape)
    st = {"vx": U(*_pts(g, ("xc", "ym", "zm"))) + z, "vy": V(*_pts(g, ("xm", "yc", "zm"))) + z,
          "vz": W(*_pts(g, ("xm", "ym", "zc"))) + z, "temp": T(*_pts(g, ("xm", "ym", "zm"))) + z,
          "VOFp": np.ones(cfg.shape), "forcx": z, "forcy": z, "forcz": z}
    load_state(o, st)
    o.interp_vof()
    o.hdnl1(); o.hdnl2(); o.hdnl3(); o.hdnlte()
    out = {}
    for name, Q, at in (("dq", U, ("xc", "ym", "zm")), ("dph", V, ("xm", "yc", "zm")),
                        ("qcap", W, ("xm", "ym", "zc")), ("htemp", T, ("xm", "ym", "zm"))):
        exact = _flux_div(U, V, W, Q, _pts(g, at)) + z
        out[name] = np.max(np.abs(o.download(name) - exact)) / np.max(np.abs(exact))
    return out


def test_explicit_terms_converge_to_the_flux_divergence(lib):
    e = [_explicit_errors(n) for n in (16, 32, 64)]
    for name in e[0]:
        r1, r2 = e[0][name] / e[1][name], e[1][name] / e[2][name]
        assert r1 > 3.3 and r2 > 3.6, (name, [x[name] for x in e])
        assert e[2][name] < 2e-2, (name, e[2][name])


def test_forcing_terms_use_their_own_box_length(lib):
    cfg = make_cfg(16, 12, betagz=0.0, xlen=6.0, ylen=5.0, zlen=4.0)
    o = oracle.context(cfg)
    rng = np.random.default_rng(9)
    z = np.zeros(cfg.shape)
    f = {q: rng.normal(size=cfg.shape) for q in ("forcx", "forcy", "forcz")}
    load_state(o, {"vx": z, "vy": z, "vz": z, "temp": z, "VOFp": np.ones(cfg.shape), **f})
    o.interp_vof()
    o.hdnl1(); o.hdnl2(); o.hdnl3()
    assert np.array_equal(o.download("dq"), f["forcx"] * 1.0 / cfg.xlen) or \
        np.max(np.abs(o.download("dq") - f["forcx"] / cfg.xlen)) < 1e-16 * 8
    assert np.max(np.abs(o.download("dph") - f["forcy"] / cfg.ylen)) < 1e-15
    assert np.max(np.abs(o.download("qcap") - f["forcz"] / cfg.zlen)) < 1e-15


def test_hdnl3_buoyancy_closed_form(lib):
    cfg = make_cfg(16, 24, betagz=0.7, tliq=0.4)
    g = _grid(cfg)
    o = oracle.context(cfg)
    z = np.zeros(cfg.shape)
    T = np.cos(g["zm"])[:, None, None] + z
    load_state(o, {"vx": z, "vy": z, "vz": z, "temp": T, "VOFp": np.ones(cfg.shape), "forcz": z})
    o.interp_vof()
    o.hdnl3()
    h = cfg.zlen / cfg.n3m
    # mean of the two cell centres either side of the w face: cos(zc) cos(h/2)
    expect = cfg.betagz * (np.cos(g["zc"])[:, None, None] * np.cos(h / 2) - cfg.tliq) + z
    assert np.max(np.abs(o.download("qcap") - expect)) < 1e-15 * 4


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_interp_vof_face_between_cells(lib, axis):
    cfg = make_cfg(16, 12)
    o = oracle.context(cfg)
    i0 = 5
    idx = np.arange(cfg.shape[2 - axis])
    step = (idx >= i0).astype(np.float64)     # cells i0.. are fluid
    shp = [1, 1, 1]
    shp[2 - axis] = -1
    vofp = np.broadcast_to(step.reshape(shp), cfg.shape).copy()
    o.upload("VOFp", vofp)
    o.interp_vof()
    face = o.download(("VOFx", "VOFy", "VOFz")[axis])
    prof = np.moveaxis(face, 2 - axis, 0)[:, 0, 0]
    # face index i (0-based) sits between cells i-1 and i
    assert prof[i0] == 0.5 and prof[i0 + 1] == 1.0 and prof[i0 - 1] == 0.0
    assert prof[0] == 0.5     # the periodic seam between the last (fluid) and first (solid) cell


# ------------------------------------------------------------ MLS supports
def test_mls_supports_centred_on_each_components_lattice(lib):
    cfg = make_cfg(32, 24)
    o = oracle.context(cfg)
    with_body(o, 2, 1.1)
    b = o.body()
    b.pos_cm[0] += 0.037    # off the symmetric centre
    b.pos_cm[1] -= 0.051
    o.set_body(b)
    o.findCentroidIndices()
    o.mlsWeight()
    bar = o.lag("tri_bar")
    h = (cfg.xlen / cfg.n1m, cfg.ylen / cfg.n2m, cfg.zlen / cfg.n3m)
    # lattice offsets: faces (c) at n h, centres (m) at (n + 1/2) h
    lattices = {"ptxAB_q1": (0.0, 0.5, 0.5), "ptxAB_q2": (0.5, 0.0, 0.5), "ptxAB_q3": (0.5, 0.5, 0.0), "ptxAB_temp": (0.5, 0.5, 0.5)}
    off = np.array([-1, 0, 1])
    for name, sh in lattices.items():
        phi = o.lag(name)                                   # (nf, 27), (k, j, i) order, i fastest
        pts = []
        for ax in range(3):
            c = np.floor(bar[:, ax] / h[ax] - sh[ax] + 0.5)     # nearest lattice point (unwrapped)
            pts.append((c[:, None] + off[None, :] + sh[ax]) * h[ax])
        X = np.broadcast_to(pts[0][:, None, None, :], (len(bar), 3, 3, 3)).reshape(len(bar), 27)
        Y = np.broadcast_to(pts[1][:, None, :, None], (len(bar), 3, 3, 3)).reshape(len(bar), 27)
        Z = np.broadcast_to(pts[2][:, :, None, None], (len(bar), 3, 3, 3)).reshape(len(bar), 27)
        for ax, P in enumerate((X, Y, Z)):
            err = np.max(np.abs(np.sum(phi * P, axis=1) - bar[:, ax]))
            assert err < 1e-8, (name, ax, err)


# ------------------------------------------------------------ cfl / inqpr
def test_cfl_single_face_value(lib):
    cfg = make_cfg(16, 12, xlen=6.0, ylen=5.0, zlen=4.0)
    d = (cfg.n1m / cfg.xlen, cfg.n2m / cfg.ylen, cfg.n3m / cfg.zlen)
    o = oracle.context(cfg)
    z = np.zeros(cfg.shape)
    load_state(o, {"vx": z, "vy": z, "vz": z})
    assert o.cfl() == 1e-8          # the floor cfl.f90:9 starts from
    for ax, name in enumerate(("vx", "vy", "vz")):
        a = z.copy()
        a[3, 4, 5] = -0.8
        load_state(o, {"vx": z, "vy": z, "vz": z, name: a})
        assert abs(o.cfl() - 0.8 * 0.5 * d[ax]) < 1e-15, name


def test_inqpr_profile(lib):
    cfg = make_cfg(64, tliq=1.0, tsol=0.2)
    o = oracle.context(cfg)
    with_body(o, 3, 1.0)
    o.inqpr()
    T = o.download("temp")
    for q in ("vx", "vy", "vz"):
        assert np.max(np.abs(o.download(q))) == 0.0
    g = _grid(cfg)
    pc = np.array(o.body().pos_cm[:])
    r = np.sqrt((g["xm"][None, None, :] - pc[0]) ** 2 + (g["ym"][None, :, None] - pc[1]) ** 2 + (g["zm"][:, None, None] - pc[2]) ** 2)
    order = np.argsort(r.ravel())
    rs, ts = r.ravel()[order], T.ravel()[order]
    assert np.all(np.diff(ts) >= -1e-15)                       # radial and non-decreasing outward
    assert abs(ts[0] - 0.2) < 1e-3 and abs(ts[-1] - 1.0) < 1e-12
    mid = 0.5 * (1.0 + 0.2)
    k = np.searchsorted(ts, mid)
    r_mid = rs[k - 1] + (mid - ts[k - 1]) * (rs[k] - rs[k - 1]) / (ts[k] - ts[k - 1])
    h = cfg.xlen / cfg.n1m
    assert abs(r_mid - 1.0) < 0.02 * h
    near = np.abs(rs - 1.0) < 0.1 * h
    slope = np.polyfit(rs[near], ts[near], 1)[0]
    assert abs(slope - (1.0 - 0.2) * (cfg.n1m / cfg.xlen) / 4.0) < 0.01 * slope


def test_solxk_array_form_equals_field_form(lib):
    # solxk(q, betadx) on an array section with the reference's padding
    # column / row (ld1 = n1m + 1) adds the same z solution as the field form
    cfg = make_cfg(16, 12)
    o = oracle.context(cfg)
    rng = np.random.default_rng(3)
    rhs, q = rng.normal(size=cfg.shape), rng.normal(size=cfg.shape)
    o.upload("rhs", rhs)
    o.upload("vx", q)
    o.solxk("vx", 0.4)
    padded = np.full((cfg.n3m, cfg.n2m + 1, cfg.n1m + 1), 5.0)
    padded[:, :cfg.n2m, :cfg.n1m] = q
    o.upload("rhs", rhs)
    o.solxk_array(padded, 0.4)
    assert np.array_equal(padded[:, :cfg.n2m, :cfg.n1m], o.download("vx"))
    assert np.all(padded[:, cfg.n2m, :] == 5.0) and np.all(padded[:, :, cfg.n1m] == 5.0)


# ------------------------------------------------------ get_dt / dissipation
def test_get_dt_closed_forms(lib):
    cfg = make_cfg(16, 12, ren=50.0, xlen=6.0, ylen=5.0, zlen=4.0)
    d = (cfg.n1m / cfg.xlen, cfg.n2m / cfg.ylen, cfg.n3m / cfg.zlen)
    o = oracle.context(cfg)
    z = np.zeros(cfg.shape)
    dt, cflfix = 0.01, 0.7
    load_state(o, {"vx": z, "vy": z, "vz": z})
    o.set_dt(dt)
    # at rest the viscous restriction 6 dt dx1^2 / Re sets dt (cfl.f90:54-55)
    assert abs(o.get_dt(cflfix) - cflfix * cfg.ren / (6.0 * d[0] ** 2)) < 1e-15
    for ax, name in enumerate(("vx", "vy", "vz")):
        a = z.copy()
        a[2, 3, 4] = -3.0
        load_state(o, {"vx": z, "vy": z, "vz": z, name: a})
        # a single face value s: |s| dt dx_axis (not averaged, cfl.f90:48-50)
        assert abs(o.get_dt(cflfix) - cflfix / (3.0 * d[ax])) < 1e-15 * cflfix / (3.0 * d[ax]) * 4, name


def _diss_case(cfg, st):
    o = oracle.context(cfg)
    z = np.zeros(cfg.shape)
    load_state(o, {"vx": z, "vy": z, "vz": z, "temp": z, "VOFp": np.ones(cfg.shape), **st})
    return o.dissipation()


def test_dissipation_closed_forms(lib):
    cfg = make_cfg(32, 24, ren=40.0, prandtl=0.7, xlen=6.0, ylen=5.0, zlen=4.0)
    g = _grid(cfg)
    z = np.zeros(cfg.shape)
    # uniform translation: no dissipation, k = |U|^2 / 2 (diss.f90:107-111)
    U = (0.3, -0.4, 0.2)
    r = _diss_case(cfg, {"vx": z + U[0], "vy": z + U[1], "vz": z + U[2], "temp": z + 0.7})
    assert r["diss"] == 0.0 and r["chi"] == 0.0
    assert abs(r["kenerg"] - 0.5 * sum(x * x for x in U)) < 1e-12 * r["kenerg"] and r["sumvof"] == z.size
    # shear u = A sin(k y): the central difference of diss.f90:67-69 gives
    # du/dy = A cos(k y) sin(k h) / h, so diss = (du/dy)^2 / Re averages to
    # A^2 sin^2(k h) / (2 h^2 Re); T = B sin(k x) gives chi the same form with Pe
    A, B, m = 0.8, 0.5, 3
    ky, kx = 2 * np.pi * m / cfg.ylen, 2 * np.pi * m / cfg.xlen
    hy, hx = cfg.ylen / cfg.n2m, cfg.xlen / cfg.n1m
    u = A * np.sin(ky * g["ym"])[None, :, None] + z
    T = B * np.sin(kx * g["xm"])[None, None, :] + z
    r = _diss_case(cfg, {"vx": u, "temp": T})
    pec = cfg.ren * cfg.prandtl
    assert abs(r["diss"] - A * A * np.sin(ky * hy) ** 2 / (2 * hy * hy * cfg.ren)) < 1e-12 * r["diss"]
    assert abs(r["chi"] - B * B * np.sin(kx * hx) ** 2 / (2 * hx * hx * pec)) < 1e-12 * r["chi"]
    assert abs(r["kenerg"] - A * A / 4.0) < 1e-12 * r["kenerg"]
    # the derived scales obey their definitions (Pope 2000, eqs. 6.56-6.65):
    # Re_lambda = sqrt(20/3 Re_L), eta = (nu^3/eps)^(1/4), L = k^(3/2)/eps
    nu = 1.0 / cfg.ren
    assert abs(r["re_lambda"] - np.sqrt(20.0 / 3.0 * r["re_l"])) < 1e-12 * r["re_lambda"]
    assert abs(r["eta"] - (nu ** 3 / r["diss"]) ** 0.25) < 1e-14 * r["eta"]
    assert abs(r["l_int"] - r["kenerg"] ** 1.5 / r["diss"]) < 1e-12 * r["l_int"]
    assert abs(r["urms"] - np.sqrt(2.0 / 3.0 * r["kenerg"])) < 1e-15
    assert abs(r["kmax_eta"] - np.pi * cfg.n1m / cfg.xlen * r["eta"]) < 1e-14 * r["kmax_eta"]
    # only fully fluid cells (VOFp == 1) are averaged (diss.f90:144-153)
    vof = np.ones(cfg.shape)
    vof[:, :, : cfg.n1m // 2] = 0.999
    o = oracle.context(cfg)
    load_state(o, {"vx": u, "vy": z, "vz": z, "temp": T, "VOFp": vof})
    half = o.dissipation()
    assert half["sumvof"] == z.size // 2
    assert abs(half["diss"] - r["diss"]) < 1e-12 * r["diss"]     # u depends on y only


# ------------------------------------------------------ sphereTagging (tagType 0)
def _sphere_case(n, centre=(0.03, -0.05, 0.02), rad=1.0):
    cfg = make_cfg(n, tagtype=0, imelt=0)
    o = oracle.context(cfg)
    with_body(o, 3, rad)
    b = o.body()
    for d in range(3):
        b.pos_cm[d] += centre[d]
    o.set_body(b)
    o.set_lag("xyzv", o.lag("xyzv") + np.array(centre))
    o.tagCells()
    return cfg, o, np.array(o.body().pos_cm[:])


def test_sphere_tagging_volume_and_staggering(lib):
    """The Kempe-Froehlich fractions of sphereTagging.f90 add up to the
    sphere's volume at second order, and the solid of each staggered field,
    weighted at that field's own points (VOFx at the u faces, ...), is centred
    on the sphere (a half-cell slip of the corner sets moves it by h/2)."""
    errs = []
    for n in (24, 48):
        cfg, o, pc = _sphere_case(n)
        h = cfg.xlen / n
        (xm, ym, zm), (xc, yc, zc) = inputs.grid_coords(cfg)
        vol = 4.0 / 3.0 * np.pi
        e = []
        for name, pts in (("VOFp", (xm, ym, zm)), ("VOFx", (xc, ym, zm)), ("VOFy", (xm, yc, zm)), ("VOFz", (xm, ym, zc))):
            s = 1.0 - o.download(name)
            assert s.min() >= -1e-15 and s.max() <= 1.0 + 1e-15
            e.append(abs(s.sum() * h ** 3 - vol) / vol)
            X = np.broadcast_to(pts[0][None, None, :], s.shape)
            Y = np.broadcast_to(pts[1][None, :, None], s.shape)
            Z = np.broadcast_to(pts[2][:, None, None], s.shape)
            cen = np.array([(s * X).sum(), (s * Y).sum(), (s * Z).sum()]) / s.sum()
            assert np.max(np.abs(cen - pc)) < 0.02 * h, (name, cen - pc)
        errs.append(max(e))
        # far outside untouched, deep inside solid
        r = np.sqrt((xm[None, None, :] - pc[0]) ** 2 + (ym[None, :, None] - pc[1]) ** 2 + (zm[:, None, None] - pc[2]) ** 2)
        vofp = o.download("VOFp")
        assert np.all(vofp[r > 1.0 + 2 * h] == 1.0) and np.all(vofp[r < 1.0 - 2 * h] == 0.0)
    assert errs[1] < errs[0] / 3.0 and errs[1] < 1e-2, errs
