"""Python host mirror of the melt_IBM substep interface over the C ABI of
libmeltgpu.so (include/melt_ibm.h).

`MeltContext` holds what the reference keeps in module globals (param.f90) and
exposes the reference's subroutines by name: `tschem`, `tagCells`, `hdnl1`,
`hdnl2`, `hdnl3`, `hdnlte`, `invtr1`..`invtrte`, `solxi`, `solxj`, `solxk`,
`divg`, `phcalc`, `updvp`, `prcalc`, `findCentroidIndices`, `mlsWeight`,
`mlsForce`, `velforce`, `tempforce`, `particle`.  A non-zero status from the
library raises `MeltError` carrying the library's message, as the reference
stops with a message on its fatal paths.

The binding is written against a ctypes library and a symbol prefix, so the same
class drives any library exporting this ABI.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field, fields
from pathlib import Path
from typing import Optional

import numpy as np

PKG = Path(__file__).resolve().parent
LIBPATH = PKG / "libmeltgpu.so"

# ---------------------------------------------------------------- ids
FIELDS = ["vx", "vy", "vz", "temp", "pr", "ru1", "ru2", "ru3", "rut", "forcx", "forcy", "forcz",
          "VOFp", "VOFx", "VOFy", "VOFz", "sdf", "dq", "dph", "qcap", "htemp", "rhs",
          "for_xc", "for_yc", "for_zc", "for_temp"]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}

LAG = ["xyzv", "vert_nor", "Avert", "qw_oVert", "qw_iVert", "vmelt", "vmelt_m1", "dxyzv_s",
       "tri_bar", "tri_nor", "sur", "vel_tri", "cfac", "skewness", "qw_o", "qw_i",
       "ptxAB_q1", "ptxAB_q2", "ptxAB_q3", "ptxAB_temp", "dxyz_s", "press_n_tri", "tau_n", "eLengths"]
LAG_ID = {n: i for i, n in enumerate(LAG)}
_LAG_SHAPE = {  # (per-entity width, entity) entity 0 vertex, 1 face, 2 edge
    "xyzv": (3, 0), "vert_nor": (3, 0), "Avert": (1, 0), "qw_oVert": (1, 0), "qw_iVert": (1, 0),
    "vmelt": (3, 0), "vmelt_m1": (3, 0), "dxyzv_s": (3, 0),
    "tri_bar": (3, 1), "tri_nor": (3, 1), "sur": (1, 1), "vel_tri": (3, 1), "cfac": (1, 1),
    "skewness": (1, 1), "qw_o": (1, 1), "qw_i": (1, 1), "ptxAB_q1": (27, 1), "ptxAB_q2": (27, 1),
    "ptxAB_q3": (27, 1), "ptxAB_temp": (27, 1), "dxyz_s": (3, 1), "press_n_tri": (3, 1), "tau_n": (3, 1),
    "eLengths": (1, 2)}

STATUS = {0: "ok", 1: "bad argument", 2: "HIP error", 3: "singular MLS moment matrix", 4: "body volume below threshold",
          5: "pencil buffer overflow", 6: "remeshing failed", 7: "ray recursion", 8: "no geometry",
          9: "window overflow"}


class MeltError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}:{STATUS.get(code, '?')}] {msg}")
        self.code = code


class _Config(C.Structure):
    _fields_ = [("n1m", C.c_int32), ("n2m", C.c_int32), ("n3m", C.c_int32), ("precision", C.c_int32),
                ("xlen", C.c_double), ("ylen", C.c_double), ("zlen", C.c_double),
                ("ren", C.c_double), ("prandtl", C.c_double), ("betagz", C.c_double),
                ("tmelt", C.c_double), ("tliq", C.c_double), ("tsol", C.c_double), ("latheat", C.c_double),
                ("cpliquid", C.c_double), ("dens_ratio", C.c_double),
                ("imlsfor", C.c_int32), ("imlsstr", C.c_int32), ("imelt", C.c_int32), ("iremesh", C.c_int32),
                ("tagtype", C.c_int32), ("forcing", C.c_int32),
                ("perc_ethresh", C.c_double), ("v_on_ve_perc", C.c_double)]


class Body(C.Structure):
    """melt_body: rigid-body and Lagrangian scalars (mls_param)."""
    _fields_ = [("pos_cm", C.c_double * 3), ("vel_cm", C.c_double * 3), ("omega_c", C.c_double * 3), ("quat", C.c_double * 4),
                ("inert_tensor", C.c_double * 9), ("volume", C.c_double), ("surface", C.c_double),
                ("int_prn_dA", C.c_double * 3), ("int_tau_dA", C.c_double * 3), ("int_r_x_prn_dA", C.c_double * 3),
                ("int_r_x_tau_dA", C.c_double * 3), ("int_prn_dA_m1", C.c_double * 3), ("int_tau_dA_m1", C.c_double * 3),
                ("int_r_x_prn_dA_m1", C.c_double * 3), ("int_r_x_tau_dA_m1", C.c_double * 3),
                ("fp", C.c_double * 3), ("ftau", C.c_double * 3), ("torq_p", C.c_double * 3), ("torq_tau", C.c_double * 3),
                ("h_eulerian", C.c_double), ("celvol", C.c_double), ("e_thresh", C.c_double), ("v_thresh", C.c_double),
                ("rm_flag", C.c_int32), ("initial_tag", C.c_int32)]

    def as_dict(self) -> dict:
        out = {}
        for name, typ in self._fields_:
            v = getattr(self, name)
            out[name] = np.array(v[:]) if hasattr(v, "__len__") else v
        return out


class _Layout(C.Structure):
    _fields_ = [("ld1", C.c_int32), ("ld2", C.c_int32), ("kghost", C.c_int32), ("pad", C.c_int32)]


@dataclass
class MeltConfig:
    """Input-deck values (bou.in / part.in / HITForcing.in, papero.f90:31-71)."""
    n1m: int = 64
    n2m: int = 64
    n3m: int = 64
    xlen: float = 2.0 * np.pi
    ylen: float = 2.0 * np.pi
    zlen: float = 2.0 * np.pi
    ren: float = 500.0
    prandtl: float = 1.0
    betagz: float = 0.0
    tmelt: float = 0.0
    tliq: float = 1.0
    tsol: float = 0.0
    latheat: float = 1.0
    cpliquid: float = 1.0
    dens_ratio: float = 1.0
    imlsfor: int = 1
    imlsstr: int = 1
    imelt: int = 1
    iremesh: int = 0
    tagtype: int = 1
    forcing: int = 0
    perc_ethresh: float = 0.7
    v_on_ve_perc: float = 27.0
    precision: int = 0

    def to_c(self) -> _Config:
        c = _Config()
        for f in fields(self):
            setattr(c, f.name, getattr(self, f.name))
        return c

    @property
    def shape(self) -> tuple[int, int, int]:
        """numpy shape (k, j, i) of a field: i fastest as in the reference."""
        return (self.n3m, self.n2m, self.n1m)


REMESH_KEYS = ("events", "collapses", "deletions", "relaxations", "drift", "residual", "refreshes", "total_collapses")

_lib_cache: dict[str, C.CDLL] = {}


def load_library(path: Optional[os.PathLike] = None) -> C.CDLL:
    """Load libmeltgpu.so.  Raises if it has not been built (no CPU fallback)."""
    p = Path(path) if path else LIBPATH
    key = str(p)
    if key in _lib_cache:
        return _lib_cache[key]
    if not p.exists():
        raise FileNotFoundError(f"{p} is missing: build it with `python -m melt_ibm_amd.build` (hipcc, gfx950)")
    lib = C.CDLL(str(p), mode=C.RTLD_GLOBAL)
    _lib_cache[key] = lib
    return lib


def device_available(lib: Optional[C.CDLL] = None) -> bool:
    lib = lib or load_library()
    return bool(lib.melt_device_ok())


_P = C.c_void_p
_SIGS = {
    "ctx_create": (C.c_int, [C.POINTER(_Config), C.POINTER(_P)]),
    "ctx_destroy": (None, [_P]),
    "last_error": (C.c_char_p, [_P]),
    "sync": (C.c_int, [_P]),
    "upload_field": (C.c_int, [_P, C.c_int, _P, C.POINTER(_Layout)]),
    "download_field": (C.c_int, [_P, C.c_int, _P, C.POINTER(_Layout)]),
    "fill_field": (C.c_int, [_P, C.c_int, C.c_double]),
    "set_geometry": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, _P, _P, _P, C.c_double]),
    "get_body": (C.c_int, [_P, C.POINTER(Body)]),
    "set_body": (C.c_int, [_P, C.POINTER(Body)]),
    "get_lag": (C.c_int, [_P, C.c_int, _P]),
    "set_lag": (C.c_int, [_P, C.c_int, _P]),
    "get_pind": (C.c_int, [_P, _P]),
    "get_vert_of_face": (C.c_int, [_P, _P]),
    "set_dt": (C.c_int, [_P, C.c_double]),
    "set_stage": (C.c_int, [_P, C.c_int32]),
    "tschem": (C.c_int, [_P]),
    "substep": (C.c_int, [_P, C.c_int32]),
    "tagcells": (C.c_int, [_P]),
    "interp_vof": (C.c_int, [_P]),
    "hdnl1": (C.c_int, [_P]), "hdnl2": (C.c_int, [_P]), "hdnl3": (C.c_int, [_P]), "hdnlte": (C.c_int, [_P]),
    "invtr1": (C.c_int, [_P]), "invtr2": (C.c_int, [_P]), "invtr3": (C.c_int, [_P]), "invtrte": (C.c_int, [_P]),
    "solxi": (C.c_int, [_P, C.c_double]), "solxj": (C.c_int, [_P, C.c_double]), "solxk": (C.c_int, [_P, C.c_int, C.c_double]),
    "solxk_array": (C.c_int, [_P, _P, C.POINTER(_Layout), C.c_double]),
    "divg": (C.c_int, [_P]), "phcalc": (C.c_int, [_P]), "updvp": (C.c_int, [_P]), "prcalc": (C.c_int, [_P]),
    "find_centroid_indices": (C.c_int, [_P]), "mlsweight": (C.c_int, [_P]), "mlsforce": (C.c_int, [_P]),
    "velforce": (C.c_int, [_P]), "tempforce": (C.c_int, [_P]), "particle": (C.c_int, [_P]),
    "cfl": (C.c_int, [_P, C.POINTER(C.c_double)]), "divgck": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "inqpr": (C.c_int, [_P]),
    "get_dt": (C.c_int, [_P, C.c_double, C.POINTER(C.c_double)]),
    "hit_init": (C.c_int, [_P, C.c_double, C.c_double, C.c_double, C.c_uint64]),
    "hit_force": (C.c_int, [_P]),
    "hit_state": (C.c_int, [_P, _P, C.POINTER(C.c_uint64), C.c_int32]),
    "hit_nmodes": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "abc_force": (C.c_int, [_P, C.c_double]),
    "snapshot_size": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "snapshot_save": (C.c_int, [_P, _P, C.c_uint64]),
    "snapshot_load": (C.c_int, [_P, _P, C.c_uint64]),
    "dissipation": (C.c_int, [_P, _P]),
    "timing_enable": (C.c_int, [_P, C.c_int32]), "timing_select": (C.c_int, [_P, C.c_uint32]), "timing_reset": (C.c_int, [_P]),
    "timing_count": (C.c_int, []), "timing_name": (C.c_char_p, [C.c_int32]),
    "timing_read": (C.c_int, [_P, _P, _P]),
    "set_refinement": (C.c_int, [_P, C.c_int32, C.c_double]),
    "get_refinement_stats": (C.c_int, [_P, _P, _P, _P]),
    "graph_enable": (C.c_int, [_P, C.c_int32]),
    "set_tuning": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "stream": (C.c_void_p, [_P]),
    "poll": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "upload_field_device": (C.c_int, [_P, C.c_int, C.c_void_p]),
    "get_topology": (C.c_int, [_P] * 8),
    "get_remesh_stats": (C.c_int, [_P, _P]),
}
OPTIONAL = {"timing_enable", "timing_select", "timing_reset", "timing_count", "timing_name", "timing_read", "device_ok", "set_refinement",
            "get_refinement_stats", "graph_enable", "stream",
            "set_tuning", "poll", "upload_field_device", "snapshot_size", "snapshot_save", "snapshot_load"}


def bind(lib: C.CDLL, prefix: str) -> dict:
    fns = {}
    for name, (res, args) in _SIGS.items():
        sym = prefix + name
        if not hasattr(lib, sym):
            if name in OPTIONAL:
                continue
            raise AttributeError(f"library lacks {sym}")
        fn = getattr(lib, sym)
        fn.restype = res
        fn.argtypes = args
        fns[name] = fn
    return fns


class MeltContext:
    """One simulation's device-resident state plus the reference's subroutines."""

    def __init__(self, cfg: MeltConfig, lib: Optional[C.CDLL] = None, prefix: str = "melt_"):
        self.cfg = cfg
        self.lib = lib or load_library()
        self._f = bind(self.lib, prefix)
        self._ctx = _P()
        self._c = cfg.to_c()
        rc = self._f["ctx_create"](C.byref(self._c), C.byref(self._ctx))
        if rc:
            raise MeltError(rc, f"context creation failed for {cfg}")
        self.nv = self.ne = self.nf = 0
        self.dt = None

    # ------------------------------------------------------------ plumbing
    def close(self) -> None:
        if self._ctx:
            self._f["ctx_destroy"](self._ctx)
            self._ctx = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _chk(self, rc: int) -> None:
        if rc:
            msg = self._f["last_error"](self._ctx)
            raise MeltError(rc, msg.decode() if msg else "")

    def sync(self) -> None:
        self._chk(self._f["sync"](self._ctx))

    def poll(self) -> bool:
        """True when all queued work has finished (never blocks)."""
        d = C.c_int32(0)
        self._chk(self._f["poll"](self._ctx, C.byref(d)))
        return bool(d.value)

    def upload_device(self, name: str, ptr: int) -> None:
        """Upload from device memory at ptr (N1*N2*N3 doubles, i fastest, complete)."""
        self._chk(self._f["upload_field_device"](self._ctx, FIELD_ID[name], C.c_void_p(ptr)))

    # ------------------------------------------------------------ fields
    def upload(self, name: str, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != self.cfg.shape:
            raise ValueError(f"{name}: expected shape {self.cfg.shape}, got {a.shape}")
        lay = _Layout(self.cfg.n1m, self.cfg.n2m, 0, 0)
        self._chk(self._f["upload_field"](self._ctx, FIELD_ID[name], a.ctypes.data, C.byref(lay)))

    def download(self, name: str) -> np.ndarray:
        a = np.empty(self.cfg.shape, dtype=np.float64)
        lay = _Layout(self.cfg.n1m, self.cfg.n2m, 0, 0)
        self._chk(self._f["download_field"](self._ctx, FIELD_ID[name], a.ctypes.data, C.byref(lay)))
        return a

    def download_with_ghosts(self, name: str) -> np.ndarray:
        """(n3m+2, n2m, n1m) with the ghost planes update_both_ghosts would hold."""
        a = np.empty((self.cfg.n3m + 2, self.cfg.n2m, self.cfg.n1m), dtype=np.float64)
        lay = _Layout(self.cfg.n1m, self.cfg.n2m, 1, 0)
        self._chk(self._f["download_field"](self._ctx, FIELD_ID[name], a.ctypes.data, C.byref(lay)))
        return a

    def fill(self, name: str, value: float) -> None:
        self._chk(self._f["fill_field"](self._ctx, FIELD_ID[name], float(value)))

    # ------------------------------------------------------------ body
    def set_geometry(self, xyz0: np.ndarray, vert_of_edge: np.ndarray, edge_of_face: np.ndarray, rad_p: float) -> None:
        x = np.ascontiguousarray(xyz0, dtype=np.float64)
        voe = np.ascontiguousarray(vert_of_edge, dtype=np.int32)
        eof = np.ascontiguousarray(edge_of_face, dtype=np.int32)
        nv, ne, nf = x.shape[0], voe.shape[0], eof.shape[0]
        self._chk(self._f["set_geometry"](self._ctx, nv, ne, nf, x.ctypes.data, voe.ctypes.data, eof.ctypes.data, float(rad_p)))
        self.nv, self.ne, self.nf = nv, ne, nf

    def body(self) -> Body:
        b = Body()
        self._chk(self._f["get_body"](self._ctx, C.byref(b)))
        return b

    def set_body(self, b: Body) -> None:
        self._chk(self._f["set_body"](self._ctx, C.byref(b)))

    def lag(self, name: str) -> np.ndarray:
        w, ent = _LAG_SHAPE[name]
        n = (self.nv, self.nf, self.ne)[ent]
        a = np.empty((n, w) if w > 1 else (n,), dtype=np.float64)
        self._chk(self._f["get_lag"](self._ctx, LAG_ID[name], a.ctypes.data))
        return a

    def set_lag(self, name: str, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self._chk(self._f["set_lag"](self._ctx, LAG_ID[name], a.ctypes.data))

    def pind(self) -> np.ndarray:
        a = np.empty((self.nf, 6), dtype=np.int32)
        self._chk(self._f["get_pind"](self._ctx, a.ctypes.data))
        return a

    def topology(self) -> dict:
        """Connectivity (1-based) and ghost flags; remeshing edits them in place."""
        nv, ne, nf = self.nv, self.ne, self.nf
        out = {"vert_of_face": np.empty((nf, 3), np.int32), "edge_of_face": np.empty((nf, 3), np.int32),
               "vert_of_edge": np.empty((ne, 2), np.int32), "face_of_edge": np.empty((ne, 2), np.int32),
               "ghost_v": np.empty(nv, np.int32), "ghost_e": np.empty(ne, np.int32), "ghost_f": np.empty(nf, np.int32)}
        self._chk(self._f["get_topology"](self._ctx, *(a.ctypes.data for a in out.values())))
        return out

    def remesh_stats(self) -> dict:
        """Remeshing counters: events so far, and of the last event the edge collapses, vertex
        deletions, edge relaxations, largest vertex drift, volume residual and refreshes."""
        a = np.zeros(8)
        self._chk(self._f["get_remesh_stats"](self._ctx, a.ctypes.data))
        return dict(zip(REMESH_KEYS, a.tolist()))

    def vert_of_face(self) -> np.ndarray:
        a = np.empty((self.nf, 3), dtype=np.int32)
        self._chk(self._f["get_vert_of_face"](self._ctx, a.ctypes.data))
        return a

    # ------------------------------------------------------------ time
    def set_dt(self, dt: float) -> None:
        self.dt = float(dt)
        self._chk(self._f["set_dt"](self._ctx, float(dt)))

    def set_stage(self, ns: int) -> None:
        self._chk(self._f["set_stage"](self._ctx, int(ns)))

    # ------------------------------------------------- reference routines
    def tschem(self) -> None:
        self._chk(self._f["tschem"](self._ctx))

    def substep(self, ns: int) -> None:
        self._chk(self._f["substep"](self._ctx, int(ns)))

    def tagCells(self) -> None:
        self._chk(self._f["tagcells"](self._ctx))

    def interp_vof(self) -> None:
        self._chk(self._f["interp_vof"](self._ctx))

    def hdnl1(self): self._chk(self._f["hdnl1"](self._ctx))
    def hdnl2(self): self._chk(self._f["hdnl2"](self._ctx))
    def hdnl3(self): self._chk(self._f["hdnl3"](self._ctx))
    def hdnlte(self): self._chk(self._f["hdnlte"](self._ctx))
    def invtr1(self): self._chk(self._f["invtr1"](self._ctx))
    def invtr2(self): self._chk(self._f["invtr2"](self._ctx))
    def invtr3(self): self._chk(self._f["invtr3"](self._ctx))
    def invtrte(self): self._chk(self._f["invtrte"](self._ctx))

    def solxi(self, betadx: float): self._chk(self._f["solxi"](self._ctx, float(betadx)))
    def solxj(self, betadx: float): self._chk(self._f["solxj"](self._ctx, float(betadx)))
    def solxk(self, q: str, betadx: float): self._chk(self._f["solxk"](self._ctx, FIELD_ID[q], float(betadx)))

    def solxk_array(self, q: np.ndarray, betadx: float, kghost: int = 0) -> None:
        """solxk(q, betadx) on a host array q[k, j, i] (Fortran q(i, j, k)), updated in place."""
        if q.dtype != np.float64 or not q.flags.c_contiguous:
            raise ValueError("q must be a C-contiguous float64 array")
        lay = _Layout(q.shape[2], q.shape[1], kghost, 0)
        self._chk(self._f["solxk_array"](self._ctx, q.ctypes.data, C.byref(lay), float(betadx)))

    def divg(self): self._chk(self._f["divg"](self._ctx))
    def phcalc(self): self._chk(self._f["phcalc"](self._ctx))
    def updvp(self): self._chk(self._f["updvp"](self._ctx))
    def prcalc(self): self._chk(self._f["prcalc"](self._ctx))
    def findCentroidIndices(self): self._chk(self._f["find_centroid_indices"](self._ctx))
    def mlsWeight(self): self._chk(self._f["mlsweight"](self._ctx))
    def mlsForce(self): self._chk(self._f["mlsforce"](self._ctx))
    def velforce(self): self._chk(self._f["velforce"](self._ctx))
    def tempforce(self): self._chk(self._f["tempforce"](self._ctx))
    def particle(self): self._chk(self._f["particle"](self._ctx))
    def inqpr(self): self._chk(self._f["inqpr"](self._ctx))

    def cfl(self) -> float:
        v = C.c_double()
        self._chk(self._f["cfl"](self._ctx, C.byref(v)))
        return v.value

    def get_dt(self, cflfix: float) -> float:
        """get_dt (cfl.f90:33-77): the dt that brings the current CFL to cflfix."""
        d = C.c_double(0.0)
        self._chk(self._f["get_dt"](self._ctx, float(cflfix), C.byref(d)))
        return d.value

    # ------------------------------------------------------------ HIT forcing
    def hit_init(self, tl: float, epsstar: float, kf_on_kmin: float, seed: int = 0) -> None:
        """InitRandomForce (hit.f90:128-232) with HITForcing.in's TL, EPSSTAR, KF/KMIN."""
        self._chk(self._f["hit_init"](self._ctx, float(tl), float(epsstar), float(kf_on_kmin), int(seed)))

    def hit_force(self) -> None:
        """CalcHITRandomForce (hit.f90:1-126) at the current dt: forcx/forcy/forcz."""
        self._chk(self._f["hit_force"](self._ctx))

    def hit_nmodes(self) -> tuple[int, int]:
        nm, na = C.c_int32(0), C.c_int32(0)
        self._chk(self._f["hit_nmodes"](self._ctx, C.byref(nm), C.byref(na)))
        return nm.value, na.value

    def hit_state(self) -> tuple[np.ndarray, int]:
        """bhat as complex [n, m, l, dir] (Fortran bhat(dir, l, m, n)) and the steps taken."""
        nm, _ = self.hit_nmodes()
        a = np.zeros(2 * 3 * nm ** 3)
        st = C.c_uint64(0)
        self._chk(self._f["hit_state"](self._ctx, a.ctypes.data, C.byref(st), 0))
        return (a[0::2] + 1j * a[1::2]).reshape(nm, nm, nm, 3), int(st.value)

    def set_hit_state(self, bhat: np.ndarray, steps: int) -> None:
        b = np.ascontiguousarray(bhat, dtype=np.complex128).reshape(-1)
        a = np.empty(2 * b.size)
        a[0::2], a[1::2] = b.real, b.imag
        st = C.c_uint64(int(steps))
        self._chk(self._f["hit_state"](self._ctx, a.ctypes.data, C.byref(st), 1))

    def abc_force(self, c_hit: float) -> None:
        """CalcABC_HITForce (hit.f90:234-265)."""
        self._chk(self._f["abc_force"](self._ctx, float(c_hit)))

    # ------------------------------------------------------- snapshot / resume
    def snapshot(self) -> bytearray:
        """The context's state between steps as one blob (melt_snapshot_save)."""
        n = C.c_uint64(0)
        self._chk(self._f["snapshot_size"](self._ctx, C.byref(n)))
        buf = bytearray(n.value)
        self._chk(self._f["snapshot_save"](self._ctx, (C.c_char * len(buf)).from_buffer(buf), n.value))
        return buf

    def restore(self, blob) -> None:
        """Resume from a snapshot taken on a context of the same configuration and mesh."""
        b = bytearray(blob)
        self._chk(self._f["snapshot_load"](self._ctx, (C.c_char * len(b)).from_buffer(b), len(b)))

    def save(self, path) -> None:
        with open(path, "wb") as f:
            f.write(self.snapshot())

    def load(self, path) -> None:
        with open(path, "rb") as f:
            self.restore(f.read())

    DISS_KEYS = ("diss", "kenerg", "urms", "eta", "lambda_t", "re_l", "re_lambda", "l_int", "chi", "kmax_eta", "sumvof")

    def dissipation(self) -> dict:
        """CalcDissipation (diss.f90:11-201): fluid averages and turbulence statistics."""
        a = np.zeros(11)
        self._chk(self._f["dissipation"](self._ctx, a.ctypes.data))
        return dict(zip(self.DISS_KEYS, (float(x) for x in a)))

    def divgck(self) -> float:
        v = C.c_double()
        self._chk(self._f["divgck"](self._ctx, C.byref(v)))
        return v.value

    # ------------------------------------------------------------ timing
    def timing(self, on: bool = True, slots: tuple[str, ...] | None = None) -> None:
        """HIP-event timing of every kernel slot, or of the named slots only
        (each timed slot puts two event records on the stream)."""
        if not on or slots is None:
            self._chk(self._f["timing_enable"](self._ctx, 1 if on else 0))
            return
        names = [self._f["timing_name"](i).decode() for i in range(self._f["timing_count"]())]
        mask = 0
        for s in slots:
            mask |= 1 << names.index(s)
        self._chk(self._f["timing_select"](self._ctx, mask))

    def timing_reset(self) -> None:
        self._chk(self._f["timing_reset"](self._ctx))

    # ------------------------------------------------------ precision form
    def set_refinement(self, max_iter: int = 1, rtol: float = 1e-4) -> None:
        """Gated FP64-residual refinement of the precision form's pressure."""
        self._chk(self._f["set_refinement"](self._ctx, int(max_iter), float(rtol)))

    def refinement_stats(self) -> tuple[int, int, float]:
        """(checks, refinements applied, last residual ratio after/before)."""
        a, b, r = C.c_int64(), C.c_int64(), C.c_double()
        self._chk(self._f["get_refinement_stats"](self._ctx, C.byref(a), C.byref(b), C.byref(r)))
        return a.value, b.value, r.value

    def graph(self, on: bool = True) -> None:
        """Replay tschem as one captured HIP graph."""
        self._chk(self._f["graph_enable"](self._ctx, 1 if on else 0))

    TUNE_KEYS = {"xf_w": 0, "xf_nt": 1, "solve_w": 2, "solve_nt": 3, "xf_z": 4, "yz_jb": 5, "streams": 6, "coef": 7, "pbuf": 8, "xf_one": 9}

    def set_tuning(self, **kw) -> None:
        """Launch-shape overrides (xf_w, xf_nt, solve_w, solve_nt, xf_z, yz_jb, streams, coef) and the ray-tagging
        pencil capacity (pbuf, BUFF_SIZE of pencil_rayTag.f90:84); 0 = default."""
        for k, v in kw.items():
            self._chk(self._f["set_tuning"](self._ctx, self.TUNE_KEYS[k], int(v)))

    def timing_read(self) -> dict:
        n = self._f["timing_count"]()
        ms = np.zeros(n, dtype=np.float64)
        cnt = np.zeros(n, dtype=np.int64)
        self._chk(self._f["timing_read"](self._ctx, ms.ctypes.data, cnt.ctypes.data))
        return {self._f["timing_name"](i).decode(): (float(ms[i]), int(cnt[i])) for i in range(n)}
