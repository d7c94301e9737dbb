"""Host mesh state and ctypes bindings of the remeshing entry points.

`melt_remesh_coarsen` and `melt_remesh_smooth` are the library's forms of
remesh_coarsen (remesh_coarsen.f90:2-258) and remesh_smooth
(remesh_smooth.f90:3-347): they edit a triangulated surface held in host
arrays with the reference's layout (1-based connectivity, ghost flags kept in
their slots).  `TriMesh` holds those arrays, built from a `geometry.Mesh` the
way set_connectivity (partaux.f90:160-267) derives vert_of_face and
face_of_edge; the same class drives the oracle (prefix "orc_").
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import geometry

_P = C.c_void_p
_COARSEN = (C.c_int, [C.c_int32] * 3 + [_P] * 3 + [C.c_double] + [_P] * 12)
_SMOOTH = (C.c_int, [C.c_double, C.c_double] + [C.c_int32] * 3 + [_P] * 12)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


@dataclass
class TriMesh:
    xyz: np.ndarray              # (nv, 3)
    vert_of_face: np.ndarray     # (nf, 3) 1-based
    edge_of_face: np.ndarray     # (nf, 3) 1-based
    vert_of_edge: np.ndarray     # (ne, 2) 1-based
    face_of_edge: np.ndarray     # (ne, 2) 1-based
    ghost_v: np.ndarray          # (nv,) int32 0/1
    ghost_e: np.ndarray
    ghost_f: np.ndarray
    anchor_v: np.ndarray
    flagged_e: np.ndarray
    tri_nor: np.ndarray = field(default=None)
    e_lengths: np.ndarray = field(default=None)
    vmelt: Optional[np.ndarray] = None

    @classmethod
    def from_mesh(cls, m: geometry.Mesh, scale: float = 1.0) -> "TriMesh":
        nv, ne, nf = m.counts
        vof = m.faces().astype(np.int32)
        foe = np.zeros((ne, 2), dtype=np.int32)
        for f in range(nf):
            for e in m.edge_of_face[f] - 1:
                foe[e, 0 if foe[e, 0] == 0 else 1] = f + 1
        t = cls(np.ascontiguousarray(m.xyz * scale, dtype=np.float64), vof, m.edge_of_face.astype(np.int32).copy(),
                m.vert_of_edge.astype(np.int32).copy(), foe, np.zeros(nv, np.int32), np.zeros(ne, np.int32), np.zeros(nf, np.int32),
                np.ones(nv, np.int32), np.zeros(ne, np.int32))
        t.update_normals()
        t.update_lengths()
        return t

    @property
    def counts(self) -> tuple[int, int, int]:
        return self.xyz.shape[0], self.vert_of_edge.shape[0], self.vert_of_face.shape[0]

    def copy(self) -> "TriMesh":
        return TriMesh(*(None if v is None else v.copy() for v in (self.xyz, self.vert_of_face, self.edge_of_face, self.vert_of_edge,
                                                                     self.face_of_edge, self.ghost_v, self.ghost_e, self.ghost_f,
                                                                     self.anchor_v, self.flagged_e, self.tri_nor, self.e_lengths,
                                                                     self.vmelt)))

    def update_normals(self) -> None:
        """update_tri_normal (triaux.f90:183-232) on the live faces."""
        f = self.vert_of_face - 1
        x = self.xyz
        n = np.cross(x[f[:, 1]] - x[f[:, 0]], x[f[:, 2]] - x[f[:, 0]])
        s = np.sqrt(n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1] + n[:, 2] * n[:, 2])
        with np.errstate(invalid="ignore", divide="ignore"):     # ghost faces may be degenerate
            nn = n / s[:, None]
        if self.tri_nor is None:
            self.tri_nor = np.zeros_like(nn)
        live = self.ghost_f == 0
        self.tri_nor[live] = nn[live]

    def update_lengths(self) -> None:
        """calculate_eLengths (triaux.f90:277-305) on the live edges."""
        d = self.xyz[self.vert_of_edge[:, 0] - 1] - self.xyz[self.vert_of_edge[:, 1] - 1]
        el = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
        if self.e_lengths is None:
            self.e_lengths = np.zeros_like(el)
        live = self.ghost_e == 0
        self.e_lengths[live] = el[live]

    def live_counts(self) -> tuple[int, int, int]:
        return int((self.ghost_v == 0).sum()), int((self.ghost_e == 0).sum()), int((self.ghost_f == 0).sum())

    def volume(self) -> float:
        """calculate_volume (triaux.f90:53-79), serial over the live faces."""
        x = self.xyz
        V = 0.0
        for i in np.nonzero(self.ghost_f == 0)[0]:
            v1, v2, v3 = self.vert_of_face[i] - 1
            V = V + (x[v1, 0] * (x[v2, 1] * x[v3, 2] - x[v2, 2] * x[v3, 1]) + x[v2, 0] * (x[v3, 1] * x[v1, 2] - x[v3, 2] * x[v1, 1]) +
                     x[v3, 0] * (x[v1, 1] * x[v2, 2] - x[v1, 2] * x[v2, 1]))
        return V / 6.0


def _bind(lib: C.CDLL, prefix: str, name: str, sig):
    fn = getattr(lib, prefix + name)
    fn.restype, fn.argtypes = sig
    return fn


def coarsen(t: TriMesh, e_thresh: float, lib: C.CDLL, prefix: str = "melt_", rm_flag: bool = True) -> dict:
    """remesh_coarsen on t in place; returns {rc, ecol, vdel, rm_flag}."""
    nv, ne, nf = t.counts
    fl = C.c_int32(1 if rm_flag else 0)
    ec, vd = C.c_int32(0), C.c_int32(0)
    vm = _ptr(t.vmelt) if t.vmelt is not None else None
    rc = _bind(lib, prefix, "remesh_coarsen", _COARSEN)(
        nf, ne, nv, _ptr(t.xyz), _ptr(t.tri_nor), _ptr(t.e_lengths), float(e_thresh), _ptr(t.vert_of_face), _ptr(t.edge_of_face),
        _ptr(t.vert_of_edge), _ptr(t.face_of_edge), _ptr(t.ghost_f), _ptr(t.ghost_e), _ptr(t.ghost_v), C.addressof(fl), _ptr(t.anchor_v),
        _ptr(t.flagged_e), vm, C.addressof(ec), C.addressof(vd))
    return {"rc": rc, "ecol": ec.value, "vdel": vd.value, "rm_flag": fl.value}


def smooth(t: TriMesh, vol_old: float, target_dv: float, lib: C.CDLL, prefix: str = "melt_") -> dict:
    """remesh_smooth on t in place; returns {rc, n_erel, drift, refresh}."""
    nv, ne, nf = t.counts
    ner, cr = C.c_int32(0), C.c_int32(0)
    dr = C.c_double(0.0)
    rc = _bind(lib, prefix, "remesh_smooth", _SMOOTH)(
        float(vol_old), float(target_dv), nv, ne, nf, _ptr(t.xyz), _ptr(t.ghost_v), _ptr(t.ghost_e), _ptr(t.ghost_f), _ptr(t.flagged_e),
        _ptr(t.vert_of_edge), _ptr(t.vert_of_face), _ptr(t.face_of_edge), _ptr(t.edge_of_face), C.addressof(ner), C.addressof(dr),
        C.addressof(cr))
    return {"rc": rc, "n_erel": ner.value, "drift": dr.value, "refresh": cr.value}
