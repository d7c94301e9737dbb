"""Triangulated-surface input for the immersed body.

- `read_gts` reads the GTS surface format the reference imports
  (set_particle_array_sizes / set_connectivity, partaux.f90:139-267): a header
  line "nv ne nf", nv vertex lines, ne edge lines (1-based vertex pairs), nf face
  lines (1-based edge triples).
- `icosphere` builds a geodesic sphere (each icosahedron face split into
  freq^2 triangles, vertices projected onto the unit sphere) in the same form:
  freq = 2^L gives the level-L subdivided icosahedron (L = 3: 642 vertices,
  1280 faces; L = 5: 10242 vertices, 20480 faces).  Faces are listed with their
  edges in counter-clockwise order seen from outside, so the normal
  (v2 - v1) x (v3 - v1) that the connectivity import derives points outwards.
- `spheroid` scales an icosphere to semi-axes (a, a, c).
"""
from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np


@dataclass
class Mesh:
    xyz: np.ndarray          # (nv, 3) float64
    vert_of_edge: np.ndarray  # (ne, 2) int32, 1-based
    edge_of_face: np.ndarray  # (nf, 3) int32, 1-based

    @property
    def counts(self) -> tuple[int, int, int]:
        return self.xyz.shape[0], self.vert_of_edge.shape[0], self.edge_of_face.shape[0]

    def faces(self) -> np.ndarray:
        """(nf, 3) 1-based vertex triples derived exactly as set_connectivity does."""
        voe = self.vert_of_edge
        out = np.empty_like(self.edge_of_face)
        for i, (e1, e2, _) in enumerate(self.edge_of_face - 1):
            a1, b1 = voe[e1]
            a2, b2 = voe[e2]
            if b1 == a2:
                out[i] = (a1, b1, b2)
            elif b1 == b2:
                out[i] = (a1, b1, a2)
            elif a1 == a2:
                out[i] = (b1, a1, b2)
            else:
                out[i] = (b1, a1, a2)
        return out


def read_gts(path: str | Path) -> Mesh:
    with open(path) as fh:
        tok = fh.readline().split()
        nv, ne, nf = int(tok[0]), int(tok[1]), int(tok[2])
        xyz = np.loadtxt(fh, max_rows=nv, dtype=np.float64).reshape(nv, 3)
        voe = np.loadtxt(fh, max_rows=ne, dtype=np.int64).reshape(ne, -1)[:, :2]
        eof = np.loadtxt(fh, max_rows=nf, dtype=np.int64).reshape(nf, -1)[:, :3]
    return Mesh(xyz, voe.astype(np.int32), eof.astype(np.int32))


def write_gts(mesh: Mesh, path: str | Path) -> None:
    nv, ne, nf = mesh.counts
    with open(path, "w") as fh:
        fh.write(f"{nv} {ne} {nf}\n")
        for x in mesh.xyz:
            fh.write(f"{x[0]:.17g} {x[1]:.17g} {x[2]:.17g}\n")
        for e in mesh.vert_of_edge:
            fh.write(f"{e[0]} {e[1]}\n")
        for f in mesh.edge_of_face:
            fh.write(f"{f[0]} {f[1]} {f[2]}\n")


def _icosahedron() -> tuple[np.ndarray, np.ndarray]:
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def icosphere(freq: int = 8) -> Mesh:
    """Geodesic sphere of unit radius with 20*freq^2 faces."""
    if freq < 1:
        raise ValueError("freq must be >= 1")
    v0, f0 = _icosahedron()
    verts: list[np.ndarray] = []
    index: dict[tuple, int] = {}

    def vid(key: tuple, pos: np.ndarray) -> int:
        if key not in index:
            index[key] = len(verts)
            verts.append(pos / np.linalg.norm(pos))
        return index[key]

    faces = []
    for a, b, c in f0:
        # barycentric lattice point (i, j) -> a + (b-a) i/n + (c-a) j/n, keyed so
        # points shared between icosahedron faces coincide
        def key(i, j):
            w = {a: freq - i - j, b: i, c: j}
            return tuple(sorted((k, v) for k, v in w.items() if v))

        grid = {}
        for i in range(freq + 1):
            for j in range(freq + 1 - i):
                p = (v0[a] * (freq - i - j) + v0[b] * i + v0[c] * j) / freq
                grid[(i, j)] = vid(key(i, j), p)
        for i in range(freq):
            for j in range(freq - i):
                faces.append((grid[(i, j)], grid[(i + 1, j)], grid[(i, j + 1)]))
                if i + j < freq - 1:
                    faces.append((grid[(i + 1, j)], grid[(i + 1, j + 1)], grid[(i, j + 1)]))
    xyz = np.array(verts)
    faces = np.array(faces, dtype=np.int64)
    # orient every face outwards
    n = np.cross(xyz[faces[:, 1]] - xyz[faces[:, 0]], xyz[faces[:, 2]] - xyz[faces[:, 0]])
    flip = np.einsum("ij,ij->i", n, xyz[faces].sum(axis=1)) < 0
    faces[flip] = faces[flip][:, [0, 2, 1]]
    return _mesh_from_faces(xyz, faces)


def _mesh_from_faces(xyz: np.ndarray, faces: np.ndarray) -> Mesh:
    edges: dict[tuple, int] = {}
    voe = []
    eof = np.empty_like(faces)
    for fi, (a, b, c) in enumerate(faces):
        for slot, (p, q) in enumerate(((a, b), (b, c), (c, a))):
            k = (min(p, q), max(p, q))
            if k not in edges:
                edges[k] = len(voe)
                voe.append((p, q))
            eof[fi, slot] = edges[k]
    return Mesh(xyz.astype(np.float64), np.array(voe, dtype=np.int32) + 1, eof.astype(np.int32) + 1)


def spheroid(a: float, c: float, freq: int) -> Mesh:
    """Icosphere scaled to semi-axes (a, a, c) (unit-radius input for rad_p = 1)."""
    m = icosphere(freq)
    xyz = m.xyz * np.array([a, a, c])
    return Mesh(xyz, m.vert_of_edge, m.edge_of_face)


def sphere_level(level: int) -> Mesh:
    """Level-L subdivided icosahedron (the reference's unitSphere_N<L>.gts shape)."""
    return icosphere(2 ** level)
