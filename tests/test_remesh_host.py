"""The library's remeshing (melt_remesh_coarsen / melt_remesh_smooth: host
code in libmeltgpu.so, by design; no GPU call) against the oracle's
restatement, bit for bit: positions, melt velocities, normals, edge lengths,
connectivity, ghost flags, anchor / flagged state and the counters, on jittered
spheres with valence-3 vertices planted among short edges (so collapses,
deletions and the stale 1-ring normal refreshes interleave), over several
thresholds and seeds.  Argument checks refuse out-of-range indices."""
import ctypes as C

import numpy as np
import pytest

import oracle
from melt_ibm_amd import geometry
from melt_ibm_amd.api import load_library
from melt_ibm_amd.remesh import TriMesh, coarsen, smooth


@pytest.fixture(scope="module")
def libs():
    try:
        lib = load_library()
    except FileNotFoundError:
        pytest.fail("libmeltgpu.so is not built")
    return lib, oracle.load()


def planted(level=3, n_split=12, seed=0, rad=0.4, amp=0.15):
    """A jittered sphere with n_split faces split at a point near their centroid
    (each split leaves a valence-3 vertex with three short edges)."""
    rng = np.random.default_rng(seed)
    m = geometry.sphere_level(level)
    h = 1.0 / 2 ** level
    x = m.xyz + amp * h * rng.standard_normal(m.xyz.shape)
    x /= np.linalg.norm(x, axis=1)[:, None]
    faces = m.faces() - 1
    pick = rng.choice(len(faces), n_split, replace=False)
    keep = np.ones(len(faces), bool)
    keep[pick] = False
    new, pts = [faces[keep]], [x]
    nv = len(x)
    for f in pick:
        a, b, c = faces[f]
        w = rng.dirichlet([4, 4, 4])
        p = w[0] * x[a] + w[1] * x[b] + w[2] * x[c]
        pts.append((p / np.linalg.norm(p))[None])
        new.append(np.array([[a, b, nv], [b, c, nv], [c, a, nv]]))
        nv += 1
    t = TriMesh.from_mesh(geometry._mesh_from_faces(np.vstack(pts), np.vstack(new)), scale=rad)
    t.vmelt = rng.standard_normal((t.counts[0], 3))
    return t


ARRAYS = ("xyz", "vert_of_face", "edge_of_face", "vert_of_edge", "face_of_edge", "ghost_v", "ghost_e", "ghost_f", "anchor_v", "flagged_e",
          "tri_nor", "e_lengths", "vmelt")


def same(a: TriMesh, b: TriMesh):
    for k in ARRAYS:
        x, y = getattr(a, k), getattr(b, k)
        assert np.array_equal(x, y, equal_nan=True), k


@pytest.mark.parametrize("seed,quant", [(0, 0.02), (0, 0.03), (2, 0.03), (5, 0.02), (7, 0.02)])
def test_coarsen_and_smooth_match_the_oracle(libs, seed, quant):
    lib, orc = libs
    t = planted(seed=seed)
    thresh = float(np.quantile(t.e_lengths, quant))
    V0 = t.volume()
    a, b = t.copy(), t.copy()
    ra, rb = coarsen(a, thresh, lib, "melt_"), coarsen(b, thresh, orc, "orc_")
    assert ra == rb and ra["rc"] == 0, (ra, rb)
    assert ra["vdel"] > 0 and ra["ecol"] > 0
    same(a, b)
    Vc = a.volume()
    sa, sb = smooth(a, V0, -(Vc - V0), lib, "melt_"), smooth(b, V0, -(Vc - V0), orc, "orc_")
    assert sa == sb and sa["rc"] == 0, (sa, sb)
    same(a, b)


@pytest.mark.parametrize("seed,quant", [(1, 0.06), (3, 0.02), (4, 0.02)])
def test_failures_match_the_oracle(libs, seed, quant):
    """Many short edges next to planted valence-3 vertices: a collapse leaves a
    vertex with two neighbours (the reference aborts, remesh_smooth.f90:422-426)
    or a degenerate valence-3 ring; both sides stop after the same edits."""
    lib, orc = libs
    t = planted(seed=seed)
    thresh = float(np.quantile(t.e_lengths, quant))
    a, b = t.copy(), t.copy()
    ra, rb = coarsen(a, thresh, lib, "melt_"), coarsen(b, thresh, orc, "orc_")
    assert ra == rb and ra["rc"] == 6, (ra, rb)
    same(a, b)


def test_bad_indices_are_refused(libs):
    lib, _ = libs
    t = planted(seed=5, n_split=2)
    t.vert_of_face[3, 1] = 0
    assert coarsen(t, 0.01, lib, "melt_")["rc"] == 1
    t = planted(seed=5, n_split=2)
    t.face_of_edge[7, 0] = t.counts[2] + 1
    assert smooth(t, 0.1, 0.0, lib, "melt_")["rc"] == 1
