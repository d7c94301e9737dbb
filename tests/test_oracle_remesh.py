"""Pinning the oracle's remeshing (remesh_coarsen.f90, remesh_smooth.f90,
particle.f90:121-211) on the CPU, against independent numpy formulations:

 - validity of the edited surface: every live face has three distinct live
   vertices and three live edges whose vertex pairs are the face's sides, every
   live edge has two distinct live faces that list it and run it in opposite
   directions (consistent orientation), V - E + F = 2, and each operation takes
   one vertex, three edges and two faces (collapse or valence-3 deletion);
 - the collapse position: the minimiser of the summed fundamental quadrics of
   the planes of the faces around both vertices (Garland & Heckbert 1997),
   solved as 3x3 normal equations in numpy, where the oracle solves the 4x4
   system of remesh_coarsen.f90:318-341 with its last row replaced; vmelt of
   the kept vertex is the mean of the two;
 - valence-3 vertex deletion gives back the face the vertex had split,
   outward-oriented, with no vertex moved;
 - the edge relaxations of Kuprat et al. (2001) preserve the enclosed volume
   (smoothing with target_DV = 0 leaves it at round-off), and with the
   coarsening's volume change as target the remeshed surface encloses the
   pre-coarsening volume to the reference's 1e-14 (particle.f90:199-211);
 - through orc_particle: a melting sphere whose shortest edges fall below
   E_thresh is remeshed in the same substep, the residual stays below 1e-14
   and the face velocities are U + omega x (centroid - x_cm).
"""
import numpy as np
import pytest

import oracle
from melt_ibm_amd import geometry
from melt_ibm_amd.remesh import TriMesh, coarsen, smooth


@pytest.fixture(scope="module")
def lib():
    return oracle.load()


def check_valid(t: TriMesh) -> None:
    gv, ge, gf = t.ghost_v, t.ghost_e, t.ghost_f
    vof, eof, voe, foe = t.vert_of_face - 1, t.edge_of_face - 1, t.vert_of_edge - 1, t.face_of_edge - 1
    live_f = np.nonzero(gf == 0)[0]
    live_e = np.nonzero(ge == 0)[0]
    live_v = np.nonzero(gv == 0)[0]
    directed = {}
    for f in live_f:
        vs = vof[f]
        assert len(set(vs)) == 3 and not gv[vs].any(), (f, vs)
        es = eof[f]
        assert len(set(es)) == 3 and not ge[es].any(), (f, es)
        sides = {frozenset((vs[0], vs[1])), frozenset((vs[1], vs[2])), frozenset((vs[2], vs[0]))}
        assert {frozenset(voe[e]) for e in es} == sides, f
        for a, b in ((vs[0], vs[1]), (vs[1], vs[2]), (vs[2], vs[0])):
            assert (a, b) not in directed, ("side traversed twice in one direction", a, b)
            directed[(a, b)] = f
    for e in live_e:
        fa, fb = foe[e]
        assert fa != fb and not gf[fa] and not gf[fb], (e, fa, fb)
        assert e in eof[fa] and e in eof[fb], e
        p, q = voe[e]
        assert {directed.get((p, q)), directed.get((q, p))} == {fa, fb}, e
    used = np.zeros(t.counts[0], bool)
    used[vof[live_f].ravel()] = True
    assert used[live_v].all() and not used[gv == 1].any()
    assert len(live_v) - len(live_e) + len(live_f) == 2


def jittered_sphere(level=3, amp=0.12, seed=0, rad=0.3):
    """A sphere of radius rad (the reference's bodies: 0.2-0.6, volumes below 1, where the
    serial volume sum rounds well below the 1e-14 residual bar) with jittered vertices."""
    m = geometry.sphere_level(level)
    rng = np.random.default_rng(seed)
    h = 1.0 / 2 ** level
    x = m.xyz + amp * h * rng.standard_normal(m.xyz.shape)
    return TriMesh.from_mesh(geometry.Mesh(x / np.linalg.norm(x, axis=1)[:, None], m.vert_of_edge, m.edge_of_face), scale=rad)


def test_icosphere_import_is_valid():
    check_valid(TriMesh.from_mesh(geometry.sphere_level(3)))


def test_coarsening_keeps_a_valid_surface(lib):
    t = jittered_sphere(3)
    V0, E0, F0 = t.live_counts()
    thresh = np.quantile(t.e_lengths, 0.05)
    out = coarsen(t, thresh, lib, "orc_")
    assert out["rc"] == 0
    ops = out["ecol"] + out["vdel"]
    assert out["ecol"] > 20
    assert t.live_counts() == (V0 - ops, E0 - 3 * ops, F0 - 2 * ops)
    check_valid(t)
    # the stored normals of the live faces are those of their current triangles, outward
    ref = t.copy()
    ref.update_normals()
    live = t.ghost_f == 0
    np.testing.assert_allclose(t.tri_nor[live], ref.tri_nor[live], atol=1e-12)
    f = t.vert_of_face[live] - 1
    cen = t.xyz[f].mean(axis=1)
    assert (np.einsum("ij,ij->i", ref.tri_nor[live], cen) > 0).all()
    if out["rm_flag"] == 0:
        t.update_lengths()
        assert (t.e_lengths[t.ghost_e == 0] >= thresh).all()


def _ring_faces(t: TriMesh, v: int) -> list[int]:
    live = np.nonzero(t.ghost_f == 0)[0]
    return [f for f in live if v in t.vert_of_face[f]]


def test_collapse_position_minimises_the_quadric(lib):
    m = geometry.sphere_level(2)
    x = m.xyz.copy()
    e = 7
    a, b = m.vert_of_edge[e] - 1
    x[b] = x[a] + 0.3 * (x[b] - x[a])          # one short edge, valences 5 and 6
    t = TriMesh.from_mesh(geometry.Mesh(x, m.vert_of_edge, m.edge_of_face))
    L = t.e_lengths
    thresh = 0.5 * np.sort(L)[1]
    assert (L <= thresh).sum() == 1 and L[e] <= thresh
    t.vmelt = np.random.default_rng(1).standard_normal((t.counts[0], 3))
    v1, v2 = t.vert_of_edge[e]
    Q = np.zeros((4, 4))
    for v in (v1, v2):
        for f in _ring_faces(t, v):
            n = t.tri_nor[f]
            p = np.append(n, -n @ t.xyz[v - 1])
            p /= np.linalg.norm(p)
            Q += np.outer(p, p)
    want = np.linalg.solve(Q[:3, :3], -Q[:3, 3])
    vm = 0.5 * (t.vmelt[v1 - 1] + t.vmelt[v2 - 1])
    faces_e = set(t.face_of_edge[e])
    out = coarsen(t, thresh, lib, "orc_")
    assert out["rc"] == 0 and out["ecol"] == 1 and out["vdel"] == 0
    np.testing.assert_allclose(t.xyz[v1 - 1], want, rtol=0, atol=1e-12)
    np.testing.assert_array_equal(t.vmelt[v1 - 1], vm)
    assert t.ghost_v[v2 - 1] == 1 and t.ghost_e[e] == 1
    assert {f + 1 for f in np.nonzero(t.ghost_f)[0]} == faces_e
    check_valid(t)


def test_valence_three_vertex_is_deleted(lib):
    m = geometry.sphere_level(2)
    faces = m.faces() - 1
    f0 = 11
    a, b, c = faces[f0]
    x = np.vstack([m.xyz, (m.xyz[a] + m.xyz[b] + m.xyz[c]) / 3.0 * 1.01])
    n = len(x) - 1
    new = np.vstack([np.delete(faces, f0, axis=0), [[a, b, n], [b, c, n], [c, a, n]]])
    t = TriMesh.from_mesh(geometry._mesh_from_faces(x, new))
    check_valid(t)
    short = [i for i, (p, q) in enumerate(t.vert_of_edge - 1) if n in (p, q)]
    thresh = 1.01 * t.e_lengths[short].max()
    assert set(np.nonzero(t.e_lengths <= thresh)[0]) == set(short)
    x0 = t.xyz.copy()
    out = coarsen(t, thresh, lib, "orc_")
    assert out["rc"] == 0 and out["vdel"] == 1 and out["ecol"] == 0
    assert t.ghost_v[n] == 1 and all(t.ghost_e[s] == 1 for s in short)
    np.testing.assert_array_equal(t.xyz, x0)
    check_valid(t)
    kept = [f for f in np.nonzero(t.ghost_f == 0)[0] if set(t.vert_of_face[f] - 1) == {a, b, c}]
    assert len(kept) == 1
    vs = t.vert_of_face[kept[0]] - 1
    nrm = np.cross(x[vs[1]] - x[vs[0]], x[vs[2]] - x[vs[0]])
    assert nrm @ x[vs].mean(axis=0) > 0


def test_edge_relaxation_preserves_volume(lib):
    t = jittered_sphere(3, seed=2)
    assert coarsen(t, np.quantile(t.e_lengths, 0.04), lib, "orc_")["rc"] == 0
    flagged = t.flagged_e.astype(bool) & (t.ghost_e == 0)
    assert flagged.sum() > 50
    V = t.volume()
    x0 = t.xyz.copy()
    out = smooth(t, V, 0.0, lib, "orc_")
    assert out["rc"] == 0 and out["n_erel"] == 10 * flagged.sum() and out["refresh"] == 0 and out["drift"] > 1e-5
    assert abs(t.volume() - V) < 1e-14
    moved = set(np.nonzero((t.xyz != x0).any(axis=1))[0])
    assert moved and moved <= set((t.vert_of_edge[flagged] - 1).ravel())
    assert t.flagged_e.sum() == 0          # re-anchored after smoothing
    check_valid(t)


def test_coarsen_then_smooth_restores_the_volume(lib):
    t = jittered_sphere(3, seed=4)
    V0 = t.volume()
    assert coarsen(t, np.quantile(t.e_lengths, 0.06), lib, "orc_")["rc"] == 0
    Vc = t.volume()
    assert abs(Vc - V0) > 1e-6
    out = smooth(t, V0, -(Vc - V0), lib, "orc_")
    assert out["rc"] == 0
    assert abs(t.volume() - V0) < 1e-14
    check_valid(t)


def test_particle_remeshes_a_melting_sphere(lib):
    from melt_ibm_amd.inputs import inqpr_temperature
    from tests.helpers import load_state, make_cfg, melt_case, with_body
    el = np.unique(TriMesh.from_mesh(geometry.sphere_level(3), scale=1.2).e_lengths)
    h = 2 * np.pi / 48
    cfg = make_cfg(48, iremesh=1, perc_ethresh=0.5 * (el[0] + el[1]) / h)     # the shortest class of edges
    o = oracle.context(cfg)
    st, dt = melt_case(cfg, 3, 1.2)
    load_state(o, st)
    with_body(o, 3, 1.2)
    assert o.body().h_eulerian == h
    o.upload("temp", inqpr_temperature(cfg, o.body().pos_cm[:], 1.2))
    o.set_dt(dt)
    o.set_stage(1)
    o.tagCells()
    o.particle()
    st = o.remesh_stats()
    assert st["events"] == 1 and st["collapses"] > 0 and abs(st["residual"]) <= 1e-14
    topo = o.topology()
    t = TriMesh(o.lag("xyzv"), topo["vert_of_face"], topo["edge_of_face"], topo["vert_of_edge"], topo["face_of_edge"],
                topo["ghost_v"], topo["ghost_e"], topo["ghost_f"], np.ones(1, np.int32), np.zeros(1, np.int32))
    check_valid(t)
    b = o.body()
    bar, vel = o.lag("tri_bar"), o.lag("vel_tri")
    live = topo["ghost_f"] == 0
    want = np.asarray(b.vel_cm[:]) + np.cross(np.asarray(b.omega_c[:]), bar - np.asarray(b.pos_cm[:]))
    np.testing.assert_allclose(vel[live], want[live], rtol=0, atol=1e-15)
    assert abs(b.volume - t.volume()) < 1e-12


def subdivided_cube(k=4, half=0.3):
    """A closed cube surface, each face a k x k grid of squares cut into two
    triangles (flat patches with valence-6 interior vertices), outward."""
    pts, index = [], {}

    def vid(p):
        key = tuple(p)
        if key not in index:
            index[key] = len(pts)
            pts.append(key)
        return index[key]

    faces = []
    for axis in range(3):
        for side in (0, k):
            u, v = [a for a in range(3) if a != axis]
            for a in range(k):
                for b in range(k):
                    quad = []
                    for da, db in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        p = [0, 0, 0]
                        p[axis], p[u], p[v] = side, a + da, b + db
                        quad.append(vid(p))
                    faces += [[quad[0], quad[1], quad[2]], [quad[0], quad[2], quad[3]]]
    x = (np.array(pts, dtype=np.float64) / k - 0.5) * 2 * half
    faces = np.array(faces)
    n = np.cross(x[faces[:, 1]] - x[faces[:, 0]], x[faces[:, 2]] - x[faces[:, 0]])
    flip = np.einsum("ij,ij->i", n, x[faces].mean(axis=1)) < 0
    faces[flip] = faces[flip][:, [0, 2, 1]]
    return geometry._mesh_from_faces(x, faces)


def test_relaxation_on_a_flat_patch_moves_to_the_coupled_centroids(lib):
    """On a flat patch the relaxation needs no normal move (h = 0), so an edge
    (v1, v2) ends where each vertex is the centroid of its ring with the other
    at its new place: n1 x1 = S1 + x2, n2 x2 = S2 + x1 (remesh_smooth.f90:173-186),
    solved here as a 2x2 system."""
    t = TriMesh.from_mesh(subdivided_cube(6))
    check_valid(t)
    x = t.xyz
    on_face = np.isclose(x[:, 2], 0.3)
    interior = on_face & (np.abs(x[:, 0]) < 0.29) & (np.abs(x[:, 1]) < 0.29)
    e = next(i for i, (p, q) in enumerate(t.vert_of_edge - 1) if interior[p] and interior[q])
    v1, v2 = t.vert_of_edge[e] - 1
    rings = {}
    for v in (v1, v2):
        nb = set()
        for f in range(t.counts[2]):
            if v in t.vert_of_face[f] - 1:
                nb |= set(t.vert_of_face[f] - 1)
        rings[v] = sorted(nb - {v})
    S1 = x[[w for w in rings[v1] if w != v2]].sum(axis=0)
    S2 = x[[w for w in rings[v2] if w != v1]].sum(axis=0)
    n1, n2 = len(rings[v1]), len(rings[v2])
    M = np.array([[n1, -1.0], [-1.0, n2]])
    want = np.linalg.solve(M, np.vstack([S1, S2]))
    V = t.volume()
    t.flagged_e[:] = 0
    t.flagged_e[e] = 1
    out = smooth(t, V, 0.0, lib, "orc_")
    assert out["rc"] == 0
    # ten sweeps over the one flagged edge: the first moves the pair to the
    # coupled centroids, the later ones find it there
    np.testing.assert_allclose(t.xyz[v1], want[0], rtol=0, atol=1e-14)
    np.testing.assert_allclose(t.xyz[v2], want[1], rtol=0, atol=1e-14)
    assert abs(t.volume() - V) < 1e-15
