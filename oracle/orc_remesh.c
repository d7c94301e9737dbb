/*
 * ORACLE -- test infrastructure only (see orc.h).
 *
 * Remeshing of the melting body, restated from remesh_coarsen.f90 and
 * remesh_smooth.f90 and the call site particle.f90:121-160:
 *  - coarsening: every live edge not longer than E_thresh is collapsed to the
 *    minimiser of the summed fundamental error quadrics of its two vertices
 *    (Garland & Heckbert 1997), or, where one of its vertices has valence 3,
 *    that vertex is deleted; up to 5 sweeps while some edge stays short;
 *  - smoothing: edge relaxations of Kuprat et al. (2001, algorithm 4) over the
 *    edges the coarsening touched, 10 sweeps, the first relaxations carrying
 *    the volume the coarsening changed (target_DV) until the volume is back.
 * Ghost entities stay in their slots, as in the reference (isGhost* flags).
 *
 * Connectivity is 1-based as in the reference; flags are 0/1.  Indices follow
 * the reference's loops exactly, including its choice of "arbitrary" starting
 * faces and the order of where-statements, since they decide which entity is
 * kept.  Where the reference would index out of bounds, stop, or loop
 * without end on a broken mesh (a 1-ring longer than its 30-entry buffers, an
 * edge left with more than two faces after a collapse, a walk around a vertex
 * that finds no next edge, a valence-3 vertex whose ring is not 3 edges and 3
 * faces, a singular QEM system), this raises MELT_ERR_REMESH instead.
 * flagged_faces starts empty (the reference leaves it uninitialised until the
 * first collapse).
 */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "orc.h"

#define RING_MAX 30

typedef struct {
  int nv, ne, nf;
  double *xyz, *tri_nor;
  int32_t *vof, *eof, *voe, *foe;
  int32_t *gv, *ge, *gf;
} rm_mesh;

#define X(d, v) m->xyz[3 * ((v) - 1) + (d) - 1]
#define NOR(d, f) m->tri_nor[3 * ((f) - 1) + (d) - 1]
#define VOF(c, f) m->vof[3 * ((f) - 1) + (c) - 1]
#define EOFC(c, f) m->eof[3 * ((f) - 1) + (c) - 1]
#define VOE(c, e) m->voe[2 * ((e) - 1) + (c) - 1]
#define FOE(c, e) m->foe[2 * ((e) - 1) + (c) - 1]

static int fail(const char* what) {
  (void)what;
  return MELT_ERR_REMESH;
}

/* gfortran dot_product of two 3-vectors: left to right */
static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* get_next_edge_of_v remesh_coarsen.f90:423-450 (= get_next_edge remesh_smooth.f90:443-474):
 * the edge of cur_face other than prev that holds v (the last such in face order) */
static int next_edge(const rm_mesh* m, int v, int cur_face, int prev, int* next) {
  int ok = 0;
  for (int i = 1; i <= 3; ++i) {
    int ei = EOFC(i, cur_face);
    if (ei != prev && (VOE(1, ei) == v || VOE(2, ei) == v)) { *next = ei; ok = 1; }
  }
  return ok ? 0 : fail("next edge");
}

/* the face across prev_edge from prev_face (the walk step shared by the 1-ring routines) */
static int other_face(const rm_mesh* m, int prev_edge, int prev_face) {
  return FOE(1, prev_edge) != prev_face ? FOE(1, prev_edge) : FOE(2, prev_edge);
}

/* calculate_eLengths triaux.f90:277-305 */
static void elengths(const rm_mesh* m, double* el, int* rm_flag, double thresh) {
  *rm_flag = 0;
  for (int i = 1; i <= m->ne; ++i) {
    if (m->ge[i - 1]) continue;
    int v1 = VOE(1, i), v2 = VOE(2, i);
    double a = X(1, v1) - X(1, v2), b = X(2, v1) - X(2, v2), c = X(3, v1) - X(3, v2);
    el[i - 1] = sqrt(a * a + b * b + c * c);
    if (el[i - 1] < thresh) *rm_flag = 1;
  }
}

/* calculate_volume triaux.f90:53-79 */
static double volume(const rm_mesh* m) {
  double V = 0.0;
  for (int i = 1; i <= m->nf; ++i) {
    if (m->gf[i - 1]) continue;
    int v1 = VOF(1, i), v2 = VOF(2, i), v3 = VOF(3, i);
    double vol = X(1, v1) * (X(2, v2) * X(3, v3) - X(3, v2) * X(2, v3)) + X(1, v2) * (X(2, v3) * X(3, v1) - X(3, v3) * X(2, v1)) +
                 X(1, v3) * (X(2, v1) * X(3, v2) - X(3, v1) * X(2, v2));
    V = V + vol;
  }
  return V / 6.0;
}

/* DGEMM('N','T',4,4,1,1,p,4,p,4,1,Q,4): Q(i,j) += p(j)*p(i) */
static void add_quadric(double Q[16], const double p[4]) {
  for (int j = 0; j < 4; ++j) {
    double temp = 1.0 * p[j];
    for (int i = 0; i < 4; ++i) Q[i + 4 * j] = Q[i + 4 * j] + temp * p[i];
  }
}

/* the fundamental quadric of face f through vertex v (remesh_coarsen.f90:285-289) */
static void face_quadric(const rm_mesh* m, double Q[16], int f, int v) {
  double p[4], n[3] = {NOR(1, f), NOR(2, f), NOR(3, f)}, x[3] = {X(1, v), X(2, v), X(3, v)};
  p[0] = n[0]; p[1] = n[1]; p[2] = n[2];
  p[3] = -dot3(n, x);
  double s = orc_norm2(p, 4);
  for (int i = 0; i < 4; ++i) p[i] = p[i] / s;
  add_quadric(Q, p);
}

/* calc_errorQuadric_of_v remesh_coarsen.f90:260-316 */
static int error_quadric(const rm_mesh* m, double Q[16], int v, int e) {
  int F1 = FOE(1, e), prev_face = F1, cur_edge = 0, prev_edge = e, steps = 0;
  face_quadric(m, Q, F1, v);
  while (cur_edge != e) {
    int cur_face = other_face(m, prev_edge, prev_face);
    if (cur_face == F1) break;
    if (++steps > RING_MAX) return fail("ring");
    face_quadric(m, Q, cur_face, v);
    if (next_edge(m, v, cur_face, prev_edge, &cur_edge)) return MELT_ERR_REMESH;
    prev_face = cur_face;
    prev_edge = cur_edge;
  }
  return 0;
}

/* DGESV(4,1,A,4,IPIV,b,4,INFO): DGETRF (partial pivoting, netlib order) and
 * DGETRS (row interchanges, unit-lower then upper triangular solves). */
static int dgesv4(double a[16], double b[4]) {
#define A(i, j) a[(i) + 4 * (j)]
  int ipiv[4], info = 0;
  for (int j = 0; j < 4; ++j) {
    int jp = j;
    double amax = fabs(A(j, j));
    for (int i = j + 1; i < 4; ++i)
      if (fabs(A(i, j)) > amax) { amax = fabs(A(i, j)); jp = i; }
    ipiv[j] = jp;
    if (A(jp, j) != 0.0) {
      if (jp != j)
        for (int c = 0; c < 4; ++c) { double t = A(j, c); A(j, c) = A(jp, c); A(jp, c) = t; }
      if (j < 3) {
        if (fabs(A(j, j)) >= 2.2250738585072014e-308) {
          double r = 1.0 / A(j, j);
          for (int i = j + 1; i < 4; ++i) A(i, j) = A(i, j) * r;
        } else {
          for (int i = j + 1; i < 4; ++i) A(i, j) = A(i, j) / A(j, j);
        }
      }
    } else if (info == 0) {
      info = j + 1;
    }
    if (j < 3)
      for (int jj = j + 1; jj < 4; ++jj) {
        if (A(j, jj) != 0.0) {
          double temp = -A(j, jj);
          for (int ii = j + 1; ii < 4; ++ii) A(ii, jj) = A(ii, jj) + A(ii, j) * temp;
        }
      }
  }
  if (info) return info;
  for (int i = 0; i < 4; ++i)
    if (ipiv[i] != i) { double t = b[i]; b[i] = b[ipiv[i]]; b[ipiv[i]] = t; }
  for (int k = 0; k < 4; ++k)
    if (b[k] != 0.0)
      for (int i = k + 1; i < 4; ++i) b[i] = b[i] - b[k] * A(i, k);
  for (int k = 3; k >= 0; --k)
    if (b[k] != 0.0) {
      b[k] = b[k] / A(k, k);
      for (int i = 0; i < k; ++i) b[i] = b[i] - b[k] * A(i, k);
    }
  return 0;
#undef A
}

/* solve_QEM_system remesh_coarsen.f90:318-341 */
static int solve_qem(double Q1[16], const double Q2[16], double* out) {
  for (int i = 0; i < 16; ++i) Q1[i] = Q1[i] + Q2[i];
  double b[4] = {0.0, 0.0, 0.0, 1.0};
  Q1[3 + 0] = 0.0; Q1[3 + 4] = 0.0; Q1[3 + 8] = 0.0; Q1[3 + 12] = 1.0;
  if (dgesv4(Q1, b)) return fail("Error in QEM system solve");
  out[0] = b[0]; out[1] = b[1]; out[2] = b[2];
  return 0;
}

/* get_ghost_and_retained_edges remesh_coarsen.f90:344-366 */
static void ghost_and_retained(const rm_mesh* m, int f, int e, int* ghost, int* kept) {
  int buf[3] = {0, 0, 0}, cnt = 0;
  for (int i = 1; i <= 3; ++i)
    if (EOFC(i, f) != e) buf[cnt++] = EOFC(i, f);
  *ghost = buf[0];
  *kept = buf[1];
}

/* update_face_connectivity remesh_coarsen.f90:368-381 */
static void update_faces(rm_mesh* m, int v1, int v2, int e1, int e2, int re1, int re2) {
  for (int i = 0; i < 3 * m->nf; ++i)
    if (m->vof[i] == v2) m->vof[i] = v1;
  for (int i = 0; i < 3 * m->nf; ++i)
    if (m->eof[i] == e1) m->eof[i] = re1;
  for (int i = 0; i < 3 * m->nf; ++i)
    if (m->eof[i] == e2) m->eof[i] = re2;
}

/* the faces of the retained edge after its ghosted twin is merged into it */
static int merge_faces(rm_mesh* m, int eg, int ek, int F) {
  int buf[4] = {FOE(1, eg), FOE(2, eg), FOE(1, ek), FOE(2, ek)}, cnt = 0;
  for (int i = 0; i < 4; ++i) {
    if (buf[i] == F) continue;
    if (cnt == 2) return fail("edge keeps more than two faces");
    FOE(cnt + 1, ek) = buf[i];
    ++cnt;
  }
  return 0;
}

/* update_edge_connectivity remesh_coarsen.f90:383-421 */
static int update_edges(rm_mesh* m, int v1, int v2, int e1, int e2, int re1, int re2, int F1, int F2) {
  for (int i = 0; i < 2 * m->ne; ++i)
    if (m->voe[i] == v2) m->voe[i] = v1;
  if (merge_faces(m, e1, re1, F1)) return MELT_ERR_REMESH;
  return merge_faces(m, e2, re2, F2);
}

/* flag_neighbours_1ring remesh_coarsen.f90:454-513 */
static int flag_ring(rm_mesh* m, int32_t* anchor, int32_t* flagged_edge, int flagged_faces[RING_MAX], int v, int e) {
  anchor[v - 1] = 0;
  flagged_edge[e - 1] = 1;
  for (int i = 0; i < RING_MAX; ++i) flagged_faces[i] = 0;
  int F1 = FOE(1, e), prev_face = F1, cur_edge = 0, prev_edge = e, cnt_f = 1;
  flagged_faces[0] = F1;
  while (cur_edge != e) {
    int cur_face = other_face(m, prev_edge, prev_face);
    if (cur_face == F1) break;
    if (cnt_f == RING_MAX) return fail("ring");
    flagged_faces[cnt_f++] = cur_face;
    if (next_edge(m, v, cur_face, prev_edge, &cur_edge)) return MELT_ERR_REMESH;
    anchor[VOE(1, cur_edge) - 1] = 0;
    anchor[VOE(2, cur_edge) - 1] = 0;
    flagged_edge[cur_edge - 1] = 1;
    prev_face = cur_face;
    prev_edge = cur_edge;
  }
  return 0;
}

/* get_edge_face_1ring remesh_coarsen.f90:515-579 */
static int edge_face_ring(const rm_mesh* m, int e_nb[RING_MAX], int f_nb[RING_MAX], int* cnt_en, int* cnt_f, int v, int e) {
  for (int i = 0; i < RING_MAX; ++i) e_nb[i] = f_nb[i] = 0;
  *cnt_en = 1;
  e_nb[0] = e;
  int F1 = FOE(1, e), prev_face = F1, cur_edge = 0, prev_edge = e;
  *cnt_f = 1;
  f_nb[0] = F1;
  while (cur_edge != e) {
    int cur_face = other_face(m, prev_edge, prev_face);
    if (cur_face == F1) break;
    if (*cnt_f == RING_MAX || *cnt_en == RING_MAX) return fail("ring");
    f_nb[(*cnt_f)++] = cur_face;
    if (next_edge(m, v, cur_face, prev_edge, &cur_edge)) return MELT_ERR_REMESH;
    e_nb[(*cnt_en)++] = cur_edge;
    prev_face = cur_face;
    prev_edge = cur_edge;
  }
  return 0;
}

/* get_vertNeighbours remesh_smooth.f90:349-441 */
static int vert_neighbours(const rm_mesh* m, int* num, int nb[RING_MAX], int v, int e, int reverse) {
  for (int i = 0; i < RING_MAX; ++i) nb[i] = 0;
  int F1 = FOE(1, e), prev_face = F1, cur_edge = 0, prev_edge = e;
  *num = 1;
  nb[0] = VOE(1, e) == v ? VOE(2, e) : VOE(1, e);
  while (cur_edge != e) {
    int cur_face = other_face(m, prev_edge, prev_face);
    if (cur_face == F1) break;
    if (*num == RING_MAX) return fail("Exceeded vertNeighbour buffer size");
    ++*num;
    if (next_edge(m, v, cur_face, prev_edge, &cur_edge)) return MELT_ERR_REMESH;
    nb[*num - 1] = VOE(1, cur_edge) == v ? VOE(2, cur_edge) : VOE(1, cur_edge);
    prev_face = cur_face;
    prev_edge = cur_edge;
  }
  if (*num < 3) return fail("Vertex v has less than 3 neighbours");
  if (reverse)
    for (int a = 1, b = *num - 1; a < b; ++a, --b) { int t = nb[a]; nb[a] = nb[b]; nb[b] = t; }
  return 0;
}

/* update_1ring_normals remesh_coarsen.f90:581-633 */
static void ring_normals(rm_mesh* m, const int flagged_faces[RING_MAX]) {
  for (int cnt = 0; cnt < RING_MAX && flagged_faces[cnt] != 0; ++cnt) {
    int i = flagged_faces[cnt];
    if (m->gf[i - 1]) continue;
    double old[3] = {NOR(1, i), NOR(2, i), NOR(3, i)};
    int v1 = VOF(1, i), v2 = VOF(2, i), v3 = VOF(3, i);
    double ve1[3], ve2[3];
    for (int d = 1; d <= 3; ++d) { ve1[d - 1] = X(d, v2) - X(d, v1); ve2[d - 1] = X(d, v3) - X(d, v1); }
    NOR(1, i) = ve1[1] * ve2[2] - ve1[2] * ve2[1];
    NOR(2, i) = ve1[2] * ve2[0] - ve1[0] * ve2[2];
    NOR(3, i) = ve1[0] * ve2[1] - ve1[1] * ve2[0];
    double s = sqrt(NOR(1, i) * NOR(1, i) + NOR(2, i) * NOR(2, i) + NOR(3, i) * NOR(3, i));
    for (int d = 1; d <= 3; ++d) NOR(d, i) = NOR(d, i) / s;
    double nw[3] = {NOR(1, i), NOR(2, i), NOR(3, i)};
    double sgn = dot3(nw, old);
    if (sgn < 0.0) { int t = VOF(1, i); VOF(1, i) = VOF(3, i); VOF(3, i) = t; }
    sgn = copysign(1.0, sgn);
    for (int d = 1; d <= 3; ++d) NOR(d, i) = NOR(d, i) * sgn;
  }
}

/* update_connectivity_vert_deletion remesh_coarsen.f90:637-760: v1 (valence
 * 3) goes with e and its other two edges and two of its three faces; the
 * third face fL becomes (v2, v3, v4). */
static int vert_deletion(rm_mesh* m, int v1, int v2, int* re_face, int e, const int e_nb[RING_MAX], const int f_nb[RING_MAX],
                         const int v_nb[RING_MAX], int cnt_en, int cnt_fn, int cnt_vn) {
  int local_e[2] = {0, 0}, local_v[2] = {0, 0}, incr = 0;
  for (int i = 0; i < cnt_en; ++i) {
    if (e_nb[i] == e) continue;
    if (incr == 2) return fail("vertex deletion");
    local_e[incr++] = e_nb[i];
  }
  int e1 = local_e[0], e2 = local_e[1];
  incr = 0;
  for (int i = 0; i < cnt_vn; ++i) {
    if (v_nb[i] == v2) continue;
    if (incr == 2) return fail("vertex deletion");
    local_v[incr++] = v_nb[i];
  }
  int fL = FOE(1, e), fR = FOE(2, e), f3 = 0, e3 = 0, e4 = 0, e5 = 0;
  for (int i = 0; i < cnt_fn; ++i) {
    if (f_nb[i] == fL || f_nb[i] == fR) continue;
    f3 = f_nb[i];
  }
  if (!f3) return fail("vertex deletion");
  for (int i = 1; i <= 3; ++i) {
    if (EOFC(i, fL) != e && EOFC(i, fL) != e1 && EOFC(i, fL) != e2) e3 = EOFC(i, fL);
    if (EOFC(i, fR) != e && EOFC(i, fR) != e1 && EOFC(i, fR) != e2) e4 = EOFC(i, fR);
    if (EOFC(i, f3) != e1 && EOFC(i, f3) != e2) e5 = EOFC(i, f3);
  }
  if (!e3 || !e4 || !e5) return fail("vertex deletion");
  for (int i = 1; i <= 2; ++i) {
    if (FOE(i, e3) == fR || FOE(i, e3) == f3) FOE(i, e3) = fL;
    if (FOE(i, e4) == fR || FOE(i, e4) == f3) FOE(i, e4) = fL;
    if (FOE(i, e5) == fR || FOE(i, e5) == f3) FOE(i, e5) = fL;
  }
  incr = 0;
  for (int i = 1; i <= 3; ++i) {
    if (VOF(i, fL) == v2) continue;
    if (incr == 2) return fail("vertex deletion");
    VOF(i, fL) = local_v[incr++];
  }
  int repl[2] = {e4, e5};
  incr = 0;
  for (int i = 1; i <= 3; ++i) {
    if (EOFC(i, fL) == e3) continue;
    if (incr == 2) return fail("vertex deletion");
    EOFC(i, fL) = repl[incr++];
  }
  m->gv[v1 - 1] = 1;
  m->ge[e - 1] = 1; m->ge[e1 - 1] = 1; m->ge[e2 - 1] = 1;
  m->gf[fR - 1] = 1; m->gf[f3 - 1] = 1;
  *re_face = fL;
  return 0;
}

static int euler_ok(const rm_mesh* m) {
  long V = 0, E = 0, F = 0;
  for (int i = 0; i < m->nv; ++i) V += !m->gv[i];
  for (int i = 0; i < m->ne; ++i) E += !m->ge[i];
  for (int i = 0; i < m->nf; ++i) F += !m->gf[i];
  return V - E + F == 2;
}

static long live_faces(const rm_mesh* m) {
  long F = 0;
  for (int i = 0; i < m->nf; ++i) F += !m->gf[i];
  return F;
}

/* remesh_coarsen remesh_coarsen.f90:2-258 */
static int coarsen(rm_mesh* m, double* el, double e_thresh, int32_t* rm_flag, int32_t* anchor, int32_t* flagged_edge, double* vmelt,
                   int* ecol_cnt, int* vdel_cnt) {
  int flagged_faces[RING_MAX], v1_n[RING_MAX], v2_n[RING_MAX], e_nb[RING_MAX], f_nb[RING_MAX];
  for (int i = 0; i < RING_MAX; ++i) flagged_faces[i] = 0;
  *ecol_cnt = 0;
  *vdel_cnt = 0;
  int sweep = 0, rc;
  while (*rm_flag && sweep < 5) {
    for (int e = 1; e <= m->ne; ++e) {
      if (m->ge[e - 1] || !(el[e - 1] <= e_thresh)) continue;
      int v1 = VOE(1, e), v2 = VOE(2, e), v1val, v2val;
      if ((rc = vert_neighbours(m, &v1val, v1_n, v1, e, 1)) || (rc = vert_neighbours(m, &v2val, v2_n, v2, e, 0))) return rc;
      if (v1val == 3 || v2val == 3) {
        int vd = v1val == 3 ? v1 : v2, vk = v1val == 3 ? v2 : v1, cnt_en, cnt_f, F1 = 0;
        const int* nb = v1val == 3 ? v1_n : v2_n;
        int val = v1val == 3 ? v1val : v2val;
        ++*vdel_cnt;
        if ((rc = edge_face_ring(m, e_nb, f_nb, &cnt_en, &cnt_f, vd, e))) return rc;
        if (cnt_en != 3 || cnt_f != 3) return fail("Not 3 edge or face neighbours");
        if ((rc = vert_deletion(m, vd, vk, &F1, e, e_nb, f_nb, nb, cnt_en, cnt_f, val))) return rc;
        flagged_faces[0] = F1;
        ring_normals(m, flagged_faces);
        if (!euler_ok(m)) return fail("Topology not preserved");
      } else {
        double Q1[16], Q2[16];
        for (int i = 0; i < 16; ++i) Q1[i] = Q2[i] = 0.0;
        if ((rc = error_quadric(m, Q1, v1, e)) || (rc = error_quadric(m, Q2, v2, e))) return rc;
        double vbar[3];
        if ((rc = solve_qem(Q1, Q2, vbar))) return rc;
        for (int d = 1; d <= 3; ++d) X(d, v1) = vbar[d - 1];
        if (vmelt)
          for (int d = 0; d < 3; ++d) vmelt[3 * (v1 - 1) + d] = 0.5 * (vmelt[3 * (v1 - 1) + d] + vmelt[3 * (v2 - 1) + d]);
        int F1 = FOE(1, e), F2 = FOE(2, e), e1, e2, re1, re2;
        ghost_and_retained(m, F1, e, &e1, &re1);
        ghost_and_retained(m, F2, e, &e2, &re2);
        m->gf[F1 - 1] = 1; m->gf[F2 - 1] = 1;
        m->gv[v2 - 1] = 1;
        m->ge[e - 1] = 1; m->ge[e1 - 1] = 1; m->ge[e2 - 1] = 1;
        update_faces(m, v1, v2, e1, e2, re1, re2);
        if ((rc = update_edges(m, v1, v2, e1, e2, re1, re2, F1, F2))) return rc;
        if ((rc = flag_ring(m, anchor, flagged_edge, flagged_faces, v1, re1))) return rc;
        int flag = 0;
        elengths(m, el, &flag, e_thresh);
        *rm_flag = flag;
        ring_normals(m, flagged_faces);
        if (!euler_ok(m)) return fail("Topology not preserved");
        if (live_faces(m) <= 4) return fail("Geometry reduced to tetrahedron");
        ++*ecol_cnt;
      }
    }
    ++sweep;
  }
  return 0;
}

/* one edge relaxation of remesh_smooth.f90:105-204 (and its copy :224-320);
 * returns the residual test result through *corrected */
static int relax_edge(rm_mesh* m, int i, int* corrected, double target_dv, double vol_old, double* drift, int* cnt_refresh) {
  int v1 = VOE(1, i), v2 = VOE(2, i), v1val, v2val, v1_n[RING_MAX], v2_n[RING_MAX], rc;
  double A1[3] = {0, 0, 0}, A2[3] = {0, 0, 0}, vec[3] = {0, 0, 0}, ej[3], ejp1[3], buf[3];
  if ((rc = vert_neighbours(m, &v1val, v1_n, v1, i, 1)) || (rc = vert_neighbours(m, &v2val, v2_n, v2, i, 0))) return rc;
  for (int j = 1; j <= v1val; ++j) {
    int jp1 = 1 + (j + 1 - 1) % v1val;
    for (int d = 1; d <= 3; ++d) { ej[d - 1] = X(d, v1_n[j - 1]) - X(d, v1); ejp1[d - 1] = X(d, v1_n[jp1 - 1]) - X(d, v1); }
    orc_cross(buf, ej, ejp1);
    for (int d = 0; d < 3; ++d) A1[d] = A1[d] + buf[d];
  }
  for (int j = 1; j <= v2val; ++j) {
    int jp1 = 1 + (j + 1 - 1) % v2val;
    for (int d = 1; d <= 3; ++d) { ej[d - 1] = X(d, v2_n[j - 1]) - X(d, v2); ejp1[d - 1] = X(d, v2_n[jp1 - 1]) - X(d, v2); }
    orc_cross(buf, ej, ejp1);
    for (int d = 0; d < 3; ++d) A2[d] = A2[d] + buf[d];
    if (j == v2val) for (int d = 0; d < 3; ++d) vec[d] = vec[d] + ej[d];
    else if (j == 2) for (int d = 0; d < 3; ++d) vec[d] = vec[d] - ej[d];
  }
  double x1s[3], x2s[3], dx1s[3], dx2s[3], A[3], nhat[3], dd[3];
  const double omega = 1.0, small = 1.e-15;
  for (int d = 1; d <= 3; ++d) {
    double s2 = 0.0, s1 = 0.0;
    for (int k = 2; k <= v2val; ++k) s2 = s2 + X(d, v2_n[k - 1]);
    for (int k = 2; k <= v1val; ++k) s1 = s1 + X(d, v1_n[k - 1]);
    x1s[d - 1] = (s2 + (double)v2val * s1) / ((double)(v1val * v2val) - 1.0);
  }
  for (int d = 1; d <= 3; ++d) {
    double s2 = 0.0;
    for (int k = 2; k <= v2val; ++k) s2 = s2 + X(d, v2_n[k - 1]);
    x2s[d - 1] = (x1s[d - 1] + s2) / (double)v2val;
  }
  for (int d = 1; d <= 3; ++d) { dx1s[d - 1] = omega * (x1s[d - 1] - X(d, v1)); dx2s[d - 1] = omega * (x2s[d - 1] - X(d, v2)); }
  for (int d = 0; d < 3; ++d) dd[d] = dx1s[d] - dx2s[d];
  orc_cross(buf, vec, dd);
  for (int d = 0; d < 3; ++d) A[d] = A1[d] + A2[d] + buf[d];
  double nA = orc_norm2(A, 3);
  for (int d = 0; d < 3; ++d) nhat[d] = A[d] / (nA + small);
  orc_cross(buf, vec, dx1s);
  double h = -(dot3(dx1s, A1) + dot3(dx2s, A2) + dot3(dx2s, buf));
  if (!*corrected) h = h + copysign(1.0, h) * target_dv * 6.0;
  h = h / (nA + small);
  double t1[3], t2[3];
  for (int d = 0; d < 3; ++d) { t1[d] = dx1s[d] + h * nhat[d]; t2[d] = dx2s[d] + h * nhat[d]; }
  double n1 = orc_norm2(t1, 3), n2 = orc_norm2(t2, 3);
  double dm = *drift;
  if (n1 > dm) dm = n1;
  if (n2 > dm) dm = n2;
  *drift = dm;
  for (int d = 1; d <= 3; ++d) {
    X(d, v1) = X(d, v1) + dx1s[d - 1] + h * nhat[d - 1];
    X(d, v2) = X(d, v2) + dx2s[d - 1] + h * nhat[d - 1];
  }
  if (!*corrected) {
    double vc = volume(m);
    if (fabs(vc - vol_old) < 1.e-14) *corrected = 1;
    else ++*cnt_refresh;
  }
  return 0;
}

/* remesh_smooth remesh_smooth.f90:3-347 */
static int smooth(rm_mesh* m, double vol_old, double target_dv, int32_t* flagged_edge, int* n_erel, double* drift, int* cnt_refresh) {
  int corrected = 0, rc;
  *drift = 1.0e-18;
  *n_erel = 0;
  *cnt_refresh = 0;
  for (int it = 0; it < 10; ++it)
    for (int i = 1; i <= m->ne; ++i) {
      if (!flagged_edge[i - 1] || m->ge[i - 1]) continue;
      ++*n_erel;
      if ((rc = relax_edge(m, i, &corrected, target_dv, vol_old, drift, cnt_refresh))) return rc;
    }
  if (fabs(volume(m) - vol_old) > 1.e-14) corrected = 0;
  if (!corrected)
    for (int i = 1; i <= m->ne; ++i) {
      if (m->ge[i - 1]) continue;
      ++*n_erel;
      if ((rc = relax_edge(m, i, &corrected, target_dv, vol_old, drift, cnt_refresh))) return rc;
      if (corrected) break;
    }
  for (int i = 0; i < m->ne; ++i) flagged_edge[i] = 0;
  return 0;
}

/* ------------------------------------------------------- exported mirrors */
int orc_remesh_coarsen(int32_t nf, int32_t ne, int32_t nv, double* xyz, double* tri_nor, double* e_lengths, double e_thresh,
                       int32_t* vert_of_face, int32_t* edge_of_face, int32_t* vert_of_edge, int32_t* face_of_edge, int32_t* is_ghost_face,
                       int32_t* is_ghost_edge, int32_t* is_ghost_vert, int32_t* rm_flag, int32_t* anchor_vert, int32_t* flagged_edge,
                       double* vmelt, int32_t* ecol_cnt, int32_t* vdel_cnt) {
  if (nf <= 0 || ne <= 0 || nv <= 0 || !xyz || !tri_nor || !e_lengths || !vert_of_face || !edge_of_face || !vert_of_edge || !face_of_edge ||
      !is_ghost_face || !is_ghost_edge || !is_ghost_vert || !rm_flag || !anchor_vert || !flagged_edge)
    return MELT_ERR_ARG;
  rm_mesh mm = {nv, ne, nf, xyz, tri_nor, vert_of_face, edge_of_face, vert_of_edge, face_of_edge, is_ghost_vert, is_ghost_edge, is_ghost_face};
  int ec = 0, vd = 0;
  int rc = coarsen(&mm, e_lengths, e_thresh, rm_flag, anchor_vert, flagged_edge, vmelt, &ec, &vd);
  if (ecol_cnt) *ecol_cnt = ec;
  if (vdel_cnt) *vdel_cnt = vd;
  return rc;
}

int orc_remesh_smooth(double vol_old, double target_dv, int32_t nv, int32_t ne, int32_t nf, double* xyz, const int32_t* is_ghost_vert,
                      const int32_t* is_ghost_edge, const int32_t* is_ghost_face, int32_t* flagged_edge, const int32_t* vert_of_edge,
                      const int32_t* vert_of_face, const int32_t* face_of_edge, const int32_t* edge_of_face, int32_t* n_erel, double* drift,
                      int32_t* cnt_refresh) {
  if (nf <= 0 || ne <= 0 || nv <= 0 || !xyz || !is_ghost_vert || !is_ghost_edge || !is_ghost_face || !flagged_edge || !vert_of_edge ||
      !vert_of_face || !face_of_edge || !edge_of_face)
    return MELT_ERR_ARG;
  rm_mesh mm = {nv, ne, nf, xyz, NULL, (int32_t*)vert_of_face, (int32_t*)edge_of_face, (int32_t*)vert_of_edge, (int32_t*)face_of_edge,
                (int32_t*)is_ghost_vert, (int32_t*)is_ghost_edge, (int32_t*)is_ghost_face};
  int er = 0, cr = 0;
  double dr = 0.0;
  int rc = smooth(&mm, vol_old, target_dv, flagged_edge, &er, &dr, &cr);
  if (n_erel) *n_erel = er;
  if (drift) *drift = dr;
  if (cnt_refresh) *cnt_refresh = cr;
  return rc;
}

/* particle.f90:121-160 and :185-213 on the context's mesh: measure the edges,
 * and if one is short, coarsen and smooth back to vol_melt. *did is set when
 * the mesh was remeshed; stats as orc_get_remesh_stats. */
int orc_remesh_step(orc_ctx* o, double vol_melt, int* did) {
  *did = 0;
  rm_mesh mm = {o->nv, o->ne, o->nf, o->lag[MELT_L_XYZV], o->lag[MELT_L_TRI_NOR], o->vert_of_face, o->edge_of_face, o->vert_of_edge,
                o->face_of_edge, o->ghost_v, o->ghost_e, o->ghost_f};
  rm_mesh* m = &mm;
  double* el = o->lag[MELT_L_ELENGTHS];
  int flag = 0;
  elengths(m, el, &flag, o->body.e_thresh);
  o->body.rm_flag = flag;
  if (!flag) return 0;
  *did = 1;
  /* update_tri_normal triaux.f90:183-232 */
  for (int f = 1; f <= o->nf; ++f) {
    if (o->ghost_f[f - 1]) continue;
    int v1 = VOF(1, f), v2 = VOF(2, f), v3 = VOF(3, f);
    double e1[3], e2[3], n[3];
    for (int d = 1; d <= 3; ++d) { e1[d - 1] = X(d, v2) - X(d, v1); e2[d - 1] = X(d, v3) - X(d, v1); }
    n[0] = e1[1] * e2[2] - e1[2] * e2[1];
    n[1] = e1[2] * e2[0] - e1[0] * e2[2];
    n[2] = e1[0] * e2[1] - e1[1] * e2[0];
    double s = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int d = 1; d <= 3; ++d) NOR(d, f) = n[d - 1] / s;
  }
  int32_t rmf = 1;
  int ec = 0, vd = 0, er = 0, cr = 0, rc;
  double drift = 0.0;
  if ((rc = coarsen(m, el, o->body.e_thresh, &rmf, o->anchor_v, o->flagged_e, o->lag[MELT_L_VMELT], &ec, &vd))) return rc;
  double vol_coarse = volume(m);
  double target = -(vol_coarse - vol_melt);
  if ((rc = smooth(m, vol_melt, target, o->flagged_e, &er, &drift, &cr))) return rc;
  double vol_smooth = volume(m);
  double* s = o->remesh_stats;
  s[0] += 1.0;
  s[1] = ec; s[2] = vd; s[3] = er; s[4] = drift; s[5] = vol_melt - vol_smooth; s[6] = cr;
  s[7] += ec;
  return 0;
}

int orc_get_remesh_stats(orc_ctx* o, double* out) {
  if (!o || !out) return MELT_ERR_ARG;
  memcpy(out, o->remesh_stats, sizeof o->remesh_stats);
  return 0;
}

int orc_get_topology(orc_ctx* o, int32_t* vert_of_face, int32_t* edge_of_face, int32_t* vert_of_edge, int32_t* face_of_edge,
                     int32_t* is_ghost_vert, int32_t* is_ghost_edge, int32_t* is_ghost_face) {
  if (!o || !o->has_geom) return MELT_ERR_NO_GEOMETRY;
  if (vert_of_face) memcpy(vert_of_face, o->vert_of_face, sizeof(int32_t) * 3 * (size_t)o->nf);
  if (edge_of_face) memcpy(edge_of_face, o->edge_of_face, sizeof(int32_t) * 3 * (size_t)o->nf);
  if (vert_of_edge) memcpy(vert_of_edge, o->vert_of_edge, sizeof(int32_t) * 2 * (size_t)o->ne);
  if (face_of_edge) memcpy(face_of_edge, o->face_of_edge, sizeof(int32_t) * 2 * (size_t)o->ne);
  if (is_ghost_vert) memcpy(is_ghost_vert, o->ghost_v, sizeof(int32_t) * (size_t)o->nv);
  if (is_ghost_edge) memcpy(is_ghost_edge, o->ghost_e, sizeof(int32_t) * (size_t)o->ne);
  if (is_ghost_face) memcpy(is_ghost_face, o->ghost_f, sizeof(int32_t) * (size_t)o->nf);
  return 0;
}
