// Remeshing of the melting body (particle.f90:121-211): mesh coarsening by
// edge collapse to the quadric-error minimiser or valence-3 vertex deletion
// (remesh_coarsen.f90:2-760), then volume-restoring edge relaxation
// (remesh_smooth.f90:3-474).
//
// Both are serial, data-dependent edits whose decisions (which edge goes,
// whether the volume check passes at 1e-14) depend on every rounding of the
// serial volume sum, so they run on the host, in the reference's order, on a
// host mirror of the connectivity; the device only measures the edges after
// the Newton-Euler update, and the host looks at the flag once per substep.
// When it is set, positions and melt velocities come down, the edited
// arrays go back up, and the device continues with the geometry refresh.
// Ghost entities keep their slots (isGhostFace/Edge/Vert), as in the
// reference, so every per-face, per-edge and per-vertex array keeps its
// indexing.
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "common.h"

namespace melt_remesh {

constexpr int RING = 30;   // the reference's neighbour buffers

using Ring = std::array<int, RING>;

// A triangulated surface with 0-based connectivity; -1 marks an empty slot.
struct Surface {
  int nv = 0, ne = 0, nf = 0;
  double* x = nullptr;     // 3 nv
  double* nor = nullptr;   // 3 nf
  int32_t *vof = nullptr, *eof = nullptr, *voe = nullptr, *foe = nullptr;
  int32_t *gv = nullptr, *ge = nullptr, *gf = nullptr;

  double xv(int v, int d) const { return x[3 * v + d]; }
};

struct Failure {};

// Which containers (faces or edges) of an index array hold each value,
// complete at every moment: built by a counting sort of the array, then kept
// up as values are rewritten (a rewrite of every `from` into `to` moves the
// whole set of `from` over to `to`; a single write adds its container).  Sets
// may hold containers that no longer reference the value, or twice: every
// use re-checks the array.
struct Holders {
  std::vector<int> ptr, idx;                // the counting sort at build time
  std::vector<std::vector<int>> extra;      // added since
  std::vector<unsigned char> moved;         // the build-time set went to another value
  void build(const int32_t* a, int width, int n, int nvals) {
    ptr.assign(nvals + 1, 0);
    for (int i = 0; i < width * n; ++i) ++ptr[a[i] + 1];
    for (int v = 0; v < nvals; ++v) ptr[v + 1] += ptr[v];
    idx.resize(ptr[nvals]);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < width * n; ++i) idx[fill[a[i]]++] = i / width;
    extra.assign(nvals, {});
    moved.assign(nvals, 0);
  }
  template <class F>
  void each(int v, F&& f) const {
    if (!moved[v])
      for (int p = ptr[v]; p < ptr[v + 1]; ++p) f(idx[p]);
    for (int c : extra[v]) f(c);
  }
  void add(int v, int c) { extra[v].push_back(c); }
  void move(int from, int to) {
    each(from, [&](int c) { extra[to].push_back(c); });
    moved[from] = 1;
    extra[from].clear();
  }
};

class Editor {
 public:
  explicit Editor(Surface& s) : m_(s) { flagged_faces_.fill(-1); }

  // calculate_eLengths triaux.f90:277-305
  bool measure(double* el, double thresh) const {
    bool shorter = false;
    for (int e = 0; e < m_.ne; ++e) {
      if (m_.ge[e]) continue;
      const int a = m_.voe[2 * e], b = m_.voe[2 * e + 1];
      const double dx = m_.xv(a, 0) - m_.xv(b, 0), dy = m_.xv(a, 1) - m_.xv(b, 1), dz = m_.xv(a, 2) - m_.xv(b, 2);
      el[e] = std::sqrt(dx * dx + dy * dy + dz * dz);
      if (el[e] < thresh) shorter = true;
    }
    return shorter;
  }

  // calculate_volume triaux.f90:53-79: serial over the live faces
  double volume() const {
    double V = 0.0;
    for (int f = 0; f < m_.nf; ++f) {
      if (m_.gf[f]) continue;
      const int a = m_.vof[3 * f], b = m_.vof[3 * f + 1], c = m_.vof[3 * f + 2];
      const double t = m_.xv(a, 0) * (m_.xv(b, 1) * m_.xv(c, 2) - m_.xv(b, 2) * m_.xv(c, 1)) +
                       m_.xv(b, 0) * (m_.xv(c, 1) * m_.xv(a, 2) - m_.xv(c, 2) * m_.xv(a, 1)) +
                       m_.xv(c, 0) * (m_.xv(a, 1) * m_.xv(b, 2) - m_.xv(a, 2) * m_.xv(b, 1));
      V = V + t;
    }
    return V / 6.0;
  }

  // update_tri_normal triaux.f90:183-232 on the live faces
  void face_normals() {
    for (int f = 0; f < m_.nf; ++f)
      if (!m_.gf[f]) unit_normal(f, &m_.nor[3 * f]);
  }

  // remesh_coarsen remesh_coarsen.f90:2-258; returns false on a failure the
  // reference aborts on (or would run out of bounds on)
  bool coarsen(double* el, double thresh, int32_t& rm_flag, int32_t* anchor, int32_t* flagged_edge, double* vmelt, int& ncol, int& ndel) {
    ncol = ndel = 0;
    start_counts(el, thresh);
    bool measured = false;
    try {
      for (int sweep = 0; rm_flag && sweep < 5; ++sweep) {
        for (int e = 0; e < m_.ne; ++e) {
          if (m_.ge[e] || !(el[e] <= thresh)) continue;
          const int v1 = m_.voe[2 * e], v2 = m_.voe[2 * e + 1];
          Ring n1, n2;
          const int val1 = neighbours(v1, e, true, n1), val2 = neighbours(v2, e, false, n2);
          if (val1 == 3 || val2 == 3) {
            const bool first = val1 == 3;
            ++ndel;   // counted before the deletion, as vdel_cnt (remesh_coarsen.f90:72,101)
            delete_vertex(first ? v1 : v2, first ? v2 : v1, e, first ? n1 : n2, first ? val1 : val2);
            normals_of_flagged();
            if (!euler_two()) return false;
          } else {
            hold();
            collapse(e, v1, v2, anchor, flagged_edge, vmelt);
            // calculate_eLengths after the collapse (remesh_coarsen.f90:170):
            // all live edges the first time, then only those around v1, the
            // only ones whose length changed (the same arithmetic, so the same
            // values), with the short ones counted as live edges go and come
            if (!measured) {
              measure(el, thresh);
              nshort_ = 0;
              for (int q = 0; q < m_.ne; ++q) nshort_ += !m_.ge[q] && el[q] < thresh;
              measured = true;
            } else {
              remeasure_at(v1);
            }
            rm_flag = nshort_ > 0 ? 1 : 0;
            normals_of_flagged();
            if (!euler_two() || live_faces() <= 4) return false;
            ++ncol;
          }
        }
      }
    } catch (const Failure&) {
      return false;
    }
    return true;
  }

  // remesh_smooth remesh_smooth.f90:3-347
  bool smooth(double vol_old, double target_dv, int32_t* flagged_edge, int& nrel, double& drift, int& nrefresh) {
    bool corrected = false;
    drift = 1.0e-18;
    nrel = nrefresh = 0;
    try {
      for (int it = 0; it < 10; ++it)
        for (int e = 0; e < m_.ne; ++e) {
          if (!flagged_edge[e] || m_.ge[e]) continue;
          ++nrel;
          relax(e, corrected, target_dv, vol_old, drift, nrefresh);
        }
      if (std::fabs(volume() - vol_old) > 1.e-14) corrected = false;
      if (!corrected)
        for (int e = 0; e < m_.ne; ++e) {
          if (m_.ge[e]) continue;
          ++nrel;
          relax(e, corrected, target_dv, vol_old, drift, nrefresh);
          if (corrected) break;
        }
    } catch (const Failure&) {
      return false;
    }
    for (int e = 0; e < m_.ne; ++e) flagged_edge[e] = 0;
    return true;
  }

 private:
  Surface& m_;
  Ring flagged_faces_;   // zero-terminated in the reference: -1 here

  [[noreturn]] static void fail() { throw Failure(); }

  static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
  static void cross(double* r, const double* a, const double* b) {   // triaux.f90 cross
    r[0] = a[1] * b[2] - b[1] * a[2];
    r[1] = a[2] * b[0] - b[2] * a[0];
    r[2] = a[0] * b[1] - b[0] * a[1];
  }
  // gfortran NORM2: scaled sum of squares
  static double norm2(const double* v, int n) {
    double scale = 0.0, ssq = 1.0;
    for (int i = 0; i < n; ++i) {
      if (v[i] == 0.0) continue;
      const double a = std::fabs(v[i]);
      if (scale < a) {
        const double r = scale / a;
        ssq = 1.0 + ssq * r * r;
        scale = a;
      } else {
        const double r = a / scale;
        ssq += r * r;
      }
    }
    return scale * std::sqrt(ssq);
  }

  void unit_normal(int f, double* n) const {
    const int a = m_.vof[3 * f], b = m_.vof[3 * f + 1], c = m_.vof[3 * f + 2];
    double p[3], q[3];
    for (int d = 0; d < 3; ++d) { p[d] = m_.xv(b, d) - m_.xv(a, d); q[d] = m_.xv(c, d) - m_.xv(a, d); }
    n[0] = p[1] * q[2] - p[2] * q[1];
    n[1] = p[2] * q[0] - p[0] * q[2];
    n[2] = p[0] * q[1] - p[1] * q[0];
    const double s = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int d = 0; d < 3; ++d) n[d] = n[d] / s;
  }

  int across(int edge, int face) const { return m_.foe[2 * edge] != face ? m_.foe[2 * edge] : m_.foe[2 * edge + 1]; }

  // the edge of `face` other than `prev` that holds v (the last in face order)
  int next_edge(int v, int face, int prev) const {
    int found = -1;
    for (int i = 0; i < 3; ++i) {
      const int e = m_.eof[3 * face + i];
      if (e != prev && (m_.voe[2 * e] == v || m_.voe[2 * e + 1] == v)) found = e;
    }
    if (found < 0) fail();
    return found;
  }

  // Walk the faces around v, starting across `start` from its first face:
  // visit(face, edge) for each face after the first, with the edge of that
  // face that leads on.  Every 1-ring routine of the reference has this loop;
  // the walk ends when it comes back to the first face.
  template <class Visit>
  void walk(int v, int start, Visit&& visit) const {
    const int first = m_.foe[2 * start];
    int prev_face = first, prev_edge = start, edge = -1, steps = 0;
    while (edge != start) {
      const int face = across(prev_edge, prev_face);
      if (face == first) break;
      if (++steps >= RING) fail();
      edge = next_edge(v, face, prev_edge);
      visit(face, edge);
      prev_face = face;
      prev_edge = edge;
    }
  }

  int other_end(int e, int v) const { return m_.voe[2 * e] == v ? m_.voe[2 * e + 1] : m_.voe[2 * e]; }

  // get_vertNeighbours remesh_smooth.f90:349-441: the ring of v from e; with
  // reverse, entries 2..n in reverse order
  int neighbours(int v, int e, bool reverse, Ring& out) const {
    out.fill(-1);
    int n = 0;
    out[n++] = other_end(e, v);
    walk(v, e, [&](int, int edge) { out[n++] = other_end(edge, v); });
    if (n < 3) fail();
    if (reverse)
      for (int a = 1, b = n - 1; a < b; ++a, --b) std::swap(out[a], out[b]);
    return n;
  }

  // DGEMM rank-one update Q += p p^T (column j scaled by p_j, rows in order)
  static void rank_one(double* Q, const double* p) {
    for (int j = 0; j < 4; ++j) {
      const double t = 1.0 * p[j];
      for (int i = 0; i < 4; ++i) Q[i + 4 * j] = Q[i + 4 * j] + t * p[i];
    }
  }
  void plane_quadric(double* Q, int face, int v) const {
    const double* n = &m_.nor[3 * face];
    const double xv[3] = {m_.xv(v, 0), m_.xv(v, 1), m_.xv(v, 2)};
    double p[4] = {n[0], n[1], n[2], -dot3(n, xv)};
    const double s = norm2(p, 4);
    for (double& c : p) c = c / s;
    rank_one(Q, p);
  }
  // calc_errorQuadric_of_v remesh_coarsen.f90:260-316
  void quadric(double* Q, int v, int e) const {
    plane_quadric(Q, m_.foe[2 * e], v);
    walk(v, e, [&](int face, int) { plane_quadric(Q, face, v); });
  }

  // DGESV on the 4x4 system: LU with partial pivoting (netlib DGETRF order:
  // first maximal pivot, reciprocal scaling, rank-one updates), then the
  // interchanges and the unit-lower and upper triangular solves of DGETRS
  static void lu_solve4(double* A, double* b) {
    int piv[4];
    for (int j = 0; j < 4; ++j) {
      int p = j;
      double big = std::fabs(A[j + 4 * j]);
      for (int i = j + 1; i < 4; ++i)
        if (std::fabs(A[i + 4 * j]) > big) { big = std::fabs(A[i + 4 * j]); p = i; }
      piv[j] = p;
      if (A[p + 4 * j] == 0.0) fail();   // info /= 0: "Error in QEM system solve"
      if (p != j)
        for (int c = 0; c < 4; ++c) std::swap(A[j + 4 * c], A[p + 4 * c]);
      if (j < 3) {
        if (std::fabs(A[j + 4 * j]) >= 2.2250738585072014e-308) {
          const double r = 1.0 / A[j + 4 * j];
          for (int i = j + 1; i < 4; ++i) A[i + 4 * j] = A[i + 4 * j] * r;
        } else {
          for (int i = j + 1; i < 4; ++i) A[i + 4 * j] = A[i + 4 * j] / A[j + 4 * j];
        }
        for (int c = j + 1; c < 4; ++c) {
          if (A[j + 4 * c] == 0.0) continue;
          const double t = -A[j + 4 * c];
          for (int i = j + 1; i < 4; ++i) A[i + 4 * c] = A[i + 4 * c] + A[i + 4 * j] * t;
        }
      }
    }
    for (int i = 0; i < 4; ++i)
      if (piv[i] != i) std::swap(b[i], b[piv[i]]);
    for (int k = 0; k < 4; ++k)
      if (b[k] != 0.0)
        for (int i = k + 1; i < 4; ++i) b[i] = b[i] - b[k] * A[i + 4 * k];
    for (int k = 3; k >= 0; --k)
      if (b[k] != 0.0) {
        b[k] = b[k] / A[k + 4 * k];
        for (int i = 0; i < k; ++i) b[i] = b[i] - b[k] * A[i + 4 * k];
      }
  }

  void normals_of_flagged() {   // update_1ring_normals remesh_coarsen.f90:581-633
    for (int i = 0; i < RING && flagged_faces_[i] >= 0; ++i) {
      const int f = flagged_faces_[i];
      if (m_.gf[f]) continue;
      double* n = &m_.nor[3 * f];
      const double old[3] = {n[0], n[1], n[2]};
      unit_normal(f, n);
      double s = dot3(n, old);
      if (s < 0.0) std::swap(m_.vof[3 * f], m_.vof[3 * f + 2]);
      s = std::copysign(1.0, s);
      for (int d = 0; d < 3; ++d) n[d] = n[d] * s;
    }
  }

  // Live counts (Euler's V - E + F and the live faces, remesh_coarsen.f90:
  // 93-99, 175-190), the short live edges and the ghost entities of the
  // coarsening, kept as the ghost flags change (every ghosting goes through
  // ghost_v/e/f) instead of counted over all entities after every edit.
  long V_ = 0, E_ = 0, F_ = 0, nshort_ = 0;
  double* el_ = nullptr;
  double thresh_ = 0.0;
  bool held_ = false;             // the holders below are built (at the first collapse)
  Holders vf_, ef_, ve_;          // vertex -> faces (vof), edge -> faces (eof), vertex -> edges (voe)
  void start_counts(double* el, double thresh) {
    el_ = el;
    thresh_ = thresh;
    V_ = E_ = F_ = nshort_ = 0;
    held_ = false;
    for (int i = 0; i < m_.nv; ++i) V_ += !m_.gv[i];
    for (int i = 0; i < m_.ne; ++i) E_ += !m_.ge[i];
    for (int i = 0; i < m_.nf; ++i) F_ += !m_.gf[i];
  }
  void hold() {
    if (held_) return;
    vf_.build(m_.vof, 3, m_.nf, m_.nv);
    ef_.build(m_.eof, 3, m_.nf, m_.ne);
    ve_.build(m_.voe, 2, m_.ne, m_.nv);
    held_ = true;
  }
  // every `from` in the width-wide containers held for it becomes `to`
  static void rewrite(Holders& h, int32_t* a, int width, int from, int to) {
    h.each(from, [&](int c) { for (int k = 0; k < width; ++k) if (a[width * c + k] == from) a[width * c + k] = to; });
    h.move(from, to);
  }
  void ghost_v(int v) { if (!m_.gv[v]) { m_.gv[v] = 1; --V_; } }
  void ghost_e(int e) {
    if (m_.ge[e]) return;
    m_.ge[e] = 1;
    --E_;
    if (el_[e] < thresh_) --nshort_;
  }
  void ghost_f(int f) {
    if (m_.gf[f]) return;
    m_.gf[f] = 1;
    --F_;
  }
  bool euler_two() const { return V_ - E_ + F_ == 2; }
  long live_faces() const { return F_; }

  double edge_length(int e) const {
    const int a = m_.voe[2 * e], b = m_.voe[2 * e + 1];
    const double dx = m_.xv(a, 0) - m_.xv(b, 0), dy = m_.xv(a, 1) - m_.xv(b, 1), dz = m_.xv(a, 2) - m_.xv(b, 2);
    return std::sqrt(dx * dx + dy * dy + dz * dz);
  }
  // the live edges that hold v (after a collapse onto v: every edge whose
  // length changed): new lengths, short count
  void remeasure_at(int v) {
    ve_.each(v, [&](int e) {
      if (m_.ge[e] || (m_.voe[2 * e] != v && m_.voe[2 * e + 1] != v)) return;
      const bool was = el_[e] < thresh_;
      el_[e] = edge_length(e);
      nshort_ += (long)(el_[e] < thresh_) - (long)was;
    });
  }

  // the two edges of face f other than e: (ghosted, kept)
  void others(int f, int e, int& ghost, int& kept) const {
    int b[3] = {-1, -1, -1}, n = 0;
    for (int i = 0; i < 3; ++i)
      if (m_.eof[3 * f + i] != e) b[n++] = m_.eof[3 * f + i];
    ghost = b[0];
    kept = b[1];
  }

  // the faces of `kept` after its twin `gone` is merged into it, less `face`
  void merge(int gone, int kept, int face) {
    const int b[4] = {m_.foe[2 * gone], m_.foe[2 * gone + 1], m_.foe[2 * kept], m_.foe[2 * kept + 1]};
    int n = 0;
    for (int f : b) {
      if (f == face) continue;
      if (n == 2) fail();
      m_.foe[2 * kept + n++] = f;
    }
  }

  // one edge collapse (remesh_coarsen.f90:127-184)
  void collapse(int e, int v1, int v2, int32_t* anchor, int32_t* flagged_edge, double* vmelt) {
    double Q1[16] = {}, Q2[16] = {};
    quadric(Q1, v1, e);
    quadric(Q2, v2, e);
    for (int i = 0; i < 16; ++i) Q1[i] = Q1[i] + Q2[i];     // solve_QEM_system :318-341
    Q1[3] = 0.0; Q1[7] = 0.0; Q1[11] = 0.0; Q1[15] = 1.0;
    double b[4] = {0.0, 0.0, 0.0, 1.0};
    lu_solve4(Q1, b);
    for (int d = 0; d < 3; ++d) m_.x[3 * v1 + d] = b[d];
    if (vmelt)
      for (int d = 0; d < 3; ++d) vmelt[3 * v1 + d] = 0.5 * (vmelt[3 * v1 + d] + vmelt[3 * v2 + d]);
    const int F1 = m_.foe[2 * e], F2 = m_.foe[2 * e + 1];
    int e1, re1, e2, re2;
    others(F1, e, e1, re1);
    others(F2, e, e2, re2);
    ghost_f(F1); ghost_f(F2);
    ghost_v(v2);
    ghost_e(e); ghost_e(e1); ghost_e(e2);
    // update_face_connectivity :368-381, update_edge_connectivity :383-421:
    // every v2 -> v1, e1 -> re1, e2 -> re2 in vof / eof / voe, in the
    // containers that hold them (the reference loops over all of them)
    rewrite(vf_, m_.vof, 3, v2, v1);
    rewrite(ef_, m_.eof, 3, e1, re1);
    rewrite(ef_, m_.eof, 3, e2, re2);
    rewrite(ve_, m_.voe, 2, v2, v1);
    merge(e1, re1, F1);
    merge(e2, re2, F2);
    // flag_neighbours_1ring :454-513 around v1 from re1
    anchor[v1] = 0;
    flagged_edge[re1] = 1;
    flagged_faces_.fill(-1);
    int nf = 0;
    flagged_faces_[nf++] = m_.foe[2 * re1];
    walk(v1, re1, [&](int face, int edge) {
      flagged_faces_[nf++] = face;
      anchor[m_.voe[2 * edge]] = 0;
      anchor[m_.voe[2 * edge + 1]] = 0;
      flagged_edge[edge] = 1;
    });
  }

  // valence-3 deletion (get_edge_face_1ring :515-579 and
  // update_connectivity_vert_deletion :637-760): v (3 neighbours, the ring
  // `nb` of `val` entries) goes with its three edges and two of its faces;
  // the first face of e becomes (k, v3, v4)
  void delete_vertex(int v, int k, int e, const Ring& nb, int val) {
    Ring edges, faces;
    edges.fill(-1);
    faces.fill(-1);
    int ne = 0, nf = 0;
    edges[ne++] = e;
    faces[nf++] = m_.foe[2 * e];
    walk(v, e, [&](int face, int edge) { faces[nf++] = face; edges[ne++] = edge; });
    if (ne != 3 || nf != 3) fail();
    int le[2], lv[2], n = 0;
    for (int i = 0; i < ne; ++i) if (edges[i] != e) le[n++] = edges[i];
    n = 0;
    for (int i = 0; i < val; ++i) {
      if (nb[i] == k) continue;
      if (n == 2) fail();
      lv[n++] = nb[i];
    }
    if (n != 2) fail();
    const int e1 = le[0], e2 = le[1];
    const int fL = m_.foe[2 * e], fR = m_.foe[2 * e + 1];
    int f3 = -1, e3 = -1, e4 = -1, e5 = -1;
    for (int i = 0; i < nf; ++i) if (faces[i] != fL && faces[i] != fR) f3 = faces[i];
    if (f3 < 0) fail();
    for (int i = 0; i < 3; ++i) {
      const int a = m_.eof[3 * fL + i], b = m_.eof[3 * fR + i], c = m_.eof[3 * f3 + i];
      if (a != e && a != e1 && a != e2) e3 = a;
      if (b != e && b != e1 && b != e2) e4 = b;
      if (c != e1 && c != e2) e5 = c;
    }
    if (e3 < 0 || e4 < 0 || e5 < 0) fail();
    for (int i = 0; i < 2; ++i) {
      for (int ed : {e3, e4, e5})
        if (m_.foe[2 * ed + i] == fR || m_.foe[2 * ed + i] == f3) m_.foe[2 * ed + i] = fL;
    }
    n = 0;
    for (int i = 0; i < 3; ++i) {
      if (m_.vof[3 * fL + i] == k) continue;
      if (n == 2) fail();
      if (held_) vf_.add(lv[n], fL);
      m_.vof[3 * fL + i] = lv[n++];
    }
    const int repl[2] = {e4, e5};
    n = 0;
    for (int i = 0; i < 3; ++i) {
      if (m_.eof[3 * fL + i] == e3) continue;
      if (n == 2) fail();
      if (held_) ef_.add(repl[n], fL);
      m_.eof[3 * fL + i] = repl[n++];
    }
    ghost_v(v);
    ghost_e(e); ghost_e(e1); ghost_e(e2);
    ghost_f(fR); ghost_f(f3);
    flagged_faces_[0] = fL;
  }

  // one edge relaxation (remesh_smooth.f90:105-204)
  void relax(int e, bool& corrected, double target_dv, double vol_old, double& drift, int& nrefresh) {
    const int v1 = m_.voe[2 * e], v2 = m_.voe[2 * e + 1];
    Ring n1, n2;
    const int val1 = neighbours(v1, e, true, n1), val2 = neighbours(v2, e, false, n2);
    double A1[3] = {0.0, 0.0, 0.0}, A2[3] = {0.0, 0.0, 0.0}, vec[3] = {0.0, 0.0, 0.0}, ej[3], ej1[3], c[3];
    for (int j = 0; j < val1; ++j) {
      const int a = n1[j], b = n1[(j + 1) % val1];
      for (int d = 0; d < 3; ++d) { ej[d] = m_.xv(a, d) - m_.xv(v1, d); ej1[d] = m_.xv(b, d) - m_.xv(v1, d); }
      cross(c, ej, ej1);
      for (int d = 0; d < 3; ++d) A1[d] = A1[d] + c[d];
    }
    for (int j = 0; j < val2; ++j) {
      const int a = n2[j], b = n2[(j + 1) % val2];
      for (int d = 0; d < 3; ++d) { ej[d] = m_.xv(a, d) - m_.xv(v2, d); ej1[d] = m_.xv(b, d) - m_.xv(v2, d); }
      cross(c, ej, ej1);
      for (int d = 0; d < 3; ++d) A2[d] = A2[d] + c[d];
      if (j == val2 - 1) for (int d = 0; d < 3; ++d) vec[d] = vec[d] + ej[d];
      else if (j == 1) for (int d = 0; d < 3; ++d) vec[d] = vec[d] - ej[d];
    }
    double s1[3], s2[3], d1[3], d2[3], A[3], nh[3], dd[3];
    const double omega = 1.0, tiny = 1.e-15;
    for (int d = 0; d < 3; ++d) {
      double sum2 = 0.0, sum1 = 0.0;
      for (int q = 1; q < val2; ++q) sum2 = sum2 + m_.xv(n2[q], d);
      for (int q = 1; q < val1; ++q) sum1 = sum1 + m_.xv(n1[q], d);
      s1[d] = (sum2 + (double)val2 * sum1) / ((double)(val1 * val2) - 1.0);
    }
    for (int d = 0; d < 3; ++d) {
      double sum2 = 0.0;
      for (int q = 1; q < val2; ++q) sum2 = sum2 + m_.xv(n2[q], d);
      s2[d] = (s1[d] + sum2) / (double)val2;
    }
    for (int d = 0; d < 3; ++d) { d1[d] = omega * (s1[d] - m_.xv(v1, d)); d2[d] = omega * (s2[d] - m_.xv(v2, d)); }
    for (int d = 0; d < 3; ++d) dd[d] = d1[d] - d2[d];
    cross(c, vec, dd);
    for (int d = 0; d < 3; ++d) A[d] = A1[d] + A2[d] + c[d];
    const double nA = norm2(A, 3);
    for (int d = 0; d < 3; ++d) nh[d] = A[d] / (nA + tiny);
    cross(c, vec, d1);
    double h = -(dot3(d1, A1) + dot3(d2, A2) + dot3(d2, c));
    if (!corrected) h = h + std::copysign(1.0, h) * target_dv * 6.0;
    h = h / (nA + tiny);
    double t1[3], t2[3];
    for (int d = 0; d < 3; ++d) { t1[d] = d1[d] + h * nh[d]; t2[d] = d2[d] + h * nh[d]; }
    const double m1 = norm2(t1, 3), m2 = norm2(t2, 3);
    if (m1 > drift) drift = m1;
    if (m2 > drift) drift = m2;
    for (int d = 0; d < 3; ++d) {
      m_.x[3 * v1 + d] = m_.x[3 * v1 + d] + d1[d] + h * nh[d];
      m_.x[3 * v2 + d] = m_.x[3 * v2 + d] + d2[d] + h * nh[d];
    }
    if (!corrected) {
      if (std::fabs(volume() - vol_old) < 1.e-14) corrected = true;
      else ++nrefresh;
    }
  }
};

// 1-based reference arrays <-> 0-based working copies
struct Copy0 {
  std::vector<int32_t> vof, eof, voe, foe;
  Copy0(const int32_t* vof1, const int32_t* eof1, const int32_t* voe1, const int32_t* foe1, int ne, int nf)
      : vof(vof1, vof1 + 3 * (size_t)nf), eof(eof1, eof1 + 3 * (size_t)nf), voe(voe1, voe1 + 2 * (size_t)ne), foe(foe1, foe1 + 2 * (size_t)ne) {
    for (auto* v : {&vof, &eof, &voe, &foe})
      for (auto& i : *v) i -= 1;
  }
  static void back(const std::vector<int32_t>& a, int32_t* out) {
    for (size_t i = 0; i < a.size(); ++i) out[i] = a[i] + 1;
  }
};

bool valid_ids(const int32_t* a, size_t n, int hi) {
  for (size_t i = 0; i < n; ++i)
    if (a[i] < 1 || a[i] > hi) return false;
  return true;
}

}  // namespace melt_remesh

using namespace melt_remesh;

// ------------------------------------------------- reference-form entries
extern "C" int melt_remesh_coarsen(int32_t nf, int32_t ne, int32_t nv, double* xyz, double* tri_nor, double* e_lengths, double e_thresh,
                                   int32_t* vert_of_face, int32_t* edge_of_face, int32_t* vert_of_edge, int32_t* face_of_edge,
                                   int32_t* is_ghost_face, int32_t* is_ghost_edge, int32_t* is_ghost_vert, int32_t* rm_flag,
                                   int32_t* anchor_vert, int32_t* flagged_edge, double* vmelt, int32_t* ecol_cnt, int32_t* vdel_cnt) {
  if (nf <= 0 || ne <= 0 || nv <= 0 || !xyz || !tri_nor || !e_lengths || !vert_of_face || !edge_of_face || !vert_of_edge || !face_of_edge ||
      !is_ghost_face || !is_ghost_edge || !is_ghost_vert || !rm_flag || !anchor_vert || !flagged_edge)
    return MELT_ERR_ARG;
  if (!valid_ids(vert_of_face, 3 * (size_t)nf, nv) || !valid_ids(edge_of_face, 3 * (size_t)nf, ne) || !valid_ids(vert_of_edge, 2 * (size_t)ne, nv) ||
      !valid_ids(face_of_edge, 2 * (size_t)ne, nf))
    return MELT_ERR_ARG;
  Copy0 c(vert_of_face, edge_of_face, vert_of_edge, face_of_edge, ne, nf);
  Surface s;
  s.nv = nv; s.ne = ne; s.nf = nf; s.x = xyz; s.nor = tri_nor;
  s.vof = c.vof.data(); s.eof = c.eof.data(); s.voe = c.voe.data(); s.foe = c.foe.data();
  s.gv = is_ghost_vert; s.ge = is_ghost_edge; s.gf = is_ghost_face;
  Editor ed(s);
  int ncol = 0, ndel = 0;
  const bool ok = ed.coarsen(e_lengths, e_thresh, *rm_flag, anchor_vert, flagged_edge, vmelt, ncol, ndel);
  Copy0::back(c.vof, vert_of_face); Copy0::back(c.eof, edge_of_face); Copy0::back(c.voe, vert_of_edge); Copy0::back(c.foe, face_of_edge);
  if (ecol_cnt) *ecol_cnt = ncol;
  if (vdel_cnt) *vdel_cnt = ndel;
  return ok ? 0 : MELT_ERR_REMESH;
}

extern "C" int melt_remesh_smooth(double vol_old, double target_dv, int32_t nv, int32_t ne, int32_t nf, double* xyz, const int32_t* is_ghost_vert,
                                  const int32_t* is_ghost_edge, const int32_t* is_ghost_face, int32_t* flagged_edge, const int32_t* vert_of_edge,
                                  const int32_t* vert_of_face, const int32_t* face_of_edge, const int32_t* edge_of_face, int32_t* n_erel,
                                  double* drift, int32_t* cnt_refresh) {
  if (nf <= 0 || ne <= 0 || nv <= 0 || !xyz || !is_ghost_vert || !is_ghost_edge || !is_ghost_face || !flagged_edge || !vert_of_edge ||
      !vert_of_face || !face_of_edge || !edge_of_face)
    return MELT_ERR_ARG;
  if (!valid_ids(vert_of_face, 3 * (size_t)nf, nv) || !valid_ids(edge_of_face, 3 * (size_t)nf, ne) || !valid_ids(vert_of_edge, 2 * (size_t)ne, nv) ||
      !valid_ids(face_of_edge, 2 * (size_t)ne, nf))
    return MELT_ERR_ARG;
  Copy0 c(vert_of_face, edge_of_face, vert_of_edge, face_of_edge, ne, nf);
  std::vector<int32_t> gv(is_ghost_vert, is_ghost_vert + nv), ge(is_ghost_edge, is_ghost_edge + ne), gf(is_ghost_face, is_ghost_face + nf);
  Surface s;
  s.nv = nv; s.ne = ne; s.nf = nf; s.x = xyz;
  s.vof = c.vof.data(); s.eof = c.eof.data(); s.voe = c.voe.data(); s.foe = c.foe.data();
  s.gv = gv.data(); s.ge = ge.data(); s.gf = gf.data();
  Editor ed(s);
  int nrel = 0, nref = 0;
  double dr = 0.0;
  const bool ok = ed.smooth(vol_old, target_dv, flagged_edge, nrel, dr, nref);
  if (n_erel) *n_erel = nrel;
  if (drift) *drift = dr;
  if (cnt_refresh) *cnt_refresh = nref;
  return ok ? 0 : MELT_ERR_REMESH;
}

// --------------------------------------------------- the context's body
namespace {

// vertex -> live faces in ascending face order (the order of the reference's
// face loops that accumulate into vertices)
void vertex_faces(const melt_ctx* ctx, std::vector<int>& ptr, std::vector<int>& lst) {
  ptr.assign(ctx->nv + 1, 0);
  lst.assign(3 * (size_t)ctx->nf, 0);
  for (int f = 0; f < ctx->nf; ++f)
    if (!ctx->h_gf[f])
      for (int c = 0; c < 3; ++c) ptr[ctx->h_vof[3 * f + c] + 1]++;
  for (int v = 0; v < ctx->nv; ++v) ptr[v + 1] += ptr[v];
  std::vector<int> at(ptr.begin(), ptr.end() - 1);
  for (int f = 0; f < ctx->nf; ++f)
    if (!ctx->h_gf[f])
      for (int c = 0; c < 3; ++c) lst[at[ctx->h_vof[3 * f + c]]++] = f;
}

int push_flags(melt_ctx* ctx, const std::vector<int32_t>& h, unsigned char* d, hipStream_t s, std::vector<unsigned char>& tmp) {
  tmp.resize(h.size());
  for (size_t i = 0; i < h.size(); ++i) tmp[i] = h[i] ? 1 : 0;
  MELT_HIP(hipMemcpyAsync(d, tmp.data(), tmp.size(), hipMemcpyHostToDevice, s));
  MELT_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace

// particle.f90:121-160 on the context's body, queued on ctx->stream after the
// edge flag: positions (now and after the Stefan update) and melt velocities
// come down, the host coarsens and smooths, and the edited arrays go back.
int remesh_context(melt_ctx* ctx, int* did) {
  *did = 0;
  hipStream_t s = ctx->stream;
  const int nv = ctx->nv, ne = ctx->ne, nf = ctx->nf;
  std::vector<double> x(3 * (size_t)nv), xm(3 * (size_t)nv), vm(3 * (size_t)nv), nor(3 * (size_t)nf, 0.0), el(ne, 0.0);
  MELT_HIP(hipMemcpyAsync(x.data(), ctx->d_lag[MELT_L_XYZV], sizeof(double) * x.size(), hipMemcpyDeviceToHost, s));
  MELT_HIP(hipMemcpyAsync(vm.data(), ctx->d_lag[MELT_L_VMELT], sizeof(double) * vm.size(), hipMemcpyDeviceToHost, s));
  MELT_HIP(hipMemcpyAsync(el.data(), ctx->d_lag[MELT_L_ELENGTHS], sizeof(double) * el.size(), hipMemcpyDeviceToHost, s));
  if (ctx->cfg.imelt == 1) MELT_HIP(hipMemcpyAsync(xm.data(), ctx->d_xyz_melt, sizeof(double) * xm.size(), hipMemcpyDeviceToHost, s));
  MELT_HIP(hipStreamSynchronize(s));
  Surface sf;
  sf.nv = nv; sf.ne = ne; sf.nf = nf; sf.x = x.data(); sf.nor = nor.data();
  sf.vof = ctx->h_vof.data(); sf.eof = ctx->h_eof.data(); sf.voe = ctx->h_voe.data(); sf.foe = ctx->h_foe.data();
  sf.gv = ctx->h_gv.data(); sf.ge = ctx->h_ge.data(); sf.gf = ctx->h_gf.data();
  Editor ed(sf);
  const double thresh = ctx->h_body.e_thresh;
  if (!ed.measure(el.data(), thresh)) return 0;   // calculate_eLengths on the host agrees that no edge is short
  *did = 1;
  // vol_melt (particle.f90:86-90): zero unless the body melts
  double vol_melt = 0.0;
  if (ctx->cfg.imelt == 1) {
    Surface sm = sf;
    sm.x = xm.data();
    vol_melt = Editor(sm).volume();
  }
  ed.face_normals();
  int32_t rm = 1;
  int ncol = 0, ndel = 0, nrel = 0, nref = 0;
  double drift = 0.0;
  if (!ed.coarsen(el.data(), thresh, rm, ctx->h_anchor.data(), ctx->h_flagged.data(), vm.data(), ncol, ndel))
    return melt_fail(ctx, MELT_ERR_REMESH, "remesh_coarsen", "topology not preserved, geometry reduced to a tetrahedron or QEM solve failed");
  const double vol_coarse = ed.volume();
  const double target = -(vol_coarse - vol_melt);
  if (!ed.smooth(vol_melt, target, ctx->h_flagged.data(), nrel, drift, nref))
    return melt_fail(ctx, MELT_ERR_REMESH, "remesh_smooth", "vertex ring walk failed");
  const double vol_smooth = ed.volume();
  double* st = ctx->remesh_stats;
  st[0] += 1.0;
  st[1] = ncol; st[2] = ndel; st[3] = nrel; st[4] = drift; st[5] = vol_melt - vol_smooth; st[6] = nref;
  st[7] += ncol;
  // back to the device: positions, melt velocities, connectivity, ghost flags, vertex -> face lists
  std::vector<int> ptr, lst;
  vertex_faces(ctx, ptr, lst);
  std::vector<unsigned char> tmp;
  MELT_HIP(hipMemcpyAsync(ctx->d_lag[MELT_L_XYZV], x.data(), sizeof(double) * x.size(), hipMemcpyHostToDevice, s));
  MELT_HIP(hipMemcpyAsync(ctx->d_lag[MELT_L_VMELT], vm.data(), sizeof(double) * vm.size(), hipMemcpyHostToDevice, s));
  MELT_HIP(hipMemcpyAsync(ctx->d_vof, ctx->h_vof.data(), sizeof(int) * 3 * (size_t)nf, hipMemcpyHostToDevice, s));
  MELT_HIP(hipMemcpyAsync(ctx->d_eof, ctx->h_eof.data(), sizeof(int) * 3 * (size_t)nf, hipMemcpyHostToDevice, s));
  MELT_HIP(hipMemcpyAsync(ctx->d_voe, ctx->h_voe.data(), sizeof(int) * 2 * (size_t)ne, hipMemcpyHostToDevice, s));
  MELT_HIP(hipMemcpyAsync(ctx->d_foe, ctx->h_foe.data(), sizeof(int) * 2 * (size_t)ne, hipMemcpyHostToDevice, s));
  MELT_HIP(hipMemcpyAsync(ctx->d_v2f_ptr, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice, s));
  MELT_HIP(hipMemcpyAsync(ctx->d_v2f, lst.data(), sizeof(int) * lst.size(), hipMemcpyHostToDevice, s));
  MELT_HIP(hipStreamSynchronize(s));
  if (int rc = push_flags(ctx, ctx->h_gv, ctx->d_gv, s, tmp)) return rc;
  if (int rc = push_flags(ctx, ctx->h_ge, ctx->d_ge, s, tmp)) return rc;
  if (int rc = push_flags(ctx, ctx->h_gf, ctx->d_gf, s, tmp)) return rc;
  if (std::fabs(vol_melt - vol_smooth) > 1.e-14) {   // particle.f90:199-211
    char msg[192];
    std::snprintf(msg, sizeof msg, "non-zero remeshing residual %.3e (volume %.6e, %d collapses, %d deletions, %d relaxations, %d refreshes, drift %.3e)",
                  vol_melt - vol_smooth, vol_melt, ncol, ndel, nrel, nref, drift);
    return melt_fail(ctx, MELT_ERR_REMESH, "remesh_smooth", msg);
  }
  return 0;
}

int remesh_pull_mirrors(melt_ctx* ctx) {
  if (!ctx->has_geom) return 0;
  const size_t nv = ctx->nv, ne = ctx->ne, nf = ctx->nf;
  std::vector<unsigned char> g(std::max(nv, std::max(ne, nf)));
  MELT_HIP(hipMemcpy(ctx->h_vof.data(), ctx->d_vof, sizeof(int) * 3 * nf, hipMemcpyDeviceToHost));
  MELT_HIP(hipMemcpy(ctx->h_eof.data(), ctx->d_eof, sizeof(int) * 3 * nf, hipMemcpyDeviceToHost));
  MELT_HIP(hipMemcpy(ctx->h_voe.data(), ctx->d_voe, sizeof(int) * 2 * ne, hipMemcpyDeviceToHost));
  MELT_HIP(hipMemcpy(ctx->h_foe.data(), ctx->d_foe, sizeof(int) * 2 * ne, hipMemcpyDeviceToHost));
  struct { unsigned char* d; std::vector<int32_t>* h; size_t n; } fl[3] = {{ctx->d_gv, &ctx->h_gv, nv}, {ctx->d_ge, &ctx->h_ge, ne}, {ctx->d_gf, &ctx->h_gf, nf}};
  for (auto& f : fl) {
    MELT_HIP(hipMemcpy(g.data(), f.d, f.n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < f.n; ++i) (*f.h)[i] = g[i];
  }
  std::fill(ctx->h_flagged.begin(), ctx->h_flagged.end(), 0);
  return 0;
}

extern "C" int melt_get_topology(melt_ctx* ctx, int32_t* vert_of_face, int32_t* edge_of_face, int32_t* vert_of_edge, int32_t* face_of_edge,
                                 int32_t* is_ghost_vert, int32_t* is_ghost_edge, int32_t* is_ghost_face) {
  if (!ctx) return MELT_ERR_ARG;
  if (!ctx->has_geom) return MELT_ERR_NO_GEOMETRY;
  auto out1 = [](const std::vector<int32_t>& h, int32_t* o) { if (o) for (size_t i = 0; i < h.size(); ++i) o[i] = h[i] + 1; };
  auto out0 = [](const std::vector<int32_t>& h, int32_t* o) { if (o) std::memcpy(o, h.data(), sizeof(int32_t) * h.size()); };
  out1(ctx->h_vof, vert_of_face); out1(ctx->h_eof, edge_of_face); out1(ctx->h_voe, vert_of_edge); out1(ctx->h_foe, face_of_edge);
  out0(ctx->h_gv, is_ghost_vert); out0(ctx->h_ge, is_ghost_edge); out0(ctx->h_gf, is_ghost_face);
  return 0;
}

extern "C" int melt_get_remesh_stats(melt_ctx* ctx, double* out) {
  if (!ctx || !out) return MELT_ERR_ARG;
  std::memcpy(out, ctx->remesh_stats, sizeof ctx->remesh_stats);
  return 0;
}
