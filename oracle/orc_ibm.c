/*
 * ORACLE (test infrastructure only) -- immersed-boundary, tagging and Lagrangian
 * part of the RK3 substep.  See orc.h for the rules this restatement follows.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

#define F(id, i, j, k) OF(o, id, i, j, k)
#define LAG(id) (o->lag[id])
#define V3(a, i) (&(a)[3 * (size_t)(i)])   /* 0-based element i of a (3,n) array */

/* ------------------------------------------------------------ own sine */
/* Cody-Waite reduction by pi/2 (three-part constant) and Taylor polynomials on
 * [-pi/4, pi/4] evaluated with explicit fma; valid for |x| <= 8. */
double orc_det_sin(double x) {
  const double two_over_pi = 0.63661977236758134308;
  const double p1 = 1.5707963267341256;   /* pi/2 to 33 bits */
  const double p2 = 6.077100506303966e-11;   /* next 33 bits of pi/2 */
  const double p3 = 2.0222662487959506e-21;   /* remainder */
  double kd = rint(x * two_over_pi);
  double y = ((x - kd * p1) - kd * p2) - kd * p3;
  double z = y * y;
  double s = 1.0 / 355687428096000.0;
  s = fma(s, z, -1.0 / 1307674368000.0);
  s = fma(s, z, 1.0 / 6227020800.0);
  s = fma(s, z, -1.0 / 39916800.0);
  s = fma(s, z, 1.0 / 362880.0);
  s = fma(s, z, -1.0 / 5040.0);
  s = fma(s, z, 1.0 / 120.0);
  s = fma(s, z, -1.0 / 6.0);
  double sy = fma(y * z, s, y);
  double c = -1.0 / 6402373705728000.0;
  c = fma(c, z, 1.0 / 20922789888000.0);
  c = fma(c, z, -1.0 / 87178291200.0);
  c = fma(c, z, 1.0 / 479001600.0);
  c = fma(c, z, -1.0 / 3628800.0);
  c = fma(c, z, 1.0 / 40320.0);
  c = fma(c, z, -1.0 / 720.0);
  c = fma(c, z, 1.0 / 24.0);
  c = fma(c, z, -0.5);
  double cy = fma(z, c, 1.0);
  int q = ((int)kd) & 3;
  return q == 0 ? sy : q == 1 ? cy : q == 2 ? -sy : -cy;
}

/* Own exp for the MLS Gaussian weights (argument in [-12, 0]): reduction by
 * ln2 with a two-part constant, Taylor polynomial of degree 13 (explicit fma),
 * exact scaling by 2^k. */
double orc_det_exp(double x) {
  const double log2e = 1.4426950408889634;
  const double ln2_hi = 0.6931471803691238;      /* ln2 to 32 bits */
  const double ln2_lo = 1.9082149292705877e-10;
  double kd = rint(x * log2e);
  double r = (x - kd * ln2_hi) - kd * ln2_lo;
  double p = 1.0 / 6227020800.0;
  p = fma(p, r, 1.0 / 479001600.0);
  p = fma(p, r, 1.0 / 39916800.0);
  p = fma(p, r, 1.0 / 3628800.0);
  p = fma(p, r, 1.0 / 362880.0);
  p = fma(p, r, 1.0 / 40320.0);
  p = fma(p, r, 1.0 / 5040.0);
  p = fma(p, r, 1.0 / 720.0);
  p = fma(p, r, 1.0 / 120.0);
  p = fma(p, r, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)kd);
}

/* ---------------------------------------------- LAPACK restatements */
/* dgetf2 + dtrti2 + dgetri (netlib reference, unblocked; column-major n x n).
 * Returns info of dgetrf (first zero pivot, 1-based) or 0. */
static int invn(double* a, int n) {
#define A(i, j) a[(i) + n * (j)]
  int ipiv[4], info = 0;
  for (int j = 0; j < n; ++j) {
    int jp = j;
    double amax = fabs(A(j, j));
    for (int i = j + 1; i < n; ++i)
      if (fabs(A(i, j)) > amax) { amax = fabs(A(i, j)); jp = i; }
    ipiv[j] = jp;
    if (A(jp, j) != 0.0) {
      if (jp != j)
        for (int c = 0; c < n; ++c) { double t = A(j, c); A(j, c) = A(jp, c); A(jp, c) = t; }
      if (j < n - 1) {
        if (fabs(A(j, j)) >= 2.2250738585072014e-308) {
          double r = 1.0 / A(j, j);
          for (int i = j + 1; i < n; ++i) A(i, j) = A(i, j) * r;
        } else {
          for (int i = j + 1; i < n; ++i) A(i, j) = A(i, j) / A(j, j);
        }
      }
    } else if (info == 0) {
      info = j + 1;
    }
    if (j < n - 1)
      for (int jj = j + 1; jj < n; ++jj) {
        if (A(j, jj) != 0.0) {
          double temp = -A(j, jj);
          for (int ii = j + 1; ii < n; ++ii) A(ii, jj) = A(ii, jj) + A(ii, j) * temp;
        }
      }
  }
  if (info) return info;
  /* dtrti2: inverse of upper triangle, non-unit */
  for (int j = 0; j < n; ++j) {
    A(j, j) = 1.0 / A(j, j);
    double ajj = -A(j, j);
    /* dtrmv upper, no-trans on x = A(0:j-1, j) with T = A(0:j-1,0:j-1) */
    for (int jj = 0; jj < j; ++jj) {
      if (A(jj, j) != 0.0) {
        double temp = A(jj, j);
        for (int i = 0; i < jj; ++i) A(i, j) = A(i, j) + temp * A(i, jj);
        A(jj, j) = A(jj, j) * A(jj, jj);
      }
    }
    for (int i = 0; i < j; ++i) A(i, j) = ajj * A(i, j);
  }
  /* dgetri unblocked: solve inv(A)*L = inv(U) */
  double work[4];
  for (int j = n - 1; j >= 0; --j) {
    for (int i = j + 1; i < n; ++i) { work[i] = A(i, j); A(i, j) = 0.0; }
    if (j < n - 1) {
      for (int jj = j + 1; jj < n; ++jj) {
        if (work[jj] != 0.0) {
          double temp = -1.0 * work[jj];
          for (int i = 0; i < n; ++i) A(i, j) = A(i, j) + temp * A(i, jj);
        }
      }
    }
  }
  for (int j = n - 2; j >= 0; --j) {
    int jp = ipiv[j];
    if (jp != j)
      for (int i = 0; i < n; ++i) { double t = A(i, j); A(i, j) = A(i, jp); A(i, jp) = t; }
  }
  return 0;
#undef A
}
int orc_inv4(double a[16]) { return invn(a, 4); }
int orc_inv3(double a[9]) { return invn(a, 3); }

/* DGEMM('N','T',4,4,1,w,p,4,p,4,1,A,4): A(i,j) += (w*p(j))*p(i)  (netlib order) */
static void rank1(double A[16], double w, const double p[4]) {
  for (int j = 0; j < 4; ++j) {
    double temp = w * p[j];
    for (int i = 0; i < 4; ++i) A[i + 4 * j] = A[i + 4 * j] + temp * p[i];
  }
}

/* ----------------------------------------------------------- geom.f90 */
static double mls_gaussian(double r, double rcoef) {  /* geom.f90:24-36 */
  if (r > 1.0) return 0.0;
  double t = r / rcoef;
  return orc_det_exp(-(t * t));
}
static double mls_gauss_deriv(double r, double rcoef) {  /* geom.f90:38-53 */
  if (r > 1.0) return 0.0;
  double t = r / rcoef;
  return -2.0 * r / (rcoef * rcoef) * orc_det_exp(-(t * t));
}
static const double WCON = 0.3, WSCL = 1.5;  /* initmls.f90:12-13 */

/* ==================================================================== tagging */
static void get_bbox_inds(orc_ctx* o, int ind[3][2]) {  /* tagCells.f90:31-62 */
  double lim[3][2];
  for (int d = 0; d < 3; ++d) {
    double mn = 0, mx = 0;
    int first = 1;
    for (int v = 0; v < o->nv; ++v) {
      if (o->ghost_v[v]) continue;
      double x = LAG(MELT_L_XYZV)[3 * v + d];
      if (first) { mn = mx = x; first = 0; }
      else { if (x < mn) mn = x; if (x > mx) mx = x; }
    }
    lim[d][0] = mn; lim[d][1] = mx;
  }
  for (int d = 0; d < 3; ++d) {
    ind[d][0] = (int)floor(lim[d][0] * o->dx1) + 1 - 2;
    ind[d][1] = (int)floor(lim[d][1] * o->dx1) + 1 + 2;
  }
}

static int edge_intersects_square(double y1, double z1, double y2, double z2, double ymin, double ymax, double zmin, double zmax) {
  double t, y, z;  /* rayAux.f90:139-190 */
  if (y1 != y2) {
    t = (ymin - y1) / (y2 - y1);
    if (t >= 0.0 && t <= 1.0) { z = z1 + t * (z2 - z1); if (z >= zmin && z <= zmax) return 1; }
    t = (ymax - y1) / (y2 - y1);
    if (t >= 0.0 && t <= 1.0) { z = z1 + t * (z2 - z1); if (z >= zmin && z <= zmax) return 1; }
  }
  if (z1 != z2) {
    t = (zmin - z1) / (z2 - z1);
    if (t >= 0.0 && t <= 1.0) { y = y1 + t * (y2 - y1); if (y >= ymin && y <= ymax) return 1; }
    t = (zmax - z1) / (z2 - z1);
    if (t >= 0.0 && t <= 1.0) { y = y1 + t * (y2 - y1); if (y >= ymin && y <= ymax) return 1; }
  }
  return 0;
}

static int intersects_yz_range(const double* V0, const double* V1, const double* V2, double ymin, double ymax, double zmin, double zmax) {
  double yv[3] = {V0[1], V1[1], V2[1]}, zv[3] = {V0[2], V1[2], V2[2]};  /* rayAux.f90:86-113 */
  for (int i = 0; i < 3; ++i)
    if (yv[i] >= ymin && yv[i] <= ymax && zv[i] >= zmin && zv[i] <= zmax) return 1;
  double ymx = fmax(fmax(yv[0], yv[1]), yv[2]), ymn = fmin(fmin(yv[0], yv[1]), yv[2]);
  double zmx = fmax(fmax(zv[0], zv[1]), zv[2]), zmn = fmin(fmin(zv[0], zv[1]), zv[2]);
  if (ymx < ymin || ymn > ymax || zmx < zmin || zmn > zmax) return 0;
  for (int i = 0; i < 3; ++i) {  /* triangle_square_edges_intersection rayAux.f90:115-137 */
    int j = (i + 1) % 3;
    if (edge_intersects_square(yv[i], zv[i], yv[j], zv[j], ymin, ymax, zmin, zmax)) return 1;
  }
  return 0;
}

static int triangleXIntersect(const double* V0, const double* V1, const double* V2, double xmin, double xmax) {
  double mn = fmin(fmin(V0[0], V1[0]), V2[0]), mx = fmax(fmax(V0[0], V1[0]), V2[0]);  /* rayAux.f90:192-218 */
  if (mx < xmin || mn > xmax) return 0;
  return (V0[0] >= xmin && V0[0] <= xmax) || (V1[0] >= xmin && V1[0] <= xmax) || (V2[0] >= xmin && V2[0] <= xmax);
}

/* rayAux.f90:9-83, recursive Moller-Trumbore.  depth guards the recursion the
 * reference leaves unbounded. */
static void ray_tri(orc_ctx* o, int* intersect, double ip[3], const double C[3], const double Q[3], const double* V0, const double* V1,
                    const double* V2, int depth) {
  const double EPS = 1.e-18;
  double r[3], e1[3], e2[3], s[3], p[3], q1[3];
  *intersect = 1;
  for (int d = 0; d < 3; ++d) { r[d] = Q[d] - C[d]; e1[d] = V1[d] - V0[d]; e2[d] = V2[d] - V0[d]; s[d] = C[d] - V0[d]; }
  orc_cross(p, r, e2);
  double a = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
  if (a > -EPS && a <= EPS) {
    if (depth >= 8) { orc_set_error(o, MELT_ERR_RAY_RECURSION, "ray tagging: parallel-ray recursion did not terminate"); return; }
    double Cp[3] = {C[0], C[1] + 0.1 / o->dx2, C[2] - 0.1 / o->dx3};
    ray_tri(o, intersect, ip, Cp, Q, V0, V1, V2, depth + 1);
  }
  double f = 1.0 / a;
  double u = f * (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]);
  if (u < EPS || u > 1.0) *intersect = 0;
  orc_cross(q1, s, e1);
  double v = f * (r[0] * q1[0] + r[1] * q1[1] + r[2] * q1[2]);
  if (v < EPS || (u + v) > 1.0) *intersect = 0;
  double t = f * (e2[0] * q1[0] + e2[1] * q1[1] + e2[2] * q1[2]);
  if (t < 0.0 || t > 1.0) *intersect = 0;
  if (*intersect)
    for (int d = 0; d < 3; ++d) ip[d] = C[d] + t * r[d];
}

static double rayTagQ_pencil(orc_ctx* o, const double Q[3], const int* list, int n) {  /* pencil_rayTag.f90:381-461 */
  double C[3] = {-o->cfg.xlen * 10.0, Q[1], Q[2]};
  int cnt = 0;
  for (int nt = 0; nt < n; ++nt) {
    int f = list[nt];
    if (o->ghost_f[f]) continue;
    const double* V0 = V3(LAG(MELT_L_XYZV), o->vert_of_face[3 * f + 0] - 1);
    const double* V1 = V3(LAG(MELT_L_XYZV), o->vert_of_face[3 * f + 1] - 1);
    const double* V2 = V3(LAG(MELT_L_XYZV), o->vert_of_face[3 * f + 2] - 1);
    if (triangleXIntersect(V0, V1, V2, Q[0] - 0.5 / o->dx1, Q[0] + 0.5 / o->dx1)) return 0.5;
    int hit = 0;
    double ipt[3] = {0, 0, 0};
    ray_tri(o, &hit, ipt, C, Q, V0, V1, V2, 0);
    if (hit) {
      int inside = ipt[0] >= Q[0] - 0.5 / o->dx1 && ipt[0] <= Q[0] + 0.5 / o->dx1 && ipt[1] >= Q[1] - 0.5 / o->dx2 &&
                   ipt[1] <= Q[1] + 0.5 / o->dx2 && ipt[2] >= Q[2] - 0.5 / o->dx3 && ipt[2] <= Q[2] + 0.5 / o->dx3;
      if (inside) return 0.5;
      cnt += 1;
    }
  }
  return (cnt % 2 == 0) ? 1.0 : 0.0;
}

static int pencilTag(orc_ctx* o) {  /* pencil_rayTag.f90:37-164 (one rank: naive serial branch) */
  int ind[3][2];
  get_bbox_inds(o, ind);
  for (int id = MELT_VOFP; id <= MELT_VOFZ; ++id) orc_fill_field(o, id, 1.0);
  int jb = ind[1][0], je = ind[1][1], kb = ind[2][0], ke = ind[2][1];
  int js = je - jb + 1, ks = ke - kb + 1, BUFF = o->n1;
  int* cnt = (int*)calloc((size_t)js * ks, sizeof(int));
  int* lst = (int*)calloc((size_t)js * ks * BUFF, sizeof(int));
  for (int j = jb; j <= je; ++j) {
    double ymin = orc_ym(o, j) - 0.5 / o->dx2, ymax = orc_ym(o, j) + 0.5 / o->dx2;
    for (int k = kb; k <= ke; ++k) {
      double zmin = orc_zm(o, k) - 0.5 / o->dx3, zmax = orc_zm(o, k) + 0.5 / o->dx3;
      size_t pc = (size_t)(j - jb) + (size_t)js * (k - kb);
      for (int f = 0; f < o->nf; ++f) {
        if (o->ghost_f[f]) continue;
        const double* V0 = V3(LAG(MELT_L_XYZV), o->vert_of_face[3 * f + 0] - 1);
        const double* V1 = V3(LAG(MELT_L_XYZV), o->vert_of_face[3 * f + 1] - 1);
        const double* V2 = V3(LAG(MELT_L_XYZV), o->vert_of_face[3 * f + 2] - 1);
        if (intersects_yz_range(V0, V1, V2, ymin, ymax, zmin, zmax)) {
          cnt[pc] += 1;
          if (cnt[pc] > BUFF) {
            orc_set_error(o, MELT_ERR_PENCIL_OVERFLOW, "pencil triangle count exceeded buffer size");
            free(cnt); free(lst);
            return o->status;
          }
          lst[pc * BUFF + cnt[pc] - 1] = f;
        }
      }
    }
  }
  /* pencil_tagCells pencil_rayTag.f90:189-291 */
  for (int i = ind[0][0]; i <= ind[0][1]; ++i)
    for (int j = jb; j <= je; ++j)
      for (int kk = kb; kk <= ke; ++kk) {
        double Q[3] = {orc_xm(o, i), orc_ym(o, j), orc_zm(o, kk)};
        int k = orc_mod(kk - 1, o->n3) + 1;
        size_t pc = (size_t)(j - jb) + (size_t)js * (kk - kb);
        double vof = rayTagQ_pencil(o, Q, &lst[pc * BUFF], cnt[pc]);
        int ii = orc_mod(i - 1, o->n1) + 1, jj = orc_mod(j - 1, o->n2) + 1;
        if (!(F(MELT_VOFP, ii, jj, k) < 1.0)) F(MELT_VOFP, ii, jj, k) = vof;
      }
  free(cnt); free(lst);
  return o->status;
}

/* genSDF.f90:85-234 (Eberly point-triangle distance, regions 0-6) */
static double dist_point_tri(const double P[3], const double* V0, const double* V1, const double* V2) {
  double E0[3], E1[3], DE[3];
  for (int d = 0; d < 3; ++d) { E0[d] = V1[d] - V0[d]; E1[d] = V2[d] - V0[d]; DE[d] = V0[d] - P[d]; }
  double a = E0[0] * E0[0] + E0[1] * E0[1] + E0[2] * E0[2];
  double b = E0[0] * E1[0] + E0[1] * E1[1] + E0[2] * E1[2];
  double c = E1[0] * E1[0] + E1[1] * E1[1] + E1[2] * E1[2];
  double d = E0[0] * DE[0] + E0[1] * DE[1] + E0[2] * DE[2];
  double e = E1[0] * DE[0] + E1[1] * DE[1] + E1[2] * DE[2];
  double det = a * c - b * b, s = b * e - c * d, t = b * d - a * e;
  if (s + t <= det) {
    if (s < 0.0) {
      if (t < 0.0) {
        if (d < 0.0) { t = 0.0; s = (-d >= a) ? 1.0 : -d / a; }
        else { s = 0.0; if (e >= 0.0) t = 0.0; else if (-e >= c) t = 1.0; else t = -e / c; }
      } else {
        s = 0.0;
        if (e >= 0.0) t = 0.0; else if (-e >= c) t = 1.0; else t = -e / c;
      }
    } else if (t < 0.0) {
      t = 0.0;
      if (d >= 0.0) s = 0.0; else if (-d >= a) s = 1.0; else s = -d / a;
    } else {
      double invDet = 1.0 / det;
      s = s * invDet; t = t * invDet;
    }
  } else {
    if (s < 0.0) {
      double tmp0 = b + d, tmp1 = c + e;
      if (tmp1 > tmp0) {
        double numer = tmp1 - tmp0, denom = a - 2.0 * b + c;
        if (numer >= denom) { s = 1.0; t = 0.0; } else { s = numer / denom; t = 1.0 - s; }
      } else {
        s = 0.0;
        if (tmp1 <= 0.0) t = 1.0; else if (e >= 0.0) t = 0.0; else t = -e / c;
      }
    } else if (t < 0.0) {
      double tmp0 = b + e, tmp1 = a + d;
      if (tmp1 > tmp0) {
        double numer = tmp1 - tmp0, denom = a - 2.0 * b + c;
        if (numer >= denom) { t = 1.0; s = 0.0; } else { t = numer / denom; s = 1.0 - t; }
      } else {
        t = 0.0;
        if (tmp1 <= 0.0) s = 1.0; else if (d >= 0.0) s = 0.0; else s = -d / a;
      }
    } else {
      double numer = c + e - b - d, denom = a - 2.0 * b + c;
      if (numer <= 0.0) s = 0.0; else if (numer >= denom) s = 1.0; else s = numer / denom;
      t = 1.0 - s;
    }
  }
  double cp[3];
  for (int q = 0; q < 3; ++q) cp[q] = V0[q] + s * E0[q] + t * E1[q];
  for (int q = 0; q < 3; ++q) DE[q] = cp[q] - P[q];
  double m = DE[0] * DE[0] + DE[1] * DE[1] + DE[2] * DE[2];
  return sqrt(m);
}

static double sdf_to_vof(orc_ctx* o, double sgnd) {  /* genSDF.f90:237-255 */
  double eps = 1.0 / o->dx1;
  if (sgnd <= -eps) return 0.0;
  if (sgnd >= eps) return 1.0;
  return 0.5 * (1.0 + sgnd * o->dx1 + (1.0 / o->pi) * orc_det_sin(o->pi * sgnd * o->dx1));
}

static void narrowBandTagging(orc_ctx* o) {  /* genSDF.f90:10-82 */
  const int sb = 2;
  orc_fill_field(o, MELT_SDF, 999.0);
  for (int f = 0; f < o->nf; ++f) {
    if (o->ghost_f[f]) continue;
    const double* pos = V3(LAG(MELT_L_TRI_BAR), f);
    int i = (int)floor(pos[0] * o->dx1) + 1, j = (int)floor(pos[1] * o->dx2) + 1, k = (int)floor(pos[2] * o->dx3) + 1;
    const double* V0 = V3(LAG(MELT_L_XYZV), o->vert_of_face[3 * f + 0] - 1);
    const double* V1 = V3(LAG(MELT_L_XYZV), o->vert_of_face[3 * f + 1] - 1);
    const double* V2 = V3(LAG(MELT_L_XYZV), o->vert_of_face[3 * f + 2] - 1);
    const double* nor = V3(LAG(MELT_L_TRI_NOR), f);
    for (int kk = k - sb; kk <= k + sb; ++kk) {
      int kk_mod = orc_mod(kk - 1, o->n3) + 1;
      for (int jj = j - sb; jj <= j + sb; ++jj)
        for (int ii = i - sb; ii <= i + sb; ++ii) {
          double P[3] = {orc_xm(o, ii), orc_ym(o, jj), orc_zm(o, kk)};
          double sgnd = dist_point_tri(P, V0, V1, V2);
          int ii_mod = orc_mod(ii - 1, o->n1) + 1, jj_mod = orc_mod(jj - 1, o->n1) + 1;
          if (fabs(F(MELT_SDF, ii_mod, jj_mod, kk_mod)) > sgnd) {
            double Qv[3] = {pos[0] - P[0], pos[1] - P[1], pos[2] - P[2]};
            double dotval = nor[0] * Qv[0] + nor[1] * Qv[1] + nor[2] * Qv[2];
            double sgn = signbit(dotval) ? -1.0 : 1.0;
            sgnd = sgnd * (-sgn);
            F(MELT_SDF, ii_mod, jj_mod, kk_mod) = sgnd;
            F(MELT_VOFP, ii_mod, jj_mod, kk_mod) = sdf_to_vof(o, sgnd);
          }
        }
    }
  }
  orc_interp_vof(o);
}

int orc_interp_vof(orc_ctx* o) {  /* tagCells.f90:85-111 */
  for (int kc = 1; kc <= o->n3; ++kc)
    for (int jc = 1; jc <= o->n2; ++jc)
      for (int ic = 1; ic <= o->n1; ++ic) {
        double p = F(MELT_VOFP, ic, jc, kc);
        F(MELT_VOFX, ic, jc, kc) = 0.5 * (F(MELT_VOFP, o->imv[ic], jc, kc) + p);
        F(MELT_VOFY, ic, jc, kc) = 0.5 * (F(MELT_VOFP, ic, o->jmv[jc], kc) + p);
        F(MELT_VOFZ, ic, jc, kc) = 0.5 * (F(MELT_VOFP, ic, jc, kc - 1) + p);
      }
  return 0;
}

/* ---------------------------------------------- sphereTagging (tagType 0) */
static double sphere_dist(const double x0[3], const double xc[3], double rad) {  /* geom.f90:56-80 */
  return sqrt((x0[0] - xc[0]) * (x0[0] - xc[0]) + (x0[1] - xc[1]) * (x0[1] - xc[1]) + (x0[2] - xc[2]) * (x0[2] - xc[2])) - rad;
}

/* level_setc2 / level_set12 / 22 / 32 (sphereTagging.f90:91-549): corners of
 * the control volume of the quantity; stag[d] = 1 takes the centres c-1, c of
 * axis d, 0 the faces c, c+1. */
static double level_set(orc_ctx* o, int ic, int jc, int kc, const double xgc[3], const int stag[3]) {
  double phi_tot = 0.0, alpha = 0.0;
  for (int i = (stag[0] ? ic - 1 : ic); i <= (stag[0] ? ic : ic + 1); ++i)
    for (int j = (stag[1] ? jc - 1 : jc); j <= (stag[1] ? jc : jc + 1); ++j)
      for (int k = (stag[2] ? kc - 1 : kc); k <= (stag[2] ? kc : kc + 1); ++k) {
        double x[3] = {stag[0] ? orc_xm(o, i) : orc_xc(o, i), stag[1] ? orc_ym(o, j) : orc_yc(o, j),
                       stag[2] ? orc_zm(o, k) : orc_zc(o, k)};
        double phi = sphere_dist(x, xgc, o->rad_p);
        alpha = alpha + (-phi) * (-phi >= 0.0 ? 1.0 : 0.0);   /* heaviside, geom.f90:129-138 */
        phi_tot = phi_tot + fabs(phi);
      }
  return alpha / phi_tot;
}

static int sphereTagging(orc_ctx* o) {  /* sphereTagging.f90:5-34, tagCells.f90:11-16 */
  double lim[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};   /* get_bbox_inds, tagCells.f90:31-62 */
  for (int v = 0; v < o->nv; ++v) {
    if (o->ghost_v[v]) continue;
    for (int d = 0; d < 3; ++d) {
      lim[d] = fmin(lim[d], LAG(MELT_L_XYZV)[3 * v + d]);
      lim[3 + d] = fmax(lim[3 + d], LAG(MELT_L_XYZV)[3 * v + d]);
    }
  }
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = (int)floor(lim[d] * o->dx1) + 1 - 2;
    hi[d] = (int)floor(lim[3 + d] * o->dx1) + 1 + 2;
  }
  for (int id = MELT_VOFP; id <= MELT_VOFZ; ++id) orc_fill_field(o, id, 1.0);
  static const int ids[4] = {MELT_VOFP, MELT_VOFX, MELT_VOFY, MELT_VOFZ};   /* qc2, q12, q22, q32 */
  static const int stags[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int q = 0; q < 4; ++q)
    for (int i = lo[0]; i <= hi[0]; ++i)
      for (int j = lo[1]; j <= hi[1]; ++j)
        for (int kk = lo[2]; kk <= hi[2]; ++kk) {
          double xgc[3] = {o->body.pos_cm[0], o->body.pos_cm[1], o->body.pos_cm[2]};
          int k = kk;
          if (k >= o->n3 + 1) { k = k - o->n3; xgc[2] = xgc[2] - o->cfg.zlen; }   /* get_periodic_indices */
          if (k < 1) { k = k + o->n3; xgc[2] = xgc[2] + o->cfg.zlen; }
          if (k < 1 || k > o->n3) continue;
          double alpha = level_set(o, i, j, k, xgc, stags[q]);
          int ii = orc_mod(i - 1, o->n1) + 1, jj = orc_mod(j - 1, o->n2) + 1;
          if (!(F(ids[q], ii, jj, k) < 1.0)) F(ids[q], ii, jj, k) = 1.0 - alpha;
        }
  return 0;
}

int orc_tagcells(orc_ctx* o) {  /* tagCells.f90:4-27 */
  if (!o->has_geom) return MELT_ERR_NO_GEOMETRY;
  if (o->cfg.tagtype == 0) {
    o->body.initial_tag = 0;
    return sphereTagging(o);
  }
  if (o->body.initial_tag) {
    int rc = pencilTag(o);
    if (rc) return rc;
  }
  narrowBandTagging(o);
  /* gcurv.f90:61-63: the first call (ray tagging of every cell) happens with
     initial_tag set, every later one tags the narrow band only */
  o->body.initial_tag = 0;
  return o->status;
}

/* ============================================================== MLS forcing */
int orc_find_centroid_indices(orc_ctx* o) {  /* findCentroidIndices.f90:6-53 */
  for (int f = 0; f < o->nf; ++f) {
    if (o->ghost_f[f]) continue;
    const double* pos = V3(LAG(MELT_L_TRI_BAR), f);
    int* p = &o->pind[6 * f];
    p[0] = (int)floor(pos[0] * o->dx1) + 1; p[3] = orc_fnint(pos[0] * o->dx1) + 1;
    p[1] = (int)floor(pos[1] * o->dx2) + 1; p[4] = orc_fnint(pos[1] * o->dx2) + 1;
    int k1 = (int)floor(pos[2] * o->dx3) + 1, kst = orc_fnint(pos[2] * o->dx3) + 1;
    p[2] = orc_mod(k1 - 1, o->n3) + 1; p[5] = orc_mod(kst - 1, o->n3) + 1;
  }
  return 0;
}

/* shared body of wght1/wght2/wght3/wghttemp (mlsWeight.f90:48-412): support
 * indices i0..i0+2, j0.., k0.. with grid coordinate functions gx, gy, gz */
typedef double (*coordfn)(const orc_ctx*, int);
static int mls_shape(orc_ctx* o, const double pos[3], int i0, int j0, int k0, coordfn gx, coordfn gy, coordfn gz, double out[27]) {
  double A[16] = {0}, B[4 * 27], ptx[4] = {1.0, pos[0], pos[1], pos[2]};
  int inw = 0;
  for (int k = k0; k <= k0 + 2; ++k) {
    double n3 = fabs(gz(o, k) - pos[2]) / (WSCL / o->dx3);
    double w3 = mls_gaussian(n3, WCON);
    for (int j = j0; j <= j0 + 2; ++j) {
      double n2 = fabs(gy(o, j) - pos[1]) / (WSCL / o->dx2);
      double w2 = mls_gaussian(n2, WCON);
      double w23 = w2 * w3;
      for (int i = i0; i <= i0 + 2; ++i) {
        double n1 = fabs(gx(o, i) - pos[0]) / (WSCL / o->dx1);
        double w1 = mls_gaussian(n1, WCON);
        double wtx = w1 * w23;
        double pxk[4] = {1.0, gx(o, i), gy(o, j), gz(o, k)};
        rank1(A, wtx, pxk);
        for (int q = 0; q < 4; ++q) B[q + 4 * inw] = wtx * pxk[q];
        inw++;
      }
    }
  }
  if (orc_inv4(A)) return 1;
  double ptxA[4];
  for (int j = 0; j < 4; ++j) {  /* DGEMM 1x4 = ptx * invA */
    double c = 0.0;
    for (int l = 0; l < 4; ++l) c = c + (1.0 * A[l + 4 * j]) * ptx[l];
    ptxA[j] = c;
  }
  for (int j = 0; j < 27; ++j) {  /* DGEMM 1x27 = ptxA * B */
    double c = 0.0;
    for (int l = 0; l < 4; ++l) c = c + (1.0 * B[l + 4 * j]) * ptxA[l];
    out[j] = c;
  }
  return 0;
}

int orc_mlsweight(orc_ctx* o) {  /* mlsWeight.f90:1-45 */
  for (int f = 0; f < o->nf; ++f) {
    if (o->ghost_f[f]) continue;
    const double* pos = V3(LAG(MELT_L_TRI_BAR), f);
    const int* p = &o->pind[6 * f];
    int k1 = (int)floor(pos[2] * o->dx3) + 1, kst = orc_fnint(pos[2] * o->dx3) + 1;
    int bad = 0;
    /* wght1: x = xc around ist, y = ym around j1, z = zm around unwrapped floor */
    bad |= mls_shape(o, pos, p[3] - 1, p[1] - 1, k1 - 1, orc_xc, orc_ym, orc_zm, &LAG(MELT_L_PTXAB_Q1)[27 * f]);
    /* wght2: x = xm around i1, y = yc around jst */
    bad |= mls_shape(o, pos, p[0] - 1, p[4] - 1, k1 - 1, orc_xm, orc_yc, orc_zm, &LAG(MELT_L_PTXAB_Q2)[27 * f]);
    /* wghttemp: xm/ym/zm around i1, j1, floor k */
    bad |= mls_shape(o, pos, p[0] - 1, p[1] - 1, k1 - 1, orc_xm, orc_ym, orc_zm, &LAG(MELT_L_PTXAB_TEMP)[27 * f]);
    /* wght3: xm/ym around i1,j1, zc around nint k */
    bad |= mls_shape(o, pos, p[0] - 1, p[1] - 1, kst - 1, orc_xm, orc_ym, orc_zc, &LAG(MELT_L_PTXAB_Q3)[27 * f]);
    if (bad) { orc_set_error(o, MELT_ERR_SINGULAR_MLS, "Error in dgetrf (singular MLS moment matrix)"); return o->status; }
  }
  return 0;
}

/* forc1/forc2/forc3/forctemp (forc1.f90:1-68 etc.): Um = sum q*phi over the
 * support with wrapped i,j and ghost-plane k; spread cfac*(V-Um)*phi */
static void forc_one(orc_ctx* o, int f, int fld, int out, const double* phi, double V, int i0, int j0, int kc) {
  double Um = 0.0;
  int inw = 0;
  for (int k = kc - 1; k <= kc + 1; ++k)
    for (int j = j0 - 1; j <= j0 + 1; ++j) {
      int jj = orc_mod(j - 1, o->n2) + 1;
      for (int i = i0 - 1; i <= i0 + 1; ++i) {
        int ii = orc_mod(i - 1, o->n1) + 1;
        Um = Um + F(fld, ii, jj, k) * phi[inw];
        inw++;
      }
    }
  inw = 0;
  double cf = LAG(MELT_L_CFAC)[f];
  for (int k = kc - 1; k <= kc + 1; ++k)
    for (int j = j0 - 1; j <= j0 + 1; ++j) {
      int jj = orc_mod(j - 1, o->n2) + 1;
      for (int i = i0 - 1; i <= i0 + 1; ++i) {
        int ii = orc_mod(i - 1, o->n1) + 1;
        OFG(o, out, ii, jj, k) = OFG(o, out, ii, jj, k) + cf * (V - Um) * phi[inw];
        inw++;
      }
    }
}

int orc_mlsforce(orc_ctx* o) {  /* mlsForce.f90:1-36, after for_* = 0 (particle.f90:55-58) */
  for (int id = MELT_FOR_XC; id <= MELT_FOR_TEMP; ++id) orc_fill_field(o, id, 0.0);
  for (int f = 0; f < o->nf; ++f) {
    if (o->ghost_f[f]) continue;
    const int* p = &o->pind[6 * f];
    const double* vt = V3(LAG(MELT_L_VEL_TRI), f);
    forc_one(o, f, MELT_VX, MELT_FOR_XC, &LAG(MELT_L_PTXAB_Q1)[27 * f], vt[0], p[3], p[1], p[2]);
    forc_one(o, f, MELT_VY, MELT_FOR_YC, &LAG(MELT_L_PTXAB_Q2)[27 * f], vt[1], p[0], p[4], p[2]);
    forc_one(o, f, MELT_TEMP, MELT_FOR_TEMP, &LAG(MELT_L_PTXAB_TEMP)[27 * f], o->cfg.tmelt, p[0], p[1], p[2]);
    forc_one(o, f, MELT_VZ, MELT_FOR_ZC, &LAG(MELT_L_PTXAB_Q3)[27 * f], vt[2], p[0], p[1], p[5]);
  }
  return 0;
}

static void fold_ghosts(orc_ctx* o, int id) {  /* update_add_upper_ghost then _lower (mpi_routines.f90:274-362) */
  for (int j = 1; j <= o->n2; ++j)
    for (int i = 1; i <= o->n1; ++i) OFG(o, id, i, j, o->n3) = OFG(o, id, i, j, o->n3) + OFG(o, id, i, j, 0);
  for (int j = 1; j <= o->n2; ++j)
    for (int i = 1; i <= o->n1; ++i) OFG(o, id, i, j, 1) = OFG(o, id, i, j, 1) + OFG(o, id, i, j, o->n3 + 1);
}

int orc_velforce(orc_ctx* o) {  /* velforce.f90:1-59 */
  fold_ghosts(o, MELT_FOR_XC); fold_ghosts(o, MELT_FOR_YC); fold_ghosts(o, MELT_FOR_ZC);
  double mx = 0., my = 0., mz = 0.;
  for (int kc = 1; kc <= o->n3; ++kc)
    for (int jc = 1; jc <= o->n2; ++jc)
      for (int ic = 1; ic <= o->n1; ++ic) {
        mx = mx + OFG(o, MELT_FOR_XC, ic, jc, kc);
        my = my + OFG(o, MELT_FOR_YC, ic, jc, kc);
        mz = mz + OFG(o, MELT_FOR_ZC, ic, jc, kc);
      }
  double n = (double)((long)o->n1 * o->n2 * o->n3);
  mx = mx / n; my = my / n; mz = mz / n;
  for (int kc = 1; kc <= o->n3; ++kc)
    for (int jc = 1; jc <= o->n2; ++jc)
      for (int ic = 1; ic <= o->n1; ++ic) {
        F(MELT_VX, ic, jc, kc) = F(MELT_VX, ic, jc, kc) + OFG(o, MELT_FOR_XC, ic, jc, kc) - mx;
        F(MELT_VY, ic, jc, kc) = F(MELT_VY, ic, jc, kc) + OFG(o, MELT_FOR_YC, ic, jc, kc) - my;
        F(MELT_VZ, ic, jc, kc) = F(MELT_VZ, ic, jc, kc) + OFG(o, MELT_FOR_ZC, ic, jc, kc) - mz;
      }
  return 0;
}

int orc_tempforce(orc_ctx* o) {  /* tempforce.f90:1-46 */
  fold_ghosts(o, MELT_FOR_TEMP);
  for (int kc = 1; kc <= o->n3; ++kc)
    for (int jc = 1; jc <= o->n2; ++jc)
      for (int ic = 1; ic <= o->n1; ++ic) F(MELT_TEMP, ic, jc, kc) = F(MELT_TEMP, ic, jc, kc) + OFG(o, MELT_FOR_TEMP, ic, jc, kc);
  return 0;
}

/* ========================================== MLS gradient shape functions */
/* Common body of wght_gradT/gradP/gradU/gradV/gradW (mls_heatFlux.f90:85-300,
 * mls_structLoads.f90:380-1297).  kz[3] are the grid indices used for the z
 * coordinate (the reference repeats the middle one: [k1-1, k1, k1]); ke[3] the
 * field planes [k1-1, k1, k1+1] (wrapped into 1..n3 through ghost semantics). */
static int mls_grad(orc_ctx* o, int fld, const double pos[3], int i0, int j0, const int kz[3], const int ke[3], coordfn gx, coordfn gy,
                    coordfn gz, double grad[3]) {
  double A[16] = {0}, dAx[16] = {0}, dAy[16] = {0}, dAz[16] = {0};
  double B[4 * 27], dBx[4 * 27], dBy[4 * 27], dBz[4 * 27], val[27];
  double ptx[4] = {1.0, pos[0], pos[1], pos[2]};
  const double eps = 2.220446049250313e-16;  /* EPSILON(1.0d0) */
  int inw = 0;
  for (int nk = 0; nk < 3; ++nk) {
    int k = kz[nk];
    double n3 = fabs(pos[2] - gz(o, k)) / (WSCL / o->dx3);
    double w3 = mls_gaussian(n3, WCON), dw3;
    if (n3 < eps) dw3 = 0.0;
    else { dw3 = mls_gauss_deriv(n3, WCON); dw3 = dw3 * ((pos[2] - gz(o, k)) / fabs(pos[2] - gz(o, k)) / (WSCL / o->dx3)); }
    for (int j = j0 - 1; j <= j0 + 1; ++j) {
      double n2 = fabs(pos[1] - gy(o, j)) / (WSCL / o->dx2);
      double w2 = mls_gaussian(n2, WCON);
      double w23 = w2 * w3, dw2;
      if (n2 < eps) dw2 = 0.0;
      else { dw2 = mls_gauss_deriv(n2, WCON); dw2 = dw2 * ((pos[1] - gy(o, j)) / fabs(pos[1] - gy(o, j)) / (WSCL / o->dx2)); }
      for (int i = i0 - 1; i <= i0 + 1; ++i) {
        double n1 = fabs(pos[0] - gx(o, i)) / (WSCL / o->dx1);
        double w1 = mls_gaussian(n1, WCON);
        double wtx = w1 * w23, dw1;
        if (n1 < eps) dw1 = 0.0;
        else { dw1 = mls_gauss_deriv(n1, WCON); dw1 = dw1 * ((pos[0] - gx(o, i)) / fabs(pos[0] - gx(o, i)) / (WSCL / o->dx1)); }
        double pxk[4] = {1.0, gx(o, i), gy(o, j), gz(o, k)};
        rank1(A, wtx, pxk);
        for (int q = 0; q < 4; ++q) B[q + 4 * inw] = wtx * pxk[q];
        double dWdx = dw1 * w2 * w3, dWdy = w1 * dw2 * w3, dWdz = w1 * w2 * dw3;
        rank1(dAx, dWdx, pxk); rank1(dAy, dWdy, pxk); rank1(dAz, dWdz, pxk);
        for (int q = 0; q < 4; ++q) { dBx[q + 4 * inw] = dWdx * pxk[q]; dBy[q + 4 * inw] = dWdy * pxk[q]; dBz[q + 4 * inw] = dWdz * pxk[q]; }
        int ii = orc_mod(i - 1, o->n1) + 1, jj = orc_mod(j - 1, o->n2) + 1;
        val[inw] = F(fld, ii, jj, ke[nk]);
        inw++;
      }
    }
  }
  if (orc_inv4(A)) return 1;
  double gam[4];
  for (int i = 0; i < 4; ++i) gam[i] = 0.0;  /* DGEMV N, alpha 1, beta 0 */
  for (int j = 0; j < 4; ++j) { double t = 1.0 * ptx[j]; for (int i = 0; i < 4; ++i) gam[i] = gam[i] + t * A[i + 4 * j]; }
  const double* dA[3] = {dAx, dAy, dAz};
  const double* dB[3] = {dBx, dBy, dBz};
  for (int d = 0; d < 3; ++d) {
    double pv[4] = {0.0, 0.0, 0.0, 0.0};
    pv[d + 1] = 1.0;
    for (int j = 0; j < 4; ++j) { double t = -1.0 * gam[j]; for (int i = 0; i < 4; ++i) pv[i] = pv[i] + t * dA[d][i + 4 * j]; }
    double dg[4] = {0, 0, 0, 0};
    for (int j = 0; j < 4; ++j) { double t = 1.0 * pv[j]; for (int i = 0; i < 4; ++i) dg[i] = dg[i] + t * A[i + 4 * j]; }
    double ddp[27];
    for (int j = 0; j < 27; ++j) {
      double c = 0.0;
      for (int l = 0; l < 4; ++l) c = c + (1.0 * B[l + 4 * j]) * dg[l];
      ddp[j] = c;
    }
    for (int j = 0; j < 27; ++j) {
      double c = ddp[j];
      for (int l = 0; l < 4; ++l) c = c + (1.0 * dB[d][l + 4 * j]) * gam[l];
      ddp[j] = c;
    }
    double s = 0.0;
    for (int j = 0; j < 27; ++j) s = s + ddp[j] * val[j];
    grad[d] = s;
  }
  return 0;
}

/* wght_press mls_structLoads.f90:280-378: plain shape function, z grid k1-1..k1+1 */
static int mls_press(orc_ctx* o, const double pos[3], int i0, int j0, double* out) {
  double phi[27];
  int k1 = (int)floor(pos[2] * o->dx3) + 1;
  if (mls_shape(o, pos, i0 - 1, j0 - 1, k1 - 1, orc_xm, orc_ym, orc_zm, phi)) return 1;
  int k1w = orc_mod(k1 - 1, o->n3) + 1;
  double p = 0.0;
  int inw = 0;
  for (int kk = 0; kk < 3; ++kk) {
    int k = k1w - 1 + kk;
    for (int j = j0 - 1; j <= j0 + 1; ++j) {
      int jj = orc_mod(j - 1, o->n2) + 1;
      for (int i = i0 - 1; i <= i0 + 1; ++i) {
        int ii = orc_mod(i - 1, o->n1) + 1;
        p = p + F(MELT_PR, ii, jj, k) * phi[inw];
        inw++;
      }
    }
  }
  *out = p;
  return 0;
}

static void probe_k(orc_ctx* o, double z, int use_nint, int kz[3], int ke[3]) {
  int k1 = use_nint ? orc_fnint(z * o->dx3) + 1 : (int)floor(z * o->dx3) + 1;
  kz[0] = k1 - 1; kz[1] = k1; kz[2] = k1;
  int kw = orc_mod(k1 - 1, o->n3) + 1;
  ke[0] = kw - 1; ke[1] = kw; ke[2] = kw + 1;
}

/* findProbeIndices.f90:9-76 + mls_heatFlux.f90:1-80 + faceToVert_interp :302-319 */
static int heat_flux(orc_ctx* o) {
  double pec = o->cfg.ren * o->cfg.prandtl;
  double h = o->body.h_eulerian;
  memset(LAG(MELT_L_QW_O), 0, sizeof(double) * o->nf);
  memset(LAG(MELT_L_QW_I), 0, sizeof(double) * o->nf);
  memset(LAG(MELT_L_QW_OVERT), 0, sizeof(double) * o->nv);
  memset(LAG(MELT_L_QW_IVERT), 0, sizeof(double) * o->nv);
  for (int f = 0; f < o->nf; ++f) {
    if (o->ghost_f[f]) continue;
    const double* nor = V3(LAG(MELT_L_TRI_NOR), f);
    const double* bar = V3(LAG(MELT_L_TRI_BAR), f);
    double s = orc_norm2(nor, 3);
    for (int side = 0; side < 2; ++side) {
      double pos[3];
      for (int d = 0; d < 3; ++d) pos[d] = side == 0 ? bar[d] + h * nor[d] / s : bar[d] - h * nor[d] / s;
      int i1 = (int)floor(pos[0] * o->dx1) + 1, j1 = (int)floor(pos[1] * o->dx2) + 1;
      int kz[3], ke[3];
      probe_k(o, pos[2], 0, kz, ke);
      double g[3];
      if (mls_grad(o, MELT_TEMP, pos, i1, j1, kz, ke, orc_xm, orc_ym, orc_zm, g)) {
        orc_set_error(o, MELT_ERR_SINGULAR_MLS, "Error in dgetrf (singular MLS moment matrix)");
        return o->status;
      }
      double q = nor[0] * g[0] + nor[1] * g[1] + nor[2] * g[2];
      q = q / pec;
      (side == 0 ? LAG(MELT_L_QW_O) : LAG(MELT_L_QW_I))[f] = q;
      double* qv = side == 0 ? LAG(MELT_L_QW_OVERT) : LAG(MELT_L_QW_IVERT);
      for (int c = 0; c < 3; ++c) {
        int v = o->vert_of_face[3 * f + c] - 1;
        qv[v] = qv[v] + q * (LAG(MELT_L_SUR)[f] / 3.0 / LAG(MELT_L_AVERT)[v]);
      }
    }
  }
  return 0;
}

static int struct_loads(orc_ctx* o) {  /* mls_structLoads.f90:1-260 */
  melt_body* b = &o->body;
  for (int d = 0; d < 3; ++d) {
    b->int_prn_dA_m1[d] = b->int_prn_dA[d]; b->int_tau_dA_m1[d] = b->int_tau_dA[d];
    b->int_r_x_prn_dA_m1[d] = b->int_r_x_prn_dA[d]; b->int_r_x_tau_dA_m1[d] = b->int_r_x_tau_dA[d];
  }
  double* pn = LAG(MELT_L_PRESS_N_TRI);
  double* tn = LAG(MELT_L_TAU_N);
  memset(pn, 0, sizeof(double) * 3 * o->nf);
  memset(tn, 0, sizeof(double) * 3 * o->nf);
  memset(o->r_x_prn_tri, 0, sizeof(double) * 3 * o->nf);
  memset(o->r_x_tau_n, 0, sizeof(double) * 3 * o->nf);
  double h = b->h_eulerian, ren = o->cfg.ren;
  for (int f = 0; f < o->nf; ++f) {
    if (o->ghost_f[f]) continue;
    const double* nor = V3(LAG(MELT_L_TRI_NOR), f);
    const double* bar = V3(LAG(MELT_L_TRI_BAR), f);
    double sur = LAG(MELT_L_SUR)[f];
    double s = orc_norm2(nor, 3), pos[3], rad[3];
    for (int d = 0; d < 3; ++d) pos[d] = bar[d] + h * nor[d] / s;
    for (int d = 0; d < 3; ++d) rad[d] = bar[d] - b->pos_cm[d];
    int jcent = (int)floor(pos[1] * o->dx2) + 1, jstag = orc_fnint(pos[1] * o->dx2) + 1;
    int icent = (int)floor(pos[0] * o->dx1) + 1, istag = orc_fnint(pos[0] * o->dx1) + 1;
    double pprobe, gP[3], gU[3], gV[3], gW[3];
    int kz[3], ke[3];
    int bad = mls_press(o, pos, icent, jcent, &pprobe);
    probe_k(o, pos[2], 0, kz, ke);
    bad |= mls_grad(o, MELT_PR, pos, icent, jcent, kz, ke, orc_xm, orc_ym, orc_zm, gP);
    if (bad) { orc_set_error(o, MELT_ERR_SINGULAR_MLS, "Error in dgetrf (singular MLS moment matrix)"); return o->status; }
    double ptri = pprobe - h * (gP[0] * nor[0] + gP[1] * nor[1] + gP[2] * nor[2]);
    for (int d = 0; d < 3; ++d) pn[3 * f + d] = ptri * sur * nor[d];
    double pv[3] = {ptri * nor[0], ptri * nor[1], ptri * nor[2]}, rx[3];
    orc_cross(rx, rad, pv);
    for (int d = 0; d < 3; ++d) o->r_x_prn_tri[3 * f + d] = rx[d] * sur;
    bad |= mls_grad(o, MELT_VX, pos, istag, jcent, kz, ke, orc_xc, orc_ym, orc_zm, gU);
    bad |= mls_grad(o, MELT_VY, pos, icent, jstag, kz, ke, orc_xm, orc_yc, orc_zm, gV);
    if (bad) { orc_set_error(o, MELT_ERR_SINGULAR_MLS, "Error in dgetrf (singular MLS moment matrix)"); return o->status; }
    double S11 = gU[0], S12 = gU[1], S13 = gU[2], S21 = gV[0], S22 = gV[1], S23 = gV[2];
    double t1 = (sur / ren) * (2.0 * S11 * nor[0] + (S12 + S21) * nor[1] + S13 * nor[2]);
    double t2 = (sur / ren) * ((S21 + S12) * nor[0] + 2.0 * S22 * nor[1] + S23 * nor[2]);
    double t3 = (sur / ren) * (S13 * nor[0] + S23 * nor[1]);
    tn[3 * f + 0] += t1; tn[3 * f + 1] += t2; tn[3 * f + 2] += t3;
    double tv[3] = {t1, t2, t3};
    orc_cross(rx, rad, tv);
    for (int d = 0; d < 3; ++d) o->r_x_tau_n[3 * f + d] += rx[d];
    int kzw[3], kew[3];
    probe_k(o, pos[2], 1, kzw, kew);
    if (mls_grad(o, MELT_VZ, pos, icent, jcent, kzw, kew, orc_xm, orc_ym, orc_zc, gW)) {
      orc_set_error(o, MELT_ERR_SINGULAR_MLS, "Error in dgetrf (singular MLS moment matrix)");
      return o->status;
    }
    double S31 = gW[0], S32 = gW[1], S33 = gW[2];
    t1 = (sur / ren) * (S31 * nor[2]);
    t2 = (sur / ren) * (S32 * nor[2]);
    t3 = (sur / ren) * (S31 * nor[0] + S32 * nor[1] + 2.0 * S33 * nor[2]);
    tn[3 * f + 0] += t1; tn[3 * f + 1] += t2; tn[3 * f + 2] += t3;
    double tw[3] = {t1, t2, t3};
    orc_cross(rx, rad, tw);
    for (int d = 0; d < 3; ++d) o->r_x_tau_n[3 * f + d] += rx[d];
  }
  for (int d = 0; d < 3; ++d) {
    double sp = 0.0, sr = 0.0, st = 0.0, srt = 0.0;
    for (int f = 0; f < o->nf; ++f) { sp += pn[3 * f + d]; sr += o->r_x_prn_tri[3 * f + d]; st += tn[3 * f + d]; srt += o->r_x_tau_n[3 * f + d]; }
    b->int_prn_dA[d] = sp; b->int_r_x_prn_dA[d] = sr; b->int_tau_dA[d] = st; b->int_r_x_tau_dA[d] = srt;
  }
  for (int d = 0; d < 3; ++d) {
    b->fp[d] = b->fp[d] + b->int_prn_dA[d]; b->ftau[d] = b->ftau[d] + b->int_tau_dA[d];
    b->torq_p[d] = b->torq_p[d] + b->int_r_x_prn_dA[d]; b->torq_tau[d] = b->torq_tau[d] + b->int_r_x_tau_dA[d];
  }
  return 0;
}

/* ================================================================ geometry */
static double calc_volume(orc_ctx* o) {  /* calculate_volume triaux.f90:53-80 */
  const double* x = LAG(MELT_L_XYZV);
  double V = 0.0;
  for (int f = 0; f < o->nf; ++f) {
    if (o->ghost_f[f]) continue;
    int v1 = o->vert_of_face[3 * f] - 1, v2 = o->vert_of_face[3 * f + 1] - 1, v3 = o->vert_of_face[3 * f + 2] - 1;
    double vol = x[3 * v1] * (x[3 * v2 + 1] * x[3 * v3 + 2] - x[3 * v2 + 2] * x[3 * v3 + 1]) +
                 x[3 * v2] * (x[3 * v3 + 1] * x[3 * v1 + 2] - x[3 * v3 + 2] * x[3 * v1 + 1]) +
                 x[3 * v3] * (x[3 * v1 + 1] * x[3 * v2 + 2] - x[3 * v1 + 2] * x[3 * v2 + 1]);
    V = V + vol;
  }
  return V / 6.;
}

static void calc_area(orc_ctx* o, const double* x) {  /* calculate_area triaux.f90:82-119 */
  double S = 0.0;
  for (int f = 0; f < o->nf; ++f) {
    if (o->ghost_f[f]) continue;
    int v[3] = {o->vert_of_face[3 * f] - 1, o->vert_of_face[3 * f + 1] - 1, o->vert_of_face[3 * f + 2] - 1};
    double dd[3];
    for (int e = 0; e < 3; ++e) {
      int a = v[e], b = v[(e + 1) % 3];
      double q0 = x[3 * a] - x[3 * b], q1 = x[3 * a + 1] - x[3 * b + 1], q2 = x[3 * a + 2] - x[3 * b + 2];
      dd[e] = sqrt(q0 * q0 + q1 * q1 + q2 * q2);
    }
    double sp = (dd[0] + dd[1] + dd[2]) / 2.;
    LAG(MELT_L_SUR)[f] = sqrt(sp * (sp - dd[0]) * (sp - dd[1]) * (sp - dd[2]));
    S = S + LAG(MELT_L_SUR)[f];
  }
  o->body.surface = S;
}

static void calc_vert_area(orc_ctx* o) {  /* calculate_vert_area triaux.f90:121-150 */
  double* A = LAG(MELT_L_AVERT);
  for (int v = 0; v < o->nv; ++v) A[v] = 0.;
  for (int f = 0; f < o->nf; ++f) {
    if (o->ghost_f[f]) continue;
    for (int c = 0; c < 3; ++c) { int v = o->vert_of_face[3 * f + c] - 1; A[v] = A[v] + LAG(MELT_L_SUR)[f] / 3.0; }
  }
}

static void tri_normal(orc_ctx* o, const double* x, int check_ghost) {  /* calculate_normal / update_tri_normal triaux.f90:154-232 */
  for (int f = 0; f < o->nf; ++f) {
    if (check_ghost && o->ghost_f[f]) continue;
    int v1 = o->vert_of_face[3 * f] - 1, v2 = o->vert_of_face[3 * f + 1] - 1, v3 = o->vert_of_face[3 * f + 2] - 1;
    double e1[3], e2[3], n[3];
    for (int d = 0; d < 3; ++d) { e1[d] = x[3 * v2 + d] - x[3 * v1 + d]; e2[d] = x[3 * v3 + d] - x[3 * v1 + d]; }
    n[0] = e1[1] * e2[2] - e1[2] * e2[1];
    n[1] = e1[2] * e2[0] - e1[0] * e2[2];
    n[2] = e1[0] * e2[1] - e1[1] * e2[0];
    double s = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int d = 0; d < 3; ++d) LAG(MELT_L_TRI_NOR)[3 * f + d] = n[d] / s;
  }
}

static void vert_normal(orc_ctx* o) {  /* calculate_areaWeighted_vert_normal triaux.f90:237-274 */
  double* vn = LAG(MELT_L_VERT_NOR);
  for (int i = 0; i < 3 * o->nv; ++i) vn[i] = 0.;
  for (int f = 0; f < o->nf; ++f) {
    if (o->ghost_f[f]) continue;
    for (int c = 0; c < 3; ++c) {
      int v = o->vert_of_face[3 * f + c] - 1;
      for (int d = 0; d < 3; ++d) vn[3 * v + d] = vn[3 * v + d] + LAG(MELT_L_SUR)[f] * LAG(MELT_L_TRI_NOR)[3 * f + d];
    }
  }
  for (int v = 0; v < o->nv; ++v) {
    if (o->ghost_v[v]) continue;
    double m = orc_norm2(&vn[3 * v], 3);
    for (int d = 0; d < 3; ++d) vn[3 * v + d] = vn[3 * v + d] / m;
  }
}

static void edge_lengths(orc_ctx* o, const double* x) {  /* calculate_eLengths triaux.f90:277-305 */
  o->body.rm_flag = 0;
  for (int e = 0; e < o->ne; ++e) {
    if (o->ghost_e[e]) continue;
    int v1 = o->vert_of_edge[2 * e] - 1, v2 = o->vert_of_edge[2 * e + 1] - 1;
    double a = x[3 * v1] - x[3 * v2], b = x[3 * v1 + 1] - x[3 * v2 + 1], c = x[3 * v1 + 2] - x[3 * v2 + 2];
    LAG(MELT_L_ELENGTHS)[e] = sqrt(a * a + b * b + c * c);
    if (LAG(MELT_L_ELENGTHS)[e] < o->body.e_thresh) o->body.rm_flag = 1;
  }
}

static void centroids(orc_ctx* o) {  /* calc_centroids_from_vert triaux.f90:308-323 */
  const double* x = LAG(MELT_L_XYZV);
  for (int f = 0; f < o->nf; ++f) {
    if (o->ghost_f[f]) continue;
    for (int d = 0; d < 3; ++d) {
      double s = x[3 * (o->vert_of_face[3 * f] - 1) + d] + x[3 * (o->vert_of_face[3 * f + 1] - 1) + d] + x[3 * (o->vert_of_face[3 * f + 2] - 1) + d];
      LAG(MELT_L_TRI_BAR)[3 * f + d] = s / 3.0;
    }
  }
}

static void skewness(orc_ctx* o) {  /* calculate_skewness triaux.f90:325-358 */
  /* edge_of_face is 1-based */
  for (int f = 0; f < o->nf; ++f) {
    if (o->ghost_f[f]) continue;
    double perim = LAG(MELT_L_ELENGTHS)[o->edge_of_face[3 * f] - 1] + LAG(MELT_L_ELENGTHS)[o->edge_of_face[3 * f + 1] - 1] +
                   LAG(MELT_L_ELENGTHS)[o->edge_of_face[3 * f + 2] - 1];
    double q = perim / 3.0;
    double sopt = sqrt(3.0) / 4.0 * (q * q);
    LAG(MELT_L_SKEWNESS)[f] = (sopt - LAG(MELT_L_SUR)[f]) / sopt;
  }
}

static void rigid_params(orc_ctx* o, const double* xyz, double com[3], double* vol, double I[9]) {  /* RigidAuxRoutines.f90:653-820 */
  double Cx = 0, Cy = 0, Cz = 0, Ixx = 0, Iyy = 0, Izz = 0, Ixy = 0, Ixz = 0, Iyz = 0, V = 0;
  for (int f = 0; f < o->nf; ++f) {
    if (o->ghost_f[f]) continue;
    const double* p1 = &xyz[3 * (o->vert_of_face[3 * f] - 1)];
    const double* p2 = &xyz[3 * (o->vert_of_face[3 * f + 1] - 1)];
    const double* p3 = &xyz[3 * (o->vert_of_face[3 * f + 2] - 1)];
    double x1 = p1[0], y1 = p1[1], z1 = p1[2], x2 = p2[0], y2 = p2[1], z2 = p2[2], x3 = p3[0], y3 = p3[1], z3 = p3[2];
    double v = x1 * y2 * z3 + y1 * z2 * x3 + x2 * y3 * z1 - (x3 * y2 * z1 + x2 * y1 * z3 + y3 * z2 * x1);
    V = V + v;
    double x4 = x1 + x2 + x3; Cx = Cx + (v * x4);
    double y4 = y1 + y2 + y3; Cy = Cy + (v * y4);
    double z4 = z1 + z2 + z3; Cz = Cz + (v * z4);
    Ixx = Ixx + v * (x1 * x1 + x2 * x2 + x3 * x3 + x4 * x4);
    Iyy = Iyy + v * (y1 * y1 + y2 * y2 + y3 * y3 + y4 * y4);
    Izz = Izz + v * (z1 * z1 + z2 * z2 + z3 * z3 + z4 * z4);
    Ixy = Ixy + v * (y1 * x1 + y2 * x2 + y3 * x3 + y4 * x4);
    Ixz = Ixz + v * (z1 * x1 + z2 * x2 + z3 * x3 + z4 * x4);
    Iyz = Iyz + v * (z1 * y1 + z2 * y2 + z3 * y3 + z4 * y4);
  }
  double r = 1.0 / (4.0 * V);
  Cx = Cx * r; Cy = Cy * r; Cz = Cz * r;
  V = V / 6.0;
  r = 1.0 / 120.0;
  Ixy = Ixy * r - V * Cy * Cx;
  Ixz = Ixz * r - V * Cz * Cx;
  Iyz = Iyz * r - V * Cz * Cy;
  double xx = Ixx * r - V * Cx * Cx, yy = Iyy * r - V * Cy * Cy, zz = Izz * r - V * Cz * Cz;
  Ixx = yy + zz; Iyy = zz + xx; Izz = xx + yy;
  com[0] = Cx; com[1] = Cy; com[2] = Cz;
  *vol = V;
  double dr = o->cfg.dens_ratio;
  I[0] = Ixx * dr; I[3] = -Ixy * dr; I[6] = -Ixz * dr;
  I[1] = -Ixy * dr; I[4] = Iyy * dr; I[7] = -Iyz * dr;
  I[2] = -Ixz * dr; I[5] = -Iyz * dr; I[8] = Izz * dr;
}

static void rot_matrix(double q[4], double AA[9]) {  /* RigidAuxRoutines.f90:620-651, AA column-major */
  double n = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  for (int i = 0; i < 4; ++i) q[i] = q[i] / n;
#define M(i, j) AA[(i - 1) + 3 * (j - 1)]
  M(1, 1) = q[0] * q[0] + q[1] * q[1] - q[2] * q[2] - q[3] * q[3];
  M(1, 2) = 2.0 * (q[1] * q[2] + q[0] * q[3]);
  M(1, 3) = 2.0 * (q[1] * q[3] - q[0] * q[2]);
  M(2, 1) = 2.0 * (q[1] * q[2] - q[0] * q[3]);
  M(2, 2) = q[0] * q[0] - q[1] * q[1] + q[2] * q[2] - q[3] * q[3];
  M(2, 3) = 2.0 * (q[2] * q[3] + q[0] * q[1]);
  M(3, 1) = 2.0 * (q[1] * q[3] + q[0] * q[2]);
  M(3, 2) = 2.0 * (q[2] * q[3] - q[0] * q[1]);
  M(3, 3) = q[0] * q[0] - q[1] * q[1] - q[2] * q[2] + q[3] * q[3];
#undef M
}

static void matmul3(const double* A, const double* B, double* C) {  /* column-major 3x3, k-summation from 0 */
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s = s + A[i + 3 * k] * B[k + 3 * j];
      C[i + 3 * j] = s;
    }
}
static void matvec3(const double* A, const double* x, double* y) {
  for (int i = 0; i < 3; ++i) { double s = 0.0; for (int k = 0; k < 3; ++k) s = s + A[i + 3 * k] * x[k]; y[i] = s; }
}

static void newton_euler(orc_ctx* o) {  /* RigidAuxRoutines.f90:140-310 */
  melt_body* b = &o->body;
  double vel_m1[3], pos_m1[3], om_m1[3];
  for (int d = 0; d < 3; ++d) { vel_m1[d] = b->vel_cm[d]; pos_m1[d] = b->pos_cm[d]; om_m1[d] = b->omega_c[d]; }
  double e_z[3] = {0., 0., -1.0};
  double pre_fac = 1.0 / b->volume / o->cfg.dens_ratio;
  double ga = o->ga, ro = o->ro, al = o->al, dt = o->dt;
  for (int d = 0; d < 3; ++d)
    b->vel_cm[d] = vel_m1[d] + ga * dt * pre_fac * (-b->int_prn_dA[d] + b->int_tau_dA[d]) +
                   ro * dt * pre_fac * (-b->int_prn_dA_m1[d] + b->int_tau_dA_m1[d]) + al * dt * (1.0 - (1.0 / o->cfg.dens_ratio)) * e_z[d];
  for (int d = 0; d < 3; ++d) b->pos_cm[d] = pos_m1[d] + 0.5 * al * dt * (b->vel_cm[d] + vel_m1[d]);
  double I_LHS[9], I_inv[9], I_prev[9], AA[9], AT[9], buf[9], tmp[3], rhs[3];
  memcpy(I_LHS, b->inert_tensor, sizeof I_LHS);
  double tol = 1.e-15, res = 1.0;
  int n_iter = 1;
  while (n_iter <= 10000 && res > tol) {
    memcpy(I_inv, I_LHS, sizeof I_inv);
    if (orc_inv3(I_inv)) { orc_set_error(o, MELT_ERR_ARG, "Error in dgetrf for I_inv calculation"); return; }
    matvec3(b->inert_tensor, om_m1, tmp);
    for (int d = 0; d < 3; ++d)
      rhs[d] = tmp[d] + ga * dt * (-b->int_r_x_prn_dA[d] + b->int_r_x_tau_dA[d]) + ro * dt * (-b->int_r_x_prn_dA_m1[d] + b->int_r_x_tau_dA_m1[d]);
    matvec3(I_inv, rhs, b->omega_c);
    double ob[3];
    for (int d = 0; d < 3; ++d) ob[d] = 0.5 * (b->omega_c[d] + om_m1[d]);
    double mag = orc_norm2(ob, 3);
    double ang = mag * al * dt;
    if (mag < 1.e-15) { ob[0] = ob[1] = ob[2] = 0.0; }
    else for (int d = 0; d < 3; ++d) ob[d] = ob[d] / mag;
    b->quat[0] = cos(0.5 * ang);
    b->quat[1] = sin(0.5 * ang) * ob[0];
    b->quat[2] = sin(0.5 * ang) * ob[1];
    b->quat[3] = sin(0.5 * ang) * ob[2];
    double qn = orc_norm2(b->quat, 4);
    for (int d = 0; d < 4; ++d) b->quat[d] = b->quat[d] / qn;
    rot_matrix(b->quat, AA);
    memcpy(I_prev, I_LHS, sizeof I_prev);
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) AT[i + 3 * j] = AA[j + 3 * i];
    matmul3(b->inert_tensor, AT, buf);
    matmul3(AA, buf, I_LHS);
    res = 0.0;
    for (int i = 0; i < 9; ++i) res = fmax(res, fabs(I_LHS[i] - I_prev[i]));
    n_iter++;
  }
  memcpy(b->inert_tensor, I_LHS, sizeof I_LHS);
}

static void update_part_pos(orc_ctx* o) {  /* RigidAuxRoutines.f90:1-57 (no collisions: Nparticle = 1) */
  melt_body* b = &o->body;
  double* x = LAG(MELT_L_XYZV);
  for (int v = 0; v < o->nv; ++v)
    if (!o->ghost_v[v]) for (int d = 0; d < 3; ++d) LAG(MELT_L_DXYZV_S)[3 * v + d] = x[3 * v + d] - b->pos_cm[d];
  for (int f = 0; f < o->nf; ++f)
    if (!o->ghost_f[f]) for (int d = 0; d < 3; ++d) LAG(MELT_L_DXYZ_S)[3 * f + d] = LAG(MELT_L_TRI_BAR)[3 * f + d] - b->pos_cm[d];
  newton_euler(o);
  double L[3] = {o->cfg.xlen, o->cfg.ylen, o->cfg.zlen};
  for (int d = 0; d < 3; ++d) {
    if (b->pos_cm[d] < 0.) b->pos_cm[d] = b->pos_cm[d] + L[d];
    if (b->pos_cm[d] > L[d]) b->pos_cm[d] = b->pos_cm[d] - L[d];
  }
  /* update_xyz RigidAuxRoutines.f90:846-900 */
  double AA[9], AT[9];
  rot_matrix(b->quat, AA);
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) AT[i + 3 * j] = AA[j + 3 * i];
  for (int v = 0; v < o->nv; ++v) {
    if (o->ghost_v[v]) continue;
    double t[3];
    matvec3(AT, &LAG(MELT_L_DXYZV_S)[3 * v], t);
    for (int d = 0; d < 3; ++d) x[3 * v + d] = t[d] + b->pos_cm[d];
  }
  centroids(o);
  for (int f = 0; f < o->nf; ++f) {
    if (o->ghost_f[f]) continue;
    double om[3];
    orc_cross(om, b->omega_c, &LAG(MELT_L_DXYZ_S)[3 * f]);
    for (int d = 0; d < 3; ++d) LAG(MELT_L_VEL_TRI)[3 * f + d] = b->vel_cm[d] + om[d];
  }
}

static void stefan(orc_ctx* o) {  /* stefanCondition.f90:1-61 */
  double* vm = LAG(MELT_L_VMELT);
  double* vm1 = LAG(MELT_L_VMELT_M1);
  memcpy(vm1, vm, sizeof(double) * 3 * o->nv);
  for (int v = 0; v < o->nv; ++v) {
    if (o->ghost_v[v]) continue;
    const double* n = &LAG(MELT_L_VERT_NOR)[3 * v];
    for (int d = 0; d < 3; ++d)
      vm[3 * v + d] = o->cfg.cpliquid / o->cfg.latheat * (LAG(MELT_L_QW_IVERT)[v] - LAG(MELT_L_QW_OVERT)[v]) * n[d];
    for (int d = 0; d < 3; ++d)
      LAG(MELT_L_XYZV)[3 * v + d] = LAG(MELT_L_XYZV)[3 * v + d] + o->dt * (o->ga * vm[3 * v + d] + o->ro * vm1[3 * v + d]);
  }
}

int orc_particle(orc_ctx* o) {  /* particle.f90:1-229 */
  if (!o->has_geom) return MELT_ERR_NO_GEOMETRY;
  if (o->cfg.imlsfor != 1) return 0;
  melt_body* b = &o->body;
  double vol_melt = 0.0;   /* particle.f90:32-35: zero unless the body melts */
  orc_find_centroid_indices(o);
  if (orc_mlsweight(o)) return o->status;
  orc_mlsforce(o);
  if (o->cfg.imelt == 1 && heat_flux(o)) return o->status;
  if (o->cfg.imlsstr == 1 && struct_loads(o)) return o->status;
  orc_velforce(o);
  orc_tempforce(o);
  if (o->cfg.imelt == 1) {
    stefan(o);
    b->volume = calc_volume(o);
    vol_melt = b->volume;
  }
  if (o->cfg.imlsstr == 1) {
    centroids(o);
    tri_normal(o, LAG(MELT_L_XYZV), 1);
    rigid_params(o, LAG(MELT_L_XYZV), b->pos_cm, &b->volume, b->inert_tensor);
    update_part_pos(o);
    if (o->status) return o->status;
  }
  /* remeshing particle.f90:121-160 */
  int did_remesh = 0;
  if (o->cfg.iremesh == 1) {
    int rc = orc_remesh_step(o, vol_melt, &did_remesh);
    if (rc) { orc_set_error(o, rc, "remeshing failed (topology, QEM solve or ring walk)"); return o->status; }
  }
  /* primitive geometry info particle.f90:163-183 */
  calc_area(o, LAG(MELT_L_XYZV));
  edge_lengths(o, LAG(MELT_L_XYZV));
  tri_normal(o, LAG(MELT_L_XYZV), 1);
  skewness(o);
  centroids(o);
  calc_vert_area(o);
  b->volume = calc_volume(o);
  vert_normal(o);
  rigid_params(o, LAG(MELT_L_XYZV), b->pos_cm, &b->volume, b->inert_tensor);
  if (did_remesh) {  /* update_tri_velocity RigidAuxRoutines.f90:902-928 */
    for (int f = 0; f < o->nf; ++f) {
      if (o->ghost_f[f]) continue;
      double rad[3], om[3];
      for (int d = 0; d < 3; ++d) rad[d] = LAG(MELT_L_TRI_BAR)[3 * f + d] - b->pos_cm[d];
      orc_cross(om, b->omega_c, rad);
      for (int d = 0; d < 3; ++d) LAG(MELT_L_VEL_TRI)[3 * f + d] = b->vel_cm[d] + om[d];
    }
  }
  if (o->cfg.imelt == 1)
    for (int f = 0; f < o->nf; ++f) LAG(MELT_L_CFAC)[f] = (LAG(MELT_L_SUR)[f] * b->h_eulerian) / b->celvol;
  if (did_remesh && fabs(o->remesh_stats[5]) > 1.e-14) {  /* particle.f90:199-211 */
    orc_set_error(o, MELT_ERR_REMESH, "Non-zero remeshing residual detected");
    return o->status;
  }
  if (b->volume < b->v_thresh) { orc_set_error(o, MELT_ERR_SMALL_VOLUME, "Geometry V/VE smaller than threshold"); return o->status; }
  return 0;
}

/* ============================================================ mesh import */
static const int lag_len[MELT_L_NLAG] = {
    /* per vertex */ 3, 3, 1, 1, 1, 3, 3, 3,
    /* per face   */ 3, 3, 1, 3, 1, 1, 1, 1, 27, 27, 27, 27, 3, 3, 3,
    /* per edge   */ 1};
static int lag_kind(int id) { return id <= MELT_L_DXYZV_S ? 0 : id <= MELT_L_TAU_N ? 1 : 2; }

int orc_set_geometry(orc_ctx* o, int32_t nv, int32_t ne, int32_t nf, const double* xyz0, const int32_t* voe, const int32_t* eof, double rad_p) {
  if (nv <= 0 || ne <= 0 || nf <= 0) return MELT_ERR_ARG;
  o->nv = nv; o->ne = ne; o->nf = nf;
  for (int id = 0; id < MELT_L_NLAG; ++id) {
    free(o->lag[id]);
    int k = lag_kind(id);
    size_t n = (size_t)lag_len[id] * (k == 0 ? nv : k == 1 ? nf : ne);
    o->lag[id] = (double*)calloc(n, sizeof(double));
  }
  free(o->vert_of_edge); free(o->edge_of_face); free(o->vert_of_face); free(o->face_of_edge);
  free(o->ghost_v); free(o->ghost_e); free(o->ghost_f); free(o->xyz0); free(o->r_x_prn_tri); free(o->r_x_tau_n); free(o->pind);
  o->vert_of_edge = (int*)malloc(sizeof(int) * 2 * ne);
  o->edge_of_face = (int*)malloc(sizeof(int) * 3 * nf);
  o->vert_of_face = (int*)malloc(sizeof(int) * 3 * nf);
  o->face_of_edge = (int*)calloc(2 * (size_t)ne, sizeof(int));
  o->ghost_v = (int32_t*)calloc(nv, sizeof(int32_t)); o->ghost_e = (int32_t*)calloc(ne, sizeof(int32_t)); o->ghost_f = (int32_t*)calloc(nf, sizeof(int32_t));
  free(o->anchor_v); free(o->flagged_e);
  o->anchor_v = (int32_t*)malloc(sizeof(int32_t) * nv); o->flagged_e = (int32_t*)calloc(ne, sizeof(int32_t));
  for (int v = 0; v < nv; ++v) o->anchor_v[v] = 1;
  memset(o->remesh_stats, 0, sizeof o->remesh_stats);
  o->xyz0 = (double*)malloc(sizeof(double) * 3 * nv);
  o->r_x_prn_tri = (double*)calloc(3 * (size_t)nf, sizeof(double));
  o->r_x_tau_n = (double*)calloc(3 * (size_t)nf, sizeof(double));
  o->pind = (int*)calloc(6 * (size_t)nf, sizeof(int));
  memcpy(o->vert_of_edge, voe, sizeof(int) * 2 * ne);
  memcpy(o->edge_of_face, eof, sizeof(int) * 3 * nf);
  memcpy(o->xyz0, xyz0, sizeof(double) * 3 * nv);
  /* set_connectivity partaux.f90:160-267 */
  for (int i = 0; i < nf; ++i) {
    int e1 = eof[3 * i] - 1, e2 = eof[3 * i + 1] - 1, v1, v2, v3;
    if (voe[2 * e1 + 1] == voe[2 * e2]) { v1 = voe[2 * e1]; v2 = voe[2 * e1 + 1]; v3 = voe[2 * e2 + 1]; }
    else if (voe[2 * e1 + 1] == voe[2 * e2 + 1]) { v1 = voe[2 * e1]; v2 = voe[2 * e1 + 1]; v3 = voe[2 * e2]; }
    else if (voe[2 * e1] == voe[2 * e2]) { v1 = voe[2 * e1 + 1]; v2 = voe[2 * e1]; v3 = voe[2 * e2 + 1]; }
    else { v1 = voe[2 * e1 + 1]; v2 = voe[2 * e1]; v3 = voe[2 * e2]; }
    o->vert_of_face[3 * i] = v1; o->vert_of_face[3 * i + 1] = v2; o->vert_of_face[3 * i + 2] = v3;
  }
  for (int i = 0; i < nf; ++i)
    for (int c = 0; c < 3; ++c) {
      int e = eof[3 * i + c] - 1;
      if (o->face_of_edge[2 * e] == 0) o->face_of_edge[2 * e] = i + 1;
      else if (o->face_of_edge[2 * e + 1] == 0) o->face_of_edge[2 * e + 1] = i + 1;
      else { orc_set_error(o, MELT_ERR_ARG, "Edge error"); return o->status; }
    }
  o->has_geom = 1;
  o->rad_p = rad_p;
  return orc_geometry_init(o, rad_p);
}

int orc_geometry_init(orc_ctx* o, double rad_p) {
  melt_body* b = &o->body;
  memset(b, 0, sizeof *b);
  /* setup_particles create_geo.f90:1-104 */
  for (int i = 0; i < 3 * o->nv; ++i) o->xyz0[i] = o->xyz0[i] * rad_p;
  rigid_params(o, o->xyz0, b->pos_cm, &b->volume, b->inert_tensor);
  b->pos_cm[0] = 0.5 * o->cfg.xlen; b->pos_cm[1] = 0.5 * o->cfg.ylen; b->pos_cm[2] = 0.5 * o->cfg.zlen;
  /* init_geomCoords create_geo.f90:106-132 */
  for (int v = 0; v < o->nv; ++v)
    for (int d = 0; d < 3; ++d) LAG(MELT_L_XYZV)[3 * v + d] = o->xyz0[3 * v + d] + b->pos_cm[d];
  centroids(o);
  /* tri_geo.f90:15-30 (E_thresh still 0 here) */
  edge_lengths(o, o->xyz0);
  calc_area(o, o->xyz0);
  calc_vert_area(o);
  skewness(o);
  tri_normal(o, o->xyz0, 0);
  vert_normal(o);
  /* init_mlsForce initmls.f90:1-27 */
  b->h_eulerian = (1.0 / o->dx1 + 1.0 / o->dx2 + 1.0 / o->dx3) / 3.0;
  b->celvol = 1.0 / (o->dx1 * o->dx2 * o->dx3);
  b->e_thresh = o->cfg.perc_ethresh * b->h_eulerian;
  b->v_thresh = o->cfg.v_on_ve_perc * b->celvol;
  for (int f = 0; f < o->nf; ++f) LAG(MELT_L_CFAC)[f] = (LAG(MELT_L_SUR)[f] * b->h_eulerian) / b->celvol;
  b->initial_tag = 1;
  return o->status;
}

/* ------------------------------------------------------------ accessors */
int orc_get_body(orc_ctx* o, melt_body* out) { *out = o->body; return 0; }
int orc_set_body(orc_ctx* o, const melt_body* in) { o->body = *in; return 0; }
static size_t lag_count(orc_ctx* o, int id) {
  int k = lag_kind(id);
  return (size_t)lag_len[id] * (k == 0 ? o->nv : k == 1 ? o->nf : o->ne);
}
int orc_get_lag(orc_ctx* o, int id, double* h) {
  if (id < 0 || id >= MELT_L_NLAG || !o->has_geom) return MELT_ERR_ARG;
  memcpy(h, o->lag[id], sizeof(double) * lag_count(o, id));
  return 0;
}
int orc_set_lag(orc_ctx* o, int id, const double* h) {
  if (id < 0 || id >= MELT_L_NLAG || !o->has_geom) return MELT_ERR_ARG;
  memcpy(o->lag[id], h, sizeof(double) * lag_count(o, id));
  return 0;
}
int orc_get_pind(orc_ctx* o, int32_t* h) { memcpy(h, o->pind, sizeof(int) * 6 * (size_t)o->nf); return 0; }
int orc_get_vert_of_face(orc_ctx* o, int32_t* h) { memcpy(h, o->vert_of_face, sizeof(int) * 3 * (size_t)o->nf); return 0; }
int orc_inv4_export(double* a) { return orc_inv4(a); }
double orc_det_sin_export(double x) { return orc_det_sin(x); }
double orc_det_exp_export(double x) { return orc_det_exp(x); }
