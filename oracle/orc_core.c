/*
 * ORACLE (test infrastructure only) -- Eulerian part of the RK3 substep.
 * See orc.h for the rules this restatement follows.
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

void orc_set_error(orc_ctx* o, int code, const char* msg) {
  if (o->status == 0) {
    o->status = code;
    snprintf(o->err, sizeof o->err, "%s", msg);
  }
}

/* Fortran NINT: round half away from zero. */
int orc_fnint(double x) { return (int)lround(x); }

/* gfortran NORM2: scaled sum of squares (overflow-safe form). */
double orc_norm2(const double* v, int n) {
  double scale = 0.0, res = 1.0;
  for (int i = 0; i < n; ++i) {
    if (v[i] != 0.0) {
      double ax = fabs(v[i]);
      if (scale < ax) {
        double r = scale / ax;
        res = 1.0 + res * r * r;
        scale = ax;
      } else {
        double r = ax / scale;
        res += r * r;
      }
    }
  }
  return scale * sqrt(res);
}

void orc_cross(double r[3], const double x[3], const double y[3]) {  /* triaux.f90 cross */
  r[0] = x[1] * y[2] - y[1] * x[2];
  r[1] = x[2] * y[0] - y[2] * x[0];
  r[2] = x[0] * y[1] - y[0] * x[1];
}

/* ------------------------------------------------------------------ lifecycle */
static void* xcalloc(size_t n, size_t sz) {
  void* p = calloc(n ? n : 1, sz);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}

int orc_ctx_create(const melt_config* cfg, orc_ctx** out) {
  *out = NULL;
  if (!cfg || cfg->n1m < 4 || cfg->n2m < 4 || cfg->n3m < 4 || cfg->n1m != cfg->n2m) return MELT_ERR_ARG;
  orc_ctx* o = (orc_ctx*)xcalloc(1, sizeof *o);
  o->cfg = *cfg;
  o->n1 = cfg->n1m; o->n2 = cfg->n2m; o->n3 = cfg->n3m;
  o->nc = (size_t)o->n1 * o->n2 * o->n3;
  o->pi = 2.0 * asin(1.0);                                   /* papero.f90:140 */
  /* cordin.f90:64-70: inverse grid spacings */
  o->dx1 = (double)o->n1 / cfg->xlen;
  o->dx2 = (double)o->n2 / cfg->ylen;
  o->dx3 = (double)o->n3 / cfg->zlen;
  o->dx1q = o->dx1 * o->dx1; o->dx2q = o->dx2 * o->dx2; o->dx3q = o->dx3 * o->dx3;
  /* cordin.f90:24-44 neighbour maps */
  o->imv = (int*)xcalloc(o->n1 + 1, sizeof(int)); o->ipv = (int*)xcalloc(o->n1 + 1, sizeof(int));
  o->jmv = (int*)xcalloc(o->n2 + 1, sizeof(int)); o->jpv = (int*)xcalloc(o->n2 + 1, sizeof(int));
  for (int ic = 1; ic <= o->n1; ++ic) { o->imv[ic] = ic == 1 ? o->n1 : ic - 1; o->ipv[ic] = ic == o->n1 ? 1 : ic + 1; }
  for (int jc = 1; jc <= o->n2; ++jc) { o->jmv[jc] = jc == 1 ? o->n2 : jc - 1; o->jpv[jc] = jc == o->n2 ? 1 : jc + 1; }
  /* RK3 coefficients papero.f90:116-137 */
  o->gam[0] = 8.0 / 15.0; o->gam[1] = 5.0 / 12.0; o->gam[2] = 3.0 / 4.0;
  o->rom[0] = 0.0; o->rom[1] = -17.0 / 60.0; o->rom[2] = -5.0 / 12.0;
  for (int ns = 0; ns < 3; ++ns) o->alm[ns] = o->gam[ns] + o->rom[ns];
  /* phini.f90:4-59 InitFourierMetric (1-based wavenumber index) */
  int n[3] = {o->n1, o->n2, o->n3};
  double len[3] = {cfg->xlen, cfg->ylen, cfg->zlen};
  double** ak[3] = {&o->ak1, &o->ak2, &o->ak3};
  for (int d = 0; d < 3; ++d) {
    int nm = n[d], nmh = nm / 2 + 1;
    *ak[d] = (double*)xcalloc(nm + 1, sizeof(double));
    for (int i = 1; i <= nm; ++i) {
      double ao = i <= nmh ? (double)(i - 1) * 2.0 * o->pi : -(double)(nm - i + 1) * 2.0 * o->pi;
      double r = (double)nm / len[d];
      (*ak[d])[i] = 2.0 * (1.0 - cos(ao / nm)) * (r * r);
    }
  }
  for (int id = 0; id < MELT_NFIELDS; ++id) {
    int ghost = id >= MELT_FOR_XC && id <= MELT_FOR_TEMP;
    size_t cnt = ghost ? (size_t)o->n1 * o->n2 * (o->n3 + 2) : o->nc;
    o->f[id] = (double*)xcalloc(cnt, sizeof(double));
  }
  /* single-phase defaults of init.f90:30-33 */
  for (int id = MELT_VOFP; id <= MELT_VOFZ; ++id) orc_fill_field(o, id, 1.0);
  o->dt = 0.0;
  *out = o;
  return MELT_OK;
}

void orc_ctx_destroy(orc_ctx* o) {
  if (!o) return;
  for (int id = 0; id < MELT_NFIELDS; ++id) free(o->f[id]);
  for (int id = 0; id < MELT_L_NLAG; ++id) free(o->lag[id]);
  free(o->ak1); free(o->ak2); free(o->ak3);
  free(o->imv); free(o->ipv); free(o->jmv); free(o->jpv);
  free(o->vert_of_edge); free(o->edge_of_face); free(o->vert_of_face); free(o->face_of_edge);
  free(o->ghost_v); free(o->ghost_e); free(o->ghost_f); free(o->anchor_v); free(o->flagged_e);
  free(o->xyz0); free(o->r_x_prn_tri); free(o->r_x_tau_n); free(o->pind);
  orc_hit_free(o);
  free(o);
}

const char* orc_last_error(const orc_ctx* o) { return o->err; }
int orc_sync(orc_ctx* o) { return o->status; }
int orc_poll(orc_ctx* o, int32_t* done) { (void)o; *done = 1; return 0; }  /* host work is done on return */

int orc_set_dt(orc_ctx* o, double dt) { o->dt = dt; return 0; }

int orc_set_stage(orc_ctx* o, int32_t ns) {
  if (ns < 1 || ns > 3) return MELT_ERR_ARG;
  o->beta = o->dt / o->cfg.ren * 0.5;                        /* tsch.f90:18 */
  o->al = o->alm[ns - 1]; o->ga = o->gam[ns - 1]; o->ro = o->rom[ns - 1];
  return 0;
}

/* ------------------------------------------------------------ data movement */
static double* field_interior(orc_ctx* o, int id) {
  int ghost = id >= MELT_FOR_XC && id <= MELT_FOR_TEMP;
  return ghost ? o->f[id] + (size_t)o->n1 * o->n2 : o->f[id];
}

int orc_upload_field(orc_ctx* o, int id, const double* h, const melt_layout* L) {
  if (id < 0 || id >= MELT_NFIELDS) return MELT_ERR_ARG;
  double* d = field_interior(o, id);
  for (int k = 0; k < o->n3; ++k)
    for (int j = 0; j < o->n2; ++j)
      for (int i = 0; i < o->n1; ++i)
        d[i + (size_t)o->n1 * (j + (size_t)o->n2 * k)] = h[i + (size_t)L->ld1 * (j + (size_t)L->ld2 * (k + L->kghost))];
  return 0;
}

int orc_download_field(orc_ctx* o, int id, double* h, const melt_layout* L) {
  if (id < 0 || id >= MELT_NFIELDS) return MELT_ERR_ARG;
  const double* d = field_interior(o, id);
  int k0 = L->kghost ? -1 : 0, k1 = L->kghost ? o->n3 + 1 : o->n3;
  for (int k = k0; k < k1; ++k) {
    int ks = orc_mod(k, o->n3);
    for (int j = 0; j < o->n2; ++j)
      for (int i = 0; i < o->n1; ++i)
        h[i + (size_t)L->ld1 * (j + (size_t)L->ld2 * (k + L->kghost))] = d[i + (size_t)o->n1 * (j + (size_t)o->n2 * ks)];
  }
  return 0;
}

int orc_fill_field(orc_ctx* o, int id, double v) {
  int ghost = id >= MELT_FOR_XC && id <= MELT_FOR_TEMP;
  size_t cnt = ghost ? (size_t)o->n1 * o->n2 * (o->n3 + 2) : o->nc;
  for (size_t i = 0; i < cnt; ++i) o->f[id][i] = v;
  return 0;
}

/* ---------------------------------------------------------------- hdnl1..3 */
#define F(id, i, j, k) OF(o, id, i, j, k)

int orc_hdnl1(orc_ctx* o) {  /* hdnl1.f90:1-80 */
  double udx1 = o->dx1 * 0.25, udx2 = o->dx2 * 0.25, udx3 = o->dx3 * 0.25;
#pragma omp parallel for schedule(static)
  for (int kc = 1; kc <= o->n3; ++kc) {
    int km = kc - 1, kp = kc + 1;
    for (int jc = 1; jc <= o->n2; ++jc) {
      int jm = o->jmv[jc], jp = o->jpv[jc];
      for (int ic = 1; ic <= o->n1; ++ic) {
        int im = o->imv[ic], ip = o->ipv[ic];
        double h11 = ((F(MELT_VX, ip, jc, kc) + F(MELT_VX, ic, jc, kc)) * (F(MELT_VX, ip, jc, kc) + F(MELT_VX, ic, jc, kc)) -
                      (F(MELT_VX, im, jc, kc) + F(MELT_VX, ic, jc, kc)) * (F(MELT_VX, im, jc, kc) + F(MELT_VX, ic, jc, kc))) * udx1;
        double h12 = ((F(MELT_VY, ic, jp, kc) + F(MELT_VY, im, jp, kc)) * (F(MELT_VX, ic, jp, kc) + F(MELT_VX, ic, jc, kc)) -
                      (F(MELT_VY, ic, jc, kc) + F(MELT_VY, im, jc, kc)) * (F(MELT_VX, ic, jc, kc) + F(MELT_VX, ic, jm, kc))) * udx2;
        double h13 = ((F(MELT_VZ, ic, jc, kp) + F(MELT_VZ, im, jc, kp)) * (F(MELT_VX, ic, jc, kp) + F(MELT_VX, ic, jc, kc)) -
                      (F(MELT_VZ, ic, jc, kc) + F(MELT_VZ, im, jc, kc)) * (F(MELT_VX, ic, jc, kc) + F(MELT_VX, ic, jc, km))) * udx3;
        F(MELT_DQ, ic, jc, kc) = -(h11 + h12 + h13) + F(MELT_VOFX, ic, jc, kc) * F(MELT_FORCX, ic, jc, kc) / o->cfg.xlen;
      }
    }
  }
  return 0;
}

int orc_hdnl2(orc_ctx* o) {  /* hdnl2.f90:1-82 */
  double udx1 = o->dx1 * 0.25, udx2 = o->dx2 * 0.25, udx3 = o->dx3 * 0.25;
#pragma omp parallel for schedule(static)
  for (int kc = 1; kc <= o->n3; ++kc) {
    int km = kc - 1, kp = kc + 1;
    for (int jc = 1; jc <= o->n2; ++jc) {
      int jm = o->jmv[jc], jp = o->jpv[jc];
      for (int ic = 1; ic <= o->n1; ++ic) {
        int im = o->imv[ic], ip = o->ipv[ic];
        double h21 = ((F(MELT_VY, ip, jc, kc) + F(MELT_VY, ic, jc, kc)) * (F(MELT_VX, ip, jc, kc) + F(MELT_VX, ip, jm, kc)) -
                      (F(MELT_VY, ic, jc, kc) + F(MELT_VY, im, jc, kc)) * (F(MELT_VX, ic, jc, kc) + F(MELT_VX, ic, jm, kc))) * udx1;
        double h22 = ((F(MELT_VY, ic, jp, kc) + F(MELT_VY, ic, jc, kc)) * (F(MELT_VY, ic, jp, kc) + F(MELT_VY, ic, jc, kc)) -
                      (F(MELT_VY, ic, jm, kc) + F(MELT_VY, ic, jc, kc)) * (F(MELT_VY, ic, jm, kc) + F(MELT_VY, ic, jc, kc))) * udx2;
        double h23 = ((F(MELT_VZ, ic, jc, kp) + F(MELT_VZ, ic, jm, kp)) * (F(MELT_VY, ic, jc, kp) + F(MELT_VY, ic, jc, kc)) -
                      (F(MELT_VZ, ic, jc, kc) + F(MELT_VZ, ic, jm, kc)) * (F(MELT_VY, ic, jc, kc) + F(MELT_VY, ic, jc, km))) * udx3;
        F(MELT_DPH, ic, jc, kc) = -(h21 + h22 + h23) + F(MELT_VOFY, ic, jc, kc) * F(MELT_FORCY, ic, jc, kc) / o->cfg.ylen;
      }
    }
  }
  return 0;
}

int orc_hdnl3(orc_ctx* o) {  /* hdnl3.f90:1-91 */
  double udx1 = o->dx1 * 0.25, udx2 = o->dx2 * 0.25, udx3 = o->dx3 * 0.25;
#pragma omp parallel for schedule(static)
  for (int kc = 1; kc <= o->n3; ++kc) {
    int km = kc - 1, kp = kc + 1;
    for (int jc = 1; jc <= o->n2; ++jc) {
      int jmm = o->jmv[jc], jpp = o->jpv[jc];
      for (int ic = 1; ic <= o->n1; ++ic) {
        int im = o->imv[ic], ip = o->ipv[ic];
        double h31 = (((F(MELT_VX, ip, jc, kc) + F(MELT_VX, ip, jc, km)) * (F(MELT_VZ, ip, jc, kc) + F(MELT_VZ, ic, jc, kc))) -
                      ((F(MELT_VX, ic, jc, kc) + F(MELT_VX, ic, jc, km)) * (F(MELT_VZ, ic, jc, kc) + F(MELT_VZ, im, jc, kc)))) * udx1;
        double h32 = (((F(MELT_VY, ic, jpp, kc) + F(MELT_VY, ic, jpp, km)) * (F(MELT_VZ, ic, jpp, kc) + F(MELT_VZ, ic, jc, kc))) -
                      ((F(MELT_VY, ic, jc, kc) + F(MELT_VY, ic, jc, km)) * (F(MELT_VZ, ic, jc, kc) + F(MELT_VZ, ic, jmm, kc)))) * udx2;
        double h33 = (((F(MELT_VZ, ic, jc, kp) + F(MELT_VZ, ic, jc, kc)) * (F(MELT_VZ, ic, jc, kp) + F(MELT_VZ, ic, jc, kc))) -
                      ((F(MELT_VZ, ic, jc, kc) + F(MELT_VZ, ic, jc, km)) * (F(MELT_VZ, ic, jc, kc) + F(MELT_VZ, ic, jc, km)))) * udx3;
        double T_interp = (F(MELT_TEMP, ic, jc, kc) + F(MELT_TEMP, ic, jc, km)) * 0.5;
        double fbz = o->cfg.betagz * (T_interp - o->cfg.tliq);
        F(MELT_QCAP, ic, jc, kc) = (-(h31 + h32 + h33) + fbz) + F(MELT_VOFZ, ic, jc, kc) * F(MELT_FORCZ, ic, jc, kc) / o->cfg.zlen;
      }
    }
  }
  return 0;
}

int orc_hdnlte(orc_ctx* o) {  /* hdnlte.f90:1-160 (Nparticle = 1) */
  double udx1 = o->dx1 * 0.5, udx2 = o->dx2 * 0.5, udx3 = o->dx3 * 0.5;
  const double* pc = o->body.pos_cm;
  const double* vc = o->body.vel_cm;
  const double* om = o->body.omega_c;
  double xl = o->cfg.xlen, yl = o->cfg.ylen, zl = o->cfg.zlen;
#pragma omp parallel for schedule(static)
  for (int kc = 1; kc <= o->n3; ++kc) {
    int km = kc - 1, kp = kc + 1;
    for (int jc = 1; jc <= o->n2; ++jc) {
      int jm = o->jmv[jc], jp = o->jpv[jc];
      for (int ic = 1; ic <= o->n1; ++ic) {
        int ip = o->ipv[ic], im = o->imv[ic];
        double T = F(MELT_TEMP, ic, jc, kc);
        double h31 = (F(MELT_VX, ip, jc, kc) * (F(MELT_TEMP, ip, jc, kc) + T) - F(MELT_VX, ic, jc, kc) * (T + F(MELT_TEMP, im, jc, kc))) * udx1;
        double h32 = (F(MELT_VY, ic, jp, kc) * (F(MELT_TEMP, ic, jp, kc) + T) - F(MELT_VY, ic, jc, kc) * (T + F(MELT_TEMP, ic, jm, kc))) * udx2;
        double h33 = (F(MELT_VZ, ic, jc, kp) * (F(MELT_TEMP, ic, jc, kp) + T) - F(MELT_VZ, ic, jc, kc) * (T + F(MELT_TEMP, ic, jc, km))) * udx3;
        double vofp = F(MELT_VOFP, ic, jc, kc);
        double ht = -(h31 + h32 + h33) * vofp;
        /* solid-velocity terms; r min-image wrapped for u and v, not for w */
        double x[3] = {orc_xc(o, ic), orc_ym(o, jc), orc_zm(o, kc)}, r[3];
        for (int d = 0; d < 3; ++d) r[d] = x[d] - pc[d];
        r[0] = r[0] - xl * (double)orc_fnint(r[0] / xl);
        r[1] = r[1] - yl * (double)orc_fnint(r[1] / yl);
        r[2] = r[2] - zl * (double)orc_fnint(r[2] / zl);
        double u_imh = vc[0] + om[1] * r[2] - om[2] * r[1];
        h31 = u_imh * (F(MELT_TEMP, ip, jc, kc) - F(MELT_TEMP, im, jc, kc)) * udx1;
        x[0] = orc_xm(o, ic); x[1] = orc_yc(o, jc);
        for (int d = 0; d < 3; ++d) r[d] = x[d] - pc[d];
        r[0] = r[0] - xl * (double)orc_fnint(r[0] / xl);
        r[1] = r[1] - yl * (double)orc_fnint(r[1] / yl);
        r[2] = r[2] - zl * (double)orc_fnint(r[2] / zl);
        double v_jmh = vc[1] + om[2] * r[0] - om[0] * r[2];
        h32 = v_jmh * (F(MELT_TEMP, ic, jp, kc) - F(MELT_TEMP, ic, jm, kc)) * udx2;
        x[1] = orc_ym(o, jc); x[2] = orc_zc(o, kc);
        for (int d = 0; d < 3; ++d) r[d] = x[d] - pc[d];
        double w_kmh = vc[2] + om[0] * r[1] - om[1] * r[0];
        h33 = w_kmh * (F(MELT_TEMP, ic, jc, kp) - F(MELT_TEMP, ic, jc, km)) * udx3;
        F(MELT_HTEMP, ic, jc, kc) = ht - (h31 + h32 + h33) * (1.0 - vofp);
      }
    }
  }
  return 0;
}

/* --------------------------------------------------------- tridiag_periodic */
/* tridiag_periodic.f90:1-60, arrays 1-based of length m (n1i = 1, n1f = m). */
void orc_tridiag_periodic(const double* ami, const double* aci, const double* api, double* rrr, int m,
                          double* q, double* s, double* fei) {
  int n1i = 1, n1f = m, ia = n1i + 1, ii = n1i + n1f, i = 0;
  q[n1i] = -api[n1i] / aci[n1i];
  s[n1i] = -ami[n1i] / aci[n1i];
  double fn = rrr[n1f];
  rrr[n1i] = rrr[n1i] / aci[n1i];
  for (i = ia; i <= n1f; ++i) {
    double p = 1. / (aci[i] + ami[i] * q[i - 1]);
    q[i] = -api[i] * p;
    s[i] = -ami[i] * s[i - 1] * p;
    rrr[i] = (rrr[i] - ami[i] * rrr[i - 1]) * p;
  }
  s[n1f] = 1.;
  fei[n1f] = 0.;
  for (int l = ia; l <= n1f; ++l) {
    i = ii - l;
    s[i] = s[i] + q[i] * s[i + 1];
    fei[i] = rrr[i] + q[i] * fei[i + 1];
  }
  /* i == n1i here (last value taken in the loop above) */
  rrr[n1f] = (fn - api[i] * fei[n1i] - ami[i] * fei[n1f - 1]) / (api[i] * s[n1i] + ami[i] * s[n1f - 1] + aci[i]);
  for (int l = ia; l <= n1f; ++l) {
    i = ii - l;
    rrr[i] = rrr[n1f] * s[i] + fei[i];
  }
}

/* solve along one direction for all lines of a field: dir 0 = x (solxi.f90),
 * 1 = y (solxj.f90), 2 = z (solxk.f90 without the transpose, which only moves data) */
static void solve_dir(orc_ctx* o, double* rhs, int dir, double betadx) {
  int n[3] = {o->n1, o->n2, o->n3};
  size_t str[3] = {1, (size_t)o->n1, (size_t)o->n1 * o->n2};
  int m = n[dir];
  int a = dir == 0 ? 1 : 0, b = dir == 2 ? 1 : 2;
  int na = n[a], nb = n[b];
#pragma omp parallel
  {
    double* buf = (double*)malloc(sizeof(double) * (size_t)(m + 2) * 7);
    double *am = buf, *ac = am + m + 2, *ap = ac + m + 2, *fl = ap + m + 2, *q = fl + m + 2, *s = q + m + 2, *fe = s + m + 2;
#pragma omp for schedule(static)
    for (int lb = 0; lb < nb; ++lb)
      for (int la = 0; la < na; ++la) {
        size_t base = (size_t)la * str[a] + (size_t)lb * str[b];
        for (int c = 1; c <= m; ++c) {
          double ackl_b = 1.0 / (1.0 + 2.0 * betadx);
          ap[c] = -betadx * ackl_b;
          ac[c] = 1.0;
          am[c] = -betadx * ackl_b;
          fl[c] = rhs[base + (size_t)(c - 1) * str[dir]] * ackl_b;
        }
        orc_tridiag_periodic(am, ac, ap, fl, m, q, s, fe);
        for (int c = 1; c <= m; ++c) rhs[base + (size_t)(c - 1) * str[dir]] = fl[c];
      }
    free(buf);
  }
}

int orc_solxi(orc_ctx* o, double betadx) { solve_dir(o, o->f[MELT_RHS], 0, betadx); return 0; }
int orc_solxj(orc_ctx* o, double betadx) { solve_dir(o, o->f[MELT_RHS], 1, betadx); return 0; }
int orc_solxk(orc_ctx* o, int field, double betadx) {  /* solxk.f90:1-60 */
  if (field < 0 || field >= MELT_NFIELDS) return MELT_ERR_ARG;
  solve_dir(o, o->f[MELT_RHS], 2, betadx);
  double* q = o->f[field];
  const double* r = o->f[MELT_RHS];
  for (size_t c = 0; c < o->nc; ++c) q[c] = q[c] + r[c];
  return 0;
}

/* solxk.f90:1-60 with q the caller's array section (invtr1.f90:68): rhs keeps
 * the z solution, q(ic,jc,kc) += rhs over the interior (solxk.f90:44-50). */
int orc_solxk_array(orc_ctx* o, double* q, const melt_layout* L, double betadx) {
  solve_dir(o, o->f[MELT_RHS], 2, betadx);
  const double* r = o->f[MELT_RHS];
  int ld1 = L ? L->ld1 : o->n1, ld2 = L ? L->ld2 : o->n2, kg = L ? L->kghost : 0;
  for (int k = 0; k < o->n3; ++k)
    for (int j = 0; j < o->n2; ++j)
      for (int i = 0; i < o->n1; ++i) {
        size_t h = (size_t)i + (size_t)ld1 * ((size_t)j + (size_t)ld2 * (size_t)(k + kg));
        q[h] = q[h] + r[(size_t)i + (size_t)o->n1 * ((size_t)j + (size_t)o->n2 * k)];
      }
  return 0;
}

/* ------------------------------------------------------------------ invtr* */
int orc_invtr1(orc_ctx* o) {  /* invtr1.f90:1-77 */
  double alre = o->al / o->cfg.ren, udx1 = o->dx1 * o->al;
  double udx1q = o->dx1q, udx2q = o->dx2q, udx3q = o->dx3q;
#pragma omp parallel for schedule(static)
  for (int kc = 1; kc <= o->n3; ++kc) {
    int km = kc - 1, kp = kc + 1;
    for (int jc = 1; jc <= o->n2; ++jc) {
      int jm = o->jmv[jc], jp = o->jpv[jc];
      for (int ic = 1; ic <= o->n1; ++ic) {
        int im = o->imv[ic], ip = o->ipv[ic];
        double v = F(MELT_VX, ic, jc, kc);
        double d11 = (F(MELT_VX, ip, jc, kc) - 2.0 * v + F(MELT_VX, im, jc, kc)) * udx1q;
        double d22 = (F(MELT_VX, ic, jp, kc) - 2.0 * v + F(MELT_VX, ic, jm, kc)) * udx2q;
        double d33 = (F(MELT_VX, ic, jc, kp) - 2.0 * v + F(MELT_VX, ic, jc, km)) * udx3q;
        double dcvx = d11 + d22 + d33;
        double dpx11 = (F(MELT_PR, ic, jc, kc) - F(MELT_PR, im, jc, kc)) * udx1;
        F(MELT_RHS, ic, jc, kc) = (o->ga * F(MELT_DQ, ic, jc, kc) + o->ro * F(MELT_RU1, ic, jc, kc) + alre * dcvx - dpx11) * o->dt;
        F(MELT_RU1, ic, jc, kc) = F(MELT_DQ, ic, jc, kc);
      }
    }
  }
  orc_solxi(o, o->beta * o->al * o->dx1q);
  orc_solxj(o, o->beta * o->al * o->dx2q);
  orc_solxk(o, MELT_VX, o->beta * o->al * o->dx3q);
  return 0;
}

int orc_invtr2(orc_ctx* o) {  /* invtr2.f90:1-71 */
  double alre = o->al / o->cfg.ren, udx2 = o->dx2 * o->al;
  double udx1q = o->dx1q, udx2q = o->dx2q, udx3q = o->dx3q;
#pragma omp parallel for schedule(static)
  for (int kc = 1; kc <= o->n3; ++kc) {
    int km = kc - 1, kp = kc + 1;
    for (int jc = 1; jc <= o->n2; ++jc) {
      int jm = o->jmv[jc], jp = o->jpv[jc];
      for (int ic = 1; ic <= o->n1; ++ic) {
        int im = o->imv[ic], ip = o->ipv[ic];
        double v = F(MELT_VY, ic, jc, kc);
        double d11 = (F(MELT_VY, ip, jc, kc) - 2.0 * v + F(MELT_VY, im, jc, kc)) * udx1q;
        double d22 = (F(MELT_VY, ic, jp, kc) - 2.0 * v + F(MELT_VY, ic, jm, kc)) * udx2q;
        double d33 = (F(MELT_VY, ic, jc, kp) - 2.0 * v + F(MELT_VY, ic, jc, km)) * udx3q;
        double dcvy = d11 + d22 + d33;
        double dpx22 = (F(MELT_PR, ic, jc, kc) - F(MELT_PR, ic, jm, kc)) * udx2;
        F(MELT_RHS, ic, jc, kc) = (o->ga * F(MELT_DPH, ic, jc, kc) + o->ro * F(MELT_RU2, ic, jc, kc) + alre * dcvy - dpx22) * o->dt;
        F(MELT_RU2, ic, jc, kc) = F(MELT_DPH, ic, jc, kc);
      }
    }
  }
  orc_solxi(o, o->beta * o->al * o->dx1q);
  orc_solxj(o, o->beta * o->al * o->dx2q);
  orc_solxk(o, MELT_VY, o->beta * o->al * o->dx3q);
  return 0;
}

int orc_invtr3(orc_ctx* o) {  /* invtr3.f90:1-65 */
  double alre = o->al / o->cfg.ren, udx3 = o->al * o->dx3;
  double udx1q = o->dx1q, udx2q = o->dx2q, udx3q = o->dx3q;
#pragma omp parallel for schedule(static)
  for (int kc = 1; kc <= o->n3; ++kc) {
    int km = kc - 1, kp = kc + 1;
    for (int jc = 1; jc <= o->n2; ++jc) {
      int jm = o->jmv[jc], jp = o->jpv[jc];
      for (int ic = 1; ic <= o->n1; ++ic) {
        int im = o->imv[ic], ip = o->ipv[ic];
        double v = F(MELT_VZ, ic, jc, kc);
        double dvz1 = (F(MELT_VZ, im, jc, kc) - 2.0 * v + F(MELT_VZ, ip, jc, kc)) * udx1q;
        double dvz2 = (F(MELT_VZ, ic, jm, kc) - 2.0 * v + F(MELT_VZ, ic, jp, kc)) * udx2q;
        double dvz3 = (F(MELT_VZ, ic, jc, kp) - 2.0 * v + F(MELT_VZ, ic, jc, km)) * udx3q;
        double dcvz = dvz2 + dvz3 + dvz1;
        double dpx33 = (F(MELT_PR, ic, jc, kc) - F(MELT_PR, ic, jc, km)) * udx3;
        F(MELT_RHS, ic, jc, kc) = (o->ga * F(MELT_QCAP, ic, jc, kc) + o->ro * F(MELT_RU3, ic, jc, kc) + alre * dcvz - dpx33) * o->dt;
        F(MELT_RU3, ic, jc, kc) = F(MELT_QCAP, ic, jc, kc);
      }
    }
  }
  orc_solxi(o, o->beta * o->al * o->dx1q);
  orc_solxj(o, o->beta * o->al * o->dx2q);
  orc_solxk(o, MELT_VZ, o->beta * o->al * o->dx3q);
  return 0;
}

int orc_invtrte(orc_ctx* o) {  /* invtrte.f90:1-67 */
  double pec = o->cfg.ren * o->cfg.prandtl;                  /* papero.f90:110 */
  double udx1q = o->dx1q, udx2q = o->dx2q, udx3q = o->dx3q;
#pragma omp parallel for schedule(static)
  for (int kc = 1; kc <= o->n3; ++kc) {
    int km = kc - 1, kp = kc + 1;
    for (int jc = 1; jc <= o->n2; ++jc) {
      int jm = o->jmv[jc], jp = o->jpv[jc];
      for (int ic = 1; ic <= o->n1; ++ic) {
        int im = o->imv[ic], ip = o->ipv[ic];
        double t = F(MELT_TEMP, ic, jc, kc);
        double dte1 = (F(MELT_TEMP, im, jc, kc) - 2.0 * t + F(MELT_TEMP, ip, jc, kc)) * udx1q;
        double dte2 = (F(MELT_TEMP, ic, jm, kc) - 2.0 * t + F(MELT_TEMP, ic, jp, kc)) * udx2q;
        double dte3 = (F(MELT_TEMP, ic, jc, kp) - 2.0 * t + F(MELT_TEMP, ic, jc, km)) * udx3q;
        double dcte = dte2 + dte3 + dte1;
        F(MELT_RHS, ic, jc, kc) = o->dt * (o->ga * F(MELT_HTEMP, ic, jc, kc) + o->ro * F(MELT_RUT, ic, jc, kc)) + (o->al / pec) * dcte * o->dt;
        F(MELT_RUT, ic, jc, kc) = F(MELT_HTEMP, ic, jc, kc);
      }
    }
  }
  orc_solxi(o, o->dt / pec * 0.5 * o->al * o->dx1q);
  orc_solxj(o, o->dt / pec * 0.5 * o->al * o->dx2q);
  orc_solxk(o, MELT_TEMP, o->dt / pec * 0.5 * o->al * o->dx3q);
  return 0;
}

/* ---------------------------------------------------------------- divg etc */
int orc_divg(orc_ctx* o) {  /* divg.f90:1-22 */
  double usdtal = 1.0 / (o->dt * o->al);
#pragma omp parallel for schedule(static)
  for (int kc = 1; kc <= o->n3; ++kc) {
    int kp = kc + 1;
    for (int jc = 1; jc <= o->n2; ++jc) {
      int jp = o->jpv[jc];
      for (int ic = 1; ic <= o->n1; ++ic) {
        int ip = o->ipv[ic];
        double dq = (F(MELT_VX, ip, jc, kc) - F(MELT_VX, ic, jc, kc)) * o->dx1 + (F(MELT_VY, ic, jp, kc) - F(MELT_VY, ic, jc, kc)) * o->dx2 +
                    (F(MELT_VZ, ic, jc, kp) - F(MELT_VZ, ic, jc, kc)) * o->dx3;
        F(MELT_DPH, ic, jc, kc) = dq * usdtal;
      }
    }
  }
  return 0;
}

int orc_divgck(orc_ctx* o, double* qmax) {  /* divg.f90:26-56 */
  double m = 0.0;
  for (int kc = 1; kc <= o->n3; ++kc)
    for (int jc = 1; jc <= o->n2; ++jc)
      for (int ic = 1; ic <= o->n1; ++ic) {
        int ip = o->ipv[ic], jp = o->jpv[jc], kp = kc + 1;
        double dq = (F(MELT_VX, ip, jc, kc) - F(MELT_VX, ic, jc, kc)) * o->dx1 + (F(MELT_VY, ic, jp, kc) - F(MELT_VY, ic, jc, kc)) * o->dx2 +
                    (F(MELT_VZ, ic, jc, kp) - F(MELT_VZ, ic, jc, kc)) * o->dx3;
        m = fmax(fabs(dq), m);
      }
  *qmax = m;
  return 0;
}

int orc_cfl(orc_ctx* o, double* cflm) {  /* cfl.f90:1-30 (padding planes read as the periodic copy) */
  double c = 0.00000001;
  for (int k = 1; k <= o->n3; ++k)
    for (int j = 1; j <= o->n2; ++j)
      for (int i = 1; i <= o->n1; ++i) {
        int ip = o->ipv[i], jp = o->jpv[j];
        double q = fabs((F(MELT_VX, i, j, k) + F(MELT_VX, ip, j, k)) * 0.5 * o->dx1) + fabs((F(MELT_VY, i, j, k) + F(MELT_VY, i, jp, k)) * 0.5 * o->dx2) +
                   fabs((F(MELT_VZ, i, j, k) + F(MELT_VZ, i, j, k + 1)) * 0.5 * o->dx3);
        c = fmax(c, q);
      }
  *cflm = c;
  return 0;
}

int orc_updvp(orc_ctx* o) {  /* updvp.f90:8-35 */
  double udx1 = o->al * o->dt * o->dx1, udx2 = o->al * o->dt * o->dx2, udx3 = o->al * o->dt * o->dx3;
#pragma omp parallel for schedule(static)
  for (int kc = 1; kc <= o->n3; ++kc) {
    int km = kc - 1;
    for (int jc = 1; jc <= o->n2; ++jc) {
      int jm = o->jmv[jc];
      for (int ic = 1; ic <= o->n1; ++ic) {
        int im = o->imv[ic];
        double l = F(MELT_DPH, ic, jc, kc);
        F(MELT_VX, ic, jc, kc) = F(MELT_VX, ic, jc, kc) - (l - F(MELT_DPH, im, jc, kc)) * udx1;
        F(MELT_VY, ic, jc, kc) = F(MELT_VY, ic, jc, kc) - (l - F(MELT_DPH, ic, jm, kc)) * udx2;
        F(MELT_VZ, ic, jc, kc) = F(MELT_VZ, ic, jc, kc) - (l - F(MELT_DPH, ic, jc, km)) * udx3;
      }
    }
  }
  return 0;
}

int orc_prcalc(orc_ctx* o) {  /* prcalc.f90:5-35 */
  double be = o->al * o->beta;
#pragma omp parallel for schedule(static)
  for (int kc = 1; kc <= o->n3; ++kc) {
    int kp = kc + 1, km = kc - 1;
    for (int jc = 1; jc <= o->n2; ++jc) {
      int jm = o->jmv[jc], jp = o->jpv[jc];
      for (int ic = 1; ic <= o->n1; ++ic) {
        int im = o->imv[ic], ip = o->ipv[ic];
        double d = F(MELT_DPH, ic, jc, kc);
        F(MELT_PR, ic, jc, kc) = F(MELT_PR, ic, jc, kc) + d -
                                 be * ((F(MELT_DPH, ip, jc, kc) - 2.0 * d + F(MELT_DPH, im, jc, kc)) * o->dx1q +
                                       (F(MELT_DPH, ic, jp, kc) - 2.0 * d + F(MELT_DPH, ic, jm, kc)) * o->dx2q +
                                       (F(MELT_DPH, ic, jc, kp) - 2.0 * d + F(MELT_DPH, ic, jc, km)) * o->dx3q);
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ FFT */
/* Mixed radix (2, 3) recursive decimation-in-time DFT, sign -1 forward.
 * X_k = sum_n x_n exp(sign * 2 pi i n k / N). */
static void dft_rec(const double complex* in, size_t istr, double complex* out, double complex* tmp, int n,
                    const double complex* tw, int ntot, int twmul) {
  if (n == 1) { out[0] = in[0]; return; }
  int r = (n % 2 == 0) ? 2 : (n % 3 == 0 ? 3 : n);
  int m = n / r;
  for (int q = 0; q < r; ++q) dft_rec(in + q * istr, istr * r, tmp + q * m, out, m, tw, ntot, twmul * r);
  for (int k = 0; k < m; ++k)
    for (int s = 0; s < r; ++s) {
      double complex acc = 0.0;
      for (int q = 0; q < r; ++q) {
        long e = ((long)q * (k + (long)s * m) * twmul) % ntot;
        acc += tmp[q * m + k] * tw[e];
      }
      out[k + s * m] = acc;
    }
}

static void dft_line(double complex* x, size_t str, int n, int sign, const double complex* twf, const double complex* twb,
                     double complex* a, double complex* b, double complex* c) {
  for (int i = 0; i < n; ++i) a[i] = x[(size_t)i * str];
  dft_rec(a, 1, b, c, n, sign < 0 ? twf : twb, n, 1);
  for (int i = 0; i < n; ++i) x[(size_t)i * str] = b[i];
}

static void fft3d(double complex* x, int n1, int n2, int n3, int sign) {
  int n[3] = {n1, n2, n3};
  size_t str[3] = {1, (size_t)n1, (size_t)n1 * n2};
  for (int d = 0; d < 3; ++d) {
    int m = n[d];
    double complex* twf = (double complex*)malloc(sizeof(double complex) * m);
    double complex* twb = (double complex*)malloc(sizeof(double complex) * m);
    for (int t = 0; t < m; ++t) {
      double ang = 2.0 * 3.14159265358979323846 * (double)t / (double)m;
      twf[t] = cos(ang) - I * sin(ang);
      twb[t] = cos(ang) + I * sin(ang);
    }
    int a = d == 0 ? 1 : 0, b = d == 2 ? 1 : 2;
#pragma omp parallel
    {
      double complex* w = (double complex*)malloc(sizeof(double complex) * 3 * m);
#pragma omp for schedule(static)
      for (int lb = 0; lb < n[b]; ++lb)
        for (int la = 0; la < n[a]; ++la)
          dft_line(x + (size_t)la * str[a] + (size_t)lb * str[b], str[d], m, sign, twf, twb, w, w + m, w + 2 * m);
      free(w);
    }
    free(twf); free(twb);
  }
}

int orc_phcalc(orc_ctx* o) {  /* phcalc.f90:11-97 */
  size_t nc = o->nc;
  double complex* x = (double complex*)malloc(sizeof(double complex) * nc);
  double* dph = o->f[MELT_DPH];
  for (size_t c = 0; c < nc; ++c) x[c] = dph[c];
  fft3d(x, o->n1, o->n2, o->n3, -1);
  double nrm = (double)((long)o->n1 * o->n2 * o->n3);
  for (int k = 1; k <= o->n3; ++k)
    for (int j = 1; j <= o->n2; ++j)
      for (int i = 1; i <= o->n1; ++i) {
        size_t c = (size_t)(i - 1) + (size_t)o->n1 * ((size_t)(j - 1) + (size_t)o->n2 * (size_t)(k - 1));
        double complex v = x[c] / nrm;
        v = v / (-o->ak3[k] - o->ak1[i] - o->ak2[j]);
        x[c] = v;
      }
  x[0] = 0.0;
  fft3d(x, o->n1, o->n2, o->n3, +1);
  for (size_t c = 0; c < nc; ++c) dph[c] = creal(x[c]);
  free(x);
  return 0;
}

/* ---------------------------------------------------------------- inqpr */
int orc_inqpr(orc_ctx* o) {  /* inqpr.f90:3-44 */
  orc_fill_field(o, MELT_VX, 0.0); orc_fill_field(o, MELT_VY, 0.0); orc_fill_field(o, MELT_VZ, 0.0);
  for (int kc = 1; kc <= o->n3; ++kc)
    for (int ic = 1; ic <= o->n1; ++ic)
      for (int jc = 1; jc <= o->n2; ++jc) {
        double v[3] = {orc_xm(o, ic) - o->body.pos_cm[0], orc_ym(o, jc) - o->body.pos_cm[1], orc_zm(o, kc) - o->body.pos_cm[2]};
        double rr = orc_norm2(v, 3);
        F(MELT_TEMP, ic, jc, kc) = o->cfg.tliq - (o->cfg.tliq - o->cfg.tsol) * 0.5 * (1.0 - tanh((rr - o->rad_p) * o->dx1 / 2.0));
      }
  return 0;
}

/* ------------------------------------------------------------ diagnostics */
int orc_get_dt(orc_ctx* o, double cflfix, double* dt_out) {  /* cfl.f90:33-77 */
  double buf = 0.0;
  for (int k = 1; k <= o->n3; ++k)
    for (int j = 1; j <= o->n2; ++j)
      for (int i = 1; i <= o->n1; ++i) {
        buf = fmax(buf, fabs(F(MELT_VX, i, j, k)) * o->dt * o->dx1 + fabs(F(MELT_VY, i, j, k)) * o->dt * o->dx2 +
                            fabs(F(MELT_VZ, i, j, k)) * o->dt * o->dx3);
        buf = fmax(buf, 6.0 / o->cfg.ren * o->dt * (o->dx1 * o->dx1));
      }
  *dt_out = buf > 0.0 ? o->dt * cflfix / buf : o->dt;
  return 0;
}

int orc_dissipation(orc_ctx* o, double* out) {  /* diss.f90:11-201 */
  int moving = o->has_geom && o->cfg.imlsfor == 1 && (o->cfg.imlsstr == 1 || o->cfg.imelt == 1);
  if (moving) {
    int rc = orc_tagcells(o);
    if (rc) return rc;
  }
  double diss_v = 0.0, chi_v = 0.0, kenerg = 0.0, sumvof = 0.0;
  double udx1 = o->dx1, udx2 = o->dx2, udx3 = o->dx3;
  double pec = o->cfg.ren * o->cfg.prandtl;
  for (int kc = 1; kc <= o->n3; ++kc) {
    int kp = kc + 1, km = kc - 1;
    for (int jc = 1; jc <= o->n2; ++jc) {
      int jm = o->jmv[jc], jp = o->jpv[jc];
      for (int ic = 1; ic <= o->n1; ++ic) {
        int im = o->imv[ic], ip = o->ipv[ic];
        double h11 = (F(MELT_VX, ip, jc, kc) - F(MELT_VX, ic, jc, kc)) * udx1;
        double h12 = ((F(MELT_VX, ic, jp, kc) + F(MELT_VX, ip, jp, kc)) - (F(MELT_VX, ic, jm, kc) + F(MELT_VX, ip, jm, kc))) * 0.25 * udx2;
        double h13 = ((F(MELT_VX, ic, jc, kp) + F(MELT_VX, ip, jc, kp)) - (F(MELT_VX, ic, jc, km) + F(MELT_VX, ip, jc, km))) * 0.25 * udx3;
        double h21 = ((F(MELT_VY, ip, jc, kc) + F(MELT_VY, ip, jp, kc)) - (F(MELT_VY, im, jc, kc) + F(MELT_VY, im, jp, kc))) * 0.25 * udx1;
        double h22 = (F(MELT_VY, ic, jp, kc) - F(MELT_VY, ic, jc, kc)) * udx2;
        double h23 = ((F(MELT_VY, ic, jc, kp) + F(MELT_VY, ic, jp, kp)) - (F(MELT_VY, ic, jc, km) + F(MELT_VY, ic, jp, km))) * 0.25 * udx3;
        double h31 = ((F(MELT_VZ, ip, jc, kc) + F(MELT_VZ, ip, jc, kp)) - (F(MELT_VZ, im, jc, kc) + F(MELT_VZ, im, jc, kp))) * 0.25 * udx1;
        double h32 = ((F(MELT_VZ, ic, jp, kc) + F(MELT_VZ, ic, jp, kp)) - (F(MELT_VZ, ic, jm, kc) + F(MELT_VZ, ic, jm, kp))) * 0.25 * udx2;
        double h33 = (F(MELT_VZ, ic, jc, kp) - F(MELT_VZ, ic, jc, kc)) * udx3;
        double tke = (F(MELT_VX, ic, jc, kc) * F(MELT_VX, ic, jc, kc) + F(MELT_VX, ip, jc, kc) * F(MELT_VX, ip, jc, kc) +
                      F(MELT_VY, ic, jc, kc) * F(MELT_VY, ic, jc, kc) + F(MELT_VY, ic, jp, kc) * F(MELT_VY, ic, jp, kc) +
                      F(MELT_VZ, ic, jc, kc) * F(MELT_VZ, ic, jc, kc) + F(MELT_VZ, ic, jc, kp) * F(MELT_VZ, ic, jc, kp)) * 0.5;
        tke = tke * 0.5;
        double d = h11 * h11 + h22 * h22 + h33 * h33 + 0.5 * ((h21 + h12) * (h21 + h12)) + 0.5 * ((h31 + h13) * (h31 + h13)) +
                   0.5 * ((h32 + h23) * (h32 + h23));
        d = d * 2.0 / o->cfg.ren;
        double ax = 0.5 * o->dx1 * (F(MELT_TEMP, ip, jc, kc) - F(MELT_TEMP, im, jc, kc));
        double ay = 0.5 * o->dx2 * (F(MELT_TEMP, ic, jp, kc) - F(MELT_TEMP, ic, jm, kc));
        double az = 0.5 * o->dx3 * (F(MELT_TEMP, ic, jc, kp) - F(MELT_TEMP, ic, jc, km));
        double chi = (ax * ax + ay * ay + az * az) / pec;
        if (F(MELT_VOFP, ic, jc, kc) == 1.0) {
          sumvof = sumvof + F(MELT_VOFP, ic, jc, kc);
          kenerg = kenerg + tke;
          diss_v = diss_v + d;
          chi_v = chi_v + chi;
        }
      }
    }
  }
  kenerg = kenerg / sumvof;
  double nu = 1.0 / o->cfg.ren;
  diss_v = diss_v / sumvof;
  chi_v = chi_v / sumvof;
  double urms = sqrt(2.0 / 3.0 * kenerg);
  double eta = pow(nu * nu * nu / diss_v, 0.25);
  double kmax_eta = o->pi * o->dx1 * eta;
  double re_l = kenerg * kenerg / (nu * diss_v);
  double lambda_t = sqrt(15.0 * nu * (urms * urms) / diss_v);
  double re_lam = urms * lambda_t / nu;
  double l_int = sqrt(kenerg * kenerg * kenerg) / diss_v;
  double v[11] = {diss_v, kenerg, urms, eta, lambda_t, re_l, re_lam, l_int, chi_v, kmax_eta, sumvof};
  for (int q = 0; q < 11; ++q) out[q] = v[q];
  return o->status;
}

/* ----------------------------------------------------------------- tschem */
/* one substep of tsch.f90:33-90 (ghost exchanges are implicit: periodic reads) */
int orc_substep(orc_ctx* o, int32_t ns) {
  int rc = orc_set_stage(o, ns);
  if (rc) return rc;
  int stationary = !(o->cfg.imlsfor == 1 && (o->cfg.imlsstr == 1 || o->cfg.imelt == 1));
  if (o->has_geom && o->cfg.imlsfor == 1 && !stationary) orc_tagcells(o);
  orc_hdnl1(o); orc_hdnl2(o); orc_hdnl3(o); orc_hdnlte(o);
  orc_invtr1(o); orc_invtr2(o); orc_invtr3(o); orc_invtrte(o);
  if (o->has_geom && o->cfg.imlsfor == 1) orc_particle(o);
  orc_divg(o);
  orc_phcalc(o);
  orc_updvp(o);
  orc_prcalc(o);
  return o->status;
}

int orc_tschem(orc_ctx* o) {  /* tsch.f90:1-99 */
  for (int k = 0; k < 3; ++k) { o->body.fp[k] = 0.0; o->body.ftau[k] = 0.0; o->body.torq_p[k] = 0.0; o->body.torq_tau[k] = 0.0; }
  for (int ns = 1; ns <= 3; ++ns) {
    int rc = orc_substep(o, ns);
    if (rc) return rc;
  }
  return 0;
}

/* threads the OpenMP loops of the restatement use (for the CPU baseline record) */
#ifdef _OPENMP
#include <omp.h>
int orc_threads(void) { return omp_get_max_threads(); }
#else
int orc_threads(void) { return 1; }
#endif
