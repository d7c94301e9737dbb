/*
 * ORACLE (test infrastructure only) -- HIT forcing of the time loop, hit.f90:
 * InitRandomForce (:128-232), CalcHITRandomForce (:1-126) with the
 * reference's direct summation order, CalcABC_HITForce (:234-265) and the
 * bhat continuation (:267-356).  See orc.h for the generator substitution.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

static uint64_t mix(uint64_t seed, uint64_t d) {   /* SplitMix64 at counter d */
  uint64_t z = seed + (d + 1u) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double uniform(uint64_t seed, uint64_t d) { return (double)(mix(seed, d) >> 11) * (1.0 / 9007199254740992.0); }

void orc_hit_free(orc_ctx* o) {
  free(o->hit_wave); free(o->hit_bhat); free(o->hit_fhat);
  for (int c = 0; c < 2; ++c) { free(o->hit_ex[c]); free(o->hit_ey[c]); free(o->hit_ez[c]); }
  o->hit_wave = NULL; o->hit_bhat = o->hit_fhat = NULL;
  for (int c = 0; c < 2; ++c) o->hit_ex[c] = o->hit_ey[c] = o->hit_ez[c] = NULL;
  o->hit_nm = 0;
}

#define BH(a, d, l, m, n) (a)[2 * ((d) + 3 * ((l) + o->hit_nm * ((m) + o->hit_nm * (n))))]

int orc_hit_init(orc_ctx* o, double tl, double epsstar, double kf_on_kmin, uint64_t seed) {  /* hit.f90:128-232 */
  orc_hit_free(o);
  int nbuf = (int)floor(kf_on_kmin) + 1;
  int nm = (nbuf - 1) * 2 + 1;
  o->hit_nm = nm;
  o->hit_tl = tl; o->hit_eps = epsstar; o->hit_kf = kf_on_kmin; o->hit_seed = seed; o->hit_steps = 0;
  o->hit_wave = (int*)calloc(nm, sizeof(int));
  for (int i = 1; i <= nbuf - 1; ++i) o->hit_wave[i - 1] = -(nbuf - i);
  for (int i = 0; i <= nbuf - 1; ++i) o->hit_wave[nbuf + i - 1] = i;
  size_t nb = 2 * 3 * (size_t)nm * nm * nm;
  o->hit_bhat = (double*)calloc(nb, sizeof(double));
  o->hit_fhat = (double*)calloc(nb, sizeof(double));
  int n[3] = {o->n1, o->n2, o->n3};
  double len[3] = {o->cfg.xlen, o->cfg.ylen, o->cfg.zlen};
  double** tabs[3] = {o->hit_ex, o->hit_ey, o->hit_ez};
  for (int d = 0; d < 3; ++d)
    for (int centre = 0; centre < 2; ++centre) {
      double* t = (double*)calloc(2 * (size_t)n[d] * nm, sizeof(double));
      for (int i = 1; i <= n[d]; ++i)
        for (int w = 0; w < nm; ++w) {
          double kappa = (double)o->hit_wave[w] * 2.0 * o->pi / len[d];
          double x = d == 0 ? (centre ? orc_xm(o, i) : orc_xc(o, i)) : d == 1 ? (centre ? orc_ym(o, i) : orc_yc(o, i))
                                                                              : (centre ? orc_zm(o, i) : orc_zc(o, i));
          t[2 * ((i - 1) + (size_t)n[d] * w)] = cos(kappa * x);
          t[2 * ((i - 1) + (size_t)n[d] * w) + 1] = sin(kappa * x);
        }
      tabs[d][centre] = t;
    }
  return 0;
}

int orc_hit_force(orc_ctx* o) {  /* hit.f90:1-126 */
  if (!o->hit_nm) return MELT_ERR_ARG;
  int nm = o->hit_nm;
  double pi = o->pi, dt = o->dt, tl = o->hit_tl;
  double kf = o->hit_kf * 2.0 * pi / o->cfg.xlen;
  size_t nb = 2 * 3 * (size_t)nm * nm * nm;
  memset(o->hit_fhat, 0, nb * sizeof(double));
  uint64_t draw = 0;
  int nact = 0;
  for (int l = 0; l < nm; ++l)
    for (int m = 0; m < nm; ++m)
      for (int nn = 0; nn < nm; ++nn) {
        double kl = (double)o->hit_wave[l] * 2.0 * pi / o->cfg.xlen, km = (double)o->hit_wave[m] * 2.0 * pi / o->cfg.ylen,
               kn = (double)o->hit_wave[nn] * 2.0 * pi / o->cfg.zlen;
        double kmag = sqrt(kl * kl + km * km + kn * kn);
        if (kmag > 0.0 && kmag < kf) ++nact;
      }
  draw = o->hit_steps * (uint64_t)nact * 6u;
  for (int l = 0; l < nm; ++l) {
    double kl = (double)o->hit_wave[l] * 2.0 * pi / o->cfg.xlen;
    for (int m = 0; m < nm; ++m) {
      double km = (double)o->hit_wave[m] * 2.0 * pi / o->cfg.ylen;
      for (int nn = 0; nn < nm; ++nn) {
        double kn = (double)o->hit_wave[nn] * 2.0 * pi / o->cfg.zlen;
        double kmag = sqrt(kl * kl + km * km + kn * kn);
        if (kmag > 0.0 && kmag < kf) {
          for (int d = 0; d < 3; ++d) {
            double u1 = 1.0 - uniform(o->hit_seed, draw++);
            double u2 = uniform(o->hit_seed, draw++);
            double re = sqrt(-2.0 * log(u1)) * cos(2.0 * pi * u2), im = sqrt(-2.0 * log(u1)) * sin(2.0 * pi * u2);
            double f = 1.0 - dt / tl, a = sqrt(2.0 * o->hit_eps * dt / (tl * tl));
            BH(o->hit_bhat, d, l, m, nn) = BH(o->hit_bhat, d, l, m, nn) * f + re * a;
            BH(o->hit_bhat + 1, d, l, m, nn) = BH(o->hit_bhat + 1, d, l, m, nn) * f + im * a;
          }
          double kv[3] = {kl, km, kn};
          double dr = 0.0, di = 0.0;
          for (int d = 0; d < 3; ++d) { dr += kv[d] * BH(o->hit_bhat, d, l, m, nn); di += kv[d] * BH(o->hit_bhat + 1, d, l, m, nn); }
          for (int d = 0; d < 3; ++d) {
            BH(o->hit_fhat, d, l, m, nn) = BH(o->hit_bhat, d, l, m, nn) - kv[d] * dr / (kmag * kmag);
            BH(o->hit_fhat + 1, d, l, m, nn) = BH(o->hit_bhat + 1, d, l, m, nn) - kv[d] * di / (kmag * kmag);
          }
        }
      }
    }
  }
  o->hit_steps += 1;
  /* direct summation, nn / m / l outer, k / j / i inner (hit.f90:68-122) */
  static const int cx[3] = {0, 1, 1}, cy[3] = {1, 0, 1}, cz[3] = {1, 1, 0};   /* 0: faces, 1: centres */
  static const int fid[3] = {MELT_FORCX, MELT_FORCY, MELT_FORCZ};
  for (int c = 0; c < 3; ++c) {
    double* f = o->f[fid[c]];
    memset(f, 0, o->nc * sizeof(double));
    const double* ex = o->hit_ex[cx[c]];
    const double* ey = o->hit_ey[cy[c]];
    const double* ez = o->hit_ez[cz[c]];
    for (int nn = 0; nn < nm; ++nn)
      for (int m = 0; m < nm; ++m)
        for (int l = 0; l < nm; ++l) {
          double complex F = BH(o->hit_fhat, c, l, m, nn) + I * BH(o->hit_fhat + 1, c, l, m, nn);
#pragma omp parallel for schedule(static)
          for (int k = 1; k <= o->n3; ++k)
            for (int j = 1; j <= o->n2; ++j) {
              double complex ey_ = ey[2 * ((j - 1) + (size_t)o->n2 * m)] + I * ey[2 * ((j - 1) + (size_t)o->n2 * m) + 1];
              double complex ez_ = ez[2 * ((k - 1) + (size_t)o->n3 * nn)] + I * ez[2 * ((k - 1) + (size_t)o->n3 * nn) + 1];
              double complex buff1 = ey_ * ez_;
              for (int i = 1; i <= o->n1; ++i) {
                double complex ex_ = ex[2 * ((i - 1) + (size_t)o->n1 * l)] + I * ex[2 * ((i - 1) + (size_t)o->n1 * l) + 1];
                size_t cc = (size_t)(i - 1) + (size_t)o->n1 * ((size_t)(j - 1) + (size_t)o->n2 * (k - 1));
                f[cc] = f[cc] + creal(F * ex_ * buff1);
              }
            }
        }
  }
  return 0;
}

int orc_hit_state(orc_ctx* o, double* bhat, uint64_t* steps, int32_t set) {
  if (!o->hit_nm) return MELT_ERR_ARG;
  size_t nb = 2 * 3 * (size_t)o->hit_nm * o->hit_nm * o->hit_nm;
  if (set) { memcpy(o->hit_bhat, bhat, nb * sizeof(double)); o->hit_steps = *steps; }
  else { memcpy(bhat, o->hit_bhat, nb * sizeof(double)); *steps = o->hit_steps; }
  return 0;
}

int orc_hit_nmodes(orc_ctx* o, int32_t* nmodes, int32_t* nforced) {
  *nmodes = o->hit_nm;
  int n = 0;
  double kf = o->hit_kf * 2.0 * o->pi / o->cfg.xlen;
  for (int l = 0; l < o->hit_nm; ++l)
    for (int m = 0; m < o->hit_nm; ++m)
      for (int nn = 0; nn < o->hit_nm; ++nn) {
        double kl = (double)o->hit_wave[l] * 2.0 * o->pi / o->cfg.xlen, km = (double)o->hit_wave[m] * 2.0 * o->pi / o->cfg.ylen,
               kn = (double)o->hit_wave[nn] * 2.0 * o->pi / o->cfg.zlen;
        double kmag = sqrt(kl * kl + km * km + kn * kn);
        if (kmag > 0.0 && kmag < kf) ++n;
      }
  *nforced = n;
  return 0;
}

int orc_abc_force(orc_ctx* o, double c_hit) {  /* hit.f90:234-265 */
  double C = c_hit * ((2.0 * o->pi / o->cfg.xlen) * (2.0 * o->pi / o->cfg.xlen));
  for (int k = 1; k <= o->n3; ++k)
    for (int j = 1; j <= o->n2; ++j)
      for (int i = 1; i <= o->n1; ++i) {
        size_t c = (size_t)(i - 1) + (size_t)o->n1 * ((size_t)(j - 1) + (size_t)o->n2 * (k - 1));
        o->f[MELT_FORCX][c] = C * sin(2.0 * o->pi * orc_zm(o, k) / o->cfg.zlen) + C * cos(2.0 * o->pi * orc_ym(o, j) / o->cfg.ylen);
        o->f[MELT_FORCY][c] = C * sin(2.0 * o->pi * orc_xm(o, i) / o->cfg.xlen) + C * cos(2.0 * o->pi * orc_zm(o, k) / o->cfg.zlen);
        o->f[MELT_FORCZ][c] = C * sin(2.0 * o->pi * orc_ym(o, j) / o->cfg.ylen) + C * cos(2.0 * o->pi * orc_xm(o, i) / o->cfg.xlen);
      }
  return 0;
}
