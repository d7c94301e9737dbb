/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C FP64 restatement of the melt_IBM RK3 substep (kevzhong/melt_IBM, the
 * Fortran files cited function by function).  It is the checker for the HIP
 * library libmeltgpu.so: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.  Nothing in the product links it.
 *
 * The exported API mirrors include/melt_ibm.h with the prefix orc_ instead of
 * melt_, so the same test driver can run a case on both.
 *
 * Arithmetic follows the reference expression by expression (same
 * parenthesisation, no FMA contraction: built with -ffp-contract=off), with the
 * Fortran intrinsics restated: NINT rounds half away from zero, MODULO is the
 * non-negative remainder, SIGN(1,x) takes the sign bit of x.
 *
 * Deviations (each pinned by the check named with it in tests/test_oracle.py):
 *  - FFTW (phcalc.f90) is replaced by an own mixed-radix FFT; the Poisson result is
 *    pinned by its defining property (discrete Laplacian of phi == rhs, mean 0).
 *  - LAPACK dgetrf/dgetri and BLAS dgemm/dgemv are restated from the reference
 *    (netlib) algorithms; MKL's operation order is not pinned.
 *  - sin() in sdf_to_VOF (genSDF.f90:251) is an own double-precision sine
 *    (Cody-Waite reduction + Taylor polynomials, explicit fma) so that the GPU
 *    can reproduce VOFp bit for bit; tests bound it to 1 ulp of libm sin.
 *  - exp() in the MLS Gaussian weights (geom.f90:24-53) is an own double
 *    exponential for the same reason: the 4x4 moment matrices have condition
 *    numbers near 1e5, so a 1-ulp difference in a weight shows up at 1e-11 in the
 *    shape functions.  Tests bound it to 1 ulp of libm exp.
 *  - For_* ghost-plane folding, the mean force and all face/vertex sums follow
 *    the reference's serial order exactly.
 *  - The HIT forcing's Gaussian numbers (hit.f90:39-44, gfortran random_number)
 *    are Box-Muller pairs from SplitMix64 at a counter, the generator the
 *    library uses (gfortran's stream is not reproducible elsewhere); the OU
 *    statistics and the synthesis are pinned by closed forms.
 */
#ifndef ORC_H
#define ORC_H

#include <stddef.h>
#include <stdint.h>
#include "../include/melt_ibm.h"

typedef struct orc_ctx {
  melt_config cfg;
  int n1, n2, n3;            /* n1m, n2m, n3m */
  size_t nc;                 /* n1*n2*n3 */
  double dx1, dx2, dx3, dx1q, dx2q, dx3q, pi;
  double *ak1, *ak2, *ak3;   /* modified wavenumbers, phini.f90:4-59 */
  int *imv, *ipv, *jmv, *jpv;  /* 1-based periodic neighbours, cordin.f90:24-44 */
  double *f[MELT_NFIELDS];   /* fields; FOR_* carry k = 0..n3+1 (n3+2 planes) */
  double gam[3], rom[3], alm[3];
  double al, ga, ro, dt, beta;
  /* mesh (1-based connectivity, as the reference) */
  int nv, ne, nf, has_geom;
  double rad_p;
  int *vert_of_edge, *edge_of_face, *vert_of_face, *face_of_edge;
  int32_t *ghost_v, *ghost_e, *ghost_f;   /* isGhostVert/Edge/Face as 0/1 */
  int32_t *anchor_v, *flagged_e;          /* anchorVert, flagged_edge (allotri.f90:27-28) */
  double remesh_stats[8];                 /* see orc_get_remesh_stats */
  double *lag[MELT_L_NLAG];
  double *xyz0;
  double *r_x_prn_tri, *r_x_tau_n; /* 3*nf each */
  int *pind;                 /* 6*nf */
  melt_body body;
  /* HIT forcing (hit.f90): nmodes wave numbers per direction, bhat/Fhat as
   * (re, im) pairs in the reference's (3, nm, nm, nm) order, exp(i k x) tables */
  int hit_nm;
  double hit_tl, hit_eps, hit_kf;
  uint64_t hit_seed, hit_steps;
  int *hit_wave;
  double *hit_bhat, *hit_fhat;
  double *hit_ex[2], *hit_ey[2], *hit_ez[2];  /* [0] faces (xc...), [1] centres (xm...): (N, nm) complex */
  int status;
  char err[256];
} orc_ctx;

/* grid coordinates of cordin.f90:55-99 (1-based index, any integer) */
static inline double orc_xc(const orc_ctx* o, int i) { double x = (double)(i - 1) / (double)o->n1; return o->cfg.xlen * x; }
static inline double orc_yc(const orc_ctx* o, int j) { double x = (double)(j - 1) / (double)o->n2; return o->cfg.ylen * x; }
static inline double orc_zc(const orc_ctx* o, int k) { double x = (double)(k - 1) / (double)o->n3; return o->cfg.zlen * x; }
static inline double orc_xm(const orc_ctx* o, int i) { return (orc_xc(o, i) + orc_xc(o, i + 1)) * 0.5; }
static inline double orc_ym(const orc_ctx* o, int j) { return (orc_yc(o, j) + orc_yc(o, j + 1)) * 0.5; }
static inline double orc_zm(const orc_ctx* o, int k) { return (orc_zc(o, k) + orc_zc(o, k + 1)) * 0.5; }

static inline int orc_mod(int a, int n) { int r = a % n; return r < 0 ? r + n : r; }  /* Fortran MODULO */

/* field element, 1-based i,j in 1..n, k any (periodic ghost-plane semantics) */
#define OF(o, id, i, j, k) ((o)->f[id][(size_t)((i) - 1) + (size_t)(o)->n1 * ((size_t)((j) - 1) + (size_t)(o)->n2 * (size_t)orc_mod((k) - 1, (o)->n3))])
/* force arrays with their ghost planes k = 0..n3+1 */
#define OFG(o, id, i, j, k) ((o)->f[id][(size_t)((i) - 1) + (size_t)(o)->n1 * ((size_t)((j) - 1) + (size_t)(o)->n2 * (size_t)(k))])

/* internal helpers shared by the oracle's translation units */
void orc_set_error(orc_ctx* o, int code, const char* msg);
int orc_fnint(double x);
double orc_det_sin(double x);
double orc_det_exp(double x);
int orc_inv4(double a[16]);            /* dgetrf + dgetri, column-major; 0 or info */
int orc_inv3(double a[9]);
void orc_cross(double r[3], const double x[3], const double y[3]);
double orc_norm2(const double* v, int n);

/* exported API (mirrors include/melt_ibm.h, prefix orc_) used across units */
int orc_fill_field(orc_ctx* o, int id, double v);
int orc_set_stage(orc_ctx* o, int32_t ns);

/* IBM / Lagrangian units (orc_ibm.c) */
int orc_tagcells(orc_ctx* o);
int orc_interp_vof(orc_ctx* o);
int orc_find_centroid_indices(orc_ctx* o);
int orc_mlsweight(orc_ctx* o);
int orc_mlsforce(orc_ctx* o);
int orc_velforce(orc_ctx* o);
int orc_tempforce(orc_ctx* o);
int orc_particle(orc_ctx* o);
int orc_geometry_init(orc_ctx* o, double rad_p);
/* remeshing unit (orc_remesh.c) */
int orc_remesh_step(orc_ctx* o, double vol_melt, int* did);
void orc_hit_free(orc_ctx* o);

#endif
