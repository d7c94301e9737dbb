// Lagrangian part of particle.f90 on gfx950: face -> vertex heat flux and the
// Stefan vertex update (mls_heatFlux.f90:302-319, stefanCondition.f90), surface
// loads (mls_structLoads.f90:51-260), Newton-Euler rigid-body motion
// (RigidAuxRoutines.f90), and the geometry refresh of particle.f90:163-193
// (triaux.f90).  All O(N_tri) and FP64.  Vertex sums gather over a
// vertex -> face list in ascending face order, which is the order the reference
// accumulates them in; face sums use a fixed-shape single-block reduction, so
// every result is reproducible run to run.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.h"

namespace {

constexpr int RB = 256;  // reduction block

struct LagArgs {
  double* L[MELT_L_NLAG];
  const int* vof;   // 0-based
  const int* eof;   // 0-based
  const int* voe;   // 0-based
  const int* v2f_ptr;
  const int* v2f;
  const unsigned char* gv;
  const unsigned char* ge;
  const unsigned char* gf;
  int nv, ne, nf;
};

// ---------------------------------------------------------- face sums
enum { SUM_RIGID = 1, SUM_AREA_VOL_RIGID = 3 };

__device__ inline double face_volume(const double* x, const int* vof, int f) {  // calculate_volume triaux.f90:53-80
  int v1 = vof[3 * f], v2 = vof[3 * f + 1], v3 = vof[3 * f + 2];
  return x[3 * v1] * (x[3 * v2 + 1] * x[3 * v3 + 2] - x[3 * v2 + 2] * x[3 * v3 + 1]) +
         x[3 * v2] * (x[3 * v3 + 1] * x[3 * v1 + 2] - x[3 * v3 + 2] * x[3 * v1 + 1]) +
         x[3 * v3] * (x[3 * v1 + 1] * x[3 * v2 + 2] - x[3 * v1 + 2] * x[3 * v2 + 1]);
}

// calc_rigidBody_params per-face terms (RigidAuxRoutines.f90:685-712)
__device__ inline void face_rigid(const double* xyz, const int* vof, int f, double o[10]) {
  const double* p1 = &xyz[3 * vof[3 * f]];
  const double* p2 = &xyz[3 * vof[3 * f + 1]];
  const double* p3 = &xyz[3 * vof[3 * f + 2]];
  double x1 = p1[0], y1 = p1[1], z1 = p1[2], x2 = p2[0], y2 = p2[1], z2 = p2[2], x3 = p3[0], y3 = p3[1], z3 = p3[2];
  double v = x1 * y2 * z3 + y1 * z2 * x3 + x2 * y3 * z1 - (x3 * y2 * z1 + x2 * y1 * z3 + y3 * z2 * x1);
  double x4 = x1 + x2 + x3, y4 = y1 + y2 + y3, z4 = z1 + z2 + z3;
  o[0] = v;
  o[1] = v * x4; o[2] = v * y4; o[3] = v * z4;
  o[4] = v * (x1 * x1 + x2 * x2 + x3 * x3 + x4 * x4);
  o[5] = v * (y1 * y1 + y2 * y2 + y3 * y3 + y4 * y4);
  o[6] = v * (z1 * z1 + z2 * z2 + z3 * z3 + z4 * z4);
  o[7] = v * (y1 * x1 + y2 * x2 + y3 * x3 + y4 * x4);
  o[8] = v * (z1 * x1 + z2 * x2 + z3 * x3 + z4 * x4);
  o[9] = v * (z1 * y1 + z2 * y2 + z3 * y3 + z4 * y4);
}


// Stage 1 of a deterministic two-stage sum: block b's threads walk faces
// b*RB + tid, + gridDim*RB, ... in order, then a fixed tree writes the block's
// partial to part[q * gridDim + b]; k_sum_partials adds the partials in a
// fixed tree (a second launch: finishing in the last block would need a
// device-scope fence per block, which writes back the L2 of every XCD and
// measured slower).  Output layout per mode: SUM_RIGID: [10 rigid terms];
// SUM_AREA_VOL_RIGID: [area, vol, 10 rigid terms].  (The motion update's
// sums have their own pass, k_face_sums_motion.)
// One block (blk of nblk) of stage 1.
__device__ void face_sums_block(const LagArgs& A, const double* __restrict__ xyz, int mode, double* __restrict__ part, int blk, int nblk) {
  __shared__ double sm[12][RB];
  const int ns = mode == SUM_RIGID ? 10 : 12;
  double acc[12];
  for (int q = 0; q < 12; ++q) acc[q] = 0.0;
  for (int f = blk * RB + threadIdx.x; f < A.nf; f += nblk * RB) {
    if (A.gf[f]) continue;
    if (mode == SUM_RIGID) {
      double o[10];
      face_rigid(xyz, A.vof, f, o);
      for (int q = 0; q < 10; ++q) acc[q] = acc[q] + o[q];
    } else {
      acc[0] = acc[0] + A.L[MELT_L_SUR][f];
      acc[1] = acc[1] + face_volume(xyz, A.vof, f);
      double o[10];
      face_rigid(xyz, A.vof, f, o);
      for (int q = 0; q < 10; ++q) acc[2 + q] = acc[2 + q] + o[q];
    }
  }
  for (int q = 0; q < ns; ++q) sm[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int w = RB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int q = 0; q < ns; ++q) sm[q][threadIdx.x] = sm[q][threadIdx.x] + sm[q][threadIdx.x + w];
    __syncthreads();
  }
  if ((int)threadIdx.x < ns) part[threadIdx.x * nblk + blk] = sm[threadIdx.x][0];
}
__global__ __launch_bounds__(RB) void k_face_sums(LagArgs A, const double* __restrict__ xyz, int mode, double* __restrict__ part) {
  face_sums_block(A, xyz, mode, part, blockIdx.x, gridDim.x);
}

__device__ void rigid_finish(const double* s, double dens_ratio, double com[3], double* vol, double I[9]);

// Body updates that consume a face sum (run by one thread after stage 2).
enum { FIN_NONE = 0, FIN_FINAL = 4 };

// The fixed-pattern sum of the nb <= 64 partials of quantity q into sm[q]
// (one wave per quantity, butterfly over the lanes), for q < ns; ends with a
// block barrier.
__device__ void reduce_partials(const double* __restrict__ part, int nb, int ns, double* sm) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = RB / 64;
  for (int q = wv; q < ns; q += nw) {
    double v = lane < nb ? part[q * nb + lane] : 0.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off, 64);
    if (lane == 0) sm[q] = v;
  }
  __syncthreads();
}

// mls_structLoads.f90:22-30, 224-256: the integrated loads of this substep
// from the four face sums s[0..11], the previous ones kept
__device__ void loads_finish(melt_body* b, const double* s) {
  for (int d = 0; d < 3; ++d) {
    b->int_prn_dA_m1[d] = b->int_prn_dA[d]; b->int_tau_dA_m1[d] = b->int_tau_dA[d];
    b->int_r_x_prn_dA_m1[d] = b->int_r_x_prn_dA[d]; b->int_r_x_tau_dA_m1[d] = b->int_r_x_tau_dA[d];
    b->int_prn_dA[d] = s[d]; b->int_r_x_prn_dA[d] = s[3 + d];
    b->int_tau_dA[d] = s[6 + d]; b->int_r_x_tau_dA[d] = s[9 + d];
    b->fp[d] = b->fp[d] + b->int_prn_dA[d]; b->ftau[d] = b->ftau[d] + b->int_tau_dA[d];
    b->torq_p[d] = b->torq_p[d] + b->int_r_x_prn_dA[d]; b->torq_tau[d] = b->torq_tau[d] + b->int_r_x_tau_dA[d];
  }
}

// Stage 2: out[q] = the sum of quantity q, then the body update `fin`.
__global__ __launch_bounds__(RB) void k_sum_partials(const double* __restrict__ part, int nb, int ns, double* __restrict__ out, int fin,
                                                     melt_body* body, double dens_ratio, int* err) {
  __shared__ double sm[16];
  reduce_partials(part, nb, ns, sm);
  if ((int)threadIdx.x < ns) out[threadIdx.x] = sm[threadIdx.x];
  if (threadIdx.x != 0 || fin != FIN_FINAL) return;
  melt_body* b = body;
  b->surface = sm[0];
  b->volume = sm[1] / 6.;
  rigid_finish(&sm[2], dens_ratio, b->pos_cm, &b->volume, b->inert_tensor);
  if (b->volume < b->v_thresh) raise_error(err, MELT_ERR_SMALL_VOLUME);
}

// Two-stage face sum of `mode` into out[] (d_red scratch holds the partials).
void face_sums(melt_ctx* ctx, const LagArgs& A, const double* xyz, int mode, double* out, int fin = FIN_NONE) {
  int G = std::max(1, std::min(64, (A.nf + RB - 1) / RB));
  int ns = mode == SUM_RIGID ? 10 : 12;
  double* part = ctx->d_red + 8192;
  hipLaunchKernelGGL(k_face_sums, dim3(G), dim3(RB), 0, ctx->stream, A, xyz, mode, part);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(RB), 0, ctx->stream, part, G, ns, out, fin, ctx->d_body, ctx->cfg.dens_ratio, ctx->d_err);
}

// calc_rigidBody_params epilogue (RigidAuxRoutines.f90:716-740) from the 10 sums
__device__ void rigid_finish(const double* s, double dens_ratio, double com[3], double* vol, double I[9]) {
  double V = s[0], Cx = s[1], Cy = s[2], Cz = s[3];
  double Ixx = s[4], Iyy = s[5], Izz = s[6], Ixy = s[7], Ixz = s[8], Iyz = s[9];
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
  I[0] = Ixx * dens_ratio; I[3] = -Ixy * dens_ratio; I[6] = -Ixz * dens_ratio;
  I[1] = -Ixy * dens_ratio; I[4] = Iyy * dens_ratio; I[7] = -Iyz * dens_ratio;
  I[2] = -Ixz * dens_ratio; I[5] = -Iyz * dens_ratio; I[8] = Izz * dens_ratio;
}

__global__ void k_body_rigid(melt_body* b, const double* __restrict__ s, double dens_ratio) {
  rigid_finish(s, dens_ratio, b->pos_cm, &b->volume, b->inert_tensor);
}



// geometry refresh epilogue: Surface, Volume (calculate_volume, then the
// rigid-body volume overwrites it as in particle.f90:170-182), COM, inertia,
// small-volume abort (particle.f90:218-224)

// ---------------------------------------------------------- heat / Stefan
// With pind2, also commits the heat-flux probes' cell indices that the fused
// substep staged there (the spreading had still read the old ones).
__global__ __launch_bounds__(256) void k_face_post(LagArgs A, const double* __restrict__ probe, const melt_body* __restrict__ b,
                                                   double* __restrict__ rxp, double* __restrict__ rxt, double ren, int heat, int loads,
                                                   int* __restrict__ pind, const int* __restrict__ pind2) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= A.nf || A.gf[f]) return;
  const double* pr = &probe[(size_t)PRB_N * f];
  if (pind2)
    for (int q = 0; q < 6; ++q) pind[6 * f + q] = pind2[6 * f + q];
  if (heat) { A.L[MELT_L_QW_O][f] = pr[PRB_QO]; A.L[MELT_L_QW_I][f] = pr[PRB_QI]; }
  if (!loads) return;
  const double* nor = &A.L[MELT_L_TRI_NOR][3 * f];
  const double* bar = &A.L[MELT_L_TRI_BAR][3 * f];
  double sur = A.L[MELT_L_SUR][f], h = b->h_eulerian;
  double rad[3] = {bar[0] - b->pos_cm[0], bar[1] - b->pos_cm[1], bar[2] - b->pos_cm[2]};
  double ptri = pr[PRB_P] - h * (pr[PRB_GP] * nor[0] + pr[PRB_GP + 1] * nor[1] + pr[PRB_GP + 2] * nor[2]);
  for (int d = 0; d < 3; ++d) A.L[MELT_L_PRESS_N_TRI][3 * f + d] = ptri * sur * nor[d];
  double pv[3] = {ptri * nor[0], ptri * nor[1], ptri * nor[2]}, rx[3];
  dcross(rx, rad, pv);
  for (int d = 0; d < 3; ++d) rxp[3 * f + d] = rx[d] * sur;
  double S11 = pr[PRB_GU], S12 = pr[PRB_GU + 1], S13 = pr[PRB_GU + 2];
  double S21 = pr[PRB_GV], S22 = pr[PRB_GV + 1], S23 = pr[PRB_GV + 2];
  double S31 = pr[PRB_GW], S32 = pr[PRB_GW + 1], S33 = pr[PRB_GW + 2];
  double t[3], tw[3], ra[3], rb[3];
  t[0] = (sur / ren) * (2.0 * S11 * nor[0] + (S12 + S21) * nor[1] + S13 * nor[2]);
  t[1] = (sur / ren) * ((S21 + S12) * nor[0] + 2.0 * S22 * nor[1] + S23 * nor[2]);
  t[2] = (sur / ren) * (S13 * nor[0] + S23 * nor[1]);
  tw[0] = (sur / ren) * (S31 * nor[2]);
  tw[1] = (sur / ren) * (S32 * nor[2]);
  tw[2] = (sur / ren) * (S31 * nor[0] + S32 * nor[1] + 2.0 * S33 * nor[2]);
  dcross(ra, rad, t);
  dcross(rb, rad, tw);
  for (int d = 0; d < 3; ++d) {
    A.L[MELT_L_TAU_N][3 * f + d] = (0.0 + t[d]) + tw[d];
    rxt[3 * f + d] = (0.0 + ra[d]) + rb[d];
  }
}

__global__ __launch_bounds__(256) void k_vert_heat_stefan(LagArgs A, const melt_body* __restrict__ b, double cp_over_L, double dt, double ga,
                                                          double ro, int stefan) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= A.nv) return;
  double qo = 0.0, qi = 0.0, av = A.L[MELT_L_AVERT][v];
  for (int p = A.v2f_ptr[v]; p < A.v2f_ptr[v + 1]; ++p) {
    int f = A.v2f[p];
    double w = A.L[MELT_L_SUR][f] / 3.0 / av;
    qo = qo + A.L[MELT_L_QW_O][f] * w;
    qi = qi + A.L[MELT_L_QW_I][f] * w;
  }
  A.L[MELT_L_QW_OVERT][v] = qo;
  A.L[MELT_L_QW_IVERT][v] = qi;
  if (!stefan) return;
  double* vm = &A.L[MELT_L_VMELT][3 * v];
  double* vm1 = &A.L[MELT_L_VMELT_M1][3 * v];
  for (int d = 0; d < 3; ++d) vm1[d] = vm[d];
  if (A.gv[v]) return;
  const double* n = &A.L[MELT_L_VERT_NOR][3 * v];
  for (int d = 0; d < 3; ++d) vm[d] = cp_over_L * (qi - qo) * n[d];
  double* x = &A.L[MELT_L_XYZV][3 * v];
  for (int d = 0; d < 3; ++d) x[d] = x[d] + dt * (ga * vm[d] + ro * vm1[d]);
}

// ---------------------------------------------------------- geometry
// calc_centroids_from_vert (triaux.f90:308-323) of face f from the vertices
__device__ inline void face_centroid(const LagArgs& A, int f) {
  int v1 = A.vof[3 * f], v2 = A.vof[3 * f + 1], v3 = A.vof[3 * f + 2];
  for (int d = 0; d < 3; ++d) A.L[MELT_L_TRI_BAR][3 * f + d] = (A.L[MELT_L_XYZV][3 * v1 + d] + A.L[MELT_L_XYZV][3 * v2 + d] + A.L[MELT_L_XYZV][3 * v3 + d]) / 3.0;
}
// update_tri_normal / calculate_normal of face f from the points xyz
__device__ inline void face_normal(const LagArgs& A, const double* __restrict__ xyz, int f) {
  int v1 = A.vof[3 * f], v2 = A.vof[3 * f + 1], v3 = A.vof[3 * f + 2];
  double e1[3], e2[3], n[3];
  for (int d = 0; d < 3; ++d) { e1[d] = xyz[3 * v2 + d] - xyz[3 * v1 + d]; e2[d] = xyz[3 * v3 + d] - xyz[3 * v1 + d]; }
  n[0] = e1[1] * e2[2] - e1[2] * e2[1];
  n[1] = e1[2] * e2[0] - e1[0] * e2[2];
  n[2] = e1[0] * e2[1] - e1[1] * e2[0];
  double s = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  for (int d = 0; d < 3; ++d) A.L[MELT_L_TRI_NOR][3 * f + d] = n[d] / s;
}

__global__ __launch_bounds__(256) void k_face_centroid_normal(LagArgs A, const double* __restrict__ xyz, int do_bar, int do_nor, int ghost_check) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= A.nf) return;
  bool ghost = A.gf[f] != 0;
  if (do_bar && !ghost) face_centroid(A, f);
  if (do_nor && !(ghost_check && ghost)) face_normal(A, xyz, f);
}

__global__ __launch_bounds__(256) void k_face_area(LagArgs A, const double* __restrict__ x) {  // calculate_area triaux.f90:82-119
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= A.nf || A.gf[f]) return;
  int v[3] = {A.vof[3 * f], A.vof[3 * f + 1], A.vof[3 * f + 2]};
  double dd[3];
  for (int e = 0; e < 3; ++e) {
    int a = v[e], b = v[(e + 1) % 3];
    double q0 = x[3 * a] - x[3 * b], q1 = x[3 * a + 1] - x[3 * b + 1], q2 = x[3 * a + 2] - x[3 * b + 2];
    dd[e] = sqrt(q0 * q0 + q1 * q1 + q2 * q2);
  }
  double sp = (dd[0] + dd[1] + dd[2]) / 2.;
  A.L[MELT_L_SUR][f] = sqrt(sp * (sp - dd[0]) * (sp - dd[1]) * (sp - dd[2]));
}

__device__ inline double edge_length(const LagArgs& A, const double* __restrict__ x, int e) {  // calculate_eLengths triaux.f90:277-305
  int v1 = A.voe[2 * e], v2 = A.voe[2 * e + 1];
  double a = x[3 * v1] - x[3 * v2], c = x[3 * v1 + 1] - x[3 * v2 + 1], d = x[3 * v1 + 2] - x[3 * v2 + 2];
  return sqrt(a * a + c * c + d * d);
}

__global__ __launch_bounds__(256) void k_edge_len(LagArgs A, const double* __restrict__ x, melt_body* b, double e_thresh) {  // calculate_eLengths triaux.f90:277-305
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  bool short_edge = false;
  if (e < A.ne && !A.ge[e]) {
    double l = edge_length(A, x, e);
    A.L[MELT_L_ELENGTHS][e] = l;
    short_edge = l < e_thresh;
  }
  // the remesh flag is a logical OR: one atomic per wave that saw a short edge
  if (__ballot(short_edge) && (threadIdx.x & 63) == 0) atomicExch(&b->rm_flag, 1);
}

__global__ void k_clear_rm(melt_body* b) { b->rm_flag = 0; }

// calculate_area (triaux.f90:82-119), calc_centroids_from_vert (:308-323) and
// update_tri_normal of every live face in one pass (the same expressions as
// k_face_area and k_face_centroid_normal); thread 0 also clears the remesh
// flag that k_edge_len then sets.
__global__ __launch_bounds__(256) void k_face_geom(LagArgs A, const double* __restrict__ x, melt_body* b) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f == 0) b->rm_flag = 0;
  if (f >= A.nf || A.gf[f]) return;
  int v[3] = {A.vof[3 * f], A.vof[3 * f + 1], A.vof[3 * f + 2]};
  double dd[3];
  for (int e = 0; e < 3; ++e) {
    int a = v[e], c = v[(e + 1) % 3];
    double q0 = x[3 * a] - x[3 * c], q1 = x[3 * a + 1] - x[3 * c + 1], q2 = x[3 * a + 2] - x[3 * c + 2];
    dd[e] = sqrt(q0 * q0 + q1 * q1 + q2 * q2);
  }
  double sp = (dd[0] + dd[1] + dd[2]) / 2.;
  A.L[MELT_L_SUR][f] = sqrt(sp * (sp - dd[0]) * (sp - dd[1]) * (sp - dd[2]));
  const int v1 = v[0], v2 = v[1], v3 = v[2];
  for (int d = 0; d < 3; ++d) A.L[MELT_L_TRI_BAR][3 * f + d] = (A.L[MELT_L_XYZV][3 * v1 + d] + A.L[MELT_L_XYZV][3 * v2 + d] + A.L[MELT_L_XYZV][3 * v3 + d]) / 3.0;
  double e1[3], e2[3], n[3];
  for (int d = 0; d < 3; ++d) { e1[d] = x[3 * v2 + d] - x[3 * v1 + d]; e2[d] = x[3 * v3 + d] - x[3 * v1 + d]; }
  n[0] = e1[1] * e2[2] - e1[2] * e2[1];
  n[1] = e1[2] * e2[0] - e1[0] * e2[2];
  n[2] = e1[0] * e2[1] - e1[1] * e2[0];
  double s = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  for (int d = 0; d < 3; ++d) A.L[MELT_L_TRI_NOR][3 * f + d] = n[d] / s;
}

__global__ __launch_bounds__(256) void k_skewness(LagArgs A) {  // calculate_skewness triaux.f90:325-358
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= A.nf || A.gf[f]) return;
  const double* el = A.L[MELT_L_ELENGTHS];
  double perim = el[A.eof[3 * f]] + el[A.eof[3 * f + 1]] + el[A.eof[3 * f + 2]];
  double q = perim / 3.0;
  double sopt = sqrt(3.0) / 4.0 * (q * q);
  A.L[MELT_L_SKEWNESS][f] = (sopt - A.L[MELT_L_SUR][f]) / sopt;
}

// vertex areas (triaux.f90:121-150) and area-weighted normals (:237-274)
__global__ __launch_bounds__(256) void k_vert_geom(LagArgs A) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= A.nv) return;
  double av = 0.0, n[3] = {0.0, 0.0, 0.0};
  for (int p = A.v2f_ptr[v]; p < A.v2f_ptr[v + 1]; ++p) {
    int f = A.v2f[p];
    double s = A.L[MELT_L_SUR][f];
    av = av + s / 3.0;
    for (int d = 0; d < 3; ++d) n[d] = n[d] + s * A.L[MELT_L_TRI_NOR][3 * f + d];
  }
  A.L[MELT_L_AVERT][v] = av;
  if (!A.gv[v]) {
    double m = fnorm2(n, 3);
    for (int d = 0; d < 3; ++d) n[d] = n[d] / m;
  }
  for (int d = 0; d < 3; ++d) A.L[MELT_L_VERT_NOR][3 * v + d] = n[d];
}

__global__ __launch_bounds__(256) void k_cfac(LagArgs A, const melt_body* __restrict__ b) {  // particle.f90:189-193, initmls.f90:25
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= A.nf) return;
  A.L[MELT_L_CFAC][f] = (A.L[MELT_L_SUR][f] * b->h_eulerian) / b->celvol;
}

// skewness of face f from its edge lengths el[0..2]
__device__ inline void face_skewness(const LagArgs& A, int f, const double el[3]) {
  double perim = el[0] + el[1] + el[2];
  double q = perim / 3.0;
  double sopt = sqrt(3.0) / 4.0 * (q * q);
  A.L[MELT_L_SKEWNESS][f] = (sopt - A.L[MELT_L_SUR][f]) / sopt;
}
// area and normal of vertex v from its faces
__device__ void vert_post_geom(const LagArgs& A, int v) {
  double av = 0.0, n[3] = {0.0, 0.0, 0.0};
  for (int p = A.v2f_ptr[v]; p < A.v2f_ptr[v + 1]; ++p) {
    int f = A.v2f[p];
    double s = A.L[MELT_L_SUR][f];
    av = av + s / 3.0;
    for (int d = 0; d < 3; ++d) n[d] = n[d] + s * A.L[MELT_L_TRI_NOR][3 * f + d];
  }
  A.L[MELT_L_AVERT][v] = av;
  if (!A.gv[v]) {
    double m = fnorm2(n, 3);
    for (int d = 0; d < 3; ++d) n[d] = n[d] / m;
  }
  for (int d = 0; d < 3; ++d) A.L[MELT_L_VERT_NOR][3 * v + d] = n[d];
}

// The geometry refresh after k_face_geom in one launch: blocks [0, G) are
// stage 1 of the refresh's face sums (surface, volume, rigid-body terms;
// the same partition as k_face_sums with G blocks, so the same partials);
// the others take one edge, face or vertex per thread: the edge lengths and
// the remesh flag as k_edge_len, cfac and skewness of the faces
// (calculate_skewness triaux.f90:325-358), the vertex areas and normals.  A face measures its own edges with k_edge_len's
// expression (the same values) instead of reading them back, as the edges
// are measured in the same launch.
__global__ __launch_bounds__(RB) void k_refresh_tail(LagArgs A, const double* __restrict__ x, melt_body* b, double e_thresh, int do_cfac, int G,
                                                     double* __restrict__ part) {
  if ((int)blockIdx.x < G) {
    face_sums_block(A, x, SUM_AREA_VOL_RIGID, part, blockIdx.x, G);
    return;
  }
  int t = ((int)blockIdx.x - G) * RB + threadIdx.x;
  bool short_edge = false;
  if (t < A.ne && !A.ge[t]) {
    double l = edge_length(A, x, t);
    A.L[MELT_L_ELENGTHS][t] = l;
    short_edge = l < e_thresh;
  }
  // the remesh flag is a logical OR: one atomic per wave that saw a short edge
  if (__ballot(short_edge) && (threadIdx.x & 63) == 0) atomicExch(&b->rm_flag, 1);
  t -= A.ne;
  if (t < 0) return;
  if (t < A.nf) {
    const int f = t;
    if (do_cfac) A.L[MELT_L_CFAC][f] = (A.L[MELT_L_SUR][f] * b->h_eulerian) / b->celvol;
    if (A.gf[f]) return;
    double e3[3];
    for (int k = 0; k < 3; ++k) {
      const int e = A.eof[3 * f + k];
      e3[k] = A.ge[e] ? A.L[MELT_L_ELENGTHS][e] : edge_length(A, x, e);
    }
    face_skewness(A, f, e3);
    return;
  }
  const int v = t - A.nf;
  if (v < A.nv) vert_post_geom(A, v);
}

// ---------------------------------------------------------- rigid body

__device__ void rot_matrix(double q[4], double AA[9]) {  // calc_rot_matrix RigidAuxRoutines.f90:620-651 (column-major)
  double n = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  for (int i = 0; i < 4; ++i) q[i] = q[i] / n;
  AA[0] = q[0] * q[0] + q[1] * q[1] - q[2] * q[2] - q[3] * q[3];
  AA[3] = 2.0 * (q[1] * q[2] + q[0] * q[3]);
  AA[6] = 2.0 * (q[1] * q[3] - q[0] * q[2]);
  AA[1] = 2.0 * (q[1] * q[2] - q[0] * q[3]);
  AA[4] = q[0] * q[0] - q[1] * q[1] + q[2] * q[2] - q[3] * q[3];
  AA[7] = 2.0 * (q[2] * q[3] + q[0] * q[1]);
  AA[2] = 2.0 * (q[1] * q[3] + q[0] * q[2]);
  AA[5] = 2.0 * (q[2] * q[3] - q[0] * q[1]);
  AA[8] = q[0] * q[0] - q[1] * q[1] - q[2] * q[2] + q[3] * q[3];
}

__device__ void mm3(const double* A, const double* B, double* C) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s = s + A[i + 3 * k] * B[k + 3 * j];
      C[i + 3 * j] = s;
    }
}
__device__ void mv3(const double* A, const double* x, double* y) {
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    for (int k = 0; k < 3; ++k) s = s + A[i + 3 * k] * x[k];
    y[i] = s;
  }
}

// newton_euler (RigidAuxRoutines.f90:140-310), the periodic COM wrap of
// update_part_pos (:44-51) and the rotation matrix update_xyz uses (:855).
__device__ void newton_update(melt_body* b, double ga, double ro, double al, double dt, double dens_ratio, double xlen, double ylen,
                              double zlen, double* __restrict__ aa_out, int* err) {
  double vel_m1[3], pos_m1[3], om_m1[3];
  for (int d = 0; d < 3; ++d) { vel_m1[d] = b->vel_cm[d]; pos_m1[d] = b->pos_cm[d]; om_m1[d] = b->omega_c[d]; }
  double e_z[3] = {0., 0., -1.0};
  double pre_fac = 1.0 / b->volume / dens_ratio;
  for (int d = 0; d < 3; ++d)
    b->vel_cm[d] = vel_m1[d] + ga * dt * pre_fac * (-b->int_prn_dA[d] + b->int_tau_dA[d]) +
                   ro * dt * pre_fac * (-b->int_prn_dA_m1[d] + b->int_tau_dA_m1[d]) + al * dt * (1.0 - (1.0 / dens_ratio)) * e_z[d];
  for (int d = 0; d < 3; ++d) b->pos_cm[d] = pos_m1[d] + 0.5 * al * dt * (b->vel_cm[d] + vel_m1[d]);
  double I_ij[9], I_LHS[9], I_inv[9], I_prev[9], AA[9], AT[9], buf[9], tmp[3], rhs[3];
  for (int i = 0; i < 9; ++i) { I_ij[i] = b->inert_tensor[i]; I_LHS[i] = I_ij[i]; }
  double res = 1.0;
  int n_iter = 1;
  while (n_iter <= 10000 && res > 1.e-15) {
    for (int i = 0; i < 9; ++i) I_inv[i] = I_LHS[i];
    if (lu_inverse<3>(I_inv)) { raise_error(err, MELT_ERR_ARG); return; }
    mv3(I_ij, om_m1, tmp);
    for (int d = 0; d < 3; ++d)
      rhs[d] = tmp[d] + ga * dt * (-b->int_r_x_prn_dA[d] + b->int_r_x_tau_dA[d]) + ro * dt * (-b->int_r_x_prn_dA_m1[d] + b->int_r_x_tau_dA_m1[d]);
    mv3(I_inv, rhs, b->omega_c);
    double ob[3];
    for (int d = 0; d < 3; ++d) ob[d] = 0.5 * (b->omega_c[d] + om_m1[d]);
    double mag = fnorm2(ob, 3);
    double ang = mag * al * dt;
    if (mag < 1.e-15) { ob[0] = ob[1] = ob[2] = 0.0; }
    else for (int d = 0; d < 3; ++d) ob[d] = ob[d] / mag;
    b->quat[0] = cos(0.5 * ang);
    b->quat[1] = sin(0.5 * ang) * ob[0];
    b->quat[2] = sin(0.5 * ang) * ob[1];
    b->quat[3] = sin(0.5 * ang) * ob[2];
    double qn = fnorm2(b->quat, 4);
    for (int d = 0; d < 4; ++d) b->quat[d] = b->quat[d] / qn;
    rot_matrix(b->quat, AA);
    for (int i = 0; i < 9; ++i) I_prev[i] = I_LHS[i];
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) AT[i + 3 * j] = AA[j + 3 * i];
    mm3(I_ij, AT, buf);
    mm3(AA, buf, I_LHS);
    res = 0.0;
    for (int i = 0; i < 9; ++i) res = fmax(res, fabs(I_LHS[i] - I_prev[i]));
    n_iter++;
  }
  for (int i = 0; i < 9; ++i) b->inert_tensor[i] = I_LHS[i];
  double L[3] = {xlen, ylen, zlen};
  for (int d = 0; d < 3; ++d) {
    if (b->pos_cm[d] < 0.) b->pos_cm[d] = b->pos_cm[d] + L[d];
    if (b->pos_cm[d] > L[d]) b->pos_cm[d] = b->pos_cm[d] - L[d];
  }
  rot_matrix(b->quat, AA);
  for (int i = 0; i < 9; ++i) aa_out[i] = AA[i];
}

// The face sums of the motion update in one pass over the faces: the loads
// (mls_structLoads.f90:224-256; L), the volume after melting
// (particle.f90:86-90; V) and the rigid-body terms of calc_rigidBody_params
// (RigidAuxRoutines.f90:685-712; with L, as the motion needs them), and with
// L the centroids and normals of the live faces from the melted vertices
// (calc_centroids_from_vert, update_tri_normal).  The loads read only the
// probe results, which the Stefan update before this pass does not touch.
// Every quantity is accumulated over the same faces per thread and reduced
// by the same tree as in its own k_face_sums pass, so each sum is
// bit-identical to it: one launch where there were three, and one finishing
// launch (k_sum_motion) instead of three.  Partials: [loads 12][volume 1]
// [rigid 10], those present, quantity-major as in k_face_sums.
template <bool L, bool V>
__global__ __launch_bounds__(RB) void k_face_sums_motion(LagArgs A, const double* __restrict__ xyz, const double* __restrict__ rxp,
                                                         const double* __restrict__ rxt, double* __restrict__ part) {
  constexpr int QV = L ? 12 : 0, QG = QV + (V ? 1 : 0), NS = QG + (L ? 10 : 0);
  __shared__ double sm[NS][RB];
  double acc[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) acc[q] = 0.0;
  for (int f = blockIdx.x * RB + threadIdx.x; f < A.nf; f += gridDim.x * RB) {
    if (A.gf[f]) continue;
    if constexpr (L) {
      for (int d = 0; d < 3; ++d) {
        acc[d] = acc[d] + A.L[MELT_L_PRESS_N_TRI][3 * f + d];
        acc[3 + d] = acc[3 + d] + rxp[3 * f + d];
        acc[6 + d] = acc[6 + d] + A.L[MELT_L_TAU_N][3 * f + d];
        acc[9 + d] = acc[9 + d] + rxt[3 * f + d];
      }
    }
    if constexpr (V) acc[QV] = acc[QV] + face_volume(xyz, A.vof, f);
    if constexpr (L) {
      double o[10];
      face_rigid(xyz, A.vof, f, o);
#pragma unroll
      for (int q = 0; q < 10; ++q) acc[QG + q] = acc[QG + q] + o[q];
      face_centroid(A, f);
      face_normal(A, xyz, f);
    }
  }
#pragma unroll
  for (int q = 0; q < NS; ++q) sm[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int w = RB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int q = 0; q < NS; ++q) sm[q][threadIdx.x] = sm[q][threadIdx.x] + sm[q][threadIdx.x + w];
    __syncthreads();
  }
  if ((int)threadIdx.x < NS) part[threadIdx.x * gridDim.x + blockIdx.x] = sm[threadIdx.x][0];
}

// Stage 2 of k_face_sums_motion and the body updates that consume it, in the
// reference's order on one thread: the loads, the volume after melting, the
// rigid-body terms (centre of mass, inertia) and, with L, the Newton-Euler
// step (its rotation into aa_out).  com_out keeps the centre of mass the step
// started from, which the relative positions are taken against.
template <bool L, bool V>
__global__ __launch_bounds__(RB) void k_sum_motion(const double* __restrict__ part, int nb, double* __restrict__ out, melt_body* body,
                                                   double dens_ratio, double ga, double ro, double al, double dt, double xlen, double ylen,
                                                   double zlen, double* __restrict__ aa_out, double* __restrict__ com_out, int* err) {
  constexpr int QV = L ? 12 : 0, QG = QV + (V ? 1 : 0), NS = QG + (L ? 10 : 0);
  __shared__ double sm[NS];
  reduce_partials(part, nb, NS, sm);
  if ((int)threadIdx.x < NS) out[threadIdx.x] = sm[threadIdx.x];
  if (threadIdx.x != 0) return;
  melt_body* b = body;
  if constexpr (L) loads_finish(b, sm);
  if constexpr (V) b->volume = sm[QV] / 6.;
  if constexpr (L) {
    rigid_finish(&sm[QG], dens_ratio, b->pos_cm, &b->volume, b->inert_tensor);
    for (int d = 0; d < 3; ++d) com_out[d] = b->pos_cm[d];
    newton_update(b, ga, ro, al, dt, dens_ratio, xlen, ylen, zlen, aa_out, err);
  }
}

// Relative vertex and face positions against the centre of mass com before
// the Newton-Euler step (RigidAuxRoutines.f90:20-35), and the vertices moved
// by its rotation AA and new centre (update_xyz, :858-862), in one pass.
__global__ __launch_bounds__(256) void k_rel_update(LagArgs A, const melt_body* __restrict__ b, const double* __restrict__ com,
                                                    const double* __restrict__ AA) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < A.nv) {
    if (A.gv[t]) return;
    double* r = &A.L[MELT_L_DXYZV_S][3 * t];
    double* x = &A.L[MELT_L_XYZV][3 * t];
    double rel[3], AT[9], u[3];
    for (int d = 0; d < 3; ++d) rel[d] = x[d] - com[d];
    for (int d = 0; d < 3; ++d) r[d] = rel[d];
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) AT[i + 3 * j] = AA[j + 3 * i];
    mv3(AT, rel, u);
    for (int d = 0; d < 3; ++d) x[d] = u[d] + b->pos_cm[d];
  } else if (t < A.nv + A.nf) {
    int f = t - A.nv;
    if (!A.gf[f])
      for (int d = 0; d < 3; ++d) A.L[MELT_L_DXYZ_S][3 * f + d] = A.L[MELT_L_TRI_BAR][3 * f + d] - com[d];
  }
}

__global__ __launch_bounds__(256) void k_face_vel(LagArgs A, const melt_body* __restrict__ b) {  // update_xyz :864-872
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= A.nf || A.gf[f]) return;
  int v1 = A.vof[3 * f], v2 = A.vof[3 * f + 1], v3 = A.vof[3 * f + 2];
  for (int d = 0; d < 3; ++d) A.L[MELT_L_TRI_BAR][3 * f + d] = (A.L[MELT_L_XYZV][3 * v1 + d] + A.L[MELT_L_XYZV][3 * v2 + d] + A.L[MELT_L_XYZV][3 * v3 + d]) / 3.0;
  double om[3];
  dcross(om, b->omega_c, &A.L[MELT_L_DXYZ_S][3 * f]);
  for (int d = 0; d < 3; ++d) A.L[MELT_L_VEL_TRI][3 * f + d] = b->vel_cm[d] + om[d];
}

// setup_particles (create_geo.f90:46-104) after calc_rigidBody_params, and the
// init_mlsForce constants (initmls.f90:15-24, evaluated on the host)
__global__ void k_geom_init_body(melt_body* b, double xlen, double ylen, double zlen, double h_eul, double celvol, double e_thresh,
                                 double v_thresh) {
  b->pos_cm[0] = 0.5 * xlen; b->pos_cm[1] = 0.5 * ylen; b->pos_cm[2] = 0.5 * zlen;
  for (int d = 0; d < 3; ++d) { b->omega_c[d] = 0.0; b->vel_cm[d] = 0.0; }
  for (int d = 0; d < 4; ++d) b->quat[d] = 0.0;
  b->h_eulerian = h_eul;
  b->celvol = celvol;
  b->e_thresh = e_thresh;
  b->v_thresh = v_thresh;
  b->initial_tag = 1;
}

__global__ __launch_bounds__(256) void k_translate(LagArgs A, const double* __restrict__ xyz0, const melt_body* __restrict__ b) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= A.nv) return;
  for (int d = 0; d < 3; ++d) A.L[MELT_L_XYZV][3 * v + d] = xyz0[3 * v + d] + b->pos_cm[d];
}

__global__ void k_surface(melt_body* b, const double* __restrict__ s) { b->surface = s[0]; }

inline unsigned nb(size_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

LagArgs lag_args(melt_ctx* ctx) {
  LagArgs A;
  for (int i = 0; i < MELT_L_NLAG; ++i) A.L[i] = ctx->d_lag[i];
  A.vof = ctx->d_vof; A.eof = ctx->d_eof; A.voe = ctx->d_voe; A.v2f_ptr = ctx->d_v2f_ptr; A.v2f = ctx->d_v2f;
  A.gv = ctx->d_gv; A.ge = ctx->d_ge; A.gf = ctx->d_gf;
  A.nv = ctx->nv; A.ne = ctx->ne; A.nf = ctx->nf;
  return A;
}

// geometry refresh of particle.f90:163-193
int refresh_geometry(melt_ctx* ctx, LagArgs& A) {
  const double* x = ctx->d_lag[MELT_L_XYZV];
  hipStream_t s = ctx->stream;
  hipLaunchKernelGGL(k_face_geom, dim3(nb(ctx->nf, 256)), dim3(256), 0, s, A, x, ctx->d_body);
  // work that needs only the refreshed faces (the fused substep's next
  // tagging) forks here; the rest of the refresh does not touch what it reads
  if (const std::function<int()>* hook = ctx->after_faces) {
    ctx->after_faces = nullptr;
    if (int rc = (*hook)()) return rc;
  }
  // edge lengths, face and vertex post-processing and the face sums: one
  // launch, then the sums' finish (surface, volume, COM, inertia, abort)
  const int G = std::max(1, std::min(64, (A.nf + RB - 1) / RB));
  double* part = ctx->d_red + 8192;
  hipLaunchKernelGGL(k_refresh_tail, dim3(G + nb((size_t)ctx->ne + ctx->nf + ctx->nv, RB)), dim3(RB), 0, s, A, x, ctx->d_body,
                     ctx->h_body.e_thresh, ctx->cfg.imelt == 1 ? 1 : 0, G, part);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(RB), 0, s, part, G, 12, ctx->d_red, (int)FIN_FINAL, ctx->d_body, ctx->cfg.dens_ratio,
                     ctx->d_err);
  MELT_HIP(hipGetLastError());
  return 0;
}

}  // namespace

// update_tri_velocity (RigidAuxRoutines.f90:902-928): after remeshing, the
// face velocities from the refreshed centroids and centre of mass
__global__ __launch_bounds__(256) void k_tri_velocity(LagArgs A, const melt_body* __restrict__ b) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= A.nf || A.gf[f]) return;
  double rad[3], om[3];
  for (int d = 0; d < 3; ++d) rad[d] = A.L[MELT_L_TRI_BAR][3 * f + d] - b->pos_cm[d];
  om[0] = b->omega_c[1] * rad[2] - rad[1] * b->omega_c[2];
  om[1] = b->omega_c[2] * rad[0] - rad[2] * b->omega_c[0];
  om[2] = b->omega_c[0] * rad[1] - rad[0] * b->omega_c[1];
  for (int d = 0; d < 3; ++d) A.L[MELT_L_VEL_TRI][3 * f + d] = b->vel_cm[d] + om[d];
}

// the rest of particle.f90 after the edges were measured (IREMESH = 1):
// remesh on the host when an edge is short, then the geometry refresh
static int lagr_finish(melt_ctx* ctx) {
  if (ctx->cfg.iremesh != 1) return 0;
  LagArgs A = lag_args(ctx);
  MELT_HIP(hipEventSynchronize(ctx->ev_rm));
  int did = 0;
  if (*ctx->h_rm) {
    if (int rc = remesh_context(ctx, &did)) return rc;
    A = lag_args(ctx);
  }
  if (int rc = refresh_geometry(ctx, A)) return rc;
  if (did) {
    hipLaunchKernelGGL(k_tri_velocity, dim3(nb(ctx->nf, 256)), dim3(256), 0, ctx->stream, A, ctx->d_body);
    MELT_HIP(hipGetLastError());
  }
  return 0;
}

// k_face_sums_motion and k_sum_motion: the sums into d_red[0..22], the
// Newton-Euler rotation into d_red[32..40], the centre of mass it started
// from into d_red[48..50].
template <bool L, bool V>
static void motion_sums(melt_ctx* ctx, const LagArgs& A) {
  const int G = std::max(1, std::min(64, (A.nf + RB - 1) / RB));
  double* part = ctx->d_red + 8192;
  hipLaunchKernelGGL((k_face_sums_motion<L, V>), dim3(G), dim3(RB), 0, ctx->stream, A, ctx->d_lag[MELT_L_XYZV], ctx->d_rxp, ctx->d_rxt,
                     part);
  hipLaunchKernelGGL((k_sum_motion<L, V>), dim3(1), dim3(RB), 0, ctx->stream, part, G, ctx->d_red, ctx->d_body, ctx->cfg.dens_ratio, ctx->ga,
                     ctx->ro, ctx->al, ctx->dt, ctx->cfg.xlen, ctx->cfg.ylen, ctx->cfg.zlen, ctx->d_red + 32, ctx->d_red + 48, ctx->d_err);
}

// particle.f90:62-193 after the MLS interpolation: heat flux -> vertices,
// loads, Stefan update, Newton-Euler motion, (remeshing,) geometry refresh.
int lagr_after_force(melt_ctx* ctx, bool commit_pind) {
  LagArgs A = lag_args(ctx);
  hipStream_t s = ctx->stream;
  int heat = ctx->cfg.imelt == 1, loads = ctx->cfg.imlsstr == 1;
  if (heat || loads)
    hipLaunchKernelGGL(k_face_post, dim3(nb(ctx->nf, 256)), dim3(256), 0, s, A, ctx->d_probe, ctx->d_body, ctx->d_rxp, ctx->d_rxt, ctx->cfg.ren,
                       heat, loads, ctx->d_pind, commit_pind && heat ? ctx->d_pind2 : nullptr);
  if (heat) {
    hipLaunchKernelGGL(k_vert_heat_stefan, dim3(nb(ctx->nv, 256)), dim3(256), 0, s, A, ctx->d_body, ctx->cfg.cpliquid / ctx->cfg.latheat, ctx->dt,
                       ctx->ga, ctx->ro, 1);
    // the volume after melting is the one the remeshing restores (particle.f90:86-90)
    if (ctx->cfg.iremesh == 1)
      MELT_HIP(hipMemcpyAsync(ctx->d_xyz_melt, ctx->d_lag[MELT_L_XYZV], sizeof(double) * 3 * ctx->nv, hipMemcpyDeviceToDevice, s));
  }
  // the loads (the reference sums them before the Stefan update, which does
  // not touch what they read), the volume after melting, the rigid-body terms
  // and the Newton-Euler step: one face pass and one finishing launch
  if (loads && heat) motion_sums<true, true>(ctx, A);
  else if (loads) motion_sums<true, false>(ctx, A);
  else if (heat) motion_sums<false, true>(ctx, A);
  if (loads) {
    hipLaunchKernelGGL(k_rel_update, dim3(nb((size_t)ctx->nv + ctx->nf, 256)), dim3(256), 0, s, A, ctx->d_body, ctx->d_red + 48, ctx->d_red + 32);
    hipLaunchKernelGGL(k_face_vel, dim3(nb(ctx->nf, 256)), dim3(256), 0, s, A, ctx->d_body);
  }
  MELT_HIP(hipGetLastError());
  if (ctx->cfg.iremesh != 1) return refresh_geometry(ctx, A);
  // calculate_eLengths (particle.f90:128): the short-edge flag goes to the host
  hipLaunchKernelGGL(k_clear_rm, dim3(1), dim3(1), 0, s, ctx->d_body);
  hipLaunchKernelGGL(k_edge_len, dim3(nb(ctx->ne, 256)), dim3(256), 0, s, A, ctx->d_lag[MELT_L_XYZV], ctx->d_body, ctx->h_body.e_thresh);
  MELT_HIP(hipGetLastError());
  MELT_HIP(hipMemcpyAsync(ctx->h_rm, &ctx->d_body->rm_flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  MELT_HIP(hipEventRecord(ctx->ev_rm, s));
  return lagr_finish(ctx);
}

// tri_geo.f90:15-30 + setup_particles + init_geomCoords + init_mlsForce on the
// scaled unit mesh xyz0 (device copy d_xyz0), in the reference's order.
static int geometry_init_device(melt_ctx* ctx, const double* d_xyz0) {
  LagArgs A = lag_args(ctx);
  hipStream_t s = ctx->stream;
  const melt_body& h = ctx->h_body;
  face_sums(ctx, A, d_xyz0, (int)SUM_RIGID, ctx->d_red);
  hipLaunchKernelGGL(k_body_rigid, dim3(1), dim3(1), 0, s, ctx->d_body, ctx->d_red, ctx->cfg.dens_ratio);
  hipLaunchKernelGGL(k_geom_init_body, dim3(1), dim3(1), 0, s, ctx->d_body, ctx->cfg.xlen, ctx->cfg.ylen, ctx->cfg.zlen, h.h_eulerian, h.celvol,
                     h.e_thresh, h.v_thresh);
  hipLaunchKernelGGL(k_translate, dim3(nb(ctx->nv, 256)), dim3(256), 0, s, A, d_xyz0, ctx->d_body);
  hipLaunchKernelGGL(k_face_centroid_normal, dim3(nb(ctx->nf, 256)), dim3(256), 0, s, A, d_xyz0, 1, 0, 1);
  hipLaunchKernelGGL(k_clear_rm, dim3(1), dim3(1), 0, s, ctx->d_body);
  // E_thresh is still unset when tri_geo measures the edges
  hipLaunchKernelGGL(k_edge_len, dim3(nb(ctx->ne, 256)), dim3(256), 0, s, A, d_xyz0, ctx->d_body, 0.0);
  hipLaunchKernelGGL(k_face_area, dim3(nb(ctx->nf, 256)), dim3(256), 0, s, A, d_xyz0);
  face_sums(ctx, A, d_xyz0, (int)SUM_AREA_VOL_RIGID, ctx->d_red + 16);
  hipLaunchKernelGGL(k_surface, dim3(1), dim3(1), 0, s, ctx->d_body, ctx->d_red + 16);
  hipLaunchKernelGGL(k_skewness, dim3(nb(ctx->nf, 256)), dim3(256), 0, s, A);
  hipLaunchKernelGGL(k_face_centroid_normal, dim3(nb(ctx->nf, 256)), dim3(256), 0, s, A, d_xyz0, 0, 1, 0);
  hipLaunchKernelGGL(k_vert_geom, dim3(nb(ctx->nv, 256)), dim3(256), 0, s, A);
  hipLaunchKernelGGL(k_cfac, dim3(nb(ctx->nf, 256)), dim3(256), 0, s, A, ctx->d_body);
  MELT_HIP(hipGetLastError());
  MELT_HIP(hipStreamSynchronize(s));
  return 0;
}

// Host-side import: connectivity (set_connectivity partaux.f90:160-267),
// vertex -> face lists, device allocations, window capacities.
int lagr_alloc(melt_ctx* ctx, const double* xyz0, const int* voe, const int* eof, const int* vof_in, double rad_p) {
  (void)vof_in;
  int nv = ctx->nv, ne = ctx->ne, nf = ctx->nf;
  std::vector<int> vof(3 * (size_t)nf), v0(2 * (size_t)ne), e0(3 * (size_t)nf), foe(2 * (size_t)ne, 0);
  for (int i = 0; i < nf; ++i) {
    int e1 = eof[3 * i] - 1, e2 = eof[3 * i + 1] - 1, v1, v2, v3;
    if (e1 < 0 || e1 >= ne || e2 < 0 || e2 >= ne) return melt_fail(ctx, MELT_ERR_ARG, "set_geometry", "edge index out of range");
    if (voe[2 * e1 + 1] == voe[2 * e2]) { v1 = voe[2 * e1]; v2 = voe[2 * e1 + 1]; v3 = voe[2 * e2 + 1]; }
    else if (voe[2 * e1 + 1] == voe[2 * e2 + 1]) { v1 = voe[2 * e1]; v2 = voe[2 * e1 + 1]; v3 = voe[2 * e2]; }
    else if (voe[2 * e1] == voe[2 * e2]) { v1 = voe[2 * e1 + 1]; v2 = voe[2 * e1]; v3 = voe[2 * e2 + 1]; }
    else { v1 = voe[2 * e1 + 1]; v2 = voe[2 * e1]; v3 = voe[2 * e2]; }
    vof[3 * i] = v1 - 1; vof[3 * i + 1] = v2 - 1; vof[3 * i + 2] = v3 - 1;
    for (int c = 0; c < 3; ++c) {
      int v = vof[3 * i + c];
      if (v < 0 || v >= nv) return melt_fail(ctx, MELT_ERR_ARG, "set_geometry", "vertex index out of range");
    }
    for (int c = 0; c < 3; ++c) {
      int e = eof[3 * i + c] - 1;
      if (e < 0 || e >= ne) return melt_fail(ctx, MELT_ERR_ARG, "set_geometry", "edge index out of range");
      e0[3 * i + c] = e;
      if (foe[2 * e] == 0) foe[2 * e] = i + 1;
      else if (foe[2 * e + 1] == 0) foe[2 * e + 1] = i + 1;
      else return melt_fail(ctx, MELT_ERR_ARG, "set_geometry", "Edge error (edge shared by more than two faces)");
    }
  }
  for (int e = 0; e < ne; ++e) {
    v0[2 * e] = voe[2 * e] - 1; v0[2 * e + 1] = voe[2 * e + 1] - 1;
    if (v0[2 * e] < 0 || v0[2 * e] >= nv || v0[2 * e + 1] < 0 || v0[2 * e + 1] >= nv)
      return melt_fail(ctx, MELT_ERR_ARG, "set_geometry", "vertex index out of range");
  }
  std::vector<int> ptr(nv + 1, 0), lst(3 * (size_t)nf);
  for (int i = 0; i < 3 * nf; ++i) ptr[vof[i] + 1]++;
  for (int v = 0; v < nv; ++v) ptr[v + 1] += ptr[v];
  std::vector<int> fill(ptr.begin(), ptr.end() - 1);
  for (int f = 0; f < nf; ++f)
    for (int c = 0; c < 3; ++c) lst[fill[vof[3 * f + c]]++] = f;  // ascending f per vertex
  // scaled mesh (setup_particles: xyz0 = xyz0 * rad_p)
  std::vector<double> xs(3 * (size_t)nv);
  for (size_t i = 0; i < xs.size(); ++i) xs[i] = xyz0[i] * rad_p;
  // window capacities from the bounding sphere of the body
  double c[3] = {0, 0, 0};
  for (int v = 0; v < nv; ++v) for (int d = 0; d < 3; ++d) c[d] += xs[3 * v + d] / nv;
  double R = 0.0;
  for (int v = 0; v < nv; ++v) {
    double r2 = 0;
    for (int d = 0; d < 3; ++d) r2 += (xs[3 * v + d] - c[d]) * (xs[3 * v + d] - c[d]);
    R = std::max(R, std::sqrt(r2));
  }
  int nn[3] = {ctx->n1, ctx->n2, ctx->n3};
  double dx[3] = {ctx->dx1, ctx->dx2, ctx->dx3};
  for (int d = 0; d < 3; ++d) {
    int e = (int)std::ceil(2.0 * R * dx[d]) + 10;
    ctx->fwin[d] = std::min(e, nn[d]);
    ctx->twin[d] = std::min(e, nn[d]);
  }
  ctx->fwin_cells = (size_t)ctx->fwin[0] * ctx->fwin[1] * ctx->fwin[2];
  ctx->twin_cells = (size_t)ctx->twin[0] * ctx->twin[1] * ctx->twin[2];

  auto dalloc = [&](void** p, size_t bytes) -> int {
    MELT_HIP(hipMalloc(p, bytes ? bytes : 8));
    MELT_HIP(hipMemset(*p, 0, bytes ? bytes : 8));
    return 0;
  };
  static const int lag_len[MELT_L_NLAG] = {3, 3, 1, 1, 1, 3, 3, 3, 3, 3, 1, 3, 1, 1, 1, 1, 27, 27, 27, 27, 3, 3, 3, 1};
  for (int id = 0; id < MELT_L_NLAG; ++id) {
    size_t cnt = id <= MELT_L_DXYZV_S ? nv : id <= MELT_L_TAU_N ? nf : ne;
    if (dalloc((void**)&ctx->d_lag[id], sizeof(double) * lag_len[id] * cnt)) return MELT_ERR_HIP;
  }
  int rc = 0;
  rc |= dalloc((void**)&ctx->d_vof, sizeof(int) * 3 * nf);
  rc |= dalloc((void**)&ctx->d_eof, sizeof(int) * 3 * nf);
  rc |= dalloc((void**)&ctx->d_voe, sizeof(int) * 2 * ne);
  rc |= dalloc((void**)&ctx->d_v2f_ptr, sizeof(int) * (nv + 1));
  rc |= dalloc((void**)&ctx->d_v2f, sizeof(int) * 3 * nf);
  rc |= dalloc((void**)&ctx->d_gv, nv);
  rc |= dalloc((void**)&ctx->d_ge, ne);
  rc |= dalloc((void**)&ctx->d_gf, nf);
  rc |= dalloc((void**)&ctx->d_pind, sizeof(int) * 6 * nf);
  rc |= dalloc((void**)&ctx->d_pind2, sizeof(int) * 6 * nf);
  rc |= dalloc((void**)&ctx->d_probe, sizeof(double) * PRB_N * nf);
  rc |= dalloc((void**)&ctx->d_rxp, sizeof(double) * 3 * nf);
  rc |= dalloc((void**)&ctx->d_rxt, sizeof(double) * 3 * nf);
  rc |= dalloc((void**)&ctx->d_acc, sizeof(unsigned long long) * 8 * ctx->fwin_cells);
  rc |= dalloc((void**)&ctx->d_tkey, sizeof(unsigned long long) * ctx->twin_cells);
  rc |= dalloc((void**)&ctx->d_tface, sizeof(int) * ctx->twin_cells);
  rc |= dalloc((void**)&ctx->d_tdist, sizeof(double) * 125 * nf);
  rc |= dalloc((void**)&ctx->d_pcoef, sizeof(double) * 27 * 3 * 6 * nf);
  rc |= dalloc((void**)&ctx->d_foe, sizeof(int) * 2 * ne);
  if (ctx->cfg.iremesh == 1) rc |= dalloc((void**)&ctx->d_xyz_melt, sizeof(double) * 3 * nv);
  if (rc) return MELT_ERR_HIP;
  if (ctx->cfg.iremesh == 1) {
    if (!ctx->h_rm) MELT_HIP(hipHostMalloc((void**)&ctx->h_rm, sizeof(int32_t)));
    if (!ctx->ev_rm) MELT_HIP(hipEventCreateWithFlags(&ctx->ev_rm, hipEventDisableTiming));
  }
  // host mirror of the connectivity (remeshing edits it there first)
  ctx->h_vof.assign(vof.begin(), vof.end());
  ctx->h_eof.assign(e0.begin(), e0.end());
  ctx->h_voe.assign(v0.begin(), v0.end());
  ctx->h_foe.resize(2 * (size_t)ne);
  for (size_t i = 0; i < ctx->h_foe.size(); ++i) ctx->h_foe[i] = foe[i] - 1;
  ctx->h_gv.assign(nv, 0); ctx->h_ge.assign(ne, 0); ctx->h_gf.assign(nf, 0);
  ctx->h_anchor.assign(nv, 1); ctx->h_flagged.assign(ne, 0);
  for (double& v : ctx->remesh_stats) v = 0.0;
  {
    uint64_t h = 1469598103934665603ull;
    const unsigned char* b = reinterpret_cast<const unsigned char*>(vof.data());
    for (size_t i = 0; i < sizeof(int) * vof.size(); ++i) { h ^= b[i]; h *= 1099511628211ull; }
    ctx->mesh_hash0 = h;
  }
  MELT_HIP(hipDeviceSynchronize());  // the zero fills above are not ordered with the context's stream
  if (int rc2 = ibm_window_reset(ctx)) return rc2;
  MELT_HIP(hipMemcpy(ctx->d_vof, vof.data(), sizeof(int) * 3 * nf, hipMemcpyHostToDevice));
  MELT_HIP(hipMemcpy(ctx->d_eof, e0.data(), sizeof(int) * 3 * nf, hipMemcpyHostToDevice));
  MELT_HIP(hipMemcpy(ctx->d_voe, v0.data(), sizeof(int) * 2 * ne, hipMemcpyHostToDevice));
  MELT_HIP(hipMemcpy(ctx->d_v2f_ptr, ptr.data(), sizeof(int) * (nv + 1), hipMemcpyHostToDevice));
  MELT_HIP(hipMemcpy(ctx->d_v2f, lst.data(), sizeof(int) * 3 * nf, hipMemcpyHostToDevice));
  MELT_HIP(hipMemcpy(ctx->d_foe, ctx->h_foe.data(), sizeof(int) * 2 * ne, hipMemcpyHostToDevice));
  double* d_xyz0 = nullptr;
  MELT_HIP(hipMalloc(&d_xyz0, sizeof(double) * 3 * nv));
  MELT_HIP(hipMemcpy(d_xyz0, xs.data(), sizeof(double) * 3 * nv, hipMemcpyHostToDevice));
  rc = geometry_init_device(ctx, d_xyz0);
  hipFree(d_xyz0);
  return rc;
}
