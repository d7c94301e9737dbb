// Immersed-boundary kernels of the RK3 substep on gfx950:
//  - narrow-band signed-distance tagging (genSDF.f90) with a deterministic
//    arg-min (smallest distance, then lowest face index: the reference's
//    "first face wins"), so tags equal the serial CPU result bit for bit;
//  - init-time ray tagging (pencil_rayTag.f90, rayAux.f90);
//  - MLS shape functions (mlsWeight.f90), interpolation and force spreading
//    (mlsForce.f90, forc1-3.f90, forctemp.f90).  Spreading accumulates each
//    contribution as a 128-bit fixed-point integer, so the sum over faces is
//    exact and independent of scheduling order;
//  - MLS gradient probes for the heat flux (mls_heatFlux.f90) and the surface
//    loads (mls_structLoads.f90).
// All of this scales with the number of triangles, not with N^3.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

#include "common.h"
#include "fixed128.h"

namespace {

constexpr double WCON = 0.3, WSCL = 1.5;  // initmls.f90:12-13

__device__ inline double gco(const Grid& g, int dim, int centre, int i) {
  if (dim == 0) return centre ? gxm(g, i) : gxc(g, i);
  if (dim == 1) return centre ? gym(g, i) : gyc(g, i);
  return centre ? gzm(g, i) : gzc(g, i);
}

__device__ inline double mls_gaussian(double r) {  // geom.f90:24-36
  if (r > 1.0) return 0.0;
  double t = r / WCON;
  return det_exp(-(t * t));
}
__device__ inline double mls_gauss_deriv(double r) {  // geom.f90:38-53
  if (r > 1.0) return 0.0;
  double t = r / WCON;
  return -2.0 * r / (WCON * WCON) * det_exp(-(t * t));
}

// A(i,j) += (w p_j) p_i : DGEMM('N','T',4,4,1,w,p,4,p,4,1,A,4)
__device__ inline void rank1(double A[16], double w, const double p[4]) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    double temp = w * p[j];
#pragma unroll
    for (int i = 0; i < 4; ++i) A[i + 4 * j] = A[i + 4 * j] + temp * p[i];
  }
}

// MLS shape function over the 3x3x3 support i0..i0+2, j0.., k0.. (wght1/2/3/temp,
// mlsWeight.f90:48-412); c[d] selects face (0) or centre (1) coordinates.
// sink(inw, phi) receives the 27 values in the reference's k, j, i order.
template <class Sink>
__device__ int mls_phi(const Grid& g, const double pos[3], int i0, int j0, int k0, const int c[3], Sink sink) {
  double A[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) A[q] = 0.0;
  // separable Gaussian factors, evaluated once per axis
  double xs[3], ys[3], zs[3], wx[3], wy[3], wz[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    xs[q] = gco(g, 0, c[0], i0 + q);
    wx[q] = mls_gaussian(fabs(xs[q] - pos[0]) / (WSCL / g.dx1));
    ys[q] = gco(g, 1, c[1], j0 + q);
    wy[q] = mls_gaussian(fabs(ys[q] - pos[1]) / (WSCL / g.dx2));
    zs[q] = gco(g, 2, c[2], k0 + q);
    wz[q] = mls_gaussian(fabs(zs[q] - pos[2]) / (WSCL / g.dx3));
  }
#pragma unroll 1
  for (int kq = 0; kq < 3; ++kq)
#pragma unroll
    for (int jq = 0; jq < 3; ++jq) {
      double w23 = wy[jq] * wz[kq];
#pragma unroll
      for (int iq = 0; iq < 3; ++iq) {
        double wtx = wx[iq] * w23;
        double p[4] = {1.0, xs[iq], ys[jq], zs[kq]};
        rank1(A, wtx, p);
      }
    }
  int info = lu_inverse<4>(A);
  if (info) return info;
  double ptx[4] = {1.0, pos[0], pos[1], pos[2]}, ptxA[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    double s = 0.0;
#pragma unroll
    for (int l = 0; l < 4; ++l) s = s + (1.0 * A[l + 4 * j]) * ptx[l];
    ptxA[j] = s;
  }
#pragma unroll 1
  for (int kq = 0; kq < 3; ++kq)
#pragma unroll
    for (int jq = 0; jq < 3; ++jq) {
      double w23 = wy[jq] * wz[kq];
#pragma unroll
      for (int iq = 0; iq < 3; ++iq) {
        double wtx = wx[iq] * w23;
        double b[4] = {wtx * 1.0, wtx * xs[iq], wtx * ys[jq], wtx * zs[kq]};
        double s = 0.0;
#pragma unroll
        for (int l = 0; l < 4; ++l) s = s + (1.0 * b[l]) * ptxA[l];
        sink(9 * kq + 3 * jq + iq, s);
      }
    }
  return 0;
}

__device__ inline double dweight(double pc, double gcrd, double nrm, double scale) {
  if (nrm < 2.220446049250313e-16) return 0.0;  // EPSILON(1.0d0)
  double d = mls_gauss_deriv(nrm);
  return d * ((pc - gcrd) / fabs(pc - gcrd) / scale);
}

// MLS gradient at a probe (wght_gradT mls_heatFlux.f90:85-300 and
// wght_gradP/U/V/W mls_structLoads.f90:380-1297), computed by the 4 lanes of
// one probe: lane 0 accumulates the moment matrix A, lane 1+d the derivative
// matrix dA/dx_d; after A is shared through shA (16 doubles, block barrier),
// lanes 1..3 each invert A and form gradient component d in the reference's
// point order, so every value rounds as in the single-thread formulation.
// kz: grid indices used for the z coordinate (the reference uses k1-1, k1,
// k1); ke: field planes (1-based, ghost-plane semantics) k1-1, k1, k1+1.
// Must be reached by every thread of the block (it contains a barrier);
// active = 0 lanes only take part in the barrier.  Returns the LU info of A
// on lanes 1..3 (0 elsewhere); *grad_d receives component lane-1.
template <typename R>
__device__ __forceinline__ int mls_grad_lanes(const Grid& g, const R* __restrict__ fld, const double pos[3], int ic, int jc, const int kz[3],
                              const int ke[3], const int c[3], int lane, bool active, double* shA, double* grad_d,
                              double* coef = nullptr, size_t cstride = 0) {
  const double sx = WSCL / g.dx1, sy = WSCL / g.dx2, sz = WSCL / g.dx3;
  // the field planes, read once: a per-plane pick from ke[] inside the rolled
  // loop is folded into one load at a selected offset, which keeps the
  // caller's probe-site record in scratch memory
  const int ke0 = ke[0], ke1 = ke[1], ke2 = ke[2];
  double xs[3], ys[3], zs[3], w1[3], w2[3], w3[3], dw1[3], dw2[3], dw3[3];
  double M[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) M[q] = 0.0;
  if (active) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      xs[q] = gco(g, 0, c[0], ic - 1 + q);
      ys[q] = gco(g, 1, c[1], jc - 1 + q);
      zs[q] = gco(g, 2, c[2], kz[q]);
    }
    // separable Gaussian factors and derivatives of the nine (axis, point)
    // pairs: pair p = 3 axis + point is evaluated on lane p % 4 of the
    // probe's 4-lane group only (one exponential per pair instead of one per
    // pair and lane) and read by the others from there
    double mw[3], md[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int pp = lane + 4 * r, ax = pp / 3, q = pp - 3 * ax;
      // selects, not an indexed pick (which would put the arrays in scratch)
      const double cx = q == 0 ? xs[0] : q == 1 ? xs[1] : xs[2], cy = q == 0 ? ys[0] : q == 1 ? ys[1] : ys[2];
      const double cz = q == 0 ? zs[0] : q == 1 ? zs[1] : zs[2];
      const double crd = ax == 0 ? cx : ax == 1 ? cy : cz;
      const double pc = ax == 0 ? pos[0] : ax == 1 ? pos[1] : pos[2];
      const double sc = ax == 0 ? sx : ax == 1 ? sy : sz;
      const double n = fabs(pc - crd) / sc;
      mw[r] = pp < 9 ? mls_gaussian(n) : 0.0;
      md[r] = pp < 9 ? dweight(pc, crd, n, sc) : 0.0;
    }
    double wv[9], dv[9];
#pragma unroll
    for (int pp = 0; pp < 9; ++pp) {
      wv[pp] = __shfl(mw[pp / 4], pp % 4, 4);
      dv[pp] = __shfl(md[pp / 4], pp % 4, 4);
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      w1[q] = wv[q]; dw1[q] = dv[q];
      w2[q] = wv[3 + q]; dw2[q] = dv[3 + q];
      w3[q] = wv[6 + q]; dw3[q] = dv[6 + q];
    }
#pragma unroll 1
    for (int nk = 0; nk < 3; ++nk) {
      // the z loop stays rolled (code size); selecting its plane's values
      // keeps the arrays in registers (an index by nk would put them in scratch)
      const double zk = nk == 0 ? zs[0] : nk == 1 ? zs[1] : zs[2];
      const double w3k = nk == 0 ? w3[0] : nk == 1 ? w3[1] : w3[2];
      const double dw3k = nk == 0 ? dw3[0] : nk == 1 ? dw3[1] : dw3[2];
#pragma unroll
      for (int jq = 0; jq < 3; ++jq)
#pragma unroll
        for (int iq = 0; iq < 3; ++iq) {
          double p[4] = {1.0, xs[iq], ys[jq], zk};
          double wgt;
          if (lane == 0) wgt = w1[iq] * (w2[jq] * w3k);
          else if (lane == 1) wgt = dw1[iq] * w2[jq] * w3k;
          else if (lane == 2) wgt = w1[iq] * dw2[jq] * w3k;
          else wgt = w1[iq] * w2[jq] * dw3k;
          rank1(M, wgt, p);
        }
    }
    if (lane == 0)
#pragma unroll
      for (int q = 0; q < 16; ++q) shA[q] = M[q];
  }
  __syncthreads();
  if (!active || lane == 0) return 0;
  double A[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) A[q] = shA[q];
  int info = lu_inverse<4>(A);
  if (info) return info;
  double ptx[4] = {1.0, pos[0], pos[1], pos[2]}, gam[4], dg[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) gam[i] = 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    double t = 1.0 * ptx[j];
#pragma unroll
    for (int i = 0; i < 4; ++i) gam[i] = gam[i] + t * A[i + 4 * j];
  }
  const int d = lane - 1;
  double pv[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) pv[i] = i == d + 1 ? 1.0 : 0.0;   // unit vector e_(d+1), in registers
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    double t = -1.0 * gam[j];
#pragma unroll
    for (int i = 0; i < 4; ++i) pv[i] = pv[i] + t * M[i + 4 * j];
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) dg[i] = 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    double t = 1.0 * pv[j];
#pragma unroll
    for (int i = 0; i < 4; ++i) dg[i] = dg[i] + t * A[i + 4 * j];
  }
  double s = 0.0;
#pragma unroll 1
  for (int nk = 0; nk < 3; ++nk) {
    const double zk = nk == 0 ? zs[0] : nk == 1 ? zs[1] : zs[2];
    const double w3k = nk == 0 ? w3[0] : nk == 1 ? w3[1] : w3[2];
    const double dw3k = nk == 0 ? dw3[0] : nk == 1 ? dw3[1] : dw3[2];
    const int kk = nk == 0 ? ke0 : nk == 1 ? ke1 : ke2;
#pragma unroll
    for (int jq = 0; jq < 3; ++jq)
#pragma unroll
      for (int iq = 0; iq < 3; ++iq) {
        double wtx = w1[iq] * (w2[jq] * w3k);
        double p[4] = {1.0, xs[iq], ys[jq], zk};
        double dw = d == 0 ? dw1[iq] * w2[jq] * w3k : d == 1 ? w1[iq] * dw2[jq] * w3k : w1[iq] * w2[jq] * dw3k;
        double acc = 0.0;
#pragma unroll
        for (int l = 0; l < 4; ++l) acc = acc + (1.0 * (wtx * p[l])) * dg[l];
#pragma unroll
        for (int l = 0; l < 4; ++l) acc = acc + (1.0 * (dw * p[l])) * gam[l];
        if (coef) {   // geometry only: the point's shape-function derivative
          coef[(size_t)(9 * nk + 3 * jq + iq) * cstride] = acc;
        } else {
          double val = fld[widx(g, ic - 1 + iq, jc - 1 + jq, kk)];
          s = s + acc * val;
        }
      }
  }
  *grad_d = s;
  return 0;
}

// ---------------------------------------------------- 128-bit accumulation
// acc_split / acc_value (fixed128.h) hold each contribution exactly as a
// 64.64 fixed-point integer; the atomics add the two words with carry.  The
// high word changes only when the running sum crosses a multiple of 2^64
// (for a small negative contribution its all-ones high word and the carry
// cancel), so its atomic is issued only when hi + carry is not zero.
__device__ inline void acc_add(unsigned long long* w /* [lo, hi] */, double v) {
  unsigned long long lo, hi;
  acc_split(v, lo, hi);
  if ((lo | hi) == 0) return;
  unsigned long long old = atomicAdd(&w[0], lo);
  unsigned long long carry = (old + lo < old) ? 1ull : 0ull;
  if (hi + carry) atomicAdd(&w[1], hi + carry);
}

__device__ inline long long win_slot(const int org[3], const int E[3], int iu, int ju, int ku) {
  int a = iu - org[0], b = ju - org[1], c = ku - org[2];
  if (a < 0 || b < 0 || c < 0 || a >= E[0] || b >= E[1] || c >= E[2]) return -1;
  return (long long)a + (long long)E[0] * ((long long)b + (long long)E[1] * c);
}

struct Win3 { int e[3]; };

// ------------------------------------------------------------ windows
// Face base indices floor(x*dx)+1 (unwrapped) bound every cell the MLS
// supports (base-1..base+2) and the tagging band (base-2..base+2) touch.
// Min / max of the face base indices over all live faces: one block per 256
// faces, combined with integer atomics (order-independent).
__device__ void window_box_block(const Grid& g, const double* __restrict__ bar, const unsigned char* __restrict__ gf, int nf, melt_window* win,
                                 int blk) {
  __shared__ int mn[3], mx[3];
  if (threadIdx.x < 3) { mn[threadIdx.x] = INT_MAX; mx[threadIdx.x] = INT_MIN; }
  __syncthreads();
  int f = blk * blockDim.x + threadIdx.x;
  if (f < nf && !gf[f]) {
    const double dx[3] = {g.dx1, g.dx2, g.dx3};
    for (int d = 0; d < 3; ++d) {
      int b = (int)floor(bar[3 * f + d] * dx[d]) + 1;
      atomicMin(&mn[d], b);
      atomicMax(&mx[d], b);
    }
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    atomicMin(&win->bmin[threadIdx.x], mn[threadIdx.x]);
    atomicMax(&win->bmax[threadIdx.x], mx[threadIdx.x]);
  }
}
__global__ __launch_bounds__(256) void k_window_box(Grid g, const double* __restrict__ bar, const unsigned char* __restrict__ gf, int nf,
                                                    melt_window* win) {
  window_box_block(g, bar, gf, nf, win, blockIdx.x);
}

// Initial state of the windows: no box placed yet, empty extents.
__global__ void k_window_reset(melt_window* win, int n) {
  for (int w = 0; w < n; ++w) {
    for (int d = 0; d < 3; ++d) {
      win[w].org[d] = 0; win[w].prev_org[d] = 0;
      win[w].bmin[d] = INT_MAX; win[w].bmax[d] = INT_MIN;
    }
    win[w].prev_valid = 0;
    win[w].pad = 0;
  }
}

// Place the forcing (which & 1: win[0]) and / or tagging (which & 2: win[1])
// box from the face extent reduced into win[0].bmin/bmax, then reset the
// extent for the next call.
__global__ void k_window_final(melt_window* win, Win3 ef, Win3 et, int which, int* err) {
  int mn[3], mx[3];
  for (int d = 0; d < 3; ++d) { mn[d] = win[0].bmin[d]; mx[d] = win[0].bmax[d]; win[0].bmin[d] = INT_MAX; win[0].bmax[d] = INT_MIN; }
  if (which & 1) {  // forcing window
    melt_window& w = win[0];
    for (int d = 0; d < 3; ++d) {
      w.org[d] = mn[d] - 2;
      if (mx[d] + 3 - w.org[d] >= ef.e[d]) raise_error(err, MELT_ERR_WINDOW);
    }
  }
  if (which & 2) {  // tagging window
    melt_window& w = win[1];
    if (w.prev_valid >= 0) {
      for (int d = 0; d < 3; ++d) w.prev_org[d] = w.org[d];
      w.prev_valid = w.pad;
    }
    for (int d = 0; d < 3; ++d) {
      w.org[d] = mn[d] - 2;
      if (mx[d] + 2 - w.org[d] >= et.e[d]) raise_error(err, MELT_ERR_WINDOW);
    }
    w.pad = 1;  // current box is valid from now on
  }
}

// ------------------------------------------------------------ narrow band
// genSDF.f90:85-234, Eberly's point-triangle distance.
__device__ double dist_point_tri(const double P[3], const double* V0, const double* V1, const double* V2) {
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
        else { s = 0.0; t = (e >= 0.0) ? 0.0 : ((-e >= c) ? 1.0 : -e / c); }
      } else {
        s = 0.0; t = (e >= 0.0) ? 0.0 : ((-e >= c) ? 1.0 : -e / c);
      }
    } else if (t < 0.0) {
      t = 0.0; s = (d >= 0.0) ? 0.0 : ((-d >= a) ? 1.0 : -d / a);
    } else {
      double inv = 1.0 / det;
      s = s * inv; t = t * inv;
    }
  } else {
    if (s < 0.0) {
      double t0 = b + d, t1 = c + e;
      if (t1 > t0) {
        double num = t1 - t0, den = a - 2.0 * b + c;
        if (num >= den) { s = 1.0; t = 0.0; } else { s = num / den; t = 1.0 - s; }
      } else {
        s = 0.0; t = (t1 <= 0.0) ? 1.0 : ((e >= 0.0) ? 0.0 : -e / c);
      }
    } else if (t < 0.0) {
      double t0 = b + e, t1 = a + d;
      if (t1 > t0) {
        double num = t1 - t0, den = a - 2.0 * b + c;
        if (num >= den) { t = 1.0; s = 0.0; } else { t = num / den; s = 1.0 - t; }
      } else {
        t = 0.0; s = (t1 <= 0.0) ? 1.0 : ((d >= 0.0) ? 0.0 : -d / a);
      }
    } else {
      double num = c + e - b - d, den = a - 2.0 * b + c;
      s = (num <= 0.0) ? 0.0 : ((num >= den) ? 1.0 : num / den);
      t = 1.0 - s;
    }
  }
  double q[3];
  for (int k = 0; k < 3; ++k) q[k] = (V0[k] + s * E0[k] + t * E1[k]) - P[k];
  return sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
}

// The tagging's set-up in one launch: blocks [0, nbox) reduce the face
// extent (as k_window_box), the other 256 set sdf to 999 over the window the
// previous tagging wrote (that tagging's cells are re-tagged from scratch)
// and clear the per-cell keys.  This runs before k_window_final moves the
// tagging window, so it takes the previous window from where
// k_window_final's update leaves it (prev_org := org, prev_valid := pad,
// while prev_valid >= 0).
template <typename R>
__global__ __launch_bounds__(256) void k_tag_prepare(Grid g, const double* __restrict__ bar, const unsigned char* __restrict__ gf, int nf,
                                                     int nbox, melt_window* win, R* __restrict__ sdf, Win3 et,
                                                     unsigned long long* __restrict__ key, int* __restrict__ face, size_t ncell) {
  if ((int)blockIdx.x < nbox) {
    window_box_block(g, bar, gf, nf, win, blockIdx.x);
    return;
  }
  const size_t t0 = (size_t)(blockIdx.x - nbox) * blockDim.x + threadIdx.x, step = (size_t)(gridDim.x - nbox) * blockDim.x;
  const melt_window& w = win[1];
  const bool moved = w.prev_valid >= 0;
  if ((moved ? w.pad : w.prev_valid) > 0) {
    const int o0 = moved ? w.org[0] : w.prev_org[0], o1 = moved ? w.org[1] : w.prev_org[1], o2 = moved ? w.org[2] : w.prev_org[2];
    const size_t n = (size_t)et.e[0] * et.e[1] * et.e[2];
    for (size_t s = t0; s < n; s += step) {
      int a = (int)(s % et.e[0]), b = (int)((s / et.e[0]) % et.e[1]), c = (int)(s / ((size_t)et.e[0] * et.e[1]));
      sdf[widx(g, o0 + a, o1 + b, o2 + c)] = R(999.0);
    }
  }
  for (size_t s = t0; s < ncell; s += step) {
    key[s] = ~0ull;
    face[s] = INT_MAX;
  }
}

struct MeshArgs {
  const double* __restrict__ xyzv;
  const double* __restrict__ bar;
  const double* __restrict__ nor;
  const int* __restrict__ vof;
  const unsigned char* __restrict__ gf;
  int nf;
};

// mode 0: distance of every (face, band cell), kept in dist[], and the min
// per cell; 1: lowest face among the minima; 2: write sdf and VOFp
template <typename R>
__global__ __launch_bounds__(256) void k_tag_pass(Grid g, MeshArgs M, const melt_window* __restrict__ win, Win3 et,
                                                  unsigned long long* __restrict__ key, int* __restrict__ face, double* __restrict__ dst,
                                                  R* __restrict__ sdf, R* __restrict__ vofp, int mode, double pi, int* err) {
  size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)M.nf * 125) return;
  int f = (int)(id / 125), q = (int)(id % 125);
  if (M.gf[f]) return;
  const double* pos = &M.bar[3 * f];
  int i = (int)floor(pos[0] * g.dx1) + 1 + (q % 5) - 2;
  int j = (int)floor(pos[1] * g.dx2) + 1 + (q / 5) % 5 - 2;
  int k = (int)floor(pos[2] * g.dx3) + 1 + q / 25 - 2;
  long long s = win_slot(win[1].org, et.e, i, j, k);
  if (s < 0) { raise_error(err, MELT_ERR_WINDOW); return; }
  if (mode == 0) {
    const double* V0 = &M.xyzv[3 * M.vof[3 * f + 0]];
    const double* V1 = &M.xyzv[3 * M.vof[3 * f + 1]];
    const double* V2 = &M.xyzv[3 * M.vof[3 * f + 2]];
    double P[3] = {gxm(g, i), gym(g, j), gzm(g, k)};
    double dist = dist_point_tri(P, V0, V1, V2);
    dst[id] = dist;
    atomicMin(&key[s], (unsigned long long)__double_as_longlong(dist));
    return;
  }
  const double dist = dst[id];
  if (mode == 1) {
    if (key[s] == (unsigned long long)__double_as_longlong(dist)) atomicMin(&face[s], f);
  } else if (face[s] == f) {
    double P[3] = {gxm(g, i), gym(g, j), gzm(g, k)};
    const double* n = &M.nor[3 * f];
    double dot = n[0] * (pos[0] - P[0]) + n[1] * (pos[1] - P[1]) + n[2] * (pos[2] - P[2]);
    double sgn = signbit(dot) ? -1.0 : 1.0;
    double sg = dist * (-sgn);
    size_t c = widx(g, i, j, k);
    sdf[c] = (R)sg;
    double eps = 1.0 / g.dx1, v;  // sdf_to_VOF genSDF.f90:237-255
    if (sg <= -eps) v = 0.0;
    else if (sg >= eps) v = 1.0;
    else v = 0.5 * (1.0 + sg * g.dx1 + (1.0 / pi) * det_sin(pi * sg * g.dx1));
    vofp[c] = (R)v;
  }
}

// ------------------------------------------------------------ ray tagging
__device__ bool edge_sq(double y1, double z1, double y2, double z2, double ymin, double ymax, double zmin, double zmax) {
  double t, y, z;  // rayAux.f90:139-190
  if (y1 != y2) {
    t = (ymin - y1) / (y2 - y1);
    if (t >= 0.0 && t <= 1.0) { z = z1 + t * (z2 - z1); if (z >= zmin && z <= zmax) return true; }
    t = (ymax - y1) / (y2 - y1);
    if (t >= 0.0 && t <= 1.0) { z = z1 + t * (z2 - z1); if (z >= zmin && z <= zmax) return true; }
  }
  if (z1 != z2) {
    t = (zmin - z1) / (z2 - z1);
    if (t >= 0.0 && t <= 1.0) { y = y1 + t * (y2 - y1); if (y >= ymin && y <= ymax) return true; }
    t = (zmax - z1) / (z2 - z1);
    if (t >= 0.0 && t <= 1.0) { y = y1 + t * (y2 - y1); if (y >= ymin && y <= ymax) return true; }
  }
  return false;
}

__device__ bool yz_range(const double* V0, const double* V1, const double* V2, double ymin, double ymax, double zmin, double zmax) {
  double yv[3] = {V0[1], V1[1], V2[1]}, zv[3] = {V0[2], V1[2], V2[2]};  // rayAux.f90:86-113
  for (int i = 0; i < 3; ++i)
    if (yv[i] >= ymin && yv[i] <= ymax && zv[i] >= zmin && zv[i] <= zmax) return true;
  double ymx = fmax(fmax(yv[0], yv[1]), yv[2]), ymn = fmin(fmin(yv[0], yv[1]), yv[2]);
  double zmx = fmax(fmax(zv[0], zv[1]), zv[2]), zmn = fmin(fmin(zv[0], zv[1]), zv[2]);
  if (ymx < ymin || ymn > ymax || zmx < zmin || zmn > zmax) return false;
  for (int i = 0; i < 3; ++i) {
    int j = (i + 1) % 3;
    if (edge_sq(yv[i], zv[i], yv[j], zv[j], ymin, ymax, zmin, zmax)) return true;
  }
  return false;
}

struct Box { int lo[3], hi[3]; };

__global__ __launch_bounds__(256) void k_pencil_build(Grid g, MeshArgs M, Box bx, int* __restrict__ count, int* __restrict__ list, int buff,
                                                      int* err) {
  int js = bx.hi[1] - bx.lo[1] + 1, ks = bx.hi[2] - bx.lo[2] + 1;
  size_t np = (size_t)js * ks;
  size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= np * M.nf) return;
  int f = (int)(id % M.nf);
  size_t p = id / M.nf;
  if (M.gf[f]) return;
  int j = bx.lo[1] + (int)(p % js), k = bx.lo[2] + (int)(p / js);
  double ymin = gym(g, j) - 0.5 / g.dx2, ymax = gym(g, j) + 0.5 / g.dx2;
  double zmin = gzm(g, k) - 0.5 / g.dx3, zmax = gzm(g, k) + 0.5 / g.dx3;
  const double* V0 = &M.xyzv[3 * M.vof[3 * f + 0]];
  const double* V1 = &M.xyzv[3 * M.vof[3 * f + 1]];
  const double* V2 = &M.xyzv[3 * M.vof[3 * f + 2]];
  if (yz_range(V0, V1, V2, ymin, ymax, zmin, zmax)) {
    int n = atomicAdd(&count[p], 1);
    if (n < buff) list[p * buff + n] = f;
    else raise_error(err, MELT_ERR_PENCIL_OVERFLOW);
  }
}

// rayTriangle_intersect (rayAux.f90:9-83).  A near-parallel ray re-enters with
// an origin shifted by (0, 0.1/dx2, -0.1/dx3); the caller then re-applies its own
// barycentric tests on top of the callee's verdict.  Evaluated here level by
// level from the innermost call outwards.
__device__ bool ray_tri(const Grid& g, const double C0[3], const double Q[3], const double* V0, const double* V1, const double* V2,
                        double ip[3], int* err) {
  const double EPS = 1.e-18;
  const int MAXD = 8;
  double Cs[MAXD + 1][3];
  int L = 0;
  Cs[0][0] = C0[0]; Cs[0][1] = C0[1]; Cs[0][2] = C0[2];
  double e1[3], e2[3];
  for (int d = 0; d < 3; ++d) { e1[d] = V1[d] - V0[d]; e2[d] = V2[d] - V0[d]; }
  // find the first level whose ray is not parallel
  for (;;) {
    double r[3] = {Q[0] - Cs[L][0], Q[1] - Cs[L][1], Q[2] - Cs[L][2]}, p[3];
    dcross(p, r, e2);
    double a = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    if (!(a > -EPS && a <= EPS)) break;
    if (L == MAXD) { raise_error(err, MELT_ERR_RAY_RECURSION); return false; }
    Cs[L + 1][0] = Cs[L][0]; Cs[L + 1][1] = Cs[L][1] + 0.1 / g.dx2; Cs[L + 1][2] = Cs[L][2] - 0.1 / g.dx3;
    ++L;
  }
  bool hit = true;
  for (int lev = L; lev >= 0; --lev) {
    const double* C = Cs[lev];
    double r[3], s[3], p[3], q1[3];
    for (int d = 0; d < 3; ++d) { r[d] = Q[d] - C[d]; s[d] = C[d] - V0[d]; }
    dcross(p, r, e2);
    double a = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    if (lev == L) hit = true;  // intersect = .true. at entry of the innermost call
    double f = 1.0 / a;
    double u = f * (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]);
    if (u < EPS || u > 1.0) hit = false;
    dcross(q1, s, e1);
    double v = f * (r[0] * q1[0] + r[1] * q1[1] + r[2] * q1[2]);
    if (v < EPS || (u + v) > 1.0) hit = false;
    double t = f * (e2[0] * q1[0] + e2[1] * q1[1] + e2[2] * q1[2]);
    if (t < 0.0 || t > 1.0) hit = false;
    if (hit)
      for (int d = 0; d < 3; ++d) ip[d] = C[d] + t * r[d];
  }
  return hit;
}

__device__ double ray_tag_cell(const Grid& g, const MeshArgs& M, double xlen, const double Q[3], const int* list, int n, int* err) {
  double C[3] = {-xlen * 10.0, Q[1], Q[2]};  // pencil_rayTag.f90:402
  int cnt = 0;
  double xmin = Q[0] - 0.5 / g.dx1, xmax = Q[0] + 0.5 / g.dx1;
  for (int t = 0; t < n; ++t) {
    int f = list[t];
    if (M.gf[f]) continue;
    const double* V0 = &M.xyzv[3 * M.vof[3 * f + 0]];
    const double* V1 = &M.xyzv[3 * M.vof[3 * f + 1]];
    const double* V2 = &M.xyzv[3 * M.vof[3 * f + 2]];
    double mn = fmin(fmin(V0[0], V1[0]), V2[0]), mx = fmax(fmax(V0[0], V1[0]), V2[0]);  // triangleXIntersect
    if (!(mx < xmin || mn > xmax)) {
      if ((V0[0] >= xmin && V0[0] <= xmax) || (V1[0] >= xmin && V1[0] <= xmax) || (V2[0] >= xmin && V2[0] <= xmax)) return 0.5;
    }
    double ip[3] = {0.0, 0.0, 0.0};
    if (ray_tri(g, C, Q, V0, V1, V2, ip, err)) {
      bool inside = ip[0] >= xmin && ip[0] <= xmax && ip[1] >= Q[1] - 0.5 / g.dx2 && ip[1] <= Q[1] + 0.5 / g.dx2 && ip[2] >= Q[2] - 0.5 / g.dx3 &&
                    ip[2] <= Q[2] + 0.5 / g.dx3;
      if (inside) return 0.5;
      cnt += 1;
    }
  }
  return (cnt % 2 == 0) ? 1.0 : 0.0;
}

// pencil_tagCells (pencil_rayTag.f90:189-291): every wrapped cell of the box
// visits its unwrapped preimages in the reference's order (i, then j, then k)
// and keeps the first value below 1.
template <typename R>
__global__ __launch_bounds__(256) void k_pencil_tag(Grid g, MeshArgs M, Box bx, const int* __restrict__ count, const int* __restrict__ list,
                                                    int buff, R* __restrict__ vofp, double xlen, int* err) {
  int ext[3], cov[3];
  int nn[3] = {g.n1, g.n2, g.n3};
  for (int d = 0; d < 3; ++d) { ext[d] = bx.hi[d] - bx.lo[d] + 1; cov[d] = min(ext[d], nn[d]); }
  size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)cov[0] * cov[1] * cov[2]) return;
  int ta = (int)(id % cov[0]), tb = (int)((id / cov[0]) % cov[1]), tc = (int)(id / ((size_t)cov[0] * cov[1]));
  int js = ext[1];
  double cur = 1.0;
  for (int i = bx.lo[0] + ta; i <= bx.hi[0]; i += g.n1)
    for (int j = bx.lo[1] + tb; j <= bx.hi[1]; j += g.n2)
      for (int kk = bx.lo[2] + tc; kk <= bx.hi[2]; kk += g.n3) {
        double Q[3] = {gxm(g, i), gym(g, j), gzm(g, kk)};
        size_t p = (size_t)(j - bx.lo[1]) + (size_t)js * (kk - bx.lo[2]);
        int n = min(count[p], buff);
        double v = ray_tag_cell(g, M, xlen, Q, &list[p * buff], n, err);
        if (!(cur < 1.0)) cur = v;
      }
  vofp[widx(g, bx.lo[0] + ta, bx.lo[1] + tb, bx.lo[2] + tc)] = (R)cur;
}

__global__ __launch_bounds__(256) void k_vbox(const double* __restrict__ xyzv, const unsigned char* __restrict__ gv, int nv, double* out) {
  __shared__ double mn[3][256], mx[3][256];
  double a[3] = {INFINITY, INFINITY, INFINITY}, b[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int v = threadIdx.x; v < nv; v += blockDim.x) {
    if (gv[v]) continue;
    for (int d = 0; d < 3; ++d) { a[d] = fmin(a[d], xyzv[3 * v + d]); b[d] = fmax(b[d], xyzv[3 * v + d]); }
  }
  for (int d = 0; d < 3; ++d) { mn[d][threadIdx.x] = a[d]; mx[d][threadIdx.x] = b[d]; }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int d = 0; d < 3; ++d) {
        mn[d][threadIdx.x] = fmin(mn[d][threadIdx.x], mn[d][threadIdx.x + w]);
        mx[d][threadIdx.x] = fmax(mx[d][threadIdx.x], mx[d][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int d = 0; d < 3; ++d) { out[d] = mn[d][0]; out[3 + d] = mx[d][0]; }
}

template <typename R>
__global__ __launch_bounds__(256) void k_fill(R* __restrict__ p, double v, size_t n) {
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) p[c] = (R)v;
}

// ------------------------------------------------------------ MLS forcing
__global__ __launch_bounds__(256) void k_centroid_idx(Grid g, const double* __restrict__ bar, const unsigned char* __restrict__ gf, int nf,
                                                      int* __restrict__ pind) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;  // findCentroidIndices.f90:16-48
  if (f >= nf || gf[f]) return;
  const double* p = &bar[3 * f];
  int* o = &pind[6 * f];
  o[0] = (int)floor(p[0] * g.dx1) + 1; o[3] = fnint(p[0] * g.dx1) + 1;
  o[1] = (int)floor(p[1] * g.dx2) + 1; o[4] = fnint(p[1] * g.dx2) + 1;
  int k1 = (int)floor(p[2] * g.dx3) + 1, kst = fnint(p[2] * g.dx3) + 1;
  o[2] = pmod(k1 - 1, g.n3) + 1; o[5] = pmod(kst - 1, g.n3) + 1;
}

struct PhiOut { double* q[4]; };  // ptxAB_q1, ptxAB_q2, ptxAB_temp, ptxAB_q3

// component c: 0 = wght1 (u), 1 = wght2 (v), 2 = wghttemp (T), 3 = wght3 (w).
// With find, the face's cell indices are formed here as k_centroid_idx
// forms them (findCentroidIndices.f90:16-48) and component 0 stores them;
// otherwise they are read from pind.
__global__ __launch_bounds__(64) void k_mls_weights(Grid g, const double* __restrict__ bar, const unsigned char* __restrict__ gf, int nf,
                                                    int* __restrict__ pind, int find, PhiOut out, int* err) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nf * 4) return;
  int f = id >> 2, comp = id & 3;
  if (gf[f]) return;
  const double* pos = &bar[3 * f];
  int k1 = (int)floor(pos[2] * g.dx3) + 1, kst = fnint(pos[2] * g.dx3) + 1;
  int p[6];
  if (find) {
    p[0] = (int)floor(pos[0] * g.dx1) + 1; p[3] = fnint(pos[0] * g.dx1) + 1;
    p[1] = (int)floor(pos[1] * g.dx2) + 1; p[4] = fnint(pos[1] * g.dx2) + 1;
    p[2] = pmod(k1 - 1, g.n3) + 1; p[5] = pmod(kst - 1, g.n3) + 1;
    if (comp == 0)
      for (int q = 0; q < 6; ++q) pind[6 * f + q] = p[q];
  } else {
    for (int q = 0; q < 6; ++q) p[q] = pind[6 * f + q];
  }
  int i0, j0, k0, c[3];
  if (comp == 0) { i0 = p[3] - 1; j0 = p[1] - 1; k0 = k1 - 1; c[0] = 0; c[1] = 1; c[2] = 1; }
  else if (comp == 1) { i0 = p[0] - 1; j0 = p[4] - 1; k0 = k1 - 1; c[0] = 1; c[1] = 0; c[2] = 1; }
  else if (comp == 2) { i0 = p[0] - 1; j0 = p[1] - 1; k0 = k1 - 1; c[0] = 1; c[1] = 1; c[2] = 1; }
  else { i0 = p[0] - 1; j0 = p[1] - 1; k0 = kst - 1; c[0] = 1; c[1] = 1; c[2] = 0; }
  double* o = &out.q[comp][27 * (size_t)f];
  if (mls_phi(g, pos, i0, j0, k0, c, [&](int q, double v) { o[q] = v; })) raise_error(err, MELT_ERR_SINGULAR_MLS);
}

template <typename R>
struct ForceArgs {
  const R* fld[4];        // vx, vy, temp, vz
  const double* phi[4];
  const double* vel_tri;
  const double* cfac;
  double tmelt;
};

// forc1/forc2/forctemp/forc3: Um = sum q phi; spread cfac (V - Um) phi
// forc1/forc2/forctemp/forc3 (forc1.f90:28-65): Um = sum q phi, spread
// cfac (V - Um) phi.  32 lanes per (face, component): lane l < 27 loads stencil
// point l and its weight, every lane forms Um from the 27 products in the
// reference's point order (shuffles, so Um rounds exactly as the serial loop),
// and lane l adds its contribution to the exact 64.64 accumulators.
template <typename R>
__global__ __launch_bounds__(256) void k_mls_force(Grid g, ForceArgs<R> F, const double* __restrict__ bar, const unsigned char* __restrict__ gf,
                                                   int nf, const int* __restrict__ pind, const melt_window* __restrict__ win, Win3 ef,
                                                   unsigned long long* __restrict__ acc, size_t cells, int* err) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int item = gid >> 5, l = threadIdx.x & 31, lbase = threadIdx.x & 32;
  const int f = item >> 2, comp = item & 3;
  // every lane takes part in the shuffles; dead items / lanes carry zeros
  const bool live = f < nf && !gf[f];
  int i0 = 0, j0 = 0, kw = 0, ku = 0;
  if (live) {
    const int* p = &pind[6 * f];
    const double* pos = &bar[3 * f];
    int k1u = (int)floor(pos[2] * g.dx3) + 1, kstu = fnint(pos[2] * g.dx3) + 1;
    if (comp == 0) { i0 = p[3]; j0 = p[1]; kw = p[2]; ku = k1u; }
    else if (comp == 1) { i0 = p[0]; j0 = p[4]; kw = p[2]; ku = k1u; }
    else if (comp == 2) { i0 = p[0]; j0 = p[1]; kw = p[2]; ku = k1u; }
    else { i0 = p[0]; j0 = p[1]; kw = p[5]; ku = kstu; }
  }
  const bool act = live && l < 27;
  const int di = l % 3 - 1, dj = (l / 3) % 3 - 1, dk = l / 9 - 1;
  double ph = 0.0, prod = 0.0;
  if (act) {
    ph = F.phi[comp][27 * (size_t)f + l];
    prod = (double)F.fld[comp][widx(g, i0 + di, j0 + dj, kw + dk)] * ph;
  }
  double Um = 0.0;
#pragma unroll
  for (int q = 0; q < 27; ++q) Um = Um + __shfl(prod, lbase + q, 64);
  if (!act) return;
  double V = comp == 2 ? F.tmelt : F.vel_tri[3 * f + (comp == 0 ? 0 : comp == 1 ? 1 : 2)];
  double v = F.cfac[f] * (V - Um) * ph;
  long long sl = win_slot(win[0].org, ef.e, i0 + di, j0 + dj, ku + dk);
  if (sl < 0) { raise_error(err, MELT_ERR_WINDOW); return; }
  acc_add(&acc[((size_t)comp * cells + (size_t)sl) * 2], v);
}

// Drain the accumulators: mode 0 adds the forces to the fields (velocity
// without its mean, which the projection subtracts later) and sums the
// velocity forces exactly per block; mode 1 stores them into the for_* fields.
template <typename R>
__global__ __launch_bounds__(256) void k_force_apply(Grid g, const melt_window* __restrict__ win, Win3 ef, unsigned long long* __restrict__ acc,
                                                     size_t cells, R* __restrict__ vx, R* __restrict__ vy, R* __restrict__ t,
                                                     R* __restrict__ vz, int mode, unsigned long long* __restrict__ part) {
  __shared__ unsigned long long slo[3][256], shi[3][256];
  unsigned long long mlo[3] = {0, 0, 0}, mhi[3] = {0, 0, 0};
  R* dst[4] = {vx, vy, t, vz};
  for (size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x; s < cells; s += (size_t)gridDim.x * blockDim.x) {
    int a = (int)(s % ef.e[0]), b = (int)((s / ef.e[0]) % ef.e[1]), c = (int)(s / ((size_t)ef.e[0] * ef.e[1]));
    size_t cell = widx(g, win[0].org[0] + a, win[0].org[1] + b, win[0].org[2] + c);
    for (int comp = 0; comp < 4; ++comp) {
      unsigned long long* w = acc + ((size_t)comp * cells + s) * 2;
      unsigned long long lo = w[0], hi = w[1];
      if ((lo | hi) == 0) continue;
      w[0] = 0; w[1] = 0;
      double v = acc_value(lo, hi);
      if (mode == 0) dst[comp][cell] = dst[comp][cell] + v;
      else dst[comp][cell] = v;
      if (comp != 2) {
        int m = comp == 3 ? 2 : comp;
        acc_add_words(mlo[m], mhi[m], lo, hi);
      }
    }
  }
  if (!part) return;
  for (int m = 0; m < 3; ++m) { slo[m][threadIdx.x] = mlo[m]; shi[m][threadIdx.x] = mhi[m]; }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int m = 0; m < 3; ++m) {
        unsigned long long l1 = slo[m][threadIdx.x], l2 = slo[m][threadIdx.x + w];
        unsigned long long nl = l1 + l2;
        shi[m][threadIdx.x] = shi[m][threadIdx.x] + shi[m][threadIdx.x + w] + (nl < l1 ? 1ull : 0ull);
        slo[m][threadIdx.x] = nl;
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int m = 0; m < 3; ++m) { part[(blockIdx.x * 3 + m) * 2] = slo[m][0]; part[(blockIdx.x * 3 + m) * 2 + 1] = shi[m][0]; }
}

// Exact sum of the per-block partial force sums (integer adds: any order
// gives the same words), divided by the cell count.
__global__ __launch_bounds__(256) void k_fmean_final(const unsigned long long* __restrict__ part, int nb, double ncell,
                                                     double* __restrict__ mean) {
  __shared__ unsigned long long slo[3][256], shi[3][256];
  unsigned long long lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    for (int m = 0; m < 3; ++m) acc_add_words(lo[m], hi[m], part[(b * 3 + m) * 2], part[(b * 3 + m) * 2 + 1]);
  for (int m = 0; m < 3; ++m) { slo[m][threadIdx.x] = lo[m]; shi[m][threadIdx.x] = hi[m]; }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int m = 0; m < 3; ++m) {
        unsigned long long l = slo[m][threadIdx.x], h = shi[m][threadIdx.x];
        acc_add_words(l, h, slo[m][threadIdx.x + w], shi[m][threadIdx.x + w]);
        slo[m][threadIdx.x] = l; shi[m][threadIdx.x] = h;
      }
    __syncthreads();
  }
  if (threadIdx.x < 3) mean[threadIdx.x] = acc_value(slo[threadIdx.x][0], shi[threadIdx.x][0]) / ncell;
}

// velforce.f90:21-51 on full fields (per-routine entry point)
template <typename R>
__global__ __launch_bounds__(256) void k_sum3(const R* __restrict__ a, const R* __restrict__ b, const R* __restrict__ c,
                                              size_t n, double* __restrict__ part) {
  __shared__ double s[3][256];
  double x = 0.0, y = 0.0, z = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    x = x + a[i]; y = y + b[i]; z = z + c[i];
  }
  s[0][threadIdx.x] = x; s[1][threadIdx.x] = y; s[2][threadIdx.x] = z;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int m = 0; m < 3; ++m) s[m][threadIdx.x] = s[m][threadIdx.x] + s[m][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int m = 0; m < 3; ++m) part[blockIdx.x * 3 + m] = s[m][0];
}
__global__ void k_sum3_final(const double* __restrict__ part, int nb, double ncell, double* __restrict__ mean) {
  int m = threadIdx.x;
  if (m >= 3) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s = s + part[b * 3 + m];
  mean[m] = s / ncell;
}
template <typename R>
__global__ __launch_bounds__(256) void k_add_force(R* __restrict__ q, const R* __restrict__ f, const double* __restrict__ mean, int m,
                                                   size_t n) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  q[c] = mean ? q[c] + f[c] - mean[m] : q[c] + f[c];
}

template <typename R>
__global__ __launch_bounds__(256) void k_sub_mean(R* __restrict__ q, const double* __restrict__ mean, int m, size_t n) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  q[c] = q[c] - mean[m];
}

// ------------------------------------------------------------ probes
// kind 0/1: +/- temperature probe (findProbeIndices.f90 + wght_gradT);
// kind 2: pressure (wght_press); 3..6: grad p, grad u, grad v, grad w.
template <typename R>
struct ProbeArgs {
  const R* temp;
  const R* pr;
  const R* vx;
  const R* vy;
  const R* vz;
  const double* bar;
  const double* nor;
  const unsigned char* gf;
  int nf;
  double h, pec;
  int heat, loads;  // imelt == 1 (kinds 0, 1) / imlsstr == 1 (kinds 2..6)
};

// Probe point of face f on the fluid (kind != 1) or solid (kind 1) side:
// bar +- h n/|n| (mls_heatFlux.f90:33-60, mls_structLoads.f90:40-80).
template <typename R>
__device__ inline void probe_pos(const ProbeArgs<R>& P, int f, int kind, double pos[3]) {
  const double* nor = &P.nor[3 * f];
  const double* bar = &P.bar[3 * f];
  double s = fnorm2(nor, 3);
  double sg = kind == 1 ? -1.0 : 1.0;
  for (int d = 0; d < 3; ++d) pos[d] = sg > 0 ? bar[d] + P.h * nor[d] / s : bar[d] - P.h * nor[d] / s;
}

// Where gradient probe `kind` of face f samples: its position, the stencil
// origin (ic, jc), the z coordinates' grid indices kz and field planes ke
// (findProbeIndices.f90:9-76), the staggering of the field and the field.
template <typename R>
struct ProbeSite {
  double pos[3];
  int ic, jc, kz[3], ke[3], cc[3];
  int icent, jcent, kw;
  const R* fld;
};
template <typename R>
__device__ inline void probe_site(const Grid& g, const ProbeArgs<R>& P, int f, int kind, ProbeSite<R>& S) {
  probe_pos(P, f, kind, S.pos);
  const double* pos = S.pos;
  int icent = (int)floor(pos[0] * g.dx1) + 1, jcent = (int)floor(pos[1] * g.dx2) + 1;
  int istag = fnint(pos[0] * g.dx1) + 1, jstag = fnint(pos[1] * g.dx2) + 1;
  int k1 = (int)floor(pos[2] * g.dx3) + 1, kst = fnint(pos[2] * g.dx3) + 1;
  int kw = pmod(k1 - 1, g.n3) + 1, kws = pmod(kst - 1, g.n3) + 1;
  S.icent = icent; S.jcent = jcent; S.kw = kw;
  S.ic = icent; S.jc = jcent;
  S.cc[0] = S.cc[1] = S.cc[2] = 1;
  S.kz[0] = k1 - 1; S.kz[1] = k1; S.kz[2] = k1;
  S.ke[0] = kw - 1; S.ke[1] = kw; S.ke[2] = kw + 1;
  if (kind <= 1) {
    S.fld = P.temp;
  } else if (kind == 3) {
    S.fld = P.pr;
  } else if (kind == 4) {
    S.fld = P.vx; S.ic = istag; S.cc[0] = 0;
  } else if (kind == 5) {
    S.fld = P.vy; S.jc = jstag; S.cc[1] = 0;
  } else {
    S.fld = P.vz; S.cc[2] = 0;
    S.kz[0] = kst - 1; S.kz[1] = kst; S.kz[2] = kst;
    S.ke[0] = kws - 1; S.ke[1] = kws; S.ke[2] = kws + 1;
  }
}

__device__ __forceinline__ int grad_kind(int q) { return q < 2 ? q : q + 1; }  // 0, 1, 3, 4, 5, 6

// Slots of the stored shape-function derivatives.  The pressure-gradient
// probe (kind 3) samples the same point as the fluid-side temperature probe
// (kind 0, bar + h n) on the same cell-centred stencil, so with both heat flux
// and loads on it shares kind 0's slot: kinds 0, 1, 4, 5, 6.
__device__ __forceinline__ int coef_kind(int cs, int heat, int loads) {
  return heat && loads ? (cs < 2 ? cs : cs + 2) : heat ? cs : cs + 3;
}
__device__ __forceinline__ int coef_slot(int kind, int heat, int loads) {
  return heat && loads ? (kind < 2 ? kind : kind == 3 ? 0 : kind - 2) : heat ? kind : kind - 3;
}

__device__ inline double* probe_out(double* out, int f, int kind, int d) {
  double* o = &out[(size_t)PRB_N * f];
  return &o[(kind == 0 ? PRB_GT0 : kind == 1 ? PRB_GT1 : kind == 3 ? PRB_GP : kind == 4 ? PRB_GU : kind == 5 ? PRB_GV : PRB_GW) + d];
}

// Gradient probes: kinds 0/1 (temperature, fluid/solid side), 3 (pressure),
// 4/5/6 (u, v, w); item = kind slot x nf + face, 4 lanes per item.  With
// coef (the fused substep) only the geometry part runs: lane 1+d stores the
// 27 shape-function derivatives of component d at coef[q * 3 nitems + 3 item
// + d], and k_probes_apply later sums them against the field in the same
// order; without it the lanes sum against the field directly.  Either way
// the values round as in the single-thread formulation.
template <typename R>
__global__ __launch_bounds__(256) void k_probes_grad(Grid g, ProbeArgs<R> P, double* __restrict__ out, int* __restrict__ pind, int klo,
                                                     int nkinds, double* __restrict__ coef, int* err) {
  __shared__ double shA[64][16];
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  const int item = id >> 2, lane = id & 3, slot = threadIdx.x >> 2;
  const int nitem = P.nf * nkinds;   // with coef: items over the coefficient slots
  const bool inrange = item < nitem;
  const int f = inrange ? item % P.nf : 0;
  const int q = inrange ? item / P.nf : 0;
  const int kind = coef ? coef_kind(q, P.heat, P.loads) : grad_kind(klo + q);
  const bool active = inrange && !P.gf[f];
  ProbeSite<R> S;
  S.pos[0] = S.pos[1] = S.pos[2] = 0.0;
  S.ic = S.jc = 1;
  for (int q = 0; q < 3; ++q) { S.kz[q] = 1; S.ke[q] = 1; S.cc[q] = 1; }
  S.fld = P.temp;
  if (active) {
    probe_site(g, P, f, kind, S);
    if (!coef && kind <= 1 && lane == 0) { int* pi = &pind[6 * f + 3 * kind]; pi[0] = S.icent; pi[1] = S.jcent; pi[2] = S.kw; }
  }
  double gd = 0.0;
  const size_t cs = (size_t)3 * nitem;
  double* cf = coef && lane > 0 ? coef + (size_t)3 * item + (lane - 1) : nullptr;
  int info = mls_grad_lanes(g, S.fld, S.pos, S.ic, S.jc, S.kz, S.ke, S.cc, lane, active, shA[slot], &gd, cf, cs);
  if (!active || lane == 0) return;
  if (info) { raise_error(err, MELT_ERR_SINGULAR_MLS); return; }
  if (!coef) *probe_out(out, f, kind, lane - 1) = gd;
}

// The field part of the gradient probes of the fused substep: one thread per
// (item, component) sums the stored shape-function derivatives against the
// field values in the reference's point order (and records the heat-flux
// probes' cell indices, as the probes do).
template <typename R>
__global__ __launch_bounds__(256) void k_probes_apply(Grid g, ProbeArgs<R> P, const double* __restrict__ coef, double* __restrict__ out,
                                                      int* __restrict__ pind, int klo, int nkinds, int ncoef) {
  const int nitem = P.nf * nkinds;
  const int id = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (id >= 3 * nitem) return;
  const int item = id / 3, d = id - 3 * item;
  const int f = item % P.nf, kind = grad_kind(klo + item / P.nf);
  if (P.gf[f]) return;
  ProbeSite<R> S;
  probe_site(g, P, f, kind, S);
  if (kind <= 1 && d == 0) { int* pi = &pind[6 * f + 3 * kind]; pi[0] = S.icent; pi[1] = S.jcent; pi[2] = S.kw; }
  const size_t cs = (size_t)3 * P.nf * ncoef;
  const double* c = coef + (size_t)3 * ((size_t)coef_slot(kind, P.heat, P.loads) * P.nf + f) + d;
  double s = 0.0;
#pragma unroll 1
  for (int nk = 0; nk < 3; ++nk) {
    const int kk = nk == 0 ? S.ke[0] : nk == 1 ? S.ke[1] : S.ke[2];   // registers, not scratch
#pragma unroll
    for (int jq = 0; jq < 3; ++jq)
#pragma unroll
      for (int iq = 0; iq < 3; ++iq) {
        double acc = c[(size_t)(9 * nk + 3 * jq + iq) * cs];
        double val = S.fld[widx(g, S.ic - 1 + iq, S.jc - 1 + jq, kk)];
        s = s + acc * val;
      }
  }
  *probe_out(out, f, kind, d) = s;
}

// Pressure interpolation probe (kind 2) and the heat-flux dot products of
// kinds 0/1 (q = n . grad T / Pe), one thread per face.
template <typename R>
__global__ __launch_bounds__(64) void k_probes_point(Grid g, ProbeArgs<R> P, double* __restrict__ out, int* err) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= P.nf || P.gf[f]) return;
  double* o = &out[(size_t)PRB_N * f];
  if (P.heat) {
    const double* nor = &P.nor[3 * f];
    for (int kind = 0; kind < 2; ++kind) {
      const double* gr = &o[kind == 0 ? PRB_GT0 : PRB_GT1];
      double q = nor[0] * gr[0] + nor[1] * gr[1] + nor[2] * gr[2];
      o[kind == 0 ? PRB_QO : PRB_QI] = q / P.pec;
    }
  }
  if (!P.loads) return;
  double pos[3];
  probe_pos(P, f, 2, pos);
  int icent = (int)floor(pos[0] * g.dx1) + 1, jcent = (int)floor(pos[1] * g.dx2) + 1;
  int k1 = (int)floor(pos[2] * g.dx3) + 1;
  int kw = pmod(k1 - 1, g.n3) + 1;
  const int cc[3] = {1, 1, 1};
  double pp = 0.0;
  const R* pr = P.pr;
  int info = mls_phi(g, pos, icent - 1, jcent - 1, k1 - 1, cc, [&](int q, double phi) {
    int kk = q / 9, j = jcent - 1 + (q / 3) % 3, i = icent - 1 + q % 3;
    pp = pp + pr[widx(g, i, j, kw - 1 + kk)] * phi;
  });
  if (info) { raise_error(err, MELT_ERR_SINGULAR_MLS); return; }
  o[PRB_P] = pp;
}

inline unsigned nb(size_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

MeshArgs mesh_args(melt_ctx* ctx) {
  MeshArgs M;
  M.xyzv = ctx->d_lag[MELT_L_XYZV]; M.bar = ctx->d_lag[MELT_L_TRI_BAR]; M.nor = ctx->d_lag[MELT_L_TRI_NOR];
  M.vof = ctx->d_vof; M.gf = ctx->d_gf; M.nf = ctx->nf;
  return M;
}

}  // namespace

// ----------------------------------------------------------------- host
// sphereTagging (tagType 0, sphereTagging.f90:5-549): the level set of
// Kempe & Froehlich (2012) for a rigid sphere of radius rad_p.  A thread owns
// one (i, j, kk) of the bounding box of get_bbox_inds (tagCells.f90:31-62,
// lim from k_vbox, all three axes scaled by dx1 as there), unwrapped in x and
// y; kk is folded into the domain with the centre shifted by zlen
// (get_periodic_indices, tagCells.f90:68-84).  Each of VOFp, VOFx, VOFy, VOFz
// is 1 - alpha over the 8 corners of its control volume (level_setc2, 12, 22,
// 32), alpha = sum (-phi) H(-phi) / sum |phi|, phi the distance to the centre
// minus rad_p (geom.f90:56-80), corners in the reference's i, j, k order.
__device__ inline double sphere_phi(double x, double y, double z, const double c[3], double rad) {
  const double dx = x - c[0], dy = y - c[1], dz = z - c[2];
  return sqrt(dx * dx + dy * dy + dz * dz) - rad;
}

template <int XS, int YS, int ZS>   // per axis: 0 corners at the faces (c, c+1), 1 at the centres (c-1, c)
__device__ inline double sphere_alpha(const Grid& g, int ic, int jc, int kc, const double c[3], double rad) {
  double alpha = 0.0, tot = 0.0;
  for (int a = 0; a < 2; ++a) {
    const int i = XS ? ic - 1 + a : ic + a;
    const double x = XS ? gxm(g, i) : gxc(g, i);
    for (int b = 0; b < 2; ++b) {
      const int j = YS ? jc - 1 + b : jc + b;
      const double y = YS ? gym(g, j) : gyc(g, j);
      for (int d = 0; d < 2; ++d) {
        const int k = ZS ? kc - 1 + d : kc + d;
        const double z = ZS ? gzm(g, k) : gzc(g, k);
        const double phi = sphere_phi(x, y, z, c, rad);
        alpha = alpha + (-phi) * (-phi >= 0.0 ? 1.0 : 0.0);
        tot = tot + fabs(phi);
      }
    }
  }
  return alpha / tot;
}

template <typename R>
__global__ __launch_bounds__(256) void k_sphere_tag(Grid g, const double* __restrict__ lim, const melt_body* __restrict__ body, double rad,
                                                    R* __restrict__ vofp, R* __restrict__ vofx, R* __restrict__ vofy, R* __restrict__ vofz,
                                                    int* err) {
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = (int)floor(lim[d] * g.dx1) + 1 - 2;
    hi[d] = (int)floor(lim[3 + d] * g.dx1) + 1 + 2;
  }
  const int ext[3] = {hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0 && (ext[0] > g.n1 || ext[1] > g.n2 || ext[2] > g.n3)) raise_error(err, MELT_ERR_WINDOW);  // a cell would be visited twice
  if (t >= (size_t)g.n1 * g.n2 * g.n3) return;
  const int di = (int)(t % g.n1), dj = (int)((t / g.n1) % g.n2), dk = (int)(t / ((size_t)g.n1 * g.n2));
  if (di >= ext[0] || dj >= ext[1] || dk >= ext[2]) return;
  const int i = lo[0] + di, j = lo[1] + dj;
  int k = lo[2] + dk;
  double c[3] = {body->pos_cm[0], body->pos_cm[1], body->pos_cm[2]};
  if (k >= g.n3 + 1) { k = k - g.n3; c[2] = c[2] - g.zlen; }
  if (k < 1) { k = k + g.n3; c[2] = c[2] + g.zlen; }
  if (k < 1 || k > g.n3) return;
  const size_t o = (size_t)pmod(i - 1, g.n1) + (size_t)g.n1 * ((size_t)pmod(j - 1, g.n2) + (size_t)g.n2 * (k - 1));
  vofp[o] = (R)(1.0 - sphere_alpha<0, 0, 0>(g, i, j, k, c, rad));
  vofx[o] = (R)(1.0 - sphere_alpha<1, 0, 0>(g, i, j, k, c, rad));
  vofy[o] = (R)(1.0 - sphere_alpha<0, 1, 0>(g, i, j, k, c, rad));
  vofz[o] = (R)(1.0 - sphere_alpha<0, 0, 1>(g, i, j, k, c, rad));
}

template <typename R>
static int sphere_tag_t(melt_ctx* ctx) {
  const Grid g = make_grid(ctx);
  Win3 ef{{ctx->fwin[0], ctx->fwin[1], ctx->fwin[2]}}, et{{ctx->twin[0], ctx->twin[1], ctx->twin[2]}};
  hipLaunchKernelGGL(k_window_box, dim3(nb(ctx->nf, 256)), dim3(256), 0, ctx->stream, g, ctx->d_lag[MELT_L_TRI_BAR], ctx->d_gf, ctx->nf,
                     ctx->d_win);
  hipLaunchKernelGGL(k_window_final, dim3(1), dim3(1), 0, ctx->stream, ctx->d_win, ef, et, ctx->force_window_from_tag ? 3 : 2,
                     ctx->d_err);
  hipLaunchKernelGGL(k_vbox, dim3(1), dim3(256), 0, ctx->stream, ctx->d_lag[MELT_L_XYZV], ctx->d_gv, ctx->nv, ctx->d_sbox);
  for (int id = MELT_VOFP; id <= MELT_VOFZ; ++id)
    hipLaunchKernelGGL(k_fill<R>, dim3(1024), dim3(256), 0, ctx->stream, fld<R>(ctx, id), 1.0, ctx->nc);
  hipLaunchKernelGGL(k_sphere_tag<R>, dim3(nb(ctx->nc, 256)), dim3(256), 0, ctx->stream, g, ctx->d_sbox, ctx->d_body, ctx->rad_p,
                     fld<R>(ctx, MELT_VOFP), fld<R>(ctx, MELT_VOFX), fld<R>(ctx, MELT_VOFY), fld<R>(ctx, MELT_VOFZ), ctx->d_err);
  MELT_HIP(hipGetLastError());
  if (ctx->h_body.initial_tag) {   // no ray tagging in this mode: the flag only gates graph capture
    ctx->h_body.initial_tag = 0;
    MELT_HIP(hipMemcpyAsync(&ctx->d_body->initial_tag, &ctx->h_body.initial_tag, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    MELT_HIP(hipStreamSynchronize(ctx->stream));
  }
  return 0;
}

template <typename R>
static int ray_tag_init_t(melt_ctx* ctx) {
  Grid g = make_grid(ctx);
  hipLaunchKernelGGL(k_vbox, dim3(1), dim3(256), 0, ctx->stream, ctx->d_lag[MELT_L_XYZV], ctx->d_gv, ctx->nv, ctx->d_sbox);
  MELT_HIP(hipGetLastError());
  double lim[6];
  MELT_HIP(hipMemcpyAsync(lim, ctx->d_sbox, sizeof lim, hipMemcpyDeviceToHost, ctx->stream));
  MELT_HIP(hipStreamSynchronize(ctx->stream));
  Box bx;  // get_bbox_inds tagCells.f90:31-62 (the reference scales all three axes by dx1)
  for (int d = 0; d < 3; ++d) {
    bx.lo[d] = (int)floor(lim[d] * ctx->dx1) + 1 - 2;
    bx.hi[d] = (int)floor(lim[3 + d] * ctx->dx1) + 1 + 2;
  }
  for (int id = MELT_VOFP; id <= MELT_VOFZ; ++id)
    hipLaunchKernelGGL(k_fill<R>, dim3(1024), dim3(256), 0, ctx->stream, fld<R>(ctx, id), 1.0, ctx->nc);
  MELT_HIP(hipGetLastError());
  int js = bx.hi[1] - bx.lo[1] + 1, ks = bx.hi[2] - bx.lo[2] + 1;
  // BUFF_SIZE: n1m as the reference sets it (pencil_rayTag.f90:84) unless the
  // caller sized it (tuning key pbuf); tags do not depend on it below overflow
  int buff = ctx->tune[TUNE_PBUF] > 0 ? ctx->tune[TUNE_PBUF] : ctx->n1;
  size_t np = (size_t)js * ks;
  if (np * buff > ctx->plist_cap) {
    if (ctx->d_plist) hipFree(ctx->d_plist);
    if (ctx->d_pcount) hipFree(ctx->d_pcount);
    MELT_HIP(hipMalloc(&ctx->d_plist, sizeof(int) * np * buff));
    MELT_HIP(hipMalloc(&ctx->d_pcount, sizeof(int) * np));
    ctx->plist_cap = np * buff;
  }
  MELT_HIP(hipMemsetAsync(ctx->d_pcount, 0, sizeof(int) * np, ctx->stream));
  MeshArgs M = mesh_args(ctx);
  hipLaunchKernelGGL(k_pencil_build, dim3(nb(np * ctx->nf, 256)), dim3(256), 0, ctx->stream, g, M, bx, ctx->d_pcount, ctx->d_plist, buff,
                     ctx->d_err);
  MELT_HIP(hipGetLastError());
  size_t cov = 1;
  int nn[3] = {ctx->n1, ctx->n2, ctx->n3};
  for (int d = 0; d < 3; ++d) cov *= (size_t)std::min(bx.hi[d] - bx.lo[d] + 1, nn[d]);
  hipLaunchKernelGGL(k_pencil_tag<R>, dim3(nb(cov, 256)), dim3(256), 0, ctx->stream, g, M, bx, ctx->d_pcount, ctx->d_plist, buff,
                     fld<R>(ctx, MELT_VOFP), ctx->cfg.xlen, ctx->d_err);
  MELT_HIP(hipGetLastError());
  return 0;
}

template <typename R>
static int ibm_tag_t(melt_ctx* ctx) {  // tagCells.f90:4-27 (tagType 1) + interp_vof when not fused
  if (!ctx->has_geom) return melt_fail(ctx, MELT_ERR_NO_GEOMETRY, "tagCells", "no geometry");
  if (ctx->cfg.tagtype == 0) return sphere_tag_t<R>(ctx);
  Grid g = make_grid(ctx);
  if (ctx->h_body.initial_tag) {
    int rc = ray_tag_init_t<R>(ctx);
    if (rc) return rc;
    ctx->h_body.initial_tag = 0;
    MELT_HIP(hipMemcpyAsync(&ctx->d_body->initial_tag, &ctx->h_body.initial_tag, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  }
  Win3 ef{{ctx->fwin[0], ctx->fwin[1], ctx->fwin[2]}}, et{{ctx->twin[0], ctx->twin[1], ctx->twin[2]}};
  // face extent, reset of the previous window, cleared keys: one launch
  const int nbox = (int)nb(ctx->nf, 256);
  hipLaunchKernelGGL(k_tag_prepare<R>, dim3(nbox + 256), dim3(256), 0, ctx->stream, g, ctx->d_lag[MELT_L_TRI_BAR], ctx->d_gf, ctx->nf, nbox,
                     ctx->d_win, fld<R>(ctx, MELT_SDF), et, ctx->d_tkey, ctx->d_tface, ctx->twin_cells);
  // the fused substep places the forcing window from the same extent (the
  // faces do not move between tagging and forcing)
  hipLaunchKernelGGL(k_window_final, dim3(1), dim3(1), 0, ctx->stream, ctx->d_win, ef, et, ctx->force_window_from_tag ? 3 : 2,
                     ctx->d_err);
  MeshArgs M = mesh_args(ctx);
  unsigned blocks = nb((size_t)ctx->nf * 125, 256);
  for (int mode = 0; mode < 3; ++mode)
    hipLaunchKernelGGL(k_tag_pass<R>, dim3(blocks), dim3(256), 0, ctx->stream, g, M, ctx->d_win, et, ctx->d_tkey, ctx->d_tface, ctx->d_tdist,
                       fld<R>(ctx, MELT_SDF), fld<R>(ctx, MELT_VOFP), mode, ctx->pi, ctx->d_err);
  MELT_HIP(hipGetLastError());
  return 0;
}

int ibm_find_centroid_indices(melt_ctx* ctx) {
  hipLaunchKernelGGL(k_centroid_idx, dim3(nb(ctx->nf, 256)), dim3(256), 0, ctx->stream, make_grid(ctx), ctx->d_lag[MELT_L_TRI_BAR], ctx->d_gf,
                     ctx->nf, ctx->d_pind);
  MELT_HIP(hipGetLastError());
  return 0;
}

int ibm_weights(melt_ctx* ctx, bool find) {
  PhiOut o{{ctx->d_lag[MELT_L_PTXAB_Q1], ctx->d_lag[MELT_L_PTXAB_Q2], ctx->d_lag[MELT_L_PTXAB_TEMP], ctx->d_lag[MELT_L_PTXAB_Q3]}};
  hipLaunchKernelGGL(k_mls_weights, dim3(nb((size_t)ctx->nf * 4, 64)), dim3(64), 0, ctx->stream, make_grid(ctx), ctx->d_lag[MELT_L_TRI_BAR],
                     ctx->d_gf, ctx->nf, ctx->d_pind, find ? 1 : 0, o, ctx->d_err);
  MELT_HIP(hipGetLastError());
  return 0;
}

// Interpolation + spreading into the exact accumulators (mlsForce.f90, forc*).
template <typename R>
static int ibm_force_scatter_t(melt_ctx* ctx) {
  Grid g = make_grid(ctx);
  Win3 ef{{ctx->fwin[0], ctx->fwin[1], ctx->fwin[2]}}, et{{ctx->twin[0], ctx->twin[1], ctx->twin[2]}};
  if (!ctx->force_window_from_tag) {
    hipLaunchKernelGGL(k_window_box, dim3(nb(ctx->nf, 256)), dim3(256), 0, ctx->stream, g, ctx->d_lag[MELT_L_TRI_BAR], ctx->d_gf, ctx->nf,
                       ctx->d_win);
    hipLaunchKernelGGL(k_window_final, dim3(1), dim3(1), 0, ctx->stream, ctx->d_win, ef, et, 1, ctx->d_err);
  }
  ForceArgs<R> F;
  F.fld[0] = fld<R>(ctx, MELT_VX); F.fld[1] = fld<R>(ctx, MELT_VY); F.fld[2] = fld<R>(ctx, MELT_TEMP); F.fld[3] = fld<R>(ctx, MELT_VZ);
  F.phi[0] = ctx->d_lag[MELT_L_PTXAB_Q1]; F.phi[1] = ctx->d_lag[MELT_L_PTXAB_Q2];
  F.phi[2] = ctx->d_lag[MELT_L_PTXAB_TEMP]; F.phi[3] = ctx->d_lag[MELT_L_PTXAB_Q3];
  F.vel_tri = ctx->d_lag[MELT_L_VEL_TRI]; F.cfac = ctx->d_lag[MELT_L_CFAC]; F.tmelt = ctx->cfg.tmelt;
  hipLaunchKernelGGL(k_mls_force<R>, dim3(nb((size_t)ctx->nf * 4 * 32, 256)), dim3(256), 0, ctx->stream, g, F, ctx->d_lag[MELT_L_TRI_BAR], ctx->d_gf,
                     ctx->nf, ctx->d_pind, ctx->d_win, ef, ctx->d_acc, ctx->fwin_cells, ctx->d_err);
  MELT_HIP(hipGetLastError());
  return 0;
}

// Drain the accumulators.  apply_now = 1 (fused substep): forces go straight
// into u, v, w, T and the exact mean of the velocity forces lands in d_fmean
// for the projection (velforce.f90:21-51 with the uniform shift deferred);
// 0: the folded for_* fields are written (per-routine API, mlsForce output).
template <typename R>
static int ibm_force_drain_t(melt_ctx* ctx, int apply_now) {
  Grid g = make_grid(ctx);
  Win3 ef{{ctx->fwin[0], ctx->fwin[1], ctx->fwin[2]}};
  unsigned blocks = (unsigned)std::min<size_t>(512, nb(ctx->fwin_cells, 256));
  unsigned long long* part = reinterpret_cast<unsigned long long*>(ctx->d_red + 64);
  if (apply_now) {
    hipLaunchKernelGGL(k_force_apply<R>, dim3(blocks), dim3(256), 0, ctx->stream, g, ctx->d_win, ef, ctx->d_acc, ctx->fwin_cells, fld<R>(ctx, MELT_VX),
                       fld<R>(ctx, MELT_VY), fld<R>(ctx, MELT_TEMP), fld<R>(ctx, MELT_VZ), 0, part);
    hipLaunchKernelGGL(k_fmean_final, dim3(1), dim3(256), 0, ctx->stream, part, (int)blocks, (double)((long long)ctx->n1 * ctx->n2 * ctx->n3),
                       ctx->d_fmean);
  } else {
    for (int id = MELT_FOR_XC; id <= MELT_FOR_TEMP; ++id)
      hipLaunchKernelGGL(k_fill<R>, dim3(1024), dim3(256), 0, ctx->stream, fld<R>(ctx, id), 0.0, ctx->nc);
    hipLaunchKernelGGL(k_force_apply<R>, dim3(blocks), dim3(256), 0, ctx->stream, g, ctx->d_win, ef, ctx->d_acc, ctx->fwin_cells,
                       fld<R>(ctx, MELT_FOR_XC), fld<R>(ctx, MELT_FOR_YC), fld<R>(ctx, MELT_FOR_TEMP), fld<R>(ctx, MELT_FOR_ZC), 1, nullptr);
  }
  MELT_HIP(hipGetLastError());
  return 0;
}

// u -= fbar over the whole field (the shift the fused substep defers to the
// projection), for callers that stop after particle.
template <typename R>
static int ibm_sub_mean_t(melt_ctx* ctx) {
  const int q[3] = {MELT_VX, MELT_VY, MELT_VZ};
  for (int m = 0; m < 3; ++m)
    hipLaunchKernelGGL(k_sub_mean<R>, dim3(nb(ctx->nc, 256)), dim3(256), 0, ctx->stream, fld<R>(ctx, q[m]), ctx->d_fmean, m, ctx->nc);
  MELT_HIP(hipGetLastError());
  return 0;
}

template <typename R>
static int ibm_velforce_t(melt_ctx* ctx, int /*defer*/) {  // velforce.f90 on full for_* fields
  unsigned blocks = (unsigned)std::min<size_t>(1024, nb(ctx->nc, 256));
  double* part = ctx->d_red + 64;
  hipLaunchKernelGGL(k_sum3<R>, dim3(blocks), dim3(256), 0, ctx->stream, fld<R>(ctx, MELT_FOR_XC), fld<R>(ctx, MELT_FOR_YC), fld<R>(ctx, MELT_FOR_ZC), ctx->nc,
                     part);
  hipLaunchKernelGGL(k_sum3_final, dim3(1), dim3(64), 0, ctx->stream, part, (int)blocks, (double)((long long)ctx->n1 * ctx->n2 * ctx->n3),
                     ctx->d_fmean);
  const int q[3] = {MELT_VX, MELT_VY, MELT_VZ}, fo[3] = {MELT_FOR_XC, MELT_FOR_YC, MELT_FOR_ZC};
  for (int m = 0; m < 3; ++m)
    hipLaunchKernelGGL(k_add_force<R>, dim3(nb(ctx->nc, 256)), dim3(256), 0, ctx->stream, fld<R>(ctx, q[m]), fld<R>(ctx, fo[m]), ctx->d_fmean, m, ctx->nc);
  MELT_HIP(hipGetLastError());
  return 0;
}

template <typename R>
static int ibm_tempforce_t(melt_ctx* ctx) {  // tempforce.f90
  hipLaunchKernelGGL(k_add_force<R>, dim3(nb(ctx->nc, 256)), dim3(256), 0, ctx->stream, fld<R>(ctx, MELT_TEMP), fld<R>(ctx, MELT_FOR_TEMP), nullptr, 0,
                     ctx->nc);
  MELT_HIP(hipGetLastError());
  return 0;
}

template <typename R>
static ProbeArgs<R> probe_args(melt_ctx* ctx) {
  ProbeArgs<R> P;
  P.temp = fld<R>(ctx, MELT_TEMP); P.pr = fld<R>(ctx, MELT_PR); P.vx = fld<R>(ctx, MELT_VX); P.vy = fld<R>(ctx, MELT_VY); P.vz = fld<R>(ctx, MELT_VZ);
  P.bar = ctx->d_lag[MELT_L_TRI_BAR]; P.nor = ctx->d_lag[MELT_L_TRI_NOR]; P.gf = ctx->d_gf; P.nf = ctx->nf;
  P.h = ctx->h_body.h_eulerian; P.pec = ctx->cfg.ren * ctx->cfg.prandtl;
  P.heat = ctx->cfg.imelt == 1;
  P.loads = ctx->cfg.imlsstr == 1;
  return P;
}

template <typename R>
static int probe_coef_slots(const ProbeArgs<R>& P) { return P.heat && P.loads ? 5 : P.heat ? 2 : P.loads ? 4 : 0; }

// The geometry part of the gradient probes (shape-function derivatives into
// d_pcoef): reads only the faces and the grid, so the fused substep runs it
// beside the explicit pass.
template <typename R>
static int ibm_probe_coefs_t(melt_ctx* ctx) {
  ProbeArgs<R> P = probe_args<R>(ctx);
  int klo = P.heat ? 0 : 2, khi = P.loads ? 6 : 2;
  const int ncoef = probe_coef_slots(P);
  if (khi > klo) {
    size_t threads = (size_t)ctx->nf * ncoef * 4;
    hipLaunchKernelGGL(k_probes_grad<R>, dim3(nb(threads, 256)), dim3(256), 0, ctx->stream, make_grid(ctx), P, ctx->d_probe, ctx->d_pind,
                       klo, ncoef, ctx->d_pcoef, ctx->d_err);
  }
  MELT_HIP(hipGetLastError());
  return 0;
}

template <typename R>
static int ibm_probes_t(melt_ctx* ctx, int* pind_out, bool coefs_ready) {
  ProbeArgs<R> P = probe_args<R>(ctx);
  int klo = P.heat ? 0 : 2, khi = P.loads ? 6 : 2;
  int* pind = pind_out ? pind_out : ctx->d_pind;
  if (khi > klo && coefs_ready) {
    size_t threads = (size_t)ctx->nf * (khi - klo) * 3;
    hipLaunchKernelGGL(k_probes_apply<R>, dim3(nb(threads, 256)), dim3(256), 0, ctx->stream, make_grid(ctx), P, ctx->d_pcoef, ctx->d_probe,
                       pind, klo, khi - klo, probe_coef_slots(P));
  } else if (khi > klo) {
    size_t threads = (size_t)ctx->nf * (khi - klo) * 4;
    hipLaunchKernelGGL(k_probes_grad<R>, dim3(nb(threads, 256)), dim3(256), 0, ctx->stream, make_grid(ctx), P, ctx->d_probe, pind, klo,
                       khi - klo, nullptr, ctx->d_err);
  }
  if (P.heat || P.loads)
    hipLaunchKernelGGL(k_probes_point<R>, dim3(nb(ctx->nf, 64)), dim3(64), 0, ctx->stream, make_grid(ctx), P, ctx->d_probe, ctx->d_err);
  MELT_HIP(hipGetLastError());
  return 0;
}

int ibm_window_reset(melt_ctx* ctx) {
  hipLaunchKernelGGL(k_window_reset, dim3(1), dim3(1), 0, ctx->stream, ctx->d_win, 2);
  MELT_HIP(hipGetLastError());
  return 0;
}

// ------------------------------------------------- precision dispatch
int ibm_tag(melt_ctx* ctx) { return ctx->fp32 ? ibm_tag_t<float>(ctx) : ibm_tag_t<double>(ctx); }
int ibm_force_scatter(melt_ctx* ctx) { return ctx->fp32 ? ibm_force_scatter_t<float>(ctx) : ibm_force_scatter_t<double>(ctx); }
int ibm_force_drain(melt_ctx* ctx, int apply_now) { return ctx->fp32 ? ibm_force_drain_t<float>(ctx, apply_now) : ibm_force_drain_t<double>(ctx, apply_now); }
int ibm_sub_mean(melt_ctx* ctx) { return ctx->fp32 ? ibm_sub_mean_t<float>(ctx) : ibm_sub_mean_t<double>(ctx); }
int ibm_velforce(melt_ctx* ctx, int defer) { return ctx->fp32 ? ibm_velforce_t<float>(ctx, defer) : ibm_velforce_t<double>(ctx, defer); }
int ibm_tempforce(melt_ctx* ctx) { return ctx->fp32 ? ibm_tempforce_t<float>(ctx) : ibm_tempforce_t<double>(ctx); }
int ibm_probes(melt_ctx* ctx, int* pind_out, bool coefs_ready) {
  return ctx->fp32 ? ibm_probes_t<float>(ctx, pind_out, coefs_ready) : ibm_probes_t<double>(ctx, pind_out, coefs_ready);
}
int ibm_probe_coefs(melt_ctx* ctx) { return ctx->fp32 ? ibm_probe_coefs_t<float>(ctx) : ibm_probe_coefs_t<double>(ctx); }
