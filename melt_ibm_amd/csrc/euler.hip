// Eulerian part of the RK3 substep on gfx950: explicit advection/diffusion
// stencils (hdnl1-3, hdnlte, invtr1-3, invtrte right-hand sides), the
// Crank-Nicolson periodic line solves (solxi/solxj/solxk) as circulant parallel
// cyclic reduction in LDS, the divergence, and the projection (updvp, prcalc).
//
// All of it is HBM-bound (about 1 flop per byte), so the kernels are written for
// coalesced unit-stride access and to read each input field once per pass.
#include <algorithm>
#include <cstring>

#include "common.h"

namespace {

// One cell (0-based coordinates) with the element offsets to its wrapped
// neighbours.  Fields hold at most 1024^3 elements, so 32-bit offsets suffice.
struct Cell {
  int c, i, j, k;
  int im, ip, jm, jp, km, kp;
  __device__ __forceinline__ int at(int a, int b, int d) const {
    return c + (a < 0 ? im : a > 0 ? ip : 0) + (b < 0 ? jm : b > 0 ? jp : 0) + (d < 0 ? km : d > 0 ? kp : 0);
  }
};

__device__ __forceinline__ Cell cell_at(const Grid& g, int i, int j, int k) {
  Cell s;
  const int n12 = g.n1 * g.n2;
  s.i = i; s.j = j; s.k = k;
  s.c = i + g.n1 * j + n12 * k;
  s.im = i == 0 ? g.n1 - 1 : -1;
  s.ip = i == g.n1 - 1 ? 1 - g.n1 : 1;
  s.jm = j == 0 ? (g.n2 - 1) * g.n1 : -g.n1;
  s.jp = j == g.n2 - 1 ? (1 - g.n2) * g.n1 : g.n1;
  s.km = k == 0 ? (g.n3 - 1) * n12 : -n12;
  s.kp = k == g.n3 - 1 ? (1 - g.n3) * n12 : n12;
  return s;
}

__device__ __forceinline__ Cell make_cell(const Grid& g, int c) {
  int i = c % g.n1, r = c / g.n1;
  return cell_at(g, i, r % g.n2, r / g.n2);
}

// Field storage type R (double; float in the precision form).  The stencil
// arithmetic runs in R; geometry (rigid-body velocity at the cell faces) in
// double.  For R = double every expression is the reference's, in its order.
template <typename R>
struct EulerArgs {
  const R* __restrict__ vx;
  const R* __restrict__ vy;
  const R* __restrict__ vz;
  const R* __restrict__ temp;
  const R* __restrict__ pr;
  const R* __restrict__ vofp;
  const R* __restrict__ forcx;
  const R* __restrict__ forcy;
  const R* __restrict__ forcz;
  const melt_body* __restrict__ body;
  Grid g;
  double betagz, tliq;
};

// hdnl1.f90:32-74
template <typename R>
__device__ inline R h1_val(const EulerArgs<R>& A, const Cell& s, R vofx) {
  const R* vx = A.vx; const R* vy = A.vy; const R* vz = A.vz;
  R udx1 = R(A.g.dx1) * R(0.25), udx2 = R(A.g.dx2) * R(0.25), udx3 = R(A.g.dx3) * R(0.25);
  R h11 = ((vx[s.at(1, 0, 0)] + vx[s.at(0, 0, 0)]) * (vx[s.at(1, 0, 0)] + vx[s.at(0, 0, 0)]) -
           (vx[s.at(-1, 0, 0)] + vx[s.at(0, 0, 0)]) * (vx[s.at(-1, 0, 0)] + vx[s.at(0, 0, 0)])) * udx1;
  R h12 = ((vy[s.at(0, 1, 0)] + vy[s.at(-1, 1, 0)]) * (vx[s.at(0, 1, 0)] + vx[s.at(0, 0, 0)]) -
           (vy[s.at(0, 0, 0)] + vy[s.at(-1, 0, 0)]) * (vx[s.at(0, 0, 0)] + vx[s.at(0, -1, 0)])) * udx2;
  R h13 = ((vz[s.at(0, 0, 1)] + vz[s.at(-1, 0, 1)]) * (vx[s.at(0, 0, 1)] + vx[s.at(0, 0, 0)]) -
           (vz[s.at(0, 0, 0)] + vz[s.at(-1, 0, 0)]) * (vx[s.at(0, 0, 0)] + vx[s.at(0, 0, -1)])) * udx3;
  return -(h11 + h12 + h13) + vofx * A.forcx[s.at(0, 0, 0)] / R(A.g.xlen);
}

// hdnl2.f90:33-75
template <typename R>
__device__ inline R h2_val(const EulerArgs<R>& A, const Cell& s, R vofy) {
  const R* vx = A.vx; const R* vy = A.vy; const R* vz = A.vz;
  R udx1 = R(A.g.dx1) * R(0.25), udx2 = R(A.g.dx2) * R(0.25), udx3 = R(A.g.dx3) * R(0.25);
  R h21 = ((vy[s.at(1, 0, 0)] + vy[s.at(0, 0, 0)]) * (vx[s.at(1, 0, 0)] + vx[s.at(1, -1, 0)]) -
           (vy[s.at(0, 0, 0)] + vy[s.at(-1, 0, 0)]) * (vx[s.at(0, 0, 0)] + vx[s.at(0, -1, 0)])) * udx1;
  R h22 = ((vy[s.at(0, 1, 0)] + vy[s.at(0, 0, 0)]) * (vy[s.at(0, 1, 0)] + vy[s.at(0, 0, 0)]) -
           (vy[s.at(0, -1, 0)] + vy[s.at(0, 0, 0)]) * (vy[s.at(0, -1, 0)] + vy[s.at(0, 0, 0)])) * udx2;
  R h23 = ((vz[s.at(0, 0, 1)] + vz[s.at(0, -1, 1)]) * (vy[s.at(0, 0, 1)] + vy[s.at(0, 0, 0)]) -
           (vz[s.at(0, 0, 0)] + vz[s.at(0, -1, 0)]) * (vy[s.at(0, 0, 0)] + vy[s.at(0, 0, -1)])) * udx3;
  return -(h21 + h22 + h23) + vofy * A.forcy[s.at(0, 0, 0)] / R(A.g.ylen);
}

// hdnl3.f90:38-82
template <typename R>
__device__ inline R h3_val(const EulerArgs<R>& A, const Cell& s, R vofz) {
  const R* vx = A.vx; const R* vy = A.vy; const R* vz = A.vz;
  R udx1 = R(A.g.dx1) * R(0.25), udx2 = R(A.g.dx2) * R(0.25), udx3 = R(A.g.dx3) * R(0.25);
  R h31 = (((vx[s.at(1, 0, 0)] + vx[s.at(1, 0, -1)]) * (vz[s.at(1, 0, 0)] + vz[s.at(0, 0, 0)])) -
           ((vx[s.at(0, 0, 0)] + vx[s.at(0, 0, -1)]) * (vz[s.at(0, 0, 0)] + vz[s.at(-1, 0, 0)]))) * udx1;
  R h32 = (((vy[s.at(0, 1, 0)] + vy[s.at(0, 1, -1)]) * (vz[s.at(0, 1, 0)] + vz[s.at(0, 0, 0)])) -
           ((vy[s.at(0, 0, 0)] + vy[s.at(0, 0, -1)]) * (vz[s.at(0, 0, 0)] + vz[s.at(0, -1, 0)]))) * udx2;
  R h33 = (((vz[s.at(0, 0, 1)] + vz[s.at(0, 0, 0)]) * (vz[s.at(0, 0, 1)] + vz[s.at(0, 0, 0)])) -
           ((vz[s.at(0, 0, 0)] + vz[s.at(0, 0, -1)]) * (vz[s.at(0, 0, 0)] + vz[s.at(0, 0, -1)]))) * udx3;
  R T_interp = (A.temp[s.at(0, 0, 0)] + A.temp[s.at(0, 0, -1)]) * R(0.5);
  R fbz = R(A.betagz) * (T_interp - R(A.tliq));
  return (-(h31 + h32 + h33) + fbz) + vofz * A.forcz[s.at(0, 0, 0)] / R(A.g.zlen);
}

// hdnlte.f90:45-158 (Nparticle = 1)
template <typename R>
__device__ inline R ht_val(const EulerArgs<R>& A, const Cell& s, R vofp) {
  const R* vx = A.vx; const R* vy = A.vy; const R* vz = A.vz; const R* t = A.temp;
  const Grid& g = A.g;
  R udx1 = R(g.dx1) * R(0.5), udx2 = R(g.dx2) * R(0.5), udx3 = R(g.dx3) * R(0.5);
  R T = t[s.at(0, 0, 0)];
  R h31 = (vx[s.at(1, 0, 0)] * (t[s.at(1, 0, 0)] + T) - vx[s.at(0, 0, 0)] * (T + t[s.at(-1, 0, 0)])) * udx1;
  R h32 = (vy[s.at(0, 1, 0)] * (t[s.at(0, 1, 0)] + T) - vy[s.at(0, 0, 0)] * (T + t[s.at(0, -1, 0)])) * udx2;
  R h33 = (vz[s.at(0, 0, 1)] * (t[s.at(0, 0, 1)] + T) - vz[s.at(0, 0, 0)] * (T + t[s.at(0, 0, -1)])) * udx3;
  R ht = -(h31 + h32 + h33) * vofp;
  const double* pc = A.body->pos_cm;
  const double* vc = A.body->vel_cm;
  const double* om = A.body->omega_c;
  int ic = s.i + 1, jc = s.j + 1, kc = s.k + 1;
  double r0 = gxc(g, ic) - pc[0], r1 = gym(g, jc) - pc[1], r2 = gzm(g, kc) - pc[2];
  r0 = r0 - g.xlen * (double)fnint(r0 / g.xlen);
  r1 = r1 - g.ylen * (double)fnint(r1 / g.ylen);
  r2 = r2 - g.zlen * (double)fnint(r2 / g.zlen);
  double u_imh = vc[0] + om[1] * r2 - om[2] * r1;
  h31 = R(u_imh) * (t[s.at(1, 0, 0)] - t[s.at(-1, 0, 0)]) * udx1;
  r0 = gxm(g, ic) - pc[0]; r1 = gyc(g, jc) - pc[1]; r2 = gzm(g, kc) - pc[2];
  r0 = r0 - g.xlen * (double)fnint(r0 / g.xlen);
  r1 = r1 - g.ylen * (double)fnint(r1 / g.ylen);
  r2 = r2 - g.zlen * (double)fnint(r2 / g.zlen);
  double v_jmh = vc[1] + om[2] * r0 - om[0] * r2;
  h32 = R(v_jmh) * (t[s.at(0, 1, 0)] - t[s.at(0, -1, 0)]) * udx2;
  r0 = gxm(g, ic) - pc[0]; r1 = gym(g, jc) - pc[1]; r2 = gzc(g, kc) - pc[2];
  double w_kmh = vc[2] + om[0] * r1 - om[1] * r0;
  h33 = R(w_kmh) * (t[s.at(0, 0, 1)] - t[s.at(0, 0, -1)]) * udx3;
  return ht - (h31 + h32 + h33) * (R(1) - vofp);
}

template <typename R>
struct RhsCoef {
  R ga, ro, al, dt, alre, dx1q, dx2q, dx3q, dx1, dx2, dx3, pec;
};

// invtr1.f90:40-60
template <typename R>
__device__ inline R rhs_u(const RhsCoef<R>& c, const R* __restrict__ q, const R* __restrict__ pr, const Cell& s, R h, R ru) {
  R v = q[s.at(0, 0, 0)];
  R d11 = (q[s.at(1, 0, 0)] - R(2) * v + q[s.at(-1, 0, 0)]) * c.dx1q;
  R d22 = (q[s.at(0, 1, 0)] - R(2) * v + q[s.at(0, -1, 0)]) * c.dx2q;
  R d33 = (q[s.at(0, 0, 1)] - R(2) * v + q[s.at(0, 0, -1)]) * c.dx3q;
  R dc = d11 + d22 + d33;
  R dp = (pr[s.at(0, 0, 0)] - pr[s.at(-1, 0, 0)]) * (c.dx1 * c.al);
  return (c.ga * h + c.ro * ru + c.alre * dc - dp) * c.dt;
}
// invtr2.f90:37-57
template <typename R>
__device__ inline R rhs_v(const RhsCoef<R>& c, const R* __restrict__ q, const R* __restrict__ pr, const Cell& s, R h, R ru) {
  R v = q[s.at(0, 0, 0)];
  R d11 = (q[s.at(1, 0, 0)] - R(2) * v + q[s.at(-1, 0, 0)]) * c.dx1q;
  R d22 = (q[s.at(0, 1, 0)] - R(2) * v + q[s.at(0, -1, 0)]) * c.dx2q;
  R d33 = (q[s.at(0, 0, 1)] - R(2) * v + q[s.at(0, 0, -1)]) * c.dx3q;
  R dc = d11 + d22 + d33;
  R dp = (pr[s.at(0, 0, 0)] - pr[s.at(0, -1, 0)]) * (c.dx2 * c.al);
  return (c.ga * h + c.ro * ru + c.alre * dc - dp) * c.dt;
}
// invtr3.f90:35-56
template <typename R>
__device__ inline R rhs_w(const RhsCoef<R>& c, const R* __restrict__ q, const R* __restrict__ pr, const Cell& s, R h, R ru) {
  R v = q[s.at(0, 0, 0)];
  R d1 = (q[s.at(-1, 0, 0)] - R(2) * v + q[s.at(1, 0, 0)]) * c.dx1q;
  R d2 = (q[s.at(0, -1, 0)] - R(2) * v + q[s.at(0, 1, 0)]) * c.dx2q;
  R d3 = (q[s.at(0, 0, 1)] - R(2) * v + q[s.at(0, 0, -1)]) * c.dx3q;
  R dc = d2 + d3 + d1;
  R dp = (pr[s.at(0, 0, 0)] - pr[s.at(0, 0, -1)]) * (c.al * c.dx3);
  return (c.ga * h + c.ro * ru + c.alre * dc - dp) * c.dt;
}
// invtrte.f90:32-53
template <typename R>
__device__ inline R rhs_t(const RhsCoef<R>& c, const R* __restrict__ q, const Cell& s, R h, R ru) {
  R v = q[s.at(0, 0, 0)];
  R d1 = (q[s.at(-1, 0, 0)] - R(2) * v + q[s.at(1, 0, 0)]) * c.dx1q;
  R d2 = (q[s.at(0, -1, 0)] - R(2) * v + q[s.at(0, 1, 0)]) * c.dx2q;
  R d3 = (q[s.at(0, 0, 1)] - R(2) * v + q[s.at(0, 0, -1)]) * c.dx3q;
  R dc = d2 + d3 + d1;
  return c.dt * (c.ga * h + c.ro * ru) + (c.al / c.pec) * dc * c.dt;
}

template <typename R>
EulerArgs<R> make_args(melt_ctx* ctx) {
  EulerArgs<R> A;
  A.vx = fld<R>(ctx, MELT_VX); A.vy = fld<R>(ctx, MELT_VY); A.vz = fld<R>(ctx, MELT_VZ);
  A.temp = fld<R>(ctx, MELT_TEMP); A.pr = fld<R>(ctx, MELT_PR); A.vofp = fld<R>(ctx, MELT_VOFP);
  A.forcx = fld<R>(ctx, MELT_FORCX); A.forcy = fld<R>(ctx, MELT_FORCY); A.forcz = fld<R>(ctx, MELT_FORCZ);
  A.body = ctx->d_body;
  A.g = make_grid(ctx);
  A.betagz = ctx->cfg.betagz; A.tliq = ctx->cfg.tliq;
  return A;
}

template <typename R>
RhsCoef<R> make_coef(melt_ctx* ctx) {
  RhsCoef<R> c;
  c.ga = R(ctx->ga); c.ro = R(ctx->ro); c.al = R(ctx->al); c.dt = R(ctx->dt);
  c.alre = R(ctx->al / ctx->cfg.ren);
  c.dx1q = R(ctx->dx1q); c.dx2q = R(ctx->dx2q); c.dx3q = R(ctx->dx3q);
  c.dx1 = R(ctx->dx1); c.dx2 = R(ctx->dx2); c.dx3 = R(ctx->dx3);
  c.pec = R(ctx->cfg.ren * ctx->cfg.prandtl);
  return c;
}

// ------------------------------------------------------------- kernels
template <typename R>
__global__ __launch_bounds__(256) void k_hdnl(EulerArgs<R> A, R* __restrict__ out, int which, int nc) {
  unsigned b = xcd_block(blockIdx.x, gridDim.x);
  int c = (int)(b * blockDim.x + threadIdx.x);
  if (c >= nc) return;
  Cell s = make_cell(A.g, c);
  const R* vp = A.vofp;
  R p = vp[s.at(0, 0, 0)];
  R r;
  if (which == 0) r = h1_val(A, s, R(0.5) * (vp[s.at(-1, 0, 0)] + p));
  else if (which == 1) r = h2_val(A, s, R(0.5) * (vp[s.at(0, -1, 0)] + p));
  else if (which == 2) r = h3_val(A, s, R(0.5) * (vp[s.at(0, 0, -1)] + p));
  else r = ht_val(A, s, p);
  out[c] = r;
}

template <typename R>
__global__ __launch_bounds__(256) void k_invtr_rhs(Grid g, RhsCoef<R> cf, const R* __restrict__ q, const R* __restrict__ pr,
                                                   const R* __restrict__ h, R* __restrict__ ru, R* __restrict__ rhs, int which, int nc) {
  unsigned b = xcd_block(blockIdx.x, gridDim.x);
  int c = (int)(b * blockDim.x + threadIdx.x);
  if (c >= nc) return;
  Cell s = make_cell(g, c);
  R hv = h[c], rv = ru[c], r;
  if (which == 0) r = rhs_u(cf, q, pr, s, hv, rv);
  else if (which == 1) r = rhs_v(cf, q, pr, s, hv, rv);
  else if (which == 2) r = rhs_w(cf, q, pr, s, hv, rv);
  else r = rhs_t(cf, q, s, hv, rv);
  rhs[c] = r;
  ru[c] = hv;
}

// -------------------------------------------------- circulant PCR solver
// The systems of solxi/solxj/solxk are periodic tridiagonal with constant
// coefficients (diag 1, off-diagonals -G/(1+2G), rhs scaled by 1/(1+2G);
// solxi.f90:33-37): circulant.  Parallel cyclic reduction keeps a circulant
// circulant, so one scalar coefficient triple per step serves every row; the
// ratio a/b squares each step and the reduction stops once it is below 1e-17
// (after which x = d/b to round-off), or when the coupling offset reaches M/2
// (M = 2^p: closed 2x2 solve) or M/3 (M = 3*2^p: closed 3x3 circulant solve).
template <typename R>
struct SolveField {
  R* rhs;         // in/out (in only when qadd != null and store_rhs == 0)
  R* qadd;        // solxk: q += solution
  double betadx;
  int store_rhs;  // with qadd: also leave the solution in rhs (solxk.f90:40)
};
template <typename R>
struct SolveBatch {
  SolveField<R> f[4];
};

// Cyclic reduction of the W circulant systems of one LDS tile.  Element
// (row r, line c) sits at lds[XL ? c*ldx + r : r*W + c]; thread tid owns the
// elements idx = tid + q*nt, q < E (XL: r = idx % M, c = idx / M; else
// r = idx / W, c = idx % W).  d[] holds the owned, already scaled right-hand
// sides (also present in lds) and receives the solution.  The tile must be
// complete in lds and visible (barrier) on entry.
template <bool XL, int E>
__device__ __forceinline__ void pcr_core(double* lds, int M, int W, int ldx, double betadx, double* d) {
  const int tid = threadIdx.x, nt = blockDim.x;
  auto own_r = [&](int idx) { return XL ? idx % M : idx / W; };
  auto own_c = [&](int idx) { return XL ? idx / M : idx % W; };
  auto la = [&](int r, int c) { return XL ? c * ldx + r : r * W + c; };
  const double ackl_b = 1.0 / (1.0 + 2.0 * betadx);
  double a = -betadx * ackl_b, bdiag = 1.0;
  int p2 = M & (-M);          // largest power of two dividing M
  int m3 = M / p2;            // 1 or 3
  int o = 1;
  int last = (m3 == 1) ? M / 4 : p2 / 2;  // offsets 1..last are reduction steps
  bool done = false;
  while (o <= last) {
    if (fabs(a) < 1e-17 * fabs(bdiag)) { done = true; break; }
    double k = a / bdiag;
    double nd[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
      int idx = tid + q * nt;
      int r = own_r(idx), c = own_c(idx);
      int rm = r - o; if (rm < 0) rm += M;
      int rp = r + o; if (rp >= M) rp -= M;
      nd[q] = d[q] - k * (lds[la(rm, c)] + lds[la(rp, c)]);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < E; ++q) {
      int idx = tid + q * nt;
      d[q] = nd[q];
      lds[la(own_r(idx), own_c(idx))] = d[q];
    }
    __syncthreads();
    double bn = bdiag - 2.0 * a * k;
    a = -a * k;
    bdiag = bn;
    o *= 2;
  }
  if (!done && fabs(a) < 1e-17 * fabs(bdiag)) done = true;
  if (done) {
#pragma unroll
    for (int q = 0; q < E; ++q) d[q] = d[q] / bdiag;
    return;
  }
  double x[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    int idx = tid + q * nt;
    int r = own_r(idx), c = own_c(idx);
    if (m3 == 1) {
      int h = M / 2;
      int rh = r + h; if (rh >= M) rh -= M;
      double dh = lds[la(rh, c)];
      x[q] = (bdiag * d[q] - 2.0 * a * dh) / (bdiag * bdiag - 4.0 * a * a);
    } else {
      int h = M / 3;
      int r1 = r + h; if (r1 >= M) r1 -= M;
      int r2 = r1 + h; if (r2 >= M) r2 -= M;
      double D = d[q] + lds[la(r1, c)] + lds[la(r2, c)];
      x[q] = (d[q] - a * (D / (bdiag + 2.0 * a))) / (bdiag - a);
    }
  }
#pragma unroll
  for (int q = 0; q < E; ++q) d[q] = x[q];
}

// Line solves along y (DIR 1) or z (DIR 2) (and x, DIR 0, for the per-routine
// entry points): a block owns W lines of M rows, E elements per thread.
template <int DIR, int E, typename R>
__global__ __launch_bounds__(1024) void k_pcr(SolveBatch<R> batch, int M, int W, int n1, int n2, int n3) {
  extern __shared__ double lds[];
  const SolveField<R> F = batch.f[blockIdx.y];
  const int tid = threadIdx.x, nt = blockDim.x;
  int base, es, cs;
  unsigned t = xcd_block(blockIdx.x, gridDim.x);
  if (DIR == 0) {
    int ntj = n2 / W;
    int jt = t % ntj, k = t / ntj;
    base = n1 * (jt * W + n2 * k);
    es = 1; cs = n1;
  } else if (DIR == 1) {
    int nti = n1 / W;
    int it = t % nti, k = t / nti;
    base = it * W + n1 * n2 * k;
    es = n1; cs = 1;
  } else {
    int nti = n1 / W;
    int it = t % nti, j = t / nti;
    base = it * W + n1 * j;
    es = n1 * n2; cs = 1;
  }
  const int ldx = M + 1;
  constexpr bool XL = DIR == 0;
  auto own_r = [&](int idx) { return XL ? idx % M : idx / W; };
  auto own_c = [&](int idx) { return XL ? idx / M : idx % W; };
  auto la = [&](int r, int c) { return XL ? c * ldx + r : r * W + c; };
  const double ackl_b = 1.0 / (1.0 + 2.0 * F.betadx);
  double d[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    int idx = tid + q * nt;
    int r = own_r(idx), c = own_c(idx);
    d[q] = (double)F.rhs[base + r * es + c * cs] * ackl_b;
    lds[la(r, c)] = d[q];
  }
  __syncthreads();
  pcr_core<XL, E>(lds, M, W, ldx, F.betadx, d);
#pragma unroll
  for (int q = 0; q < E; ++q) {
    int idx = tid + q * nt;
    int gaddr = base + own_r(idx) * es + own_c(idx) * cs;
    if (F.qadd) {
      F.qadd[gaddr] = F.qadd[gaddr] + (R)d[q];
      if (F.store_rhs) F.rhs[gaddr] = (R)d[q];
    } else {
      F.rhs[gaddr] = (R)d[q];
    }
  }
}

// ------------------------------------------- truncated Green's function
// The exact inverse of the normalised periodic system x_i + a (x_{i-1} +
// x_{i+1}) = d_i (a = -G/(1+2G)) is a circulant with entries
//   g_m = C (rho^m + rho^(N-m)) / (1 - rho^N),  rho = -2a / (1 + sqrt(1 - 4a^2)),
//   C = 1 / (1 + 2 a rho)  (0 <= m < N, g_{-m} = g_m),
// so the solve is the symmetric filter x_i = g_0 d_i + sum_{m=1..K} g_m
// (d_{i-m} + d_{i+m}) once |g_{K+1}| is below the storage precision.  For the
// diagonally dominant systems of this path (rho ~ 1e-2 at 256^3) K is 8 in
// FP64, so the line solve becomes a short stencil: coalesced, no LDS
// reductions, no barriers.  Systems that decay more slowly (K > FIR_KMAX) use
// the cyclic reduction above.  The result equals the exact solution to
// round-off (the reference's Thomas + Sherman-Morrison is another route to it).
constexpr int FIR_KMAX = 12;
struct Green {
  int K;
  double g[FIR_KMAX + 1];
};

// x of element r of an LDS line (elements L[r], r < M, periodic).
__device__ __forceinline__ double green_lds(const Green& G, const double* L, int r, int M) {
  double x = G.g[0] * L[r];
  for (int m = 1; m <= G.K; ++m) {
    int rm = r - m; if (rm < 0) rm += M;
    int rp = r + m; if (rp >= M) rp -= M;
    x += G.g[m] * (L[rm] + L[rp]);
  }
  return x;
}

// Explicit terms + RHS of all four transported fields fused with the x-line
// solve (pass A of the substep): a block owns W complete x-lines at one z
// plane, computes rhs x4 and ru x4 per cell (E cells per thread, reading the
// state through L1/L2), keeps the scaled rhs in LDS, solves the four
// x-systems there and writes the x-solved rhs.  Saves one read and one write
// of rhs x4 per substep against separate passes.
template <int E, typename R, bool FIR>
__global__ __launch_bounds__(256) void k_explicit_xsolve(EulerArgs<R> A, RhsCoef<R> cf, R* __restrict__ ru1, R* __restrict__ ru2,
                                                         R* __restrict__ ru3, R* __restrict__ rut, R* __restrict__ r_u, R* __restrict__ r_v,
                                                         R* __restrict__ r_w, R* __restrict__ r_t, double bxu, double bxt, int W, Green gu,
                                                         Green gt) {
  extern __shared__ double lds[];
  const Grid& g = A.g;
  const int M = g.n1, ldx = M + 1, tid = threadIdx.x, nt = blockDim.x;
  unsigned t = xcd_block(blockIdx.x, gridDim.x);
  int jt, k;
  tiled_2d(t, g.n2 / W, g.n3, 16, 8, jt, k);
  const int j0 = jt * W;
  const int base = M * (j0 + g.n2 * k);
  const int plane = W * ldx;
  const double sc_u = 1.0 / (1.0 + 2.0 * bxu), sc_t = 1.0 / (1.0 + 2.0 * bxt);
#pragma unroll 1
  for (int q = 0; q < E; ++q) {
    int idx = tid + q * nt;
    int r = idx % M, c = idx / M;
    Cell s = cell_at(g, r, j0 + c, k);
    const int gc = s.c;
    const R* vp = A.vofp;
    R p = vp[gc];
    R hu = h1_val(A, s, R(0.5) * (vp[s.at(-1, 0, 0)] + p));
    R hv = h2_val(A, s, R(0.5) * (vp[s.at(0, -1, 0)] + p));
    R hw = h3_val(A, s, R(0.5) * (vp[s.at(0, 0, -1)] + p));
    R ht = ht_val(A, s, p);
    R o1 = ru1[gc], o2 = ru2[gc], o3 = ru3[gc], ot = rut[gc];
    const int l = c * ldx + r;
    lds[l] = (double)rhs_u(cf, A.vx, A.pr, s, hu, o1) * sc_u;
    lds[plane + l] = (double)rhs_v(cf, A.vy, A.pr, s, hv, o2) * sc_u;
    lds[2 * plane + l] = (double)rhs_w(cf, A.vz, A.pr, s, hw, o3) * sc_u;
    lds[3 * plane + l] = (double)rhs_t(cf, A.temp, s, ht, ot) * sc_t;
    ru1[gc] = hu; ru2[gc] = hv; ru3[gc] = hw; rut[gc] = ht;
  }
  __syncthreads();
  R* outs[4] = {r_u, r_v, r_w, r_t};
  if constexpr (FIR) {
#pragma unroll
    for (int q = 0; q < E; ++q) {
      int idx = tid + q * nt;
      int r = idx % M, c = idx / M;
      const double* L = lds + c * ldx;
#pragma unroll
      for (int f = 0; f < 4; ++f) outs[f][base + idx] = (R)green_lds(f == 3 ? gt : gu, L + f * plane, r, M);
    }
    return;
  }
#pragma unroll 1
  for (int f = 0; f < 4; ++f) {
    double* L = lds + f * plane;
    double d[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
      int idx = tid + q * nt;
      d[q] = L[(idx / M) * ldx + idx % M];
    }
    pcr_core<true, E>(L, M, W, ldx, f == 3 ? bxt : bxu, d);
#pragma unroll
    for (int q = 0; q < E; ++q) {
      int idx = tid + q * nt;
      outs[f][base + idx] = (R)d[q];
    }
  }
}

// Elements per thread the templates are instantiated for.
constexpr int kEs[] = {1, 2, 3, 4, 6, 8};

// Threads per tile of n elements: the largest divisor of n that is at most
// cap and whose element count per thread is one of kEs.
int pick_threads(int n, int cap, int want) {
  if (want > 0 && want <= cap && n % want == 0)
    for (int e : kEs)
      if (n / want == e) return want;
  for (int e : kEs)
    if (n % e == 0 && n / e <= cap) return n / e;
  return 0;
}

// y (DIR 1) or z (DIR 2) line solve as the truncated Green's function: one
// thread per element, reading src and writing dst (dst != src), or adding into
// qadd (solxk.f90:44-50).  G folds in the 1/(1+2G) scaling of the rhs.
template <typename R>
struct GreenBatch {
  const R* src[4];
  R* dst[4];
  R* qadd[4];
  int which[4];   // index into gr[] of each field's coefficients
  Green gr[2];
};

// A block is 64 consecutive i (coalesced rows) x 4 segments of GQ points of
// the line; each thread slides a register window of GQ + 2 KT values along
// its segment, so every input is read about (GQ + 2 KT) / GQ times (the halo
// rows of neighbouring segments mostly from L2).  Coefficients beyond the
// system's K are zero.  Needs n1 % 64 == 0 and N % (4 GQ) == 0.
constexpr int GQ = 16;

template <int DIR, typename R, int KT>
__global__ __launch_bounds__(256) void k_green(GreenBatch<R> B, int n1, int n2, int n3) {
  const int f = blockIdx.y;
  const int N = DIR == 1 ? n2 : n3;
  const int s = DIR == 1 ? n1 : n1 * n2;
  const int nsg = N / (4 * GQ), nit = n1 / 64;
  unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const int sg = (int)(b % nsg), rest = (int)(b / nsg);
  const int it = rest % nit, other = rest / nit;   // other: k (DIR 1) or j (DIR 2)
  const int i = it * 64 + (threadIdx.x & 63);
  const int p0 = (sg * 4 + (threadIdx.x >> 6)) * GQ;
  const int base = i + (DIR == 1 ? n1 * n2 * other : n1 * other);
  const Green& G = B.gr[B.which[f]];
  double g[KT + 1];
#pragma unroll
  for (int m = 0; m <= KT; ++m) g[m] = m <= G.K ? G.g[m] : 0.0;
  const R* __restrict__ d = B.src[f];
  double v[GQ + 2 * KT];
#pragma unroll
  for (int q = 0; q < GQ + 2 * KT; ++q) {
    int p = p0 - KT + q;
    p = p < 0 ? p + N : (p >= N ? p - N : p);
    v[q] = (double)d[base + p * s];
  }
#pragma unroll
  for (int q = 0; q < GQ; ++q) {
    double x = g[0] * v[q + KT];
#pragma unroll
    for (int m = 1; m <= KT; ++m) x += g[m] * (v[q + KT - m] + v[q + KT + m]);
    const int c = base + (p0 + q) * s;
    if (B.qadd[f]) B.qadd[f][c] = B.qadd[f][c] + (R)x;
    else B.dst[f][c] = (R)x;
  }
}

// Coefficients of the truncated Green's function of the normalised system
// with off-diagonal ratio G = betadx on N points, scaled by `scale`; false
// when more than FIR_KMAX terms are needed for relative truncation `tol`.
bool green_coeffs(double betadx, int N, double tol, double scale, Green& out) {
  long double G = betadx;
  long double a = -G / (1.0L + 2.0L * G);
  if (a == 0.0L) {
    out.K = 0;
    out.g[0] = scale;
    return true;
  }
  long double rho = -2.0L * a / (1.0L + sqrtl(1.0L - 4.0L * a * a));
  long double C = 1.0L / (1.0L + 2.0L * a * rho);
  long double rN = powl(rho, (long double)N);
  int K = 0;
  while (K < N / 2 - 1 && fabsl(powl(rho, (long double)(K + 1))) > tol) ++K;
  if (K > FIR_KMAX || 2 * K + 1 > N) return false;
  out.K = K;
  for (int m = 0; m <= K; ++m)
    out.g[m] = (double)(scale * C * (powl(rho, (long double)m) + powl(rho, (long double)(N - m))) / (1.0L - rN));
  return true;
}

// ------------------------------------------------------------ divg etc
template <typename R>
__global__ __launch_bounds__(256) void k_divg(Grid g, const R* __restrict__ vx, const R* __restrict__ vy, const R* __restrict__ vz,
                                              R* __restrict__ dph, R usdtal, int nc) {
  unsigned b = xcd_block(blockIdx.x, gridDim.x);
  int c = (int)(b * blockDim.x + threadIdx.x);
  if (c >= nc) return;
  Cell s = make_cell(g, c);
  R dq = (vx[s.at(1, 0, 0)] - vx[c]) * R(g.dx1) + (vy[s.at(0, 1, 0)] - vy[c]) * R(g.dx2) + (vz[s.at(0, 0, 1)] - vz[c]) * R(g.dx3);
  dph[c] = dq * usdtal;
}

// updvp.f90:16-31 and prcalc.f90:17-33 in one pass; with mean != null the
// mean MLS force deferred from velforce is subtracted first ((u+f)-fbar).
template <typename R>
__global__ __launch_bounds__(256) void k_project(Grid g, R* __restrict__ vx, R* __restrict__ vy, R* __restrict__ vz, R* __restrict__ pr,
                                                 const R* __restrict__ dph, const double* __restrict__ mean, R udx1, R udx2, R udx3, R be,
                                                 R dx1q, R dx2q, R dx3q, int do_u, int do_p, int nc, const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  unsigned b = xcd_block(blockIdx.x, gridDim.x);
  Cell s;
  if (g.n1 % blockDim.x == 0) {  // blocks tile whole x-lines: walk (j, k) in 16 x 8 tiles
    const int per = g.n1 / blockDim.x;
    int jj, kk;
    tiled_2d(b / per, g.n2, g.n3, 16, 8, jj, kk);
    s = cell_at(g, (int)((b % per) * blockDim.x + threadIdx.x), jj, kk);
  } else {
    int c = (int)(b * blockDim.x + threadIdx.x);
    if (c >= nc) return;
    s = make_cell(g, c);
  }
  const int c = s.c;
  R l = dph[c];
  R pim = dph[s.at(-1, 0, 0)], pjm = dph[s.at(0, -1, 0)], pkm = dph[s.at(0, 0, -1)];
  if (do_u) {
    R u = vx[c], v = vy[c], w = vz[c];
    if (mean) { u = u - R(mean[0]); v = v - R(mean[1]); w = w - R(mean[2]); }
    vx[c] = u - (l - pim) * udx1;
    vy[c] = v - (l - pjm) * udx2;
    vz[c] = w - (l - pkm) * udx3;
  }
  if (do_p) {
    R pip = dph[s.at(1, 0, 0)], pjp = dph[s.at(0, 1, 0)], pkp = dph[s.at(0, 0, 1)];
    pr[c] = pr[c] + l - be * ((pip - R(2) * l + pim) * dx1q + (pjp - R(2) * l + pjm) * dx2q + (pkp - R(2) * l + pkm) * dx3q);
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_interp_vof(Grid g, const R* __restrict__ p, R* __restrict__ vx, R* __restrict__ vy, R* __restrict__ vz,
                                                    int nc) {
  int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (c >= nc) return;
  Cell s = make_cell(g, c);
  R v = p[c];
  vx[c] = R(0.5) * (p[s.at(-1, 0, 0)] + v);
  vy[c] = R(0.5) * (p[s.at(0, -1, 0)] + v);
  vz[c] = R(0.5) * (p[s.at(0, 0, -1)] + v);
}

__global__ void k_zero_u64(unsigned long long* p) { *p = 0ull; }

// Max reductions of cfl.f90:9-25 (what = 0) and divgck divg.f90:36-50
// (what = 1), in FP64 arithmetic whatever the storage type; the result
// (non-negative) is combined by atomicMax on its bit pattern.
template <typename R>
__global__ __launch_bounds__(256) void k_reduce_max(Grid g, const R* __restrict__ vx, const R* __restrict__ vy, const R* __restrict__ vz,
                                                    int what, int nc, unsigned long long* out) {
  __shared__ double sm[256];
  double m = what == 0 ? 0.00000001 : 0.0;
  for (int c = (int)(blockIdx.x * blockDim.x + threadIdx.x); c < nc; c += (int)(gridDim.x * blockDim.x)) {
    Cell s = make_cell(g, c);
    double q;
    double x0 = vx[c], x1 = vx[s.at(1, 0, 0)], y0 = vy[c], y1 = vy[s.at(0, 1, 0)], z0 = vz[c], z1 = vz[s.at(0, 0, 1)];
    if (what == 0)
      q = fabs((x0 + x1) * 0.5 * g.dx1) + fabs((y0 + y1) * 0.5 * g.dx2) + fabs((z0 + z1) * 0.5 * g.dx3);
    else
      q = fabs((x1 - x0) * g.dx1 + (y1 - y0) * g.dx2 + (z1 - z0) * g.dx3);
    m = fmax(m, q);
  }
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + w]);
    __syncthreads();
  }
  // non-negative doubles order like their bit patterns
  if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(sm[0]));
}

// Refinement gate of the precision form: refine while the FP64 divergence of
// the projected velocity exceeds rtol times that of the provisional one.
__global__ void k_refine_gate(const unsigned long long* __restrict__ dmax, double rtol, int* gate, long long* stats) {
  double before = __longlong_as_double((long long)dmax[0]), after = __longlong_as_double((long long)dmax[1]);
  int on = after > rtol * before;
  *gate = on;
  stats[0] += 1;
  stats[1] += on;
  stats[2] = __double_as_longlong(before > 0.0 ? after / before : 0.0);
}

template <typename R>
__global__ __launch_bounds__(256) void k_inqpr(Grid g, R* __restrict__ vx, R* __restrict__ vy, R* __restrict__ vz, R* __restrict__ t,
                                               const melt_body* __restrict__ body, double tliq, double tsol, double rad_p, int nc) {
  int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (c >= nc) return;
  int i = c % g.n1 + 1, j = (c / g.n1) % g.n2 + 1, k = c / (g.n1 * g.n2) + 1;
  double v[3] = {gxm(g, i) - body->pos_cm[0], gym(g, j) - body->pos_cm[1], gzm(g, k) - body->pos_cm[2]};
  double rr = fnorm2(v, 3);
  vx[c] = R(0); vy[c] = R(0); vz[c] = R(0);
  t[c] = (R)(tliq - (tliq - tsol) * 0.5 * (1.0 - tanh((rr - rad_p) * g.dx1 / 2.0)));
}

inline unsigned nblocks(size_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ------------------------------------------------------------ launchers
template <typename R>
int launch_hdnl(melt_ctx* ctx, int which) {
  static const int out_id[4] = {MELT_DQ, MELT_DPH, MELT_QCAP, MELT_HTEMP};
  EulerArgs<R> A = make_args<R>(ctx);
  hipLaunchKernelGGL(k_hdnl<R>, dim3(nblocks(ctx->nc, 256)), dim3(256), 0, ctx->stream, A, fld<R>(ctx, out_id[which]), which, (int)ctx->nc);
  MELT_HIP(hipGetLastError());
  return 0;
}

template <typename R>
int launch_invtr_rhs(melt_ctx* ctx, int which) {
  static const int q_id[4] = {MELT_VX, MELT_VY, MELT_VZ, MELT_TEMP};
  static const int h_id[4] = {MELT_DQ, MELT_DPH, MELT_QCAP, MELT_HTEMP};
  static const int ru_id[4] = {MELT_RU1, MELT_RU2, MELT_RU3, MELT_RUT};
  RhsCoef<R> cf = make_coef<R>(ctx);
  hipLaunchKernelGGL(k_invtr_rhs<R>, dim3(nblocks(ctx->nc, 256)), dim3(256), 0, ctx->stream, make_grid(ctx), cf, fld<R>(ctx, q_id[which]),
                     fld<R>(ctx, MELT_PR), fld<R>(ctx, h_id[which]), fld<R>(ctx, ru_id[which]), fld<R>(ctx, MELT_RHS), which, (int)ctx->nc);
  MELT_HIP(hipGetLastError());
  return 0;
}

// Pass A: explicit terms + RHS + x-solve of u, v, w, T into rhs4[].
template <typename R>
int launch_explicit_xsolve(melt_ctx* ctx, double bxu, double bxt) {
  EulerArgs<R> A = make_args<R>(ctx);
  RhsCoef<R> cf = make_coef<R>(ctx);
  // W x-lines per block: the fewest that give a 256-cell tile (one cell per
  // thread); small tiles keep the LDS and register footprint low.
  int W = 0;
  if (ctx->tune[TUNE_XF_W] > 0 && ctx->n2 % ctx->tune[TUNE_XF_W] == 0) W = ctx->tune[TUNE_XF_W];
  else
    for (int w = 1; w <= ctx->n2 && !W; ++w)
      if (ctx->n2 % w == 0 && (w * ctx->n1 >= 256 || w == ctx->n2)) W = w;
  if (W * ctx->n1 > 2048) return melt_fail(ctx, MELT_ERR_ARG, "explicit_xsolve", "tile too large");
  int ncell = W * ctx->n1;
  int nt = pick_threads(ncell, 256, ctx->tune[TUNE_XF_NT]);
  if (!nt) return melt_fail(ctx, MELT_ERR_ARG, "explicit_xsolve", "unsupported line length");
  int E = ncell / nt;
  size_t shm = sizeof(double) * 4 * (size_t)W * (ctx->n1 + 1);
  unsigned blocks = (unsigned)((size_t)(ctx->n2 / W) * ctx->n3);
  R* r4[4] = {static_cast<R*>(ctx->rhs4[RHS_U]), static_cast<R*>(ctx->rhs4[RHS_V]), static_cast<R*>(ctx->rhs4[RHS_W]),
              static_cast<R*>(ctx->rhs4[RHS_T])};
  // the LDS line holds the scaled rhs, so the filter is the unscaled one
  Green gu{}, gt{};
  const double tol = ctx->fp32 ? 1e-8 : 2e-17;
  const bool fir = !ctx->no_green && green_coeffs(bxu, ctx->n1, tol, 1.0, gu) && green_coeffs(bxt, ctx->n1, tol, 1.0, gt);
#define XF_CASE(EE)                                                                                                                   \
  case EE:                                                                                                                            \
    if (fir)                                                                                                                          \
      hipLaunchKernelGGL((k_explicit_xsolve<EE, R, true>), dim3(blocks), dim3(nt), shm, ctx->stream, A, cf, fld<R>(ctx, MELT_RU1),     \
                         fld<R>(ctx, MELT_RU2), fld<R>(ctx, MELT_RU3), fld<R>(ctx, MELT_RUT), r4[0], r4[1], r4[2], r4[3], bxu, bxt, W, \
                         gu, gt);                                                                                                     \
    else                                                                                                                              \
      hipLaunchKernelGGL((k_explicit_xsolve<EE, R, false>), dim3(blocks), dim3(nt), shm, ctx->stream, A, cf, fld<R>(ctx, MELT_RU1),    \
                         fld<R>(ctx, MELT_RU2), fld<R>(ctx, MELT_RU3), fld<R>(ctx, MELT_RUT), r4[0], r4[1], r4[2], r4[3], bxu, bxt, W, \
                         gu, gt);                                                                                                     \
    break;
  switch (E) {
    XF_CASE(1) XF_CASE(2) XF_CASE(3) XF_CASE(4) XF_CASE(6) XF_CASE(8)
    default: return melt_fail(ctx, MELT_ERR_ARG, "explicit_xsolve", "unsupported tile");
  }
#undef XF_CASE
  MELT_HIP(hipGetLastError());
  return 0;
}

// dir 0/1/2 = x/y/z; fields[nfield] are solved in place, or (qadd != null)
// added into qadd[] (solxk.f90:44-50).
template <typename R>
int launch_solve(melt_ctx* ctx, int dir, void* const* fields, int nfield, const double* betadx, void* const* qadd, int store_rhs) {
  int n[3] = {ctx->n1, ctx->n2, ctx->n3};
  int M = n[dir];
  int other = dir == 0 ? ctx->n2 : ctx->n1;
  // W lines per tile: 16 by default (128-byte rows for y/z), at most 8 elements per thread
  int W = ctx->tune[TUNE_SOLVE_W] > 0 ? ctx->tune[TUNE_SOLVE_W] : 16;
  while (W > 1 && (other % W || M * W > 8 * 1024)) W /= 2;
  if (other % W) W = 1;
  int nt = pick_threads(M * W, 1024, ctx->tune[TUNE_SOLVE_NT] > 0 ? ctx->tune[TUNE_SOLVE_NT] : 256);
  if (!nt) return melt_fail(ctx, MELT_ERR_ARG, "line solve", "unsupported line length");
  int E = (M * W) / nt;
  SolveBatch<R> B{};
  for (int q = 0; q < nfield; ++q) {
    B.f[q].rhs = static_cast<R*>(fields[q]);
    B.f[q].qadd = qadd ? static_cast<R*>(qadd[q]) : nullptr;
    B.f[q].betadx = betadx[q];
    B.f[q].store_rhs = store_rhs;
  }
  size_t tiles = dir == 0 ? (size_t)(ctx->n2 / W) * ctx->n3 : dir == 1 ? (size_t)(ctx->n1 / W) * ctx->n3 : (size_t)(ctx->n1 / W) * ctx->n2;
  size_t shm = dir == 0 ? sizeof(double) * (size_t)W * (M + 1) : sizeof(double) * (size_t)W * M;
  if (shm > 64 * 1024) return melt_fail(ctx, MELT_ERR_ARG, "line solve", "tile exceeds LDS");
  dim3 grid((unsigned)tiles, nfield);
#define PCR_CASE(D, EE) \
  case EE: hipLaunchKernelGGL((k_pcr<D, EE, R>), grid, dim3(nt), shm, ctx->stream, B, M, W, ctx->n1, ctx->n2, ctx->n3); break;
#define PCR_DIR(D)                                                                                          \
  switch (E) {                                                                                              \
    PCR_CASE(D, 1) PCR_CASE(D, 2) PCR_CASE(D, 3) PCR_CASE(D, 4) PCR_CASE(D, 6) PCR_CASE(D, 8)                 \
    default: return melt_fail(ctx, MELT_ERR_ARG, "line solve", "unsupported tile");                          \
  }
  if (dir == 0) { PCR_DIR(0) }
  else if (dir == 1) { PCR_DIR(1) }
  else { PCR_DIR(2) }
#undef PCR_DIR
#undef PCR_CASE
  MELT_HIP(hipGetLastError());
  return 0;
}

// y/z solves of the fused substep as Green filters: returns 1 (and launches
// nothing) when a system needs the cyclic reduction instead.
template <typename R>
int launch_green(melt_ctx* ctx, int dir, void* const* src, void* const* dst, void* const* qadd, const double* betadx, int nfield,
                 int* used) {
  *used = 0;
  if (ctx->no_green) return 0;
  const int N = dir == 1 ? ctx->n2 : ctx->n3;
  const double tol = ctx->fp32 ? 1e-8 : 2e-17;
  GreenBatch<R> B{};
  int ng = 0;
  double bseen[2] = {0.0, 0.0};
  for (int q = 0; q < nfield; ++q) {
    int w = -1;
    for (int i = 0; i < ng; ++i)
      if (bseen[i] == betadx[q]) w = i;
    if (w < 0) {
      if (ng == 2) return 0;
      if (!green_coeffs(betadx[q], N, tol, 1.0 / (1.0 + 2.0 * betadx[q]), B.gr[ng])) return 0;
      bseen[ng] = betadx[q];
      w = ng++;
    }
    B.which[q] = w;
    B.src[q] = static_cast<const R*>(src[q]);
    B.dst[q] = dst ? static_cast<R*>(dst[q]) : nullptr;
    B.qadd[q] = qadd ? static_cast<R*>(qadd[q]) : nullptr;
  }
  int kmax = std::max(B.gr[0].K, ng > 1 ? B.gr[1].K : 0);
  int KT = kmax <= 4 ? 4 : kmax <= 8 ? 8 : 12;
  // KT < N / 2 keeps every window inside one period
  if (ctx->n1 % 64 || N % (4 * GQ) || 2 * KT >= N) return 0;
  const int nsg = N / (4 * GQ), nit = ctx->n1 / 64, other = dir == 1 ? ctx->n3 : ctx->n2;
  dim3 grid((unsigned)(nsg * nit * other), nfield);
#define GREEN_CASE(D, KK) hipLaunchKernelGGL((k_green<D, R, KK>), grid, dim3(256), 0, ctx->stream, B, ctx->n1, ctx->n2, ctx->n3)
  if (dir == 1) { if (KT == 4) GREEN_CASE(1, 4); else if (KT == 8) GREEN_CASE(1, 8); else GREEN_CASE(1, 12); }
  else { if (KT == 4) GREEN_CASE(2, 4); else if (KT == 8) GREEN_CASE(2, 8); else GREEN_CASE(2, 12); }
#undef GREEN_CASE
  MELT_HIP(hipGetLastError());
  *used = 1;
  return 0;
}

template <typename R>
int launch_divg(melt_ctx* ctx) {
  R usdtal = R(1.0 / (ctx->dt * ctx->al));
  hipLaunchKernelGGL(k_divg<R>, dim3(nblocks(ctx->nc, 256)), dim3(256), 0, ctx->stream, make_grid(ctx), fld<R>(ctx, MELT_VX),
                     fld<R>(ctx, MELT_VY), fld<R>(ctx, MELT_VZ), fld<R>(ctx, MELT_DPH), usdtal, (int)ctx->nc);
  MELT_HIP(hipGetLastError());
  return 0;
}

template <typename R>
int launch_project(melt_ctx* ctx, int do_u, int do_p, int use_mean, const int* gate) {
  R udx1 = R(ctx->al * ctx->dt * ctx->dx1), udx2 = R(ctx->al * ctx->dt * ctx->dx2), udx3 = R(ctx->al * ctx->dt * ctx->dx3);
  R be = R(ctx->al * ctx->beta);
  hipLaunchKernelGGL(k_project<R>, dim3(nblocks(ctx->nc, 256)), dim3(256), 0, ctx->stream, make_grid(ctx), fld<R>(ctx, MELT_VX),
                     fld<R>(ctx, MELT_VY), fld<R>(ctx, MELT_VZ), fld<R>(ctx, MELT_PR), fld<R>(ctx, MELT_DPH),
                     use_mean ? ctx->d_fmean : nullptr, udx1, udx2, udx3, be, R(ctx->dx1q), R(ctx->dx2q), R(ctx->dx3q), do_u, do_p,
                     (int)ctx->nc, gate);
  MELT_HIP(hipGetLastError());
  return 0;
}

template <typename R>
int launch_interp_vof(melt_ctx* ctx) {
  hipLaunchKernelGGL(k_interp_vof<R>, dim3(nblocks(ctx->nc, 256)), dim3(256), 0, ctx->stream, make_grid(ctx), fld<R>(ctx, MELT_VOFP),
                     fld<R>(ctx, MELT_VOFX), fld<R>(ctx, MELT_VOFY), fld<R>(ctx, MELT_VOFZ), (int)ctx->nc);
  MELT_HIP(hipGetLastError());
  return 0;
}

// Launch the max reduction into d (zeroed by a fill kernel first, so the
// sequence may be captured into a graph).
template <typename R>
int launch_reduce_max(melt_ctx* ctx, int what, unsigned long long* d) {
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, ctx->stream, d);
  unsigned nb = (unsigned)std::min<size_t>(1024, nblocks(ctx->nc, 256));
  hipLaunchKernelGGL(k_reduce_max<R>, dim3(nb), dim3(256), 0, ctx->stream, make_grid(ctx), fld<R>(ctx, MELT_VX), fld<R>(ctx, MELT_VY),
                     fld<R>(ctx, MELT_VZ), what, (int)ctx->nc, d);
  MELT_HIP(hipGetLastError());
  return 0;
}

template <typename R>
int launch_inqpr(melt_ctx* ctx) {
  hipLaunchKernelGGL(k_inqpr<R>, dim3(nblocks(ctx->nc, 256)), dim3(256), 0, ctx->stream, make_grid(ctx), fld<R>(ctx, MELT_VX),
                     fld<R>(ctx, MELT_VY), fld<R>(ctx, MELT_VZ), fld<R>(ctx, MELT_TEMP), ctx->d_body, ctx->cfg.tliq, ctx->cfg.tsol, ctx->rad_p,
                     (int)ctx->nc);
  MELT_HIP(hipGetLastError());
  return 0;
}

}  // namespace

// ------------------------------------------------- precision dispatch
int euler_launch_hdnl(melt_ctx* ctx, int which) { return ctx->fp32 ? launch_hdnl<float>(ctx, which) : launch_hdnl<double>(ctx, which); }
int euler_launch_invtr_rhs(melt_ctx* ctx, int which) {
  return ctx->fp32 ? launch_invtr_rhs<float>(ctx, which) : launch_invtr_rhs<double>(ctx, which);
}
int euler_launch_explicit_xsolve(melt_ctx* ctx, double bxu, double bxt) {
  return ctx->fp32 ? launch_explicit_xsolve<float>(ctx, bxu, bxt) : launch_explicit_xsolve<double>(ctx, bxu, bxt);
}
int euler_launch_solve(melt_ctx* ctx, int dir, void* const* fields, int nfield, const double* betadx, void* const* qadd, int store_rhs) {
  return ctx->fp32 ? launch_solve<float>(ctx, dir, fields, nfield, betadx, qadd, store_rhs)
                   : launch_solve<double>(ctx, dir, fields, nfield, betadx, qadd, store_rhs);
}
int euler_launch_green(melt_ctx* ctx, int dir, void* const* src, void* const* dst, void* const* qadd, const double* betadx, int nfield,
                       int* used) {
  return ctx->fp32 ? launch_green<float>(ctx, dir, src, dst, qadd, betadx, nfield, used)
                   : launch_green<double>(ctx, dir, src, dst, qadd, betadx, nfield, used);
}
int euler_launch_divg(melt_ctx* ctx) { return ctx->fp32 ? launch_divg<float>(ctx) : launch_divg<double>(ctx); }
int euler_launch_project(melt_ctx* ctx, int do_u, int do_p, int use_mean, const int* gate) {
  return ctx->fp32 ? launch_project<float>(ctx, do_u, do_p, use_mean, gate) : launch_project<double>(ctx, do_u, do_p, use_mean, gate);
}
int euler_launch_interp_vof(melt_ctx* ctx) { return ctx->fp32 ? launch_interp_vof<float>(ctx) : launch_interp_vof<double>(ctx); }
int euler_launch_reduce_max(melt_ctx* ctx, int what, unsigned long long* d) {
  return ctx->fp32 ? launch_reduce_max<float>(ctx, what, d) : launch_reduce_max<double>(ctx, what, d);
}
int euler_launch_refine_gate(melt_ctx* ctx, const unsigned long long* dmax, int* gate) {
  hipLaunchKernelGGL(k_refine_gate, dim3(1), dim3(1), 0, ctx->stream, dmax, ctx->refine_rtol, gate, ctx->d_refine_stats);
  MELT_HIP(hipGetLastError());
  return 0;
}
int euler_reduce_max(melt_ctx* ctx, int what, double* out) {
  unsigned long long* d = reinterpret_cast<unsigned long long*>(ctx->d_red);
  int rc = euler_launch_reduce_max(ctx, what, d);
  if (rc) return rc;
  unsigned long long h = 0;
  MELT_HIP(hipMemcpyAsync(&h, d, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
  MELT_HIP(hipStreamSynchronize(ctx->stream));
  double v;
  memcpy(&v, &h, sizeof v);
  *out = v;
  return 0;
}
int euler_launch_inqpr(melt_ctx* ctx) { return ctx->fp32 ? launch_inqpr<float>(ctx) : launch_inqpr<double>(ctx); }
