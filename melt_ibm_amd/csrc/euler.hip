// Eulerian part of the RK3 substep on gfx950: explicit advection/diffusion
// stencils (hdnl1-3, hdnlte, invtr1-3, invtrte right-hand sides), the
// Crank-Nicolson periodic line solves (solxi/solxj/solxk) as circulant parallel
// cyclic reduction in LDS, the divergence, and the projection (updvp, prcalc).
//
// All of it is HBM-bound (about 1 flop per byte), so the kernels are written for
// coalesced unit-stride access and to read each input field once per pass.
#include <algorithm>
#include <cstring>
#include <type_traits>

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
  template <typename T>
  __device__ __forceinline__ T ld(const T* p, int a, int b, int d) const { return p[at(a, b, d)]; }
  template <typename T>
  __device__ __forceinline__ T& ref(T* p) const { return p[c]; }
};

// A cell addressed by 32-bit byte offsets from the field base, for fields
// under 4 GiB: the field pointers stay in scalar registers and every
// neighbour load is that base plus one vector offset (the 13 distinct
// neighbour offsets are formed once and shared by all fields).  The wrapped
// steps are kept modulo 2^32.
template <typename T>
struct FCell {
  int i, j, k;
  unsigned o, xm, xp, ym, yp, zm, zp;
  __device__ __forceinline__ unsigned off(int a, int b, int d) const {
    return o + (a < 0 ? xm : a > 0 ? xp : 0u) + (b < 0 ? ym : b > 0 ? yp : 0u) + (d < 0 ? zm : d > 0 ? zp : 0u);
  }
  __device__ __forceinline__ T ld(const T* p, int a, int b, int d) const {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + off(a, b, d));
  }
  __device__ __forceinline__ T& ref(T* p) const { return *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + o); }
};

template <typename T>
__device__ __forceinline__ FCell<T> fcell_at(const Grid& g, int i, int j, int k) {
  FCell<T> s;
  const unsigned E = sizeof(T), n1 = (unsigned)g.n1, n12 = n1 * (unsigned)g.n2;
  s.i = i; s.j = j; s.k = k;
  s.o = ((unsigned)i + n1 * (unsigned)j + n12 * (unsigned)k) * E;
  s.xm = i == 0 ? (n1 - 1) * E : 0u - E;
  s.xp = i == g.n1 - 1 ? 0u - (n1 - 1) * E : E;
  s.ym = j == 0 ? (n12 - n1) * E : 0u - n1 * E;
  s.yp = j == g.n2 - 1 ? 0u - (n12 - n1) * E : n1 * E;
  s.zm = k == 0 ? (unsigned)(g.n3 - 1) * n12 * E : 0u - n12 * E;
  s.zp = k == g.n3 - 1 ? 0u - (unsigned)(g.n3 - 1) * n12 * E : n12 * E;
  return s;
}

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
  const double* __restrict__ coord;  // coordinate tables (coord_off)
  // face volume fractions as fields (sphereTagging sets them geometrically);
  // null: 1/2 (VOFp of the two cells), their value after every narrow-band
  // tagging (tagCells.f90:105-107)
  const R* __restrict__ vofx;
  const R* __restrict__ vofy;
  const R* __restrict__ vofz;
  Grid g;
  double rlen[3];                    // 1/xlen, 1/ylen, 1/zlen
  double betagz, tliq;
};

// The stencil formulas below read field F at offset (a, b, d) from the cell as
// SV(F, a, b, d); a stencil type supplies the values (straight from the
// fields, or from registers and LDS in the z-marching pass).
enum : int { FU, FV, FW, FT, FP, FVOF, FFX, FFY, FFZ, FN };
#define SV(F, a, b, d) s.template v<F, a, b, d>()

// Reads straight from the fields through a cell addressing C (Cell: element
// indices, any size; FCell: 32-bit byte offsets).
template <typename R, typename C>
struct PtrStencil {
  const R* p[FN];
  C c;
  int i, j, k;
  template <int F, int a, int b, int d>
  __device__ __forceinline__ R v() const { return c.ld(p[F], a, b, d); }
};

template <typename R, typename C>
__device__ __forceinline__ PtrStencil<R, C> ptr_stencil(const EulerArgs<R>& A, const C& c) {
  return PtrStencil<R, C>{{A.vx, A.vy, A.vz, A.temp, A.pr, A.vofp, A.forcx, A.forcy, A.forcz}, c, c.i, c.j, c.k};
}

// hdnl1.f90:32-74
template <typename R, typename S>
__device__ inline R h1_val(const EulerArgs<R>& A, const S& s, R vofx) {
  R udx1 = R(A.g.dx1) * R(0.25), udx2 = R(A.g.dx2) * R(0.25), udx3 = R(A.g.dx3) * R(0.25);
  R h11 = ((SV(FU, 1, 0, 0) + SV(FU, 0, 0, 0)) * (SV(FU, 1, 0, 0) + SV(FU, 0, 0, 0)) -
           (SV(FU, -1, 0, 0) + SV(FU, 0, 0, 0)) * (SV(FU, -1, 0, 0) + SV(FU, 0, 0, 0))) * udx1;
  R h12 = ((SV(FV, 0, 1, 0) + SV(FV, -1, 1, 0)) * (SV(FU, 0, 1, 0) + SV(FU, 0, 0, 0)) -
           (SV(FV, 0, 0, 0) + SV(FV, -1, 0, 0)) * (SV(FU, 0, 0, 0) + SV(FU, 0, -1, 0))) * udx2;
  R h13 = ((SV(FW, 0, 0, 1) + SV(FW, -1, 0, 1)) * (SV(FU, 0, 0, 1) + SV(FU, 0, 0, 0)) -
           (SV(FW, 0, 0, 0) + SV(FW, -1, 0, 0)) * (SV(FU, 0, 0, 0) + SV(FU, 0, 0, -1))) * udx3;
  return -(h11 + h12 + h13) + vofx * SV(FFX, 0, 0, 0) / R(A.g.xlen);
}

// hdnl2.f90:33-75
template <typename R, typename S>
__device__ inline R h2_val(const EulerArgs<R>& A, const S& s, R vofy) {
  R udx1 = R(A.g.dx1) * R(0.25), udx2 = R(A.g.dx2) * R(0.25), udx3 = R(A.g.dx3) * R(0.25);
  R h21 = ((SV(FV, 1, 0, 0) + SV(FV, 0, 0, 0)) * (SV(FU, 1, 0, 0) + SV(FU, 1, -1, 0)) -
           (SV(FV, 0, 0, 0) + SV(FV, -1, 0, 0)) * (SV(FU, 0, 0, 0) + SV(FU, 0, -1, 0))) * udx1;
  R h22 = ((SV(FV, 0, 1, 0) + SV(FV, 0, 0, 0)) * (SV(FV, 0, 1, 0) + SV(FV, 0, 0, 0)) -
           (SV(FV, 0, -1, 0) + SV(FV, 0, 0, 0)) * (SV(FV, 0, -1, 0) + SV(FV, 0, 0, 0))) * udx2;
  R h23 = ((SV(FW, 0, 0, 1) + SV(FW, 0, -1, 1)) * (SV(FV, 0, 0, 1) + SV(FV, 0, 0, 0)) -
           (SV(FW, 0, 0, 0) + SV(FW, 0, -1, 0)) * (SV(FV, 0, 0, 0) + SV(FV, 0, 0, -1))) * udx3;
  return -(h21 + h22 + h23) + vofy * SV(FFY, 0, 0, 0) / R(A.g.ylen);
}

// hdnl3.f90:38-82
template <typename R, typename S>
__device__ inline R h3_val(const EulerArgs<R>& A, const S& s, R vofz) {
  R udx1 = R(A.g.dx1) * R(0.25), udx2 = R(A.g.dx2) * R(0.25), udx3 = R(A.g.dx3) * R(0.25);
  R h31 = (((SV(FU, 1, 0, 0) + SV(FU, 1, 0, -1)) * (SV(FW, 1, 0, 0) + SV(FW, 0, 0, 0))) -
           ((SV(FU, 0, 0, 0) + SV(FU, 0, 0, -1)) * (SV(FW, 0, 0, 0) + SV(FW, -1, 0, 0)))) * udx1;
  R h32 = (((SV(FV, 0, 1, 0) + SV(FV, 0, 1, -1)) * (SV(FW, 0, 1, 0) + SV(FW, 0, 0, 0))) -
           ((SV(FV, 0, 0, 0) + SV(FV, 0, 0, -1)) * (SV(FW, 0, 0, 0) + SV(FW, 0, -1, 0)))) * udx2;
  R h33 = (((SV(FW, 0, 0, 1) + SV(FW, 0, 0, 0)) * (SV(FW, 0, 0, 1) + SV(FW, 0, 0, 0))) -
           ((SV(FW, 0, 0, 0) + SV(FW, 0, 0, -1)) * (SV(FW, 0, 0, 0) + SV(FW, 0, 0, -1)))) * udx3;
  R T_interp = (SV(FT, 0, 0, 0) + SV(FT, 0, 0, -1)) * R(0.5);
  R fbz = R(A.betagz) * (T_interp - R(A.tliq));
  return (-(h31 + h32 + h33) + fbz) + vofz * SV(FFZ, 0, 0, 0) / R(A.g.zlen);
}

// r - L NINT(r/L), the minimum image of hdnlte.f90:99-101, with NINT of the
// correctly rounded quotient as there: r * (1/L) is within 2 ulp of r/L and so
// rounds to the same integer unless it lies within 1e-9 of a half-integer,
// where the quotient itself is formed.
__device__ __forceinline__ double min_image(double r, double L, double rL) {
  const double q = r * rL, a = fabs(q);
  const int n = fabs(a - floor(a) - 0.5) < 1e-9 ? fnint(r / L) : fnint(q);
  return r - L * (double)n;
}

// hdnlte.f90:45-158 (Nparticle = 1)
template <typename R, typename S>
__device__ inline R ht_val(const EulerArgs<R>& A, const S& s, R vofp) {
  const Grid& g = A.g;
  R udx1 = R(g.dx1) * R(0.5), udx2 = R(g.dx2) * R(0.5), udx3 = R(g.dx3) * R(0.5);
  R T = SV(FT, 0, 0, 0);
  R h31 = (SV(FU, 1, 0, 0) * (SV(FT, 1, 0, 0) + T) - SV(FU, 0, 0, 0) * (T + SV(FT, -1, 0, 0))) * udx1;
  R h32 = (SV(FV, 0, 1, 0) * (SV(FT, 0, 1, 0) + T) - SV(FV, 0, 0, 0) * (T + SV(FT, 0, -1, 0))) * udx2;
  R h33 = (SV(FW, 0, 0, 1) * (SV(FT, 0, 0, 1) + T) - SV(FW, 0, 0, 0) * (T + SV(FT, 0, 0, -1))) * udx3;
  R ht = -(h31 + h32 + h33) * vofp;
  const double* pc = A.body->pos_cm;
  const double* vc = A.body->vel_cm;
  const double* om = A.body->omega_c;
  const double* xy = A.coord;
  const double xm = xy[coord_off(g, 0, 1) + s.i];
  const double yc = xy[coord_off(g, 1, 0) + s.j], ym = xy[coord_off(g, 1, 1) + s.j];
  const double zm = xy[coord_off(g, 2, 1) + s.k];
  // u at xc(ic): r(1) is not used; w at zc(kc): r(3) is not used and r is not wrapped
  double r1 = min_image(ym - pc[1], g.ylen, A.rlen[1]), r2 = min_image(zm - pc[2], g.zlen, A.rlen[2]);
  double u_imh = vc[0] + om[1] * r2 - om[2] * r1;
  h31 = R(u_imh) * (SV(FT, 1, 0, 0) - SV(FT, -1, 0, 0)) * udx1;
  double r0 = min_image(xm - pc[0], g.xlen, A.rlen[0]);
  double v_jmh = vc[1] + om[2] * r0 - om[0] * r2;
  h32 = R(v_jmh) * (SV(FT, 0, 1, 0) - SV(FT, 0, -1, 0)) * udx2;
  r0 = xm - pc[0]; r1 = ym - pc[1];
  double w_kmh = vc[2] + om[0] * r1 - om[1] * r0;
  h33 = R(w_kmh) * (SV(FT, 0, 0, 1) - SV(FT, 0, 0, -1)) * udx3;
  return ht - (h31 + h32 + h33) * (R(1) - vofp);
}

template <typename R>
struct RhsCoef {
  R ga, ro, al, dt, alre, dx1q, dx2q, dx3q, dx1, dx2, dx3, pec;
  R alpec;   // al / pec, formed once in R
};

// invtr1.f90:40-60
template <int FQ, typename R, typename S>
__device__ inline R rhs_u(const RhsCoef<R>& c, const S& s, R h, R ru) {
  R v = SV(FQ, 0, 0, 0);
  R d11 = (SV(FQ, 1, 0, 0) - R(2) * v + SV(FQ, -1, 0, 0)) * c.dx1q;
  R d22 = (SV(FQ, 0, 1, 0) - R(2) * v + SV(FQ, 0, -1, 0)) * c.dx2q;
  R d33 = (SV(FQ, 0, 0, 1) - R(2) * v + SV(FQ, 0, 0, -1)) * c.dx3q;
  R dc = d11 + d22 + d33;
  R dp = (SV(FP, 0, 0, 0) - SV(FP, -1, 0, 0)) * (c.dx1 * c.al);
  return (c.ga * h + c.ro * ru + c.alre * dc - dp) * c.dt;
}
// invtr2.f90:37-57
template <int FQ, typename R, typename S>
__device__ inline R rhs_v(const RhsCoef<R>& c, const S& s, R h, R ru) {
  R v = SV(FQ, 0, 0, 0);
  R d11 = (SV(FQ, 1, 0, 0) - R(2) * v + SV(FQ, -1, 0, 0)) * c.dx1q;
  R d22 = (SV(FQ, 0, 1, 0) - R(2) * v + SV(FQ, 0, -1, 0)) * c.dx2q;
  R d33 = (SV(FQ, 0, 0, 1) - R(2) * v + SV(FQ, 0, 0, -1)) * c.dx3q;
  R dc = d11 + d22 + d33;
  R dp = (SV(FP, 0, 0, 0) - SV(FP, 0, -1, 0)) * (c.dx2 * c.al);
  return (c.ga * h + c.ro * ru + c.alre * dc - dp) * c.dt;
}
// invtr3.f90:35-56
template <int FQ, typename R, typename S>
__device__ inline R rhs_w(const RhsCoef<R>& c, const S& s, R h, R ru) {
  R v = SV(FQ, 0, 0, 0);
  R d1 = (SV(FQ, -1, 0, 0) - R(2) * v + SV(FQ, 1, 0, 0)) * c.dx1q;
  R d2 = (SV(FQ, 0, -1, 0) - R(2) * v + SV(FQ, 0, 1, 0)) * c.dx2q;
  R d3 = (SV(FQ, 0, 0, 1) - R(2) * v + SV(FQ, 0, 0, -1)) * c.dx3q;
  R dc = d2 + d3 + d1;
  R dp = (SV(FP, 0, 0, 0) - SV(FP, 0, 0, -1)) * (c.al * c.dx3);
  return (c.ga * h + c.ro * ru + c.alre * dc - dp) * c.dt;
}
// invtrte.f90:32-53
template <int FQ, typename R, typename S>
__device__ inline R rhs_t(const RhsCoef<R>& c, const S& s, R h, R ru) {
  R v = SV(FQ, 0, 0, 0);
  R d1 = (SV(FQ, -1, 0, 0) - R(2) * v + SV(FQ, 1, 0, 0)) * c.dx1q;
  R d2 = (SV(FQ, 0, -1, 0) - R(2) * v + SV(FQ, 0, 1, 0)) * c.dx2q;
  R d3 = (SV(FQ, 0, 0, 1) - R(2) * v + SV(FQ, 0, 0, -1)) * c.dx3q;
  R dc = d2 + d3 + d1;
  return c.dt * (c.ga * h + c.ro * ru) + c.alpec * dc * c.dt;
}

template <typename R>
EulerArgs<R> make_args(melt_ctx* ctx) {
  EulerArgs<R> A;
  A.vx = fld<R>(ctx, MELT_VX); A.vy = fld<R>(ctx, MELT_VY); A.vz = fld<R>(ctx, MELT_VZ);
  A.temp = fld<R>(ctx, MELT_TEMP); A.pr = fld<R>(ctx, MELT_PR); A.vofp = fld<R>(ctx, MELT_VOFP);
  A.forcx = fld<R>(ctx, MELT_FORCX); A.forcy = fld<R>(ctx, MELT_FORCY); A.forcz = fld<R>(ctx, MELT_FORCZ);
  ctx->forc_free = 0;   // a pass that reads the forcing is being queued
  A.body = ctx->d_body;
  A.coord = ctx->d_coord;
  const bool faces = ctx->cfg.tagtype == 0;
  A.vofx = faces ? fld<R>(ctx, MELT_VOFX) : nullptr;
  A.vofy = faces ? fld<R>(ctx, MELT_VOFY) : nullptr;
  A.vofz = faces ? fld<R>(ctx, MELT_VOFZ) : nullptr;
  A.g = make_grid(ctx);
  A.rlen[0] = 1.0 / A.g.xlen; A.rlen[1] = 1.0 / A.g.ylen; A.rlen[2] = 1.0 / A.g.zlen;
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
  c.alpec = c.al / c.pec;
  return c;
}

// ------------------------------------------------------------- kernels
template <typename R>
__global__ __launch_bounds__(256) void k_hdnl(EulerArgs<R> A, R* __restrict__ out, int which, int nc) {
  unsigned b = xcd_block(blockIdx.x, gridDim.x);
  int c = (int)(b * blockDim.x + threadIdx.x);
  if (c >= nc) return;
  const auto s = ptr_stencil(A, make_cell(A.g, c));
  R p = SV(FVOF, 0, 0, 0);
  R r;
  if (which == 0) r = h1_val(A, s, A.vofx ? A.vofx[c] : R(0.5) * (SV(FVOF, -1, 0, 0) + p));
  else if (which == 1) r = h2_val(A, s, A.vofy ? A.vofy[c] : R(0.5) * (SV(FVOF, 0, -1, 0) + p));
  else if (which == 2) r = h3_val(A, s, A.vofz ? A.vofz[c] : R(0.5) * (SV(FVOF, 0, 0, -1) + p));
  else r = ht_val(A, s, p);
  out[c] = r;
}

template <typename R>
__global__ __launch_bounds__(256) void k_invtr_rhs(Grid g, RhsCoef<R> cf, const R* __restrict__ q, const R* __restrict__ pr,
                                                   const R* __restrict__ h, R* __restrict__ ru, R* __restrict__ rhs, int which, int nc) {
  unsigned b = xcd_block(blockIdx.x, gridDim.x);
  int c = (int)(b * blockDim.x + threadIdx.x);
  if (c >= nc) return;
  const Cell cl = make_cell(g, c);
  // the transported field q in every slot but the pressure's
  const PtrStencil<R, Cell> s{{q, q, q, q, pr, q, q, q, q}, cl, cl.i, cl.j, cl.k};
  R hv = h[c], rv = ru[c], r;
  if (which == 0) r = rhs_u<FU>(cf, s, hv, rv);
  else if (which == 1) r = rhs_v<FV>(cf, s, hv, rv);
  else if (which == 2) r = rhs_w<FW>(cf, s, hv, rv);
  else r = rhs_t<FT>(cf, s, hv, rv);
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

// x of element r of an LDS line L[0..M-1] padded with its periodic images
// L[-FIR_KMAX..-1] and L[M..M+FIR_KMAX-1], so no index wraps.
__device__ __forceinline__ double green_lds(const Green& G, const double* L, int r) {
  double x = G.g[0] * L[r];
  for (int m = 1; m <= G.K; ++m) x += G.g[m] * (L[r - m] + L[r + m]);
  return x;
}

// Explicit terms + RHS of all four transported fields fused with the x-line
// solve (pass A of the substep): a block owns W complete x-lines at one z
// plane, computes rhs x4 and ru x4 per cell (E cells per thread, reading the
// state through L1/L2), keeps the scaled rhs in LDS, solves the four
// x-systems there and writes the x-solved rhs.  Saves one read and one write
// of rhs x4 per substep against separate passes.
template <int E, typename R, bool FIR, bool FLAT>
__global__ __launch_bounds__(256) void k_explicit_xsolve(EulerArgs<R> A, RhsCoef<R> cf, R* __restrict__ ru1, R* __restrict__ ru2,
                                                         R* __restrict__ ru3, R* __restrict__ rut, R* __restrict__ r_u, R* __restrict__ r_v,
                                                         R* __restrict__ r_w, R* __restrict__ r_t, double bxu, double bxt, double sc_u,
                                                         double sc_t, int W, Green gu, Green gt) {
  extern __shared__ double lds[];
  const Grid& g = A.g;
  // LDS line stride: FIR lines carry FIR_KMAX periodic images on each side
  constexpr int H = FIR ? FIR_KMAX : 0;
  const int M = g.n1, ldx = FIR ? M + 2 * H : M + 1, tid = threadIdx.x, nt = blockDim.x;
  unsigned t = xcd_block(blockIdx.x, gridDim.x);
  int jt, k;
  tiled_2d(t, g.n2 / W, g.n3, 16, 8, jt, k);
  const int j0 = jt * W;
  const int base = M * (j0 + g.n2 * k);
  const int plane = W * ldx;
  const bool one_line = W == 1;   // then E * nt == M: idx is the x index
#pragma unroll 1
  for (int q = 0; q < E; ++q) {
    int idx = tid + q * nt;
    int r = one_line ? idx : idx % M, c = one_line ? 0 : idx / M;
    const int l = c * ldx + H + r;
    auto cell = [&](const auto& cl) {
      const auto s = ptr_stencil(A, cl);
      R p = SV(FVOF, 0, 0, 0);
      R hu = h1_val(A, s, A.vofx ? cl.ld(A.vofx, 0, 0, 0) : R(0.5) * (SV(FVOF, -1, 0, 0) + p));
      R hv = h2_val(A, s, A.vofy ? cl.ld(A.vofy, 0, 0, 0) : R(0.5) * (SV(FVOF, 0, -1, 0) + p));
      R hw = h3_val(A, s, A.vofz ? cl.ld(A.vofz, 0, 0, 0) : R(0.5) * (SV(FVOF, 0, 0, -1) + p));
      R ht = ht_val(A, s, p);
      R o1 = cl.ref(ru1), o2 = cl.ref(ru2), o3 = cl.ref(ru3), ot = cl.ref(rut);
      const double d[4] = {(double)rhs_u<FU>(cf, s, hu, o1) * sc_u, (double)rhs_v<FV>(cf, s, hv, o2) * sc_u,
                           (double)rhs_w<FW>(cf, s, hw, o3) * sc_u, (double)rhs_t<FT>(cf, s, ht, ot) * sc_t};
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        lds[f * plane + l] = d[f];
        if constexpr (FIR) {
          if (r < H) lds[f * plane + l + M] = d[f];
          if (r >= M - H) lds[f * plane + l - M] = d[f];
        }
      }
      cl.ref(ru1) = hu; cl.ref(ru2) = hv; cl.ref(ru3) = hw; cl.ref(rut) = ht;
    };
    if constexpr (FLAT) cell(fcell_at<R>(g, r, j0 + c, k));
    else cell(cell_at(g, r, j0 + c, k));
  }
  __syncthreads();
  R* outs[4] = {r_u, r_v, r_w, r_t};
  if constexpr (FIR) {
    const size_t p0 = (size_t)k * (size_t)M * g.n2 * sizeof(R);
#pragma unroll
    for (int q = 0; q < E; ++q) {
      int idx = tid + q * nt;
      int r = one_line ? idx : idx % M, c = one_line ? 0 : idx / M;
      const double* L = lds + c * ldx + H;
      const unsigned off = (unsigned)(r + M * (j0 + c)) * (unsigned)sizeof(R);
#pragma unroll
      for (int f = 0; f < 4; ++f)
        *reinterpret_cast<R*>(reinterpret_cast<char*>(outs[f]) + p0 + off) = (R)green_lds(f == 3 ? gt : gu, L + f * plane, r);
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

template <int...>
constexpr bool no_stencil_value = false;

// Stencil values of one cell of the z-marching pass: the centre column at
// k-1, k, k+1 rolls in registers, the rows j-1 and j+1 are loaded per plane,
// and the x-neighbours are read from the plane's rows published in LDS.
template <typename R>
struct ZStencil {
  R cu[3], cv[3], cw[3], ct[3], cp[2], co[2];           // (0, 0, d) at [d + 1]; p, VOFp: k-1, k
  R ujm, ujp, vjm, vjp, vjp_km, wjm, wjp, wjm_kp, tjm, tjp, pjm, ojm, fx, fy, fz;
  R u_im, u_ip, ujm_ip, ukm_ip, v_im, v_ip, vjp_im, w_im, w_ip, wkp_im, t_im, t_ip, p_im, o_im;
  int i, j, k;
  template <int F, int a, int b, int d>
  __device__ __forceinline__ R v() const {
    constexpr int key = (a + 1) * 9 + (b + 1) * 3 + (d + 1);
    constexpr int C = 13, XM = 4, XP = 22, YM = 10, YP = 16;   // keys of (0,0,0), (-1,0,0), (1,0,0), (0,-1,0), (0,1,0)
    if constexpr ((F == FU || F == FV || F == FW || F == FT) && a == 0 && b == 0) {
      return F == FU ? cu[d + 1] : F == FV ? cv[d + 1] : F == FW ? cw[d + 1] : ct[d + 1];
    } else if constexpr (F == FU) {
      if constexpr (key == XP) return u_ip;
      else if constexpr (key == XM) return u_im;
      else if constexpr (key == YM) return ujm;
      else if constexpr (key == YP) return ujp;
      else if constexpr (a == 1 && b == -1 && d == 0) return ujm_ip;
      else if constexpr (a == 1 && b == 0 && d == -1) return ukm_ip;
      else static_assert(no_stencil_value<F, a, b, d>);
    } else if constexpr (F == FV) {
      if constexpr (key == XP) return v_ip;
      else if constexpr (key == XM) return v_im;
      else if constexpr (key == YM) return vjm;
      else if constexpr (key == YP) return vjp;
      else if constexpr (a == -1 && b == 1 && d == 0) return vjp_im;
      else if constexpr (a == 0 && b == 1 && d == -1) return vjp_km;
      else static_assert(no_stencil_value<F, a, b, d>);
    } else if constexpr (F == FW) {
      if constexpr (key == XP) return w_ip;
      else if constexpr (key == XM) return w_im;
      else if constexpr (key == YM) return wjm;
      else if constexpr (key == YP) return wjp;
      else if constexpr (a == -1 && b == 0 && d == 1) return wkp_im;
      else if constexpr (a == 0 && b == -1 && d == 1) return wjm_kp;
      else static_assert(no_stencil_value<F, a, b, d>);
    } else if constexpr (F == FT) {
      if constexpr (key == XP) return t_ip;
      else if constexpr (key == XM) return t_im;
      else if constexpr (key == YM) return tjm;
      else if constexpr (key == YP) return tjp;
      else static_assert(no_stencil_value<F, a, b, d>);
    } else if constexpr (F == FP || F == FVOF) {
      if constexpr (key == C) return F == FP ? cp[1] : co[1];
      else if constexpr (a == 0 && b == 0 && d == -1) return F == FP ? cp[0] : co[0];
      else if constexpr (key == XM) return F == FP ? p_im : o_im;
      else if constexpr (key == YM) return F == FP ? pjm : ojm;
      else static_assert(no_stencil_value<F, a, b, d>);
    } else if constexpr (key == C) {
      return F == FFX ? fx : F == FFY ? fy : fz;
    } else {
      static_assert(no_stencil_value<F, a, b, d>);
    }
  }
};

// One cell of pass A: the explicit terms H of u, v, w, T (hdnl1-3, hdnlte)
// and the four right-hand sides scaled for the x-line solve.  o: ru at the
// cell; h receives the new ru values, d the scaled rhs.
template <typename R, typename S>
__device__ __forceinline__ void zcell(const EulerArgs<R>& A, const RhsCoef<R>& cf, const S& s, R o1, R o2, R o3, R ot, double sc_u,
                                      double sc_t, R h[4], double d[4]) {
  R p = SV(FVOF, 0, 0, 0);
  h[0] = h1_val(A, s, R(0.5) * (SV(FVOF, -1, 0, 0) + p));
  h[1] = h2_val(A, s, R(0.5) * (SV(FVOF, 0, -1, 0) + p));
  h[2] = h3_val(A, s, R(0.5) * (SV(FVOF, 0, 0, -1) + p));
  h[3] = ht_val(A, s, p);
  d[0] = (double)rhs_u<FU>(cf, s, h[0], o1) * sc_u;
  d[1] = (double)rhs_v<FV>(cf, s, h[1], o2) * sc_u;
  d[2] = (double)rhs_w<FW>(cf, s, h[2], o3) * sc_u;
  d[3] = (double)rhs_t<FT>(cf, s, h[3], ot) * sc_t;
}

// Pass A as a z-march: a block owns one x-line (one thread per i) of row j and
// walks D consecutive planes.  Per plane a cell loads the centre column at k+1
// (u, v, w, T) and at k (p, VOFp), the rows j-1 / j+1 it needs (u, v, w, T at
// k; w at j-1, k+1; p and VOFp at j-1), its forcing and ru: 23 loads instead
// of the 57 of a free-standing stencil, the rest rolling in registers or
// coming from LDS rows (two sets, alternating per plane, so one barrier
// separates writing a set from reading it).  Each plane then runs the x-line
// filter like k_explicit_xsolve.  Every value and expression is the same as
// there, so the results are identical.  Off: 32-bit (fields < 4 GiB) or
// 64-bit byte offsets.
template <typename R, typename Off>
__global__ __launch_bounds__(512) void k_explicit_zmarch(EulerArgs<R> A, RhsCoef<R> cf, R* __restrict__ ru1, R* __restrict__ ru2,
                                                         R* __restrict__ ru3, R* __restrict__ rut, R* __restrict__ r_u, R* __restrict__ r_v,
                                                         R* __restrict__ r_w, R* __restrict__ r_t, double sc_u, double sc_t, int D, Green gu,
                                                         Green gt) {
  extern __shared__ double lds[];
  constexpr int H = FIR_KMAX;
  enum { XU, XUJM, XV, XVJP, XWP, XT, XP, XO, NX };   // exchange rows of one set
  const Grid& g = A.g;
  const int M = g.n1, i = threadIdx.x, ldx = M + 2 * H;
  R* xr = reinterpret_cast<R*>(lds + 4 * ldx);
  auto row = [&](int set, int r) { return xr + (set * NX + r) * M; };
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const int j = (int)(b % (unsigned)g.n2), k0 = (int)(b / (unsigned)g.n2) * D;
  const int jm = j == 0 ? g.n2 - 1 : j - 1, jp = j == g.n2 - 1 ? 0 : j + 1;
  const int im = i == 0 ? M - 1 : i - 1, ip = i == M - 1 ? 0 : i + 1;
  auto off = [&](int jj, int kk) { return ((Off)i + (Off)M * ((Off)jj + (Off)g.n2 * (Off)kk)) * (Off)sizeof(R); };
  auto ld = [](const R* p, Off o) { return *reinterpret_cast<const R*>(reinterpret_cast<const char*>(p) + o); };
  auto st = [](R* p, Off o, R v) { *reinterpret_cast<R*>(reinterpret_cast<char*>(p) + o) = v; };
  // values read once (forcing, ru) and every store stream past the caches
  // (nontemporal), which leaves the L2 to the rows the neighbouring blocks
  // read again one plane later
  auto ldn = [](const R* p, Off o) { return __builtin_nontemporal_load(reinterpret_cast<const R*>(reinterpret_cast<const char*>(p) + o)); };
  auto stn = [](R* p, Off o, R v) { __builtin_nontemporal_store(v, reinterpret_cast<R*>(reinterpret_cast<char*>(p) + o)); };
  ZStencil<R> s;
  s.i = i; s.j = j;
  {
    const int km = k0 == 0 ? g.n3 - 1 : k0 - 1;
    const Off om = off(j, km), o0 = off(j, k0);
    s.cu[0] = ld(A.vx, om); s.cu[1] = ld(A.vx, o0);
    s.cv[0] = ld(A.vy, om); s.cv[1] = ld(A.vy, o0);
    s.cw[0] = ld(A.vz, om); s.cw[1] = ld(A.vz, o0);
    s.ct[0] = ld(A.temp, om); s.ct[1] = ld(A.temp, o0);
    s.cp[0] = ld(A.pr, om); s.co[0] = ld(A.vofp, om);
    s.vjp_km = ld(A.vy, off(jp, km));
    s.wjm = ld(A.vz, off(jm, k0));
    row(1, XU)[i] = s.cu[0];
    row(1, XWP)[i] = s.cw[1];
  }
  R* outs[4] = {r_u, r_v, r_w, r_t};
#pragma unroll 1
  for (int t = 0; t < D; ++t) {
    const int k = k0 + t, kp = k == g.n3 - 1 ? 0 : k + 1;
    s.k = k;
    const Off o0 = off(j, k), op = off(j, kp), ojm = off(jm, k), ojp = off(jp, k);
    s.cu[2] = ld(A.vx, op); s.cv[2] = ld(A.vy, op); s.cw[2] = ld(A.vz, op); s.ct[2] = ld(A.temp, op);
    s.cp[1] = ld(A.pr, o0); s.co[1] = ld(A.vofp, o0);
    s.ujm = ld(A.vx, ojm); s.ujp = ld(A.vx, ojp);
    s.vjm = ld(A.vy, ojm); s.vjp = ld(A.vy, ojp);
    s.wjp = ld(A.vz, ojp); s.wjm_kp = ld(A.vz, off(jm, kp));
    s.tjm = ld(A.temp, ojm); s.tjp = ld(A.temp, ojp);
    s.pjm = ld(A.pr, ojm); s.ojm = ld(A.vofp, ojm);
    s.fx = ldn(A.forcx, o0); s.fy = ldn(A.forcy, o0); s.fz = ldn(A.forcz, o0);
    const R o1 = ldn(ru1, o0), o2 = ldn(ru2, o0), o3 = ldn(ru3, o0), ot = ldn(rut, o0);
    const int cur = t & 1, prv = cur ^ 1;
    row(cur, XU)[i] = s.cu[1]; row(cur, XUJM)[i] = s.ujm; row(cur, XV)[i] = s.cv[1]; row(cur, XVJP)[i] = s.vjp;
    row(cur, XWP)[i] = s.cw[2]; row(cur, XT)[i] = s.ct[1]; row(cur, XP)[i] = s.cp[1]; row(cur, XO)[i] = s.co[1];
    __syncthreads();
    s.u_im = row(cur, XU)[im]; s.u_ip = row(cur, XU)[ip]; s.ujm_ip = row(cur, XUJM)[ip]; s.ukm_ip = row(prv, XU)[ip];
    s.v_im = row(cur, XV)[im]; s.v_ip = row(cur, XV)[ip]; s.vjp_im = row(cur, XVJP)[im];
    s.w_im = row(prv, XWP)[im]; s.w_ip = row(prv, XWP)[ip]; s.wkp_im = row(cur, XWP)[im];
    s.t_im = row(cur, XT)[im]; s.t_ip = row(cur, XT)[ip];
    s.p_im = row(cur, XP)[im]; s.o_im = row(cur, XO)[im];
    R hh[4];
    double d[4];
    zcell(A, cf, s, o1, o2, o3, ot, sc_u, sc_t, hh, d);
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      double* L = lds + f * ldx + H;
      L[i] = d[f];
      if (i < H) L[i + M] = d[f];
      if (i >= M - H) L[i - M] = d[f];
    }
    stn(ru1, o0, hh[0]); stn(ru2, o0, hh[1]); stn(ru3, o0, hh[2]); stn(rut, o0, hh[3]);
    __syncthreads();
#pragma unroll
    for (int f = 0; f < 4; ++f) stn(outs[f], o0, (R)green_lds(f == 3 ? gt : gu, lds + f * ldx + H, i));
    s.cu[0] = s.cu[1]; s.cu[1] = s.cu[2]; s.cv[0] = s.cv[1]; s.cv[1] = s.cv[2];
    s.cw[0] = s.cw[1]; s.cw[1] = s.cw[2]; s.ct[0] = s.ct[1]; s.ct[1] = s.ct[2];
    s.cp[0] = s.cp[1]; s.co[0] = s.co[1];
    s.vjp_km = s.vjp; s.wjm = s.wjm_kp;
  }
}

// The z-march with two x-neighbour cells per thread (cells 2t and 2t+1), for
// 4-byte values: every field access is one 8-byte load or store, half the
// vector memory instructions of the one-cell form, which is what bounds this
// pass.  A thread's second cell gives its first cell the +x values and the
// first gives the second the -x values; the rest come from the neighbouring
// threads' pairs published in LDS as in the one-cell form.  Same values and
// expressions per cell, so the results are identical to it.  Block: n1 / 2
// threads.
template <typename Off>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_explicit_zmarch_pair(EulerArgs<float> A, RhsCoef<float> cf, float* __restrict__ ru1,
                                                              float* __restrict__ ru2, float* __restrict__ ru3, float* __restrict__ rut,
                                                              float* __restrict__ r_u, float* __restrict__ r_v, float* __restrict__ r_w,
                                                              float* __restrict__ r_t, double sc_u, double sc_t, int D, Green gu, Green gt) {
  using R = float;
  typedef float f2 __attribute__((ext_vector_type(2)));
  typedef double d2 __attribute__((ext_vector_type(2)));
  extern __shared__ double lds[];
  constexpr int H = FIR_KMAX;
  enum { XU, XUJM, XV, XVJP, XWP, XT, XP, XO, NX };   // exchange rows of one set, one pair per thread
  const Grid& g = A.g;
  const int M = g.n1, T = M >> 1, t = threadIdx.x, ldx = M + 2 * H;
  f2* xr = reinterpret_cast<f2*>(lds + 4 * ldx);
  auto row = [&](int set, int r) { return xr + (set * NX + r) * T; };
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const int j = (int)(b % (unsigned)g.n2), k0 = (int)(b / (unsigned)g.n2) * D;
  const int jm = j == 0 ? g.n2 - 1 : j - 1, jp = j == g.n2 - 1 ? 0 : j + 1;
  const int tm = t == 0 ? T - 1 : t - 1, tp = t == T - 1 ? 0 : t + 1;
  const int i0 = 2 * t;
  auto off = [&](int jj, int kk) { return ((Off)i0 + (Off)M * ((Off)jj + (Off)g.n2 * (Off)kk)) * (Off)sizeof(R); };
  auto ld = [](const R* p, Off o) { return *reinterpret_cast<const f2*>(reinterpret_cast<const char*>(p) + o); };
  auto st = [](R* p, Off o, f2 v) { __builtin_nontemporal_store(v, reinterpret_cast<f2*>(reinterpret_cast<char*>(p) + o)); };
  auto ldn = [](const R* p, Off o) { return __builtin_nontemporal_load(reinterpret_cast<const f2*>(reinterpret_cast<const char*>(p) + o)); };
  ZStencil<R> a, c;   // cells 2t and 2t+1
  a.i = i0; c.i = i0 + 1; a.j = c.j = j;
#define PAIR_LD(field, expr) { const f2 v_ = (expr); a.field = v_.x; c.field = v_.y; }
  {
    const int km = k0 == 0 ? g.n3 - 1 : k0 - 1;
    const Off om = off(j, km), o0 = off(j, k0);
    PAIR_LD(cu[0], ld(A.vx, om)); PAIR_LD(cu[1], ld(A.vx, o0));
    PAIR_LD(cv[0], ld(A.vy, om)); PAIR_LD(cv[1], ld(A.vy, o0));
    PAIR_LD(cw[0], ld(A.vz, om)); PAIR_LD(cw[1], ld(A.vz, o0));
    PAIR_LD(ct[0], ld(A.temp, om)); PAIR_LD(ct[1], ld(A.temp, o0));
    PAIR_LD(cp[0], ld(A.pr, om)); PAIR_LD(co[0], ld(A.vofp, om));
    PAIR_LD(vjp_km, ld(A.vy, off(jp, km)));
    PAIR_LD(wjm, ld(A.vz, off(jm, k0)));
    row(1, XU)[t] = f2{a.cu[0], c.cu[0]};
    row(1, XWP)[t] = f2{a.cw[1], c.cw[1]};
  }
  R* outs[4] = {r_u, r_v, r_w, r_t};
#pragma unroll 1
  for (int q = 0; q < D; ++q) {
    const int k = k0 + q, kp = k == g.n3 - 1 ? 0 : k + 1;
    a.k = c.k = k;
    const Off o0 = off(j, k), op = off(j, kp), ojm = off(jm, k), ojp = off(jp, k);
    PAIR_LD(cu[2], ld(A.vx, op)); PAIR_LD(cv[2], ld(A.vy, op)); PAIR_LD(cw[2], ld(A.vz, op)); PAIR_LD(ct[2], ld(A.temp, op));
    PAIR_LD(cp[1], ld(A.pr, o0)); PAIR_LD(co[1], ld(A.vofp, o0));
    PAIR_LD(ujm, ld(A.vx, ojm)); PAIR_LD(ujp, ld(A.vx, ojp));
    PAIR_LD(vjm, ld(A.vy, ojm)); PAIR_LD(vjp, ld(A.vy, ojp));
    PAIR_LD(wjp, ld(A.vz, ojp)); PAIR_LD(wjm_kp, ld(A.vz, off(jm, kp)));
    PAIR_LD(tjm, ld(A.temp, ojm)); PAIR_LD(tjp, ld(A.temp, ojp));
    PAIR_LD(pjm, ld(A.pr, ojm)); PAIR_LD(ojm, ld(A.vofp, ojm));
    PAIR_LD(fx, ldn(A.forcx, o0)); PAIR_LD(fy, ldn(A.forcy, o0)); PAIR_LD(fz, ldn(A.forcz, o0));
    const f2 o1 = ldn(ru1, o0), o2 = ldn(ru2, o0), o3 = ldn(ru3, o0), ot = ldn(rut, o0);
    const int cur = q & 1, prv = cur ^ 1;
    row(cur, XU)[t] = f2{a.cu[1], c.cu[1]}; row(cur, XUJM)[t] = f2{a.ujm, c.ujm};
    row(cur, XV)[t] = f2{a.cv[1], c.cv[1]}; row(cur, XVJP)[t] = f2{a.vjp, c.vjp};
    row(cur, XWP)[t] = f2{a.cw[2], c.cw[2]}; row(cur, XT)[t] = f2{a.ct[1], c.ct[1]};
    row(cur, XP)[t] = f2{a.cp[1], c.cp[1]}; row(cur, XO)[t] = f2{a.co[1], c.co[1]};
    __syncthreads();
    // cell 2t: -x from the pair of thread t-1, +x from cell 2t+1
    a.u_im = row(cur, XU)[tm].y; a.u_ip = c.cu[1]; a.ujm_ip = c.ujm; a.ukm_ip = c.cu[0];
    a.v_im = row(cur, XV)[tm].y; a.v_ip = c.cv[1]; a.vjp_im = row(cur, XVJP)[tm].y;
    a.w_im = row(prv, XWP)[tm].y; a.w_ip = c.cw[1]; a.wkp_im = row(cur, XWP)[tm].y;
    a.t_im = row(cur, XT)[tm].y; a.t_ip = c.ct[1];
    a.p_im = row(cur, XP)[tm].y; a.o_im = row(cur, XO)[tm].y;
    // cell 2t+1: -x from cell 2t, +x from the pair of thread t+1
    c.u_im = a.cu[1]; c.u_ip = row(cur, XU)[tp].x; c.ujm_ip = row(cur, XUJM)[tp].x; c.ukm_ip = row(prv, XU)[tp].x;
    c.v_im = a.cv[1]; c.v_ip = row(cur, XV)[tp].x; c.vjp_im = a.vjp;
    c.w_im = a.cw[1]; c.w_ip = row(prv, XWP)[tp].x; c.wkp_im = a.cw[2];
    c.t_im = a.ct[1]; c.t_ip = row(cur, XT)[tp].x;
    c.p_im = a.cp[1]; c.o_im = a.co[1];
    R ha[4], hc[4];
    double da[4], dc[4];
    zcell(A, cf, a, o1.x, o2.x, o3.x, ot.x, sc_u, sc_t, ha, da);
    zcell(A, cf, c, o1.y, o2.y, o3.y, ot.y, sc_u, sc_t, hc, dc);
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      double* L = lds + f * ldx + H;
      *reinterpret_cast<d2*>(L + i0) = d2{da[f], dc[f]};
      if (i0 < H) { L[i0 + M] = da[f]; L[i0 + 1 + M] = dc[f]; }   // H even: both or neither
      if (i0 >= M - H) { L[i0 - M] = da[f]; L[i0 + 1 - M] = dc[f]; }
    }
    st(ru1, o0, f2{ha[0], hc[0]}); st(ru2, o0, f2{ha[1], hc[1]}); st(ru3, o0, f2{ha[2], hc[2]}); st(rut, o0, f2{ha[3], hc[3]});
    __syncthreads();
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const Green& G = f == 3 ? gt : gu;
      const double* L = lds + f * ldx + H;
      st(outs[f], o0, f2{(R)green_lds(G, L, i0), (R)green_lds(G, L, i0 + 1)});
    }
    a.cu[0] = a.cu[1]; a.cu[1] = a.cu[2]; a.cv[0] = a.cv[1]; a.cv[1] = a.cv[2];
    a.cw[0] = a.cw[1]; a.cw[1] = a.cw[2]; a.ct[0] = a.ct[1]; a.ct[1] = a.ct[2];
    a.cp[0] = a.cp[1]; a.co[0] = a.co[1];
    a.vjp_km = a.vjp; a.wjm = a.wjm_kp;
    c.cu[0] = c.cu[1]; c.cu[1] = c.cu[2]; c.cv[0] = c.cv[1]; c.cv[1] = c.cv[2];
    c.cw[0] = c.cw[1]; c.cw[1] = c.cw[2]; c.ct[0] = c.ct[1]; c.ct[1] = c.ct[2];
    c.cp[0] = c.cp[1]; c.co[0] = c.co[1];
    c.vjp_km = c.vjp; c.wjm = c.wjm_kp;
  }
#undef PAIR_LD
}
static_assert(FIR_KMAX % 2 == 0, "the pair z-march writes the periodic images of a pair together");

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

// A block is 64 consecutive i (coalesced rows) x 4 segments of NCH chunks of
// GQ points of the line; each thread slides a register window of GQ + 2 KT
// values along its segment, GQ new values per chunk, so every input is read
// about (NCH GQ + 2 KT) / (NCH GQ) times (the halo rows of neighbouring
// segments mostly from L2).  Coefficients beyond the system's K are zero.
// Needs n1 % 64 == 0 and N % (4 GQ NCH) == 0.
constexpr int GQ = 16;

template <int DIR, typename R, int KT, int NCH>
__global__ __launch_bounds__(256) void k_green(GreenBatch<R> B, int n1, int n2, int n3) {
  const int f = blockIdx.y;
  const int N = DIR == 1 ? n2 : n3;
  const int s = DIR == 1 ? n1 : n1 * n2;
  const int nsg = N / (4 * GQ * NCH), nit = n1 / 64;
  unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const int sg = (int)(b % nsg), rest = (int)(b / nsg);
  const int it = rest % nit, other = rest / nit;   // other: k (DIR 1) or j (DIR 2)
  const int i = it * 64 + (threadIdx.x & 63);
  const int p0 = (sg * 4 + (threadIdx.x >> 6)) * GQ * NCH;
  const int base = i + (DIR == 1 ? n1 * n2 * other : n1 * other);
  const Green& G = B.gr[B.which[f]];
  double g[KT + 1];
#pragma unroll
  for (int m = 0; m <= KT; ++m) g[m] = m <= G.K ? G.g[m] : 0.0;
  const R* __restrict__ d = B.src[f];
  auto at = [&](int p) {
    p = p < 0 ? p + N : (p >= N ? p - N : p);
    return (double)d[base + p * s];
  };
  double v[GQ + 2 * KT];   // window of positions pc - KT .. pc + GQ + KT - 1
#pragma unroll
  for (int q = 0; q < GQ + 2 * KT; ++q) v[q] = at(p0 - KT + q);
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    const int pc = p0 + ch * GQ;
#pragma unroll
    for (int q = 0; q < GQ; ++q) {
      double x = g[0] * v[q + KT];
#pragma unroll
      for (int m = 1; m <= KT; ++m) x += g[m] * (v[q + KT - m] + v[q + KT + m]);
      const int c = base + (pc + q) * s;
      if (B.qadd[f]) B.qadd[f][c] = B.qadd[f][c] + (R)x;
      else B.dst[f][c] = (R)x;
    }
    if (ch + 1 < NCH) {   // slide by GQ: the last 2 KT values become the head
#pragma unroll
      for (int q = 0; q < 2 * KT; ++q) v[q] = v[q + GQ];
#pragma unroll
      for (int q = 0; q < GQ; ++q) v[2 * KT + q] = at(pc + GQ + KT + q);
    }
  }
}

// y and z solves of one field in one pass, q += Gz * (Gy * rhs) (solxj.f90,
// solxk.f90:44-50): the y filter of every plane is taken from an LDS tile and
// fed to the z filter's register window, so the y result never goes to HBM.
template <typename R>
struct GreenYZ {
  const R* src[4];
  R* qadd[4];
  int which[4];
  Green gy[2], gz[2];
};

// C consecutive values of a field (C = 1, or 2 as one 8-byte access).
template <typename R, int C>
__device__ __forceinline__ void ld_c(const R* p, R (&o)[C]) {
  if constexpr (C == 2) {
    const HIP_vector_type<R, 2> t = *reinterpret_cast<const HIP_vector_type<R, 2>*>(p);
    o[0] = t.x; o[1] = t.y;
  } else {
    o[0] = *p;
  }
}
template <typename R, int C>
__device__ __forceinline__ void st_c(R* p, const R (&o)[C]) {
  if constexpr (C == 2) {
    HIP_vector_type<R, 2> t;
    t.x = o[0]; t.y = o[1];
    *reinterpret_cast<HIP_vector_type<R, 2>*>(p) = t;
  } else {
    *p = o[0];
  }
}
// The same for values a kernel reads or writes once: nontemporal accesses,
// which stream past the caches and leave them to the values that neighbouring
// blocks read again.
template <typename R, int C>
__device__ __forceinline__ void ld_c_nt(const R* p, R (&o)[C]) {
  if constexpr (C == 2) {
    typedef R v2 __attribute__((ext_vector_type(2)));
    const v2 t = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    o[0] = t.x; o[1] = t.y;
  } else {
    o[0] = __builtin_nontemporal_load(p);
  }
}
template <typename R, int C>
__device__ __forceinline__ void st_c_nt(R* p, const R (&o)[C]) {
  if constexpr (C == 2) {
    typedef R v2 __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store(v2{o[0], o[1]}, reinterpret_cast<v2*>(p));
  } else {
    __builtin_nontemporal_store(o[0], p);
  }
}

// A block is 64 C consecutive i x JB rows j, C (i, j) columns per thread (C =
// 2 for 4-byte values, so a wave still moves 512-byte rows), and marches
// through one z segment of seg planes (plus KT halo planes each side).  Per
// plane the block stages rows j0-KT .. j0+JB+KT-1 in one of two LDS tiles
// (one barrier per plane), each thread filters its columns along y, rounds to
// R (the value the separate y pass would store) and shifts the result into a
// register window of 2 KT + 1 planes per column, on which the z filter runs
// as in k_green.  The loads run PF planes ahead of the plane being filtered
// (a ring of PF register sets, indexed statically by unrolling PF planes at a
// time).  Needs n1 % (64 C) == 0, n2 % JB == 0, n3 % seg == 0, seg % PF == 0.
template <typename R, int KT, int JB, int C>
__global__ __launch_bounds__(64 * JB) void k_green_yz(GreenYZ<R> B, int n1, int n2, int n3, int seg) {
  constexpr int ROWS = JB + 2 * KT, RPT = (ROWS + JB - 1) / JB, W = 2 * KT + 1, PF = 2, RW = 64 * C;
  static_assert((2 * KT) % PF == 0, "prefetch ring must tile the halo planes");
  __shared__ R tile[2][ROWS * RW];
  const int f = blockIdx.y;
  const int nsg = n3 / seg, njt = n2 / JB;
  unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const int sg = (int)(b % nsg), rest = (int)(b / nsg);
  const int jt = rest % njt, it = rest / njt;
  const int lane = threadIdx.x & 63, jl = threadIdx.x >> 6;
  const int i = it * RW + C * lane, j0 = jt * JB, p0 = sg * seg;
  const int wy = B.which[f];
  double gy[KT + 1], gz[KT + 1];
#pragma unroll
  for (int m = 0; m <= KT; ++m) {
    gy[m] = m <= B.gy[wy].K ? B.gy[wy].g[m] : 0.0;
    gz[m] = m <= B.gz[wy].K ? B.gz[wy].g[m] : 0.0;
  }
  // no __restrict__: loads of src must stay behind the barriers, or the
  // compiler hoists the prefetches of many planes and runs out of registers
  const R* d = B.src[f];
  R* q = B.qadd[f];
  const size_t plane = (size_t)n1 * n2;
  int roff[RPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    int jj = j0 - KT + jl + r * JB;
    jj = jj < 0 ? jj + n2 : (jj >= n2 ? jj - n2 : jj);
    roff[r] = jj * n1 + i;
  }
  R nx[PF][RPT][C];
  auto fetch = [&](int slot, int p) {
    p = p < 0 ? p + n3 : (p >= n3 ? p - n3 : p);
    const R* dp = d + (size_t)p * plane;
#pragma unroll
    for (int r = 0; r < RPT; ++r)
      if (jl + r * JB < ROWS) ld_c<R, C>(dp + roff[r], nx[slot][r]);
  };
  int buf = 0;
  int pn = p0 - KT;   // next plane to filter; its values sit in ring slot (pn - p0 + KT) % PF
  const int pend = p0 + seg + KT;
#pragma unroll
  for (int s = 0; s < PF; ++s) fetch(s, pn + s);
  double w[C][W];   // per column: y-filtered planes p - KT .. p + KT around the output plane p
  // shifts the windows by one plane and appends the y filter of plane pn for
  // this thread's columns, then refills the ring slot with plane pn + PF
  auto advance = [&](int slot) {
#pragma unroll
    for (int cc = 0; cc < C; ++cc)
#pragma unroll
      for (int m = 0; m + 1 < W; ++m) w[cc][m] = w[cc][m + 1];
    R* T = tile[buf];
#pragma unroll
    for (int r = 0; r < RPT; ++r)
      if (jl + r * JB < ROWS) st_c<R, C>(T + (jl + r * JB) * RW + C * lane, nx[slot][r]);
    __syncthreads();
    if (pn + PF < pend) fetch(slot, pn + PF);
#pragma unroll
    for (int cc = 0; cc < C; ++cc) {
      const R* L = T + (jl + KT) * RW + C * lane + cc;
      double x = gy[0] * (double)L[0];
#pragma unroll
      for (int m = 1; m <= KT; ++m) x += gy[m] * ((double)L[-m * RW] + (double)L[m * RW]);
      w[cc][W - 1] = (double)(R)x;
    }
    buf ^= 1;
    ++pn;
  };
#pragma unroll 1
  for (int s = 0; s < 2 * KT; s += PF) {
#pragma unroll
    for (int u = 0; u < PF; ++u) advance(u);
  }
  const size_t col = (size_t)j0 * n1 + (size_t)jl * n1 + i;
#pragma unroll 1
  for (int p = p0; p < p0 + seg; p += PF) {
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      advance(u);
      const size_t c = (size_t)(p + u) * plane + col;
      R qv[C];
      ld_c_nt<R, C>(q + c, qv);
#pragma unroll
      for (int cc = 0; cc < C; ++cc) {
        double x = gz[0] * w[cc][KT];
#pragma unroll
        for (int m = 1; m <= KT; ++m) x += gz[m] * (w[cc][KT - m] + w[cc][KT + m]);
        qv[cc] = qv[cc] + (R)x;
      }
      st_c_nt<R, C>(q + c, qv);
    }
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

// k_project for 4-byte values, two x-neighbour cells per thread (8-byte
// loads and stores, a wave moves 512-byte rows); the same expressions per
// cell.  Needs n1 even; cells in x-line order.
template <typename R>
__global__ __launch_bounds__(256) void k_project2(Grid g, R* __restrict__ vx, R* __restrict__ vy, R* __restrict__ vz, R* __restrict__ pr,
                                                  const R* __restrict__ dph, const double* __restrict__ mean, R udx1, R udx2, R udx3, R be,
                                                  R dx1q, R dx2q, R dx3q, int do_u, int do_p, int nc, const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const int c = 2 * (int)(b * blockDim.x + threadIdx.x);
  if (c >= nc) return;
  const int n1 = g.n1, n2 = g.n2, n3 = g.n3, pl = n1 * n2;
  const int i = c % n1, r = c / n1, j = r % n2, k = r / n2;
  const int cjm = j == 0 ? c + (n2 - 1) * n1 : c - n1, cjp = j == n2 - 1 ? c - (n2 - 1) * n1 : c + n1;
  const int ckm = k == 0 ? c + (n3 - 1) * pl : c - pl, ckp = k == n3 - 1 ? c - (n3 - 1) * pl : c + pl;
  const int cim = i == 0 ? c + n1 - 1 : c - 1, cip2 = i + 2 == n1 ? c + 2 - n1 : c + 2;
  R l[2], pjm[2], pkm[2];
  ld_c<R, 2>(dph + c, l); ld_c<R, 2>(dph + cjm, pjm); ld_c<R, 2>(dph + ckm, pkm);
  const R pim[2] = {dph[cim], l[0]};
  if (do_u) {
    R u[2], v[2], w[2];
    ld_c_nt<R, 2>(vx + c, u); ld_c_nt<R, 2>(vy + c, v); ld_c_nt<R, 2>(vz + c, w);
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      if (mean) { u[e] = u[e] - R(mean[0]); v[e] = v[e] - R(mean[1]); w[e] = w[e] - R(mean[2]); }
      u[e] = u[e] - (l[e] - pim[e]) * udx1;
      v[e] = v[e] - (l[e] - pjm[e]) * udx2;
      w[e] = w[e] - (l[e] - pkm[e]) * udx3;
    }
    st_c_nt<R, 2>(vx + c, u); st_c_nt<R, 2>(vy + c, v); st_c_nt<R, 2>(vz + c, w);
  }
  if (do_p) {
    R pjp[2], pkp[2], p[2];
    ld_c<R, 2>(dph + cjp, pjp); ld_c<R, 2>(dph + ckp, pkp); ld_c_nt<R, 2>(pr + c, p);
    const R pip[2] = {l[1], dph[cip2]};
#pragma unroll
    for (int e = 0; e < 2; ++e)
      p[e] = p[e] + l[e] - be * ((pip[e] - R(2) * l[e] + pim[e]) * dx1q + (pjp[e] - R(2) * l[e] + pjm[e]) * dx2q + (pkp[e] - R(2) * l[e] + pkm[e]) * dx3q);
    st_c_nt<R, 2>(pr + c, p);
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

// Max reductions of cfl.f90:9-25 (what = 0), divgck divg.f90:36-50
// (what = 1) and get_dt's advective measure cfl.f90:48-51 (what = 2, with dt),
// in FP64 arithmetic whatever the storage type; the result (non-negative) is
// combined by atomicMax on its bit pattern.
template <typename R>
__global__ __launch_bounds__(256) void k_reduce_max(Grid g, const R* __restrict__ vx, const R* __restrict__ vy, const R* __restrict__ vz,
                                                    int what, double dt, int nc, unsigned long long* slots) {
  // V consecutive x cells per thread through 16-byte loads (n1 % V == 0)
  constexpr int V = 16 / sizeof(R);
  struct alignas(16) Vec { R v[V]; };
  double m = what == 0 ? 0.00000001 : 0.0;
  const int c = (int)(xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) * V;
  if (c < nc) {
    const Cell s = make_cell(g, c);
    const Vec X = *reinterpret_cast<const Vec*>(vx + c), Y0 = *reinterpret_cast<const Vec*>(vy + c);
    const Vec Y1 = *reinterpret_cast<const Vec*>(vy + s.at(0, 1, 0)), Z0 = *reinterpret_cast<const Vec*>(vz + c);
    const Vec Z1 = *reinterpret_cast<const Vec*>(vz + s.at(0, 0, 1));
    const R xl = s.i + V == g.n1 ? vx[c + V - g.n1] : vx[c + V];   // x neighbour of the last cell
#pragma unroll
    for (int e = 0; e < V; ++e) {
      double x0 = X.v[e], x1 = e + 1 < V ? X.v[e + 1] : xl, y0 = Y0.v[e], y1 = Y1.v[e], z0 = Z0.v[e], z1 = Z1.v[e];
      double q;
      if (what == 0)
        q = fabs((x0 + x1) * 0.5 * g.dx1) + fabs((y0 + y1) * 0.5 * g.dx2) + fabs((z0 + z1) * 0.5 * g.dx3);
      else if (what == 2)
        q = fabs(x0) * dt * g.dx1 + fabs(y0) * dt * g.dx2 + fabs(z0) * dt * g.dx3;
      else
        q = fabs((x1 - x0) * g.dx1 + (y1 - y0) * g.dx2 + (z1 - z0) * g.dx3);
      m = fmax(m, q);
    }
  }
  block_max_to_slots(m, slots);
}

// The REDUCE_SLOTS partial maxima into *out; the slots are cleared for the
// next reduction.  One block of REDUCE_SLOTS threads.
__global__ __launch_bounds__(REDUCE_SLOTS) void k_fold_slots(unsigned long long* __restrict__ slots, unsigned long long* __restrict__ out) {
  __shared__ unsigned long long sm[REDUCE_SLOTS];
  const int t = threadIdx.x;
  sm[t] = slots[t];
  slots[t] = 0ull;
  __syncthreads();
  for (int w = REDUCE_SLOTS / 2; w > 0; w >>= 1) {
    if (t < w) sm[t] = sm[t] > sm[t + w] ? sm[t] : sm[t + w];
    __syncthreads();
  }
  if (t == 0) *out = sm[0];
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
  const double tol = ctx->fp32 ? 1e-8 : 2e-17;
  const double sc_u = 1.0 / (1.0 + 2.0 * bxu), sc_t = 1.0 / (1.0 + 2.0 * bxt);   // solxi.f90:33-37
  Green gu{}, gt{};   // the LDS line holds the scaled rhs, so the filter is the unscaled one
  const bool fir = !ctx->no_green && green_coeffs(bxu, ctx->n1, tol, 1.0, gu) && green_coeffs(bxt, ctx->n1, tol, 1.0, gt);
  R* r4[4] = {static_cast<R*>(ctx->rhs4[RHS_U]), static_cast<R*>(ctx->rhs4[RHS_V]), static_cast<R*>(ctx->rhs4[RHS_W]),
              static_cast<R*>(ctx->rhs4[RHS_T])};
  // z-march (one x-line per block, D planes) unless line-tile shapes are asked for
  const size_t zshm = sizeof(double) * 4 * (ctx->n1 + 2 * FIR_KMAX) + sizeof(R) * 16 * (size_t)ctx->n1;
  // (with geometric face fractions, tagType 0, the line tiles read them)
  if (fir && ctx->tune[TUNE_XF_W] == 0 && ctx->tune[TUNE_XF_NT] == 0 && ctx->n1 % 64 == 0 && ctx->n1 <= 512 && zshm <= ctx->lds_max &&
      ctx->cfg.tagtype != 0) {
    int D = ctx->tune[TUNE_XF_Z];
    if (D <= 0 || ctx->n3 % D) {
      D = 1;
      // 8 planes per block measured best for one cell per thread (FP32 1-2%,
      // FP64 level or better), 4 for the two-cell form (1%)
      const bool pair = sizeof(R) == 4 && ctx->tune[TUNE_XF_ONE] == 0;
      for (int d : {8, 4, 2})
        if (!(pair && d == 8) && ctx->n3 % d == 0) { D = d; break; }
    }
    const unsigned blocks = (unsigned)((size_t)ctx->n2 * (ctx->n3 / D));
    if constexpr (std::is_same_v<R, float>) {
      if (ctx->tune[TUNE_XF_ONE] == 0) {   // two cells per thread, 8-byte accesses
        if (ctx->nc * sizeof(R) < (size_t(1) << 32))
          hipLaunchKernelGGL((k_explicit_zmarch_pair<unsigned>), dim3(blocks), dim3(ctx->n1 / 2), zshm, ctx->stream, A, cf,
                             fld<R>(ctx, MELT_RU1), fld<R>(ctx, MELT_RU2), fld<R>(ctx, MELT_RU3), fld<R>(ctx, MELT_RUT), r4[0], r4[1],
                             r4[2], r4[3], sc_u, sc_t, D, gu, gt);
        else
          hipLaunchKernelGGL((k_explicit_zmarch_pair<size_t>), dim3(blocks), dim3(ctx->n1 / 2), zshm, ctx->stream, A, cf,
                             fld<R>(ctx, MELT_RU1), fld<R>(ctx, MELT_RU2), fld<R>(ctx, MELT_RU3), fld<R>(ctx, MELT_RUT), r4[0], r4[1],
                             r4[2], r4[3], sc_u, sc_t, D, gu, gt);
        MELT_HIP(hipGetLastError());
        return 0;
      }
    }
    if (ctx->nc * sizeof(R) < (size_t(1) << 32))
      hipLaunchKernelGGL((k_explicit_zmarch<R, unsigned>), dim3(blocks), dim3(ctx->n1), zshm, ctx->stream, A, cf, fld<R>(ctx, MELT_RU1),
                         fld<R>(ctx, MELT_RU2), fld<R>(ctx, MELT_RU3), fld<R>(ctx, MELT_RUT), r4[0], r4[1], r4[2], r4[3], sc_u, sc_t, D, gu, gt);
    else
      hipLaunchKernelGGL((k_explicit_zmarch<R, size_t>), dim3(blocks), dim3(ctx->n1), zshm, ctx->stream, A, cf, fld<R>(ctx, MELT_RU1),
                         fld<R>(ctx, MELT_RU2), fld<R>(ctx, MELT_RU3), fld<R>(ctx, MELT_RUT), r4[0], r4[1], r4[2], r4[3], sc_u, sc_t, D, gu, gt);
    MELT_HIP(hipGetLastError());
    return 0;
  }
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
  size_t shm0 = sizeof(double) * 4 * (size_t)W * (ctx->n1 + 1);
  unsigned blocks = (unsigned)((size_t)(ctx->n2 / W) * ctx->n3);
  const bool flat = ctx->nc * sizeof(R) < (size_t(1) << 32);   // 32-bit byte offsets reach every element
  const size_t shm = fir ? sizeof(double) * 4 * (size_t)W * (ctx->n1 + 2 * FIR_KMAX) : shm0;
  if (shm > ctx->lds_max) return melt_fail(ctx, MELT_ERR_ARG, "explicit_xsolve", "tile exceeds LDS");
#define XF_LAUNCH(EE, FIRV, FLATV)                                                                                      \
  hipLaunchKernelGGL((k_explicit_xsolve<EE, R, FIRV, FLATV>), dim3(blocks), dim3(nt), shm, ctx->stream, A, cf,           \
                     fld<R>(ctx, MELT_RU1), fld<R>(ctx, MELT_RU2), fld<R>(ctx, MELT_RU3), fld<R>(ctx, MELT_RUT), r4[0], r4[1], \
                     r4[2], r4[3], bxu, bxt, sc_u, sc_t, W, gu, gt)
#define XF_CASE(EE)                                   \
  case EE:                                            \
    if (fir && flat) XF_LAUNCH(EE, true, true);       \
    else if (fir) XF_LAUNCH(EE, true, false);         \
    else if (flat) XF_LAUNCH(EE, false, true);        \
    else XF_LAUNCH(EE, false, false);                 \
    break;
  switch (E) {
    XF_CASE(1) XF_CASE(2) XF_CASE(3) XF_CASE(4) XF_CASE(6) XF_CASE(8)
    default: return melt_fail(ctx, MELT_ERR_ARG, "explicit_xsolve", "unsupported tile");
  }
#undef XF_CASE
#undef XF_LAUNCH
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
  if (shm > ctx->lds_max) return melt_fail(ctx, MELT_ERR_ARG, "line solve", "tile exceeds LDS");
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
  // chunks per thread along z (the longest segment that tiles the line; along
  // y single chunks measured faster, the extra blocks outweigh the halo)
  const int nch = dir == 1 ? 1 : N % (16 * GQ) == 0 ? 4 : N % (8 * GQ) == 0 ? 2 : 1;
  const int nsg = N / (4 * GQ * nch), nit = ctx->n1 / 64, other = dir == 1 ? ctx->n3 : ctx->n2;
  dim3 grid((unsigned)(nsg * nit * other), nfield);
#define GREEN_CASE(D, KK, NC) hipLaunchKernelGGL((k_green<D, R, KK, NC>), grid, dim3(256), 0, ctx->stream, B, ctx->n1, ctx->n2, ctx->n3)
#define GREEN_NCH(D, KK) \
  do { if (nch == 4) GREEN_CASE(D, KK, 4); else if (nch == 2) GREEN_CASE(D, KK, 2); else GREEN_CASE(D, KK, 1); } while (0)
  if (dir == 1) { if (KT == 4) GREEN_NCH(1, 4); else if (KT == 8) GREEN_NCH(1, 8); else GREEN_NCH(1, 12); }
  else { if (KT == 4) GREEN_NCH(2, 4); else if (KT == 8) GREEN_NCH(2, 8); else GREEN_NCH(2, 12); }
#undef GREEN_NCH
#undef GREEN_CASE
  MELT_HIP(hipGetLastError());
  *used = 1;
  return 0;
}

// The y and z solves of the fused substep as one pass (k_green_yz): returns 0
// with *used = 0 (launching nothing) when the shapes or the systems do not fit,
// and the caller runs the two passes instead.  Tuning key yz_jb: 0 default
// (8 rows per block), 8 or 16 rows, 1 = always two passes.
// planes per z segment at most: whole 256-plane lines measured fastest for
// 8-byte values; 4-byte values (two columns per thread, half the blocks) need
// the extra blocks of 128-plane segments (0.243 -> 0.169 ms at 256^3)
constexpr int YZ_SEG_MAX8 = 256, YZ_SEG_MAX4 = 128;

template <typename R>
int launch_green_yz(melt_ctx* ctx, void* const* src, void* const* qadd, const double* by, const double* bz, int nfield, int* used) {
  *used = 0;
  const int jb = ctx->tune[TUNE_YZ_JB] == 0 ? 8 : ctx->tune[TUNE_YZ_JB];
  if (ctx->no_green || (jb != 8 && jb != 16)) return 0;
  const int n1 = ctx->n1, n2 = ctx->n2, n3 = ctx->n3;
  const double tol = ctx->fp32 ? 1e-8 : 2e-17;
  GreenYZ<R> B{};
  int ng = 0;
  double seen[2][2] = {};
  for (int f = 0; f < nfield; ++f) {
    int w = -1;
    for (int g = 0; g < ng; ++g)
      if (seen[g][0] == by[f] && seen[g][1] == bz[f]) w = g;
    if (w < 0) {
      if (ng == 2) return 0;
      if (!green_coeffs(by[f], n2, tol, 1.0 / (1.0 + 2.0 * by[f]), B.gy[ng]) ||
          !green_coeffs(bz[f], n3, tol, 1.0 / (1.0 + 2.0 * bz[f]), B.gz[ng]))
        return 0;
      seen[ng][0] = by[f];
      seen[ng][1] = bz[f];
      w = ng++;
    }
    B.which[f] = w;
    B.src[f] = static_cast<const R*>(src[f]);
    B.qadd[f] = static_cast<R*>(qadd[f]);
  }
  int kmax = 0;
  for (int g = 0; g < ng; ++g) kmax = std::max(kmax, std::max(B.gy[g].K, B.gz[g].K));
  const int KT = kmax <= 4 ? 4 : kmax <= 8 ? 8 : 12;
  // two columns per thread for 4-byte values where the rows allow it
  constexpr int C2 = sizeof(R) == 4 ? 2 : 1;   // (no 2-column instantiation for 8-byte values)
  const int C = C2 == 2 && n1 % 128 == 0 ? 2 : 1;
  if (n1 % 64 || n2 % jb || n3 % GQ || 2 * KT >= n2 || 2 * KT >= n3) return 0;
  // planes per z segment: the longest of the maximum above, 64, 32, 16 that tiles the line
  int seg = GQ;
  for (int s : {sizeof(R) == 4 ? YZ_SEG_MAX4 : YZ_SEG_MAX8, 64, 32})
    if (n3 % s == 0) { seg = s; break; }
  dim3 grid((unsigned)((n3 / seg) * (n2 / jb) * (n1 / (64 * C))), nfield);
#define YZ_CASE(KK, JB, CC) hipLaunchKernelGGL((k_green_yz<R, KK, JB, CC>), grid, dim3(64 * JB), 0, ctx->stream, B, n1, n2, n3, seg)
#define YZ_JB(KK) \
  do {                                                         \
    if (C == 2) { if (jb == 16) YZ_CASE(KK, 16, C2); else YZ_CASE(KK, 8, C2); } \
    else { if (jb == 16) YZ_CASE(KK, 16, 1); else YZ_CASE(KK, 8, 1); }       \
  } while (0)
  if (KT == 4) YZ_JB(4); else if (KT == 8) YZ_JB(8); else YZ_JB(12);
#undef YZ_JB
#undef YZ_CASE
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
  if (ctx->n1 % 2 == 0) {
    hipLaunchKernelGGL(k_project2<R>, dim3(nblocks(ctx->nc / 2, 256)), dim3(256), 0, ctx->stream, make_grid(ctx), fld<R>(ctx, MELT_VX),
                       fld<R>(ctx, MELT_VY), fld<R>(ctx, MELT_VZ), fld<R>(ctx, MELT_PR), fld<R>(ctx, MELT_DPH),
                       use_mean ? ctx->d_fmean : nullptr, udx1, udx2, udx3, be, R(ctx->dx1q), R(ctx->dx2q), R(ctx->dx3q), do_u, do_p,
                       (int)ctx->nc, gate);
    MELT_HIP(hipGetLastError());
    return 0;
  }
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

// Launch the max reduction into d: one cell per thread into the partial
// slots, then one block folds them (kernels only, so the sequence may be
// captured into a graph).
template <typename R>
int launch_reduce_max(melt_ctx* ctx, int what, unsigned long long* d, double dt) {
  constexpr int V = 16 / sizeof(R);
  if (ctx->n1 % V) return melt_fail(ctx, MELT_ERR_ARG, "reduce_max", "n1m is not a multiple of the vector width");
  hipLaunchKernelGGL(k_reduce_max<R>, dim3(nblocks(ctx->nc / V, 256)), dim3(256), 0, ctx->stream, make_grid(ctx), fld<R>(ctx, MELT_VX),
                     fld<R>(ctx, MELT_VY), fld<R>(ctx, MELT_VZ), what, dt, (int)ctx->nc, ctx->d_slots);
  MELT_HIP(hipGetLastError());
  return euler_fold_slots(ctx, ctx->d_slots, d);
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
int euler_launch_green_yz(melt_ctx* ctx, void* const* src, void* const* qadd, const double* by, const double* bz, int nfield, int* used) {
  return ctx->fp32 ? launch_green_yz<float>(ctx, src, qadd, by, bz, nfield, used)
                   : launch_green_yz<double>(ctx, src, qadd, by, bz, nfield, used);
}
int euler_launch_divg(melt_ctx* ctx) { return ctx->fp32 ? launch_divg<float>(ctx) : launch_divg<double>(ctx); }
int euler_launch_project(melt_ctx* ctx, int do_u, int do_p, int use_mean, const int* gate) {
  return ctx->fp32 ? launch_project<float>(ctx, do_u, do_p, use_mean, gate) : launch_project<double>(ctx, do_u, do_p, use_mean, gate);
}
int euler_launch_interp_vof(melt_ctx* ctx) { return ctx->fp32 ? launch_interp_vof<float>(ctx) : launch_interp_vof<double>(ctx); }
int euler_launch_reduce_max(melt_ctx* ctx, int what, unsigned long long* d, double dt) {
  return ctx->fp32 ? launch_reduce_max<float>(ctx, what, d, dt) : launch_reduce_max<double>(ctx, what, d, dt);
}
int euler_fold_slots(melt_ctx* ctx, unsigned long long* slots, unsigned long long* out) {
  hipLaunchKernelGGL(k_fold_slots, dim3(1), dim3(REDUCE_SLOTS), 0, ctx->stream, slots, out);
  MELT_HIP(hipGetLastError());
  return 0;
}
int euler_launch_refine_gate(melt_ctx* ctx, const unsigned long long* dmax, int* gate) {
  hipLaunchKernelGGL(k_refine_gate, dim3(1), dim3(1), 0, ctx->stream, dmax, ctx->refine_rtol, gate, ctx->d_refine_stats);
  MELT_HIP(hipGetLastError());
  return 0;
}
int euler_reduce_max(melt_ctx* ctx, int what, double* out, double dt) {
  unsigned long long* d = reinterpret_cast<unsigned long long*>(ctx->d_red);
  int rc = euler_launch_reduce_max(ctx, what, d, dt);
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

// ------------------------------------------- allocation-free divergence check
// Scratch of melt_divgck_noalloc: the REDUCE_SLOTS partial maxima of the
// reduction (8 KiB).  Every GPU the library runs on holds its own copy until
// the process ends; host code reaches it through its device address, looked
// up once per device, never through this name.
static __device__ __align__(256) unsigned long long scratch[REDUCE_SLOTS];

#include <mutex>

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
unsigned long long* noalloc_ptr[NOALLOC_MAX_DEVICES] = {};
std::mutex noalloc_mu;

// Device address of `scratch` on the current device (hipGetSymbolAddress on
// the first call per device, with that device current).
int noalloc_scratch(unsigned long long** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) return MELT_ERR_HIP;
  std::lock_guard<std::mutex> lock(noalloc_mu);
  if (!noalloc_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MELT_ERR_HIP;
    noalloc_ptr[dev] = static_cast<unsigned long long*>(p);
  }
  *out = noalloc_ptr[dev];
  return 0;
}

__global__ __launch_bounds__(REDUCE_SLOTS) void k_clear_slots(unsigned long long* __restrict__ slots) { slots[threadIdx.x] = 0ull; }

template <typename R>
int divgck_noalloc(const void* vx, const void* vy, const void* vz, const Grid& g, hipStream_t s, double* qmax, unsigned long long* slots) {
  constexpr int V = 16 / sizeof(R);
  const size_t nc = (size_t)g.n1 * g.n2 * g.n3;
  hipLaunchKernelGGL(k_clear_slots, dim3(1), dim3(REDUCE_SLOTS), 0, s, slots);
  hipLaunchKernelGGL(k_reduce_max<R>, dim3(nblocks(nc / V, 256)), dim3(256), 0, s, g, static_cast<const R*>(vx), static_cast<const R*>(vy),
                     static_cast<const R*>(vz), 1, 0.0, (int)nc, slots);
  hipLaunchKernelGGL(k_fold_slots, dim3(1), dim3(REDUCE_SLOTS), 0, s, slots, reinterpret_cast<unsigned long long*>(qmax));
  return hipGetLastError() == hipSuccess ? 0 : MELT_ERR_HIP;
}

}  // namespace

// melt_divgck (divg.f90:26-56) without a context and without allocating: the
// three velocity components are device arrays of n1*n2*n3 values (double for
// elem_bytes 8, float for 4), i fastest, periodic in all directions; the
// maximum |div u| is written to the device double *qmax on `stream` (kernels
// only, so the call may be captured into a graph).  Limits: n1, n2, n3 >= 2,
// n1 a multiple of 16 / elem_bytes, n1*n2*n3 < 2^31.  Calls that overlap on
// one GPU race on the shared scratch.
extern "C" int melt_divgck_noalloc(const void* vx, const void* vy, const void* vz, int32_t n1, int32_t n2, int32_t n3, double xlen,
                                   double ylen, double zlen, int32_t elem_bytes, void* stream, double* qmax) {
  if (!vx || !vy || !vz || !qmax || (elem_bytes != 4 && elem_bytes != 8)) return MELT_ERR_ARG;
  if (n1 < 2 || n2 < 2 || n3 < 2 || !(xlen > 0.0) || !(ylen > 0.0) || !(zlen > 0.0)) return MELT_ERR_ARG;
  if (n1 % (16 / elem_bytes)) return MELT_ERR_ARG;
  if ((size_t)n1 * (size_t)n2 * (size_t)n3 >= ((size_t)1 << 31)) return MELT_ERR_ARG;
  unsigned long long* slots = nullptr;
  if (int rc = noalloc_scratch(&slots)) return rc;
  Grid g;
  g.n1 = n1; g.n2 = n2; g.n3 = n3;
  g.xlen = xlen; g.ylen = ylen; g.zlen = zlen;
  g.dx1 = (double)n1 / xlen; g.dx2 = (double)n2 / ylen; g.dx3 = (double)n3 / zlen;
  hipStream_t s = static_cast<hipStream_t>(stream);
  return elem_bytes == 8 ? divgck_noalloc<double>(vx, vy, vz, g, s, qmax, slots) : divgck_noalloc<float>(vx, vy, vz, g, s, qmax, slots);
}
