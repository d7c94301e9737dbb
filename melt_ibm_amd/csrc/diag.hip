// Diagnostics the reference's time loop calls around tschem (gcurv.f90:151-201):
// the time-step choice get_dt (cfl.f90:33-77) and the turbulence statistics
// of CalcDissipation (diss.f90:11-201).  Both read the fields once: get_dt is
// a max reduction (exact in any order), CalcDissipation a sum over the fluid
// cells taken in a fixed order (grid-stride per thread, a fixed tree per
// block, the block partials summed in block order), so every run on a given
// grid gives the same bits.
#include <cmath>

#include "common.h"

namespace {

constexpr int DISS_THREADS = 256;
constexpr int DISS_BLOCKS = 2048;   // 4 partials per block in ctx->d_red (16384 doubles)

// Per-cell terms of diss.f90:56-141, summed over cells with VOFp == 1
// (diss.f90:144-153): [0] tke, [1] diss, [2] chi, [3] the number of such cells.
template <typename R>
__global__ __launch_bounds__(DISS_THREADS) void k_diss_partial(Grid g, const R* __restrict__ vx, const R* __restrict__ vy,
                                                               const R* __restrict__ vz, const R* __restrict__ te,
                                                               const R* __restrict__ vof, double ren, double pec,
                                                               double* __restrict__ part) {
  const size_t nc = (size_t)g.n1 * g.n2 * g.n3;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const double udx1 = g.dx1, udx2 = g.dx2, udx3 = g.dx3;
  for (size_t c = (size_t)blockIdx.x * DISS_THREADS + threadIdx.x; c < nc; c += (size_t)gridDim.x * DISS_THREADS) {
    if ((double)vof[c] != 1.0) continue;
    const int i = (int)(c % g.n1), j = (int)((c / g.n1) % g.n2), k = (int)(c / ((size_t)g.n1 * g.n2));
    const size_t sj = g.n1, sk = (size_t)g.n1 * g.n2;
    const size_t ip = c - i + (i + 1 == g.n1 ? 0 : i + 1), im = c - i + (i == 0 ? g.n1 - 1 : i - 1);
    const size_t jp = (j + 1 == g.n2) ? c - j * sj : c + sj, jm = (j == 0) ? c + (g.n2 - 1) * sj : c - sj;
    const size_t kp = (k + 1 == g.n3) ? c - k * sk : c + sk, km = (k == 0) ? c + (g.n3 - 1) * sk : c - sk;
    auto X = [&](size_t a) { return (double)vx[a]; };
    auto Y = [&](size_t a) { return (double)vy[a]; };
    auto Z = [&](size_t a) { return (double)vz[a]; };
    auto T = [&](size_t a) { return (double)te[a]; };
    // the (ip, jp), (im, jp), ... corners: offsets combine
    const size_t dip = ip - c, dim_ = im - c, djp = jp - c, djm = jm - c, dkp = kp - c, dkm = km - c;
    const double h11 = (X(ip) - X(c)) * udx1;
    const double h12 = ((X(jp) + X(c + dip + djp)) - (X(jm) + X(c + dip + djm))) * 0.25 * udx2;
    const double h13 = ((X(kp) + X(c + dip + dkp)) - (X(km) + X(c + dip + dkm))) * 0.25 * udx3;
    const double h21 = ((Y(ip) + Y(c + dip + djp)) - (Y(im) + Y(c + dim_ + djp))) * 0.25 * udx1;
    const double h22 = (Y(jp) - Y(c)) * udx2;
    const double h23 = ((Y(kp) + Y(c + djp + dkp)) - (Y(km) + Y(c + djp + dkm))) * 0.25 * udx3;
    const double h31 = ((Z(ip) + Z(c + dip + dkp)) - (Z(im) + Z(c + dim_ + dkp))) * 0.25 * udx1;
    const double h32 = ((Z(jp) + Z(c + djp + dkp)) - (Z(jm) + Z(c + djm + dkp))) * 0.25 * udx2;
    const double h33 = (Z(kp) - Z(c)) * udx3;
    double tke = (X(c) * X(c) + X(ip) * X(ip) + Y(c) * Y(c) + Y(jp) * Y(jp) + Z(c) * Z(c) + Z(kp) * Z(kp)) * 0.5;
    tke = tke * 0.5;
    double d = h11 * h11 + h22 * h22 + h33 * h33 + 0.5 * ((h21 + h12) * (h21 + h12)) + 0.5 * ((h31 + h13) * (h31 + h13)) +
               0.5 * ((h32 + h23) * (h32 + h23));
    d = d * 2.0 / ren;
    const double ax = 0.5 * g.dx1 * (T(ip) - T(im)), ay = 0.5 * g.dx2 * (T(jp) - T(jm)), az = 0.5 * g.dx3 * (T(kp) - T(km));
    double chi = ax * ax + ay * ay + az * az;
    chi = chi / pec;
    acc[0] += tke;
    acc[1] += d;
    acc[2] += chi;
    acc[3] += 1.0;
  }
  __shared__ double red[4][DISS_THREADS / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double v = acc[q];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) red[q][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double v = 0.0;
    for (int q = 0; q < DISS_THREADS / 64; ++q) v += red[threadIdx.x][q];
    part[4 * blockIdx.x + threadIdx.x] = v;
  }
}

// The block partials in block order, one thread per quantity.
__global__ void k_diss_final(const double* __restrict__ part, int nb, double* __restrict__ out) {
  if (threadIdx.x >= 4) return;
  double v = 0.0;
  for (int b = 0; b < nb; ++b) v += part[4 * b + threadIdx.x];
  out[threadIdx.x] = v;
}

template <typename R>
int launch_diss(melt_ctx* ctx, double* sums) {
  double* part = ctx->d_red;
  double* out = ctx->d_red + 4 * DISS_BLOCKS;
  hipLaunchKernelGGL(k_diss_partial<R>, dim3(DISS_BLOCKS), dim3(DISS_THREADS), 0, ctx->stream, make_grid(ctx), fld<R>(ctx, MELT_VX),
                     fld<R>(ctx, MELT_VY), fld<R>(ctx, MELT_VZ), fld<R>(ctx, MELT_TEMP), fld<R>(ctx, MELT_VOFP), ctx->cfg.ren,
                     ctx->cfg.ren * ctx->cfg.prandtl, part);
  MELT_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_diss_final, dim3(1), dim3(64), 0, ctx->stream, part, DISS_BLOCKS, out);
  MELT_HIP(hipGetLastError());
  MELT_HIP(hipMemcpyAsync(sums, out, 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  MELT_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

}  // namespace

int diag_dissipation(melt_ctx* ctx, double* sums) { return ctx->fp32 ? launch_diss<float>(ctx, sums) : launch_diss<double>(ctx, sums); }

// get_dt (cfl.f90:33-77): the largest advective CFL of the cells at the
// current dt, at least the viscous one 6 dt dx1^2 / Re, and dt scaled to cflfix.
extern "C" int melt_get_dt(melt_ctx* ctx, double cflfix, double* dt_out) {
  if (!ctx || !dt_out || !(ctx->dt > 0.0)) return MELT_ERR_ARG;
  double adv = 0.0;
  int rc = euler_reduce_max(ctx, 2, &adv, ctx->dt);
  if (rc) return rc;
  double buf = adv;
  const double visc = 6.0 / ctx->cfg.ren * ctx->dt * (ctx->dx1 * ctx->dx1);
  if (visc > buf) buf = visc;
  *dt_out = buf > 0.0 ? ctx->dt * cflfix / buf : ctx->dt;
  return 0;
}

// CalcDissipation (diss.f90:11-201): re-tags the cells of a moving body first
// (diss.f90:47-49), sums over the fully fluid cells and forms the statistics
// the reference writes to diss.txt (diss.f90:155-197):
// out = {diss_volAvg, kenerg, urms, eta, lambda_t, Re_L, re_lam, L_int,
//        chi_volAvg, kmax_eta, sumVOF}.
extern "C" int melt_dissipation(melt_ctx* ctx, double* out) {
  if (!ctx || !out) return MELT_ERR_ARG;
  const melt_config& c = ctx->cfg;
  const bool moving = c.imlsfor == 1 && (c.imlsstr == 1 || c.imelt == 1) && ctx->has_geom;
  int rc;
  if (moving && ((rc = ibm_tag(ctx)) || (c.tagtype != 0 && (rc = euler_launch_interp_vof(ctx))))) return rc;
  double s[4];
  if ((rc = diag_dissipation(ctx, s))) return rc;
  const double sumvof = s[3];
  const double kenerg = s[0] / sumvof;
  const double nu = 1.0 / c.ren;
  const double diss = s[1] / sumvof;
  const double chi = s[2] / sumvof;
  const double urms = sqrt(2.0 / 3.0 * kenerg);
  const double eta = pow(nu * nu * nu / diss, 0.25);
  const double kmax_eta = ctx->pi * ctx->dx1 * eta;
  const double re_l = kenerg * kenerg / (nu * diss);
  const double lambda_t = sqrt(15.0 * nu * (urms * urms) / diss);
  const double re_lam = urms * lambda_t / nu;
  const double l_int = sqrt(kenerg * kenerg * kenerg) / diss;
  const double v[11] = {diss, kenerg, urms, eta, lambda_t, re_l, re_lam, l_int, chi, kmax_eta, sumvof};
  for (int q = 0; q < 11; ++q) out[q] = v[q];
  return melt_sync(ctx);
}
