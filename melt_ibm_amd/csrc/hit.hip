// HIT forcing of the reference's time loop (hit.f90), computed on the device:
//  - InitRandomForce (hit.f90:128-232): nmodes = 2 floor(KF/KMIN) + 1 wave
//    numbers -n..n per direction, bhat = 0, the complex exponentials of the
//    staggered grid coordinates;
//  - CalcHITRandomForce (hit.f90:1-126): one Ornstein-Uhlenbeck step of the
//    Eswaran & Pope coefficients bhat for every mode with 0 < |k| < kf, the
//    solenoidal projection Fhat = bhat - k (k.bhat) / |k|^2, and the forcing
//    fields forc(x) = Re sum_k Fhat(k) exp(i k.x) at each component's staggered
//    points;
//  - CalcABC_HITForce (hit.f90:234-265).
// The reference draws its Gaussian numbers from gfortran's random_number
// (Box-Muller, hit.f90:39-44), a stream no other runtime reproduces.  Here the
// uniform numbers come from SplitMix64 (Steele, Lea & Flood 2014) evaluated at
// a counter: draw d of the run is mix(seed + (d + 1) * golden), consumed in the
// reference's order (modes in l, m, n loop order, for each direction u1 then
// u2), so the stream is a pure function of (seed, step) and the device-side
// counter makes the step safe to capture in a graph.
// The synthesis is separable: a thread owns one (i, j) column, sums
// H(n) = sum_m exp(i km y_j) sum_l Fhat(l, m, n) exp(i kl x_i) once, and then
// marches a run of z planes, each value being Re sum_n H(n) exp(i kn z_k).
#include <cmath>
#include <vector>

#include "common.h"

namespace {

constexpr int HIT_ZSEG = 32;   // z planes per synthesis thread

__host__ __device__ inline unsigned long long hit_mix(unsigned long long seed, unsigned long long d) {
  unsigned long long z = seed + (d + 1ull) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ inline double hit_uniform(unsigned long long seed, unsigned long long d) {
  return (double)(hit_mix(seed, d) >> 11) * (1.0 / 9007199254740992.0);   // [0, 1)
}

struct HitArgs {
  int nm, nact;
  double tl, epsstar, dt;
  double twopi_x, twopi_y, twopi_z;   // 2 pi / len
  unsigned long long seed;
};

// One OU step of every active mode (one thread each) and the projection;
// the step counter advances once all threads have read it.
__global__ __launch_bounds__(1024) void k_hit_ou(HitArgs a, const int* __restrict__ act, double* __restrict__ bhat,
                                                 double* __restrict__ fhat, unsigned long long* __restrict__ count) {
  __shared__ unsigned long long step;
  if (threadIdx.x == 0) step = *count;
  __syncthreads();
  const int half = (a.nm - 1) / 2;
  const double fac = 1.0 - a.dt / a.tl;
  const double amp = sqrt(2.0 * a.epsstar * a.dt / (a.tl * a.tl));
  for (int q = threadIdx.x; q < a.nact; q += blockDim.x) {
    const int idx = act[q];                       // l + nm (m + nm n), 0-based
    const int l = idx % a.nm, m = (idx / a.nm) % a.nm, n = idx / (a.nm * a.nm);
    const double kl = (double)(l - half) * a.twopi_x, km = (double)(m - half) * a.twopi_y, kn = (double)(n - half) * a.twopi_z;
    const double kmag = sqrt(kl * kl + km * km + kn * kn);
    const unsigned long long d0 = (step * (unsigned long long)a.nact + (unsigned long long)q) * 6ull;
    double br[3], bi[3];
    for (int dir = 0; dir < 3; ++dir) {
      const double u1 = 1.0 - hit_uniform(a.seed, d0 + 2 * dir);    // (0, 1]: log defined
      const double u2 = hit_uniform(a.seed, d0 + 2 * dir + 1);
      const double r = sqrt(-2.0 * log(u1));
      const double re = r * cos(2.0 * M_PI * u2), im = r * sin(2.0 * M_PI * u2);
      double* b = &bhat[2 * (dir + 3 * idx)];
      br[dir] = b[0] * fac + re * amp;
      bi[dir] = b[1] * fac + im * amp;
      b[0] = br[dir];
      b[1] = bi[dir];
    }
    const double dr = kl * br[0] + km * br[1] + kn * br[2], di = kl * bi[0] + km * bi[1] + kn * bi[2];
    const double kv[3] = {kl, km, kn}, k2 = kmag * kmag;
    for (int dir = 0; dir < 3; ++dir) {
      fhat[2 * (dir + 3 * idx)] = br[dir] - kv[dir] * dr / k2;
      fhat[2 * (dir + 3 * idx) + 1] = bi[dir] - kv[dir] * di / k2;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) *count = step + 1ull;
}

// Synthesis F_c(x, y, z) = Re sum_{l,m,n} Fhat(c,l,m,n) e^{i k_l x} e^{i k_m y}
// e^{i k_n z} at each component's staggered points, in three separable
// sums.  tab holds exp(i k x) for the 6 staggered coordinate sets, each
// (nm, N) complex: [0] xc, [1] xm, [2] yc, [3] ym, [4] zc, [5] zm; component c
// sits at (xc, ym, zm), (xm, yc, zm), (xm, ym, zc).
//
// The x sum depends only on i: k_hit_xsum forms ax(c, n, m, i) =
// sum_l Fhat(c,l,m,n) e^{i k_l x_i} once per grid line instead of once per
// (i, j) column.  k_hit_synth then takes, per (i, j) column and 32-plane
// segment, the y sum of those (75 complex products for nm = 5) and per plane
// the z sum (15 real parts): about 40 FP64 FMAs per cell, so the pass is
// bound by its field stores.
template <int NM>
__global__ __launch_bounds__(256) void k_hit_xsum(int n1, const double* __restrict__ fhat, const double* __restrict__ tab,
                                                  double* __restrict__ ax) {
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= 3 * NM * NM * n1) return;
  const int i = t % n1, r = t / n1, m = r % NM, n = (r / NM) % NM, c = r / (NM * NM);
  const double* tx = c == 0 ? tab : tab + 2 * NM * n1;   // xc for u, xm for v and w
  double gr = 0.0, gi = 0.0;
#pragma unroll
  for (int l = 0; l < NM; ++l) {
    const double* F = &fhat[2 * (c + 3 * (l + NM * (m + NM * n)))];
    const double* E = &tx[2 * (l * n1 + i)];
    gr = fma(F[0], E[0], fma(-F[1], E[1], gr));
    gi = fma(F[0], E[1], fma(F[1], E[0], gi));
  }
  ax[2 * t] = gr;
  ax[2 * t + 1] = gi;
}

//
// A block covers 256 columns of one segment (bps blocks per segment), so the
// plane index is the same for the whole block and the z table is read with
// scalar loads.
template <typename R, int NM>
__global__ __launch_bounds__(256) void k_hit_synth(Grid g, const double* __restrict__ ax, const double* __restrict__ tab,
                                                   R* __restrict__ fx, R* __restrict__ fy, R* __restrict__ fz, int bps) {
  const int cols = g.n1 * g.n2;
  const int seg = (int)blockIdx.x / bps, col = ((int)blockIdx.x % bps) * 256 + (int)threadIdx.x;
  if (col >= cols) return;
  const int i = col % g.n1, j = col / g.n1;
  const int n1 = g.n1, n2 = g.n2, n3 = g.n3;
  const double* ty[3] = {tab + 2 * NM * (2 * n1) + 2 * NM * n2, tab + 2 * NM * (2 * n1), tab + 2 * NM * (2 * n1) + 2 * NM * n2};  // ym, yc, ym
  const double* zb = tab + 2 * NM * (2 * n1 + 2 * n2);
  const double* tz[3] = {zb + 2 * NM * n3, zb + 2 * NM * n3, zb};                            // zm, zm, zc
  double hr[3][NM], hi[3][NM];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
#pragma unroll
    for (int n = 0; n < NM; ++n) {
      double sr = 0.0, si = 0.0;
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        const double* A = &ax[2 * (((c * NM + n) * NM + m) * n1 + i)];
        const double* E = &ty[c][2 * (m * n2 + j)];
        sr = fma(A[0], E[0], fma(-A[1], E[1], sr));
        si = fma(A[0], E[1], fma(A[1], E[0], si));
      }
      hr[c][n] = sr;
      hi[c][n] = si;
    }
  }
  R* out[3] = {fx, fy, fz};
  const int k0 = seg * HIT_ZSEG, k1 = min(n3, k0 + HIT_ZSEG);
  for (int k = k0; k < k1; ++k) {
    const size_t o = (size_t)i + (size_t)n1 * ((size_t)j + (size_t)n2 * k);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double v = 0.0;
#pragma unroll
      for (int n = 0; n < NM; ++n) {
        const double* E = &tz[c][2 * (n * n3 + k)];
        v = fma(hr[c][n], E[0], fma(-hi[c][n], E[1], v));
      }
      __builtin_nontemporal_store((R)v, &out[c][o]);   // read once, by the next explicit pass
    }
  }
}

// CalcABC_HITForce (hit.f90:234-265).
template <typename R>
__global__ __launch_bounds__(256) void k_abc(Grid g, double C, const double* __restrict__ coord, R* __restrict__ fx, R* __restrict__ fy,
                                             R* __restrict__ fz) {
  const size_t nc = (size_t)g.n1 * g.n2 * g.n3;
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const int i = (int)(c % g.n1), j = (int)((c / g.n1) % g.n2), k = (int)(c / ((size_t)g.n1 * g.n2));
  const double pi2 = 2.0 * M_PI;
  const double xm = coord[coord_off(g, 0, 1) + i], ym = coord[coord_off(g, 1, 1) + j], zm = coord[coord_off(g, 2, 1) + k];
  fx[c] = (R)(C * sin(pi2 * zm / g.zlen) + C * cos(pi2 * ym / g.ylen));
  fy[c] = (R)(C * sin(pi2 * xm / g.xlen) + C * cos(pi2 * zm / g.zlen));
  fz[c] = (R)(C * sin(pi2 * ym / g.ylen) + C * cos(pi2 * xm / g.xlen));
}

inline unsigned nblk(size_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

int hit_free(melt_ctx* ctx) {
  for (void* p : {(void*)ctx->d_hit_bhat, (void*)ctx->d_hit_fhat, (void*)ctx->d_hit_tab, (void*)ctx->d_hit_act, (void*)ctx->d_hit_count,
                  (void*)ctx->d_hit_ax})
    if (p) hipFree(p);
  ctx->d_hit_bhat = ctx->d_hit_fhat = ctx->d_hit_tab = ctx->d_hit_ax = nullptr;
  ctx->d_hit_act = nullptr;
  ctx->d_hit_count = nullptr;
  ctx->hit_nm = 0;
  return 0;
}

// InitRandomForce (hit.f90:128-232) with the forcing parameters of HITForcing.in.
extern "C" int melt_hit_init(melt_ctx* ctx, double tl, double epsstar, double kf_on_kmin, uint64_t seed) {
  if (ctx) ctx->forc_free = 0;
  if (!ctx || !(tl > 0.0) || !(epsstar >= 0.0) || !(kf_on_kmin > 0.0)) return MELT_ERR_ARG;
  const int nbuf = (int)floor(kf_on_kmin) + 1;
  const int nm = (nbuf - 1) * 2 + 1;
  if (nm > 9) return melt_fail(ctx, MELT_ERR_ARG, "hit_init", "KF/KMIN of 5 or more (over 9 wave numbers per direction)");
  MELT_HIP(hipStreamSynchronize(ctx->stream));
  hit_free(ctx);
  const double pi = ctx->pi;
  const double kf = kf_on_kmin * 2.0 * pi / ctx->cfg.xlen;
  std::vector<int> wave(nm);
  for (int i = 0; i < nm; ++i) wave[i] = i - (nbuf - 1);
  // forced modes, in the reference's l, m, n loop order (hit.f90:28-32, 36)
  std::vector<int> act;
  for (int l = 0; l < nm; ++l)
    for (int m = 0; m < nm; ++m)
      for (int n = 0; n < nm; ++n) {
        const double kl = (double)wave[l] * 2.0 * pi / ctx->cfg.xlen, km = (double)wave[m] * 2.0 * pi / ctx->cfg.ylen,
                     kn = (double)wave[n] * 2.0 * pi / ctx->cfg.zlen;
        const double kmag = sqrt(kl * kl + km * km + kn * kn);
        if (kmag > 0.0 && kmag < kf) act.push_back(l + nm * (m + nm * n));
      }
  // exp(i kappa x) of hit.f90:171-204 on the staggered coordinates
  const Grid g = make_grid(ctx);
  const int ns[3] = {ctx->n1, ctx->n2, ctx->n3};
  const double lens[3] = {ctx->cfg.xlen, ctx->cfg.ylen, ctx->cfg.zlen};
  std::vector<double> tab;
  for (int d = 0; d < 3; ++d)
    for (int centre = 0; centre < 2; ++centre)
      for (int w = 0; w < nm; ++w)
        for (int i = 1; i <= ns[d]; ++i) {
          const double kappa = (double)wave[w] * 2.0 * pi / lens[d];
          double x = d == 0 ? (centre ? gxm(g, i) : gxc(g, i)) : d == 1 ? (centre ? gym(g, i) : gyc(g, i)) : (centre ? gzm(g, i) : gzc(g, i));
          tab.push_back(cos(kappa * x));
          tab.push_back(sin(kappa * x));
        }
  const size_t nb = sizeof(double) * 2 * 3 * (size_t)nm * nm * nm;
  MELT_HIP(hipMalloc(&ctx->d_hit_bhat, nb));
  MELT_HIP(hipMalloc(&ctx->d_hit_fhat, nb));
  MELT_HIP(hipMalloc(&ctx->d_hit_tab, sizeof(double) * tab.size()));
  MELT_HIP(hipMalloc(&ctx->d_hit_ax, sizeof(double) * 2 * 3 * (size_t)nm * nm * ctx->n1));
  MELT_HIP(hipMalloc(&ctx->d_hit_act, sizeof(int) * (act.empty() ? 1 : act.size())));
  MELT_HIP(hipMalloc(&ctx->d_hit_count, sizeof(unsigned long long)));
  MELT_HIP(hipMemset(ctx->d_hit_bhat, 0, nb));
  MELT_HIP(hipMemset(ctx->d_hit_fhat, 0, nb));
  MELT_HIP(hipMemset(ctx->d_hit_count, 0, sizeof(unsigned long long)));
  MELT_HIP(hipMemcpy(ctx->d_hit_tab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
  if (!act.empty()) MELT_HIP(hipMemcpy(ctx->d_hit_act, act.data(), sizeof(int) * act.size(), hipMemcpyHostToDevice));
  ctx->hit_nm = nm;
  ctx->hit_nact = (int)act.size();
  ctx->hit_tl = tl;
  ctx->hit_eps = epsstar;
  ctx->hit_kf = kf_on_kmin;
  ctx->hit_seed = seed;
  return 0;
}

template <typename R, int NM>
static int launch_hit_synth_nm(melt_ctx* ctx) {
  const int nseg = (ctx->n3 + HIT_ZSEG - 1) / HIT_ZSEG;
  const int bps = (int)nblk((size_t)ctx->n1 * ctx->n2, 256);
  hipLaunchKernelGGL((k_hit_xsum<NM>), dim3(nblk((size_t)3 * NM * NM * ctx->n1, 256)), dim3(256), 0, ctx->stream, ctx->n1, ctx->d_hit_fhat,
                     ctx->d_hit_tab, ctx->d_hit_ax);
  hipLaunchKernelGGL((k_hit_synth<R, NM>), dim3((unsigned)(bps * nseg)), dim3(256), 0, ctx->stream, make_grid(ctx), ctx->d_hit_ax,
                     ctx->d_hit_tab, fld<R>(ctx, MELT_FORCX), fld<R>(ctx, MELT_FORCY), fld<R>(ctx, MELT_FORCZ), bps);
  MELT_HIP(hipGetLastError());
  return 0;
}
template <typename R>
static int launch_hit_synth(melt_ctx* ctx) {
  switch (ctx->hit_nm) {
    case 1: return launch_hit_synth_nm<R, 1>(ctx);
    case 3: return launch_hit_synth_nm<R, 3>(ctx);
    case 5: return launch_hit_synth_nm<R, 5>(ctx);
    case 7: return launch_hit_synth_nm<R, 7>(ctx);
    case 9: return launch_hit_synth_nm<R, 9>(ctx);
    default: return melt_fail(ctx, MELT_ERR_ARG, "hit_force", "unsupported number of wave numbers");
  }
}

// CalcHITRandomForce (hit.f90:1-126) at the context's dt: kernels only, so a
// caller may capture it together with tschem.
extern "C" int melt_hit_force(melt_ctx* ctx) {
  if (!ctx) return MELT_ERR_ARG;
  if (!ctx->hit_nm) return melt_fail(ctx, MELT_ERR_ARG, "hit_force", "melt_hit_init has not been called");
  if (!(ctx->dt > 0.0)) return melt_fail(ctx, MELT_ERR_ARG, "hit_force", "dt not set");
  HitArgs a;
  a.nm = ctx->hit_nm;
  a.nact = ctx->hit_nact;
  a.tl = ctx->hit_tl;
  a.epsstar = ctx->hit_eps;
  a.dt = ctx->dt;
  a.twopi_x = 2.0 * ctx->pi / ctx->cfg.xlen;
  a.twopi_y = 2.0 * ctx->pi / ctx->cfg.ylen;
  a.twopi_z = 2.0 * ctx->pi / ctx->cfg.zlen;
  a.seed = ctx->hit_seed;
  // the synthesis writes only forcx/y/z and the generator's state: after a
  // tschem it starts when the last explicit pass has read the forcing, beside
  // the rest of that step
  return run_after_forcing_read(ctx, [&]() -> int {
    hipLaunchKernelGGL(k_hit_ou, dim3(1), dim3(256), 0, ctx->stream, a, ctx->d_hit_act, ctx->d_hit_bhat, ctx->d_hit_fhat, ctx->d_hit_count);
    MELT_HIP(hipGetLastError());
    return ctx->fp32 ? launch_hit_synth<float>(ctx) : launch_hit_synth<double>(ctx);
  });
}

// bhat(3, nmodes, nmodes, nmodes) as (re, im) pairs and the number of OU
// steps taken (the generator's position): WriteRandForcCoef / hdf_read_bhat
// (hit.f90:267-356), plus the position so that a resumed run continues the
// same random stream.
extern "C" int melt_hit_state(melt_ctx* ctx, double* bhat, uint64_t* steps, int32_t set) {
  if (!ctx || !bhat || !steps) return MELT_ERR_ARG;
  if (!ctx->hit_nm) return melt_fail(ctx, MELT_ERR_ARG, "hit_state", "melt_hit_init has not been called");
  const size_t nb = sizeof(double) * 2 * 3 * (size_t)ctx->hit_nm * ctx->hit_nm * ctx->hit_nm;
  MELT_HIP(hipStreamSynchronize(ctx->stream));
  if (set) {
    MELT_HIP(hipMemcpy(ctx->d_hit_bhat, bhat, nb, hipMemcpyHostToDevice));
    unsigned long long s = *steps;
    MELT_HIP(hipMemcpy(ctx->d_hit_count, &s, sizeof s, hipMemcpyHostToDevice));
  } else {
    MELT_HIP(hipMemcpy(bhat, ctx->d_hit_bhat, nb, hipMemcpyDeviceToHost));
    unsigned long long s = 0;
    MELT_HIP(hipMemcpy(&s, ctx->d_hit_count, sizeof s, hipMemcpyDeviceToHost));
    *steps = s;
  }
  return 0;
}

extern "C" int melt_hit_nmodes(melt_ctx* ctx, int32_t* nmodes, int32_t* nforced) {
  if (!ctx || !nmodes || !nforced) return MELT_ERR_ARG;
  *nmodes = ctx->hit_nm;
  *nforced = ctx->hit_nact;
  return 0;
}

template <typename R>
static int launch_abc(melt_ctx* ctx, double C) {
  hipLaunchKernelGGL(k_abc<R>, dim3(nblk(ctx->nc, 256)), dim3(256), 0, ctx->stream, make_grid(ctx), C, ctx->d_coord,
                     fld<R>(ctx, MELT_FORCX), fld<R>(ctx, MELT_FORCY), fld<R>(ctx, MELT_FORCZ));
  MELT_HIP(hipGetLastError());
  return 0;
}

extern "C" int melt_abc_force(melt_ctx* ctx, double c_hit) {
  if (!ctx) return MELT_ERR_ARG;
  const double C = c_hit * ((2.0 * ctx->pi / ctx->cfg.xlen) * (2.0 * ctx->pi / ctx->cfg.xlen));
  return ctx->fp32 ? launch_abc<float>(ctx, C) : launch_abc<double>(ctx, C);
}
