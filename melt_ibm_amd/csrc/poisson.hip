// Pressure Poisson solve of phcalc.f90:11-97 on gfx950.
//
// phi = L^{-1} dph with L the 7-point staggered Laplacian and the (0,0,0) mode
// set to zero: forward transforms in x (real-to-complex, fused with the
// divergence of divg.f90 in the fused substep), y and z, division by the
// modified wavenumbers -(ak3 + ak1 + ak2) (phini.f90:35-55) after the 1/N^3
// normalisation (phcalc.f90:48-60), and the inverse transforms.  Each transform
// is a Stockham radix-4/2/3 FFT of a tile of lines held in LDS, so each pass
// reads and writes the spectrum once.  The z pass does forward FFT, division
// and inverse FFT in one visit.
//
// Spectrum layout: complex (n1hp x N2 x N3), kx fastest, n1hp = N1/2+1 rounded
// up to a multiple of 8 so every row starts on a 128-byte boundary.
#include <cmath>
#include <cstring>

#include "common.h"

namespace {

constexpr int FFT_TILE = 2048;  // complex elements per LDS tile (32 KiB)
constexpr int FFT_MAXB = 4;     // butterflies per thread per stage (256 threads, radix >= 2)

struct FftPlan {
  int N, nst;
  int radix[24];
  const double2* tw;  // exp(-2 pi i t / N), t < N
};

__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// In-LDS Stockham FFT of nl lines: element (line l, position p) at s[l*ls + p*ps].
// lfast: consecutive butterflies run over lines (tile layouts with the line
// index contiguous) instead of positions.  sign -1 forward, +1 backward
// (unnormalised, as FFTW's FFTW_FORWARD / FFTW_BACKWARD).
__device__ void lds_fft(double2* s, int nl, int ls, int ps, bool lfast, const FftPlan& P, int sign) {
  const int N = P.N, tid = threadIdx.x, nt = blockDim.x;
  int Ns = 1;
  for (int st = 0; st < P.nst; ++st) {
    const int R = P.radix[st];
    const int NR = N / R;
    const int nbf = nl * NR;
    double2 v[FFT_MAXB][4];
#pragma unroll
    for (int q = 0; q < FFT_MAXB; ++q) {
      int b = tid + q * nt;
      if (b < nbf) {
        int l = lfast ? b % nl : b / NR;
        int j = lfast ? b / nl : b % NR;
        for (int r = 0; r < R; ++r) v[q][r] = s[l * ls + (j + r * NR) * ps];
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < FFT_MAXB; ++q) {
      int b = tid + q * nt;
      if (b < nbf) {
        int l = lfast ? b % nl : b / NR;
        int j = lfast ? b / nl : b % NR;
        int k = j % Ns;
        int step = N / (Ns * R);
        for (int r = 1; r < R; ++r) {
          double2 w = P.tw[k * r * step];
          if (sign > 0) w.y = -w.y;
          v[q][r] = cmul(v[q][r], w);
        }
        double2 y0, y1, y2, y3;
        if (R == 4) {
          double2 a0 = cadd(v[q][0], v[q][2]), a1 = csub(v[q][0], v[q][2]);
          double2 a2 = cadd(v[q][1], v[q][3]), d = csub(v[q][1], v[q][3]);
          // times -i (forward) or +i (backward)
          double2 a3 = sign < 0 ? make_double2(d.y, -d.x) : make_double2(-d.y, d.x);
          y0 = cadd(a0, a2); y1 = cadd(a1, a3); y2 = csub(a0, a2); y3 = csub(a1, a3);
        } else if (R == 2) {
          y0 = cadd(v[q][0], v[q][1]); y1 = csub(v[q][0], v[q][1]);
          y2 = y3 = make_double2(0, 0);
        } else {  // R == 3
          const double c = -0.5, sn = (sign < 0 ? -1.0 : 1.0) * 0.86602540378443864676;
          double2 sm = cadd(v[q][1], v[q][2]), df = csub(v[q][1], v[q][2]);
          y0 = cadd(v[q][0], sm);
          double2 t1 = make_double2(v[q][0].x + c * sm.x, v[q][0].y + c * sm.y);
          double2 t2 = make_double2(-sn * df.y, sn * df.x);  // i*sn*df
          y1 = cadd(t1, t2); y2 = csub(t1, t2);
          y3 = make_double2(0, 0);
        }
        int o = (j / Ns) * Ns * R + k;
        s[l * ls + o * ps] = y0;
        s[l * ls + (o + Ns) * ps] = y1;
        if (R >= 3) s[l * ls + (o + 2 * Ns) * ps] = y2;
        if (R == 4) s[l * ls + (o + 3 * Ns) * ps] = y3;
      }
    }
    __syncthreads();
    Ns *= R;
  }
}

struct SpecArgs {
  Grid g;
  int n1h, n1hp;
  double2* spec;
};

// x pass, forward: real line (divergence or dph) -> bins 0..N1/2.
__global__ __launch_bounds__(256) void k_fft_x_fwd(SpecArgs S, FftPlan P, int nl, const double* __restrict__ vx,
                                                   const double* __restrict__ vy, const double* __restrict__ vz,
                                                   const double* __restrict__ dph, double usdtal) {
  extern __shared__ double2 sh[];
  const Grid& g = S.g;
  const int N = g.n1, ls = N + 1;
  unsigned t = xcd_block(blockIdx.x, gridDim.x);
  int ntj = g.n2 / nl;
  int j0 = (t % ntj) * nl, k = t / ntj;
  for (int c = threadIdx.x; c < nl * N; c += blockDim.x) {
    int l = c / N, i = c % N, j = j0 + l;
    size_t idx = (size_t)i + (size_t)N * ((size_t)j + (size_t)g.n2 * k);
    double v;
    if (dph) {
      v = dph[idx];
    } else {  // divg.f90:17-20
      int ip = i == N - 1 ? 0 : i + 1, jp = j == g.n2 - 1 ? 0 : j + 1, kp = k == g.n3 - 1 ? 0 : k + 1;
      double dq = (vx[idx - i + ip] - vx[idx]) * g.dx1 + (vy[(size_t)i + (size_t)N * ((size_t)jp + (size_t)g.n2 * k)] - vy[idx]) * g.dx2 +
                  (vz[(size_t)i + (size_t)N * ((size_t)j + (size_t)g.n2 * kp)] - vz[idx]) * g.dx3;
      v = dq * usdtal;
    }
    sh[l * ls + i] = make_double2(v, 0.0);
  }
  __syncthreads();
  lds_fft(sh, nl, ls, 1, false, P, -1);
  for (int c = threadIdx.x; c < nl * S.n1hp; c += blockDim.x) {
    int l = c / S.n1hp, m = c % S.n1hp;
    double2 v = m < S.n1h ? sh[l * ls + m] : make_double2(0.0, 0.0);
    S.spec[(size_t)m + (size_t)S.n1hp * ((size_t)(j0 + l) + (size_t)g.n2 * k)] = v;
  }
}

// y pass: tile of nl kx-columns for one z plane.
__global__ __launch_bounds__(256) void k_fft_y(SpecArgs S, FftPlan P, int nl, int sign) {
  extern __shared__ double2 sh[];
  const Grid& g = S.g;
  const int N = g.n2;
  unsigned t = xcd_block(blockIdx.x, gridDim.x);
  int ntm = S.n1hp / nl;
  int m0 = (t % ntm) * nl, k = t / ntm;
  size_t base = (size_t)m0 + (size_t)S.n1hp * (size_t)g.n2 * k;
  for (int c = threadIdx.x; c < nl * N; c += blockDim.x) {
    int l = c % nl, p = c / nl;
    sh[p * nl + l] = S.spec[base + l + (size_t)S.n1hp * p];
  }
  __syncthreads();
  lds_fft(sh, nl, 1, nl, true, P, sign);
  for (int c = threadIdx.x; c < nl * N; c += blockDim.x) {
    int l = c % nl, p = c / nl;
    S.spec[base + l + (size_t)S.n1hp * p] = sh[p * nl + l];
  }
}

// z pass: forward FFT, 1/(N1 N2 N3), division by -(ak3+ak1+ak2) with the zero
// mode set to 0 (phcalc.f90:40-73), inverse FFT.
__global__ __launch_bounds__(256) void k_fft_z_solve(SpecArgs S, FftPlan P, int nl, const double* __restrict__ ak1,
                                                     const double* __restrict__ ak2, const double* __restrict__ ak3, double nrm) {
  extern __shared__ double2 sh[];
  const Grid& g = S.g;
  const int N = g.n3;
  unsigned t = xcd_block(blockIdx.x, gridDim.x);
  int ntm = S.n1hp / nl;
  int m0 = (t % ntm) * nl, j = t / ntm;
  size_t plane = (size_t)S.n1hp * g.n2;
  size_t base = (size_t)m0 + (size_t)S.n1hp * j;
  for (int c = threadIdx.x; c < nl * N; c += blockDim.x) {
    int l = c % nl, p = c / nl;
    sh[p * nl + l] = S.spec[base + l + plane * p];
  }
  __syncthreads();
  lds_fft(sh, nl, 1, nl, true, P, -1);
  for (int c = threadIdx.x; c < nl * N; c += blockDim.x) {
    int l = c % nl, p = c / nl, m = m0 + l;
    double2 v = sh[p * nl + l];
    if (m < S.n1h) {
      v.x = v.x / nrm; v.y = v.y / nrm;
      double ev = -ak3[p] - ak1[m] - ak2[j];
      if (m == 0 && j == 0 && p == 0) { v.x = 0.0; v.y = 0.0; }
      else { v.x = v.x / ev; v.y = v.y / ev; }
    } else {
      v = make_double2(0.0, 0.0);
    }
    sh[p * nl + l] = v;
  }
  __syncthreads();
  lds_fft(sh, nl, 1, nl, true, P, +1);
  for (int c = threadIdx.x; c < nl * N; c += blockDim.x) {
    int l = c % nl, p = c / nl;
    S.spec[base + l + plane * p] = sh[p * nl + l];
  }
}

// x pass, backward (complex-to-real, FFTW c2r semantics): rebuild the full
// Hermitian line, inverse FFT, keep the real part.
__global__ __launch_bounds__(256) void k_fft_x_bwd(SpecArgs S, FftPlan P, int nl, double* __restrict__ dph) {
  extern __shared__ double2 sh[];
  const Grid& g = S.g;
  const int N = g.n1, ls = N + 1;
  unsigned t = xcd_block(blockIdx.x, gridDim.x);
  int ntj = g.n2 / nl;
  int j0 = (t % ntj) * nl, k = t / ntj;
  for (int c = threadIdx.x; c < nl * S.n1h; c += blockDim.x) {
    int l = c / S.n1h, m = c % S.n1h;
    double2 v = S.spec[(size_t)m + (size_t)S.n1hp * ((size_t)(j0 + l) + (size_t)g.n2 * k)];
    sh[l * ls + m] = v;
    if (m > 0 && m < N - m) sh[l * ls + (N - m)] = make_double2(v.x, -v.y);
  }
  __syncthreads();
  lds_fft(sh, nl, ls, 1, false, P, +1);
  for (int c = threadIdx.x; c < nl * N; c += blockDim.x) {
    int l = c / N, i = c % N;
    dph[(size_t)i + (size_t)N * ((size_t)(j0 + l) + (size_t)g.n2 * k)] = sh[l * ls + i].x;
  }
}

bool make_plan(int N, FftPlan& P) {
  P.N = N;
  P.nst = 0;
  int m = N;
  while (m % 4 == 0) { P.radix[P.nst++] = 4; m /= 4; }
  while (m % 2 == 0) { P.radix[P.nst++] = 2; m /= 2; }
  while (m % 3 == 0) { P.radix[P.nst++] = 3; m /= 3; }
  return m == 1 && P.nst <= 24;
}

int largest_divisor_le(int n, int cap) {
  for (int d = cap; d >= 1; --d)
    if (n % d == 0) return d;
  return 1;
}

}  // namespace

int poisson_init(melt_ctx* ctx) {
  int n[3] = {ctx->n1, ctx->n2, ctx->n3};
  double2** tw[3] = {&ctx->tw1, &ctx->tw2, &ctx->tw3};
  double** ak[3] = {&ctx->ak1, &ctx->ak2, &ctx->ak3};
  double len[3] = {ctx->cfg.xlen, ctx->cfg.ylen, ctx->cfg.zlen};
  for (int d = 0; d < 3; ++d) {
    FftPlan P;
    if (!make_plan(n[d], P)) return melt_fail(ctx, MELT_ERR_ARG, "poisson_init", "grid size must be 2^a 3^b");
    std::vector<double2> h(n[d]);
    for (int t = 0; t < n[d]; ++t) {
      long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)n[d];
      h[t] = make_double2((double)cosl(ang), (double)-sinl(ang));
    }
    MELT_HIP(hipMalloc(tw[d], sizeof(double2) * n[d]));
    MELT_HIP(hipMemcpy(*tw[d], h.data(), sizeof(double2) * n[d], hipMemcpyHostToDevice));
    // phini.f90:35-55, 0-based index m -> reference index m+1
    std::vector<double> a(n[d]);
    int nm = n[d], nmh = nm / 2 + 1;
    for (int i = 1; i <= nm; ++i) {
      double ao = i <= nmh ? (double)(i - 1) * 2.0 * ctx->pi : -(double)(nm - i + 1) * 2.0 * ctx->pi;
      double r = (double)nm / len[d];
      a[i - 1] = 2.0 * (1.0 - cos(ao / nm)) * (r * r);
    }
    MELT_HIP(hipMalloc(ak[d], sizeof(double) * n[d]));
    MELT_HIP(hipMemcpy(*ak[d], a.data(), sizeof(double) * n[d], hipMemcpyHostToDevice));
  }
  size_t sz = sizeof(double2) * (size_t)ctx->n1hp * ctx->n2 * ctx->n3;
  MELT_HIP(hipMalloc(&ctx->spec, sz));
  MELT_HIP(hipMemset(ctx->spec, 0, sz));
  return 0;
}

// fused_divg = 1: dph := divergence / (al dt) is formed on the fly from the
// velocities (divg.f90) and never stored; 0: transforms MELT_DPH as phcalc.
int poisson_solve(melt_ctx* ctx, int fused_divg) {
  SpecArgs S;
  S.g = make_grid(ctx);
  S.n1h = ctx->n1h; S.n1hp = ctx->n1hp; S.spec = ctx->spec;
  FftPlan Px, Py, Pz;
  make_plan(ctx->n1, Px); make_plan(ctx->n2, Py); make_plan(ctx->n3, Pz);
  Px.tw = ctx->tw1; Py.tw = ctx->tw2; Pz.tw = ctx->tw3;
  hipEvent_t ev = nullptr;

  int nlx = largest_divisor_le(ctx->n2, std::max(1, FFT_TILE / ctx->n1));
  size_t shx = sizeof(double2) * (size_t)nlx * (ctx->n1 + 1);
  double usdtal = 1.0 / (ctx->dt * ctx->al);
  timing_begin(ctx, SLOT_DIVG_FFTX, &ev);
  hipLaunchKernelGGL(k_fft_x_fwd, dim3((unsigned)((ctx->n2 / nlx) * ctx->n3)), dim3(256), shx, ctx->stream, S, Px, nlx, ctx->f[MELT_VX],
                     ctx->f[MELT_VY], ctx->f[MELT_VZ], fused_divg ? nullptr : ctx->f[MELT_DPH], usdtal);
  MELT_HIP(hipGetLastError());
  timing_end(ctx, SLOT_DIVG_FFTX, ev);

  int nly = largest_divisor_le(ctx->n1hp, std::max(1, FFT_TILE / ctx->n2));
  size_t shy = sizeof(double2) * (size_t)nly * ctx->n2;
  timing_begin(ctx, SLOT_FFTY, &ev);
  hipLaunchKernelGGL(k_fft_y, dim3((unsigned)((ctx->n1hp / nly) * ctx->n3)), dim3(256), shy, ctx->stream, S, Py, nly, -1);
  MELT_HIP(hipGetLastError());
  timing_end(ctx, SLOT_FFTY, ev);

  int nlz = largest_divisor_le(ctx->n1hp, std::max(1, FFT_TILE / ctx->n3));
  size_t shz = sizeof(double2) * (size_t)nlz * ctx->n3;
  double nrm = (double)((long long)ctx->n1 * ctx->n2 * ctx->n3);
  timing_begin(ctx, SLOT_FFTZ_SOLVE, &ev);
  hipLaunchKernelGGL(k_fft_z_solve, dim3((unsigned)((ctx->n1hp / nlz) * ctx->n2)), dim3(256), shz, ctx->stream, S, Pz, nlz, ctx->ak1,
                     ctx->ak2, ctx->ak3, nrm);
  MELT_HIP(hipGetLastError());
  timing_end(ctx, SLOT_FFTZ_SOLVE, ev);

  timing_begin(ctx, SLOT_IFFTY, &ev);
  hipLaunchKernelGGL(k_fft_y, dim3((unsigned)((ctx->n1hp / nly) * ctx->n3)), dim3(256), shy, ctx->stream, S, Py, nly, +1);
  MELT_HIP(hipGetLastError());
  timing_end(ctx, SLOT_IFFTY, ev);

  timing_begin(ctx, SLOT_IFFTX, &ev);
  hipLaunchKernelGGL(k_fft_x_bwd, dim3((unsigned)((ctx->n2 / nlx) * ctx->n3)), dim3(256), shx, ctx->stream, S, Px, nlx, ctx->f[MELT_DPH]);
  MELT_HIP(hipGetLastError());
  timing_end(ctx, SLOT_IFFTX, ev);
  return 0;
}
