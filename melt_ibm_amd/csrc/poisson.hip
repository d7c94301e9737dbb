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
// The real transforms along x pack two real lines into one complex line
// (z = a + i b) and separate the two half spectra afterwards (inverse: the
// two Hermitian half spectra are packed the same way), which halves their LDS
// work.  Radix stages are templates, so every butterfly lives in registers.
//
// Spectrum layout: complex (n1hp x N2 x N3), kx fastest, n1hp = N1/2+1 rounded
// up to a multiple of 8 so every row starts on a 128-byte boundary.
#include <cmath>
#include <cstring>

#include "common.h"

namespace {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T x, y;
};
template <typename T>
__device__ __forceinline__ cx<T> mkc(T a, T b) {
  cx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return mkc<T>(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return mkc<T>(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) { return mkc<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

constexpr int FFT_NT = 256;   // threads per FFT block
constexpr int FFT_EPT = 8;    // LDS elements per thread: a tile holds <= 2048 complex values

struct FftPlan {
  int N, nst;
  int radix[24];
};

// One Stockham stage of radix R over the tile: element (line l, position p)
// at s[l*ls + p*ps].  LFAST: consecutive butterflies run over lines (tile
// layouts with the line index contiguous) instead of positions.
template <typename T, int R, bool LFAST>
__device__ __forceinline__ void fft_stage(cx<T>* s, int nl, int ls, int ps, int N, int Ns, const cx<T>* __restrict__ tw, bool inv) {
  constexpr int Q = (FFT_EPT + R - 1) / R;  // butterflies per thread
  const int NR = N / R, nbf = nl * NR, tid = threadIdx.x, nt = blockDim.x;
  const int step = N / (Ns * R);
  cx<T> v[Q][R];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    int b = tid + q * nt;
    if (b < nbf) {
      int l = LFAST ? b % nl : b / NR;
      int j = LFAST ? b / nl : b % NR;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = s[l * ls + (j + r * NR) * ps];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    int b = tid + q * nt;
    if (b < nbf) {
      int l = LFAST ? b % nl : b / NR;
      int j = LFAST ? b / nl : b % NR;
      int k = j % Ns;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cx<T> w = tw[k * r * step];
        if (inv) w.y = -w.y;
        v[q][r] = cmul(v[q][r], w);
      }
      cx<T> y[R];
      if constexpr (R == 4) {
        cx<T> a0 = cadd(v[q][0], v[q][2]), a1 = csub(v[q][0], v[q][2]);
        cx<T> a2 = cadd(v[q][1], v[q][3]), d = csub(v[q][1], v[q][3]);
        cx<T> a3 = inv ? mkc<T>(-d.y, d.x) : mkc<T>(d.y, -d.x);  // +i d (backward) or -i d (forward)
        y[0] = cadd(a0, a2); y[1] = cadd(a1, a3); y[2] = csub(a0, a2); y[3] = csub(a1, a3);
      } else if constexpr (R == 2) {
        y[0] = cadd(v[q][0], v[q][1]); y[1] = csub(v[q][0], v[q][1]);
      } else {
        const T c = T(-0.5), sn = (inv ? T(1) : T(-1)) * T(0.86602540378443864676);
        cx<T> sm = cadd(v[q][1], v[q][2]), df = csub(v[q][1], v[q][2]);
        y[0] = cadd(v[q][0], sm);
        cx<T> t1 = mkc<T>(v[q][0].x + c * sm.x, v[q][0].y + c * sm.y);
        cx<T> t2 = mkc<T>(-sn * df.y, sn * df.x);  // i sn df
        y[1] = cadd(t1, t2); y[2] = csub(t1, t2);
      }
      int o = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) s[l * ls + (o + r * Ns) * ps] = y[r];
    }
  }
  __syncthreads();
}

// Unnormalised in-LDS FFT of nl lines (forward: exp(-i...), as FFTW_FORWARD).
// The tile must hold at most FFT_EPT * blockDim.x elements, be complete in LDS
// and visible (barrier) on entry; it is visible again on return.
template <typename T, bool LFAST>
__device__ void lds_fft(cx<T>* s, int nl, int ls, int ps, const FftPlan& P, const cx<T>* __restrict__ tw, bool inv) {
  int Ns = 1;
  for (int st = 0; st < P.nst; ++st) {
    const int R = P.radix[st];
    if (R == 4) fft_stage<T, 4, LFAST>(s, nl, ls, ps, P.N, Ns, tw, inv);
    else if (R == 2) fft_stage<T, 2, LFAST>(s, nl, ls, ps, P.N, Ns, tw, inv);
    else fft_stage<T, 3, LFAST>(s, nl, ls, ps, P.N, Ns, tw, inv);
    Ns *= R;
  }
}

template <typename T>
struct SpecArgs {
  Grid g;
  int n1h, n1hp;
  cx<T>* spec;
};

// x pass, forward: nl real lines (divergence divg.f90:17-20, or a stored field)
// -> bins 0..N1/2 of each, two lines per complex transform.  div_slots !=
// null: max |div u| in FP64 (divg.f90:36-50) is folded into those slots too.
template <typename Rf, typename T>
__global__ __launch_bounds__(FFT_NT) void k_fft_x_fwd(SpecArgs<T> S, FftPlan P, const cx<T>* __restrict__ tw, int nl,
                                                      const Rf* __restrict__ vx, const Rf* __restrict__ vy, const Rf* __restrict__ vz,
                                                      const Rf* __restrict__ src, T usdtal, const int* __restrict__ gate,
                                                      unsigned long long* __restrict__ div_slots) {
  if (gate && *gate == 0) return;
  extern __shared__ unsigned char smem[];
  cx<T>* sh = reinterpret_cast<cx<T>*>(smem);
  T* shr = reinterpret_cast<T*>(smem);
  const Grid& g = S.g;
  const int N = g.n1, ls = N + 1;
  unsigned t = xcd_block(blockIdx.x, gridDim.x);
  int jt, k;
  tiled_2d(t, g.n2 / nl, g.n3, 16, 8, jt, k);  // the divergence reads lines j+1 and planes k+1
  const int j0 = jt * nl;
  const T dx1 = (T)g.dx1, dx2 = (T)g.dx2, dx3 = (T)g.dx3;
  const int kp = k == g.n3 - 1 ? 0 : k + 1;
  double dmax = 0.0;
  for (int c = threadIdx.x; c < nl * N; c += blockDim.x) {
    int l = c / N, i = c % N, j = j0 + l;
    size_t idx = (size_t)i + (size_t)N * ((size_t)j + (size_t)g.n2 * k);
    T v;
    if (src) {
      v = (T)src[idx];
    } else {
      int ip = i == N - 1 ? 0 : i + 1, jp = j == g.n2 - 1 ? 0 : j + 1;
      const Rf x0 = vx[idx], x1 = vx[idx - i + ip], y0 = vy[idx], y1 = vy[(size_t)i + (size_t)N * ((size_t)jp + (size_t)g.n2 * k)];
      const Rf z0 = vz[idx], z1 = vz[(size_t)i + (size_t)N * ((size_t)j + (size_t)g.n2 * kp)];
      T dq = ((T)x1 - (T)x0) * dx1 + ((T)y1 - (T)y0) * dx2 + ((T)z1 - (T)z0) * dx3;
      v = dq * usdtal;
      if (div_slots)
        dmax = fmax(dmax, fabs(((double)x1 - (double)x0) * g.dx1 + ((double)y1 - (double)y0) * g.dx2 + ((double)z1 - (double)z0) * g.dx3));
    }
    shr[2 * ((l >> 1) * ls + i) + (l & 1)] = v;  // line 2p -> real part, 2p+1 -> imaginary part
  }
  if (div_slots) block_max_to_slots(dmax, div_slots);
  __syncthreads();
  lds_fft<T, false>(sh, nl / 2, ls, 1, P, tw, false);
  const int np = nl / 2;
  for (int c = threadIdx.x; c < np * S.n1hp; c += blockDim.x) {
    int pl = c / S.n1hp, m = c % S.n1hp;
    cx<T> a = mkc<T>(0, 0), b = mkc<T>(0, 0);
    if (m < S.n1h) {
      cx<T> z = sh[pl * ls + m], zr = sh[pl * ls + (m == 0 ? 0 : N - m)];
      // A = (Z[m] + conj Z[N-m]) / 2,  B = (Z[m] - conj Z[N-m]) / (2i)
      a = mkc<T>((z.x + zr.x) * T(0.5), (z.y - zr.y) * T(0.5));
      b = mkc<T>((z.y + zr.y) * T(0.5), (zr.x - z.x) * T(0.5));
    }
    size_t row = (size_t)S.n1hp * ((size_t)(j0 + 2 * pl) + (size_t)g.n2 * k);
    S.spec[row + m] = a;
    S.spec[row + S.n1hp + m] = b;
  }
}

// y pass: tile of nl kx-columns (the last tile may be partial) for one z plane.
template <typename T>
__global__ __launch_bounds__(FFT_NT) void k_fft_y(SpecArgs<T> S, FftPlan P, const cx<T>* __restrict__ tw, int nl, int inv,
                                                  const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  extern __shared__ unsigned char smem[];
  cx<T>* sh = reinterpret_cast<cx<T>*>(smem);
  const Grid& g = S.g;
  const int N = g.n2;
  unsigned t = xcd_block(blockIdx.x, gridDim.x);
  int ntm = (S.n1hp + nl - 1) / nl;
  int m0 = (t % ntm) * nl, k = t / ntm;
  size_t base = (size_t)m0 + (size_t)S.n1hp * (size_t)g.n2 * k;
  for (int c = threadIdx.x; c < nl * N; c += blockDim.x) {
    int l = c % nl, p = c / nl;
    sh[p * nl + l] = m0 + l < S.n1hp ? S.spec[base + l + (size_t)S.n1hp * p] : mkc<T>(0, 0);
  }
  __syncthreads();
  lds_fft<T, true>(sh, nl, 1, nl, P, tw, inv != 0);
  for (int c = threadIdx.x; c < nl * N; c += blockDim.x) {
    int l = c % nl, p = c / nl;
    if (m0 + l < S.n1hp) S.spec[base + l + (size_t)S.n1hp * p] = sh[p * nl + l];
  }
}

// z pass: forward FFT, 1/(N1 N2 N3), division by -(ak3+ak1+ak2) with the zero
// mode set to 0 (phcalc.f90:40-73), inverse FFT.
template <typename T>
__global__ __launch_bounds__(FFT_NT) void k_fft_z_solve(SpecArgs<T> S, FftPlan P, const cx<T>* __restrict__ tw, int nl,
                                                        const double* __restrict__ ak1, const double* __restrict__ ak2,
                                                        const double* __restrict__ ak3, double nrm, const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  extern __shared__ unsigned char smem[];
  cx<T>* sh = reinterpret_cast<cx<T>*>(smem);
  const Grid& g = S.g;
  const int N = g.n3;
  unsigned t = xcd_block(blockIdx.x, gridDim.x);
  int ntm = (S.n1hp + nl - 1) / nl;
  int m0 = (t % ntm) * nl, j = t / ntm;
  size_t plane = (size_t)S.n1hp * g.n2;
  size_t base = (size_t)m0 + (size_t)S.n1hp * j;
  for (int c = threadIdx.x; c < nl * N; c += blockDim.x) {
    int l = c % nl, p = c / nl;
    sh[p * nl + l] = m0 + l < S.n1hp ? S.spec[base + l + plane * p] : mkc<T>(0, 0);
  }
  __syncthreads();
  lds_fft<T, true>(sh, nl, 1, nl, P, tw, false);
  for (int c = threadIdx.x; c < nl * N; c += blockDim.x) {
    int l = c % nl, p = c / nl, m = m0 + l;
    cx<T> v = sh[p * nl + l];
    if (m < S.n1h) {
      double ev = -ak3[p] - ak1[m] - ak2[j];
      double vx = (double)v.x / nrm, vy = (double)v.y / nrm;
      if (m == 0 && j == 0 && p == 0) { vx = 0.0; vy = 0.0; }
      else { vx = vx / ev; vy = vy / ev; }
      v = mkc<T>((T)vx, (T)vy);
    } else {
      v = mkc<T>(0, 0);
    }
    sh[p * nl + l] = v;
  }
  __syncthreads();
  lds_fft<T, true>(sh, nl, 1, nl, P, tw, true);
  for (int c = threadIdx.x; c < nl * N; c += blockDim.x) {
    int l = c % nl, p = c / nl;
    if (m0 + l < S.n1hp) S.spec[base + l + plane * p] = sh[p * nl + l];
  }
}

// x pass, backward (complex-to-real with FFTW c2r semantics: the imaginary
// parts of bins 0 and N1/2 are ignored): the half spectra A, B of lines 2p,
// 2p+1 are packed as Z = A + iB over the full Hermitian range, one inverse
// FFT gives line 2p in the real and line 2p+1 in the imaginary part.
template <typename Rf, typename T>
__global__ __launch_bounds__(FFT_NT) void k_fft_x_bwd(SpecArgs<T> S, FftPlan P, const cx<T>* __restrict__ tw, int nl, Rf* __restrict__ dst,
                                                      const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  extern __shared__ unsigned char smem[];
  cx<T>* sh = reinterpret_cast<cx<T>*>(smem);
  const Grid& g = S.g;
  const int N = g.n1, ls = N + 1, np = nl / 2;
  unsigned t = xcd_block(blockIdx.x, gridDim.x);
  int ntj = g.n2 / nl;
  int j0 = (t % ntj) * nl, k = t / ntj;
  for (int c = threadIdx.x; c < np * S.n1h; c += blockDim.x) {
    int pl = c / S.n1h, m = c % S.n1h;
    size_t row = (size_t)S.n1hp * ((size_t)(j0 + 2 * pl) + (size_t)g.n2 * k);
    cx<T> a = S.spec[row + m], b = S.spec[row + S.n1hp + m];
    if (m == 0 || 2 * m == N) { a.y = T(0); b.y = T(0); }
    sh[pl * ls + m] = mkc<T>(a.x - b.y, a.y + b.x);                         // A + iB
    if (m > 0 && 2 * m < N) sh[pl * ls + (N - m)] = mkc<T>(a.x + b.y, b.x - a.y);  // conj(A) + i conj(B)
  }
  __syncthreads();
  lds_fft<T, false>(sh, np, ls, 1, P, tw, true);
  for (int c = threadIdx.x; c < nl * N; c += blockDim.x) {
    int l = c / N, i = c % N;
    cx<T> z = sh[(l >> 1) * ls + i];
    dst[(size_t)i + (size_t)N * ((size_t)(j0 + l) + (size_t)g.n2 * k)] = (Rf)((l & 1) ? z.y : z.x);
  }
}

// ---------------------------------------------------------------- register FFTs
// For lines of N = P*P points (P = 8 or 16) the y and z passes use the
// two-level form X[k1 + P k2] = sum_b W_P^(b k2) W_N^(b k1) sum_a x[b + P a] W_P^(a k1):
// thread b of a line loads x[b + P a] straight from HBM, does a P-point DFT in
// registers, scales by W_N^(b k1) (twiddles of N staged in LDS), and one LDS
// transpose hands it the P values of its k1 for the second P-point DFT, after
// which it holds X[k1 + P k2] for every k2.  The inverse runs the same steps
// from that layout in reverse.  One exchange (two barriers) per transform
// instead of one per radix stage.

// cos(pi m / 8) and sin(pi m / 8), m = 0..15 (W_16^m = cos - i sin forward)
template <typename T>
__device__ __forceinline__ cx<T> w16(int m, bool inv) {
  const T c1 = T(0.92387953251128675613), s1 = T(0.38268343236508977173), r = T(0.70710678118654752440);
  const T cs[16] = {T(1), c1, r, s1, T(0), -s1, -r, -c1, T(-1), -c1, -r, -s1, T(0), s1, r, c1};
  const T sn[16] = {T(0), s1, r, c1, T(1), c1, r, s1, T(0), -s1, -r, -c1, T(-1), -c1, -r, -s1};
  return mkc<T>(cs[m], inv ? sn[m] : -sn[m]);
}

// v * W_16^M with the exact forms for M = 0, 4, 8, 12
template <typename T, int M>
__device__ __forceinline__ cx<T> mul_w16(cx<T> v, bool inv) {
  constexpr int m = M % 16;
  if constexpr (m == 0) return v;
  else if constexpr (m == 8) return mkc<T>(-v.x, -v.y);
  else if constexpr (m == 4) return inv ? mkc<T>(-v.y, v.x) : mkc<T>(v.y, -v.x);    // +i v or -i v
  else if constexpr (m == 12) return inv ? mkc<T>(v.y, -v.x) : mkc<T>(-v.y, v.x);
  else return cmul(v, w16<T>(m, inv));
}

// 4-point DFT of (v[o], v[o+s], v[o+2s], v[o+3s]) in place, natural order.
template <typename T>
__device__ __forceinline__ void dft4(cx<T>& x0, cx<T>& x1, cx<T>& x2, cx<T>& x3, bool inv) {
  cx<T> a0 = cadd(x0, x2), a1 = csub(x0, x2), a2 = cadd(x1, x3), d = csub(x1, x3);
  cx<T> a3 = inv ? mkc<T>(-d.y, d.x) : mkc<T>(d.y, -d.x);
  x0 = cadd(a0, a2); x1 = cadd(a1, a3); x2 = csub(a0, a2); x3 = csub(a1, a3);
}

// P-point DFT (P = 8, 16) of v[0..P-1] in registers; result in natural order.
template <typename T, int P>
__device__ __forceinline__ void dft_reg(cx<T> (&v)[P], bool inv) {
  if constexpr (P == 16) {
    // a = 4 a1 + a0: DFT4 over a1, twiddle W_16^(a0 k0), DFT4 over a0
#pragma unroll
    for (int a0 = 0; a0 < 4; ++a0) dft4(v[a0], v[a0 + 4], v[a0 + 8], v[a0 + 12], inv);   // v[a0 + 4 k0]
    v[5] = mul_w16<T, 1>(v[5], inv); v[9] = mul_w16<T, 2>(v[9], inv); v[13] = mul_w16<T, 3>(v[13], inv);
    v[6] = mul_w16<T, 2>(v[6], inv); v[10] = mul_w16<T, 4>(v[10], inv); v[14] = mul_w16<T, 6>(v[14], inv);
    v[7] = mul_w16<T, 3>(v[7], inv); v[11] = mul_w16<T, 6>(v[11], inv); v[15] = mul_w16<T, 9>(v[15], inv);
    cx<T> o[16];
#pragma unroll
    for (int k0 = 0; k0 < 4; ++k0) {
      cx<T> x0 = v[4 * k0], x1 = v[4 * k0 + 1], x2 = v[4 * k0 + 2], x3 = v[4 * k0 + 3];
      dft4(x0, x1, x2, x3, inv);
      o[k0] = x0; o[k0 + 4] = x1; o[k0 + 8] = x2; o[k0 + 12] = x3;
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = o[q];
  } else {
    static_assert(P == 8, "register DFT of 8 or 16 points");
    // a = 2 a1 + a0: DFT4 over a1, twiddle W_8^(a0 k0) = W_16^(2 a0 k0), DFT2 over a0
    dft4(v[0], v[2], v[4], v[6], inv);
    dft4(v[1], v[3], v[5], v[7], inv);
    v[3] = mul_w16<T, 2>(v[3], inv); v[5] = mul_w16<T, 4>(v[5], inv); v[7] = mul_w16<T, 6>(v[7], inv);
    cx<T> o[8];
#pragma unroll
    for (int k0 = 0; k0 < 4; ++k0) {
      cx<T> e = v[2 * k0], f = v[2 * k0 + 1];
      o[k0] = cadd(e, f); o[k0 + 4] = csub(e, f);
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = o[q];
  }
}

constexpr int RF_NC = 8;   // kx columns per register-FFT block (8 x 16 B = one 128-byte line per row)
constexpr int RF_NCZ = 8;  // the same for the z pass (16 and 32 measured slower)

// LDS image of one block of NC lines: the exchange tile (NC x P rows of
// pitch P+1) and the N twiddles.
template <typename T, int P, int NC>
struct RegFftLds {
  static constexpr int lines = NC;
  // The transposes move the real and then the imaginary parts through one
  // tile of T (half the LDS of a complex tile, so more blocks fit per CU).
  // Column c's P x (P+1) tile starts at c * CS; CS = 2 mod 16 puts the 16
  // lanes of a 64-bit access in distinct bank pairs in both thread layouts
  // (y/z: 8 columns x 2 threads per lane group; x: 16 threads of one column).
  static constexpr int CS = P * (P + 1) + ((2 - P * (P + 1) % 16) + 16) % 16;
  T ex[NC * CS];
  cx<T> tw[P * P];
};

// v[q] of thread b of column c becomes v[b] of thread q (a P x P transpose
// within the column), one component at a time.  The tile must be free on
// entry (callers put a barrier after any earlier read of it).
template <typename T, int P, int NC>
__device__ __forceinline__ void reg_transpose(cx<T> (&v)[P], RegFftLds<T, P, NC>& L, int c, int b) {
  T* e = L.ex + c * L.CS;
#pragma unroll
  for (int k1 = 0; k1 < P; ++k1) e[k1 * (P + 1) + b] = v[k1].x;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < P; ++q) v[q].x = e[b * (P + 1) + q];
  __syncthreads();
#pragma unroll
  for (int k1 = 0; k1 < P; ++k1) e[k1 * (P + 1) + b] = v[k1].y;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < P; ++q) v[q].y = e[b * (P + 1) + q];
}

// One transform of the lines (column c, thread b) held as v[a] = x[b + P a]:
// on return v[k2] = X[k1 + P k2] for k1 = b.
template <typename T, int P, int NC>
__device__ __forceinline__ void reg_fft_fwd(cx<T> (&v)[P], RegFftLds<T, P, NC>& L, int c, int b, bool inv) {
  constexpr int N = P * P;
  dft_reg<T, P>(v, inv);
#pragma unroll
  for (int k1 = 1; k1 < P; ++k1) {
    cx<T> w = L.tw[(b * k1) % N];
    if (inv) w.y = -w.y;
    v[k1] = cmul(v[k1], w);
  }
  reg_transpose<T, P, NC>(v, L, c, b);
  dft_reg<T, P>(v, inv);
}

// Inverse of that layout: v[k2] = X[k1 + P k2] with k1 = b in, v[n2] =
// x[b + P n2] out (unnormalised, like FFTW_BACKWARD when inv, else forward).
template <typename T, int P, int NC>
__device__ __forceinline__ void reg_fft_back(cx<T> (&v)[P], RegFftLds<T, P, NC>& L, int c, int b, bool inv) {
  constexpr int N = P * P;
  dft_reg<T, P>(v, inv);   // over k2: v[n1] = Z[k1 = b][n1]
#pragma unroll
  for (int n1 = 1; n1 < P; ++n1) {
    cx<T> w = L.tw[(b * n1) % N];
    if (inv) w.y = -w.y;
    v[n1] = cmul(v[n1], w);
  }
  __syncthreads();   // every thread has read its forward exchange
  reg_transpose<T, P, NC>(v, L, c, b);
  dft_reg<T, P>(v, inv);   // over k1: v[n2] = x[b + P n2]
}

// y pass (N2 = P*P) with register FFTs: RF_NC kx columns x P threads per block.
template <typename T, int P>
__global__ __launch_bounds__(RF_NC * P) void k_fft_y_reg(SpecArgs<T> S, const cx<T>* __restrict__ tw, int inv, const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  __shared__ RegFftLds<T, P, RF_NC> L;
  constexpr int N = P * P;
  const int t = threadIdx.x, c = t % RF_NC, b = t / RF_NC;
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const int ntm = S.n1hp / RF_NC;
  const int m0 = (int)(blk % (unsigned)ntm) * RF_NC, k = (int)(blk / (unsigned)ntm);
  for (int q = t; q < N; q += RF_NC * P) L.tw[q] = tw[q];
  cx<T>* base = S.spec + (size_t)m0 + c + (size_t)S.n1hp * N * k;
  const size_t rs = (size_t)S.n1hp;
  cx<T> v[P];
#pragma unroll
  for (int a = 0; a < P; ++a) v[a] = base[rs * (b + P * a)];
  __syncthreads();   // twiddles
  reg_fft_fwd<T, P>(v, L, c, b, inv != 0);
#pragma unroll
  for (int k2 = 0; k2 < P; ++k2) base[rs * (b + P * k2)] = v[k2];
}

// z pass with register FFTs: forward FFT, 1/(N1 N2 N3), division by
// -(ak3+ak1+ak2) with the zero mode set to 0 (phcalc.f90:40-73), inverse FFT.
template <typename T, int P, int NC = RF_NCZ>
__global__ __launch_bounds__(NC * P) void k_fft_z_solve_reg(SpecArgs<T> S, const cx<T>* __restrict__ tw, const double* __restrict__ ak1,
                                                               const double* __restrict__ ak2, const double* __restrict__ ak3, double nrm,
                                                               const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  __shared__ RegFftLds<T, P, NC> L;
  constexpr int N = P * P;
  const int t = threadIdx.x, c = t % NC, b = t / NC;
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const int ntm = (S.n1hp + NC - 1) / NC;   // the last block may overhang the row
  const int m0 = (int)(blk % (unsigned)ntm) * NC, j = (int)(blk / (unsigned)ntm);
  for (int q = t; q < N; q += NC * P) L.tw[q] = tw[q];
  const int m = m0 + c;
  const bool col = m < S.n1hp;   // columns past the row transform zeros and store nothing
  cx<T>* base = S.spec + (size_t)(col ? m : 0) + (size_t)S.n1hp * j;
  const size_t rs = (size_t)S.n1hp * S.g.n2;
  cx<T> v[P];
#pragma unroll
  for (int a = 0; a < P; ++a) v[a] = col ? base[rs * (b + P * a)] : mkc<T>(0, 0);
  __syncthreads();
  reg_fft_fwd<T, P>(v, L, c, b, false);
#pragma unroll
  for (int k2 = 0; k2 < P; ++k2) {
    const int p = b + P * k2;
    if (m < S.n1h) {
      double ev = -ak3[p] - ak1[m] - ak2[j];
      double vx = (double)v[k2].x / nrm, vy = (double)v[k2].y / nrm;
      if (m == 0 && j == 0 && p == 0) { vx = 0.0; vy = 0.0; }
      else { vx = vx / ev; vy = vy / ev; }
      v[k2] = mkc<T>((T)vx, (T)vy);
    } else {
      v[k2] = mkc<T>(0, 0);
    }
  }
  reg_fft_back<T, P>(v, L, c, b, true);
#pragma unroll
  for (int n2 = 0; n2 < P; ++n2)
    if (col) base[rs * (b + P * n2)] = v[n2];
}

// x pass, forward, with register FFTs (N1 = P*P): RF_XT threads per block,
// thread (c, b) owns line pair c of the block (lines j, j+1 packed as real and
// imaginary part, divergence formed at load as k_fft_x_fwd) and points
// b + P a.  After the transform the two half spectra are separated with the
// mirror bins N-m, exchanged through LDS.
constexpr int RF_XT = 128;

template <typename Rf, typename T, int P>
__global__ __launch_bounds__(RF_XT) void k_fft_x_fwd_reg(SpecArgs<T> S, const cx<T>* __restrict__ tw, const Rf* __restrict__ vx,
                                                         const Rf* __restrict__ vy, const Rf* __restrict__ vz, const Rf* __restrict__ src,
                                                         T usdtal, const int* __restrict__ gate, unsigned long long* __restrict__ div_slots) {
  if (gate && *gate == 0) return;
  constexpr int N = P * P, NPB = RF_XT / P;   // line pairs per block
  __shared__ RegFftLds<T, P, NPB> L;
  static_assert(NPB * N <= (int)(sizeof(L.ex) / sizeof(L.ex[0])), "mirror exchange fits the transpose tile");
  const Grid& g = S.g;
  const int t = threadIdx.x, b = t % P, c = t / P;
  int jt, k;
  tiled_2d(xcd_block(blockIdx.x, gridDim.x), g.n2 / (2 * NPB), g.n3, 16, 8, jt, k);  // the divergence reads rows j+1 and planes k+1
  const int j = jt * 2 * NPB + 2 * c, kp = k == g.n3 - 1 ? 0 : k + 1;
  for (int q = t; q < N; q += RF_XT) L.tw[q] = tw[q];
  const T dx1 = (T)g.dx1, dx2 = (T)g.dx2, dx3 = (T)g.dx3;
  double dmax = 0.0;
  cx<T> v[P];
#pragma unroll
  for (int a = 0; a < P; ++a) {
    const int i = b + P * a, ip = i == N - 1 ? 0 : i + 1;
    T z[2];
#pragma unroll
    for (int l = 0; l < 2; ++l) {
      const int jj = j + l, jp = jj == g.n2 - 1 ? 0 : jj + 1;
      const size_t idx = (size_t)i + (size_t)N * ((size_t)jj + (size_t)g.n2 * k);
      if (src) {
        z[l] = (T)src[idx];
      } else {
        const Rf x0 = vx[idx], x1 = vx[idx - i + ip], y0 = vy[idx], y1 = vy[(size_t)i + (size_t)N * ((size_t)jp + (size_t)g.n2 * k)];
        const Rf z0 = vz[idx], z1 = vz[(size_t)i + (size_t)N * ((size_t)jj + (size_t)g.n2 * kp)];
        T dq = ((T)x1 - (T)x0) * dx1 + ((T)y1 - (T)y0) * dx2 + ((T)z1 - (T)z0) * dx3;
        z[l] = dq * usdtal;
        if (div_slots)
          dmax = fmax(dmax, fabs(((double)x1 - (double)x0) * g.dx1 + ((double)y1 - (double)y0) * g.dx2 + ((double)z1 - (double)z0) * g.dx3));
      }
    }
    v[a] = mkc<T>(z[0], z[1]);
  }
  if (div_slots) block_max_to_slots(dmax, div_slots);
  __syncthreads();   // twiddles
  reg_fft_fwd<T, P>(v, L, c, b, false);   // v[k2] = Z[b + P k2]
  // mirror bins Z[N-m] of the line pair through LDS, one component at a time
  T* zs = L.ex + c * N;
  T zrx[P];
  __syncthreads();   // the transpose tile is read
#pragma unroll
  for (int k2 = 0; k2 < P; ++k2) zs[b + P * k2] = v[k2].x;
  __syncthreads();
#pragma unroll
  for (int k2 = 0; k2 < P; ++k2) {
    const int m = b + P * k2;
    zrx[k2] = zs[m == 0 ? 0 : N - m];
  }
  __syncthreads();
#pragma unroll
  for (int k2 = 0; k2 < P; ++k2) zs[b + P * k2] = v[k2].y;
  __syncthreads();
  const size_t row = (size_t)S.n1hp * ((size_t)j + (size_t)g.n2 * k);
#pragma unroll
  for (int k2 = 0; k2 < P; ++k2) {
    const int m = b + P * k2;
    if (m >= S.n1hp) break;
    cx<T> A = mkc<T>(0, 0), B = mkc<T>(0, 0);
    if (m < S.n1h) {
      const cx<T> z = v[k2], zr = mkc<T>(zrx[k2], zs[m == 0 ? 0 : N - m]);
      // A = (Z[m] + conj Z[N-m]) / 2,  B = (Z[m] - conj Z[N-m]) / (2i)
      A = mkc<T>((z.x + zr.x) * T(0.5), (z.y - zr.y) * T(0.5));
      B = mkc<T>((z.y + zr.y) * T(0.5), (zr.x - z.x) * T(0.5));
    }
    S.spec[row + m] = A;
    S.spec[row + S.n1hp + m] = B;
  }
}

// x pass, backward, with register FFTs (FFTW c2r semantics as k_fft_x_bwd):
// thread (c, b) packs Z = A + iB of line pair c for bins b + P k2 (the bins
// above N/2 as conj(A) + i conj(B) of the mirror bin), transforms, and writes
// line j from the real and line j+1 from the imaginary part.
template <typename Rf, typename T, int P>
__global__ __launch_bounds__(RF_XT) void k_fft_x_bwd_reg(SpecArgs<T> S, const cx<T>* __restrict__ tw, Rf* __restrict__ dst,
                                                         const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  constexpr int N = P * P, NPB = RF_XT / P;
  __shared__ RegFftLds<T, P, NPB> L;
  const Grid& g = S.g;
  const int t = threadIdx.x, b = t % P, c = t / P;
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const int ntj = g.n2 / (2 * NPB);
  const int j = (int)(blk % (unsigned)ntj) * 2 * NPB + 2 * c, k = (int)(blk / (unsigned)ntj);
  for (int q = t; q < N; q += RF_XT) L.tw[q] = tw[q];
  const size_t row = (size_t)S.n1hp * ((size_t)j + (size_t)g.n2 * k);
  cx<T> v[P];
#pragma unroll
  for (int k2 = 0; k2 < P; ++k2) {
    const int m = b + P * k2;
    if (m < S.n1h) {
      cx<T> A = S.spec[row + m], B = S.spec[row + S.n1hp + m];
      if (m == 0 || 2 * m == N) { A.y = T(0); B.y = T(0); }
      v[k2] = mkc<T>(A.x - B.y, A.y + B.x);   // A + iB
    } else {
      const cx<T> A = S.spec[row + (N - m)], B = S.spec[row + S.n1hp + (N - m)];
      v[k2] = mkc<T>(A.x + B.y, B.x - A.y);   // conj(A) + i conj(B)
    }
  }
  __syncthreads();   // twiddles
  reg_fft_back<T, P>(v, L, c, b, true);   // v[n2] = z[b + P n2]
  const size_t o0 = (size_t)N * ((size_t)j + (size_t)g.n2 * k);
#pragma unroll
  for (int n2 = 0; n2 < P; ++n2) {
    const int i = b + P * n2;
    dst[o0 + i] = (Rf)v[n2].x;
    dst[o0 + N + i] = (Rf)v[n2].y;
  }
}

// P with P*P = n, or 0
int reg_fft_p(int n) { return n == 256 ? 16 : n == 64 ? 8 : 0; }

bool make_plan(int N, FftPlan& P) {
  P.N = N;
  P.nst = 0;
  int m = N;
  while (m % 4 == 0) { P.radix[P.nst++] = 4; m /= 4; }
  while (m % 2 == 0) { P.radix[P.nst++] = 2; m /= 2; }
  while (m % 3 == 0) { P.radix[P.nst++] = 3; m /= 3; }
  return m == 1 && P.nst <= 24;
}

// Largest even divisor of n2 with (nl/2) * N1 complex values per tile.
int x_lines(int n1, int n2) {
  int cap = std::max(2, 2 * (FFT_EPT * FFT_NT) / n1);
  int best = 2;
  for (int d = 2; d <= std::min(cap, n2); d += 2)
    if (n2 % d == 0) best = d;
  return best;
}

}  // namespace

template <typename T>
int upload_twiddles(melt_ctx* ctx, int n, void** dst) {
  std::vector<cx<T>> h(n);
  for (int t = 0; t < n; ++t) {
    long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)n;
    h[t].x = (T)cosl(ang);
    h[t].y = (T)-sinl(ang);
  }
  MELT_HIP(hipMalloc(dst, sizeof(cx<T>) * n));
  MELT_HIP(hipMemcpy(*dst, h.data(), sizeof(cx<T>) * n, hipMemcpyHostToDevice));
  return 0;
}

int poisson_init(melt_ctx* ctx) {
  int n[3] = {ctx->n1, ctx->n2, ctx->n3};
  void** tw[3] = {&ctx->tw1, &ctx->tw2, &ctx->tw3};
  double** ak[3] = {&ctx->ak1, &ctx->ak2, &ctx->ak3};
  double len[3] = {ctx->cfg.xlen, ctx->cfg.ylen, ctx->cfg.zlen};
  if (ctx->n2 % 2) return melt_fail(ctx, MELT_ERR_ARG, "poisson_init", "n2m must be even");
  for (int d = 0; d < 3; ++d) {
    FftPlan P;
    if (!make_plan(n[d], P)) return melt_fail(ctx, MELT_ERR_ARG, "poisson_init", "grid size must be 2^a 3^b");
    int rc = ctx->spec64 ? upload_twiddles<double>(ctx, n[d], tw[d]) : upload_twiddles<float>(ctx, n[d], tw[d]);
    if (rc) return rc;
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
  size_t sz = 2 * (ctx->spec64 ? sizeof(double) : sizeof(float)) * (size_t)ctx->n1hp * ctx->n2 * ctx->n3;
  MELT_HIP(hipMalloc(&ctx->spec, sz));
  MELT_HIP(hipMemset(ctx->spec, 0, sz));
  return 0;
}

// fused_divg = 1: dph := divergence / (al dt) is formed on the fly from the
// velocities (divg.f90) and never stored; 0: transforms MELT_DPH as phcalc.
// gate != null: every kernel returns at once unless *gate != 0 (the
// refinement solve of the precision form).  Rf is the field type, T the
// spectrum's (and the transforms' arithmetic).
template <typename Rf, typename T>
int solve_t(melt_ctx* ctx, int fused_divg, const int* gate, unsigned long long* div_slots) {
  SpecArgs<T> S;
  S.g = make_grid(ctx);
  S.n1h = ctx->n1h; S.n1hp = ctx->n1hp; S.spec = static_cast<cx<T>*>(ctx->spec);
  FftPlan Px, Py, Pz;
  make_plan(ctx->n1, Px); make_plan(ctx->n2, Py); make_plan(ctx->n3, Pz);
  const cx<T>* twx = static_cast<const cx<T>*>(ctx->tw1);
  const cx<T>* twy = static_cast<const cx<T>*>(ctx->tw2);
  const cx<T>* twz = static_cast<const cx<T>*>(ctx->tw3);
  hipEvent_t ev = nullptr;

  int nlx = x_lines(ctx->n1, ctx->n2);
  size_t shx = sizeof(cx<T>) * (size_t)(nlx / 2) * (ctx->n1 + 1);
  T usdtal = T(1.0 / (ctx->dt * ctx->al));
  unsigned gx = (unsigned)((ctx->n2 / nlx) * ctx->n3);
  // x lines of 64 or 256 points: register FFTs (RF_XT threads, RF_XT / P line pairs per block)
  const int px = reg_fft_p(ctx->n1);
  const unsigned grx = px ? (unsigned)((ctx->n2 / (2 * RF_XT / px)) * ctx->n3) : 0;
  const Rf* src = fused_divg ? nullptr : fld<Rf>(ctx, MELT_DPH);
  unsigned long long* dsl = fused_divg ? div_slots : nullptr;
  if (!gate) timing_begin(ctx, SLOT_DIVG_FFTX, &ev);
  if (px == 16)
    hipLaunchKernelGGL((k_fft_x_fwd_reg<Rf, T, 16>), dim3(grx), dim3(RF_XT), 0, ctx->stream, S, twx, fld<Rf>(ctx, MELT_VX), fld<Rf>(ctx, MELT_VY),
                       fld<Rf>(ctx, MELT_VZ), src, usdtal, gate, dsl);
  else if (px == 8)
    hipLaunchKernelGGL((k_fft_x_fwd_reg<Rf, T, 8>), dim3(grx), dim3(RF_XT), 0, ctx->stream, S, twx, fld<Rf>(ctx, MELT_VX), fld<Rf>(ctx, MELT_VY),
                       fld<Rf>(ctx, MELT_VZ), src, usdtal, gate, dsl);
  else
    hipLaunchKernelGGL((k_fft_x_fwd<Rf, T>), dim3(gx), dim3(FFT_NT), shx, ctx->stream, S, Px, twx, nlx, fld<Rf>(ctx, MELT_VX), fld<Rf>(ctx, MELT_VY),
                       fld<Rf>(ctx, MELT_VZ), src, usdtal, gate, dsl);
  MELT_HIP(hipGetLastError());
  if (!gate) timing_end(ctx, SLOT_DIVG_FFTX, ev);

  int nly = std::max(1, std::min(ctx->n1hp, (FFT_EPT * FFT_NT) / ctx->n2));
  size_t shy = sizeof(cx<T>) * (size_t)nly * ctx->n2;
  unsigned gy = (unsigned)(((ctx->n1hp + nly - 1) / nly) * ctx->n3);
  // y and z lines of 64 or 256 points: register FFTs (RF_NC columns per block)
  const int py = reg_fft_p(ctx->n2), pz = reg_fft_p(ctx->n3);
  const unsigned gry = (unsigned)((ctx->n1hp / RF_NC) * ctx->n3), grz = (unsigned)(((ctx->n1hp + RF_NCZ - 1) / RF_NCZ) * ctx->n2);
  auto fft_y = [&](int inv) {
    if (py == 16) hipLaunchKernelGGL((k_fft_y_reg<T, 16>), dim3(gry), dim3(RF_NC * 16), 0, ctx->stream, S, twy, inv, gate);
    else if (py == 8) hipLaunchKernelGGL((k_fft_y_reg<T, 8>), dim3(gry), dim3(RF_NC * 8), 0, ctx->stream, S, twy, inv, gate);
    else hipLaunchKernelGGL((k_fft_y<T>), dim3(gy), dim3(FFT_NT), shy, ctx->stream, S, Py, twy, nly, inv, gate);
  };
  if (!gate) timing_begin(ctx, SLOT_FFTY, &ev);
  fft_y(0);
  MELT_HIP(hipGetLastError());
  if (!gate) timing_end(ctx, SLOT_FFTY, ev);

  int nlz = std::max(1, std::min(ctx->n1hp, (FFT_EPT * FFT_NT) / ctx->n3));
  size_t shz = sizeof(cx<T>) * (size_t)nlz * ctx->n3;
  unsigned gz = (unsigned)(((ctx->n1hp + nlz - 1) / nlz) * ctx->n2);
  double nrm = (double)((long long)ctx->n1 * ctx->n2 * ctx->n3);
  if (!gate) timing_begin(ctx, SLOT_FFTZ_SOLVE, &ev);
  if (pz == 16)
    hipLaunchKernelGGL((k_fft_z_solve_reg<T, 16>), dim3(grz), dim3(RF_NCZ * 16), 0, ctx->stream, S, twz, ctx->ak1, ctx->ak2, ctx->ak3, nrm, gate);
  else if (pz == 8)
    hipLaunchKernelGGL((k_fft_z_solve_reg<T, 8>), dim3(grz), dim3(RF_NCZ * 8), 0, ctx->stream, S, twz, ctx->ak1, ctx->ak2, ctx->ak3, nrm, gate);
  else
    hipLaunchKernelGGL((k_fft_z_solve<T>), dim3(gz), dim3(FFT_NT), shz, ctx->stream, S, Pz, twz, nlz, ctx->ak1, ctx->ak2, ctx->ak3, nrm, gate);
  MELT_HIP(hipGetLastError());
  if (!gate) timing_end(ctx, SLOT_FFTZ_SOLVE, ev);

  if (!gate) timing_begin(ctx, SLOT_IFFTY, &ev);
  fft_y(1);
  MELT_HIP(hipGetLastError());
  if (!gate) timing_end(ctx, SLOT_IFFTY, ev);

  if (!gate) timing_begin(ctx, SLOT_IFFTX, &ev);
  if (px == 16)
    hipLaunchKernelGGL((k_fft_x_bwd_reg<Rf, T, 16>), dim3(grx), dim3(RF_XT), 0, ctx->stream, S, twx, fld<Rf>(ctx, MELT_DPH), gate);
  else if (px == 8)
    hipLaunchKernelGGL((k_fft_x_bwd_reg<Rf, T, 8>), dim3(grx), dim3(RF_XT), 0, ctx->stream, S, twx, fld<Rf>(ctx, MELT_DPH), gate);
  else
    hipLaunchKernelGGL((k_fft_x_bwd<Rf, T>), dim3(gx), dim3(FFT_NT), shx, ctx->stream, S, Px, twx, nlx, fld<Rf>(ctx, MELT_DPH), gate);
  MELT_HIP(hipGetLastError());
  if (!gate) timing_end(ctx, SLOT_IFFTX, ev);
  return 0;
}

int poisson_solve(melt_ctx* ctx, int fused_divg, const int* gate, unsigned long long* div_slots) {
  if (!ctx->fp32) return solve_t<double, double>(ctx, fused_divg, gate, div_slots);
  return ctx->spec64 ? solve_t<float, double>(ctx, fused_divg, gate, div_slots) : solve_t<float, float>(ctx, fused_divg, gate, div_slots);
}
