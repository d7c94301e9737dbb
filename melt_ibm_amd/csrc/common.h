// Shared definitions of libmeltgpu: the device-resident context that replaces the
// reference's module globals (param.f90:4-274) and device helpers used by every
// kernel file.  All HIP code is compiled with -ffp-contract=off so that each
// expression rounds exactly as the reference's Fortran does.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "../../include/melt_ibm.h"

// ------------------------------------------------------------------ errors
#define MELT_HIP(call)                                                                 \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) return melt_fail(ctx, MELT_ERR_HIP, #call, hipGetErrorString(e_)); \
  } while (0)

// ------------------------------------------------------------- field store
// FP64 fields are N1*N2*N3 doubles, i fastest; the spectrum of the pressure
// solve is (n1hp complex) x N2 x N3 with n1hp = N1/2+1 rounded up to 8.
enum { RHS_U = 0, RHS_V, RHS_W, RHS_T };

// Per-kernel timing slots of the fused substep (names in api.hip).
enum melt_slot {
  SLOT_TAG = 0, SLOT_EXPLICIT, SLOT_SOLVE_X, SLOT_SOLVE_Y, SLOT_SOLVE_Z, SLOT_IBM, SLOT_LAGR,
  SLOT_DIVG_FFTX, SLOT_FFTY, SLOT_FFTZ_SOLVE, SLOT_IFFTY, SLOT_IFFTX, SLOT_PROJECT,
  SLOT_PH_TAG, SLOT_PH_EULER, SLOT_PH_MLS, SLOT_PH_PRESSURE, SLOT_SUBSTEP, SLOT_REFINE,
  SLOT_COUNT
};

// Device-side record of the immersed-boundary windows (dense index boxes that
// hold every cell touched by the 3x3x3 MLS supports / the 5x5x5 tagging band).
struct melt_window {
  int org[3];      // unwrapped 1-based origin of the current box
  int prev_org[3]; // box of the previous tagging call (its sdf must be reset)
  int prev_valid;
  int pad;
  int bmin[3], bmax[3];  // running min / max of the face base indices (reset after use)
};

// Per-face probe results of the MLS gradient / interpolation evaluations.
enum {
  PRB_QO = 0, PRB_QI, PRB_P, PRB_GP, PRB_GU = PRB_GP + 3, PRB_GV = PRB_GU + 3, PRB_GW = PRB_GV + 3,
  PRB_GT0 = PRB_GW + 3, PRB_GT1 = PRB_GT0 + 3,  // temperature gradients behind QO / QI
  PRB_N = PRB_GT1 + 3
};

// Launch-shape overrides (0 = built-in choice), set from MELT_TUNE at
// context creation ("xf_w=4,solve_w=16,...") for tile sweeps.
enum melt_tune { TUNE_XF_W = 0, TUNE_XF_NT, TUNE_SOLVE_W, TUNE_SOLVE_NT, TUNE_XF_Z, TUNE_YZ_JB, TUNE_STREAMS, TUNE_COEF, TUNE_PBUF, TUNE_XF_ONE, TUNE_N };

struct melt_ctx {
  melt_config cfg;
  int tune[TUNE_N] = {};
  int n1, n2, n3;
  size_t nc;
  int n1h, n1hp;
  double dx1, dx2, dx3, dx1q, dx2q, dx3q, pi;
  double gam[3], rom[3], alm[3];
  double al = 0, ga = 0, ro = 0, dt = 0, beta = 0;
  int stage = 0;
  hipStream_t stream = nullptr;
  // second stream of the fused substep: the Lagrangian update runs there,
  // concurrently with the pressure phase (fork / join events)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // third stream: the next substep's tagging, forked from the side stream as
  // soon as the faces are refreshed, so that it runs beside the rest of the
  // geometry refresh and the probe geometry stage; joined with the side stream
  hipStream_t side2 = nullptr;
  hipEvent_t ev_fork2 = nullptr, ev_join2 = nullptr;
  int join2_pending = 0;
  // called once by the next geometry refresh right after the face centroids
  // and normals are written (set by the fused substep, consumed by the call)
  const std::function<int()>* after_faces = nullptr;
  // HIT forcing synthesis beside the end of the step: ev_forc is recorded on
  // the main stream after the last explicit pass (the last reader of
  // forcx/y/z); while forc_free is set, nothing queued on the main stream
  // after it touches the forcing, so melt_hit_force may run on side2 from
  // ev_forc on (the main stream then waits for ev_hit)
  hipEvent_t ev_forc = nullptr, ev_hit = nullptr;
  int forc_free = 0;
  // graph replay of one RK3 step (melt_graph_enable): captured on first use,
  // dropped whenever a host-side launch parameter (dt, refinement) changes
  int use_graph = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int device = 0;
  size_t lds_max = 64 * 1024;   // dynamic LDS a workgroup may use (hipDeviceAttributeMaxSharedMemoryPerBlock)

  // Eulerian fields: double, or float in the precision form (fp32 = 1); fld<R>()
  int fp32 = 0;
  size_t esz = sizeof(double);
  // Poisson spectrum, twiddles and FFT arithmetic in FP64 (the FP64 form); the
  // precision form keeps them in FP32 (an FP64 spectrum there was measured to
  // leave its pressure difference to FP64 unchanged, at 0.28 ms per step)
  int spec64 = 1;
  void* f[MELT_NFIELDS] = {};
  void* rhs4[4] = {};
  void* tmp4[4] = {};      // output of the y-direction Green filters (the z pass reads it)
  int no_green = 0;        // MELT_TUNE green=0: cyclic reduction for every line solve
  void* spec = nullptr;                                    // complex, double if spec64 else float
  double *ak1 = nullptr, *ak2 = nullptr, *ak3 = nullptr;   // 0-based wavenumber index
  void *tw1 = nullptr, *tw2 = nullptr, *tw3 = nullptr;     // exp(-2 pi i t / N), complex of the field type
  // precision-form pressure refinement: up to refine_max correction solves while
  // the FP64 divergence after projection exceeds refine_rtol times the one before
  int refine_max = 0;                    // opt-in (melt_set_refinement)
  double refine_rtol = 1e-4;
  int* d_gate = nullptr;                 // device flag read by the gated refinement kernels
  long long* d_refine_stats = nullptr;   // [checks, refinements, bits of the last ratio]

  // body / mesh
  int has_geom = 0;
  int nv = 0, ne = 0, nf = 0;
  double rad_p = 0;
  melt_body* d_body = nullptr;   // device scalars
  melt_body h_body{};            // host mirror (refreshed by melt_get_body)
  int *d_vof = nullptr, *d_eof = nullptr, *d_voe = nullptr;  // 0-based connectivity
  int *d_v2f_ptr = nullptr, *d_v2f = nullptr;                 // vertex -> faces (ascending)
  unsigned char *d_gv = nullptr, *d_ge = nullptr, *d_gf = nullptr;
  int* d_foe = nullptr;          // faces of each edge (0-based), kept for snapshots of a remeshed body
  double* d_lag[MELT_L_NLAG] = {};
  // remeshing (IREMESH = 1, particle.f90:121-160): host mirror of the 0-based
  // connectivity and ghost flags (authoritative; the device copies follow each
  // edit), anchorVert / flagged_edge, the positions after the Stefan update
  // (the volume the smoothing restores), the device's short-edge flag read
  // through pinned memory, and the counters of melt_get_remesh_stats
  std::vector<int32_t> h_vof, h_eof, h_voe, h_foe, h_gv, h_ge, h_gf, h_anchor, h_flagged;
  double* d_xyz_melt = nullptr;
  int32_t* h_rm = nullptr;
  hipEvent_t ev_rm = nullptr;
  double remesh_stats[8] = {};
  uint64_t mesh_hash0 = 0;       // FNV-1a of the imported vert_of_face (snapshot compatibility)
  int* d_pind = nullptr;
  int* d_pind2 = nullptr;        // probe indices of the fused substep, copied into d_pind after the spreading
  double* d_probe = nullptr;     // PRB_N * nf
  double *d_rxp = nullptr, *d_rxt = nullptr;  // r_x_prn_tri, r_x_tau_n (3*nf)
  double* d_red = nullptr;       // scratch for deterministic reductions
  double* d_sbox = nullptr;      // vertex bounding box of the tagging (its own: the tagging may run beside the reductions)
  int* d_vbox = nullptr;         // vertex bounding box (6 ints) for ray tagging
  unsigned long long* d_slots = nullptr;  // REDUCE_SLOTS partial maxima (zero between uses)
  double* d_coord = nullptr;     // xc, xm, yc, ym, zc, zm tables (coord_off)

  // forcing accumulators: 128-bit fixed point (hi:int64, lo:uint64) per
  // component over the MLS window (E1*E2*E3 cells)
  int fwin[3] = {0, 0, 0};
  size_t fwin_cells = 0;
  unsigned long long* d_acc = nullptr;  // 4 comps * 2 words * cells
  double* d_fmean = nullptr;            // 3 means of the velocity forces (deferred shift)
  // narrow-band tagging window
  int twin[3] = {0, 0, 0};
  size_t twin_cells = 0;
  unsigned long long* d_tkey = nullptr;
  int* d_tface = nullptr;
  double* d_tdist = nullptr;   // the 125 band distances of every face (first tagging pass)
  double* d_pcoef = nullptr;   // gradient probes' shape-function derivatives: 27 x 3 x 6 per face
  melt_window* d_win = nullptr;
  // ray tagging scratch (init only)
  int* d_pcount = nullptr;
  int* d_plist = nullptr;
  size_t plist_cap = 0;

  // HIT forcing (hit.f90): bhat and Fhat (3 x nm^3 complex), exp(i k x) of the
  // staggered coordinates, the forced modes, the generator's step counter
  int hit_nm = 0, hit_nact = 0;
  double hit_tl = 0, hit_eps = 0, hit_kf = 0;
  unsigned long long hit_seed = 0;
  double *d_hit_bhat = nullptr, *d_hit_fhat = nullptr, *d_hit_tab = nullptr;
  double* d_hit_ax = nullptr;   // the x sums of the synthesis, 3 nm^2 n1 complex
  int* d_hit_act = nullptr;
  unsigned long long* d_hit_count = nullptr;

  int* d_err = nullptr;    // device error word (first error wins)
  int* h_err = nullptr;    // pinned mirror
  int defer_mean = 0;      // fused path: mean force applied in the projection
  int force_window_from_tag = 0;  // fused substep: tagging also places the forcing window
  int tag_ahead = 0;              // the next substep's tagging is already queued (side stream)
  int coef_ahead = 0;             // and the geometry part of its gradient probes
  int graph_uses_ahead = 0;       // the captured step's first substep reads coefficients queued before it

  // timing
  int timing = 0;
  uint32_t timing_mask = 0;   // slots timed while timing is on (bit s = slot s)
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
  size_t ev_next = 0;
  double t_ms[SLOT_COUNT] = {};
  long long t_n[SLOT_COUNT] = {};

  std::string err;
};

int melt_fail(melt_ctx* ctx, int code, const char* what, const char* detail);

// Typed view of an Eulerian field (R = double, or float in the precision form).
template <typename R>
inline R* fld(melt_ctx* c, int id) { return static_cast<R*>(c->f[id]); }

// ---------------------------------------------------------- device helpers
struct Grid {
  int n1, n2, n3;
  double xlen, ylen, zlen;
  double dx1, dx2, dx3;
};

__host__ __device__ inline Grid make_grid(const melt_ctx* c) {
  Grid g;
  g.n1 = c->n1; g.n2 = c->n2; g.n3 = c->n3;
  g.xlen = c->cfg.xlen; g.ylen = c->cfg.ylen; g.zlen = c->cfg.zlen;
  g.dx1 = c->dx1; g.dx2 = c->dx2; g.dx3 = c->dx3;
  return g;
}

// Cell-face and cell-centre coordinates of cordin.f90:55-99 (1-based index).
__host__ __device__ inline double gxc(const Grid& g, int i) { double x = (double)(i - 1) / (double)g.n1; return g.xlen * x; }
__host__ __device__ inline double gyc(const Grid& g, int j) { double x = (double)(j - 1) / (double)g.n2; return g.ylen * x; }
__host__ __device__ inline double gzc(const Grid& g, int k) { double x = (double)(k - 1) / (double)g.n3; return g.zlen * x; }
__host__ __device__ inline double gxm(const Grid& g, int i) { return (gxc(g, i) + gxc(g, i + 1)) * 0.5; }
__host__ __device__ inline double gym(const Grid& g, int j) { return (gyc(g, j) + gyc(g, j + 1)) * 0.5; }
__host__ __device__ inline double gzm(const Grid& g, int k) { return (gzc(g, k) + gzc(g, k + 1)) * 0.5; }

// The same coordinates as tables (cordin.f90 keeps them as arrays too), for
// kernels that would otherwise divide per cell: xc[0..n1], xm[0..n1-1],
// yc[0..n2], ym, zc[0..n3], zm, 0-based (xc[i] = gxc(g, i + 1)), laid out one
// after the other from coord_off(g, axis, centre).
__host__ __device__ inline int coord_off(const Grid& g, int axis, int centre) {
  int o = 0;
  if (axis > 0) o += 2 * g.n1 + 1;
  if (axis > 1) o += 2 * g.n2 + 1;
  const int n = axis == 0 ? g.n1 : axis == 1 ? g.n2 : g.n3;
  return o + (centre ? n + 1 : 0);
}
__host__ __device__ inline int coord_count(const Grid& g) { return 2 * (g.n1 + g.n2 + g.n3) + 3; }

// Non-negative remainder (Fortran MODULO).
__host__ __device__ inline int pmod(int a, int n) { int r = a % n; return r < 0 ? r + n : r; }
// Fortran NINT (round half away from zero).
__device__ inline int fnint(double x) { return (int)round(x); }

__device__ inline size_t cidx(const Grid& g, int i0, int j0, int k0) {  // 0-based, in range
  return (size_t)i0 + (size_t)g.n1 * ((size_t)j0 + (size_t)g.n2 * (size_t)k0);
}
// 1-based Fortran indices with periodic wrap in all three directions.
__device__ inline size_t widx(const Grid& g, int i, int j, int k) {
  return cidx(g, pmod(i - 1, g.n1), pmod(j - 1, g.n2), pmod(k - 1, g.n3));
}

// gfortran NORM2 (scaled sum of squares).
__device__ inline double fnorm2(const double* v, int n) {
  double scale = 0.0, res = 1.0;
  for (int i = 0; i < n; ++i) {
    if (v[i] != 0.0) {
      double ax = fabs(v[i]);
      if (scale < ax) { double r = scale / ax; res = 1.0 + res * r * r; scale = ax; }
      else { double r = ax / scale; res += r * r; }
    }
  }
  return scale * sqrt(res);
}

__device__ inline void dcross(double r[3], const double x[3], const double y[3]) {
  r[0] = x[1] * y[2] - y[1] * x[2];
  r[1] = x[2] * y[0] - y[2] * x[0];
  r[2] = x[0] * y[1] - y[0] * x[1];
}

// Double-precision sine for |x| <= 8: Cody-Waite reduction by pi/2 with a
// three-part constant, Taylor polynomials on [-pi/4, pi/4] (explicit fma).
// Used in the SDF -> VOF map so that cell tags are reproducible bit for bit.
__device__ inline double det_sin(double x) {
  const double two_over_pi = 0.63661977236758134308;
  const double p1 = 1.5707963267341256;
  const double p2 = 6.077100506303966e-11;
  const double p3 = 2.0222662487959506e-21;
  double kd = rint(x * two_over_pi);
  double y = ((x - kd * p1) - kd * p2) - kd * p3;
  double z = y * y;
  double s = 1.0 / 355687428096000.0;
  s = fma(s, z, -1.0 / 1307674368000.0);
  s = fma(s, z, 1.0 / 6227020800.0);
  s = fma(s, z, -1.0 / 39916800.0);
  s = fma(s, z, 1.0 / 362880.0);
  s = fma(s, z, -1.0 / 5040.0);
  s = fma(s, z, 1.0 / 120.0);
  s = fma(s, z, -1.0 / 6.0);
  double sy = fma(y * z, s, y);
  double c = -1.0 / 6402373705728000.0;
  c = fma(c, z, 1.0 / 20922789888000.0);
  c = fma(c, z, -1.0 / 87178291200.0);
  c = fma(c, z, 1.0 / 479001600.0);
  c = fma(c, z, -1.0 / 3628800.0);
  c = fma(c, z, 1.0 / 40320.0);
  c = fma(c, z, -1.0 / 720.0);
  c = fma(c, z, 1.0 / 24.0);
  c = fma(c, z, -0.5);
  double cy = fma(z, c, 1.0);
  int q = ((int)kd) & 3;
  return q == 0 ? sy : q == 1 ? cy : q == 2 ? -sy : -cy;
}

// Own exp for the MLS Gaussian weights (argument in [-12, 0]): reduction by
// ln2 with a two-part constant, Taylor polynomial of degree 13 (explicit fma),
// exact scaling by 2^k; the same bits on every device and on the host side.
__device__ inline double det_exp(double x) {
  const double log2e = 1.4426950408889634;
  const double ln2_hi = 0.6931471803691238;      // ln2 to 32 bits
  const double ln2_lo = 1.9082149292705877e-10;
  double kd = rint(x * log2e);
  double r = (x - kd * ln2_hi) - kd * ln2_lo;
  double p = 1.0 / 6227020800.0;
  p = fma(p, r, 1.0 / 479001600.0);
  p = fma(p, r, 1.0 / 39916800.0);
  p = fma(p, r, 1.0 / 3628800.0);
  p = fma(p, r, 1.0 / 362880.0);
  p = fma(p, r, 1.0 / 40320.0);
  p = fma(p, r, 1.0 / 5040.0);
  p = fma(p, r, 1.0 / 720.0);
  p = fma(p, r, 1.0 / 120.0);
  p = fma(p, r, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)kd);
}

// In-register LU with partial pivoting followed by the inverse from the LU
// factors (the dgetrf/dgetri sequence the reference calls, mlsWeight.f90:120-126),
// column-major n x n.  Returns 0 or the 1-based index of the first zero pivot.
template <int N>
__device__ inline int lu_inverse(double a[N * N]) {
#define A_(i, j) a[(i) + N * (j)]
  int ipiv[N];
  int info = 0;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    int jp = j;
    double amax = fabs(A_(j, j));
#pragma unroll
    for (int i = j + 1; i < N; ++i) {
      double v = fabs(A_(i, j));
      if (v > amax) { amax = v; jp = i; }
    }
    ipiv[j] = jp;
    // row swap through predicated selects keeps the array in registers
    double pivot = A_(j, j);
#pragma unroll
    for (int i = j + 1; i < N; ++i) pivot = (jp == i) ? A_(i, j) : pivot;
    if (pivot != 0.0) {
#pragma unroll
      for (int c = 0; c < N; ++c) {
        double rj = A_(j, c), rp = rj;
#pragma unroll
        for (int i = j + 1; i < N; ++i) rp = (jp == i) ? A_(i, c) : rp;
#pragma unroll
        for (int i = j + 1; i < N; ++i) A_(i, c) = (jp == i) ? rj : A_(i, c);
        A_(j, c) = rp;
      }
      if (j < N - 1) {
        if (fabs(A_(j, j)) >= 2.2250738585072014e-308) {
          double r = 1.0 / A_(j, j);
#pragma unroll
          for (int i = j + 1; i < N; ++i) A_(i, j) = A_(i, j) * r;
        } else {
#pragma unroll
          for (int i = j + 1; i < N; ++i) A_(i, j) = A_(i, j) / A_(j, j);
        }
      }
    } else if (info == 0) {
      info = j + 1;
    }
    if (j < N - 1) {
#pragma unroll
      for (int jj = j + 1; jj < N; ++jj) {
        if (A_(j, jj) != 0.0) {
          double temp = -A_(j, jj);
#pragma unroll
          for (int ii = j + 1; ii < N; ++ii) A_(ii, jj) = A_(ii, jj) + A_(ii, j) * temp;
        }
      }
    }
  }
  if (info) return info;
#pragma unroll
  for (int j = 0; j < N; ++j) {  // inverse of U (upper, non-unit)
    A_(j, j) = 1.0 / A_(j, j);
    double ajj = -A_(j, j);
#pragma unroll
    for (int jj = 0; jj < j; ++jj) {
      if (A_(jj, j) != 0.0) {
        double temp = A_(jj, j);
#pragma unroll
        for (int i = 0; i < jj; ++i) A_(i, j) = A_(i, j) + temp * A_(i, jj);
        A_(jj, j) = A_(jj, j) * A_(jj, jj);
      }
    }
#pragma unroll
    for (int i = 0; i < j; ++i) A_(i, j) = ajj * A_(i, j);
  }
  double work[N];
#pragma unroll
  for (int j = N - 1; j >= 0; --j) {  // solve inv(A) L = inv(U)
#pragma unroll
    for (int i = j + 1; i < N; ++i) { work[i] = A_(i, j); A_(i, j) = 0.0; }
#pragma unroll
    for (int jj = j + 1; jj < N; ++jj) {
      if (work[jj] != 0.0) {
        double temp = -1.0 * work[jj];
#pragma unroll
        for (int i = 0; i < N; ++i) A_(i, j) = A_(i, j) + temp * A_(i, jj);
      }
    }
  }
#pragma unroll
  for (int j = N - 2; j >= 0; --j) {  // column interchanges
    int jp = ipiv[j];
#pragma unroll
    for (int c = j + 1; c < N; ++c) {
      if (jp == c) {
#pragma unroll
        for (int i = 0; i < N; ++i) { double t = A_(i, j); A_(i, j) = A_(i, c); A_(i, c) = t; }
      }
    }
  }
  return 0;
#undef A_
}

// Device error word: first error code wins.  A blown-up body sends every
// thread of a tagging or forcing pass here; reading the word first keeps them
// from queueing compare-and-swaps on one address once it is set (3M of them
// cost 35 ms per pass).
__device__ inline void raise_error(int* err, int code) {
  if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) atomicCAS(err, 0, code);
}

// Max reductions of non-negative doubles: their bit patterns order like the
// values, so partial maxima combine with integer atomicMax.  Each block folds
// its maximum into slot blockIdx % REDUCE_SLOTS (few atomics per address);
// euler_fold_slots turns the slots into one value and clears them.
constexpr int REDUCE_SLOTS = 1024;
__device__ inline void block_max_to_slots(double v, unsigned long long* slots) {
  __shared__ double wm[16];
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if (lane == 0) wm[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wm[0];
    for (int q = 1; q < nw; ++q) m = fmax(m, wm[q]);
    atomicMax(&slots[blockIdx.x % REDUCE_SLOTS], (unsigned long long)__double_as_longlong(m));
  }
}

// XCD-aware block remap: consecutive logical blocks land on the same XCD (the
// hardware deals workgroups round-robin over the 8 XCDs), so neighbouring tiles
// share an L2.
__device__ inline unsigned xcd_block(unsigned b, unsigned nb) {
  const unsigned nx = 8;
  unsigned per = nb / nx, rem = nb % nx;
  unsigned x = b % nx, q = b / nx;
  // blocks with x < rem own per+1 slots
  unsigned base = x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per;
  return base + q;
}

// Tiled order of (a, b) line indices (na x nb lines): consecutive logical
// blocks walk a ta x tb tile before moving on, so the lines a block's
// stencils share with its neighbours in both directions are loaded into the
// same L2 at about the same time.  ta | na and tb | nb (else 1 is used).
__device__ __forceinline__ void tiled_2d(unsigned t, int na, int nb, int ta, int tb, int& a, int& b) {
  if (na % ta) ta = 1;
  if (nb % tb) tb = 1;
  const unsigned per = (unsigned)(ta * tb);
  unsigned tile = t / per, r = t % per;
  int nta = na / ta;
  a = (int)(tile % nta) * ta + (int)(r % ta);
  b = (int)(tile / nta) * tb + (int)(r / ta);
}

// host-side helpers implemented in the kernel files
int euler_launch_hdnl(melt_ctx* ctx, int which);
int euler_launch_invtr_rhs(melt_ctx* ctx, int which);
int euler_launch_solve(melt_ctx* ctx, int dir, void* const* fields, int nfield, const double* betadx, void* const* qadd, int store_rhs);
int euler_launch_explicit_xsolve(melt_ctx* ctx, double bxu, double bxt);
int euler_launch_divg(melt_ctx* ctx);
int euler_launch_green(melt_ctx* ctx, int dir, void* const* src, void* const* dst, void* const* qadd, const double* betadx, int nfield,
                       int* used);
int euler_launch_green_yz(melt_ctx* ctx, void* const* src, void* const* qadd, const double* by, const double* bz, int nfield, int* used);
int euler_launch_project(melt_ctx* ctx, int do_u, int do_p, int use_mean, const int* gate = nullptr);
int euler_launch_reduce_max(melt_ctx* ctx, int what, unsigned long long* d, double dt = 0.0);
int euler_launch_refine_gate(melt_ctx* ctx, const unsigned long long* dmax, int* gate);
int euler_launch_interp_vof(melt_ctx* ctx);
int euler_reduce_max(melt_ctx* ctx, int what, double* out, double dt = 0.0);
int diag_dissipation(melt_ctx* ctx, double* sums);
int hit_free(melt_ctx* ctx);   // fluid-cell sums of tke, diss, chi and the cell count
int euler_launch_inqpr(melt_ctx* ctx);
int poisson_init(melt_ctx* ctx);
// div_slots != null: the forward x pass also folds max |div u| (FP64, as
// divgck) into those REDUCE_SLOTS slots.
int poisson_solve(melt_ctx* ctx, int fused_divg, const int* gate = nullptr, unsigned long long* div_slots = nullptr);
int euler_fold_slots(melt_ctx* ctx, unsigned long long* slots, unsigned long long* out);
int ibm_tag(melt_ctx* ctx);
int ibm_find_centroid_indices(melt_ctx* ctx);
int ibm_weights(melt_ctx* ctx, bool find = false);   // find: form the cell indices too
int ibm_force_scatter(melt_ctx* ctx);
int ibm_force_drain(melt_ctx* ctx, int apply_now);
int ibm_sub_mean(melt_ctx* ctx);
int ibm_velforce(melt_ctx* ctx, int defer);
int ibm_tempforce(melt_ctx* ctx);
// pind_out: where the heat-flux probes' cell indices go (null: d_pind, as
// mls_heatFlux does); lagr_after_force(ctx, true) copies d_pind2 into d_pind.
int ibm_probes(melt_ctx* ctx, int* pind_out = nullptr, bool coefs_ready = false);
// geometry part of the gradient probes into d_pcoef (then ibm_probes(..., true))
int ibm_probe_coefs(melt_ctx* ctx);
int ibm_window_reset(melt_ctx* ctx);
// with commit_pind (the fused substep) it also commits the staged heat-flux probe indices.
// With IREMESH = 1 the host reads the device's short-edge flag after the
// Newton-Euler update (waiting for the queue up to there) and remeshes on the
// host before the geometry refresh.
int lagr_after_force(melt_ctx* ctx, bool commit_pind = false);
bool capturing(melt_ctx* ctx);
int run_after_forcing_read(melt_ctx* ctx, const std::function<int()>& work);
// host remeshing of the context's body (remesh.hip): *did = 1 when it ran
int remesh_context(melt_ctx* ctx, int* did);
// host mirrors from the device connectivity (after a snapshot load)
int remesh_pull_mirrors(melt_ctx* ctx);
int lagr_alloc(melt_ctx* ctx, const double* xyz0, const int* voe, const int* eof, const int* vof, double rad_p);
void timing_begin(melt_ctx* ctx, int slot, hipEvent_t* ev);
void timing_end(melt_ctx* ctx, int slot, hipEvent_t ev);
