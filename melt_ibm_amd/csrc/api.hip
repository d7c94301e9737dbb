// C ABI of libmeltgpu (include/melt_ibm.h): context lifecycle, host <-> device
// transfers, the drop-in entry points with the reference's subroutine semantics,
// and the fused RK3 substep.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>

#include "common.h"

int melt_fail(melt_ctx* ctx, int code, const char* what, const char* detail) {
  if (ctx) ctx->err = std::string(what) + ": " + (detail ? detail : "");
  return code;
}

namespace {

const char* slot_names[SLOT_COUNT] = {
    "tag_narrowband", "explicit_xsolve", "solve_x", "solve_y", "solve_z_update", "mls_force", "lagrangian",
    "divg_fft_x", "fft_y", "fft_z_solve", "ifft_y", "ifft_x", "project",
    "phase_tagging", "phase_eulerian", "phase_mls", "phase_pressure", "substep", "refine"};

const char* err_text(int code) {
  switch (code) {
    case MELT_ERR_SINGULAR_MLS: return "Error in dgetrf (singular MLS moment matrix)";
    case MELT_ERR_SMALL_VOLUME: return "Geometry V/VE is smaller than threshold";
    case MELT_ERR_PENCIL_OVERFLOW: return "pencil triangle count has exceeded buffer size";
    case MELT_ERR_RAY_RECURSION: return "parallel-ray recursion did not terminate";
    case MELT_ERR_WINDOW: return "immersed body outgrew the device window";
    case MELT_ERR_ARG: return "invalid argument in device code";
    default: return "device error";
  }
}

bool size_ok(int n) {  // 2^a * 3^b, b <= 1 (FFT and circulant PCR closed forms)
  if (n < 8) return false;
  int m = n;
  while (m % 2 == 0) m /= 2;
  return m == 1 || m == 3;
}

__global__ void k_reset_loads(melt_body* b) {  // tsch.f90:19-23
  for (int d = 0; d < 3; ++d) { b->fp[d] = 0.0; b->ftau[d] = 0.0; b->torq_p[d] = 0.0; b->torq_tau[d] = 0.0; }
}

template <typename R>
__global__ void k_fill_d(R* __restrict__ p, double v, size_t n) {
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) p[c] = (R)v;
}

// Fill a field with a value (fill kernel, not a memset: safe under capture).
void fill_field(melt_ctx* ctx, int id, double v) {
  if (ctx->fp32) hipLaunchKernelGGL(k_fill_d<float>, dim3(1024), dim3(256), 0, ctx->stream, fld<float>(ctx, id), v, ctx->nc);
  else hipLaunchKernelGGL(k_fill_d<double>, dim3(1024), dim3(256), 0, ctx->stream, fld<double>(ctx, id), v, ctx->nc);
}

int check_ctx(melt_ctx* ctx) { return ctx ? 0 : MELT_ERR_ARG; }

// Drop a captured step graph (its kernels carry dt and the stage coefficients
// as arguments).
void graph_drop(melt_ctx* ctx) {
  if (ctx->graph_exec) hipGraphExecDestroy(ctx->graph_exec);
  if (ctx->graph) hipGraphDestroy(ctx->graph);
  ctx->graph_exec = nullptr;
  ctx->graph = nullptr;
  ctx->graph_uses_ahead = 0;
}

// tschem queues the next call's first tagging and probe geometry stage at the
// end of its last substep.  A call that changes what those read (the faces,
// the body) or what the tagging writes (VOFp, sdf) cancels them, and drops a
// captured step whose first substep relies on them (also when the queue was
// already consumed outside the graph); the next substep then does both itself.
void cancel_ahead(melt_ctx* ctx) {
  const bool queued = ctx->tag_ahead || ctx->coef_ahead;
  ctx->tag_ahead = 0;
  ctx->coef_ahead = 0;
  if (queued || ctx->graph_uses_ahead) graph_drop(ctx);
}
bool tag_field(int field) {
  return field == MELT_VOFP || field == MELT_VOFX || field == MELT_VOFY || field == MELT_VOFZ || field == MELT_SDF;
}

}  // namespace

// --------------------------------------------------------------- timing
void timing_begin(melt_ctx* ctx, int slot, hipEvent_t* ev) {
  *ev = nullptr;
  if (!ctx->timing || !((ctx->timing_mask >> slot) & 1u)) return;
  if (ctx->ev_next >= ctx->ev_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    ctx->ev_pool.push_back(e);
  }
  *ev = ctx->ev_pool[ctx->ev_next++];
  hipEventRecord(*ev, ctx->stream);
}

void timing_end(melt_ctx* ctx, int slot, hipEvent_t ev) {
  if (!ctx->timing || !ev) return;
  if (ctx->ev_next >= ctx->ev_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    ctx->ev_pool.push_back(e);
  }
  hipEvent_t e2 = ctx->ev_pool[ctx->ev_next++];
  hipEventRecord(e2, ctx->stream);
  ctx->ev_pending.push_back({slot, {ev, e2}});
}

static void timing_flush(melt_ctx* ctx) {
  if (ctx->ev_pending.empty()) return;
  hipStreamSynchronize(ctx->stream);
  for (auto& p : ctx->ev_pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess) {
      ctx->t_ms[p.first] += ms;
      ctx->t_n[p.first] += 1;
    }
  }
  ctx->ev_pending.clear();
  ctx->ev_next = 0;
}

// ------------------------------------------------------------ lifecycle
extern "C" int melt_device_ok(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n > 0 ? 1 : 0;
}

extern "C" void melt_ctx_destroy(melt_ctx* ctx) {
  if (!ctx) return;
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  if (ctx->side) hipStreamSynchronize(ctx->side);   // work left there after an error
  if (ctx->side2) hipStreamSynchronize(ctx->side2);
  graph_drop(ctx);
  for (int i = 0; i < MELT_NFIELDS; ++i) if (ctx->f[i]) hipFree(ctx->f[i]);
  for (int i = 1; i < 4; ++i) if (ctx->rhs4[i]) hipFree(ctx->rhs4[i]);
  for (int i = 0; i < 4; ++i) if (ctx->tmp4[i]) hipFree(ctx->tmp4[i]);
  void* ps[] = {ctx->spec, ctx->ak1, ctx->ak2, ctx->ak3, ctx->tw1, ctx->tw2, ctx->tw3, ctx->d_body, ctx->d_vof, ctx->d_eof, ctx->d_voe,
                ctx->d_v2f_ptr, ctx->d_v2f, ctx->d_gv, ctx->d_ge, ctx->d_gf, ctx->d_pind, ctx->d_pind2, ctx->d_probe, ctx->d_rxp, ctx->d_rxt, ctx->d_red, ctx->d_sbox,
                ctx->d_vbox, ctx->d_slots, ctx->d_coord, ctx->d_acc, ctx->d_fmean, ctx->d_tkey, ctx->d_tface, ctx->d_tdist, ctx->d_pcoef, ctx->d_win, ctx->d_pcount, ctx->d_plist, ctx->d_err, ctx->d_gate, ctx->d_refine_stats,
                ctx->d_foe, ctx->d_xyz_melt};
  for (void* p : ps) if (p) hipFree(p);
  for (int i = 0; i < MELT_L_NLAG; ++i) if (ctx->d_lag[i]) hipFree(ctx->d_lag[i]);
  hit_free(ctx);
  if (ctx->h_err) hipHostFree(ctx->h_err);
  if (ctx->h_rm) hipHostFree(ctx->h_rm);
  if (ctx->ev_rm) hipEventDestroy(ctx->ev_rm);
  for (auto e : ctx->ev_pool) hipEventDestroy(e);
  if (ctx->side) hipStreamSynchronize(ctx->side);
  if (ctx->side2) hipStreamSynchronize(ctx->side2);
  for (hipEvent_t e : {ctx->ev_fork, ctx->ev_join, ctx->ev_fork2, ctx->ev_join2, ctx->ev_forc, ctx->ev_hit}) if (e) hipEventDestroy(e);
  if (ctx->side) hipStreamDestroy(ctx->side);
  if (ctx->side2) hipStreamDestroy(ctx->side2);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" int melt_ctx_create(const melt_config* cfg, melt_ctx** out) {
  if (!out) return MELT_ERR_ARG;
  *out = nullptr;
  if (!cfg || cfg->n1m != cfg->n2m || !size_ok(cfg->n1m) || !size_ok(cfg->n2m) || !size_ok(cfg->n3m) || cfg->n1m > 1024 || cfg->n3m > 1024)
    return MELT_ERR_ARG;
  if (cfg->precision != 0 && cfg->precision != 1) return MELT_ERR_ARG;
  melt_ctx* ctx = new melt_ctx();
  ctx->cfg = *cfg;
  ctx->fp32 = cfg->precision == 1;
  ctx->esz = ctx->fp32 ? sizeof(float) : sizeof(double);
  ctx->spec64 = !ctx->fp32;
  ctx->n1 = cfg->n1m; ctx->n2 = cfg->n2m; ctx->n3 = cfg->n3m;
  ctx->nc = (size_t)ctx->n1 * ctx->n2 * ctx->n3;
  ctx->n1h = ctx->n1 / 2 + 1;
  ctx->n1hp = (ctx->n1h + 7) & ~7;
  ctx->pi = 2.0 * asin(1.0);
  ctx->dx1 = (double)ctx->n1 / cfg->xlen; ctx->dx2 = (double)ctx->n2 / cfg->ylen; ctx->dx3 = (double)ctx->n3 / cfg->zlen;
  ctx->dx1q = ctx->dx1 * ctx->dx1; ctx->dx2q = ctx->dx2 * ctx->dx2; ctx->dx3q = ctx->dx3 * ctx->dx3;
  ctx->gam[0] = 8.0 / 15.0; ctx->gam[1] = 5.0 / 12.0; ctx->gam[2] = 3.0 / 4.0;
  ctx->rom[0] = 0.0; ctx->rom[1] = -17.0 / 60.0; ctx->rom[2] = -5.0 / 12.0;
  for (int i = 0; i < 3; ++i) ctx->alm[i] = ctx->gam[i] + ctx->rom[i];
  int rc = 0;
  do {
    if (hipGetDevice(&ctx->device) != hipSuccess) { rc = melt_fail(ctx, MELT_ERR_HIP, "ctx_create", "no HIP device"); break; }
    {
      int lds = 0;
      if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) == hipSuccess && lds > 0)
        ctx->lds_max = (size_t)lds;
    }
    if (const char* tv = getenv("MELT_TUNE")) {
      static const char* keys[TUNE_N] = {"xf_w", "xf_nt", "solve_w", "solve_nt", "xf_z", "yz_jb", "streams", "coef", "pbuf", "xf_one"};
      std::string spec(tv);
      size_t pos = 0;
      while (pos < spec.size()) {
        size_t end = spec.find(',', pos);
        if (end == std::string::npos) end = spec.size();
        std::string item = spec.substr(pos, end - pos);
        size_t eq = item.find('=');
        if (eq != std::string::npos) {
          for (int t = 0; t < TUNE_N; ++t)
            if (item.substr(0, eq) == keys[t]) ctx->tune[t] = atoi(item.c_str() + eq + 1);
          if (item.substr(0, eq) == "green") ctx->no_green = atoi(item.c_str() + eq + 1) == 0;
        }
        pos = end + 1;
      }
    }
    // the side stream is created on first use (one_stream() contexts never
    // create it): HIP places streams on its hardware queues round-robin in
    // creation order, so many contexts on one GPU spread their main streams
    // over all queues only if no side streams are interleaved among them
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_fork2, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_join2, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_forc, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_hit, hipEventDisableTiming) != hipSuccess) {
      rc = MELT_ERR_HIP;
      break;
    }
    for (int i = 0; i < MELT_NFIELDS && !rc; ++i) {
      if (hipMalloc(&ctx->f[i], ctx->esz * ctx->nc) != hipSuccess || hipMemset(ctx->f[i], 0, ctx->esz * ctx->nc) != hipSuccess)
        rc = melt_fail(ctx, MELT_ERR_HIP, "ctx_create", "field allocation failed");
    }
    if (rc) break;
    ctx->rhs4[0] = ctx->f[MELT_RHS];
    for (int i = 1; i < 4 && !rc; ++i)
      if (hipMalloc(&ctx->rhs4[i], ctx->esz * ctx->nc) != hipSuccess) rc = melt_fail(ctx, MELT_ERR_HIP, "ctx_create", "rhs allocation");
    for (int i = 0; i < 4 && !rc; ++i)
      if (hipMalloc(&ctx->tmp4[i], ctx->esz * ctx->nc) != hipSuccess) rc = melt_fail(ctx, MELT_ERR_HIP, "ctx_create", "rhs allocation");
    if (rc) break;
    if (hipMalloc(&ctx->d_body, sizeof(melt_body)) != hipSuccess || hipMemset(ctx->d_body, 0, sizeof(melt_body)) != hipSuccess ||
        hipMalloc(&ctx->d_red, sizeof(double) * 16384) != hipSuccess || hipMalloc(&ctx->d_sbox, sizeof(double) * 8) != hipSuccess || hipMalloc(&ctx->d_fmean, sizeof(double) * 4) != hipSuccess ||
        hipMemset(ctx->d_fmean, 0, sizeof(double) * 4) != hipSuccess || hipMalloc(&ctx->d_err, sizeof(int)) != hipSuccess ||
        hipMemset(ctx->d_err, 0, sizeof(int)) != hipSuccess || hipHostMalloc(&ctx->h_err, sizeof(int)) != hipSuccess ||
        hipMalloc(&ctx->d_win, sizeof(melt_window) * 2) != hipSuccess || hipMemset(ctx->d_win, 0, sizeof(melt_window) * 2) != hipSuccess ||
        hipMalloc(&ctx->d_gate, sizeof(int)) != hipSuccess || hipMemset(ctx->d_gate, 0, sizeof(int)) != hipSuccess ||
        hipMalloc(&ctx->d_refine_stats, sizeof(long long) * 8) != hipSuccess ||
        hipMalloc(&ctx->d_slots, sizeof(unsigned long long) * REDUCE_SLOTS) != hipSuccess ||
        hipMemset(ctx->d_slots, 0, sizeof(unsigned long long) * REDUCE_SLOTS) != hipSuccess ||
        hipMemset(ctx->d_refine_stats, 0, sizeof(long long) * 8) != hipSuccess) {
      rc = melt_fail(ctx, MELT_ERR_HIP, "ctx_create", "scratch allocation failed");
      break;
    }
    {
      const Grid g = make_grid(ctx);
      std::vector<double> tab((size_t)coord_count(g));
      const int ns[3] = {g.n1, g.n2, g.n3};
      for (int a = 0; a < 3; ++a)
        for (int i = 1; i <= ns[a] + 1; ++i) {
          tab[coord_off(g, a, 0) + i - 1] = a == 0 ? gxc(g, i) : a == 1 ? gyc(g, i) : gzc(g, i);
          if (i <= ns[a]) tab[coord_off(g, a, 1) + i - 1] = a == 0 ? gxm(g, i) : a == 1 ? gym(g, i) : gzm(g, i);
        }
      if (hipMalloc(&ctx->d_coord, sizeof(double) * tab.size()) != hipSuccess ||
          hipMemcpy(ctx->d_coord, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice) != hipSuccess) {
        rc = melt_fail(ctx, MELT_ERR_HIP, "ctx_create", "coordinate table");
        break;
      }
    }
    rc = poisson_init(ctx);
    if (rc) break;
    // the body state the reference starts from (module zeros) plus the MLS constants
    melt_body& b = ctx->h_body;
    memset(&b, 0, sizeof b);
    b.h_eulerian = (1.0 / ctx->dx1 + 1.0 / ctx->dx2 + 1.0 / ctx->dx3) / 3.0;
    b.celvol = 1.0 / (ctx->dx1 * ctx->dx2 * ctx->dx3);
    b.e_thresh = cfg->perc_ethresh * b.h_eulerian;
    b.v_thresh = cfg->v_on_ve_perc * b.celvol;
    if (hipMemcpy(ctx->d_body, &b, sizeof b, hipMemcpyHostToDevice) != hipSuccess) rc = MELT_ERR_HIP;
    // single-phase defaults of init.f90:30-33 (after the zero fills above,
    // which are not ordered with this non-blocking stream)
    if (hipDeviceSynchronize() != hipSuccess) { rc = melt_fail(ctx, MELT_ERR_HIP, "ctx_create", "initial fill failed"); break; }
    for (int id : {MELT_VOFX, MELT_VOFY, MELT_VOFZ, MELT_VOFP}) fill_field(ctx, id, 1.0);
    if ((rc = ibm_window_reset(ctx))) break;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) rc = melt_fail(ctx, MELT_ERR_HIP, "ctx_create", "initial fill failed");
  } while (0);
  if (rc) {
    melt_ctx_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return MELT_OK;
}

extern "C" const char* melt_last_error(const melt_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int melt_sync(melt_ctx* ctx) {
  if (check_ctx(ctx)) return MELT_ERR_ARG;
  MELT_HIP(hipMemcpyAsync(ctx->h_err, ctx->d_err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  MELT_HIP(hipStreamSynchronize(ctx->stream));
  int code = *ctx->h_err;
  if (code) {
    *ctx->h_err = 0;
    MELT_HIP(hipMemset(ctx->d_err, 0, sizeof(int)));
    return melt_fail(ctx, code, "device", err_text(code));
  }
  return 0;
}

// ---------------------------------------------------------- data movement
// Host arrays are double; the precision form stores fields as float (the
// conversion happens here, on the host).
static int copy_to_device(melt_ctx* ctx, int field, const double* src) {
  if (!ctx->fp32) {
    MELT_HIP(hipMemcpy(ctx->f[field], src, sizeof(double) * ctx->nc, hipMemcpyHostToDevice));
    return 0;
  }
  std::vector<float> t(ctx->nc);
  for (size_t c = 0; c < ctx->nc; ++c) t[c] = (float)src[c];
  MELT_HIP(hipMemcpy(ctx->f[field], t.data(), sizeof(float) * ctx->nc, hipMemcpyHostToDevice));
  return 0;
}

static int copy_from_device(melt_ctx* ctx, int field, double* dst) {
  if (!ctx->fp32) {
    MELT_HIP(hipMemcpy(dst, ctx->f[field], sizeof(double) * ctx->nc, hipMemcpyDeviceToHost));
    return 0;
  }
  std::vector<float> t(ctx->nc);
  MELT_HIP(hipMemcpy(t.data(), ctx->f[field], sizeof(float) * ctx->nc, hipMemcpyDeviceToHost));
  for (size_t c = 0; c < ctx->nc; ++c) dst[c] = (double)t[c];
  return 0;
}

extern "C" int melt_upload_field(melt_ctx* ctx, int field, const double* host, const melt_layout* lay) {
  if (ctx) ctx->forc_free = 0;   // it may touch the forcing on the main stream
  if (check_ctx(ctx) || field < 0 || field >= MELT_NFIELDS || !host) return MELT_ERR_ARG;
  if (tag_field(field)) cancel_ahead(ctx);
  melt_layout L = lay ? *lay : melt_layout{ctx->n1, ctx->n2, 0, 0};
  if (L.ld1 < ctx->n1 || L.ld2 < ctx->n2) return melt_fail(ctx, MELT_ERR_ARG, "upload", "leading dimension too small");
  MELT_HIP(hipStreamSynchronize(ctx->stream));
  if (L.ld1 == ctx->n1 && L.ld2 == ctx->n2 && L.kghost == 0) return copy_to_device(ctx, field, host);
  std::vector<double> tmp(ctx->nc);
  for (int k = 0; k < ctx->n3; ++k)
    for (int j = 0; j < ctx->n2; ++j)
      memcpy(&tmp[(size_t)ctx->n1 * (j + (size_t)ctx->n2 * k)], &host[(size_t)L.ld1 * (j + (size_t)L.ld2 * (k + L.kghost))],
             sizeof(double) * ctx->n1);
  return copy_to_device(ctx, field, tmp.data());
}

extern "C" int melt_download_field(melt_ctx* ctx, int field, double* host, const melt_layout* lay) {
  if (ctx) ctx->forc_free = 0;   // it may touch the forcing on the main stream
  if (check_ctx(ctx) || field < 0 || field >= MELT_NFIELDS || !host) return MELT_ERR_ARG;
  melt_layout L = lay ? *lay : melt_layout{ctx->n1, ctx->n2, 0, 0};
  if (L.ld1 < ctx->n1 || L.ld2 < ctx->n2) return melt_fail(ctx, MELT_ERR_ARG, "download", "leading dimension too small");
  MELT_HIP(hipStreamSynchronize(ctx->stream));
  if (L.ld1 == ctx->n1 && L.ld2 == ctx->n2 && L.kghost == 0) return copy_from_device(ctx, field, host);
  std::vector<double> tmp(ctx->nc);
  int rc = copy_from_device(ctx, field, tmp.data());
  if (rc) return rc;
  int k0 = L.kghost ? -1 : 0, k1 = L.kghost ? ctx->n3 + 1 : ctx->n3;
  for (int k = k0; k < k1; ++k) {
    int ks = pmod(k, ctx->n3);
    for (int j = 0; j < ctx->n2; ++j)
      memcpy(&host[(size_t)L.ld1 * (j + (size_t)L.ld2 * (k + L.kghost))], &tmp[(size_t)ctx->n1 * (j + (size_t)ctx->n2 * ks)],
             sizeof(double) * ctx->n1);
  }
  return 0;
}

template <typename R>
__global__ void k_from_double(R* __restrict__ dst, const double* __restrict__ src, size_t n) {
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) dst[c] = (R)src[c];
}

// Device-to-device upload: src is device memory holding N1*N2*N3 doubles
// (i fastest) that is complete before the call; the precision form stores
// them rounded to float.  Returns after the copy has finished.
extern "C" int melt_upload_field_device(melt_ctx* ctx, int field, const double* src) {
  if (ctx) ctx->forc_free = 0;   // it may touch the forcing on the main stream
  if (check_ctx(ctx) || field < 0 || field >= MELT_NFIELDS || !src) return MELT_ERR_ARG;
  if (tag_field(field)) cancel_ahead(ctx);
  if (ctx->fp32) hipLaunchKernelGGL(k_from_double<float>, dim3(2048), dim3(256), 0, ctx->stream, fld<float>(ctx, field), src, ctx->nc);
  else hipLaunchKernelGGL(k_from_double<double>, dim3(2048), dim3(256), 0, ctx->stream, fld<double>(ctx, field), src, ctx->nc);
  MELT_HIP(hipGetLastError());
  MELT_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

// Non-blocking completion test of the work queued on the context (its side
// stream joins the main stream before every substep ends).
extern "C" int melt_poll(melt_ctx* ctx, int32_t* done) {
  if (check_ctx(ctx) || !done) return MELT_ERR_ARG;
  hipError_t e = hipStreamQuery(ctx->stream);
  if (e == hipSuccess) { *done = 1; return 0; }
  if (e == hipErrorNotReady) { *done = 0; return 0; }
  return melt_fail(ctx, MELT_ERR_HIP, "hipStreamQuery", hipGetErrorString(e));
}

extern "C" int melt_fill_field(melt_ctx* ctx, int field, double value) {
  if (check_ctx(ctx) || field < 0 || field >= MELT_NFIELDS) return MELT_ERR_ARG;
  if (tag_field(field)) cancel_ahead(ctx);
  fill_field(ctx, field, value);
  MELT_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------- geometry / body
extern "C" int melt_set_geometry(melt_ctx* ctx, int32_t nv, int32_t ne, int32_t nf, const double* xyz0, const int32_t* vert_of_edge,
                                 const int32_t* edge_of_face, double rad_p) {
  if (check_ctx(ctx) || nv <= 0 || ne <= 0 || nf <= 0 || !xyz0 || !vert_of_edge || !edge_of_face) return MELT_ERR_ARG;
  if (ctx->has_geom) return melt_fail(ctx, MELT_ERR_ARG, "set_geometry", "geometry already set for this context");
  ctx->nv = nv; ctx->ne = ne; ctx->nf = nf;
  ctx->rad_p = rad_p;
  MELT_HIP(hipStreamSynchronize(ctx->stream));
  int rc = lagr_alloc(ctx, xyz0, vert_of_edge, edge_of_face, nullptr, rad_p);
  if (rc) return rc;
  fill_field(ctx, MELT_SDF, 999.0);
  MELT_HIP(hipGetLastError());
  ctx->has_geom = 1;
  MELT_HIP(hipMemcpyAsync(&ctx->h_body, ctx->d_body, sizeof(melt_body), hipMemcpyDeviceToHost, ctx->stream));
  MELT_HIP(hipStreamSynchronize(ctx->stream));
  return melt_sync(ctx);
}

extern "C" int melt_get_body(melt_ctx* ctx, melt_body* out) {
  if (check_ctx(ctx) || !out) return MELT_ERR_ARG;
  MELT_HIP(hipMemcpyAsync(&ctx->h_body, ctx->d_body, sizeof(melt_body), hipMemcpyDeviceToHost, ctx->stream));
  MELT_HIP(hipStreamSynchronize(ctx->stream));
  *out = ctx->h_body;
  return 0;
}

extern "C" int melt_set_body(melt_ctx* ctx, const melt_body* in) {
  if (check_ctx(ctx) || !in) return MELT_ERR_ARG;
  cancel_ahead(ctx);
  ctx->h_body = *in;
  MELT_HIP(hipMemcpyAsync(ctx->d_body, &ctx->h_body, sizeof(melt_body), hipMemcpyHostToDevice, ctx->stream));
  MELT_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

static size_t lag_count(melt_ctx* ctx, int id) {
  static const int len[MELT_L_NLAG] = {3, 3, 1, 1, 1, 3, 3, 3, 3, 3, 1, 3, 1, 1, 1, 1, 27, 27, 27, 27, 3, 3, 3, 1};
  size_t n = id <= MELT_L_DXYZV_S ? ctx->nv : id <= MELT_L_TAU_N ? ctx->nf : ctx->ne;
  return (size_t)len[id] * n;
}

extern "C" int melt_get_lag(melt_ctx* ctx, int id, double* host) {
  if (check_ctx(ctx) || id < 0 || id >= MELT_L_NLAG || !host || !ctx->has_geom) return MELT_ERR_ARG;
  MELT_HIP(hipStreamSynchronize(ctx->stream));
  MELT_HIP(hipMemcpy(host, ctx->d_lag[id], sizeof(double) * lag_count(ctx, id), hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int melt_set_lag(melt_ctx* ctx, int id, const double* host) {
  if (check_ctx(ctx) || id < 0 || id >= MELT_L_NLAG || !host || !ctx->has_geom) return MELT_ERR_ARG;
  cancel_ahead(ctx);
  MELT_HIP(hipStreamSynchronize(ctx->stream));
  MELT_HIP(hipMemcpy(ctx->d_lag[id], host, sizeof(double) * lag_count(ctx, id), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int melt_get_pind(melt_ctx* ctx, int32_t* host) {
  if (check_ctx(ctx) || !host || !ctx->has_geom) return MELT_ERR_ARG;
  MELT_HIP(hipStreamSynchronize(ctx->stream));
  MELT_HIP(hipMemcpy(host, ctx->d_pind, sizeof(int) * 6 * (size_t)ctx->nf, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int melt_get_vert_of_face(melt_ctx* ctx, int32_t* host) {
  if (check_ctx(ctx) || !host || !ctx->has_geom) return MELT_ERR_ARG;
  MELT_HIP(hipStreamSynchronize(ctx->stream));
  MELT_HIP(hipMemcpy(host, ctx->d_vof, sizeof(int) * 3 * (size_t)ctx->nf, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < 3 * (size_t)ctx->nf; ++i) host[i] += 1;
  return 0;
}

// ------------------------------------------------------------ time control
extern "C" int melt_set_dt(melt_ctx* ctx, double dt) {
  if (check_ctx(ctx) || !(dt > 0.0)) return MELT_ERR_ARG;
  if (dt != ctx->dt) graph_drop(ctx);
  ctx->dt = dt;
  return 0;
}

extern "C" int melt_set_stage(melt_ctx* ctx, int32_t ns) {
  if (check_ctx(ctx) || ns < 1 || ns > 3) return MELT_ERR_ARG;
  ctx->stage = ns;
  ctx->beta = ctx->dt / ctx->cfg.ren * 0.5;  // tsch.f90:18
  ctx->al = ctx->alm[ns - 1]; ctx->ga = ctx->gam[ns - 1]; ctx->ro = ctx->rom[ns - 1];
  return 0;
}

static int need_stage(melt_ctx* ctx) {
  if (ctx->stage < 1 || !(ctx->dt > 0.0)) return melt_fail(ctx, MELT_ERR_ARG, "stage", "call melt_set_dt and melt_set_stage first");
  return 0;
}

// ---------------------------------------------------- per-routine entries
extern "C" int melt_tagcells(melt_ctx* ctx) {
  if (check_ctx(ctx)) return MELT_ERR_ARG;
  cancel_ahead(ctx);
  int rc = ibm_tag(ctx);
  if (rc) return rc;
  // sphereTagging sets VOFx/y/z itself; interp_vof belongs to the narrow band (genSDF.f90:80)
  return ctx->cfg.tagtype == 0 ? 0 : euler_launch_interp_vof(ctx);
}
extern "C" int melt_interp_vof(melt_ctx* ctx) { return check_ctx(ctx) ? MELT_ERR_ARG : euler_launch_interp_vof(ctx); }

static int hdnl(melt_ctx* ctx, int which) {
  if (check_ctx(ctx)) return MELT_ERR_ARG;
  return euler_launch_hdnl(ctx, which);
}
extern "C" int melt_hdnl1(melt_ctx* ctx) { return hdnl(ctx, 0); }
extern "C" int melt_hdnl2(melt_ctx* ctx) { return hdnl(ctx, 1); }
extern "C" int melt_hdnl3(melt_ctx* ctx) { return hdnl(ctx, 2); }
extern "C" int melt_hdnlte(melt_ctx* ctx) { return hdnl(ctx, 3); }

static double betadx_q(melt_ctx* ctx, int which, int dir) {  // invtr1.f90:66-68, invtrte.f90:61-63
  double dq = dir == 0 ? ctx->dx1q : dir == 1 ? ctx->dx2q : ctx->dx3q;
  if (which < 3) return ctx->beta * ctx->al * dq;
  double pec = ctx->cfg.ren * ctx->cfg.prandtl;
  return ctx->dt / pec * 0.5 * ctx->al * dq;
}

static int invtr(melt_ctx* ctx, int which) {
  if (check_ctx(ctx)) return MELT_ERR_ARG;
  int rc = need_stage(ctx);
  if (rc) return rc;
  static const int q_id[4] = {MELT_VX, MELT_VY, MELT_VZ, MELT_TEMP};
  rc = euler_launch_invtr_rhs(ctx, which);
  if (rc) return rc;
  void* r[1] = {ctx->f[MELT_RHS]};
  double b0 = betadx_q(ctx, which, 0), b1 = betadx_q(ctx, which, 1), b2 = betadx_q(ctx, which, 2);
  if ((rc = euler_launch_solve(ctx, 0, r, 1, &b0, nullptr, 1))) return rc;
  if ((rc = euler_launch_solve(ctx, 1, r, 1, &b1, nullptr, 1))) return rc;
  void* q[1] = {ctx->f[q_id[which]]};
  return euler_launch_solve(ctx, 2, r, 1, &b2, q, 1);
}
extern "C" int melt_invtr1(melt_ctx* ctx) { return invtr(ctx, 0); }
extern "C" int melt_invtr2(melt_ctx* ctx) { return invtr(ctx, 1); }
extern "C" int melt_invtr3(melt_ctx* ctx) { return invtr(ctx, 2); }
extern "C" int melt_invtrte(melt_ctx* ctx) { return invtr(ctx, 3); }

extern "C" int melt_solxi(melt_ctx* ctx, double betadx) {
  if (check_ctx(ctx)) return MELT_ERR_ARG;
  void* r[1] = {ctx->f[MELT_RHS]};
  return euler_launch_solve(ctx, 0, r, 1, &betadx, nullptr, 1);
}
extern "C" int melt_solxj(melt_ctx* ctx, double betadx) {
  if (check_ctx(ctx)) return MELT_ERR_ARG;
  void* r[1] = {ctx->f[MELT_RHS]};
  return euler_launch_solve(ctx, 1, r, 1, &betadx, nullptr, 1);
}
extern "C" int melt_solxk(melt_ctx* ctx, int field, double betadx) {
  if (check_ctx(ctx) || field < 0 || field >= MELT_NFIELDS || field == MELT_RHS) return MELT_ERR_ARG;
  void* r[1] = {ctx->f[MELT_RHS]};
  void* q[1] = {ctx->f[field]};
  return euler_launch_solve(ctx, 2, r, 1, &betadx, q, 1);
}

// The reference's own form, solxk(q, betadx) with q an array section
// (solxk.f90:1,7; invtr1.f90:68): the z-line solves run on the device rhs
// (left holding the solution, as the reference's rhs is) and q += rhs is
// applied to the caller's array through its layout (solxk.f90:44-50).
extern "C" int melt_solxk_array(melt_ctx* ctx, double* q, const melt_layout* lay, double betadx) {
  if (check_ctx(ctx) || !q) return MELT_ERR_ARG;
  melt_layout L = lay ? *lay : melt_layout{ctx->n1, ctx->n2, 0, 0};
  if (L.ld1 < ctx->n1 || L.ld2 < ctx->n2 || L.kghost < 0) return melt_fail(ctx, MELT_ERR_ARG, "solxk", "bad layout");
  void* r[1] = {ctx->f[MELT_RHS]};
  int rc = euler_launch_solve(ctx, 2, r, 1, &betadx, nullptr, 1);
  if (rc) return rc;
  MELT_HIP(hipStreamSynchronize(ctx->stream));
  std::vector<double> tmp(ctx->nc);
  if ((rc = copy_from_device(ctx, MELT_RHS, tmp.data()))) return rc;
  for (int k = 0; k < ctx->n3; ++k)
    for (int j = 0; j < ctx->n2; ++j) {
      double* row = q + (size_t)L.ld1 * (j + (size_t)L.ld2 * (k + L.kghost));
      const double* src = &tmp[(size_t)ctx->n1 * (j + (size_t)ctx->n2 * k)];
      for (int i = 0; i < ctx->n1; ++i) row[i] = row[i] + src[i];
    }
  return 0;
}

extern "C" int melt_divg(melt_ctx* ctx) {
  if (check_ctx(ctx)) return MELT_ERR_ARG;
  int rc = need_stage(ctx);
  return rc ? rc : euler_launch_divg(ctx);
}
extern "C" int melt_phcalc(melt_ctx* ctx) { return check_ctx(ctx) ? MELT_ERR_ARG : poisson_solve(ctx, 0); }
extern "C" int melt_updvp(melt_ctx* ctx) {
  if (check_ctx(ctx)) return MELT_ERR_ARG;
  int rc = need_stage(ctx);
  return rc ? rc : euler_launch_project(ctx, 1, 0, 0);
}
extern "C" int melt_prcalc(melt_ctx* ctx) {
  if (check_ctx(ctx)) return MELT_ERR_ARG;
  int rc = need_stage(ctx);
  return rc ? rc : euler_launch_project(ctx, 0, 1, 0);
}

static int need_geom(melt_ctx* ctx) {
  return ctx->has_geom ? 0 : melt_fail(ctx, MELT_ERR_NO_GEOMETRY, "immersed boundary", "melt_set_geometry has not been called");
}
extern "C" int melt_find_centroid_indices(melt_ctx* ctx) {
  if (check_ctx(ctx)) return MELT_ERR_ARG;
  int rc = need_geom(ctx);
  return rc ? rc : ibm_find_centroid_indices(ctx);
}
extern "C" int melt_mlsweight(melt_ctx* ctx) {
  if (check_ctx(ctx)) return MELT_ERR_ARG;
  int rc = need_geom(ctx);
  return rc ? rc : ibm_weights(ctx);
}
extern "C" int melt_mlsforce(melt_ctx* ctx) {
  if (check_ctx(ctx)) return MELT_ERR_ARG;
  int rc = need_geom(ctx);
  if (rc || (rc = ibm_force_scatter(ctx))) return rc;
  return ibm_force_drain(ctx, 0);
}
extern "C" int melt_velforce(melt_ctx* ctx) { return check_ctx(ctx) ? MELT_ERR_ARG : ibm_velforce(ctx, 0); }
extern "C" int melt_tempforce(melt_ctx* ctx) { return check_ctx(ctx) ? MELT_ERR_ARG : ibm_tempforce(ctx); }

// particle.f90:38-193 (IREMESH = 0).  defer: leave the uniform mean-force shift
// to the projection (fused substep) instead of applying it here.
static int particle_impl(melt_ctx* ctx, int defer) {
  if (ctx->cfg.imlsfor != 1) return 0;
  int rc;
  if ((rc = ibm_find_centroid_indices(ctx))) return rc;
  if ((rc = ibm_weights(ctx))) return rc;
  if ((rc = ibm_force_scatter(ctx))) return rc;
  if ((rc = ibm_probes(ctx))) return rc;      // heat flux and loads read the pre-force fields
  if ((rc = ibm_force_drain(ctx, 1))) return rc;
  if (!defer && (rc = ibm_sub_mean(ctx))) return rc;
  return lagr_after_force(ctx);
}

extern "C" int melt_particle(melt_ctx* ctx) {
  if (check_ctx(ctx)) return MELT_ERR_ARG;
  cancel_ahead(ctx);
  int rc = need_geom(ctx);
  if (rc || (rc = need_stage(ctx))) return rc;
  return particle_impl(ctx, 0);
}

// ------------------------------------------------------------ fused path
// Queue `work` (launches on ctx->stream) on the side stream after everything
// queued on the main stream so far; side_join makes the main stream wait for
// it.  Fork and join are stream events, so a captured step keeps them as
// graph dependencies.
// Tuning key streams = 1 keeps all work on the main stream (many contexts
// sharing one GPU get their concurrency from each other).
static bool one_stream(const melt_ctx* ctx) { return ctx->tune[TUNE_STREAMS] == 1; }
// Where the geometry stage of the gradient probes runs: 1 = queued on the side
// stream after the previous substep's Lagrangian update and next tagging
// (beside the pressure solve), 2 = on the side stream at the start of the
// substep that uses it (beside the explicit pass).
static int coef_place(const melt_ctx* ctx) { return ctx->tune[TUNE_COEF] ? ctx->tune[TUNE_COEF] : 1; }
bool capturing(melt_ctx* ctx) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(ctx->stream, &st) != hipSuccess || st != hipStreamCaptureStatusNone;
}
static int ensure_side(melt_ctx* ctx) {
  if (!ctx->side && !one_stream(ctx)) MELT_HIP(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
  if (!ctx->side2 && !one_stream(ctx)) MELT_HIP(hipStreamCreateWithFlags(&ctx->side2, hipStreamNonBlocking));
  return 0;
}
template <class F>
static int side_run(melt_ctx* ctx, F&& work) {
  if (one_stream(ctx)) return work();
  if (int rc = ensure_side(ctx)) return rc;
  MELT_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
  MELT_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
  hipStream_t main_stream = ctx->stream;
  ctx->stream = ctx->side;
  int rc = work();
  ctx->stream = main_stream;
  if (rc) return rc;
  MELT_HIP(hipEventRecord(ctx->ev_join, ctx->side));
  return 0;
}
// The fork point of later side-stream work: the side stream waits for
// everything queued on the main stream so far, and work queued on it later
// (side_continue) starts from here, whatever the main stream gets meanwhile.
static int side_fork(melt_ctx* ctx) {
  if (one_stream(ctx)) return 0;
  if (int rc = ensure_side(ctx)) return rc;
  MELT_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
  MELT_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
  return 0;
}
// More work on the side stream after its last fork (no new fork).
template <class F>
static int side_continue(melt_ctx* ctx, F&& work) {
  if (one_stream(ctx)) return work();
  hipStream_t main_stream = ctx->stream;
  ctx->stream = ctx->side;
  int rc = work();
  ctx->stream = main_stream;
  if (rc) return rc;
  MELT_HIP(hipEventRecord(ctx->ev_join, ctx->side));
  return 0;
}
static int side_join(melt_ctx* ctx) {
  if (one_stream(ctx)) return 0;
  MELT_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
  if (ctx->join2_pending) {
    MELT_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_join2, 0));
    ctx->join2_pending = 0;
  }
  return 0;
}
// Queue `work`, which writes only the forcing and its own state, on the third
// stream from the last explicit pass's read of the forcing (ev_forc) when
// nothing on the main stream has touched the forcing since; the main stream
// waits for it.  Otherwise on the main stream.
int run_after_forcing_read(melt_ctx* ctx, const std::function<int()>& work) {
  if (!ctx->forc_free || one_stream(ctx) || !ctx->side2 || capturing(ctx)) return work();
  ctx->forc_free = 0;
  MELT_HIP(hipStreamWaitEvent(ctx->side2, ctx->ev_forc, 0));
  hipStream_t cur = ctx->stream;
  ctx->stream = ctx->side2;
  int rc = work();
  ctx->stream = cur;
  MELT_HIP(hipEventRecord(ctx->ev_hit, ctx->side2));
  MELT_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_hit, 0));
  return rc;
}
// Queue `work` on the third stream, forked from the stream now current (the
// side stream, inside side_continue); side_join waits for it.
template <class F>
static int side2_run(melt_ctx* ctx, F&& work) {
  MELT_HIP(hipEventRecord(ctx->ev_fork2, ctx->stream));
  MELT_HIP(hipStreamWaitEvent(ctx->side2, ctx->ev_fork2, 0));
  hipStream_t cur = ctx->stream;
  ctx->stream = ctx->side2;
  int rc = work();
  ctx->stream = cur;
  if (rc) return rc;
  MELT_HIP(hipEventRecord(ctx->ev_join2, ctx->side2));
  ctx->join2_pending = 1;
  return 0;
}

// One fused substep.  With tag_next (tschem's first two substeps, where no
// caller code runs before the next one) the tagging of the next substep is
// queued during this substep's Lagrangian update, on a third stream forked
// as soon as the geometry refresh has written the face centroids and normals:
// it reads only the vertices and faces and writes only VOFp, sdf, its own
// scratch and the window box, none of which the rest of the refresh, the
// probe stage or the pressure phase touches, so it overlaps them all; so
// does, with coef_next (every substep of tschem, also for the next call's
// first substep: a caller's change of the faces or body in between cancels
// it), the geometry part of the next substep's gradient probes (side stream,
// after the refresh; it writes only its coefficient buffer).  The tagging
// itself is not queued across calls: VOFp and sdf after tschem hold the last
// substep's tagging, as in the reference.
static int substep_impl(melt_ctx* ctx, int32_t ns, bool tag_next, bool coef_next) {
  if (check_ctx(ctx)) return MELT_ERR_ARG;
  int rc = melt_set_stage(ctx, ns);
  if (rc || (rc = need_stage(ctx))) return rc;
  const melt_config& c = ctx->cfg;
  bool ibm = c.imlsfor == 1 && ctx->has_geom;
  bool moving = ibm && (c.imlsstr == 1 || c.imelt == 1);  // papero.f90:95-100
  hipEvent_t e_sub = nullptr, e_ph = nullptr, e = nullptr;
  timing_begin(ctx, SLOT_SUBSTEP, &e_sub);

  timing_begin(ctx, SLOT_PH_TAG, &e_ph);
  if (moving && ctx->tag_ahead) {
    ctx->force_window_from_tag = 1;   // placed by the tagging queued last substep
  } else if (moving) {
    timing_begin(ctx, SLOT_TAG, &e);
    ctx->force_window_from_tag = 1;
    rc = ibm_tag(ctx);
    if (rc) { ctx->force_window_from_tag = 0; return rc; }
    timing_end(ctx, SLOT_TAG, e);
  }
  ctx->tag_ahead = 0;
  timing_end(ctx, SLOT_PH_TAG, e_ph);
  // the geometry part of the gradient probes reads only the faces (final
  // since the last Lagrangian update) and the grid: queued with the tagging
  // ahead of time, else on the side stream beside the explicit pass
  const bool probes = ibm && (c.imelt == 1 || c.imlsstr == 1);
  if (probes && !ctx->coef_ahead && (rc = side_run(ctx, [&] { return ibm_probe_coefs(ctx); }))) return rc;
  ctx->coef_ahead = 0;

  timing_begin(ctx, SLOT_PH_EULER, &e_ph);
  timing_begin(ctx, SLOT_EXPLICIT, &e);
  if ((rc = euler_launch_explicit_xsolve(ctx, betadx_q(ctx, 0, 0), betadx_q(ctx, 3, 0)))) return rc;
  timing_end(ctx, SLOT_EXPLICIT, e);
  if (ctx->hit_nm && !one_stream(ctx) && !capturing(ctx)) {
    MELT_HIP(hipEventRecord(ctx->ev_forc, ctx->stream));
    ctx->forc_free = 1;
  }
  void* r[4] = {ctx->rhs4[RHS_U], ctx->rhs4[RHS_V], ctx->rhs4[RHS_W], ctx->rhs4[RHS_T]};
  void* q[4] = {ctx->f[MELT_VX], ctx->f[MELT_VY], ctx->f[MELT_VZ], ctx->f[MELT_TEMP]};
  double b[4], bz[4];
  for (int m = 0; m < 4; ++m) b[m] = betadx_q(ctx, m, 1);
  for (int m = 0; m < 4; ++m) bz[m] = betadx_q(ctx, m, 2);
  // y and z in one pass: q += Gz Gy rhs
  timing_begin(ctx, SLOT_SOLVE_Y, &e);
  int used_yz = 0;
  if ((rc = euler_launch_green_yz(ctx, r, q, b, bz, 4, &used_yz))) return rc;
  // otherwise y: Green filter rhs -> tmp (or cyclic reduction in place)
  int used_y = 0;
  if (!used_yz && (rc = euler_launch_green(ctx, 1, r, ctx->tmp4, nullptr, b, 4, &used_y))) return rc;
  if (!used_yz && !used_y && (rc = euler_launch_solve(ctx, 1, r, 4, b, nullptr, 0))) return rc;
  timing_end(ctx, SLOT_SOLVE_Y, e);
  // z: q += solution (solxk.f90:44-50)
  void* const* zsrc = used_y ? ctx->tmp4 : r;
  timing_begin(ctx, SLOT_SOLVE_Z, &e);
  int used_z = 0;
  if (!used_yz && (rc = euler_launch_green(ctx, 2, zsrc, nullptr, q, bz, 4, &used_z))) return rc;
  if (!used_yz && !used_z && (rc = euler_launch_solve(ctx, 2, zsrc, 4, bz, q, 0))) return rc;
  timing_end(ctx, SLOT_SOLVE_Z, e);
  timing_end(ctx, SLOT_PH_EULER, e_ph);

  timing_begin(ctx, SLOT_PH_MLS, &e_ph);
  // what the side stream queues after the Lagrangian update: the next
  // substep's tagging (tag_next) and probe geometry stage (coef_next).  The
  // tagging needs only the refreshed faces (positions, centroids, normals):
  // with a second side stream it forks there as soon as the refresh has
  // written them, beside the rest of the refresh and the probe stage
  auto tag_now = [&]() -> int {
    hipEvent_t et = nullptr;   // its own event: it may run inside the Lagrangian slot
    timing_begin(ctx, SLOT_TAG, &et);
    ctx->force_window_from_tag = 1;
    int r = ibm_tag(ctx);
    ctx->force_window_from_tag = 0;
    timing_end(ctx, SLOT_TAG, et);
    if (!r) ctx->tag_ahead = 1;
    return r;
  };
  const std::function<int()> tag_fork = [&]() -> int { return side2_run(ctx, tag_now); };
  auto queue_coef = [&]() -> int {
    int r = ibm_probe_coefs(ctx);
    if (!r) ctx->coef_ahead = 1;
    return r;
  };
  // within a tschem call (tag_next) the main stream does not wait for the
  // probe geometry stage: only the next substep's probes read its output, and
  // they run on the side stream after it; the call's last substep joins it
  const bool coef_after_join = tag_next && coef_next && probes && !one_stream(ctx);
  auto queue_next = [&]() -> int {
    int r = 0;
    if (tag_next && moving && !ctx->tag_ahead) r = tag_now();
    if (!r && coef_next && probes && !coef_after_join) r = queue_coef();
    return r;
  };
  if (ibm) {
    timing_begin(ctx, SLOT_IBM, &e);
    if ((rc = ibm_weights(ctx, true))) return rc;   // cell indices and shape functions
    // the gradient probes read the same pre-force fields as the spreading and
    // write only their own outputs (their cell indices into d_pind2, since the
    // spreading still reads d_pind): side stream, alongside the spreading
    if ((rc = side_run(ctx, [&] { return ibm_probes(ctx, ctx->d_pind2, probes); }))) return rc;
    rc = ibm_force_scatter(ctx);
    ctx->force_window_from_tag = 0;
    if (rc || (rc = side_join(ctx)) || (rc = ibm_force_drain(ctx, 1))) return rc;
    timing_end(ctx, SLOT_IBM, e);
    // The Lagrangian update after the forcing (loads, heat flux, Stefan
    // melting, Newton-Euler motion, remeshing, geometry refresh) reads only
    // the probe results and the body arrays, which the pressure phase does
    // not touch: it runs on the side stream from here, concurrently with the
    // pressure solve, and rejoins before the substep ends (the next tagging
    // needs the geometry).  Its 20-30 short launches are queued after the
    // pressure phase's, so that the host has queued the pressure solve before
    // the device needs it (queued first, they left the main stream idle for
    // 0.1-0.23 ms per substep while the host submitted them).
    if ((rc = side_fork(ctx))) return rc;
  }
  timing_end(ctx, SLOT_PH_MLS, e_ph);

  ctx->force_window_from_tag = 0;
  timing_begin(ctx, SLOT_PH_PRESSURE, &e_ph);
  // the precision form measures the FP64 divergence of the provisional field
  // (before) and of the projected one (after) for its refinement gate
  unsigned long long* dmax = reinterpret_cast<unsigned long long*>(ctx->d_refine_stats + 4);
  const bool refine = ctx->fp32 && ctx->refine_max > 0;
  if ((rc = poisson_solve(ctx, 1, nullptr, refine ? ctx->d_slots : nullptr))) return rc;
  if (refine && (rc = euler_fold_slots(ctx, ctx->d_slots, dmax))) return rc;
  timing_begin(ctx, SLOT_PROJECT, &e);
  if ((rc = euler_launch_project(ctx, 1, 1, ibm ? 1 : 0))) return rc;
  timing_end(ctx, SLOT_PROJECT, e);
  if (refine) {
    // iterative refinement: while the FP64 residual div(u) exceeds rtol times
    // the provisional one, solve for the correction of the FP32 solve and
    // project again (updvp/prcalc are linear in phi); gated on the device
    timing_begin(ctx, SLOT_REFINE, &e);
    for (int it = 0; it < ctx->refine_max; ++it) {
      if ((rc = euler_launch_reduce_max(ctx, 1, dmax + 1)) || (rc = euler_launch_refine_gate(ctx, dmax, ctx->d_gate)) ||
          (rc = poisson_solve(ctx, 1, ctx->d_gate)) || (rc = euler_launch_project(ctx, 1, 1, 0, ctx->d_gate)))
        return rc;
    }
    timing_end(ctx, SLOT_REFINE, e);
  }
  timing_end(ctx, SLOT_PH_PRESSURE, e_ph);
  // the Lagrangian update from the fork point above (with remeshing the host
  // waits for the edge flag here, while the pressure solve runs), then the
  // next substep's tagging and probe geometry stage
  if (ibm && (rc = side_continue(ctx, [&] {
                timing_begin(ctx, SLOT_LAGR, &e);
                if (tag_next && moving && !one_stream(ctx)) ctx->after_faces = &tag_fork;
                int r = lagr_after_force(ctx, true);
                ctx->after_faces = nullptr;
                timing_end(ctx, SLOT_LAGR, e);
                return r ? r : queue_next();
              })))
    return rc;
  if (ibm && coef_after_join) {   // on the side stream after its join point
    // (measured: waiting for the next tagging first, so that the two do not
    // share the CUs, moves the stage beside the next explicit pass, which it
    // then slows by as much as the tagging gains)
    hipStream_t main_stream = ctx->stream;
    ctx->stream = ctx->side;
    rc = queue_coef();
    ctx->stream = main_stream;
    if (rc) return rc;
  }
  if (ibm && (rc = side_join(ctx))) return rc;
  timing_end(ctx, SLOT_SUBSTEP, e_sub);
  if (ctx->timing && ctx->ev_pending.size() > 4096) timing_flush(ctx);
  return 0;
}

// A substep outside tschem consumes the queued probe coefficients and moves the
// faces without queueing new ones, so a captured step that starts from queued
// coefficients no longer holds.
extern "C" int melt_substep(melt_ctx* ctx, int32_t ns) {
  if (check_ctx(ctx)) return MELT_ERR_ARG;
  if (ctx->graph_uses_ahead) graph_drop(ctx);
  return substep_impl(ctx, ns, false, false);
}

static int tschem_launch(melt_ctx* ctx) {
  if (ctx->has_geom) {
    hipLaunchKernelGGL(k_reset_loads, dim3(1), dim3(1), 0, ctx->stream, ctx->d_body);
    MELT_HIP(hipGetLastError());
  }
  for (int ns = 1; ns <= 3; ++ns) {
    int rc = substep_impl(ctx, ns, ns < 3 && !one_stream(ctx), !one_stream(ctx) && coef_place(ctx) == 1);
    if (rc) {
      ctx->tag_ahead = 0;
      ctx->coef_ahead = 0;
      if (ctx->graph_uses_ahead) graph_drop(ctx);
      return rc;
    }
  }
  return 0;
}

extern "C" int melt_tschem(melt_ctx* ctx) {
  if (check_ctx(ctx)) return MELT_ERR_ARG;
  if (!(ctx->dt > 0.0)) return melt_fail(ctx, MELT_ERR_ARG, "tschem", "dt not set");
  // the first tagging (ray casting, host round trip) happens outside graphs
  // the first tagging (ray casting) and remeshing (a host decision every substep) are not captured
  bool graph = ctx->use_graph && !ctx->timing && !(ctx->has_geom && ctx->h_body.initial_tag) && ctx->cfg.iremesh != 1;
  if (!graph) return tschem_launch(ctx);
  if (!ctx->graph_exec) {
    if (int rc = ensure_side(ctx)) return rc;   // not created inside the capture
    const int uses_ahead = ctx->coef_ahead;
    MELT_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = tschem_launch(ctx);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc) { if (g) hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return melt_fail(ctx, MELT_ERR_HIP, "tschem graph capture", hipGetErrorString(e));
    ctx->graph = g;
    MELT_HIP(hipGraphInstantiate(&ctx->graph_exec, ctx->graph, nullptr, nullptr, 0));
    ctx->graph_uses_ahead = uses_ahead;
  }
  MELT_HIP(hipGraphLaunch(ctx->graph_exec, ctx->stream));
  return 0;
}

// Replay tschem as one captured HIP graph (per context, per dt).
extern "C" int melt_graph_enable(melt_ctx* ctx, int32_t on) {
  if (check_ctx(ctx)) return MELT_ERR_ARG;
  if (!on) graph_drop(ctx);
  ctx->use_graph = on ? 1 : 0;
  return 0;
}

// Launch-shape override (tile sweeps / autotuning): key 0 xf_w, 1 xf_nt,
// 2 solve_w, 3 solve_nt; value 0 restores the built-in choice.
extern "C" int melt_set_tuning(melt_ctx* ctx, int32_t key, int32_t value) {
  if (check_ctx(ctx) || key < 0 || key >= TUNE_N || value < 0) return MELT_ERR_ARG;
  graph_drop(ctx);
  ctx->tune[key] = value;
  return 0;
}

// The context's HIP stream (hipStream_t), for callers that order their own work
// against it or overlap several contexts.
extern "C" void* melt_stream(melt_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

// ------------------------------------------------------------ diagnostics
extern "C" int melt_cfl(melt_ctx* ctx, double* cflm) {
  if (check_ctx(ctx) || !cflm) return MELT_ERR_ARG;
  return euler_reduce_max(ctx, 0, cflm);
}
extern "C" int melt_divgck(melt_ctx* ctx, double* qmax) {
  if (check_ctx(ctx) || !qmax) return MELT_ERR_ARG;
  return euler_reduce_max(ctx, 1, qmax);
}
extern "C" int melt_inqpr(melt_ctx* ctx) {
  if (check_ctx(ctx)) return MELT_ERR_ARG;
  return euler_launch_inqpr(ctx);
}

// ------------------------------------------------------------ timing API
extern "C" int melt_timing_enable(melt_ctx* ctx, int32_t on) {
  return melt_timing_select(ctx, on ? (1u << SLOT_COUNT) - 1u : 0u);
}
extern "C" int melt_timing_select(melt_ctx* ctx, uint32_t slot_mask) {
  if (check_ctx(ctx) || (slot_mask >> SLOT_COUNT)) return MELT_ERR_ARG;
  if (!slot_mask) timing_flush(ctx);
  ctx->timing = slot_mask ? 1 : 0;
  ctx->timing_mask = slot_mask;
  return 0;
}
extern "C" int melt_timing_reset(melt_ctx* ctx) {
  if (check_ctx(ctx)) return MELT_ERR_ARG;
  timing_flush(ctx);
  for (int i = 0; i < SLOT_COUNT; ++i) { ctx->t_ms[i] = 0; ctx->t_n[i] = 0; }
  return 0;
}
extern "C" int melt_timing_count(void) { return SLOT_COUNT; }
extern "C" const char* melt_timing_name(int32_t slot) { return (slot >= 0 && slot < SLOT_COUNT) ? slot_names[slot] : ""; }
extern "C" int melt_timing_read(melt_ctx* ctx, double* ms, int64_t* samples) {
  if (check_ctx(ctx) || !ms || !samples) return MELT_ERR_ARG;
  timing_flush(ctx);
  for (int i = 0; i < SLOT_COUNT; ++i) { ms[i] = ctx->t_ms[i]; samples[i] = ctx->t_n[i]; }
  return 0;
}

// ------------------------------------------------------ precision form
extern "C" int melt_set_refinement(melt_ctx* ctx, int32_t max_iter, double rtol) {
  if (check_ctx(ctx) || max_iter < 0 || max_iter > 8 || !(rtol >= 0.0)) return MELT_ERR_ARG;
  graph_drop(ctx);
  ctx->refine_max = max_iter;
  ctx->refine_rtol = rtol;
  return 0;
}

extern "C" int melt_get_refinement_stats(melt_ctx* ctx, int64_t* checks, int64_t* refined, double* last_ratio) {
  if (check_ctx(ctx) || !checks || !refined || !last_ratio) return MELT_ERR_ARG;
  long long h[3];
  MELT_HIP(hipMemcpyAsync(h, ctx->d_refine_stats, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
  MELT_HIP(hipStreamSynchronize(ctx->stream));
  *checks = h[0];
  *refined = h[1];
  memcpy(last_ratio, &h[2], sizeof(double));
  return 0;
}

// ------------------------------------------------------- snapshot / resume
// The continuation files of the reference (mpi_write_continua /
// mpi_read_continua, mpi_routines.f90:1223-1631: the fields; continua_particle
// / import_particles, create_geo.f90:135-409: the body and its vertices; the
// bhat file, hit.f90:267-356) as one versioned host blob of the context's
// state at a step boundary: header, body scalars, immersed-boundary windows,
// the fields that carry state between steps (stored at the context's
// precision), every Lagrangian array, the probe indices, the ghost flags and
// the connectivity (which remeshing edits), and the HIT forcing state with its
// generator position.  A context created
// with the same configuration and mesh resumes from it bit for bit.
namespace {

constexpr char SNAP_MAGIC[8] = {'M', 'E', 'L', 'T', 'S', 'N', 'P', '1'};
constexpr uint32_t SNAP_VERSION = 2;
const int SNAP_FIELDS[] = {MELT_VX, MELT_VY, MELT_VZ, MELT_TEMP, MELT_PR, MELT_RU1, MELT_RU2, MELT_RU3, MELT_RUT,
                           MELT_FORCX, MELT_FORCY, MELT_FORCZ, MELT_VOFP, MELT_VOFX, MELT_VOFY, MELT_VOFZ, MELT_SDF};

struct SnapHeader {
  char magic[8];
  uint32_t version, pad0;
  melt_config cfg;
  int32_t has_geom, nv, ne, nf;
  int32_t hit_nm, pad1;
  uint64_t bytes, mesh_hash;
  double dt, rad_p, hit_tl, hit_eps, hit_kf;
  uint64_t hit_seed, hit_steps;
};

// One entry per stored device array: pointer and byte count.
std::vector<std::pair<void*, size_t>> snap_arrays(melt_ctx* ctx) {
  std::vector<std::pair<void*, size_t>> a;
  a.push_back({ctx->d_body, sizeof(melt_body)});
  a.push_back({ctx->d_win, sizeof(melt_window) * 2});
  for (int id : SNAP_FIELDS) a.push_back({ctx->f[id], ctx->esz * ctx->nc});
  if (ctx->has_geom) {
    for (int id = 0; id < MELT_L_NLAG; ++id) a.push_back({ctx->d_lag[id], sizeof(double) * lag_count(ctx, id)});
    a.push_back({ctx->d_pind, sizeof(int) * 6 * (size_t)ctx->nf});
    a.push_back({ctx->d_gv, (size_t)ctx->nv});
    a.push_back({ctx->d_ge, (size_t)ctx->ne});
    a.push_back({ctx->d_gf, (size_t)ctx->nf});
    // the connectivity as remeshing left it
    a.push_back({ctx->d_vof, sizeof(int) * 3 * (size_t)ctx->nf});
    a.push_back({ctx->d_eof, sizeof(int) * 3 * (size_t)ctx->nf});
    a.push_back({ctx->d_voe, sizeof(int) * 2 * (size_t)ctx->ne});
    a.push_back({ctx->d_foe, sizeof(int) * 2 * (size_t)ctx->ne});
    a.push_back({ctx->d_v2f_ptr, sizeof(int) * ((size_t)ctx->nv + 1)});
    a.push_back({ctx->d_v2f, sizeof(int) * 3 * (size_t)ctx->nf});
  }
  if (ctx->hit_nm) a.push_back({ctx->d_hit_bhat, sizeof(double) * 2 * 3 * (size_t)ctx->hit_nm * ctx->hit_nm * ctx->hit_nm});
  return a;
}

// the mesh the context imported (remeshing later edits it; the blob carries that)
int snap_mesh_hash(melt_ctx* ctx, uint64_t* h) {
  *h = ctx->has_geom ? ctx->mesh_hash0 : 0;
  return 0;
}

}  // namespace

extern "C" int melt_snapshot_size(melt_ctx* ctx, uint64_t* bytes) {
  if (check_ctx(ctx) || !bytes) return MELT_ERR_ARG;
  uint64_t n = sizeof(SnapHeader);
  for (auto& e : snap_arrays(ctx)) n += e.second;
  *bytes = n;
  return 0;
}

extern "C" int melt_snapshot_save(melt_ctx* ctx, void* buf, uint64_t bytes) {
  if (ctx) ctx->forc_free = 0;   // it may touch the forcing on the main stream
  if (check_ctx(ctx) || !buf) return MELT_ERR_ARG;
  uint64_t need = 0;
  melt_snapshot_size(ctx, &need);
  if (bytes < need) return melt_fail(ctx, MELT_ERR_ARG, "snapshot_save", "buffer smaller than melt_snapshot_size");
  MELT_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->side) MELT_HIP(hipStreamSynchronize(ctx->side));
  SnapHeader h;
  memset(&h, 0, sizeof h);
  memcpy(h.magic, SNAP_MAGIC, 8);
  h.version = SNAP_VERSION;
  h.cfg = ctx->cfg;
  h.has_geom = ctx->has_geom; h.nv = ctx->nv; h.ne = ctx->ne; h.nf = ctx->nf;
  h.hit_nm = ctx->hit_nm;
  h.bytes = need;
  if (int rc = snap_mesh_hash(ctx, &h.mesh_hash)) return rc;
  h.dt = ctx->dt; h.rad_p = ctx->rad_p;
  h.hit_tl = ctx->hit_tl; h.hit_eps = ctx->hit_eps; h.hit_kf = ctx->hit_kf; h.hit_seed = ctx->hit_seed;
  if (ctx->hit_nm) {
    unsigned long long s = 0;
    MELT_HIP(hipMemcpy(&s, ctx->d_hit_count, sizeof s, hipMemcpyDeviceToHost));
    h.hit_steps = s;
  }
  char* p = static_cast<char*>(buf);
  memcpy(p, &h, sizeof h);
  p += sizeof h;
  for (auto& e : snap_arrays(ctx)) {
    MELT_HIP(hipMemcpy(p, e.first, e.second, hipMemcpyDeviceToHost));
    p += e.second;
  }
  return 0;
}

extern "C" int melt_snapshot_load(melt_ctx* ctx, const void* buf, uint64_t bytes) {
  if (ctx) ctx->forc_free = 0;   // it may touch the forcing on the main stream
  if (check_ctx(ctx) || !buf || bytes < sizeof(SnapHeader)) return MELT_ERR_ARG;
  SnapHeader h;
  memcpy(&h, buf, sizeof h);
  if (memcmp(h.magic, SNAP_MAGIC, 8) || h.version != SNAP_VERSION)
    return melt_fail(ctx, MELT_ERR_ARG, "snapshot_load", "not a snapshot of this library version");
  if (memcmp(&h.cfg, &ctx->cfg, sizeof(melt_config)))
    return melt_fail(ctx, MELT_ERR_ARG, "snapshot_load", "the context's configuration differs from the snapshot's");
  if (h.has_geom != ctx->has_geom || h.nv != ctx->nv || h.ne != ctx->ne || h.nf != ctx->nf)
    return melt_fail(ctx, MELT_ERR_ARG, "snapshot_load", "the context's mesh differs from the snapshot's");
  uint64_t mh = 0;
  if (int rc = snap_mesh_hash(ctx, &mh)) return rc;
  if (mh != h.mesh_hash) return melt_fail(ctx, MELT_ERR_ARG, "snapshot_load", "the context's mesh connectivity differs from the snapshot's");
  if (h.hit_nm && (h.hit_nm != ctx->hit_nm || h.hit_kf != ctx->hit_kf)) {
    if (int rc = melt_hit_init(ctx, h.hit_tl, h.hit_eps, h.hit_kf, h.hit_seed)) return rc;
  }
  if (h.hit_nm) { ctx->hit_tl = h.hit_tl; ctx->hit_eps = h.hit_eps; ctx->hit_seed = h.hit_seed; }
  uint64_t need = 0;
  melt_snapshot_size(ctx, &need);
  if (bytes < need || h.bytes != need) return melt_fail(ctx, MELT_ERR_ARG, "snapshot_load", "snapshot size does not match the context");
  cancel_ahead(ctx);
  graph_drop(ctx);
  MELT_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->side) MELT_HIP(hipStreamSynchronize(ctx->side));
  const char* p = static_cast<const char*>(buf) + sizeof h;
  for (auto& e : snap_arrays(ctx)) {
    MELT_HIP(hipMemcpy(e.first, p, e.second, hipMemcpyHostToDevice));
    p += e.second;
  }
  if (h.hit_nm) {
    unsigned long long s = h.hit_steps;
    MELT_HIP(hipMemcpy(ctx->d_hit_count, &s, sizeof s, hipMemcpyHostToDevice));
  }
  ctx->dt = h.dt;
  MELT_HIP(hipMemcpy(&ctx->h_body, ctx->d_body, sizeof(melt_body), hipMemcpyDeviceToHost));
  return remesh_pull_mirrors(ctx);
}
