/*
 * melt_ibm.h -- C ABI of libmeltgpu.so, the MI355X (gfx950) implementation of the
 * melt_IBM fluid / immersed-boundary RK3 substep.
 *
 * The reference (kevzhong/melt_IBM) is an MPI Fortran program whose hot-path
 * subroutines take no arguments and communicate through module globals
 * (param.f90:4-274).  Here those globals live in a device-resident context
 * (melt_ctx); every drop-in entry point below operates on that context exactly as
 * the reference subroutine operates on its module state.  The Fortran binding
 * (melt_ibm_amd/fortran/melt_ibm_f.f90) wraps these with the reference's names.
 *
 * Conventions
 *  - All entry points return 0 (MELT_OK) on success, a MELT_ERR_* code otherwise;
 *    melt_last_error() gives the message.  Work is enqueued on the context's HIP
 *    stream; melt_sync() waits and reports errors raised by device code (the
 *    reference's fatal paths: singular MLS moment matrix, small body volume).
 *  - Field arrays are N1*N2*N3 doubles, i fastest (Fortran order), without the
 *    reference's padding/ghost planes; periodicity is index arithmetic.  The
 *    layout argument of melt_upload_field/melt_download_field lets a caller pass a
 *    Fortran module array with its (1:n1,1:n2,kstart-1:kend+1) shape directly.
 *  - Mesh connectivity is passed 1-based, as read from a GTS file.
 *  - One host thread drives one context.  Independent contexts are independent.
 */
#ifndef MELT_IBM_H
#define MELT_IBM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- errors */
enum melt_status {
  MELT_OK = 0,
  MELT_ERR_ARG = 1,             /* bad argument / unsupported size              */
  MELT_ERR_HIP = 2,             /* HIP runtime error                            */
  MELT_ERR_SINGULAR_MLS = 3,    /* dgetrf info /= 0: mlsWeight.f90:121-124      */
  MELT_ERR_SMALL_VOLUME = 4,    /* Volume < V_thresh: particle.f90:218-224      */
  MELT_ERR_PENCIL_OVERFLOW = 5, /* > BUFF_SIZE faces in a pencil: pencil_rayTag.f90:126-130 */
  MELT_ERR_REMESH = 6,          /* remeshing failed: topology not preserved, geometry reduced to a
                                   tetrahedron, singular QEM system or volume residual > 1e-14
                                   (remesh_coarsen.f90:69-253, particle.f90:199-211) */
  MELT_ERR_RAY_RECURSION = 7,   /* parallel ray recursion did not terminate: rayAux.f90:37-44 */
  MELT_ERR_NO_GEOMETRY = 8,     /* immersed-boundary call before melt_set_geometry */
  MELT_ERR_WINDOW = 9           /* body outgrew the device window sized at melt_set_geometry */
};

/* ------------------------------------------------------------ field ids */
enum melt_field_id {
  MELT_VX = 0, MELT_VY, MELT_VZ, MELT_TEMP, MELT_PR,       /* local_arrays, param.f90:87-96 */
  MELT_RU1, MELT_RU2, MELT_RU3, MELT_RUT,                  /* RK3 explicit-term history   */
  MELT_FORCX, MELT_FORCY, MELT_FORCZ,                      /* HIT forcing (input)         */
  MELT_VOFP, MELT_VOFX, MELT_VOFY, MELT_VOFZ, MELT_SDF,    /* param.f90:31-36             */
  MELT_DQ, MELT_DPH, MELT_QCAP, MELT_HTEMP, MELT_RHS,      /* hdnl* outputs, Poisson, rhs */
  MELT_FOR_XC, MELT_FOR_YC, MELT_FOR_ZC, MELT_FOR_TEMP,    /* mls_local, param.f90:265-268 (ghost-folded) */
  MELT_NFIELDS
};

/* ------------------------------------------------- Lagrangian array ids */
enum melt_lag_id {
  /* per vertex (nv) */
  MELT_L_XYZV = 0,     /* 3*nv  xyzv(3,nv)                 */
  MELT_L_VERT_NOR,     /* 3*nv  vert_nor                   */
  MELT_L_AVERT,        /*   nv  Avert                      */
  MELT_L_QW_OVERT,     /*   nv  qw_oVert                   */
  MELT_L_QW_IVERT,     /*   nv  qw_iVert                   */
  MELT_L_VMELT,        /* 3*nv  vmelt                      */
  MELT_L_VMELT_M1,     /* 3*nv  vmelt_m1                   */
  MELT_L_DXYZV_S,      /* 3*nv  dxyzv_s                    */
  /* per face (nf) */
  MELT_L_TRI_BAR,      /* 3*nf  tri_bar                    */
  MELT_L_TRI_NOR,      /* 3*nf  tri_nor                    */
  MELT_L_SUR,          /*   nf  sur                        */
  MELT_L_VEL_TRI,      /* 3*nf  vel_tri                    */
  MELT_L_CFAC,         /*   nf  cfac                       */
  MELT_L_SKEWNESS,     /*   nf  skewness                   */
  MELT_L_QW_O,         /*   nf  qw_o                       */
  MELT_L_QW_I,         /*   nf  qw_i                       */
  MELT_L_PTXAB_Q1,     /* 27*nf ptxAB_q1(27,nf)            */
  MELT_L_PTXAB_Q2,     /* 27*nf                            */
  MELT_L_PTXAB_Q3,     /* 27*nf                            */
  MELT_L_PTXAB_TEMP,   /* 27*nf                            */
  MELT_L_DXYZ_S,       /* 3*nf  dxyz_s                     */
  MELT_L_PRESS_N_TRI,  /* 3*nf  press_n_tri                */
  MELT_L_TAU_N,        /* 3*nf  (tau_n1,tau_n2,tau_n3) per face */
  /* per edge (ne) */
  MELT_L_ELENGTHS,     /*   ne  eLengths                   */
  MELT_L_NLAG
};

/* -------------------------------------------------------------- config */
/* Input-deck values (bou.in, part.in, HITForcing.in as read at papero.f90:31-71). */
typedef struct melt_config {
  int32_t n1m, n2m, n3m;   /* cells per side; the reference requires n1m == n2m   */
  int32_t precision;       /* 0: FP64 throughout; 1: mixed (FP32 bulk, FP64 refinement) */
  double xlen, ylen, zlen;
  double ren, prandtl, betagz;
  double tmelt, tliq, tsol, latheat, cpliquid;
  double dens_ratio;       /* part.in rhop/rhof                                    */
  int32_t imlsfor, imlsstr, imelt, iremesh;
  int32_t tagtype;         /* 1: ray tagging at init + narrow band (genSDF.f90)    */
  int32_t forcing;         /* 1: forcx/forcy/forcz are applied (hdnl1.f90:74)      */
  double perc_ethresh, v_on_ve_perc;
} melt_config;

/* Rigid-body / Lagrangian scalars (mls_param, param.f90:160-262). */
typedef struct melt_body {
  double pos_cm[3], vel_cm[3], omega_c[3], quat[4];
  double inert_tensor[9];                       /* InertTensor(3,3), column-major */
  double volume, surface;
  double int_prn_dA[3], int_tau_dA[3], int_r_x_prn_dA[3], int_r_x_tau_dA[3];
  double int_prn_dA_m1[3], int_tau_dA_m1[3], int_r_x_prn_dA_m1[3], int_r_x_tau_dA_m1[3];
  double fp[3], ftau[3], torq_p[3], torq_tau[3]; /* tsch.f90:20-23 accumulators  */
  double h_eulerian, celvol, e_thresh, v_thresh; /* initmls.f90:15-24             */
  int32_t rm_flag, initial_tag;
} melt_body;

/* Host-array layout for field copies: element (i,j,k) (1-based, k in 1..n3m) is
 * at host[(i-1) + ld1*((j-1) + ld2*(k-1+kghost))].  kghost = 1 means the array
 * has the reference's ghost planes k=0 and k=n3m+1; melt_download_field then
 * fills them with the periodic copies update_both_ghosts would hold
 * (mpi_routines.f90:139-210 on one rank).  ld1 >= n1m, ld2 >= n2m. */
typedef struct melt_layout {
  int32_t ld1, ld2, kghost, pad;
} melt_layout;

typedef struct melt_ctx melt_ctx;

/* ------------------------------------------------------------ lifecycle */
int  melt_ctx_create(const melt_config* cfg, melt_ctx** out);
void melt_ctx_destroy(melt_ctx* ctx);
const char* melt_last_error(const melt_ctx* ctx);
int  melt_sync(melt_ctx* ctx);
/* No counterpart in the reference: *done = 1 when all work queued on the
 * context has finished, 0 while some is pending (never blocks), for a caller
 * that schedules several contexts itself; device errors are then read by
 * melt_sync, which returns at once. */
int  melt_poll(melt_ctx* ctx, int32_t* done);
int  melt_device_ok(void);   /* 1 if a HIP device is present */

/* ------------------------------------------------------- data movement */
int melt_upload_field(melt_ctx* ctx, int field, const double* host, const melt_layout* lay);
int melt_download_field(melt_ctx* ctx, int field, double* host, const melt_layout* lay);
int melt_fill_field(melt_ctx* ctx, int field, double value);
/* No counterpart in the reference (its fields live in host memory): upload
 * from device memory, src holding N1*N2*N3 doubles (i fastest) that are
 * complete before the call, for callers that synthesise initial states on the
 * GPU.  Returns once the copy is done. */
int melt_upload_field_device(melt_ctx* ctx, int field, const double* src);

/* Mesh import: set_connectivity + setup_particles + init_geomCoords + the
 * geometry set-up of tri_geo + init_mlsForce (partaux.f90:160-267,
 * create_geo.f90:1-132, tri_geo.f90:1-50, initmls.f90:1-27).  xyz0 is the
 * unit mesh as read from the GTS file (3*nv, xyz interleaved), scaled by rad_p and
 * centred in the box; vert_of_edge (2*ne) and edge_of_face (3*nf) are 1-based. */
int melt_set_geometry(melt_ctx* ctx, int32_t nv, int32_t ne, int32_t nf,
                      const double* xyz0, const int32_t* vert_of_edge,
                      const int32_t* edge_of_face, double rad_p);
int melt_get_body(melt_ctx* ctx, melt_body* out);
int melt_set_body(melt_ctx* ctx, const melt_body* in);
int melt_get_lag(melt_ctx* ctx, int id, double* host);
int melt_set_lag(melt_ctx* ctx, int id, const double* host);
int melt_get_pind(melt_ctx* ctx, int32_t* host);           /* pind(6,nf)             */
int melt_get_vert_of_face(melt_ctx* ctx, int32_t* host);   /* 3*nf, 1-based          */
/* The connectivity as remeshing leaves it (allotri.f90:13-28: vert_of_face,
 * edge_of_face 3*nf, vert_of_edge, face_of_edge 2*ne, 1-based) and the ghost
 * flags isGhostVert/Edge/Face (0/1); any pointer may be null. */
int melt_get_topology(melt_ctx* ctx, int32_t* vert_of_face, int32_t* edge_of_face, int32_t* vert_of_edge, int32_t* face_of_edge,
                      int32_t* is_ghost_vert, int32_t* is_ghost_edge, int32_t* is_ghost_face);
/* Remeshing counters (the figures writeRemeshStats reports, particle.f90:196):
 * out[8] = {events so far, edge collapses, vertex deletions, edge relaxations,
 * largest vertex drift, vol_melt - vol_smooth, volume refreshes (all of the
 * last event), collapses so far}. */
int melt_get_remesh_stats(melt_ctx* ctx, double* out);

/* ------------------------------------------------------------ remeshing */
/* remesh_coarsen (remesh_coarsen.f90:2-258) on host arrays with the
 * reference's layout: 1-based connectivity, logical arrays as int32 0/1,
 * tri_nor(3,nf) and eLengths(ne) as particle.f90:127-133 leave them, rm_flag
 * in/out.  vmelt (3*nv, may be null) is mls_param's vmelt(:,:,1), averaged
 * onto the kept vertex of each collapse.  Counts of edge collapses and
 * valence-3 vertex deletions go to *ecol_cnt, *vdel_cnt.  Runs on the host.
 * MELT_ERR_REMESH where the reference aborts (topology not preserved,
 * geometry reduced to a tetrahedron, QEM solve) or would overrun its 30-entry
 * ring buffers. */
int melt_remesh_coarsen(int32_t nf, int32_t ne, int32_t nv, double* xyz, double* tri_nor, double* e_lengths, double e_thresh,
                        int32_t* vert_of_face, int32_t* edge_of_face, int32_t* vert_of_edge, int32_t* face_of_edge, int32_t* is_ghost_face,
                        int32_t* is_ghost_edge, int32_t* is_ghost_vert, int32_t* rm_flag, int32_t* anchor_vert, int32_t* flagged_edge,
                        double* vmelt, int32_t* ecol_cnt, int32_t* vdel_cnt);
/* remesh_smooth (remesh_smooth.f90:3-347): volume-restoring edge relaxation
 * of the flagged edges toward vol_old with the volume change target_dv; the
 * relaxation count, largest vertex drift and volume refreshes come back as
 * n_erel, drift, cnt_refresh (the reference's outputs).  Runs on the host. */
int melt_remesh_smooth(double vol_old, double target_dv, int32_t nv, int32_t ne, int32_t nf, double* xyz, const int32_t* is_ghost_vert,
                       const int32_t* is_ghost_edge, const int32_t* is_ghost_face, int32_t* flagged_edge, const int32_t* vert_of_edge,
                       const int32_t* vert_of_face, const int32_t* face_of_edge, const int32_t* edge_of_face, int32_t* n_erel, double* drift,
                       int32_t* cnt_refresh);

/* --------------------------------------------------------- time control */
int melt_set_dt(melt_ctx* ctx, double dt);       /* param dt                      */
int melt_set_stage(melt_ctx* ctx, int32_t ns);   /* al,ga,ro = alm,gam,rom(ns), beta = dt/ren*0.5
                                                    (papero.f90:116-137, tsch.f90:18,33-35) */

/* ------------------------------------------------------ drop-in entries */
/* Fused RK3 step: 3 substeps of tsch.f90:32-91. */
int melt_tschem(melt_ctx* ctx);
/* One fused substep ns (1..3) of tsch.f90:33-90. */
int melt_substep(melt_ctx* ctx, int32_t ns);

int melt_tagcells(melt_ctx* ctx);     /* tagCells.f90:4-27 (tagType 1)              */
int melt_interp_vof(melt_ctx* ctx);   /* tagCells.f90:85-111                        */
int melt_hdnl1(melt_ctx* ctx);        /* hdnl1.f90:1-80  -> MELT_DQ                 */
int melt_hdnl2(melt_ctx* ctx);        /* hdnl2.f90:1-82  -> MELT_DPH                */
int melt_hdnl3(melt_ctx* ctx);        /* hdnl3.f90:1-91  -> MELT_QCAP               */
int melt_hdnlte(melt_ctx* ctx);       /* hdnlte.f90:1-160 -> MELT_HTEMP             */
int melt_invtr1(melt_ctx* ctx);       /* invtr1.f90:1-77 (incl. solxi/solxj/solxk)  */
int melt_invtr2(melt_ctx* ctx);       /* invtr2.f90:1-71                            */
int melt_invtr3(melt_ctx* ctx);       /* invtr3.f90:1-65                            */
int melt_invtrte(melt_ctx* ctx);      /* invtrte.f90:1-67                           */
int melt_solxi(melt_ctx* ctx, double betadx);              /* solxi.f90:1-52 on MELT_RHS */
int melt_solxj(melt_ctx* ctx, double betadx);              /* solxj.f90:1-40 on MELT_RHS */
int melt_solxk(melt_ctx* ctx, int field, double betadx);   /* solxk.f90:1-60: q += solve(rhs), q device-resident */
/* solxk.f90:1-60 in the reference's own form solxk(q, betadx) (call sites
 * invtr1.f90:68, invtr2.f90:64, invtr3.f90:62, invtrte.f90:63): the z solves run
 * on the device rhs, which keeps the solution, and q += rhs is applied to the
 * caller's host array q(1:ld1,1:ld2,...) described by lay (null: n1m x n2m x n3m). */
int melt_solxk_array(melt_ctx* ctx, double* q, const melt_layout* lay, double betadx);
int melt_divg(melt_ctx* ctx);         /* divg.f90:1-22   -> MELT_DPH                */
int melt_phcalc(melt_ctx* ctx);       /* phcalc.f90:11-97 in place on MELT_DPH      */
int melt_updvp(melt_ctx* ctx);        /* updvp.f90:8-35                             */
int melt_prcalc(melt_ctx* ctx);       /* prcalc.f90:5-35                            */
int melt_find_centroid_indices(melt_ctx* ctx); /* findCentroidIndices.f90:6-53      */
int melt_mlsweight(melt_ctx* ctx);    /* mlsWeight.f90:1-45 (wght1/2/3/temp)        */
int melt_mlsforce(melt_ctx* ctx);     /* mlsForce.f90:1-36 (forc1/2/3/temp), zeroes for_* first */
int melt_velforce(melt_ctx* ctx);     /* velforce.f90:1-59                          */
int melt_tempforce(melt_ctx* ctx);    /* tempforce.f90:1-46                         */
int melt_particle(melt_ctx* ctx);     /* particle.f90:1-229                         */

/* Diagnostics the reference's time loop (gcurv.f90) calls around tschem. */
int melt_cfl(melt_ctx* ctx, double* cflm);     /* cfl.f90:1-30 (max over cells)     */
int melt_divgck(melt_ctx* ctx, double* qmax);  /* divg.f90:26-56                    */
/* No counterpart in the reference.  melt_divgck for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops): no context and
 * no hipMalloc; vx, vy, vz are device arrays of n1*n2*n3 values (double for
 * elem_bytes 8, float for 4), i fastest, periodic; max |div u| goes to the
 * device double *qmax, enqueued on stream (a hipStream_t, null for the null
 * stream), kernels only.  Its scratch is a fixed 8 KiB device array per GPU
 * held until the process ends: calls that overlap on one GPU race.  Limits:
 * n1, n2, n3 >= 2, n1 a multiple of 16 / elem_bytes, n1*n2*n3 < 2^31
 * (MELT_ERR_ARG otherwise). */
int melt_divgck_noalloc(const void* vx, const void* vy, const void* vz, int32_t n1, int32_t n2, int32_t n3, double xlen,
                        double ylen, double zlen, int32_t elem_bytes, void* stream, double* qmax);
int melt_inqpr(melt_ctx* ctx);                 /* inqpr.f90:3-44                    */
/* get_dt (cfl.f90:33-77): *dt_out = dt * cflfix / max(max over cells of
 * |u| dt dx1 + |v| dt dx2 + |w| dt dx3, 6 dt dx1^2 / Re) at the context's dt
 * (melt_set_dt); the caller sets the new dt, as gcurv.f90:151-157 does. */
int melt_get_dt(melt_ctx* ctx, double cflfix, double* dt_out);
/* CalcDissipation (diss.f90:11-201): re-tags a moving body's cells, then the
 * fluid-cell (VOFp == 1) averages and the statistics diss.txt holds:
 * out[11] = {diss_volAvg, kenerg, urms, eta, lambda_t, Re_L, re_lam, L_int,
 *            chi_volAvg, kmax_eta, sumVOF}. */
int melt_dissipation(melt_ctx* ctx, double* out);

/* HIT forcing of the time loop (hit.f90), outside tschem as in gcurv.f90:163-170.
 * melt_hit_init = InitRandomForce (hit.f90:128-232) with HITForcing.in's TL,
 * EPSSTAR and KF/KMIN (nmodes = 2 floor(KF/KMIN) + 1 <= 9); melt_hit_force =
 * CalcHITRandomForce (hit.f90:1-126) at the context's dt: one Ornstein-Uhlenbeck
 * step of bhat for the modes with 0 < |k| < kf, the solenoidal projection and
 * forcx/forcy/forcz on the device (kernels only: capturable).  The Gaussian
 * numbers are Box-Muller pairs from SplitMix64 at a counter (seed, step) in the
 * reference's draw order, in place of gfortran's random_number stream.
 * melt_hit_state gets (set = 0) or sets (set = 1) bhat(3,nm,nm,nm) as (re,im)
 * pairs and the number of steps taken (WriteRandForcCoef / hdf_read_bhat,
 * hit.f90:267-356, plus the generator position).  melt_abc_force =
 * CalcABC_HITForce (hit.f90:234-265). */
int melt_hit_init(melt_ctx* ctx, double tl, double epsstar, double kf_on_kmin, uint64_t seed);
int melt_hit_force(melt_ctx* ctx);
int melt_hit_state(melt_ctx* ctx, double* bhat, uint64_t* steps, int32_t set);
int melt_hit_nmodes(melt_ctx* ctx, int32_t* nmodes, int32_t* nforced);
int melt_abc_force(melt_ctx* ctx, double c_hit);

/* Snapshot and resume (the continuation files of mpi_write_continua /
 * mpi_read_continua, mpi_routines.f90:1223-1631, continua_particle /
 * import_particles, create_geo.f90:135-409, and the bhat file, hit.f90:267-356)
 * as one versioned host blob: the fields that carry state between steps (at
 * the context's precision), the body scalars, every Lagrangian array, the
 * immersed-boundary windows, dt and the HIT forcing state with its generator
 * position.  melt_snapshot_load needs a context with the same configuration
 * and mesh (melt_set_geometry with the same connectivity); the resumed context
 * continues bit for bit.  Call between steps. */
int melt_snapshot_size(melt_ctx* ctx, uint64_t* bytes);
int melt_snapshot_save(melt_ctx* ctx, void* buf, uint64_t bytes);
int melt_snapshot_load(melt_ctx* ctx, const void* buf, uint64_t bytes);

/* Phase timing with HIP events around the four TIMED buckets of tsch.f90:37-84
 * (tagging, Eulerian solve, MLS/particle, pressure) and around each fused-path
 * kernel.  melt_timing_read returns the accumulated milliseconds per kernel slot
 * and the number of samples; names from melt_timing_name. */
int melt_timing_enable(melt_ctx* ctx, int32_t on);
/* Time only the slots in slot_mask (bit s = slot s, s < melt_timing_count());
 * 0 turns timing off.  Each timed slot adds two event records to the stream,
 * so a caller that wants one kernel's duration over a timed loop selects that
 * slot alone.  MELT_ERR_ARG for bits beyond the last slot. */
int melt_timing_select(melt_ctx* ctx, uint32_t slot_mask);
int melt_timing_reset(melt_ctx* ctx);
int melt_timing_count(void);
const char* melt_timing_name(int32_t slot);
int melt_timing_read(melt_ctx* ctx, double* ms, int64_t* samples);

/* Precision form (melt_config.precision = 1; no counterpart in the reference,
 * which is FP64 only).  Eulerian fields are stored as FP32 and the stencils,
 * FFTs and MLS spreading run in FP32; the periodic line reductions, the MLS
 * moment inversions, the Lagrangian / Newton-Euler / Stefan updates and the
 * Poisson residual run in FP64.  After each projection the FP64 divergence of
 * the projected velocity is compared with that of the provisional one; while
 * it exceeds rtol times it (at most max_iter times) a correction Poisson solve
 * and projection are applied.  Off by default (max_iter 0: the measured ratio
 * is about 1e-6, far below rtol); melt_set_refinement(ctx, 1, 1e-4) turns it
 * on.  Host arrays
 * of upload/download stay double in both forms. */
int melt_set_refinement(melt_ctx* ctx, int32_t max_iter, double rtol);

/* Graph replay (no counterpart in the reference, which has no device): with
 * on = 1, melt_tschem captures one RK3 step into a HIP graph on first use and
 * replays it afterwards; melt_set_dt with a new dt and melt_set_refinement
 * drop the graph (their values are launch arguments).  Not used while timing
 * is enabled or before the first tagging.  melt_stream returns the context's
 * hipStream_t, so that a caller can overlap several contexts. */
int melt_graph_enable(melt_ctx* ctx, int32_t on);
/* Launch-shape override used by the autotuner (melt_ibm_amd/tune.py): key 0 =
 * x-lines per block of the fused explicit/x-solve pass, 1 = its threads per
 * block (either one selects the line-tile form of that pass instead of the
 * z-march), 2 = lines per block of the y/z line solves, 3 = their threads per
 * block, 4 = planes per block of the z-marching explicit pass, 5 = j-rows per
 * block (8 or 16) of the pass that does the y and z solves together, 1 there
 * = two separate passes, 6 = streams: 1 keeps all of the context's work on its
 * one stream (for many contexts sharing a GPU; default: the Lagrangian update,
 * probes and next tagging run on a second stream), 7 = where the geometry
 * stage of the gradient probes runs: 1 after the previous substep's Lagrangian
 * update (beside the pressure solve), 2 at the start of its own substep
 * (beside the explicit pass), 8 = faces a ray-tagging pencil may hold
 * (BUFF_SIZE of pencil_rayTag.f90:84, n1m there; more faces raise
 * MELT_ERR_PENCIL_OVERFLOW as the reference aborts), 9 = 1 keeps one cell per
 * thread in the z-marching explicit pass of the precision form (default: two
 * x-neighbour cells per thread there); value 0 = built-in choice.
 * Results do not depend on it (key 8: as long as no pencil overflows). */
int melt_set_tuning(melt_ctx* ctx, int32_t key, int32_t value);
void* melt_stream(melt_ctx* ctx);
int melt_get_refinement_stats(melt_ctx* ctx, int64_t* checks, int64_t* refined, double* last_ratio);

#ifdef __cplusplus
}
#endif
#endif /* MELT_IBM_H */
