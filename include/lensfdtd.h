/* lensfdtd.h - C ABI of the MI355X (gfx950) Yee-FDTD step engine.
 *
 * One context = one Yee domain resident in HBM (fields, CPML psi slabs, per-cell
 * update coefficients, source planes, monitor accumulators) plus one HIP stream.
 * The reference drives its solvers per array operation from Python
 * (planar_beads/fused2d.py:105-126 te_update_H/te_update_E/tm_update_H/tm_update_E,
 * planar_beads/boundaries.py:186-268 CPML_TE/CPML_TM.update_H/update_E/terminate,
 * planar_3d/boundaries3d.py:305-362 CPML_3D.step/step_H/step_E); this ABI replaces
 * those calls and the per-step work of the time loops (fdtd2d_core.py:781-873,
 * :1403-1474; fdtd3d_core.py:1105-1197, :1616-1656) at loop level: lf_run advances
 * a chunk of steps described by a table of per-step scalars.
 *
 * Conventions: arrays passed in or out are host pointers to C-order float32 (or the
 * stated type) arrays with the reference's shapes; the library copies them and keeps
 * no pointer after return.  Every function returns 0 on success and a negative code
 * on error; lf_last_error() gives the message (thread-local).  Calls on one context
 * are not thread-safe; distinct contexts may be used from distinct threads.
 */
#ifndef LENSFDTD_H
#define LENSFDTD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lf_ctx lf_ctx;

/* One time step's scalars (120 bytes, packed); filled by the host with the
 * reference's expressions (env: fdtd2d_core.py:787; phasors: :839, flux.py:32-33;
 * drive: fdtd2d_core.py:1426, fdtd3d_core.py:1627; split DFT: fdtd3d_core.py:1638). */
typedef struct lf_step {
    double env, ssin, scos;      /* envelope, env*sin(wt), env*cos(wt)            */
    double pr, pi;               /* exp(-i w t)                                   */
    double hr, hi;               /* exp(-i w (t - dt/2))                          */
    double dr, di;               /* exp(+i w t)                                   */
    float wt32, e32;             /* float32(w t), float32(env)                    */
    float pr32, pi32;            /* complex64(exp(-i w t))                        */
    float hr32, hi32;            /* complex64(exp(-i w (t - dt/2)))               */
    float qr32, qi32;            /* float32(cos w t), -float32(sin w t)           */
    float dr32, di32;            /* complex64(exp(+i w t))                        */
    unsigned int mask;           /* bit m set: monitor m accumulates this step    */
    unsigned int row;            /* row of the plane recorders and TF/SF series   */
} lf_step;

enum { LF_TE2D = 0, LF_TM2D = 1, LF_YEE3D = 2 };
enum { LF_MODE_REF32 = 0, LF_MODE_FP64ACC = 1 };
/* field / coefficient ids */
enum { LF_EX = 0, LF_EY, LF_EZ, LF_HX, LF_HY, LF_HZ,
       LF_CEE, LF_CEH,                 /* 2D TE  Ce_E, Ce_H  (Nx,Ny)              */
       LF_CAX, LF_CAY,                 /* 2D TM  Ca_ex, Ca_ey                     */
       LF_CEX, LF_CEY, LF_CEZ,         /* 3D     ce_x, ce_y, ce_z                 */
       LF_DAX, LF_DAY, LF_DAZ,         /* loss factors D (TM 2D: x,y; 3D: x,y,z)  */
       LF_NARRAYS };
/* monitor kinds / options */
enum { LF_SQ_TE = 0, LF_SQ_TM = 1, LF_SQ_TM_EX2 = 2, LF_SQ_3D_S = 3, LF_SQ_3D_P = 4 };
enum { LF_SRC_EX = 0, LF_SRC_EY = 1, LF_SRC_EZ = 2, LF_SRC_EY_NODE = 3, LF_SRC_EX_NODE = 4 };
enum { LF_PH_P = 0, LF_PH_Q = 1 };            /* complex64 phasor: (pr32,pi32) or (qr32,qi32) */
enum { LF_ACC_C64P = 0, LF_ACC_C64 = 1, LF_ACC_C128 = 2 };
enum { LF_DRIVE_C128 = 0, LF_DRIVE_C64 = 1 };

/* Context lifetime.  dims: nx, ny (2D) or nx, ny, nz (3D; pass nz = 1 for 2D). */
int lf_create(int kind, int mode, int nx, int ny, int nz, lf_ctx **out);
int lf_destroy(lf_ctx *ctx);
const char *lf_last_error(void);
int lf_device_count(int *count);
int lf_device_name(char *buf, int len);

/* Setup.  lf_upload/lf_download move one field or coefficient (reference shape). */
int lf_upload(lf_ctx *ctx, int array_id, const float *host, long long count);
int lf_download(lf_ctx *ctx, int array_id, float *host, long long count);
int lf_set_ch(lf_ctx *ctx, float ch);
/* 2D coefficient array assembled on the device from its compact form (instead of
 * uploading the full array): row j of column i is band[i*rows + j - j0] for
 * j0 <= j < j0 + rows, else table[xclass[i]*ny_c + j]; xclass has the array's x extent,
 * table ncls rows of its y extent.  Replaces the host-side fill of the per-cell arrays by
 * build_materials + _coeffs / _tm_coeffs + _apply_bead_tm + _tm_loss (fdtd2d_core.py:122-349,
 * :445-455) and their transfer; the package computes the compact form with the same
 * expressions on sub-grids (window_coeffs). */
int lf_build_coeff(lf_ctx *ctx, int array_id, const unsigned char *xclass, int ncls,
                   const float *table, int j0, int rows, const float *band);
/* CPML profiles of one axis (boundaries.py:104-151, boundaries3d.py:150-194):
 * e_* on the n integer nodes, h_* on the n-1 half nodes; PEC on the PML ends. */
int lf_set_cpml(lf_ctx *ctx, int axis, const float *e_kappa, const float *e_b, const float *e_c,
                const float *h_kappa, const float *h_b, const float *h_c, int pec_lo, int pec_hi);
int lf_finalize(lf_ctx *ctx);   /* allocate psi slabs (after every lf_set_cpml)  */
/* Huygens source pair on planes ix (ph1) and ix+1 (ph2) of component comp
 * (vcsel.py:151-156); amp/ph arrays have the component's plane shape. */
int lf_set_source(lf_ctx *ctx, int comp, int ix, int plane_count, const float *amp,
                  const float *ph1, const float *ph2);
/* Stitch overlap imposition F[:n] = w d(t) + (1-w) F[:n] (fdtd2d_core.py:1426-1427,
 * fdtd3d_core.py:1626-1630); drives: for each comp in comps, a complex64 array
 * (interleaved re,im) of shape (n, plane) stored back to back. */
int lf_set_drive(lf_ctx *ctx, int ncells, const float *wtap, int ncomp, const int *comps,
                 const float *drives_c64, int math);
/* Secondary absorbers, selected instead of the CPML on some faces (the run then advances
 * one step per launch and applies them after the E update and the source):
 * lf_set_sponge: graded sponge of n cells, factors g[0..n-1] (g[0] outermost) on the faces
 *   of bitmask faces (bit 2*axis + hi), all six fields, 3D only; replaces _sponge_profile /
 *   _apply_sponge (fdtd3d_core.py:397-414) as called at fdtd3d_core.py:1133-1135, 1635-1637.
 * lf_add_mur: first-order Mur on face (axis, hi) of E component comp, coefficient m
 *   (n == 1 scalar, else one per cell of the boundary plane), faces applied in the order
 *   added; flags bit 0: F0_old = current value (the 2D inline form), bit 1: zero the
 *   plane's cells on PEC faces afterwards.  Replaces Mur1_3D.capture/apply
 *   (boundaries3d.py:99-140, called at fdtd3d_core.py:1112-1132) and the 2D inline Mur of
 *   fdtd2d_core.py:805-810, 831-833 (stitch :1422-1433, :1458-1469).  Call after
 *   lf_set_cpml of every axis. */
int lf_set_sponge(lf_ctx *ctx, int n, const float *g, unsigned faces);
int lf_add_mur(lf_ctx *ctx, int comp, int axis, int hi, const float *m, long long n, int flags);

/* Monitors; each returns its id (0..31) in *id. */
int lf_add_box(lf_ctx *ctx, const int *ranges18, int *id);  /* 3 comps x 3 axes x (lo,hi) */
int lf_add_sq(lf_ctx *ctx, int form, int nx_acc, int *id);
int lf_add_dft(lf_ctx *ctx, int src, int x0, int x1, int kplane, int phasor, int acc, int *id);
int lf_add_flux(lf_ctx *ctx, int il, int *id);
int lf_monitor_bytes(lf_ctx *ctx, int id, long long *bytes);
int lf_monitor_read(lf_ctx *ctx, int id, void *dst, long long bytes);
int lf_monitor_write(lf_ctx *ctx, int id, const void *src, long long bytes);
int lf_monitor_zero(lf_ctx *ctx, int id);
/* Steady-state test on the device (fdtd2d_core.py:859-873, fdtd3d_core.py:1161-1175):
 * out[0] = sum |A/da - B/db|^2 and out[1] = sum |A/da|^2 in FP64 over the DFT monitors
 * id_a (A) and id_b (B) of the same shape and format; the reference reads both phasor
 * windows to compute norm(a - b) / norm(a). */
int lf_monitor_change(lf_ctx *ctx, int id_a, int id_b, double da, double db, double *out);

/* ---- Time-domain TF/SF stitch (run_beads_stitched_td, fdtd2d_core.py:1586-1863;
 * run_3d_stitched_td, fdtd3d_core.py:1785-2084).  A context with TF/SF corrections or
 * plane recorders advances one step per launch.
 * lf_add_box_form: lf_add_box with a 2D form, 0 = the node-grid I_E of lf_add_box, 1 (2D
 *   TM) = the node Ey^2 alone; replaces the detector sums of fdtd2d_core.py:1811-1815.
 * lf_add_sx: FP64 sum over the masked steps of the instantaneous net flux in +x through x
 *   plane il (H averaged onto the E plane); replaces _sx_plane_2d (fdtd2d_core.py:489-497)
 *   and _sx_plane_3d (fdtd3d_core.py:762-767) as summed at fdtd2d_core.py:1804-1818 and
 *   fdtd3d_core.py:1990-1999.
 * lf_add_rec: x plane il of component comp copied into row lf_step.row of a rows x plane
 *   buffer on the steps of its mask bit, phase 0 at the start of the step (E(n)), phase 1
 *   after the H update and its TF/SF corrections (H(n+1/2)); replaces the rec_E / rec_H
 *   writes of fdtd2d_core.py:1767-1776 and fdtd3d_core.py:1968-1977.
 * lf_add_tfsf / lf_load_tfsf: F = F +- c inc on x plane il over [j0,j1) x [k0,k1), c = Ch
 *   (coef = -1, H components, after the H update) or coefficient array coef at the cell (E
 *   components, after the E update and source), inc = row lf_step.row of the series
 *   loaded with lf_load_tfsf (nrows x plane floats from row 0); replaces the TF/SF lines
 *   fdtd2d_core.py:1771, 1779, 1784, 1790 and fdtd3d_core.py:1974-1975, 1985-1986.
 * lf_field_maxabs: max |F| of a component; replaces xp.max(xp.abs(...)) of the stop test
 *   (fdtd2d_core.py:1829-1835, fdtd3d_core.py:2000-2004). */
int lf_add_box_form(lf_ctx *ctx, const int *ranges18, int form, int *id);
int lf_add_sx(lf_ctx *ctx, int il, int *id);
int lf_add_rec(lf_ctx *ctx, int comp, int il, int phase, int rows, int *id);
int lf_add_tfsf(lf_ctx *ctx, int comp, int il, int sign, int coef, int j0, int j1, int k0, int k1,
                int rows, int *id);
int lf_load_tfsf(lf_ctx *ctx, int id, const float *rows, int nrows);
int lf_field_maxabs(lf_ctx *ctx, int comp, float *out);

/* Advance nsteps steps; steps[s] holds the scalars of step s. */
int lf_run(lf_ctx *ctx, const lf_step *steps, int nsteps);
int lf_sync(lf_ctx *ctx);
/* Average duration [ms] of the Yee update kernels of the last lf_run, measured with
 * HIP events on the context's stream (0 if timing is off); lf_set_timing(ctx, 1) on. */
int lf_set_timing(lf_ctx *ctx, int on);
/* Kernel variant (no reference counterpart: an engine tuning knob): 0 = single-pass
 * H+E kernels where the kind has one (2D TM, 3D), 1 = separate H and E kernels (default);
 * lx = x planes per wavefront chunk of the single-pass kernels.  Results are bit-identical. */
int lf_set_variant(lf_ctx *ctx, int variant, int lx);
/* Owned y rows per wavefront of the 3D single-pass kernel: 2 or 4 (tuning knob, no
 * reference counterpart). */
int lf_set_tile(lf_ctx *ctx, int ty);
int lf_kernel_ms(lf_ctx *ctx, double *update_ms, double *total_ms, int *nsteps);
/* Temporal blocking (no reference counterpart: an engine tuning knob; the reference
 * advances one step per call, boundaries.py:240-258).  depth = steps advanced per pass
 * (1 = off, 2, 4, 6 or 8) over runs of steps whose monitor mask is 0; lx = x planes per
 * wavefront chunk.  Results are bit-identical to depth 1.  Used by 2D TM. */
int lf_set_temporal(lf_ctx *ctx, int depth, int lx);
/* Depth of the monitored 2D passes (no reference counterpart: an engine tuning knob):
 * runs of `depth` steps (1 = off, 2, 3 or 4) whose monitors are volume phasors (at most
 * two), one intensity sum and up to four flux planes advance in one pass that also sums
 * those monitors (the reference sums them after every step, fdtd2d_core.py:832-845,
 * flux.py:253-291).  Every accumulator receives the same additions in the same order, so
 * results are bit-identical to depth 1.  2D TE and TM; 3D accepts only 1. */
int lf_set_temporal_mon(lf_ctx *ctx, int depth);
/* Blocked passes of the last lf_run with timing on: HIP-event time (ms) of the runs of
 * consecutive plain passes (events at each run's ends, none between its passes), number
 * of those passes, current depth. */
int lf_tb_stats(lf_ctx *ctx, double *ms, int *launches, int *depth);
/* Launch order of the classes of a 2D blocked pass (no reference counterpart: an engine
 * tuning knob): class `cls` (0 slab-free, 1 y-slab rows, 2 x-slab planes, 3 excitation
 * planes) is dispatched first, the others behind it on side streams; 4 launches every
 * class in turn on the context stream without events (small windows, where queuing a
 * pass costs the host more than the device needs for it).  Results are bit-identical for
 * every order.  3D accepts only 0. */
int lf_set_tb_lead(lf_ctx *ctx, int cls);
/* Autotuner: times every candidate temporal-block depth and x chunk (2D: depth 2/4/6/8 x
 * chunk 32/64/128, then the lead class 0..4 at the fastest of those; 3D: depth 2/3 x
 * chunk 32/48/64/96) on this context with `reps` passes of monitor-free copies of `row`,
 * restores every field and psi slab, and keeps the fastest (results are bit-identical
 * for every choice).  No counterpart in the reference, whose step loop has no blocking;
 * it tunes the lf_set_temporal and lf_set_tb_lead settings. */
int lf_autotune(lf_ctx *ctx, const lf_step *row, int reps, int *depth, int *lx, int *lead,
                double *ms_per_step);

/* ---- Batch engine: many independent 2D domains advanced together.  No reference
 * counterpart: the reference runs sweep instances and requests one after another, each
 * through run_beads / run_beads_stitched (run_simulation.py:290-431).  Instances keep
 * their own fields, coefficients, sources and monitors (set up with the calls above);
 * every temporally blocked pass is ONE launch per wavefront class covering all instances
 * that take it, the others step one step per launch.  Results are bit-identical to
 * running each context alone. */
typedef struct lf_batch lf_batch;
int lf_batch_create(lf_batch **out);
/* finalized 2D context, same precision mode for all instances */
int lf_batch_add(lf_batch *b, lf_ctx *ctx);
/* depth 1 (off), 2, 4, 6 or 8 (default 6); lx = x planes per wavefront chunk */
int lf_batch_set_temporal(lf_batch *b, int depth, int lx);
int lf_batch_set_timing(lf_batch *b, int on);
/* steps: shared step table; masks: ninst x nsteps monitor masks or NULL (steps[].mask);
 * nact: steps per instance in this call (<= nsteps) or NULL (all nsteps) */
int lf_batch_run(lf_batch *b, const lf_step *steps, int nsteps, const unsigned int *masks,
                 const int *nact);
/* Resident step table for lf_batch_run_at (uploaded once; instances that started at
 * different times read it at their own offsets). */
int lf_batch_set_table(lf_batch *b, const lf_step *table, int ntable);
/* instance i takes table[offs[i] + s] for s < nact[i] <= nsteps (offs / nact may be NULL:
 * 0 / nsteps); masks: ninst x nsteps indexed by s, or NULL (table masks) */
int lf_batch_run_at(lf_batch *b, const int *offs, int nsteps, const unsigned int *masks,
                    const int *nact);
/* last lf_batch_run with timing on: HIP-event time (ms) and blocked passes */
int lf_batch_stats(lf_batch *b, double *ms, int *passes);
int lf_batch_destroy(lf_batch *b);

/* The same reduction over caller-owned DEVICE arrays a, b of n complex elements (acc:
 * LF_ACC_C128 double2, LF_ACC_C64 / LF_ACC_C64P float2); out: 2 doubles in device
 * memory, written on `stream`.  The partials live in a library buffer per (device,
 * stream). */
int lf_norm_diff(const void *a, const void *b, long long n, int acc, double da, double db,
                 double *out, void *stream);

/* lf_norm_diff that allocates nothing (no hipMalloc / hipFree, no allocating library
 * call): the partials live in a static 1 MiB device array, one per GPU, held until the
 * process ends; n is at most 2^28 elements (larger n: error, use lf_norm_diff).  No
 * counterpart in the reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Calls that overlap on
 * one GPU share the array and race. */
int lf_norm_diff_noalloc(const void *a, const void *b, long long n, int acc, double da, double db,
                         double *out, void *stream);

/* ---- Operator level: single CPML updates on caller-owned DEVICE arrays, in the
 * reference's own C-order shapes (no padding), in place, enqueued on `stream` (a
 * hipStream_t, NULL = the null stream).  The drop-in for a caller that keeps its own
 * time loop over device arrays; lf_run is the fast path.  Per element the float32
 * operations are the reference's, in its order. */
typedef struct lf_op2d {
    /* 1-D CPML profiles on the device: x on the Nx nodes (e) and Nx-1 half nodes (h),
     * y likewise (boundaries.py:104-151) */
    const float *bex, *cex, *kex, *bhx, *chx, *khx;
    const float *bey, *cey, *key, *bhy, *chy, *khy;
    /* the two psi arrays of the update (full-domain, the reference's shapes):
     * te_update_H psi_Hxy, psi_Hyx; te_update_E psi_Ezx, psi_Ezy;
     * tm_update_H psi_Hzx, psi_Hzy; tm_update_E psi_Exy, psi_Eyx */
    float *psi0, *psi1;
} lf_op2d;
/* fused2d.te_update_H (planar_beads/fused2d.py:105-108; boundaries.py:186-195):
 * Ez (Nx,Ny), Hx (Nx,Ny-1), Hy (Nx-1,Ny) */
int lf_op_te_update_H(const lf_op2d *p, int Nx, int Ny, const float *Ez, float *Hx, float *Hy,
                      float ch, void *stream);
/* fused2d.te_update_E (fused2d.py:111-114; boundaries.py:197-205) */
int lf_op_te_update_E(const lf_op2d *p, int Nx, int Ny, float *Ez, const float *Hx, const float *Hy,
                      const float *ce_E, const float *ce_H, void *stream);
/* fused2d.tm_update_H (fused2d.py:117-120; boundaries.py:240-248):
 * Hz (Nx-1,Ny-1), Ex (Nx-1,Ny), Ey (Nx,Ny-1) */
int lf_op_tm_update_H(const lf_op2d *p, int Nx, int Ny, const float *Ex, const float *Ey, float *Hz,
                      float ch, void *stream);
/* fused2d.tm_update_E (fused2d.py:123-126; boundaries.py:250-258) */
int lf_op_tm_update_E(const lf_op2d *p, int Nx, int Ny, float *Ex, float *Ey, const float *Hz,
                      const float *Ca_ex, const float *Ca_ey, void *stream);
/* CPML_TE.terminate / CPML_TM.terminate (boundaries.py:207-215, :260-268): PEC on the
 * ends that carry a PML, pec = {x lo, x hi, y lo, y hi} */
int lf_op_te_terminate(int Nx, int Ny, float *Ez, const int pec[4], void *stream);
int lf_op_tm_terminate(int Nx, int Ny, float *Ex, float *Ey, const int pec[4], void *stream);

/* CPML_3D (planar_3d/boundaries3d.py:262-362): the 12 psi slabs live in the object
 * (device memory, the PML slabs only, as _PsiStrip :201-251 keeps them). */
typedef struct lf_op3d lf_op3d;
int lf_op3d_create(int Nx, int Ny, int Nz, lf_op3d **out);
/* axis 0..2: e_* on the N nodes, h_* on the N-1 half nodes, (kappa, b, c) as
 * _cpml_bc_1d builds them (boundaries3d.py:150-194); PEC on the axis' faces */
int lf_op3d_set_cpml(lf_op3d *op, int axis, const float *e_kappa, const float *e_b, const float *e_c,
                     const float *h_kappa, const float *h_b, const float *h_c, int pec_lo, int pec_hi);
int lf_op3d_finalize(lf_op3d *op);
/* CPML_3D.step_H (boundaries3d.py:311-330): Ex (Nx-1,Ny,Nz), Ey (Nx,Ny-1,Nz),
 * Ez (Nx,Ny,Nz-1), Hx (Nx,Ny-1,Nz-1), Hy (Nx-1,Ny,Nz-1), Hz (Nx-1,Ny-1,Nz) */
int lf_op3d_step_H(lf_op3d *op, const float *Ex, const float *Ey, const float *Ez, float *Hx,
                   float *Hy, float *Hz, float ch, void *stream);
/* CPML_3D.step_E (boundaries3d.py:332-362): ce_* on the E grids; da_* the staggered loss
 * factors (fdtd3d_core.py:606-613) or all three NULL */
int lf_op3d_step_E(lf_op3d *op, float *Ex, float *Ey, float *Ez, const float *Hx, const float *Hy,
                   const float *Hz, const float *ce_x, const float *ce_y, const float *ce_z,
                   const float *da_x, const float *da_y, const float *da_z, void *stream);
int lf_op3d_destroy(lf_op3d *op);

#ifdef __cplusplus
}
#endif
#endif
