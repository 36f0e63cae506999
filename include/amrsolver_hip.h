/*
 * amrsolver_hip: MI355X (gfx950) implementation of amrsolver's per-box
 * right-hand side and SSPRK stage update, behind a C ABI.
 *
 * Every entry point returns AMR_OK (0) or a negative status; amr_last_error()
 * gives the message of the last failing call on the calling thread.  All
 * device work is enqueued on the context's stream and is asynchronous unless
 * stated otherwise.  A context (and the box tables made from it) must not be
 * used from two host threads at once.
 *
 * Array layout ("FAB", AMReX FArrayBox/Array4 convention): Fortran order, i
 * fastest, then j, then k, component outermost.  An amr_fab_t describes one
 * box's array: `data` points at element (lo-ng, lo-ng, lo-ng, 0), strides are
 * in elements.  Several boxes may share one allocation (a level-wide array)
 * by pointing into it with the level's strides; their ghost cells are then
 * their neighbours' valid cells.
 *
 * Conserved components: 0 rho, 1 rho*u, 2 rho*v, 3 rho*w, 4 rho*E
 * (exm/Sods/prob.h:16-21).  Ghost width NGHOST = 3 (exm/Sods/prob.h:39).
 */
#ifndef AMRSOLVER_HIP_H_
#define AMRSOLVER_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AMR_OK 0
#define AMR_ERR_ARG (-1)      /* invalid argument / unsupported combination */
#define AMR_ERR_HIP (-2)      /* HIP runtime error                         */
#define AMR_ERR_STATE (-3)    /* non-physical state detected (rho<=0, NaN)  */

#define AMR_NCONS 5
#define AMR_NGHOST 3

/* Closure object: the POD form of PROB::ProbClosures =
 * closures_derived_t<Visc,Cond,Thermo> (src/cls/Closures.h:172-182) with
 * calorifically_perfect_gas_t (Closures.h:72-82), visc_const_t/visc_suth_t
 * (:103-120), cond_const_t/cond_suth_t (:130-166).  cond_law 2 is the
 * constant-Prandtl conductivity k = mu*cp/Pr. */
typedef struct {
  double gamma, mw, Ru, cv, cp, Rspec, Cshock, Cdamp;
  int32_t visc_law; /* 0 constant, 1 Sutherland */
  double visc_ref, Tvisc_ref;
  int32_t cond_law; /* 0 constant, 1 Sutherland, 2 mu*cp/Pr */
  double cond_ref, Tcond_ref, Pr;
} amr_closures_t;

/* The cns.* scheme switches read by CNS::read_params (src/CNS.cpp:85-153). */
typedef struct {
  int32_t rhs_euler, rhs_visc, rhs_source;
  int32_t flux_euler; /* 0 MUSCL+Riemann, 1 KEEP, 2 WENO5-JS with LF split */
  int32_t order_keep; /* 2, 4, 6 */
  int32_t art_diss;   /* 1: JST dissipation (CNS_advance.cpp:235-265) */
  int32_t order_rk, stages_rk;
} amr_scheme_t;

/* Level geometry (AMReX Geometry + the cns.lo_bc/hi_bc codes 0..6;
 * src/set/CNS_setup.cpp:18-31).  Code 6 is the y wall of exm/Channel: its
 * ghost cells come from the problem's bcnormal (exm/Channel/prob.h:157-187,
 * an isothermal no-slip wall; see amr_problem_t) and the KEEP and viscous
 * fluxes next to it take the reference's wall forms. */
typedef struct {
  int32_t dom_lo[3], dom_hi[3];
  int32_t periodic[3];
  int32_t lo_bc[3], hi_bc[3];
  double dx[3];
  double prob_lo[3];
  double prob_hi[3]; /* Geometry::ProbHi (read by the wall boundary) */
} amr_geom_t;

/* Problem parameters the path reads (the ProbParm fields of
 * exm/Channel/prob.h:71-98): the wall temperature of bc 6 walls and the x
 * body force that the channel's user_source adds (rho*fx to rho*u, (rho*u)*fx
 * to rho*E; exm/Channel/prob.h:189-193). */
typedef struct {
  double wall_T;       /* ProbParm::Tw */
  double body_force_x; /* ProbParm::fx; 0: no body force */
} amr_problem_t;

/* One box's array (FArrayBox / Array4). */
typedef struct {
  int32_t lo[3], hi[3]; /* valid box, inclusive */
  int32_t ng;           /* ghost cells the array holds around the valid box */
  int32_t pad_;
  double* data;         /* device pointer to element (lo-ng, lo-ng, lo-ng, comp 0) */
  int64_t jstride, kstride, nstride; /* element strides */
} amr_fab_t;

typedef struct amr_ctx_s* amr_ctx_t;
typedef struct amr_boxtab_s* amr_boxtab_t;

/* Last error message of the calling thread ("" if none). */
const char* amr_last_error(void);

/* Device, stream and tuning table.  `stream` is a hipStream_t (NULL: the
 * context creates its own).  Replaces the AMReX runtime state the CNS level
 * relies on (amrex::Initialize, The_Arena, Gpu::gpuStream). */
int amr_ctx_create(int device, void* stream, amr_ctx_t* out);
/* Releases the caller's handle; the context lives on until its last box
 * table is destroyed (either order is safe). */
int amr_ctx_destroy(amr_ctx_t ctx);
int amr_ctx_synchronize(amr_ctx_t ctx);

/* Box table of one level: the BoxArray plus the per-level workspace that
 * CNS::post_regrid allocates (src/CNS.cpp:432-467).  `boxes` holds
 * lo[3],hi[3] per box.  Rebuild after a regrid. */
int amr_boxtab_create(amr_ctx_t ctx, const amr_geom_t* geom, int nbox, const int32_t* boxes,
                      amr_boxtab_t* out);
int amr_boxtab_destroy(amr_boxtab_t bt);

/* Ghost fill of one level from its own valid data (single-level FillPatch:
 * FillBoundary + periodic images + foextrap/reflect physical BCs,
 * src/set/bcs.cpp:95-106, src/set/CNS_setup.cpp:18-31).  fabs[nbox] need
 * ng >= 3. */
int amr_fill_boundary(amr_boxtab_t bt, const amr_fab_t* fabs);

/* Ghost fill of a fine AMR level: FillPatch of a level above the coarsest,
 * which AMReX runs as FillPatchTwoLevels with the interpolater CNS
 * registers for State_Type, lincc_interp = CellConservativeLinear with
 * linear limiting (src/set/CNS_setup.cpp:112-114; FillPatch calls
 * src/tim/CNS_advance.cpp:54-161).  In order: same-level copies and periodic
 * images; every ghost cell inside the domain that no fine box covers from
 * the coarse level's data at the fill time, a_old*crse_old + b_new*crse_new
 * (crse_old alone when b_new == 0, crse_new alone when a_old == 0; AMReX's
 * weights are a = (t1 - t)/(t1 - t0), b = (t - t0)/(t1 - t0), or one state
 * alone within 1e-3 (t1 - t0) of its time), by conservative linear
 * interpolation; physical BCs.  Only the `ngrow` (1..3) ghost layers next to
 * each box are filled, as FillPatch's nghost argument does (the fused stage
 * kernel of amr_rk_stage reads 2; the other schemes need 3).  `crse` is the
 * coarse level's box table,
 * `fine` this level's; fine_state needs ng >= 3, the coarse states ng >= 1
 * with their ghost cells filled (amr_fill_boundary or amr_fill_two_levels of
 * the coarse level).  ratio 2 or 4.  The coarse stencil (coarse cell and its
 * six neighbours) of every interpolated cell must lie in the coarse level or
 * across a foextrap/reflect boundary; otherwise, or next to ext_dir (bc code
 * 1 or 6) boundaries, AMR_ERR_ARG. */
int amr_fill_two_levels(amr_boxtab_t fine, amr_boxtab_t crse, int ratio, double a_old,
                        const amr_fab_t* crse_old, double b_new, const amr_fab_t* crse_new, int ngrow,
                        const amr_fab_t* fine_state);

/* amr_fill_two_levels without the same-level copies (the coarse-fine
 * interpolation and physical BCs only), for a stage input whose stage's fused
 * kernel reads the cells other boxes hold straight from their arrays
 * (amr_step_plan on that input: info[3] == 3; src/tim/CNS_advance.cpp:54-161
 * fills those ghost cells by copies).  Other stage inputs need
 * amr_fill_two_levels.  On a level with no owner map for such reads it is
 * amr_fill_two_levels.  No reference counterpart. */
int amr_fill_two_levels_cf(amr_boxtab_t fine, amr_boxtab_t crse, int ratio, double a_old,
                           const amr_fab_t* crse_old, double b_new, const amr_fab_t* crse_new, int ngrow,
                           const amr_fab_t* fine_state);

/* CNS::avgDown (src/CNS.cpp:513-530), i.e. amrex::average_down: every
 * coarse cell under a fine box becomes the mean of its ratio^3 children
 * (summed x fastest, times 1/ratio^3); other coarse cells are untouched.
 * Fine boxes must be coarsenable by the ratio (2 or 4). */
int amr_average_down(amr_boxtab_t fine, amr_boxtab_t crse, int ratio, const amr_fab_t* fine_state,
                     const amr_fab_t* crse_state);

/* Error estimation (CNS::errorEst, src/CNS.cpp:469-507: the problem's
 * user_tagging on every valid cell of every box of a FillPatch'ed state).
 * kind AMR_TAG_DENGRAD: exm/Sods/prob.h:187-197, |rho(i+1) - rho(i-1)| /
 * rho(i) > threshold tags cells i .. i+3 (the reference's dengrad_threshold
 * is 0.5; tags past the domain's hi end are dropped, where the reference
 * writes past its TagBox).  AMR_TAG_WALLBAND: exm/Channel/prob.h:199-209,
 * cells with y = (j + 1/2) dy below factor*h or above prob_hi[1] - factor*h
 * (the reference: factor 0.5, h = ProbParm::h).  AMR_TAG_VORT: |curl u| >
 * threshold, with u = (rho u)/rho and central differences (the criterion of
 * the TGV workload's refined levels; the reference's TGV tags nothing). */
#define AMR_TAG_DENGRAD 1
#define AMR_TAG_WALLBAND 2
#define AMR_TAG_VORT 3
typedef struct {
  int32_t kind;
  int32_t pad_;
  double threshold; /* DENGRAD, VORT */
  double h;         /* WALLBAND */
  double factor;    /* WALLBAND */
} amr_tagger_t;

/* Tags of one level into `tags`, a device byte array over the level's
 * domain (dom_lo .. dom_hi, x fastest): 1 tagged, 0 not (the array is
 * cleared first).  `state` needs ng >= 1 with its ghost cells filled
 * (amr_fill_boundary / amr_fill_two_levels, as errorEst's FillPatch). */
int amr_tag_cells(amr_boxtab_t bt, const amr_tagger_t* tagger, const amr_fab_t* state, uint8_t* tags);

/* The tag map reduced to blocks of cb^3 cells (the clustering unit, the
 * blocking factor over the refinement ratio in coarse cells): out[b] = 1
 * when any cell of block b is tagged; out is a device byte array of
 * (n0/cb)*(n1/cb)*(n2/cb) blocks, x fastest; the domain extents must be
 * multiples of cb. */
int amr_tag_blocks(amr_boxtab_t bt, const uint8_t* tags, int cb, uint8_t* out);

/* CNS::init(AmrLevel& old) (src/CNS.cpp:155-167): the state of a regridded
 * level, FillPatch(old, S_new, 0 ghost cells, cur_time).  Every valid cell of
 * `new_state` (boxes of `fine_new`) takes the value of the old level's box
 * that holds it (`fine_old`, `old_state`); the others are interpolated from
 * the coarse level at the fill time as amr_fill_two_levels does (weights
 * a_old, b_new; conservative linear interpolation).  fine_old NULL: every
 * cell from the coarse level (CNS::init(), FillCoarsePatch, src/CNS.cpp:169-
 * 179).  The coarse states need ng >= 1 with ghost cells filled and every
 * interpolated cell's coarse stencil in the coarse level; ghost cells of
 * new_state are not written.  No reference counterpart beyond the calls
 * cited: Amr::regrid drives it. */
int amr_regrid_fill(amr_boxtab_t fine_new, amr_boxtab_t fine_old, const amr_fab_t* old_state, amr_boxtab_t crse,
                    int ratio, double a_old, const amr_fab_t* crse_old, double b_new, const amr_fab_t* crse_new,
                    const amr_fab_t* new_state);

/* The numerical face fluxes numflxmf[3] that CNS::compute_rhs assembles
 * (src/tim/CNS_advance.cpp:218-346): per direction d the scheme's Euler flux
 * (CentralKEEP::FluxKEEP with Flux_2nd_Order_KEEP next to bc 6 walls,
 * HiRes::FluxWENO or Riemann::Flux; with JSTflux when art_diss = 1)
 * plus ViscousNumericalFluxes (with ViscousWallFluxes at
 * walls), at every face of every box normal to d.  flux_d[b] describes box
 * b's face box: lo = the box's lo, hi = its hi + 1 in direction d (as the
 * Array4 of AMReX's face-centred numflxmf); ghost layers allowed, none
 * written.  `state` needs its ghost cells filled (ng >= 3).  The assembly
 * dSdt = dxinv[0]*(fx(i) - fx(i+1)) + dxinv[1]*(fy(j) - fy(j+1)) +
 * dxinv[2]*(fz(k) - fz(k+1)) of these faces is amr_rhs's result before the
 * source, bit for bit.  For a caller that assembles itself or feeds flux
 * registers (the reference's commented-out refluxing,
 * src/tim/CNS_advance.cpp:356-371). */
int amr_face_fluxes(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch, const amr_fab_t* state,
                    const amr_fab_t* flux_x, const amr_fab_t* flux_y, const amr_fab_t* flux_z);

/* Plot variables of one level (the plotMF of CNS::writePlotFile,
 * src/CNS.cpp:644-874): into `out` (valid cells), the five state components
 * when state_vars != 0, then the derived variables whose bits are set in
 * derive_mask, in the order src/set/CNS_setup.cpp:156-175 registers them:
 * bit 0 pressure (derpres), 1 temperature (dertemp), 2..4 x/y/z velocity
 * (dervel), computed as src/cls/CNS_K.h:76-129 computes them.  `out` holds
 * 5*(state_vars != 0) + popcount(derive_mask) components and must not alias
 * `state`. */
int amr_plot_vars(amr_boxtab_t bt, const amr_closures_t* cls, const amr_fab_t* state, int state_vars,
                  int derive_mask, const amr_fab_t* out);

/* AMReX plotfile of levels 0..nlev-1 (CNS::writePlotFile with
 * Amr::writePlotFile's layout, src/CNS.cpp:644-874): `dir`/Header
 * (HyperCLaw-V1.1: variable names, time, finest level, problem extent,
 * ratios, domains, level steps, cell sizes, then per level its boxes in
 * physical coordinates), and per level `dir`/Level_l/Cell_H (VisMF header:
 * boxes, FAB offsets, per-FAB minima and maxima) and one data file
 * Level_l/Cell_D_00000 (per box a FAB header line and the valid cells as
 * little-endian doubles, component-major, x fastest).  plot[l] are level
 * l's plot variables (amr_plot_vars), ncomp of them named `names`;
 * level_steps[l] = steps taken by level l; ref_ratio[l] between levels l and
 * l+1.  Synchronous; allocates a device staging buffer per level. */
int amr_write_plotfile(const char* dir, int nlev, const amr_boxtab_t* levels, const amr_fab_t* const* plot,
                       int ncomp, const char* const* names, double time, const int32_t* level_steps,
                       const int32_t* ref_ratio);

/* dSdt = L(state) for every box: CNS::compute_rhs (src/tim/CNS_advance.cpp:186-427)
 * with the per-scheme Flux functions (Riemann::Flux, Riemann.h:339;
 * HiRes::FluxWENO, High_resolution.h:279; CentralKEEP::FluxKEEP,
 * CentralKEEP.h:189).  `state` must be ghost-filled (ng >= 3); dsdt needs
 * ng >= 0.  This is the order_rk = 0 unit. */
int amr_rhs(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch,
            const amr_fab_t* state, const amr_fab_t* dsdt);

/* One fused RHS + stage combination (the compute_rhs call plus the
 * MultiFab::LinComb/Saxpy/Xpay that follow it in CNS::advance,
 * src/tim/CNS_advance.cpp:51-181):
 *   mode 0: out = a*s1 + b*(c*sin + d*L(sin))
 *   mode 1: out = (a*s1 + c*sin) + d*L(sin)
 *   mode 2: out = L(sin)
 * `sin` must be ghost-filled.  `out` may alias `s1` (same cell is read then
 * written by one thread) but must not alias `sin`.  s1 is not read when
 * a == 0.  amr_stage_coeffs() gives (mode,a,b,c,d) for each stage of the
 * reference's schemes. */
int amr_rk_stage(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch, int mode,
                 double a, double b, double c, double d, const amr_fab_t* s1, const amr_fab_t* sin,
                 const amr_fab_t* out);

/* Stage count and coefficients of stage `stage` (0-based) of the scheme
 * (order_rk, stages_rk) for step size dt, in the form amr_rk_stage takes. */
int amr_stage_coeffs(int order_rk, int stages_rk, double dt, int stage, int* nstages, int* mode,
                     double* a, double* b, double* c, double* d);

/* One time step of the level: CNS::advance (src/tim/CNS_advance.cpp:14-183)
 * = per stage FillPatch + compute_rhs + stage update.  `state` (ng >= 3)
 * holds U^n on entry and U^{n+1} on exit (valid cells; ghost cells are not
 * guaranteed to be filled on exit: a periodic level held as one array is
 * advanced without ghost fills, its halos read from periodic images).  Uses
 * two ghosted workspace arrays owned by the box table. */
int amr_advance(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch,
                const amr_fab_t* state, double time, double dt);

/* nsteps steps captured once into a hipGraph and replayed (fixed dt).
 * Capture happens on the first call for a given (state, scheme, dt,
 * nsteps); later calls replay. */
int amr_advance_graph(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch,
                      const amr_fab_t* state, double dt, int nsteps);

/* ---- Device-side time-step control (CNS::computeInitialDt /
 * CNS::computeNewDt, src/CNS.cpp:282-377, with CNS::estTimeStep,
 * src/CNS.cpp:380-395).  Every problem of a box table -- the level, or each
 * problem of a batch -- keeps a clock in device memory: its time, its last
 * time step and its step count. */

/* Sets every problem's time; the next step of each takes the CFL estimate
 * itself, as computeInitialDt does.  Synchronous. */
int amr_clock_set(amr_boxtab_t bt, double time);

/* The clocks (nprob entries: 1 for a level, one per problem of a batch;
 * any output may be NULL).  AMR_ERR_STATE when a problem met a non-physical
 * state (CFL estimate not positive or NaN); its steps stopped there.
 * Synchronous. */
int amr_clock_get(amr_boxtab_t bt, double* time, double* dt, int32_t* steps);

/* nsteps steps under device-side time-step control, with no host
 * synchronisation: per step and problem the estimate cfl * min cns_estdt,
 * the step (the estimate on the first step after amr_clock_set, later
 * min(estimate, 1.1 * previous step), cut to stop_time - time when time +
 * step > stop_time - 0.001 step; stop_time < 0: no stop time), then
 * CNS::advance with that step.  A problem within 1e-14 of stop_time takes
 * no further steps (its arrays are left untouched). */
int amr_advance_cfl(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch, const amr_fab_t* state,
                    double cfl, double stop_time, int nsteps);

/* amr_advance_cfl captured into a hipGraph on the first call for a given
 * (state, closures, scheme, cfl, stop_time, nsteps) and replayed after. */
int amr_advance_cfl_graph(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch,
                          const amr_fab_t* state, double cfl, double stop_time, int nsteps);

/* Time steps of a subcycled hierarchy (computeNewDt's multi-level rule,
 * src/CNS.cpp:319-377; computeInitialDt, :283-316, when first != 0),
 * evaluated on the device: per level l the estimate cfl * min cns_estdt of
 * states[l], limited by 1.1 * dt_level[l] (not when first); dt0 = min over
 * levels of nfactor_l * that, nfactor_l = n_cycle[0] * ... * n_cycle[l];
 * cut to stop_time - cur_time as above (not when first); dt_level[l] =
 * dt0 / nfactor_l on exit.  1 <= nlev <= 8, levels on one device.
 * Synchronous. */
int amr_new_dt_levels(int nlev, const amr_boxtab_t* levels, const amr_closures_t* cls,
                      const amr_fab_t* const* states, const int32_t* n_cycle, double cfl, double cur_time,
                      double stop_time, int first, double* dt_level);

/* CFL time step: cfl * min over valid cells of cns_estdt (src/cls/CNS_K.h:12-47,
 * CNS::estTimeStep src/CNS.cpp:380-395).  Synchronous. */
int amr_est_dt(amr_boxtab_t bt, const amr_closures_t* cls, const amr_fab_t* state, double cfl,
               double* dt_out);

/* amr_est_dt without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (tight stepping loops, code
 * that must not call hipMalloc).  The box descriptors (at most 8192 boxes,
 * else AMR_ERR_ARG) and the per-block partial minima live in one fixed
 * device array of the library, about 330 KB per GPU, held until the process
 * ends, instead of memory the library allocates; the call makes no
 * hipMalloc/hipFree.  All calls on one GPU share that array: calls that
 * overlap on one GPU (two host threads, or contexts with different streams)
 * race.  Same result and errors as amr_est_dt.  Synchronous. */
int amr_est_dt_noalloc(amr_boxtab_t bt, const amr_closures_t* cls, const amr_fab_t* state, double cfl,
                       double* dt_out);

/* Batch of `nprob` independent problems advanced together: problem b is
 * the single box geoms[b].dom_lo..dom_hi with its own cell size, periodicity
 * and boundary codes.  Every entry point that takes a box table accepts a
 * batch (one amr_fab_t per problem).  Replaces running many CNS problems
 * one after another (no reference counterpart for the batching itself). */
int amr_batch_create(amr_ctx_t ctx, int nprob, const amr_geom_t* geoms, amr_boxtab_t* out);

/* One CNS::advance step of every problem with its own time step dt[b];
 * problems with dt[b] == 0 are left untouched.  All problems share one
 * launch per kernel. */
int amr_batch_advance(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch,
                      const amr_fab_t* state, const double* dt);

/* Per-problem CFL time step cfl * min cns_estdt (src/cls/CNS_K.h:12-47) into
 * dt_out[nprob].  Synchronous. */
int amr_batch_est_dt(amr_boxtab_t bt, const amr_closures_t* cls, const amr_fab_t* state, double cfl,
                     double* dt_out);

/* Tiling of the fused stage kernel for a scheme on every box table of this
 * context: (tx,ty,tz) = (0,0,0) automatic (each table's autotuned choice,
 * else a default), tx = -1 forces the
 * pass-structured kernels, otherwise one of the compiled shapes
 * (16x16, 32x8, 8x8 columns; tz planes per workgroup).  No reference
 * counterpart: AMReX picks launch shapes itself (ParallelFor). */
int amr_set_tiling(amr_ctx_t ctx, const amr_scheme_t* sch, int tx, int ty, int tz);

/* As amr_set_tiling, choosing also the form of the fused kernel: 0 the
 * first, 1 the self-loading one (16x16 tiles of a viscous KEEP 2/4 scheme:
 * each thread converts the cells of the next plane it reads itself, which
 * saves one workgroup barrier per plane).  The forms give the same bits.
 * No reference counterpart. */
int amr_set_tiling_form(amr_ctx_t ctx, const amr_scheme_t* sch, int tx, int ty, int tz, int form);

/* As amr_set_tiling with tx > 0, but the fused kernel keeps to tiles per box
 * also where the state is one array over the level (no reference
 * counterpart). */
int amr_set_box_tiling(amr_ctx_t ctx, const amr_scheme_t* sch, int tx, int ty, int tz);

/* How a step of `sch` on `state` runs under the tiling in force: info[0..2]
 * the fused kernel's tile (tx,ty,tz; tx = -1: pass-structured kernels),
 * info[3] 0 pass-structured, 1 tiles per box, 2 tiles over the union of the
 * boxes, 3 tiles per box that read the cells other boxes hold (periodic
 * images included) straight from those boxes' arrays, so that a step fills
 * no same-level ghost cells (levels whose boxes keep 3 cells away from every
 * non-periodic domain boundary, KEEP order 2 or 4, the boxes' arrays within
 * 2^31 elements of each other and not one periodic array over the domain),
 * info[4] 1 when the stages after the first read the ghost cells no
 * box covers straight from U^n (else a copy pass per step into the stage
 * workspaces, where there are such cells), info[5] the fused kernel's form
 * (0 first, 1 self-loading; see amr_set_tiling_form).  No reference
 * counterpart. */
int amr_step_plan(amr_boxtab_t bt, const amr_scheme_t* sch, const amr_fab_t* state, int32_t info[6]);

/* Time a whole step (fills and stages, the last one written to a workspace)
 * on `state` (ghost-filled, ng >= 3) under every compiled tiling that fits
 * the level -- tiles per box and, where the state is one array over the
 * level, tiles over the union of the boxes -- and keep the fastest for this
 * scheme on this box table (it clears an explicit amr_set_tiling for the
 * scheme).  Writes the choice to tiling[3] (tx,ty,tz; tx = -1 when only the
 * pass-structured kernels apply).  Synchronous; leaves `state` unchanged. */
int amr_autotune(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch,
                 const amr_fab_t* state, int32_t tiling[3]);

/* Source for schemes with rhs_source = 1: source[nbox] (NCONS components on
 * the valid cells, any ghost width) is added to dSdt after the flux
 * assembly, where CNS::compute_rhs calls the problem's user_source
 * (src/tim/CNS_advance.cpp:397-407; e.g. exm/MMS_euler/prob.h:219-283, whose
 * source depends only on position).  The arrays are read at every call, so
 * the caller may update them in place; NULL unregisters.  Synchronous. */
int amr_set_source(amr_boxtab_t bt, const amr_fab_t* source);

/* A problem's user_source as HIP source text (CNS::compute_rhs calls
 * user_source(i, j, k, statefab, dsdtfab, probparm, closures, dx) on every
 * valid cell after the flux assembly, src/tim/CNS_advance.cpp:397-407; e.g.
 * exm/Channel/prob.h:189-193), compiled at run time for this GPU (hipRTC,
 * -ffp-contract=off).  `code` defines
 *   __device__ void user_source(int i, int j, int k, const StateView& state,
 *                               const RhsView& rhs, const double* pparm,
 *                               const amr_closures_t& cls, const double* dx);
 * state(i,j,k,n) reads the stage input (ghost cells filled, as the
 * reference's statefab), rhs(i,j,k,n) is the cell's dSdt (a double&), pparm
 * a device copy of params[nparams] (the ProbParm values the code reads), dx
 * the level's cell size; URHO, UMX, UMY, UMZ, UET name the components and
 * amr_closures_t is this header's struct.  With a user source set, schemes
 * with rhs_source = 1 add it (in place of amr_set_source's array and the
 * problem's body force) in amr_rhs, amr_rk_stage, amr_advance and the
 * captured steps of amr_advance_graph, on the pass-structured kernels: dSdt
 * is assembled, the source applied, then the stage combined.  code NULL
 * removes it.  A compile error returns AMR_ERR_ARG with the compiler's log in
 * amr_last_error().  No reference counterpart beyond the hook: the
 * reference compiles user_source into its build. */
int amr_set_user_source(amr_boxtab_t bt, const char* code, const double* params, int nparams);

/* Problem parameters of a box table (walls, body force) and the closures
 * the wall fill uses (gamma, Rspec: bcnormal's ProbClosures); problem NULL
 * clears them.  With rhs_source = 1 the body force is added after the
 * registered source array (if any).  Tables with bc 6 walls need this call
 * before any fill or stage.  Synchronous. */
int amr_set_problem(amr_boxtab_t bt, const amr_closures_t* cls, const amr_problem_t* problem);

/* ---- Reverse mode (no reference counterpart: the reference has no adjoint;
 * its lib/autodiff submodule is empty).  Implemented for the KEEP
 * (flux_euler = 1, orders 2/4/6) and viscous right-hand sides of
 * CNS::compute_rhs (src/tim/CNS_advance.cpp:186-427) and every stage scheme
 * of CNS::advance.  Adjoint arrays hold NCONS components on the valid cells
 * (any ghost width).  The ghost-fill transposes add in a fixed order with no
 * atomics, so every run gives the same bits. */

/* grad = (dL/dU)^T lam for the map U (valid cells) -> L(fill(U)) of amr_rhs.
 * `state` must be ghost-filled (ng >= 3). */
int amr_vjp_rhs(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch, const amr_fab_t* state,
                const amr_fab_t* lam, const amr_fab_t* grad);

/* Reverse of amr_rk_stage(mode, a, b, c, d, s1, sin, out) with respect to
 * the valid cells of s1 and sin, given lam_out = adjoint of out:
 *   lam_sin  = b (c lam_out + d J^T lam_out)   (mode 0)
 *            = c lam_out + d J^T lam_out        (mode 1)
 *            = J^T lam_out                      (mode 2)
 *   assigned, or added to lam_sin when `accumulate` is nonzero;
 *   lam_s1 += a lam_out (modes 0 and 1; lam_s1 may be NULL).
 * J = dL/dU at `sin` (ghost-filled, ng >= 3).  lam_sin may alias lam_s1 (the
 * lam_sin update comes first); neither may share storage with lam_out
 * (AMR_ERR_ARG), and no adjoint array may alias `sin`. */
int amr_vjp_stage(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch, int mode, double a, double b,
                  double c, double d, const amr_fab_t* sin, const amr_fab_t* lam_out, const amr_fab_t* lam_s1,
                  const amr_fab_t* lam_sin, int accumulate);

/* Gradient of an objective of the state after `nsteps` CNS::advance steps
 * of fixed dt from state0 (ng >= 3, left unchanged), with respect to the
 * valid cells of state0.  The forward stores a checkpoint every `kcheck`
 * steps and the stage inputs of the last segment; the reverse sweep
 * recomputes the stage inputs of every other segment from its checkpoint.
 *   objective 0: J = sum over valid cells of |m|^2 / (2 rho) dx dy dz
 *   objective 1: J = sum over valid cells and components of weights . U
 * Writes grad (valid cells) and *J_out.  Synchronous.  Workspace: about
 * (nsteps/kcheck + kcheck * stages) state arrays. */
int amr_adjoint_run(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch, const amr_fab_t* state0,
                    double dt, int nsteps, int kcheck, int objective, const amr_fab_t* weights,
                    const amr_fab_t* grad, double* J_out);

/* The rest of CNS::printTotal (src/CNS.cpp:532-621) for `state` (valid
 * cells): primitive minima and maxima in the order rho, u, v, w, T, p
 * (cons2prim of `state`; the reference reads the primitives of the step's
 * last stage input), per direction d wave_max[2d] = max |u_d + c| and
 * wave_max[2d+1] = max |u_d - c| with c = sqrt(gamma R T), and cfl[d] =
 * max(|u_d + c|, |u_d - c|) * dt / dx_d over the cells (pointCFL,
 * src/cls/CNS_K.h:49-59), plus the totals of amr_totals.  Any output may be
 * NULL.  Synchronous. */
int amr_diagnostics(amr_boxtab_t bt, const amr_closures_t* cls, const amr_fab_t* state, double dt, double totals[5],
                    double prim_min[6], double prim_max[6], double wave_max[6], double cfl[3]);

/* AMR_ERR_STATE (and *bad_box = the lowest such box, else -1) when a valid
 * cell of `state` holds a NaN or infinite value, rho <= 0 or a pressure
 * (gamma - 1)(E - |m|^2/(2 rho)) <= 0.  No reference counterpart (the
 * reference clamps silently).  Synchronous. */
int amr_check_state(amr_boxtab_t bt, const amr_closures_t* cls, const amr_fab_t* state, int32_t* bad_box);

/* on != 0: amr_advance, amr_batch_advance and amr_advance_cfl run
 * amr_check_state on the state after their steps (making them
 * synchronous) and return its AMR_ERR_STATE.  Off by default. */
int amr_set_state_check(amr_boxtab_t bt, int on);

/* Conserved totals sum(U_n) * ProbSize (CNS::printTotal, src/CNS.cpp:537-542).
 * Synchronous. */
int amr_totals(amr_boxtab_t bt, const amr_fab_t* state, double out[5]);

#ifdef __cplusplus
}
#endif
#endif
