// Internal interface between the C-ABI layer (capi.cpp) and the kernel
// translation units.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <array>
#include <map>
#include <vector>

#include "amrsolver_hip.h"
#include "cns_point.h"

namespace amr {

struct BoxDev {
  int lo[3], hi[3];
  double dxi[3];  // inverse cell size of the box's level or problem
  double dx[3];   // cell size (as given; not recomputed from dxi)
};

// Block -> (box, block within the box) map of one launch region, so that
// launches over uneven boxes carry no empty blocks.
struct BlockRef {
  int box, blk;
};

// The level's box list on host and device.
struct BoxTable {
  int nbox = 0;
  std::vector<BoxDev> h_boxes;
  BoxDev* d_boxes = nullptr;
  // region (grow lo/hi per direction) -> device block map and its length
  mutable std::map<std::array<int, 6>, std::pair<BlockRef*, int>> block_maps;
};

// Directional flux passes (WENO or MUSCL without viscous fluxes): brick =
// 64 cells along x from x0 by 4 along y (passes x, y) or z (pass z) from b0,
// at third coordinate c (z, or y for pass z) of box `box`.  wrap > 0: the box
// spans a periodic domain of that many cells along the pass direction, and
// the pass reads periodic images there instead of ghost cells.
struct WenoBrick {
  int box, x0, b0, c, wrap;
};
// Where the x pass of a box reads the cells beyond its x ends (2 per box,
// left then right): kind 0 its ghost cells; 1 box `src` at x + shift (a
// neighbour, or a periodic image); 2 the physical boundary's rule per
// component (type 2 foextrap, 3 reflect even, 4 reflect odd, about the
// boundary cell bnd, from the box's own valid cells).
struct XSide {
  int kind, src, shift, bnd;
  int type[5];
};
struct WenoWork {
  WenoBrick* d[3] = {nullptr, nullptr, nullptr};
  int n[3] = {0, 0, 0};
  // set when every box spans its periodic y and z extents and both x sides
  // have kind 1 or 2: the passes then read no ghost cell (no fill needed)
  const XSide* xside = nullptr;
};
int weno_brick_count(const BoxTable& bt, int d);
void weno_bricks(const BoxTable& bt, int d, const int* wrap, WenoBrick* out);

// Per-box workspace views of the unfused pass structure (device arrays of
// nbox Fab each).
struct UnfusedWork {
  WenoWork weno;            // brick tables of the directional flux passes
  Fab* prims = nullptr;     // NPRIM comps on grow(bx,3)
  Fab* nf[3] = {};          // NCONS comps on [lo, hi+1]^3
  Fab* pv[3] = {};          // NCONS comps on grow(bx,1)   (viscous)
  Fab* pe[3] = {};          // NCONS comps on grow(bx,3)   (WENO point fluxes)
  Fab* lam = nullptr;       // 3 comps on grow(bx,3)       (WENO eigenvalues)
  Fab* slope = nullptr;     // NCONS comps on grow(bx,1)   (MUSCL slopes)
  Fab* jst = nullptr;       // 6 comps on grow(bx,1): JST sensor (0-2), spectral radius (3-5)
};

// Periodic halo wrap for a state that is one array over the whole periodic
// domain: indices outside [lo, lo+len) read their periodic image, so the
// kernel needs no ghost fill.
struct WrapArgs {
  int on;
  int lo[3], len[3];
};

// Level tiling: when every state array of a stage is one array over the
// level (boxes are windows), tiles are laid over the union of the boxes
// instead of box by box, so they cross box boundaries.  [lo, hi] is the
// bounding box of the boxes; bits (null when the bounding box is fully
// covered) marks covered blocks of 2^sh cells a side, x fastest, nb blocks
// per direction.  Only covered cells are written.
struct Cover {
  int on;
  int lo[3], hi[3];
  int nb[3];
  int sh;
  const unsigned* bits;
};

// Owner map of a level for direct neighbour reads by the fused kernel's
// tiles per box: the domain [dlo, dlo+len) in blocks of 2^sh cells a side
// (nb per direction, x fastest), owner[block] = the box holding the block or
// -1; per[d] marks periodic directions (halo coordinates wrap there).  Every
// box boundary lies on a block boundary, and no box lies within NGHOST cells
// of a non-periodic domain boundary (its halo cells are all inside the
// domain or periodic images).
struct NbMap {
  int on;
  int dlo[3], len[3], per[3], nb[3];
  int sh;
  const int* owner;
  int ustride;  // set per launch: every box's stage-input array has the same strides
};

// Ghost-fill tasks.  Copy: dst(i,j,k) = src(i - sh) for (i,j,k) in [lo,hi].
struct CopyTask {
  int dst, src;
  int lo[3], hi[3];
  int sh[3];
  int64_t start;  // exclusive prefix of cell counts
};
// Block map of a task list: workgroup w copies cells [first, first + TASK_BLK)
// of task `task` (cells numbered x fastest), so no lane searches the list.
constexpr int TASK_BLK = 256;
struct TaskBlock {
  int task, first;
};
// Cell (i,j,k) of task region [lo,hi] for lane `tid` of block `tb` (cells
// numbered x fastest); false past the region's last cell.
template <class T>
__device__ __forceinline__ bool task_cell(const T& t, const TaskBlock& tb, int tid, int c[3]) {
  const int nx = t.hi[0] - t.lo[0] + 1, ny = t.hi[1] - t.lo[1] + 1, nz = t.hi[2] - t.lo[2] + 1;
  const int r = tb.first + tid;
  if (r >= nx * ny * nz) return false;
  const int r2 = r / nx;
  c[0] = t.lo[0] + (r - r2 * nx);
  c[1] = t.lo[1] + r2 % ny;
  c[2] = t.lo[2] + r2 / ny;
  return true;
}

// Physical BC: for cells [lo,hi] of box `box` in direction `dir`, side
// `side` (0 lo, 1 hi) with domain boundary index `bnd`; per-component type
// 0 nothing, 2 foextrap, 3 reflect_even, 4 reflect_odd, 5 the isothermal
// no-slip wall of bc code 6 (all components at once; plo/phi are the
// problem's lower and upper bounds in `dir`).
struct BcTask {
  int box, dir, side, bnd;
  int lo[3], hi[3];
  int type[5];
  int64_t start;
  double plo, phi;
};
constexpr int BC_WALL = 5;
// Wall parameters of the fill: wall temperature and the gas constants.
struct WallFill {
  double Tw, gamma, Rspec;
};

// Flux corrections next to bc 6 walls (y): task regions of boxes touching a
// wall, side 0 (lower) or 1 (upper).  KEEP tasks cover the three y-face
// planes next to the wall (the kernel keeps order_keep/2 of them), viscous
// tasks the ghost row of cell fluxes outside the wall.
struct WallTask {
  int box, side;
  int lo[3], hi[3];
};
struct WallArgs {
  const WallTask* keep = nullptr;
  const TaskBlock* keep_blk = nullptr;
  int nkeep_blk = 0;
  const WallTask* visc = nullptr;
  const TaskBlock* visc_blk = nullptr;
  int nvisc_blk = 0;
  int dlo = 0, dhi = 0;  // domain bounds in y
};

inline void keep_coefficients(int order, double co[3]) {
  // src/CNS.cpp:243-251; src/rhs/CentralKEEP.h:196-202
  if (order == 6) {
    co[0] = 6.0 / 4; co[1] = -6.0 / 20; co[2] = 2.0 / 60;
  } else if (order == 4) {
    co[0] = 4.0 / 3; co[1] = -2.0 / 12; co[2] = 0.0;
  } else {
    co[0] = 1.0; co[1] = 0.0; co[2] = 0.0;
  }
}

// Builds the block maps of every region the pass-structured launches use
// (called once when a box table is created; launches only look them up).
hipError_t prepare_block_maps(const BoxTable& bt);
hipError_t launch_unfused_stage(const BoxTable& bt, const UnfusedWork& w, const Closures& c,
                                const amr_scheme_t& s, const StageArgs& sa, const WallArgs& wa, const Fab* s1,
                                const Fab* sin, const Fab* out, hipStream_t st, const Fab* const* faces = nullptr);
// (faces non-null: the face fluxes of every direction into faces[d] over the
// boxes' faces normal to d, instead of their assembly into `out`)
// out = the stage combination (StageArgs) of s1, sin and the dSdt in `rhs`,
// on the valid cells of every box.
hipError_t launch_combine(const BoxTable& bt, const StageArgs& sa, const Fab* s1, const Fab* sin, const Fab* rhs,
                          const Fab* out, hipStream_t st);

// The three directional passes of a stage (flux_euler 2: WENO, 0: MUSCL):
// pass x and y accumulate the flux differences in `acc`, pass z adds its
// own, the source and the stage combination and writes `out`.  With `pv`
// (the cell-centred viscous fluxes per direction on grow(bx,1)) each pass
// adds the face averages of its direction's viscous fluxes to its faces.
hipError_t launch_dir_flux(int flux_euler, const WenoWork& ww, const BoxDev* boxes, const Closures& c,
                           const StageArgs& sa, const Fab* sin, const Fab* acc, const Fab* s1, const Fab* out,
                           hipStream_t st, const Fab* const* pv = nullptr);
// cons2prim of `sin` into w.prims on grow(bx,3) for every box (halo cells
// from their periodic images when wa.on).
hipError_t launch_prims(const BoxTable& bt, const UnfusedWork& w, const Closures& c, const double* box_dt,
                        const Fab* sin, const WrapArgs& wa, hipStream_t st);

// Reverse mode of the KEEP + viscous RHS (adjoint.hip).  Workspace per box:
// e = 16 comps on grow(bx,1) (viscous), g = NCONS comps on grow(bx,3).
struct AdjWork {
  Fab* e = nullptr;
  Fab* g = nullptr;
  // domain-wide primitives and viscous pre-pass arrays of a periodic level
  // held as one array (every box's entry maps the same storage)
  Fab* lq = nullptr;
  Fab* le = nullptr;
};
bool adjoint_supported(const amr_scheme_t& s);
// Transposed ghost fill of the adjoint workspace g (host-side view of
// device task lists): physical-boundary tasks per direction and side, their
// `start` counting lines along the direction; same-level copies in colour
// classes whose source regions do not overlap, `start` counting cells
// within the class.
constexpr int FOLD_MAX_CLASSES = 32;
struct FoldTasks {
  BcTask* bc[3][2];
  int nbc[3][2];
  int64_t nlines[3][2];
  CopyTask* classes[FOLD_MAX_CLASSES];
  int class_tasks[FOLD_MAX_CLASSES];
  int64_t class_cells[FOLD_MAX_CLASSES];
  int nclass;
};
// Reverse of one stage (modes as amr_rk_stage) at the ghost-filled input
// sin: lam_sin (=|+=) stage terms of lam_out through the RHS transpose, the
// transposed ghost fill included; lam_s1 updated with a lam_out (s1_mode 0
// none, 1 +=, 2 =).  lam_sin and lam_s1 must not alias lam_out.
hipError_t launch_adj_stage(const BoxTable& bt, const UnfusedWork& w, const AdjWork& aw, const Closures& c,
                            const amr_scheme_t& s, const StageArgs& sa, const WrapArgs& wa, const Fab* sin, const Fab* lam_out, const Fab* lam_s1, const Fab* lam_sin,
                            bool accumulate, int s1_mode, const FoldTasks& ft, hipStream_t st);
// Objective of the final state and its gradient lam (kind 0 kinetic energy,
// kind 1 linear with weights); result[0] = J.
hipError_t launch_objective(const BoxTable& bt, int kind, const Fab* st, const Fab* wts, const Fab* lam,
                            double* partial, double* result, hipStream_t s);

// Ghost fill over block-mapped task lists (nblk workgroups).
// packed block map (workgroups span consecutive cells of the task list)
hipError_t launch_fill_copy(const CopyTask* tasks, int ntask, const TaskBlock* blocks, int nblk, const Fab* fabs,
                            hipStream_t st);
hipError_t launch_fill_bc(const BcTask* tasks, const TaskBlock* blocks, int nblk, const Fab* fabs,
                          const BoxDev* boxes, const WallFill& wf, hipStream_t st);

// dst[T.dst](c) = src[T.src](c - sh) over the task regions, into one or two
// FAB sets (dst1 may be null): the source is read once for both.
hipError_t launch_copy_regions(const CopyTask* tasks, const TaskBlock* blocks, int nblk, const Fab* src,
                               const Fab* dst0, const Fab* dst1, hipStream_t st);

// Coarse-fine ghost cells of a fine level from the coarse level: per task,
// dst = fine box, src = coarse box, lo/hi = coarse cells (inside box src),
// flo/fhi = the fine cells to fill (images inside the domain), sh = periodic
// shift (fine cell w is stored at w + sh).  Coarse data a*c0 + b*c1 (c0 alone
// when blend is false), ratio r.
struct InterpTask {
  int dst, src;
  int lo[3], hi[3];
  int flo[3], fhi[3];
  int sh[3];
  int64_t start;
};
hipError_t launch_interp_cf(const InterpTask* tasks, const TaskBlock* blocks, int nblk, const Fab* c0, const Fab* c1,
                            bool blend, double a, double b, int r, const Fab* fine, hipStream_t st);
// Coarse cells under fine boxes := mean of their r^3 children: per task,
// dst = coarse box, src = fine box, lo/hi = coarse cells.
hipError_t launch_avg_down(const CopyTask* tasks, const TaskBlock* blocks, int nblk, const Fab* fine,
                           const Fab* crse, int r, hipStream_t st);

// dst = src on the valid cells of every box.
hipError_t launch_copy_valid(const BoxTable& bt, const Fab* src, const Fab* dst, const double* box_dt,
                             hipStream_t st);
// Plot variables (state components, then derived pressure, temperature and
// velocities selected by `mask` bits 0..4) of every box's valid cells.
hipError_t launch_plot_vars(const BoxTable& bt, const Fab* src, const Fab* dst, int state_vars, int mask,
                            const Closures& c, hipStream_t st);
// Valid cells of every box into one buffer (box b at off[b], component-major).
hipError_t launch_pack_valid(const BoxTable& bt, const Fab* src, int ncomp, const int64_t* off, double* out,
                             hipStream_t st);

// Reductions over valid cells.  `partial` holds at least reduce_partials()
// doubles (x5 for totals); the result lands in result[0] (dt) or result[0..4].
int reduce_blocks(const BoxTable& bt);
hipError_t launch_estdt(const BoxTable& bt, const Fab* st, const Closures& c, double* partial,
                        double* result, hipStream_t s);
// per-box minimum of cns_estdt: result[b] for every box (estdt_box_blocks
// partial minima per box in `partial`; result may be null).  With clocks
// (clk; batch: one per box, else one for the table) the boxes of stopped
// problems are skipped (their partials hold the largest double).
struct Clock;
int estdt_box_blocks(const BoxTable& bt);
hipError_t launch_estdt_boxes(const BoxTable& bt, const Fab* st, const Closures& c, double* partial,
                              double* result, hipStream_t s, const Clock* clk = nullptr, bool batch = false,
                              double stop_time = -1.0);
// Device-side time-step control: the clock of each problem of a box table
// (the level, or each problem of a batch).  first = 1: the next step takes
// the CFL estimate itself (computeInitialDt); bad = 1: a non-physical state
// stopped the problem.
struct Clock {
  double t, dt;
  int32_t steps, first, bad, pad_;
};
// A problem that takes no more steps: stopped by a non-physical state, or
// within 1e-14 of stop_time (stop_time >= 0).
__host__ __device__ inline bool clock_done(const Clock& k, double stop) {
  return k.bad != 0 || (stop >= 0.0 && k.t >= stop - 1e-14);
}
// Per box (est: nper partial minima per box of cns_estdt, without the CFL
// factor, as launch_estdt_boxes writes them to `partial`) the time step of
// its problem (batch: box b is problem b; else one problem) under the policy
// of CNS::computeNewDt, written to box_dt; clocks advance.
hipError_t launch_dt_policy(int nbox, bool batch, const double* est, int nper, Clock* clk, double cfl,
                            double stop_time, double* box_dt, hipStream_t s);
// computeNewDt's multi-level rule (src/CNS.cpp:319-377) for up to
// MAX_DT_LEVELS levels: est[l] points at level l's minimum of cns_estdt;
// dt_level (device, nlev entries) holds the previous steps on entry
// (ignored when first != 0) and the new ones on exit.
constexpr int MAX_DT_LEVELS = 8;
struct DtLevels {
  const double* est[MAX_DT_LEVELS];
  int32_t ncycle[MAX_DT_LEVELS];
  int32_t nlev, first;
  double cfl, cur_time, stop_time;
};
hipError_t launch_dt_levels(const DtLevels& a, double* dt_level, hipStream_t s);

// No-allocation CFL reduction (amr_est_dt_noalloc).  Its scratch is one
// file-scope device array of NOALLOC_BYTES per GPU: the descriptors of at
// most NOALLOC_MAX_BOXES boxes, then NOALLOC_MAX_BLOCKS per-block partial
// minima, then the result.  noalloc_scratch_address() gives the current
// device's copy of the array (taken once per device at context creation).
constexpr int NOALLOC_MAX_BOXES = 8192;
constexpr int NOALLOC_MAX_BLOCKS = 1024;  // reduce_blocks() never exceeds it
constexpr size_t NOALLOC_DESC_BYTES = size_t(NOALLOC_MAX_BOXES) * sizeof(Fab);
constexpr size_t NOALLOC_BYTES = NOALLOC_DESC_BYTES + size_t(NOALLOC_MAX_BLOCKS) * sizeof(double) + 256;
hipError_t noalloc_scratch_address(void** ptr);
// k_estdt and k_min_final on the descriptors already copied to the start of
// `scratch`; the minimum (without the CFL factor) lands in the result slot,
// whose address *result_out receives
hipError_t launch_estdt_scratch(const BoxTable& bt, const Closures& c, void* scratch, double** result_out,
                                hipStream_t s);
// printTotal diagnostics: result[0..5] primitive minima, [6..11] maxima,
// [12 + 3d] max |u_d + c|, [13 + 3d] max |u_d - c|, [14 + 3d] CFL number;
// partial holds 21 doubles per reduction block.
hipError_t launch_diag(const BoxTable& bt, const Fab* st, const Closures& c, double dt, double* partial,
                       double* result, hipStream_t s);
// flag[0] (device int, preset to nbox) := lowest box holding a non-physical cell
hipError_t launch_state_check(const BoxTable& bt, const Fab* st, const Closures& c, int* flag, hipStream_t s);

// Cell tagging (amr_tag_cells): a byte per cell of the level's domain
// [dlo, dlo + n), x fastest.
struct TagArgs {
  int kind;
  int dlo[3], n[3];
  double thr, h, factor, prob_hi1;
  double dx[3];
};
// Clears the map, then tags from every box's valid cells (ghost cells of the
// state filled, one layer read).
hipError_t launch_tag_cells(const BoxTable& bt, const Fab* st, const TagArgs& ta, unsigned char* tags,
                            hipStream_t s);
// out[block] = any tag in the block of cb^3 cells (blocks x fastest).
hipError_t launch_tag_blocks(const unsigned char* tags, const TagArgs& ta, int cb, unsigned char* out,
                             hipStream_t s);
hipError_t launch_totals(const BoxTable& bt, const Fab* st, double* partial, double* result,
                         hipStream_t s);

// Fused KEEP (order 2/4) + viscous stage kernel over a tile table.
struct TileDev {
  int box;
  int i0, j0, k0;  // tile origin (absolute index)
  int nz;          // planes in this tile
};
struct FusedConfig {
  int tx, ty, tz;      // tile extents
  int box_tiles = 0;   // 1: tiles per box even where level tiles would apply
  int form = 0;        // 0 first form, 1 self-loading form (16x16 viscous; no barrier after the plane store)
};
// Ghost cells of a level that no box covers (coarse-fine ghost cells) keep
// the caller's values of U^n through a step.  Where every stage array is one
// array over the level, stages after the first read those cells straight
// from U^n instead of a copy in their input: bit set in `bits` (blocks of
// 2^sh cells from `lo`, nb blocks per direction, x fastest) = the block's
// cells are such ghost cells (periodic images resolved; physical-boundary
// and covered cells are clear).  u0: U^n's FABs, strides equal to the stage
// input's.  on = 0: every cell comes from the stage input.
struct UncovSrc {
  int on;
  int lo[3], nb[3], sh;
  const unsigned* bits;
  const Fab* u0;
};
// Plane blocks a tile's loaded planes may span (one bit each of a per-lane
// mask).
constexpr int UNCOV_MAX_ZBLK = 32;
bool fused_supported(const amr_scheme_t& s, int tx, int ty);
// the self-loading form of the fused kernel (FusedConfig::form 1) applies
bool fused_self_supported(const amr_scheme_t& s, int tx, int ty);
hipError_t launch_fused_stage(const amr_scheme_t& s, const FusedConfig& cfg, const WrapArgs& wa, const Cover& cov,
                              const UncovSrc& us, const TileDev* tiles, int ntiles, const BoxDev* boxes,
                              const Closures& c, const StageArgs& sa, const Fab* s1, const Fab* sin, const Fab* out,
                              hipStream_t st, const NbMap& nbm = NbMap{});

}  // namespace amr
