// C-ABI layer: contexts, box tables, descriptor tables, workspace, ghost-fill
// task lists, stage dispatch (fused or pass-structured kernels), CNS::advance
// sequencing, hipGraph capture and reductions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cerrno>
#include <fstream>
#include <sstream>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "amrsolver_hip.h"
#include "kernels.h"
#include "user_source.h"

using namespace amr;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& m) {
  g_err = m;
  return code;
}

#define HIPCHK(x)                                                                              \
  do {                                                                                         \
    hipError_t e_ = (x);                                                                       \
    if (e_ != hipSuccess) return fail(AMR_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

int64_t cells(const int lo[3], const int hi[3]) {
  return int64_t(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
}

// Packed block map of a task list: workgroup w takes cells [w*TASK_BLK,
// (w+1)*TASK_BLK) of the concatenated list (tasks numbered by `start`), so
// small tasks share workgroups; `task` is the task holding the first cell.
template <class T>
int upload_packed_blocks(const std::vector<T>& tasks, int64_t ncell, TaskBlock** d, int* nblk) {
  if (ncell > int64_t(INT32_MAX) - TASK_BLK) return fail(AMR_ERR_ARG, "ghost-fill task list exceeds 2^31 cells");
  std::vector<TaskBlock> h;
  int t = 0;
  for (int64_t g = 0; g < ncell; g += TASK_BLK) {
    while (t + 1 < int(tasks.size()) && tasks[t + 1].start <= g) ++t;
    h.push_back(TaskBlock{t, int(g)});
  }
  *nblk = int(h.size());
  if (!h.empty()) {
    HIPCHK(hipMalloc(d, sizeof(TaskBlock) * h.size()));
    HIPCHK(hipMemcpy(*d, h.data(), sizeof(TaskBlock) * h.size(), hipMemcpyHostToDevice));
  }
  return AMR_OK;
}

// Block map of a task list: TASK_BLK cells of one task per workgroup.
template <class T>
int upload_blocks(const std::vector<T>& tasks, TaskBlock** d, int* nblk) {
  std::vector<TaskBlock> h;
  for (int t = 0; t < int(tasks.size()); ++t) {
    const int64_t n = cells(tasks[t].lo, tasks[t].hi);
    if (n > int64_t(INT32_MAX) - TASK_BLK) return fail(AMR_ERR_ARG, "a ghost-fill region exceeds 2^31 cells");
    for (int64_t f = 0; f < n; f += TASK_BLK) h.push_back(TaskBlock{t, int(f)});
  }
  if (h.size() > size_t(INT32_MAX)) return fail(AMR_ERR_ARG, "ghost-fill task list too long");
  *nblk = int(h.size());
  if (!h.empty()) {
    HIPCHK(hipMalloc(d, sizeof(TaskBlock) * h.size()));
    HIPCHK(hipMemcpy(*d, h.data(), sizeof(TaskBlock) * h.size(), hipMemcpyHostToDevice));
  }
  return AMR_OK;
}

// amr_fab_t for a compact array of `nc` components on [lo-ng, hi+ng]
amr_fab_t compact_fab(const int lo[3], const int hi[3], int ng, double* base) {
  amr_fab_t f;
  std::memset(&f, 0, sizeof f);
  for (int d = 0; d < 3; ++d) {
    f.lo[d] = lo[d];
    f.hi[d] = hi[d];
  }
  f.ng = ng;
  f.data = base;
  int64_t nx = hi[0] - lo[0] + 1 + 2 * ng, ny = hi[1] - lo[1] + 1 + 2 * ng, nz = hi[2] - lo[2] + 1 + 2 * ng;
  f.jstride = nx;
  f.kstride = nx * ny;
  f.nstride = nx * ny * nz;
  return f;
}
int64_t compact_len(const int lo[3], const int hi[3], int ng, int nc) {
  int64_t nx = hi[0] - lo[0] + 1 + 2 * ng, ny = hi[1] - lo[1] + 1 + 2 * ng, nz = hi[2] - lo[2] + 1 + 2 * ng;
  return nx * ny * nz * nc;
}

// Address of cell (0,0,0) comp 0 plus strides: two descriptors with the same
// mapping address the same memory.
struct Mapping {
  intptr_t origin;
  int64_t js, ks, ns;
  bool operator==(const Mapping& o) const {
    return origin == o.origin && js == o.js && ks == o.ks && ns == o.ns;
  }
};
Mapping mapping_of(const amr_fab_t& f) {
  Fab v = make_fab(f);
  return Mapping{intptr_t(f.data) + intptr_t(v.off * 8), f.jstride, f.kstride, f.nstride};
}

}  // namespace

struct amr_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  void* scratch = nullptr;  // this GPU's no-allocation scratch array (device address)
  bool own_stream = false;
  // the caller's handle plus one per live box table: the context (and its
  // stream) is freed when the last of them goes, in whatever order the
  // caller destroys them
  std::atomic<int> refs{1};
  // fused-kernel tiling set explicitly per scheme key (amr_set_tiling); it
  // takes precedence over the per-box-table autotuned choice
  std::map<std::string, FusedConfig> tuned;
};

namespace {

struct DescSet {
  std::vector<amr_fab_t> host;
  Fab* d = nullptr;
  int ustride = -1;  // every box's array has the same strides (-1: not yet known)
  int span32 = -1;   // all the boxes' arrays lie within 2^31 elements (-1: not yet known)
  bool fill_built = false;
  CopyTask* d_copy = nullptr;
  int ncopy = 0;
  int64_t ncopy_cells = 0;
  TaskBlock* d_copy_blk = nullptr;
  int ncopy_blk = 0;
  BcTask* d_bc[3] = {nullptr, nullptr, nullptr};
  int nbc[3] = {0, 0, 0};
  int64_t nbc_cells[3] = {0, 0, 0};
  TaskBlock* d_bc_blk[3] = {nullptr, nullptr, nullptr};
  int nbc_blk[3] = {0, 0, 0};
  std::vector<CopyTask> h_copy;  // host copies of the fill tasks
  std::vector<BcTask> h_bc[3];
  int64_t last_use = 0;
  int wrap = -1;   // one array over the whole periodic domain (-1 unknown)
  int level = -1;  // one array holding the boxes' bounding box (-1 unknown)
  // fill tasks of the inner 1 or 2 ghost layers only (the fused stage
  // kernel reads two); their Fab table is this set's `d`
  std::unique_ptr<DescSet> narrow[2];
  ~DescSet() {
    if (d) (void)hipFree(d);
    if (d_copy) (void)hipFree(d_copy);
    if (d_copy_blk) (void)hipFree(d_copy_blk);
    for (int i = 0; i < 3; ++i) {
      if (d_bc[i]) (void)hipFree(d_bc[i]);
      if (d_bc_blk[i]) (void)hipFree(d_bc_blk[i]);
    }
  }
};

struct TileTable {
  TileDev* d = nullptr;
  int n = 0;
  int max_zblk = 0;  // most blocks of the uncovered-cell map a tile's loaded planes span
  ~TileTable() {
    if (d) (void)hipFree(d);
  }
};

struct GraphEntry {
  std::vector<char> key;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  ~GraphEntry() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

// Buffers of one checkpointed adjoint run: checkpoints ck[nseg] (state at
// the start of every segment of `k` steps), the stage inputs sb of the steps
// of one segment (sb(n, 0) = start of step n, sb(0, 0) = the checkpoint),
// the final state, and three adjoint arrays on the valid cells.
// Arrays of a checkpointed adjoint run: nseg checkpoints (U at the start of
// every k-step segment), the final state, and two sets of stage inputs of one
// segment (krows * nst - 1 arrays each; segment seg uses set seg % 2), so the
// forward sweep leaves the last two segments in place and one segment can be
// recomputed while the previous one is reversed.
struct AdjRun {
  std::vector<char> key;
  int nsteps = 0, k = 0, nst = 0, nseg = 0, krows = 0;
  std::vector<double*> mem;
  std::vector<std::unique_ptr<DescSet>> states;
  std::unique_ptr<DescSet> lam[3];
  std::vector<hipEvent_t> ev;  // reverse done (nseg), recompute done (nseg)
  int slots() const { return krows * nst - 1; }
  DescSet* ck(int seg) { return states[seg].get(); }
  DescSet* uf() { return states[nseg].get(); }
  DescSet* sb(int seg, int n, int s) {
    if (n == 0 && s == 0) return ck(seg);
    return states[nseg + 1 + (seg % 2) * slots() + n * nst + s - 1].get();
  }
  ~AdjRun() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    for (double* p : mem) (void)hipFree(p);
  }
};

}  // namespace

// Coarse-fine tables of a fine level against one coarse level (its serial
// number and the ratio): interpolation tasks (fine ghost regions per coarse
// box) and average-down tasks (coarse regions per fine box).
struct CfTable {
  uint64_t crse_serial = 0;
  int ratio = 0;
  int ngrow = 0;  // ghost width the interpolation tasks cover
  bool interp_built = false, avg_built = false;
  InterpTask* d_interp = nullptr;
  TaskBlock* d_interp_blk = nullptr;
  int ninterp_blk = 0;
  int64_t ninterp_cells = 0;  // coarse cells (lanes) of the interpolation tasks
  CopyTask* d_avg = nullptr;
  TaskBlock* d_avg_blk = nullptr;
  int navg_blk = 0;
  ~CfTable() {
    if (d_interp) (void)hipFree(d_interp);
    if (d_interp_blk) (void)hipFree(d_interp_blk);
    if (d_avg) (void)hipFree(d_avg);
    if (d_avg_blk) (void)hipFree(d_avg_blk);
  }
};

struct amr_boxtab_s {
  amr_ctx_t ctx = nullptr;
  amr_geom_t geom;               // the level's geometry (first problem for a batch)
  std::vector<amr_geom_t> geoms;  // per box: its level's or its own problem's geometry
  bool batch = false;             // independent problems, one box each
  double* d_box_dt = nullptr;     // per-box time steps of a batch step
  Clock* d_clock = nullptr;       // device time-step control (nbox entries; a level uses the first)
  double* d_diag = nullptr;       // amr_diagnostics partials (allocated at its first call)
  int* d_flag = nullptr;          // amr_check_state result
  bool state_check = false;       // amr_set_state_check: check the state after every step call
  double* d_dt_levels = nullptr;  // amr_new_dt_levels: per-level steps (MAX_DT_LEVELS + 1)
  double* h_box_dt = nullptr;     // pinned staging for d_box_dt
  hipEvent_t dt_copied = nullptr;
  BoxTable bt;
  std::vector<std::unique_ptr<DescSet>> descs;
  std::vector<std::unique_ptr<DescSet>> face_descs;  // face-flux arrays (amr_face_fluxes)
  // run-time compiled user_source (amr_set_user_source) and its dSdt buffer
  hipModule_t usrc_mod = nullptr;
  hipFunction_t usrc_fn = nullptr;
  double* usrc_params = nullptr;
  double* usrc_tmp = nullptr;
  Fab* usrc_tmp_fabs = nullptr;
  int64_t use_clock = 0;
  // pass-structured workspace
  double* uw_mem = nullptr;
  int uw_kinds = 0;  // bitmask of allocated kinds
  UnfusedWork uw;
  std::vector<void*> uw_tables;
  // advance workspace (two ghosted state copies mirroring the caller layout)
  double* ws_mem[2] = {nullptr, nullptr};
  std::vector<amr_fab_t> ws_fabs[2];
  std::vector<amr_fab_t> ws_key;
  // reductions
  double* d_partial = nullptr;
  double* d_result = nullptr;
  double* h_result = nullptr;
  // fused tiles per config
  std::map<std::tuple<int, int, int, int>, std::unique_ptr<TileTable>> tiles;  // (tx, ty, tz, level tiling)
  Cover cover{};                      // level tiling (on = 0: not available)
  std::vector<unsigned> cover_bits;   // host copy of cover.bits
  unsigned* d_cover_bits = nullptr;
  NbMap nbmap{};            // owner map for direct neighbour reads (on = 0: not available)
  int* d_owner = nullptr;
  XSide* d_xside = nullptr;  // x sides of the directional passes (null: they need ghost fills)
  unsigned* d_all_covered = nullptr;
  // problem parameters (bc 6 walls, body force) and the wall flux tasks
  bool has_problem = false;
  amr_problem_t problem{};
  WallFill wall_fill{};
  bool walls = false;  // a y boundary has bc code 6
  WallTask* d_wall_tasks = nullptr;
  TaskBlock* d_wall_blk = nullptr;
  WallArgs wall_args{};  // one all-ones word: the bitmap of a fully covered tile set
  std::vector<std::unique_ptr<GraphEntry>> graphs;
  // adjoint workspace: g (NCONS on grow 3) and e (16 on grow 1) per box
  double* adj_mem = nullptr;
  int adj_visc = -1;
  AdjWork aw;
  Fab* adj_e_table = nullptr;
  std::unique_ptr<DescSet> adj_g;
  FoldTasks fold{};                 // transposed fill of adj_g (device task lists)
  std::vector<void*> fold_mem;
  // domain-wide reverse-mode arrays of a periodic level held as one array
  double* adj_lmem = nullptr;
  Fab* adj_lq_table = nullptr;
  Fab* adj_le_table = nullptr;
  std::unique_ptr<AdjRun> run;
  hipStream_t aux = nullptr;  // second stream (adjoint runs: recompute beside the reverse sweep)
  // autotuned tiling per scheme key for this box table
  std::map<std::string, FusedConfig> tuned;
  // source added to dSdt for schemes with rhs_source = 1 (amr_set_source)
  std::unique_ptr<DescSet> src;
  // ghost cells of each box covered by no box of the level (coarse-fine
  // ghosts of a refined level): regions of box b as tasks b -> b
  CopyTask* d_uncov = nullptr;
  int nuncov = 0;
  int64_t nuncov_cells = 0;
  TaskBlock* d_uncov_blk = nullptr;
  int nuncov_blk = 0;
  // blocks of those ghost cells for stages that read them from U^n
  // (srcmap.on = 0: unavailable; u0 is set per launch)
  UncovSrc srcmap{};
  unsigned* d_srcmap_bits = nullptr;
  uint64_t serial = 0;            // distinct per box table of the process
  // boxes by 32-cell bins of the domain (level tables): neighbour queries
  std::vector<std::vector<int>> bins;
  int nbin[3] = {0, 0, 0};
  // coarse-fine tables (this table as the fine level), per coarse table
  // serial, ratio and ghost width (0 for the average-down tasks)
  std::map<std::array<uint64_t, 3>, std::unique_ptr<CfTable>> cf;
};

namespace {

int check_scheme(const amr_scheme_t& s) {
  if (s.art_diss != 0 && s.art_diss != 1) return fail(AMR_ERR_ARG, "art_diss must be 0 or 1");
  if (s.rhs_source != 0 && s.rhs_source != 1) return fail(AMR_ERR_ARG, "rhs_source must be 0 or 1");
  if (!s.rhs_euler && s.rhs_visc)
    return fail(AMR_ERR_ARG, "rhs_euler=0 with rhs_visc=1 is not supported (the reference accumulates into an "
                             "uninitialised flux array)");
  if (s.rhs_euler && (s.flux_euler < 0 || s.flux_euler > 2)) return fail(AMR_ERR_ARG, "flux_euler must be 0, 1 or 2");
  if (s.rhs_euler && s.flux_euler == 1 && !(s.order_keep == 2 || s.order_keep == 4 || s.order_keep == 6))
    return fail(AMR_ERR_ARG, "order_keep must be 2, 4 or 6");
  return AMR_OK;
}

int check_closures(const amr_closures_t& c) {
  if (!(c.gamma > 1.0) || !(c.Rspec > 0.0) || !(c.cv > 0.0))
    return fail(AMR_ERR_ARG, "closures: gamma > 1, Rspec > 0 and cv > 0 required");
  if (c.visc_law < 0 || c.visc_law > 1) return fail(AMR_ERR_ARG, "closures: visc_law must be 0 or 1");
  if (c.cond_law < 0 || c.cond_law > 2) return fail(AMR_ERR_ARG, "closures: cond_law must be 0, 1 or 2");
  if (c.cond_law == 2 && !(c.Pr > 0.0)) return fail(AMR_ERR_ARG, "closures: Pr > 0 required for cond_law 2");
  return AMR_OK;
}

// Checks that fabs[nbox] describe the boxes of the table with >= min_ng ghosts.
// Captured steps hold the device addresses of descriptor tables, workspaces
// and source arrays: every path that frees or reallocates one of them drops
// the graphs first (after the stream has drained), so no replay can touch
// freed memory; the next amr_advance_graph call captures again.
void drop_graphs(amr_boxtab_s* bt) {
  if (bt->graphs.empty()) return;
  (void)hipStreamSynchronize(bt->ctx->stream);
  bt->graphs.clear();
}

int check_fabs(amr_boxtab_s* bt, const amr_fab_t* fabs, int min_ng) {
  int nb = bt->bt.nbox;
  for (int b = 0; b < nb; ++b) {
    const amr_fab_t& f = fabs[b];
    const BoxDev& B = bt->bt.h_boxes[b];
    for (int d = 0; d < 3; ++d)
      if (f.lo[d] != B.lo[d] || f.hi[d] != B.hi[d])
        return fail(AMR_ERR_ARG, "fab " + std::to_string(b) + " does not match box " + std::to_string(b));
    if (f.ng < min_ng)
      return fail(AMR_ERR_ARG, "fab " + std::to_string(b) + " has " + std::to_string(f.ng) + " ghost cells, " +
                                   std::to_string(min_ng) + " required");
    if (!f.data) return fail(AMR_ERR_ARG, "fab " + std::to_string(b) + " has a null data pointer");
    int64_t nx = f.hi[0] - f.lo[0] + 1 + 2 * f.ng, ny = f.hi[1] - f.lo[1] + 1 + 2 * f.ng;
    int64_t nz = f.hi[2] - f.lo[2] + 1 + 2 * f.ng;
    if (f.jstride < nx || f.kstride < f.jstride * ny || f.nstride < f.kstride * nz)
      return fail(AMR_ERR_ARG, "fab " + std::to_string(b) + " strides do not cover its box");
  }
  return AMR_OK;
}

// A descriptor set (host copy + device Fab table) for fabs[nbox].
int new_desc(amr_boxtab_s* bt, const amr_fab_t* fabs, std::unique_ptr<DescSet>* out) {
  int nb = bt->bt.nbox;
  auto ds = std::make_unique<DescSet>();
  ds->host.assign(fabs, fabs + nb);
  std::vector<Fab> h(nb);
  for (int b = 0; b < nb; ++b) h[b] = make_fab(fabs[b]);
  HIPCHK(hipMalloc(&ds->d, sizeof(Fab) * nb));
  HIPCHK(hipMemcpy(ds->d, h.data(), sizeof(Fab) * nb, hipMemcpyHostToDevice));
  *out = std::move(ds);
  return AMR_OK;
}

// Descriptor set for caller-provided `fabs` (nbox entries), uploaded once and
// cached (least recently used of 24 evicted).
int get_desc(amr_boxtab_s* bt, const amr_fab_t* fabs, int min_ng, DescSet** out) {
  int nb = bt->bt.nbox;
  size_t bytes = sizeof(amr_fab_t) * nb;
  for (auto& d : bt->descs) {
    if (std::memcmp(d->host.data(), fabs, bytes) == 0) {
      if (d->host[0].ng < min_ng) break;
      d->last_use = ++bt->use_clock;
      *out = d.get();
      return AMR_OK;
    }
  }
  int rc = check_fabs(bt, fabs, min_ng);
  if (rc) return rc;
  if (bt->descs.size() >= 24) {
    auto it = std::min_element(bt->descs.begin(), bt->descs.end(),
                               [](const auto& a, const auto& b2) { return a->last_use < b2->last_use; });
    drop_graphs(bt);  // a captured step may use the evicted table
    HIPCHK(hipStreamSynchronize(bt->ctx->stream));
    bt->descs.erase(it);
  }
  std::unique_ptr<DescSet> ds;
  if ((rc = new_desc(bt, fabs, &ds))) return rc;
  ds->last_use = ++bt->use_clock;
  *out = ds.get();
  bt->descs.push_back(std::move(ds));
  return AMR_OK;
}

// BC code -> boundary type per component (src/set/CNS_setup.cpp:18-31):
// 0 int_dir (nothing), 1 ext_dir, 2 foextrap, 3 reflect_even, 4 reflect_odd
int bc_type(int code, int comp, int dir) {
  static const int scalar_bc[7] = {0, 1, 2, 3, 3, 3, 1};
  static const int norm_vel_bc[7] = {0, 1, 2, 4, 4, 4, 1};
  static const int tang_vel_bc[7] = {0, 1, 2, 3, 3, 4, 1};
  if (code < 0 || code > 6) return -1;
  if (comp == URHO || comp == UET) return scalar_bc[code];
  return (comp - UMX == dir) ? norm_vel_bc[code] : tang_vel_bc[code];
}

// Boxes of the table that may intersect grow(box b, w) or a periodic image
// of it, in increasing order (every box of a batch's problem: b itself).
constexpr int BIN = 32;
std::vector<int> neighbour_boxes(amr_boxtab_s* bt, int b, int w) {
  if (bt->batch) return {b};
  const amr_geom_t& g = bt->geoms[0];
  int len[3];
  for (int d = 0; d < 3; ++d) len[d] = g.dom_hi[d] - g.dom_lo[d] + 1;
  if (bt->bins.empty()) {
    for (int d = 0; d < 3; ++d) bt->nbin[d] = (len[d] + BIN - 1) / BIN;
    bt->bins.resize(size_t(bt->nbin[0]) * bt->nbin[1] * bt->nbin[2]);
    for (int s = 0; s < bt->bt.nbox; ++s) {
      const BoxDev& S = bt->bt.h_boxes[s];
      int lo[3], hi[3];
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::max(0, (std::max(S.lo[d], g.dom_lo[d]) - g.dom_lo[d]) / BIN);
        hi[d] = std::min(bt->nbin[d] - 1, (std::min(S.hi[d], g.dom_hi[d]) - g.dom_lo[d]) / BIN);
      }
      for (int k = lo[2]; k <= hi[2]; ++k)
        for (int j = lo[1]; j <= hi[1]; ++j)
          for (int i = lo[0]; i <= hi[0]; ++i)
            bt->bins[(size_t(k) * bt->nbin[1] + j) * bt->nbin[0] + i].push_back(s);
    }
  }
  // bins touched per direction by the grown box's cells (periodic images taken)
  std::vector<int> bl[3];
  const BoxDev& B = bt->bt.h_boxes[b];
  for (int d = 0; d < 3; ++d) {
    std::vector<char> on(bt->nbin[d], 0);
    const int lo = B.lo[d] - w, hi = B.hi[d] + w;
    if (g.periodic[d] && hi - lo + 1 >= len[d]) {
      std::fill(on.begin(), on.end(), 1);
    } else {
      for (int c = lo; c <= hi; c += 1) {
        int x = c - g.dom_lo[d];
        if (g.periodic[d]) x = (x % len[d] + len[d]) % len[d];
        else if (x < 0 || x >= len[d]) continue;
        on[x / BIN] = 1;
      }
    }
    for (int i = 0; i < bt->nbin[d]; ++i)
      if (on[i]) bl[d].push_back(i);
  }
  std::vector<int> out;
  for (int k : bl[2])
    for (int j : bl[1])
      for (int i : bl[0]) {
        const auto& v = bt->bins[(size_t(k) * bt->nbin[1] + j) * bt->nbin[0] + i];
        out.insert(out.end(), v.begin(), v.end());
      }
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
  return out;
}

int build_fill(amr_boxtab_s* bt, DescSet* ds) {
  if (ds->fill_built) return AMR_OK;
  int nb = bt->bt.nbox;
  std::vector<Mapping> map(nb);
  for (int b = 0; b < nb; ++b) map[b] = mapping_of(ds->host[b]);
  std::vector<CopyTask> copies;
  int64_t ncell = 0;
  for (int b = 0; b < nb; ++b) {
    const amr_fab_t& fb = ds->host[b];
    const amr_geom_t& g = bt->geoms[b];
    int len[3];
    for (int d = 0; d < 3; ++d) len[d] = g.dom_hi[d] - g.dom_lo[d] + 1;
    int G[2][3];
    for (int d = 0; d < 3; ++d) {
      G[0][d] = fb.lo[d] - fb.ng;
      G[1][d] = fb.hi[d] + fb.ng;
    }
    for (int s : neighbour_boxes(bt, b, fb.ng)) {  // a batch's problems share no data
      const BoxDev& S = bt->bt.h_boxes[s];
      for (int sx = -1; sx <= 1; ++sx)
        for (int sy = -1; sy <= 1; ++sy)
          for (int sz = -1; sz <= 1; ++sz) {
            int sv[3] = {sx, sy, sz};
            bool ok = true;
            int sh[3];
            for (int d = 0; d < 3; ++d) {
              if (sv[d] != 0 && !g.periodic[d]) ok = false;
              sh[d] = sv[d] * len[d];
            }
            if (!ok) continue;
            bool zero = sx == 0 && sy == 0 && sz == 0;
            if (zero && (s == b || map[s] == map[b])) continue;
            CopyTask t;
            bool empty = false;
            for (int d = 0; d < 3; ++d) {
              t.lo[d] = std::max(G[0][d], S.lo[d] + sh[d]);
              t.hi[d] = std::min(G[1][d], S.hi[d] + sh[d]);
              if (t.lo[d] > t.hi[d]) empty = true;
              t.sh[d] = sh[d];
            }
            if (empty) continue;
            t.dst = b;
            t.src = s;
            t.start = ncell;
            ncell += cells(t.lo, t.hi);
            copies.push_back(t);
          }
    }
  }
  ds->ncopy = int(copies.size());
  ds->ncopy_cells = ncell;
  ds->h_copy = copies;
  if (!copies.empty()) {
    HIPCHK(hipMalloc(&ds->d_copy, sizeof(CopyTask) * copies.size()));
    HIPCHK(hipMemcpy(ds->d_copy, copies.data(), sizeof(CopyTask) * copies.size(), hipMemcpyHostToDevice));
  }
  int rc = upload_packed_blocks(copies, ncell, &ds->d_copy_blk, &ds->ncopy_blk);
  if (rc) return rc;
  for (int d = 0; d < 3; ++d) {
    std::vector<BcTask> tasks;
    int64_t nc = 0;
    for (int b = 0; b < nb; ++b) {
      const amr_fab_t& fb = ds->host[b];
      const amr_geom_t& g = bt->geoms[b];
      if (g.periodic[d]) continue;
      for (int side = 0; side < 2; ++side) {
        BcTask t;
        for (int e = 0; e < 3; ++e) {
          t.lo[e] = fb.lo[e] - fb.ng;
          t.hi[e] = fb.hi[e] + fb.ng;
        }
        if (side == 0) {
          if (t.lo[d] >= g.dom_lo[d]) continue;
          t.hi[d] = g.dom_lo[d] - 1;
          t.bnd = g.dom_lo[d];
        } else {
          if (t.hi[d] <= g.dom_hi[d]) continue;
          t.lo[d] = g.dom_hi[d] + 1;
          t.bnd = g.dom_hi[d];
        }
        int code = side == 0 ? g.lo_bc[d] : g.hi_bc[d];
        t.plo = g.prob_lo[d];
        t.phi = g.prob_hi[d];
        if (code == 6) {  // the channel's isothermal no-slip wall (y only)
          if (d != 1) return fail(AMR_ERR_ARG, "bc code 6 walls are supported in y only, as in the reference");
          for (int n = 0; n < 5; ++n) t.type[n] = BC_WALL;
          t.box = b;
          t.dir = d;
          t.side = side;
          t.start = nc;
          nc += cells(t.lo, t.hi);
          tasks.push_back(t);
          continue;
        }
        bool any = false;
        for (int n = 0; n < 5; ++n) {
          int ty = bc_type(code, n, d);
          if (ty < 0) return fail(AMR_ERR_ARG, "invalid physical boundary code");
          if (ty == 1)
            return fail(AMR_ERR_ARG, "ext_dir boundaries (inflow / user) need a problem bcnormal; not supported");
          t.type[n] = ty;
          any = any || ty != 0;
        }
        if (!any) continue;
        t.box = b;
        t.dir = d;
        t.side = side;
        t.start = nc;
        nc += cells(t.lo, t.hi);
        tasks.push_back(t);
      }
    }
    ds->nbc[d] = int(tasks.size());
    ds->nbc_cells[d] = nc;
    ds->h_bc[d] = tasks;
    if (!tasks.empty()) {
      HIPCHK(hipMalloc(&ds->d_bc[d], sizeof(BcTask) * tasks.size()));
      HIPCHK(hipMemcpy(ds->d_bc[d], tasks.data(), sizeof(BcTask) * tasks.size(), hipMemcpyHostToDevice));
    }
    if ((rc = upload_blocks(tasks, &ds->d_bc_blk[d], &ds->nbc_blk[d]))) return rc;
  }
  ds->fill_built = true;
  return AMR_OK;
}

// bc_only: physical boundaries only (the periodic copies of a batch, whose
// problems are single boxes, when the stage reads periodic images itself).
// Fill tasks limited to the two inner ghost layers: built from a copy of
// the descriptors with ng = 2 (same element mapping), launched on the real
// descriptors.
int build_fill_narrow(amr_boxtab_s* bt, DescSet* ds, int width = 2) {
  if (width < 1 || width > 2) return fail(AMR_ERR_ARG, "internal: narrow fills are 1 or 2 layers wide");
  if (ds->narrow[width - 1]) return AMR_OK;
  auto t = std::make_unique<DescSet>();
  t->host = ds->host;
  for (amr_fab_t& f : t->host) {
    const int k = f.ng - std::min(f.ng, width);
    f.data += k * (1 + f.jstride + f.kstride);
    f.ng -= k;
  }
  int rc = build_fill(bt, t.get());
  if (rc) return rc;
  ds->narrow[width - 1] = std::move(t);
  return AMR_OK;
}

// width 1 or 2: fill only that many inner ghost layers (build_fill_narrow
// first); 0: every ghost layer of the descriptors
int enqueue_fill(amr_boxtab_s* bt, DescSet* ds, bool bc_only = false, hipStream_t stream = nullptr,
                 int width = 0) {
  hipStream_t st = stream ? stream : bt->ctx->stream;
  if (bt->walls && !bt->has_problem)
    return fail(AMR_ERR_ARG, "bc code 6 walls need amr_set_problem (wall temperature) first");
  const DescSet* tk = (width == 1 || width == 2) ? ds->narrow[width - 1].get() : ds;
  if (!tk) return fail(AMR_ERR_ARG, "internal: narrow fill tasks not built");
  if (!bc_only) HIPCHK(launch_fill_copy(tk->d_copy, tk->ncopy, tk->d_copy_blk, tk->ncopy_blk, ds->d, st));
  for (int d = 0; d < 3; ++d)
    HIPCHK(launch_fill_bc(tk->d_bc[d], tk->d_bc_blk[d], tk->nbc_blk[d], ds->d, bt->bt.d_boxes, bt->wall_fill, st));
  return AMR_OK;
}

// Flux corrections next to bc 6 walls in y (src/rhs/CentralKEEP.h:233-286
// order reduction; src/tim/CNS_advance.cpp:303-327 viscous wall fluxes): per
// box touching a wall, the three y-face planes next to it and the ghost row of
// cell fluxes outside it, over the box's nodal (faces) or grown-by-1 (cell
// fluxes) extent in x and z.  Built with the box table.
int build_walls(amr_boxtab_s* bt) {
  const amr_geom_t& g = bt->geom;
  bt->walls = !bt->batch && !g.periodic[1] && (g.lo_bc[1] == 6 || g.hi_bc[1] == 6);
  for (int d = 0; d < 3; d += 2)
    if (!bt->batch && !g.periodic[d] && (g.lo_bc[d] == 6 || g.hi_bc[d] == 6))
      return fail(AMR_ERR_ARG, "bc code 6 walls are supported in y only, as in the reference");
  if (bt->batch) {
    for (const amr_geom_t& gg : bt->geoms)
      for (int d = 0; d < 3; ++d)
        if (!gg.periodic[d] && (gg.lo_bc[d] == 6 || gg.hi_bc[d] == 6))
          return fail(AMR_ERR_ARG, "bc code 6 walls are not supported in batches of independent problems");
    return AMR_OK;
  }
  if (!bt->walls) return AMR_OK;
  if (!(g.prob_hi[1] > g.prob_lo[1])) return fail(AMR_ERR_ARG, "bc code 6 walls need prob_hi in the geometry");
  std::vector<WallTask> keep, visc;
  for (int b = 0; b < bt->bt.nbox; ++b) {
    const BoxDev& B = bt->bt.h_boxes[b];
    for (int side = 0; side < 2; ++side) {
      const bool wall = side == 0 ? (g.lo_bc[1] == 6 && B.lo[1] == g.dom_lo[1]) : (g.hi_bc[1] == 6 && B.hi[1] == g.dom_hi[1]);
      if (!wall) continue;
      WallTask k{b, side, {B.lo[0], 0, B.lo[2]}, {B.hi[0] + 1, 0, B.hi[2] + 1}};
      // (clipped to the box's own faces lo..hi+1: the flux arrays hold no
      // more, and faces outside them do not enter the box's assembly)
      if (side == 0) {
        k.lo[1] = B.lo[1];
        k.hi[1] = std::min(B.lo[1] + 2, B.hi[1] + 1);
      } else {
        k.lo[1] = std::max(B.hi[1] - 2, B.lo[1]);
        k.hi[1] = B.hi[1];
      }
      keep.push_back(k);
      const int jj = side == 0 ? B.lo[1] - 1 : B.hi[1] + 1;
      visc.push_back(WallTask{b, side, {B.lo[0] - 1, jj, B.lo[2] - 1}, {B.hi[0] + 1, jj, B.hi[2] + 1}});
    }
  }
  std::vector<WallTask> all(keep);
  all.insert(all.end(), visc.begin(), visc.end());
  std::vector<TaskBlock> blk;
  int nkeep_blk = 0;
  for (int t = 0; t < int(all.size()); ++t) {
    const int64_t n = cells(all[t].lo, all[t].hi);
    for (int64_t f = 0; f < n; f += TASK_BLK) blk.push_back(TaskBlock{t, int(f)});
    if (t + 1 == int(keep.size())) nkeep_blk = int(blk.size());
  }
  HIPCHK(hipMalloc(&bt->d_wall_tasks, sizeof(WallTask) * std::max<size_t>(all.size(), 1)));
  HIPCHK(hipMemcpy(bt->d_wall_tasks, all.data(), sizeof(WallTask) * all.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&bt->d_wall_blk, sizeof(TaskBlock) * std::max<size_t>(blk.size(), 1)));
  HIPCHK(hipMemcpy(bt->d_wall_blk, blk.data(), sizeof(TaskBlock) * blk.size(), hipMemcpyHostToDevice));
  WallArgs& w = bt->wall_args;
  w.keep = bt->d_wall_tasks;
  w.keep_blk = bt->d_wall_blk;
  w.nkeep_blk = nkeep_blk;
  w.visc = bt->d_wall_tasks;  // block entries index the combined task list
  w.visc_blk = bt->d_wall_blk + nkeep_blk;
  w.nvisc_blk = int(blk.size()) - nkeep_blk;
  w.dlo = g.dom_lo[1];
  w.dhi = g.dom_hi[1];
  return AMR_OK;
}

struct IBox {
  int lo[3], hi[3];
};

bool ibox_empty(const IBox& a) { return a.lo[0] > a.hi[0] || a.lo[1] > a.hi[1] || a.lo[2] > a.hi[2]; }

// a minus b as up to six disjoint boxes
void ibox_subtract(const IBox& a, const IBox& b, std::vector<IBox>& out) {
  IBox x;
  for (int d = 0; d < 3; ++d) {
    x.lo[d] = std::max(a.lo[d], b.lo[d]);
    x.hi[d] = std::min(a.hi[d], b.hi[d]);
  }
  if (ibox_empty(x)) {
    out.push_back(a);
    return;
  }
  IBox rest = a;
  for (int d = 0; d < 3; ++d) {
    if (rest.lo[d] < x.lo[d]) {
      IBox p = rest;
      p.hi[d] = x.lo[d] - 1;
      out.push_back(p);
      rest.lo[d] = x.lo[d];
    }
    if (rest.hi[d] > x.hi[d]) {
      IBox p = rest;
      p.lo[d] = x.hi[d] + 1;
      out.push_back(p);
      rest.hi[d] = x.hi[d];
    }
  }
}

// Ghost cells (width NGHOST) of every box that no box of the same level or
// problem covers, periodic images included, and that lie inside the domain
// in every non-periodic direction (outside it the physical BCs fill them).
std::vector<CopyTask> uncovered_pieces(amr_boxtab_s* bt, int width, int64_t* ncell_out);

int build_uncovered(amr_boxtab_s* bt) {
  int64_t ncell = 0;
  std::vector<CopyTask> tasks = uncovered_pieces(bt, NGHOST, &ncell);
  bt->nuncov = int(tasks.size());
  bt->nuncov_cells = ncell;
  if (!tasks.empty()) {
    HIPCHK(hipMalloc(&bt->d_uncov, sizeof(CopyTask) * tasks.size()));
    HIPCHK(hipMemcpy(bt->d_uncov, tasks.data(), sizeof(CopyTask) * tasks.size(), hipMemcpyHostToDevice));
  }
  return upload_blocks(tasks, &bt->d_uncov_blk, &bt->nuncov_blk);
}

// Ghost cells (width `width`) of every box that no box of the same level or
// problem covers, periodic images included, inside the domain in every
// non-periodic direction: regions of box b as tasks b -> b.
std::vector<CopyTask> uncovered_pieces(amr_boxtab_s* bt, int width, int64_t* ncell_out) {
  const int nb = bt->bt.nbox;
  std::vector<CopyTask> tasks;
  int64_t ncell = 0;
  for (int b = 0; b < nb; ++b) {
    const BoxDev& B = bt->bt.h_boxes[b];
    const amr_geom_t& g = bt->geoms[b];
    int len[3];
    for (int d = 0; d < 3; ++d) len[d] = g.dom_hi[d] - g.dom_lo[d] + 1;
    IBox gb;
    for (int d = 0; d < 3; ++d) {
      gb.lo[d] = B.lo[d] - width;
      gb.hi[d] = B.hi[d] + width;
      if (!g.periodic[d]) {
        gb.lo[d] = std::max(gb.lo[d], g.dom_lo[d]);
        gb.hi[d] = std::min(gb.hi[d], g.dom_hi[d]);
      }
    }
    std::vector<IBox> pieces{gb}, next;
    for (int s : neighbour_boxes(bt, b, width)) {
      if (pieces.empty()) break;
      const BoxDev& S = bt->bt.h_boxes[s];
      for (int sx = -1; sx <= 1; ++sx)
        for (int sy = -1; sy <= 1; ++sy)
          for (int sz = -1; sz <= 1; ++sz) {
            const int sv[3] = {sx, sy, sz};
            bool ok = true;
            IBox img;
            for (int d = 0; d < 3; ++d) {
              if (sv[d] != 0 && !g.periodic[d]) ok = false;
              img.lo[d] = S.lo[d] + sv[d] * len[d];
              img.hi[d] = S.hi[d] + sv[d] * len[d];
            }
            if (!ok) continue;
            next.clear();
            for (const IBox& p : pieces) ibox_subtract(p, img, next);
            pieces.swap(next);
          }
    }
    for (const IBox& p : pieces) {
      CopyTask t;
      t.dst = t.src = b;
      for (int d = 0; d < 3; ++d) {
        t.lo[d] = p.lo[d];
        t.hi[d] = p.hi[d];
        t.sh[d] = 0;
      }
      t.start = ncell;
      ncell += cells(t.lo, t.hi);
      tasks.push_back(t);
    }
  }
  *ncell_out = ncell;
  return tasks;
}

// Uncovered ghost cells of `from` into `to` and, if given, `to2` (no-op on
// levels that cover their own ghost cells).
int copy_uncovered(amr_boxtab_s* bt, const DescSet* from, const DescSet* to, const DescSet* to2 = nullptr) {
  if (bt->nuncov_cells == 0) return AMR_OK;
  HIPCHK(launch_copy_regions(bt->d_uncov, bt->d_uncov_blk, bt->nuncov_blk, from->d, to->d, to2 ? to2->d : nullptr,
                             bt->ctx->stream));
  return AMR_OK;
}

enum { UW_PRIMS = 1, UW_VISC = 2, UW_WENO = 4, UW_SLOPE = 8, UW_JST = 16 };

int ensure_unfused(amr_boxtab_s* bt, const amr_scheme_t& s) {
  int need = UW_PRIMS;
  if (s.rhs_visc) need |= UW_VISC;
  if (s.rhs_euler && s.flux_euler == 2) need |= UW_WENO;
  if (s.rhs_euler && s.flux_euler == 0) need |= UW_SLOPE;
  if (s.rhs_euler && s.art_diss) need |= UW_JST;
  need |= bt->uw_kinds;
  if ((need & ~bt->uw_kinds) == 0) return AMR_OK;
  // (re)allocate everything needed in one block
  drop_graphs(bt);
  HIPCHK(hipStreamSynchronize(bt->ctx->stream));
  if (bt->uw_mem) HIPCHK(hipFree(bt->uw_mem));
  for (void* p : bt->uw_tables) HIPCHK(hipFree(p));
  bt->uw_tables.clear();
  bt->uw_mem = nullptr;
  int nb = bt->bt.nbox;
  struct Arr {
    int ng, nc, nodal;
  };
  std::vector<Arr> kinds;  // order: prims, nf0..2, [pv0..2], [pe0..2, lam], [slope], [jst]
  kinds.push_back({3, NPRIM, 0});
  for (int d = 0; d < 3; ++d) kinds.push_back({0, NCONS, 1});
  if (need & UW_VISC)
    for (int d = 0; d < 3; ++d) kinds.push_back({1, NCONS, 0});
  if (need & UW_WENO) {
    for (int d = 0; d < 3; ++d) kinds.push_back({3, NCONS, 0});
    kinds.push_back({3, 3, 0});
  }
  if (need & UW_SLOPE) kinds.push_back({1, NCONS, 0});
  if (need & UW_JST) kinds.push_back({1, 6, 0});
  int64_t total = 0;
  for (auto& k : kinds)
    for (int b = 0; b < nb; ++b) {
      BoxDev B = bt->bt.h_boxes[b];
      if (k.nodal)
        for (int d = 0; d < 3; ++d) B.hi[d] += 1;
      total += compact_len(B.lo, B.hi, k.ng, k.nc);
    }
  HIPCHK(hipMalloc(&bt->uw_mem, sizeof(double) * std::max<int64_t>(total, 1)));
  HIPCHK(hipMemset(bt->uw_mem, 0, sizeof(double) * std::max<int64_t>(total, 1)));
  int64_t off = 0;
  std::vector<Fab*> tables;
  for (auto& k : kinds) {
    std::vector<Fab> h(nb);
    for (int b = 0; b < nb; ++b) {
      BoxDev B = bt->bt.h_boxes[b];
      if (k.nodal)
        for (int d = 0; d < 3; ++d) B.hi[d] += 1;
      h[b] = make_fab(compact_fab(B.lo, B.hi, k.ng, bt->uw_mem + off));
      off += compact_len(B.lo, B.hi, k.ng, k.nc);
    }
    Fab* d = nullptr;
    HIPCHK(hipMalloc(&d, sizeof(Fab) * nb));
    HIPCHK(hipMemcpy(d, h.data(), sizeof(Fab) * nb, hipMemcpyHostToDevice));
    bt->uw_tables.push_back(d);
    tables.push_back(d);
  }
  size_t t = 0;
  bt->uw = UnfusedWork();
  bt->uw.prims = tables[t++];
  for (int d = 0; d < 3; ++d) bt->uw.nf[d] = tables[t++];
  if (need & UW_VISC)
    for (int d = 0; d < 3; ++d) bt->uw.pv[d] = tables[t++];
  if (need & UW_WENO) {
    for (int d = 0; d < 3; ++d) bt->uw.pe[d] = tables[t++];
    bt->uw.lam = tables[t++];
  }
  if (need & UW_SLOPE) bt->uw.slope = tables[t++];
  if (need & UW_JST) bt->uw.jst = tables[t++];
  if (need & (UW_WENO | UW_SLOPE)) {
    for (int d = 0; d < 3; ++d) {
      // a box spanning a periodic direction is its own periodic image there
      std::vector<int> wrap(size_t(nb), 0);
      for (int b = 0; b < nb; ++b) {
        const amr_geom_t& g = bt->geoms[b];
        const BoxDev& B = bt->bt.h_boxes[b];
        if (g.periodic[d] && B.lo[d] == g.dom_lo[d] && B.hi[d] == g.dom_hi[d]) wrap[b] = B.hi[d] - B.lo[d] + 1;
      }
      const int n = weno_brick_count(bt->bt, d);
      std::vector<WenoBrick> h(size_t(std::max(n, 1)));
      weno_bricks(bt->bt, d, wrap.data(), h.data());
      WenoBrick* dv = nullptr;
      HIPCHK(hipMalloc(&dv, sizeof(WenoBrick) * h.size()));
      HIPCHK(hipMemcpy(dv, h.data(), sizeof(WenoBrick) * h.size(), hipMemcpyHostToDevice));
      bt->uw_tables.push_back(dv);
      bt->uw.weno.d[d] = dv;
      bt->uw.weno.n[d] = n;
    }
    bt->uw.weno.xside = bt->d_xside;
  }
  bt->uw_kinds = need;
  return AMR_OK;
}

// Schemes the directional flux passes run (unfused.hip).
bool dir_flux_scheme(const amr_scheme_t& s) {
  return s.rhs_euler && (s.flux_euler == 2 || s.flux_euler == 0) && !s.rhs_visc && !s.art_diss;
}

std::string scheme_key(const amr_scheme_t& s) {
  char buf[96];
  std::snprintf(buf, sizeof buf, "e%d v%d f%d k%d", s.rhs_euler, s.rhs_visc, s.flux_euler, s.order_keep);
  return buf;
}

// Tile table for the fused kernel; blockIdx -> tile is permuted so that
// consecutive tiles of a box land on one XCD (blocks are dealt round-robin
// over the 8 XCDs, each with its own L2).
bool cover_bit(const amr_boxtab_s* bt, int64_t bx, int64_t by, int64_t bz) {
  const Cover& c = bt->cover;
  if (bt->cover_bits.empty()) return true;
  const int64_t bit = (bz * c.nb[1] + by) * c.nb[0] + bx;
  return (bt->cover_bits[size_t(bit >> 5)] >> (bit & 31)) & 1u;
}

// Tiles of the fused kernel: box by box, or (level tiling) TX x TY columns on
// a grid anchored at the bounding box's corner, split into runs of planes
// that hold covered cells, each run cut into near-equal pieces of at most
// tz planes.
int get_tiles(amr_boxtab_s* bt, const FusedConfig& cfg, bool level, TileTable** out) {
  auto key = std::make_tuple(cfg.tx, cfg.ty, cfg.tz, level ? 1 : 0);
  auto it = bt->tiles.find(key);
  if (it != bt->tiles.end()) {
    *out = it->second.get();
    return AMR_OK;
  }
  std::vector<TileDev> logical;
  if (level) {
    const Cover& c = bt->cover;
    const int sh = c.sh;
    for (int j0 = c.lo[1]; j0 <= c.hi[1]; j0 += cfg.ty)
      for (int i0 = c.lo[0]; i0 <= c.hi[0]; i0 += cfg.tx) {
        const int bx0 = (i0 - c.lo[0]) >> sh, bx1 = std::min(i0 + cfg.tx - 1 - c.lo[0], c.hi[0] - c.lo[0]) >> sh;
        const int by0 = (j0 - c.lo[1]) >> sh, by1 = std::min(j0 + cfg.ty - 1 - c.lo[1], c.hi[1] - c.lo[1]) >> sh;
        auto plane_block_used = [&](int bz) {
          for (int by = by0; by <= by1; ++by)
            for (int bx = bx0; bx <= bx1; ++bx)
              if (cover_bit(bt, bx, by, bz)) return true;
          return false;
        };
        int bz = 0;
        while (bz < c.nb[2]) {
          if (!plane_block_used(bz)) {
            ++bz;
            continue;
          }
          int ez = bz;
          while (ez + 1 < c.nb[2] && plane_block_used(ez + 1)) ++ez;
          const int k0 = c.lo[2] + (bz << sh);
          const int k1 = std::min(c.hi[2], c.lo[2] + ((ez + 1) << sh) - 1);
          const int run = k1 - k0 + 1;
          const int pieces = (run + cfg.tz - 1) / cfg.tz;
          int k = k0;
          for (int p = 0; p < pieces; ++p) {
            const int nz = run / pieces + (p < run % pieces ? 1 : 0);
            logical.push_back(TileDev{0, i0, j0, k, nz});
            k += nz;
          }
          bz = ez + 1;
        }
      }
  } else {
    for (int b = 0; b < bt->bt.nbox; ++b) {
      const BoxDev& B = bt->bt.h_boxes[b];
      for (int k0 = B.lo[2]; k0 <= B.hi[2]; k0 += cfg.tz)
        for (int j0 = B.lo[1]; j0 <= B.hi[1]; j0 += cfg.ty)
          for (int i0 = B.lo[0]; i0 <= B.hi[0]; i0 += cfg.tx) {
            TileDev t;
            t.box = b;
            t.i0 = i0;
            t.j0 = j0;
            t.k0 = k0;
            t.nz = std::min(cfg.tz, B.hi[2] - k0 + 1);
            logical.push_back(t);
          }
    }
  }
  // Block t runs on XCD t % 8, and each XCD works through its share on its
  // own 32 CUs, so the launch ends when the most loaded XCD does.  XCD x gets
  // a contiguous range of the spatially ordered tiles (halo re-reads stay in
  // its L2) holding 1/8 of the work, a tile costing one plane iteration per
  // plane plus three for the pipeline; ranges are padded with empty tiles
  // (nz = 0, which the kernel skips) to a common length, and each range runs
  // its longest tiles first, so the workgroups that start last are short.
  const int NXCD = 8;
  {
    double tot = 0.0;
    for (auto& t : logical) tot += t.nz + 3;
    std::vector<std::vector<TileDev>> parts(NXCD);
    double acc = 0.0;
    int x = 0;
    for (auto& t : logical) {
      if (x < NXCD - 1 && acc + 0.5 * (t.nz + 3) > tot * (x + 1) / NXCD) ++x;
      parts[x].push_back(t);
      acc += t.nz + 3;
    }
    size_t mx = 0;
    for (auto& p : parts) mx = std::max(mx, p.size());
    logical.clear();
    for (auto& p : parts) {
      std::stable_sort(p.begin(), p.end(), [](const TileDev& a, const TileDev& b) { return a.nz > b.nz; });
      p.resize(mx, TileDev{0, 0, 0, 0, 0});
      logical.insert(logical.end(), p.begin(), p.end());
    }
  }
  int n = int(logical.size());
  std::vector<TileDev> table(n);
  const int per = n / NXCD;
  int max_zblk = 0;
  if (bt->srcmap.on)
    for (const TileDev& t : logical)
      if (t.nz > 0) {
        const int lo2 = bt->srcmap.lo[2], sh = bt->srcmap.sh;
        max_zblk = std::max(max_zblk, ((t.k0 + t.nz + 2 - lo2) >> sh) - ((t.k0 - 3 - lo2) >> sh) + 1);
      }
  std::vector<int> order;
  order.reserve(n);
  for (int t = 0; t < per * NXCD; ++t) {
    int x = t % NXCD, r = t / NXCD;
    int l = x * per + r;
    if (l < n) order.push_back(l);
  }
  for (int t = 0; t < n; ++t) table[t] = logical[order[t]];
  auto tt = std::make_unique<TileTable>();
  tt->n = n;
  tt->max_zblk = max_zblk;
  HIPCHK(hipMalloc(&tt->d, sizeof(TileDev) * std::max(n, 1)));
  HIPCHK(hipMemcpy(tt->d, table.data(), sizeof(TileDev) * n, hipMemcpyHostToDevice));
  *out = tt.get();
  bt->tiles[key] = std::move(tt);
  return AMR_OK;
}

FusedConfig default_config(const amr_boxtab_s* bt) {
  // pick the largest tile that does not exceed the smallest box extent
  int mnx = 1 << 30, mny = 1 << 30;
  for (auto& B : bt->bt.h_boxes) {
    mnx = std::min(mnx, B.hi[0] - B.lo[0] + 1);
    mny = std::min(mny, B.hi[1] - B.lo[1] + 1);
  }
  FusedConfig c{16, 16, 64};
  if (mnx < 16 || mny < 16) c = FusedConfig{8, 8, 64};
  return c;
}

// The fused kernel takes any box (tiles overhanging a box in x or y are
// masked); it skips only shapes whose tiles would be mostly idle.
bool fused_fits(const amr_boxtab_s* bt, const FusedConfig& cfg, bool level = false) {
  if (level) return true;  // level tiles idle only along the union's boundary
  int64_t cells_ = 0, lanes = 0;
  for (auto& B : bt->bt.h_boxes) {
    const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1, nz = B.hi[2] - B.lo[2] + 1;
    cells_ += int64_t(nx) * ny * nz;
    lanes += int64_t((nx + cfg.tx - 1) / cfg.tx) * cfg.tx * ((ny + cfg.ty - 1) / cfg.ty) * cfg.ty * nz;
  }
  return 2 * cells_ >= lanes;
}

FusedConfig choose_config(const amr_boxtab_s* bt, const amr_scheme_t& sch) {
  const std::string key = scheme_key(sch);
  auto tk = bt->ctx->tuned.find(key);
  if (tk != bt->ctx->tuned.end()) return tk->second;
  auto tb = bt->tuned.find(key);
  if (tb != bt->tuned.end()) return tb->second;
  return default_config(bt);
}

// Whether the fused stage kernel handles this scheme on this table.
bool fused_path(const amr_boxtab_s* bt, const amr_scheme_t& sch, FusedConfig* out, bool level = false) {
  FusedConfig cfg = choose_config(bt, sch);
  if (out) *out = cfg;
  if (bt->walls) return false;  // the wall flux corrections run in the pass-structured kernels
  if (sch.rhs_source && bt->usrc_fn) return false;  // the user's source runs between assembly and combination
  return cfg.tx > 0 && fused_supported(sch, cfg.tx, cfg.ty) && fused_fits(bt, cfg, level);
}

// Level tiling needs the boxes' bounding box [lo, hi] and a bitmap of its
// covered blocks; the block edge is the largest power of two (at most 2^6)
// dividing every box boundary's offset from lo.  Built with the box table.
int build_cover(amr_boxtab_s* bt) {
  bt->cover = Cover{};
  bt->cover_bits.clear();
  if (!bt->d_all_covered) {
    const unsigned ones = ~0u;
    HIPCHK(hipMalloc(&bt->d_all_covered, sizeof(unsigned)));
    HIPCHK(hipMemcpy(bt->d_all_covered, &ones, sizeof(unsigned), hipMemcpyHostToDevice));
  }
  if (bt->batch || bt->bt.nbox == 0) return AMR_OK;
  const auto& B = bt->bt.h_boxes;
  Cover c{};
  for (int d = 0; d < 3; ++d) {
    c.lo[d] = B[0].lo[d];
    c.hi[d] = B[0].hi[d];
  }
  for (const BoxDev& b : B)
    for (int d = 0; d < 3; ++d) {
      c.lo[d] = std::min(c.lo[d], b.lo[d]);
      c.hi[d] = std::max(c.hi[d], b.hi[d]);
    }
  unsigned g = 1u << 6;
  for (const BoxDev& b : B)
    for (int d = 0; d < 3; ++d) g |= unsigned(b.lo[d] - c.lo[d]) | unsigned(b.hi[d] + 1 - c.lo[d]);
  c.sh = __builtin_ctz(g);
  int64_t nbits = 1, bbox = 1, sum = 0;
  for (int d = 0; d < 3; ++d) {
    c.nb[d] = ((c.hi[d] - c.lo[d] + 1) + (1 << c.sh) - 1) >> c.sh;
    nbits *= c.nb[d];
    bbox *= c.hi[d] - c.lo[d] + 1;
  }
  for (const BoxDev& b : B) sum += cells(b.lo, b.hi);
  if (nbits > (int64_t(1) << 28)) return AMR_OK;  // bitmap too large: box tiling only
  c.on = 1;
  c.bits = nullptr;
  if (sum != bbox) {
    bt->cover_bits.assign(size_t((nbits + 31) / 32), 0u);
    for (const BoxDev& b : B)
      for (int bz = (b.lo[2] - c.lo[2]) >> c.sh; bz <= (b.hi[2] - c.lo[2]) >> c.sh; ++bz)
        for (int by = (b.lo[1] - c.lo[1]) >> c.sh; by <= (b.hi[1] - c.lo[1]) >> c.sh; ++by)
          for (int bx = (b.lo[0] - c.lo[0]) >> c.sh; bx <= (b.hi[0] - c.lo[0]) >> c.sh; ++bx) {
            const int64_t bit = (int64_t(bz) * c.nb[1] + by) * c.nb[0] + bx;
            bt->cover_bits[size_t(bit >> 5)] |= 1u << (bit & 31);
          }
    HIPCHK(hipMalloc(&bt->d_cover_bits, sizeof(unsigned) * bt->cover_bits.size()));
    HIPCHK(hipMemcpy(bt->d_cover_bits, bt->cover_bits.data(), sizeof(unsigned) * bt->cover_bits.size(),
                     hipMemcpyHostToDevice));
    c.bits = bt->d_cover_bits;
  }
  bt->cover = c;
  return AMR_OK;
}

// Owner map of the level's domain for the fused kernel's direct neighbour
// reads (NbMap): blocks of 2^sh cells a side, sh the largest power of two
// (at most 2^6) dividing every box boundary's and every periodic length's
// offset from the domain's lower corner.  Built with the box table for a
// level whose boxes keep NGHOST cells away from every non-periodic domain
// boundary (their halos are then inside the domain or periodic images);
// off for batches and for maps over 2^24 blocks.
int build_nbmap(amr_boxtab_s* bt) {
  bt->nbmap = NbMap{};
  if (bt->batch || bt->bt.nbox == 0) return AMR_OK;
  const amr_geom_t& g = bt->geom;
  const auto& B = bt->bt.h_boxes;
  NbMap m{};
  unsigned msk = 1u << 6;
  for (int d = 0; d < 3; ++d) {
    m.dlo[d] = g.dom_lo[d];
    m.len[d] = g.dom_hi[d] - g.dom_lo[d] + 1;
    m.per[d] = g.periodic[d] ? 1 : 0;
    if (m.per[d] && m.len[d] < 2 * NGHOST) return AMR_OK;  // a halo would wrap more than once
    if (m.per[d]) msk |= unsigned(m.len[d]);
    for (const BoxDev& b : B) {
      if (b.lo[d] < g.dom_lo[d] || b.hi[d] > g.dom_hi[d]) return AMR_OK;
      if (!m.per[d] && (b.lo[d] - g.dom_lo[d] < NGHOST || g.dom_hi[d] - b.hi[d] < NGHOST)) return AMR_OK;
      msk |= unsigned(b.lo[d] - g.dom_lo[d]) | unsigned(b.hi[d] + 1 - g.dom_lo[d]);
    }
  }
  m.sh = __builtin_ctz(msk);
  int64_t nblk = 1;
  for (int d = 0; d < 3; ++d) {
    m.nb[d] = (m.len[d] + (1 << m.sh) - 1) >> m.sh;
    nblk *= m.nb[d];
  }
  if (nblk > (int64_t(1) << 24)) return AMR_OK;
  std::vector<int> own(size_t(nblk), -1);
  for (int i = 0; i < bt->bt.nbox; ++i) {
    const BoxDev& b = B[i];
    for (int bz = (b.lo[2] - m.dlo[2]) >> m.sh; bz <= (b.hi[2] - m.dlo[2]) >> m.sh; ++bz)
      for (int by = (b.lo[1] - m.dlo[1]) >> m.sh; by <= (b.hi[1] - m.dlo[1]) >> m.sh; ++by)
        for (int bx = (b.lo[0] - m.dlo[0]) >> m.sh; bx <= (b.hi[0] - m.dlo[0]) >> m.sh; ++bx)
          own[size_t((int64_t(bz) * m.nb[1] + by) * m.nb[0] + bx)] = i;
  }
  if (bt->d_owner) HIPCHK(hipFree(bt->d_owner));
  bt->d_owner = nullptr;
  HIPCHK(hipMalloc(&bt->d_owner, sizeof(int) * own.size()));
  HIPCHK(hipMemcpy(bt->d_owner, own.data(), sizeof(int) * own.size(), hipMemcpyHostToDevice));
  m.owner = bt->d_owner;
  m.on = 1;
  bt->nbmap = m;
  return AMR_OK;
}

// Side tables of the directional x pass (XSide) for a table whose boxes all
// span their problem's periodic y and z extents (the y and z passes read
// periodic images of their own cells): each x end of a box is a periodic
// image of the box itself (the pass wraps), one neighbour box covering the
// NGHOST cells beyond it (or its periodic image), or a physical boundary
// whose rule is foextrap or reflect for every component.  Otherwise no
// table: the passes read ghost cells.  Built with the box table.
int build_xsides(amr_boxtab_s* bt) {
  const int nb = bt->bt.nbox;
  const auto& B = bt->bt.h_boxes;
  std::vector<XSide> xs(size_t(2 * nb));
  for (int b = 0; b < nb; ++b) {
    const amr_geom_t& g = bt->geoms[b];
    for (int d = 1; d < 3; ++d)
      if (!g.periodic[d] || B[b].lo[d] != g.dom_lo[d] || B[b].hi[d] != g.dom_hi[d]) return AMR_OK;
  }
  // a level's boxes now all span the same y and z extents: disjoint x
  // intervals, searched by their lower ends
  std::vector<int> by_lo;
  if (!bt->batch) {
    by_lo.resize(size_t(nb));
    for (int b = 0; b < nb; ++b) by_lo[size_t(b)] = b;
    std::sort(by_lo.begin(), by_lo.end(), [&](int a, int c) { return B[a].lo[0] < B[c].lo[0]; });
  }
  for (int b = 0; b < nb; ++b) {
    const amr_geom_t& g = bt->geoms[b];
    const int len = g.dom_hi[0] - g.dom_lo[0] + 1;
    if (g.periodic[0] && B[b].lo[0] == g.dom_lo[0] && B[b].hi[0] == g.dom_hi[0]) continue;  // wraps
    if (B[b].hi[0] - B[b].lo[0] + 1 < NGHOST) return AMR_OK;
    for (int side = 0; side < 2; ++side) {
      XSide& X = xs[size_t(2 * b + side)];
      // the cells beyond this end: [e0, e1]
      int e0 = side == 0 ? B[b].lo[0] - NGHOST : B[b].hi[0] + 1;
      int e1 = e0 + NGHOST - 1;
      const bool outside = side == 0 ? e1 < g.dom_lo[0] : e0 > g.dom_hi[0];
      if (outside && !g.periodic[0]) {
        const int code = side == 0 ? g.lo_bc[0] : g.hi_bc[0];
        X.kind = 2;
        X.bnd = side == 0 ? g.dom_lo[0] : g.dom_hi[0];
        for (int n = 0; n < 5; ++n) {
          X.type[n] = bc_type(code, n, 0);
          if (X.type[n] < 2 || X.type[n] > 4) return AMR_OK;
        }
        continue;
      }
      X.shift = 0;
      if (outside) {  // periodic image
        X.shift = side == 0 ? len : -len;
        e0 += X.shift;
        e1 += X.shift;
      }
      if (e0 < g.dom_lo[0] || e1 > g.dom_hi[0]) return AMR_OK;
      X.src = -1;
      if (!bt->batch) {
        // the last box starting at or before e0
        auto it = std::upper_bound(by_lo.begin(), by_lo.end(), e0, [&](int v, int o) { return v < B[o].lo[0]; });
        if (it != by_lo.begin() && B[*(it - 1)].hi[0] >= e1) X.src = *(it - 1);
      }
      if (X.src < 0) return AMR_OK;
      X.kind = 1;
    }
  }
  HIPCHK(hipMalloc(&bt->d_xside, sizeof(XSide) * xs.size()));
  HIPCHK(hipMemcpy(bt->d_xside, xs.data(), sizeof(XSide) * xs.size(), hipMemcpyHostToDevice));
  return AMR_OK;
}

// Block map of the ghost cells no box covers (UncovSrc): blocks of 2^sh
// cells, sh the cover's block exponent lowered until the domain boundaries
// fall on block boundaries too (then every cell of a block shares the block's
// class: inside a box or a periodic image of one, outside the domain in a
// direction with a physical boundary, or such a ghost cell); the map spans
// the boxes' bounding box grown by at least NGHOST cells.
int build_srcmap(amr_boxtab_s* bt) {
  bt->srcmap = UncovSrc{};
  if (!bt->cover.on || bt->nuncov_cells == 0) return AMR_OK;
  const Cover& c = bt->cover;
  const amr_geom_t& g = bt->geom;
  int sh = c.sh;
  auto aligned = [&](int sh_) {
    for (int d = 0; d < 3; ++d)
      if (((g.dom_lo[d] - c.lo[d]) & ((1 << sh_) - 1)) || ((g.dom_hi[d] + 1 - c.lo[d]) & ((1 << sh_) - 1)))
        return false;
    return true;
  };
  while (sh > 0 && !aligned(sh)) --sh;
  const int B = 1 << sh, G = (NGHOST + B - 1) / B;
  UncovSrc m{};
  int64_t nbits = 1;
  for (int d = 0; d < 3; ++d) {
    m.lo[d] = c.lo[d] - G * B;
    m.nb[d] = ((c.hi[d] - c.lo[d] + 1) + B - 1) / B + 2 * G;
    nbits *= m.nb[d];
  }
  if (nbits > (int64_t(1) << 28)) return AMR_OK;
  m.sh = sh;
  std::vector<unsigned> bits(size_t((nbits + 31) / 32), 0u);
  for (int bz = 0; bz < m.nb[2]; ++bz)
    for (int by = 0; by < m.nb[1]; ++by)
      for (int bx = 0; bx < m.nb[0]; ++bx) {
        const int b3[3] = {bx, by, bz};
        int cell[3];
        bool physical = false;
        for (int d = 0; d < 3; ++d) {
          int x = m.lo[d] + b3[d] * B;
          const int len = g.dom_hi[d] - g.dom_lo[d] + 1;
          if (x < g.dom_lo[d] || x > g.dom_hi[d]) {
            if (!g.periodic[d]) physical = true;
            x = g.dom_lo[d] + (((x - g.dom_lo[d]) % len) + len) % len;
          }
          cell[d] = x;
        }
        if (physical) continue;
        bool inside = true;
        for (int d = 0; d < 3; ++d) inside = inside && cell[d] >= c.lo[d] && cell[d] <= c.hi[d];
        const bool covered = inside && cover_bit(bt, (cell[0] - c.lo[0]) >> c.sh, (cell[1] - c.lo[1]) >> c.sh,
                                                 (cell[2] - c.lo[2]) >> c.sh);
        if (!covered) {
          const int64_t bit = (int64_t(bz) * m.nb[1] + by) * m.nb[0] + bx;
          bits[size_t(bit >> 5)] |= 1u << (bit & 31);
        }
      }
  HIPCHK(hipMalloc(&bt->d_srcmap_bits, sizeof(unsigned) * bits.size()));
  HIPCHK(hipMemcpy(bt->d_srcmap_bits, bits.data(), sizeof(unsigned) * bits.size(), hipMemcpyHostToDevice));
  m.bits = bt->d_srcmap_bits;
  m.on = 1;
  bt->srcmap = m;
  return AMR_OK;
}

// Every box of `ds` is a window of one array (same mapping) whose address
// range also holds what level tiles touch: the boxes' bounding box grown by
// NGHOST in x and y and by two planes in z.  The address range of an array
// is taken as the span of its boxes' windows (first to last element), all
// of which lie in the caller's allocation.
bool level_ok(const amr_boxtab_s* bt, DescSet* ds) {
  if (ds->level >= 0) return ds->level == 1;
  ds->level = 0;
  if (!bt->cover.on) return false;
  const int nb = bt->bt.nbox;
  const Mapping m0 = mapping_of(ds->host[0]);
  intptr_t lo = INTPTR_MAX, hi = 0;
  for (int b = 0; b < nb; ++b) {
    const amr_fab_t& f = ds->host[b];
    if (!(mapping_of(f) == m0)) return false;
    int64_t nx = f.hi[0] - f.lo[0] + 1 + 2 * f.ng, ny = f.hi[1] - f.lo[1] + 1 + 2 * f.ng;
    int64_t nz = f.hi[2] - f.lo[2] + 1 + 2 * f.ng;
    int64_t last = (nx - 1) + (ny - 1) * f.jstride + (nz - 1) * f.kstride + (NCONS - 1) * f.nstride;
    lo = std::min(lo, intptr_t(f.data));
    hi = std::max(hi, intptr_t(f.data + last + 1));
  }
  const Cover& c = bt->cover;
  const intptr_t a0 = m0.origin + 8 * intptr_t((c.lo[0] - NGHOST) + (c.lo[1] - NGHOST) * m0.js +
                                               (c.lo[2] - 2) * m0.ks);
  const intptr_t a1 = m0.origin + 8 * intptr_t((c.hi[0] + NGHOST) + (c.hi[1] + NGHOST) * m0.js +
                                               (c.hi[2] + 2) * m0.ks + (NCONS - 1) * m0.ns) + 8;
  if (a0 < lo || a1 > hi) return false;
  ds->level = 1;
  return true;
}

// Whether `ds` is one array over the whole periodic domain (every box a
// window of it, the boxes tiling the domain): the fused kernel then reads
// periodic images of halo cells itself and needs no ghost fill.
bool wrap_ok(const amr_boxtab_s* bt, DescSet* ds) {
  if (ds->wrap >= 0) return ds->wrap == 1;
  ds->wrap = 0;
  if (bt->batch) return false;
  const amr_geom_t& g = bt->geom;
  int64_t dom = 1;
  for (int d = 0; d < 3; ++d) {
    if (!g.periodic[d]) return false;
    dom *= g.dom_hi[d] - g.dom_lo[d] + 1;
  }
  const Mapping m0 = mapping_of(ds->host[0]);
  int64_t sum = 0;
  for (int b = 0; b < bt->bt.nbox; ++b) {
    if (!(mapping_of(ds->host[b]) == m0)) return false;
    sum += cells(bt->bt.h_boxes[b].lo, bt->bt.h_boxes[b].hi);
  }
  if (sum != dom) return false;
  ds->wrap = 1;
  return true;
}

WrapArgs wrap_args(const amr_boxtab_s* bt, DescSet* ds) {
  WrapArgs wa{0, {0, 0, 0}, {1, 1, 1}};
  if (wrap_ok(bt, ds)) {
    wa.on = 1;
    for (int d = 0; d < 3; ++d) {
      wa.lo[d] = bt->geom.dom_lo[d];
      wa.len[d] = bt->geom.dom_hi[d] - bt->geom.dom_lo[d] + 1;
    }
  }
  return wa;
}

// The fused kernel's view of the coverage: the level's bitmap under level
// tiling, else (and for a fully covered level) one all-ones word that every
// cell maps to (block edge 2^30).
Cover kernel_cover(const amr_boxtab_s* bt, bool level) {
  Cover c{};
  if (level) c = bt->cover;
  c.on = level ? 1 : 0;
  if (!c.bits) {
    c.bits = bt->d_all_covered;
    c.sh = 30;
    for (int d = 0; d < 3; ++d) c.nb[d] = 1;
  }
  return c;
}

// Level tiling for a fused stage: every array of the stage is one array over
// the level, and the level is not a periodic domain-wide array (that case
// reads periodic images through box tiles).
bool level_possible(amr_boxtab_s* bt, const amr_scheme_t& sch, DescSet* ds1, DescSet* dsin, DescSet* dout) {
  if (!bt->cover.on || wrap_ok(bt, dsin)) return false;
  if (!level_ok(bt, dsin) || !level_ok(bt, dout) || (ds1 && !level_ok(bt, ds1))) return false;
  if (sch.rhs_source && !(bt->src && level_ok(bt, bt->src.get()))) return false;
  return true;
}

// Level tiles for a fused stage where they apply and the tiling in force does
// not ask for box tiles (the autotuner times both: level tiles compute whole
// tile columns where the union covers only part of them).
bool use_level_tiles(amr_boxtab_s* bt, const amr_scheme_t& sch, DescSet* ds1, DescSet* dsin, DescSet* dout) {
  return !choose_config(bt, sch).box_tiles && level_possible(bt, sch, ds1, dsin, dout);
}

// Whether every box's array of `ds` has the same strides, and whether all of
// them lie within 2^31 elements of each other (the fused kernel's direct
// reads address a neighbour's cells by a 32-bit element offset from the own
// box's array).
void desc_layout(DescSet* ds) {
  if (ds->ustride < 0) {
    ds->ustride = 1;
    intptr_t lo = INTPTR_MAX, hi = 0;
    for (const amr_fab_t& f : ds->host) {
      if (f.jstride != ds->host[0].jstride || f.kstride != ds->host[0].kstride ||
          f.nstride != ds->host[0].nstride)
        ds->ustride = 0;
      const int64_t nx = f.hi[0] - f.lo[0] + 1 + 2 * f.ng, ny = f.hi[1] - f.lo[1] + 1 + 2 * f.ng;
      const int64_t nz = f.hi[2] - f.lo[2] + 1 + 2 * f.ng;
      const int64_t last = (nx - 1) + (ny - 1) * f.jstride + (nz - 1) * f.kstride + (NCONS - 1) * f.nstride;
      // the array's index origin (cell (0,0,0), which the kernels address
      // from) and its last element
      const int64_t org = -(int64_t(f.lo[0] - f.ng) + int64_t(f.lo[1] - f.ng) * f.jstride +
                            int64_t(f.lo[2] - f.ng) * f.kstride);
      lo = std::min({lo, intptr_t(f.data), intptr_t(f.data + org)});
      hi = std::max({hi, intptr_t(f.data + last + 1), intptr_t(f.data + org)});
    }
    ds->span32 = int64_t(hi - lo) / int64_t(sizeof(double)) < (int64_t(1) << 31) - 1 ? 1 : 0;
  }
}

// Whether the fused stages with tiles per box read the halo cells other
// boxes hold from those boxes directly (the stage inputs then need no
// same-level ghost copies): KEEP order 2 or 4 on a level with an owner map,
// stage inputs other than one periodic array over the domain (whose tiles
// read the periodic images of its own cells, which costs less).
bool nb_direct(amr_boxtab_s* bt, const amr_scheme_t& sch, bool level, DescSet* dsin) {
  desc_layout(dsin);
  return !level && bt->nbmap.on && sch.order_keep <= 4 && dsin->span32 == 1 &&
         fused_path(bt, sch, nullptr, level) && !wrap_ok(bt, dsin);
}

void free_user_source(amr_boxtab_s* bt) {
  if (bt->usrc_mod) (void)hipModuleUnload(bt->usrc_mod);
  if (bt->usrc_params) (void)hipFree(bt->usrc_params);
  if (bt->usrc_tmp) (void)hipFree(bt->usrc_tmp);
  if (bt->usrc_tmp_fabs) (void)hipFree(bt->usrc_tmp_fabs);
  bt->usrc_mod = nullptr;
  bt->usrc_fn = nullptr;
  bt->usrc_params = nullptr;
  bt->usrc_tmp = nullptr;
  bt->usrc_tmp_fabs = nullptr;
}

// A stage whose source is the run-time compiled user_source: dSdt assembled
// by the pass-structured kernels into a buffer, the user's source added to
// it cell by cell, then the stage combination (CNS::compute_rhs calls
// user_source after the assembly, src/tim/CNS_advance.cpp:375-407).
int user_source_stage(amr_boxtab_s* bt, const amr_closures_t& cls, const amr_scheme_t& sch, const StageArgs& sa_in,
                      DescSet* ds1, DescSet* dsin, DescSet* dout, hipStream_t st) {
  int rc;
  const int nb = bt->bt.nbox;
  if (!bt->usrc_tmp) {
    int64_t total = 0;
    for (auto& B : bt->bt.h_boxes) total += compact_len(B.lo, B.hi, 0, NCONS);
    HIPCHK(hipMalloc(&bt->usrc_tmp, sizeof(double) * std::max<int64_t>(total, 1)));
    std::vector<Fab> h(nb);
    int64_t off = 0;
    for (int b = 0; b < nb; ++b) {
      const BoxDev& B = bt->bt.h_boxes[b];
      h[b] = make_fab(compact_fab(B.lo, B.hi, 0, bt->usrc_tmp + off));
      off += compact_len(B.lo, B.hi, 0, NCONS);
    }
    HIPCHK(hipMalloc(&bt->usrc_tmp_fabs, sizeof(Fab) * nb));
    HIPCHK(hipMemcpy(bt->usrc_tmp_fabs, h.data(), sizeof(Fab) * nb, hipMemcpyHostToDevice));
  }
  if ((rc = ensure_unfused(bt, sch))) return rc;
  StageArgs sa = sa_in;
  sa.src = nullptr;
  sa.bfx = 0.0;
  sa.rhs_only = 1;
  HIPCHK(launch_unfused_stage(bt->bt, bt->uw, make_closures(cls), sch, sa, bt->wall_args, nullptr, dsin->d,
                              bt->usrc_tmp_fabs, st));
  int64_t mx = 1;
  for (auto& B : bt->bt.h_boxes) mx = std::max<int64_t>(mx, cells(B.lo, B.hi));
  const unsigned gx = unsigned(std::min<int64_t>((mx + 255) / 256, 1024));
  const BoxDev* boxes = bt->bt.d_boxes;
  const Fab* sfab = dsin->d;
  const Fab* rfab = bt->usrc_tmp_fabs;
  const double* pp = bt->usrc_params;
  amr_closures_t c = cls;
  const double* bdt = sa_in.box_dt;
  void* args[] = {&boxes, &sfab, &rfab, &pp, &c, &bdt};
  HIPCHK(hipModuleLaunchKernel(bt->usrc_fn, gx, unsigned(nb), 1, 256, 1, 1, 0, st, args, nullptr));
  HIPCHK(launch_combine(bt->bt, sa_in, ds1 ? ds1->d : nullptr, dsin->d, bt->usrc_tmp_fabs, dout->d, st));
  return AMR_OK;
}

int do_stage(amr_boxtab_s* bt, const amr_closures_t& cls, const amr_scheme_t& sch, const StageArgs& sa_in,
             DescSet* ds1, DescSet* dsin, DescSet* dout, const Fab* uncov_u0 = nullptr,
             hipStream_t stream = nullptr) {
  if (sch.rhs_source && bt->usrc_fn) {
    if (uncov_u0) return fail(AMR_ERR_ARG, "internal: uncovered ghost cells from U^n with a user source");
    return user_source_stage(bt, cls, sch, sa_in, ds1, dsin, dout, stream ? stream : bt->ctx->stream);
  }
  StageArgs sa = sa_in;
  if (sch.rhs_source) {
    const bool body = bt->has_problem && bt->problem.body_force_x != 0.0;
    if (!bt->src && !body)
      return fail(AMR_ERR_ARG, "rhs_source=1 needs a source (amr_set_source) or a body force (amr_set_problem)");
    sa.src = bt->src ? bt->src->d : nullptr;
    sa.bfx = body ? bt->problem.body_force_x : 0.0;
  }
  Closures c = make_closures(cls);
  hipStream_t st = stream ? stream : bt->ctx->stream;
  const Fab* s1 = ds1 ? ds1->d : nullptr;
  FusedConfig cfg;
  const bool level = use_level_tiles(bt, sch, ds1, dsin, dout);
  if (fused_path(bt, sch, &cfg, level)) {
    TileTable* tt = nullptr;
    int rc = get_tiles(bt, cfg, level, &tt);
    if (rc) return rc;
    const WrapArgs wa = wrap_args(bt, dsin);
    UncovSrc us{};
    if (uncov_u0) {
      if (!(bt->srcmap.on && !wa.on && tt->max_zblk <= UNCOV_MAX_ZBLK))
        return fail(AMR_ERR_ARG, "internal: uncovered ghost cells from U^n requested where the kernel cannot");
      us = bt->srcmap;
      us.u0 = uncov_u0;
    }
    const Cover cov = kernel_cover(bt, level);
    NbMap nbm = !uncov_u0 && nb_direct(bt, sch, level, dsin) ? bt->nbmap : NbMap{};
    nbm.ustride = nbm.on && dsin->ustride == 1 ? 1 : 0;
    HIPCHK(launch_fused_stage(sch, cfg, wa, cov, us, tt->d, tt->n, bt->bt.d_boxes, c, sa, s1, dsin->d, dout->d,
                              st, nbm));
    return AMR_OK;
  }
  if (uncov_u0) return fail(AMR_ERR_ARG, "internal: uncovered ghost cells from U^n requested off the fused path");
  int rc = ensure_unfused(bt, sch);
  if (rc) return rc;
  HIPCHK(launch_unfused_stage(bt->bt, bt->uw, c, sch, sa, bt->wall_args, s1, dsin->d, dout->d, st));
  return AMR_OK;
}

// A state array with the layout of `state`: a level-wide allocation stays
// level-wide (so its ghost fill stays cheap), separate FABs become compact
// FABs.  *mem receives the allocation.
int alloc_mirror(amr_boxtab_s* bt, const amr_fab_t* state, double** mem, std::vector<amr_fab_t>& fabs) {
  int nb = bt->bt.nbox;
  fabs.assign(nb, amr_fab_t());
  intptr_t lo = INTPTR_MAX, hi = 0;
  int64_t sum = 0;
  for (int b = 0; b < nb; ++b) {
    const amr_fab_t& f = state[b];
    int64_t nx = f.hi[0] - f.lo[0] + 1 + 2 * f.ng, ny = f.hi[1] - f.lo[1] + 1 + 2 * f.ng;
    int64_t nz = f.hi[2] - f.lo[2] + 1 + 2 * f.ng;
    int64_t last = (nx - 1) + (ny - 1) * f.jstride + (nz - 1) * f.kstride + (NCONS - 1) * f.nstride;
    intptr_t a = intptr_t(f.data), e = intptr_t(f.data + last + 1);
    lo = std::min(lo, a);
    hi = std::max(hi, e);
    sum += compact_len(f.lo, f.hi, f.ng, NCONS);
  }
  int64_t span = int64_t(hi - lo) / 8;
  // windows of one array (every box maps cell (0,0,0) to the same address)
  // are mirrored whatever the span: their ghost cells are neighbours' cells
  bool shared = true;
  const Mapping m0 = mapping_of(state[0]);
  for (int b = 1; b < nb && shared; ++b) shared = mapping_of(state[b]) == m0;
  bool mirror = shared || span <= sum + sum / 2;
  if (mirror) {
    // the mirror keeps each element's offset within 128-byte lines (a
    // padded level array's rows stay line-aligned in the workspaces)
    const int64_t shift = int64_t(lo & 127) / 8;
    HIPCHK(hipMalloc(mem, sizeof(double) * (span + shift)));
    for (int b = 0; b < nb; ++b) {
      fabs[b] = state[b];
      fabs[b].data = *mem + shift + (intptr_t(state[b].data) - lo) / 8;
    }
  } else {
    HIPCHK(hipMalloc(mem, sizeof(double) * sum));
    int64_t off = 0;
    for (int b = 0; b < nb; ++b) {
      const amr_fab_t& f = state[b];
      fabs[b] = compact_fab(f.lo, f.hi, f.ng, *mem + off);
      off += compact_len(f.lo, f.hi, f.ng, NCONS);
    }
  }
  return AMR_OK;
}

// Workspace state arrays (two) mirroring the layout of `state`.  The key is
// the layout relative to the lowest address of the state's arrays, so states
// of one layout (several level-wide arrays of a level, say) share the
// workspaces; another layout reallocates them.
int ensure_ws(amr_boxtab_s* bt, const amr_fab_t* state) {
  int nb = bt->bt.nbox;
  intptr_t lo = INTPTR_MAX;
  for (int b = 0; b < nb; ++b) lo = std::min(lo, intptr_t(state[b].data));
  std::vector<amr_fab_t> key(state, state + nb);
  for (auto& f : key) f.data = (double*)(intptr_t(f.data) - lo);
  if (int(bt->ws_key.size()) == nb && std::memcmp(bt->ws_key.data(), key.data(), sizeof(amr_fab_t) * nb) == 0)
    return AMR_OK;
  drop_graphs(bt);  // captured steps write the workspaces being replaced
  HIPCHK(hipStreamSynchronize(bt->ctx->stream));
  for (int w = 0; w < 2; ++w) {
    if (bt->ws_mem[w]) HIPCHK(hipFree(bt->ws_mem[w]));
    bt->ws_mem[w] = nullptr;
  }
  bt->ws_key.clear();
  for (int w = 0; w < 2; ++w) {
    int rc = alloc_mirror(bt, state, &bt->ws_mem[w], bt->ws_fabs[w]);
    if (rc) return rc;
  }
  bt->ws_key = std::move(key);
  return AMR_OK;
}

int stage_coeffs(int order_rk, int stages_rk, double dt, int stage, int* nst, StageCoef* sc) {
  if (!stage_coef_of(order_rk, stages_rk, dt, stage, sc, nst))
    return fail(AMR_ERR_ARG, "unsupported order_rk/stages_rk combination or stage index");
  return AMR_OK;
}

// Whether the stages of a step from `dst` (U^n) through the workspaces
// dw[0..1] after the first can read the ghost cells no box covers straight
// from U^n (UncovSrc) instead of a copy in their inputs: every stage runs the
// fused kernel with one kind of tiles, whose loaded planes span at most
// UNCOV_MAX_ZBLK blocks of the map, and each box's arrays in dst and the
// workspaces have equal strides.
int uncovered_from_un(amr_boxtab_s* bt, const amr_scheme_t& sch, DescSet* dst, DescSet* const dw[2], bool* ok) {
  *ok = false;
  if (!bt->srcmap.on || wrap_ok(bt, dst)) return AMR_OK;
  const bool level = use_level_tiles(bt, sch, dst, dst, dw[0]);
  if (level && !level_ok(bt, dw[1])) return AMR_OK;
  // stages that read the other boxes' cells directly read uncovered ghost
  // cells from their own arrays (the copy pass fills them)
  if (nb_direct(bt, sch, level, dw[0]) || nb_direct(bt, sch, level, dw[1])) return AMR_OK;
  FusedConfig cfg;
  if (!fused_path(bt, sch, &cfg, level)) return AMR_OK;
  TileTable* tt = nullptr;
  int rc = get_tiles(bt, cfg, level, &tt);
  if (rc) return rc;
  if (tt->max_zblk > UNCOV_MAX_ZBLK) return AMR_OK;
  for (int b = 0; b < bt->bt.nbox; ++b) {
    const amr_fab_t& f = dst->host[b];
    for (int w = 0; w < 2; ++w) {
      const amr_fab_t& g = dw[w]->host[b];
      if (g.jstride != f.jstride || g.kstride != f.kstride || g.nstride != f.nstride) return AMR_OK;
    }
  }
  *ok = true;
  return AMR_OK;
}

// One step (CNS::advance).  scratch_out: the autotuner's timed step, which
// leaves `state` untouched: no ghost fill of `state`, and the last stage
// writes a workspace instead of `state`.
int advance_impl(amr_boxtab_s* bt, const amr_closures_t& cls, const amr_scheme_t& sch, const amr_fab_t* state,
                 double dt, const double* d_box_dt = nullptr, bool scratch_out = false) {
  int rc;
  int nst = 0;
  StageCoef sc;
  if ((rc = stage_coeffs(sch.order_rk, sch.stages_rk, dt, 0, &nst, &sc))) return rc;
  if ((rc = ensure_ws(bt, state))) return rc;
  DescSet *dst = nullptr, *dw[2] = {nullptr, nullptr};
  if ((rc = get_desc(bt, state, NGHOST, &dst))) return rc;
  for (int w = 0; w < 2; ++w)
    if ((rc = get_desc(bt, bt->ws_fabs[w].data(), NGHOST, &dw[w]))) return rc;
  if ((rc = build_fill(bt, dst))) return rc;
  for (int w = 0; w < 2; ++w)
    if ((rc = build_fill(bt, dw[w]))) return rc;
  // the fused stage kernel (KEEP 2/4, halo 2) reads two ghost layers: its
  // stage inputs are filled that far only
  const bool narrow = fused_path(bt, sch, nullptr) && sch.order_keep <= 4;
  if (narrow) {
    if ((rc = build_fill_narrow(bt, dst))) return rc;
    for (int w = 0; w < 2; ++w)
      if ((rc = build_fill_narrow(bt, dw[w]))) return rc;
  }
  // stage inputs keep the caller's values in ghost cells no box covers:
  // read from U^n by the fused stages, else copied into the workspaces
  bool direct = false;
  if (nst > 1 && bt->nuncov_cells > 0 && (rc = uncovered_from_un(bt, sch, dst, dw, &direct))) return rc;
  if (!direct && (rc = copy_uncovered(bt, dst, dw[0], dw[1]))) return rc;
  for (int s = 0; s < nst; ++s) {
    if ((rc = stage_coeffs(sch.order_rk, sch.stages_rk, dt, s, &nst, &sc))) return rc;
    DescSet* sin = s == 0 ? dst : dw[(s - 1) % 2];
    DescSet* out = (s == nst - 1 && s > 0 && !scratch_out) ? dst : dw[s % 2];
    // a periodic domain-wide array needs no fill on the fused path; the
    // directional flux passes read the periodic images of a batch's problems
    // themselves
    const bool fused = fused_path(bt, sch, nullptr);
    // nor do fused stages whose tiles read the other boxes' cells directly
    // (do_stage's choice: not with uncovered ghost cells read from U^n), nor
    // the inviscid directional passes where they read x ends through the
    // side tables
    const Fab* u0 = direct && s > 0 ? dst->d : nullptr;
    const bool direct_nb = (!u0 && nb_direct(bt, sch, use_level_tiles(bt, sch, dst, sin, out), sin)) ||
                           (!fused && dir_flux_scheme(sch) && bt->d_xside != nullptr && !bt->walls &&
                            !(sch.rhs_source && bt->usrc_fn));
    if (!(fused && wrap_ok(bt, sin)) && !direct_nb && !(scratch_out && s == 0))
      if ((rc = enqueue_fill(bt, sin, bt->batch && !fused && dir_flux_scheme(sch), nullptr, narrow && fused ? 2 : 0)))
        return rc;
    StageArgs sa{sc, d_box_dt, sch.order_rk, sch.stages_rk, s};
    if ((rc = do_stage(bt, cls, sch, sa, dst, sin, out, u0))) return rc;
  }
  if (nst == 1 && !scratch_out) {
    // single-stage schemes wrote into workspace 0; copy its valid cells back
    HIPCHK(launch_copy_valid(bt->bt, dw[0]->d, dst->d, d_box_dt, bt->ctx->stream));
  }
  return AMR_OK;
}

// Every table and workspace a step of `sch` on `state` uses, built before a
// graph capture (a plain step on a scratch copy would change the state, so
// the tables are built directly): descriptor sets with their fill tasks,
// tile tables for box and level tiles, or the pass-structured workspace.
int prepare_step_tables(amr_boxtab_s* bt, const amr_scheme_t& sch, const amr_fab_t* state) {
  int rc;
  if ((rc = ensure_ws(bt, state))) return rc;
  DescSet* ds = nullptr;
  if ((rc = get_desc(bt, state, NGHOST, &ds)) || (rc = build_fill(bt, ds))) return rc;
  // the two-layer fills of the fused stages (as advance_impl decides)
  const bool narrow = fused_path(bt, sch, nullptr) && sch.order_keep <= 4;
  if (narrow && (rc = build_fill_narrow(bt, ds))) return rc;
  for (int w = 0; w < 2; ++w) {
    DescSet* dw = nullptr;
    if ((rc = get_desc(bt, bt->ws_fabs[w].data(), NGHOST, &dw)) || (rc = build_fill(bt, dw))) return rc;
    if (narrow && (rc = build_fill_narrow(bt, dw))) return rc;
  }
  FusedConfig cfg = choose_config(bt, sch);
  const bool fused = cfg.tx > 0 && fused_supported(sch, cfg.tx, cfg.ty) && !bt->walls;
  TileTable* tt = nullptr;
  if (fused && fused_fits(bt, cfg) && (rc = get_tiles(bt, cfg, false, &tt))) return rc;
  if (fused && bt->cover.on && (rc = get_tiles(bt, cfg, true, &tt))) return rc;
  if (!(fused && fused_fits(bt, cfg)) && (rc = ensure_unfused(bt, sch))) return rc;
  return AMR_OK;
}

// One step under device-side time-step control, enqueued only: the per-box
// CFL minima, the time-step rule into the per-box steps, then the step.
int cfl_step(amr_boxtab_s* bt, const amr_closures_t& cls, const amr_scheme_t& sch, const amr_fab_t* state,
             double cfl, double stop_time) {
  int rc;
  DescSet* ds = nullptr;
  if ((rc = get_desc(bt, state, NGHOST, &ds))) return rc;
  hipStream_t st = bt->ctx->stream;
  // per-box partial minima straight into the policy kernel (no separate
  // per-box reduction)
  HIPCHK(launch_estdt_boxes(bt->bt, ds->d, make_closures(cls), bt->d_partial, nullptr, st, bt->d_clock, bt->batch,
                            stop_time));
  HIPCHK(launch_dt_policy(bt->bt.nbox, bt->batch, bt->d_partial, estdt_box_blocks(bt->bt), bt->d_clock, cfl,
                          stop_time, bt->d_box_dt, st));
  return advance_impl(bt, cls, sch, state, 1.0, bt->d_box_dt);
}

int nprob_of(const amr_boxtab_s* bt) { return bt->batch ? bt->bt.nbox : 1; }

// Non-physical cells of `state` (NaN or infinite components, rho <= 0 or
// p <= 0): AMR_ERR_STATE naming the lowest such box.  Synchronous.
int check_state_impl(amr_boxtab_s* bt, const amr_closures_t& cls, const amr_fab_t* state, int32_t* bad_box) {
  int rc;
  DescSet* ds = nullptr;
  if ((rc = get_desc(bt, state, 0, &ds))) return rc;
  hipStream_t st = bt->ctx->stream;
  const int nb = bt->bt.nbox;
  HIPCHK(hipMemcpyAsync(bt->d_flag, &nb, sizeof(int), hipMemcpyHostToDevice, st));
  HIPCHK(launch_state_check(bt->bt, ds->d, make_closures(cls), bt->d_flag, st));
  int first = nb;
  HIPCHK(hipMemcpyAsync(&first, bt->d_flag, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (bad_box) *bad_box = first < nb ? first : -1;
  if (first < nb)
    return fail(AMR_ERR_STATE, "non-physical state in box " + std::to_string(first) +
                                   " (NaN or infinite value, rho <= 0 or p <= 0)");
  return AMR_OK;
}

// the check amr_set_state_check asks for, after a step call
int after_step(amr_boxtab_s* bt, const amr_closures_t& cls, const amr_fab_t* state, int rc) {
  if (rc || !bt->state_check) return rc;
  return check_state_impl(bt, cls, state, nullptr);
}

// ------------------------------------------------------------ adjoint
constexpr int ADJ_NE = 16;

void free_fold(amr_boxtab_s* bt) {
  for (void* p : bt->fold_mem) (void)hipFree(p);
  bt->fold_mem.clear();
  bt->fold = FoldTasks{};
}

// The transposed ghost fill of the adjoint workspace without atomics:
// physical-boundary tasks split by side, one lane per line along the
// direction; copy tasks greedily coloured so that no two tasks of a class
// read overlapping source cells of one box (in task order, so the classes
// and their launch order are fixed).
int build_fold(amr_boxtab_s* bt, const DescSet* g) {
  free_fold(bt);
  FoldTasks& f = bt->fold;
  auto upload = [&](const void* h, size_t bytes, void** d) -> int {
    HIPCHK(hipMalloc(d, bytes));
    bt->fold_mem.push_back(*d);
    HIPCHK(hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice));
    return AMR_OK;
  };
  int rc;
  for (int d = 0; d < 3; ++d)
    for (int side = 0; side < 2; ++side) {
      std::vector<BcTask> t;
      int64_t n = 0;
      for (const BcTask& b : g->h_bc[d]) {
        if (b.side != side) continue;
        BcTask c = b;
        c.start = n;
        n += cells(b.lo, b.hi) / (b.hi[d] - b.lo[d] + 1);
        t.push_back(c);
      }
      f.nbc[d][side] = int(t.size());
      f.nlines[d][side] = n;
      f.bc[d][side] = nullptr;
      if (!t.empty() && (rc = upload(t.data(), sizeof(BcTask) * t.size(), (void**)&f.bc[d][side]))) return rc;
    }
  const std::vector<CopyTask>& cp = g->h_copy;
  std::vector<int> cls(cp.size(), -1);
  int ncls = 0;
  for (size_t i = 0; i < cp.size(); ++i) {
    unsigned used = 0;
    for (size_t j = 0; j < i; ++j) {
      if (cp[j].src != cp[i].src) continue;
      bool overlap = true;
      for (int d = 0; d < 3; ++d)
        if (cp[i].hi[d] - cp[i].sh[d] < cp[j].lo[d] - cp[j].sh[d] ||
            cp[j].hi[d] - cp[j].sh[d] < cp[i].lo[d] - cp[i].sh[d])
          overlap = false;
      if (overlap) used |= 1u << cls[j];
    }
    int c = 0;
    while (c < FOLD_MAX_CLASSES && ((used >> c) & 1u)) ++c;
    if (c >= FOLD_MAX_CLASSES)
      return fail(AMR_ERR_ARG, "reverse mode: a cell is the source of more than 32 ghost copies");
    cls[i] = c;
    ncls = std::max(ncls, c + 1);
  }
  f.nclass = ncls;
  for (int c = 0; c < ncls; ++c) {
    std::vector<CopyTask> t;
    int64_t n = 0;
    for (size_t i = 0; i < cp.size(); ++i)
      if (cls[i] == c) {
        CopyTask x = cp[i];
        x.start = n;
        n += cells(x.lo, x.hi);
        t.push_back(x);
      }
    f.class_tasks[c] = int(t.size());
    f.class_cells[c] = n;
    if ((rc = upload(t.data(), sizeof(CopyTask) * t.size(), (void**)&f.classes[c]))) return rc;
  }
  return AMR_OK;
}

int ensure_adjoint(amr_boxtab_s* bt, const amr_scheme_t& s) {
  const int visc = s.rhs_visc ? 1 : 0;
  if (bt->adj_mem && bt->adj_visc >= visc) return AMR_OK;
  HIPCHK(hipStreamSynchronize(bt->ctx->stream));
  free_fold(bt);
  bt->adj_g.reset();
  if (bt->adj_e_table) HIPCHK(hipFree(bt->adj_e_table));
  if (bt->adj_mem) HIPCHK(hipFree(bt->adj_mem));
  bt->adj_e_table = nullptr;
  bt->adj_mem = nullptr;
  bt->aw = AdjWork();
  const int nb = bt->bt.nbox;
  int64_t total = 0;
  for (int b = 0; b < nb; ++b) {
    const BoxDev& B = bt->bt.h_boxes[b];
    total += compact_len(B.lo, B.hi, NGHOST, NCONS);
    if (visc) total += compact_len(B.lo, B.hi, 1, ADJ_NE);
  }
  HIPCHK(hipMalloc(&bt->adj_mem, sizeof(double) * total));
  std::vector<amr_fab_t> gf(nb);
  std::vector<Fab> he(nb);
  int64_t off = 0;
  for (int b = 0; b < nb; ++b) {
    const BoxDev& B = bt->bt.h_boxes[b];
    gf[b] = compact_fab(B.lo, B.hi, NGHOST, bt->adj_mem + off);
    off += compact_len(B.lo, B.hi, NGHOST, NCONS);
  }
  if (visc)
    for (int b = 0; b < nb; ++b) {
      const BoxDev& B = bt->bt.h_boxes[b];
      he[b] = make_fab(compact_fab(B.lo, B.hi, 1, bt->adj_mem + off));
      off += compact_len(B.lo, B.hi, 1, ADJ_NE);
    }
  int rc = new_desc(bt, gf.data(), &bt->adj_g);
  if (rc) return rc;
  if ((rc = build_fill(bt, bt->adj_g.get()))) return rc;
  if ((rc = build_fold(bt, bt->adj_g.get()))) return rc;
  bt->aw.g = bt->adj_g->d;
  if (visc) {
    HIPCHK(hipMalloc(&bt->adj_e_table, sizeof(Fab) * nb));
    HIPCHK(hipMemcpy(bt->adj_e_table, he.data(), sizeof(Fab) * nb, hipMemcpyHostToDevice));
    bt->aw.e = bt->adj_e_table;
  }
  bt->adj_visc = visc;
  return AMR_OK;
}

// Descriptor of one array over the level's domain (no ghost cells) with nc
// components at `base`.
amr_fab_t domain_fab(const amr_boxtab_s* bt, double* base) {
  return compact_fab(bt->geom.dom_lo, bt->geom.dom_hi, 0, base);
}

// Domain-wide primitives and viscous pre-pass arrays for the reverse stage of
// a periodic level held as one array.
int ensure_adjoint_level(amr_boxtab_s* bt, const amr_scheme_t& s) {
  const bool visc = s.rhs_visc != 0;
  if (bt->adj_lmem && (bt->adj_le_table || !visc)) {
    bt->aw.lq = bt->adj_lq_table;
    bt->aw.le = bt->adj_le_table;
    return AMR_OK;
  }
  HIPCHK(hipStreamSynchronize(bt->ctx->stream));
  if (bt->adj_lq_table) HIPCHK(hipFree(bt->adj_lq_table));
  if (bt->adj_le_table) HIPCHK(hipFree(bt->adj_le_table));
  if (bt->adj_lmem) HIPCHK(hipFree(bt->adj_lmem));
  bt->adj_lq_table = bt->adj_le_table = nullptr;
  bt->adj_lmem = nullptr;
  const int nb = bt->bt.nbox;
  const int64_t n = cells(bt->geom.dom_lo, bt->geom.dom_hi);
  HIPCHK(hipMalloc(&bt->adj_lmem, sizeof(double) * n * (NPRIM + (visc ? ADJ_NE : 0))));
  std::vector<Fab> h(nb, make_fab(domain_fab(bt, bt->adj_lmem)));
  HIPCHK(hipMalloc(&bt->adj_lq_table, sizeof(Fab) * nb));
  HIPCHK(hipMemcpy(bt->adj_lq_table, h.data(), sizeof(Fab) * nb, hipMemcpyHostToDevice));
  if (visc) {
    std::vector<Fab> he(nb, make_fab(domain_fab(bt, bt->adj_lmem + n * NPRIM)));
    HIPCHK(hipMalloc(&bt->adj_le_table, sizeof(Fab) * nb));
    HIPCHK(hipMemcpy(bt->adj_le_table, he.data(), sizeof(Fab) * nb, hipMemcpyHostToDevice));
  }
  bt->aw.lq = bt->adj_lq_table;
  bt->aw.le = bt->adj_le_table;
  return AMR_OK;
}

// Reverse of one stage: lam_sin (=|+=) stage terms of lam_out through the
// RHS VJP at `sin` (ghost-filled); lam_s1 gets a lam_out (s1_mode 1: +=,
// 2: =).  lam_sin and lam_s1 must not alias lam_out.
int adj_stage(amr_boxtab_s* bt, const amr_closures_t& cls, const amr_scheme_t& sch, const StageArgs& sa,
              DescSet* sin, DescSet* lam_out, const DescSet* lam_s1, const DescSet* lam_sin, bool acc,
              int s1_mode) {
  int rc;
  if ((rc = ensure_unfused(bt, sch))) return rc;
  if ((rc = ensure_adjoint(bt, sch))) return rc;
  const Closures c = make_closures(cls);
  const FoldTasks& ft = bt->fold;
  const WrapArgs wa = wrap_args(bt, sin);
  // the domain-wide path needs lam_out as one periodic array as well
  const bool level = wa.on && wrap_ok(bt, lam_out);
  if (level && (rc = ensure_adjoint_level(bt, sch))) return rc;
  AdjWork aw = bt->aw;
  if (!level) aw.lq = aw.le = nullptr;
  HIPCHK(launch_adj_stage(bt->bt, bt->uw, aw, c, sch, sa, wa, sin->d, lam_out->d,
                          lam_s1 ? lam_s1->d : nullptr, lam_sin->d, acc, lam_s1 ? s1_mode : 0, ft,
                          bt->ctx->stream));
  return AMR_OK;
}

// Whether the storage spans of two FAB sets overlap.
bool fabs_overlap(const amr_fab_t* a, const amr_fab_t* b, int nb) {
  // per-box address spans (box arrays of two sets may interleave in
  // memory, so one span per set would report overlaps that do not exist);
  // sorted by start, an overlap between the sets is a span of one set that
  // starts before the furthest end reached by the other set's earlier spans
  struct Span {
    intptr_t lo, hi;
    int set;
  };
  std::vector<Span> v;
  v.reserve(2 * size_t(nb));
  for (int set = 0; set < 2; ++set) {
    const amr_fab_t* f = set == 0 ? a : b;
    for (int i = 0; i < nb; ++i) {
      int64_t nx = f[i].hi[0] - f[i].lo[0] + 1 + 2 * f[i].ng, ny = f[i].hi[1] - f[i].lo[1] + 1 + 2 * f[i].ng;
      int64_t nz = f[i].hi[2] - f[i].lo[2] + 1 + 2 * f[i].ng;
      int64_t last = (nx - 1) + (ny - 1) * f[i].jstride + (nz - 1) * f[i].kstride + (NCONS - 1) * f[i].nstride;
      v.push_back(Span{intptr_t(f[i].data), intptr_t(f[i].data + last + 1), set});
    }
  }
  std::sort(v.begin(), v.end(), [](const Span& x, const Span& y) { return x.lo < y.lo; });
  intptr_t reach[2] = {INTPTR_MIN, INTPTR_MIN};
  for (const Span& x : v) {
    if (x.lo < reach[1 - x.set]) return true;
    reach[x.set] = std::max(reach[x.set], x.hi);
  }
  return false;
}

int check_adjoint(const amr_scheme_t& s) {
  if (!adjoint_supported(s))
    return fail(AMR_ERR_ARG,
                "the adjoint covers the KEEP (flux_euler=1) and viscous right-hand sides without JST only");
  return AMR_OK;
}

// Box-table conditions of the reverse mode: no bc 6 walls and no
// state-dependent (body force) source.
int check_adjoint_table(const amr_boxtab_s* bt, const amr_scheme_t& s) {
  if (bt->walls) return fail(AMR_ERR_ARG, "the adjoint does not cover bc code 6 walls");
  if (s.rhs_source && bt->has_problem && bt->problem.body_force_x != 0.0)
    return fail(AMR_ERR_ARG, "the adjoint does not cover the body-force source");
  return AMR_OK;
}

int ensure_run(amr_boxtab_s* bt, const amr_fab_t* state, int nsteps, int k, int nst) {
  const int nb = bt->bt.nbox;
  std::vector<char> key(sizeof(amr_fab_t) * nb + 3 * sizeof(int));
  std::memcpy(key.data(), state, sizeof(amr_fab_t) * nb);
  int tail[3] = {nsteps, k, nst};
  std::memcpy(key.data() + sizeof(amr_fab_t) * nb, tail, sizeof tail);
  if (bt->run && bt->run->key == key) return AMR_OK;
  HIPCHK(hipStreamSynchronize(bt->ctx->stream));
  bt->run.reset();
  auto R = std::make_unique<AdjRun>();
  R->key = key;
  R->nsteps = nsteps;
  R->k = k;
  R->nst = nst;
  R->nseg = (nsteps + k - 1) / k;
  R->krows = std::min(k, nsteps);
  const int nstates = R->nseg + 1 + (R->nseg > 1 ? 2 : 1) * (R->krows * nst - 1);
  int rc;
  for (int i = 0; i < nstates; ++i) {
    double* m = nullptr;
    std::vector<amr_fab_t> f;
    rc = alloc_mirror(bt, state, &m, f);
    if (m) R->mem.push_back(m);
    if (rc) return rc;
    std::unique_ptr<DescSet> ds;
    if ((rc = new_desc(bt, f.data(), &ds))) return rc;
    if ((rc = build_fill(bt, ds.get()))) return rc;
    R->states.push_back(std::move(ds));
  }
  // adjoint arrays: windows of one domain-wide array when the state is a
  // periodic level held as one array (the reverse stage then gathers across
  // boxes), else one compact array per box
  const bool level = wrap_ok(bt, R->states[0].get());
  int64_t per = 0;
  for (int b = 0; b < nb; ++b) per += compact_len(bt->bt.h_boxes[b].lo, bt->bt.h_boxes[b].hi, 0, NCONS);
  double* lm = nullptr;
  HIPCHK(hipMalloc(&lm, sizeof(double) * per * 3));
  R->mem.push_back(lm);
  for (int w = 0; w < 3; ++w) {
    std::vector<amr_fab_t> f(nb);
    int64_t off = 0;
    const amr_fab_t whole = domain_fab(bt, lm + w * per);
    for (int b = 0; b < nb; ++b) {
      const BoxDev& B = bt->bt.h_boxes[b];
      if (level) {
        f[b] = whole;
        for (int d = 0; d < 3; ++d) {
          f[b].lo[d] = B.lo[d];
          f[b].hi[d] = B.hi[d];
        }
        f[b].data = whole.data + (B.lo[0] - bt->geom.dom_lo[0]) + (B.lo[1] - bt->geom.dom_lo[1]) * whole.jstride +
                    (B.lo[2] - bt->geom.dom_lo[2]) * whole.kstride;
      } else {
        f[b] = compact_fab(B.lo, B.hi, 0, lm + w * per + off);
        off += compact_len(B.lo, B.hi, 0, NCONS);
      }
    }
    if ((rc = new_desc(bt, f.data(), &R->lam[w]))) return rc;
  }
  bt->run = std::move(R);
  return AMR_OK;
}

int check_geom(const amr_geom_t& g) {
  for (int d = 0; d < 3; ++d) {
    if (g.dom_hi[d] < g.dom_lo[d]) return fail(AMR_ERR_ARG, "empty domain");
    if (!(g.dx[d] > 0.0)) return fail(AMR_ERR_ARG, "dx must be positive");
    if (g.periodic[d] && g.dom_hi[d] - g.dom_lo[d] + 1 < NGHOST)
      return fail(AMR_ERR_ARG, "periodic direction shorter than the ghost width");
  }
  return AMR_OK;
}

int finish_boxtab(amr_boxtab_s* bt);

}  // namespace

// ======================================================================= ABI
namespace {

TagArgs tag_args(const amr_boxtab_s* bt) {
  const amr_geom_t& g = bt->geoms[0];
  TagArgs ta;
  std::memset(&ta, 0, sizeof ta);
  for (int d = 0; d < 3; ++d) {
    ta.dlo[d] = g.dom_lo[d];
    ta.n[d] = g.dom_hi[d] - g.dom_lo[d] + 1;
    ta.dx[d] = g.dx[d];
  }
  ta.prob_hi1 = g.prob_hi[1];
  return ta;
}

// Cells of box `nb` (of the new level) that no box of the old level holds, as
// rectangles: x-runs of each row, merged over consecutive rows and planes
// with the same extents.
void uncovered_by(const BoxDev& nb, int dst, const std::vector<BoxDev>* old, std::vector<CopyTask>& out) {
  const int nx = nb.hi[0] - nb.lo[0] + 1, ny = nb.hi[1] - nb.lo[1] + 1, nz = nb.hi[2] - nb.lo[2] + 1;
  std::vector<unsigned char> cov(size_t(nx) * ny * nz, 0);
  if (old)
    for (const BoxDev& o : *old) {
      int lo[3], hi[3];
      bool empty = false;
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::max(nb.lo[d], o.lo[d]);
        hi[d] = std::min(nb.hi[d], o.hi[d]);
        empty = empty || lo[d] > hi[d];
      }
      if (empty) continue;
      for (int k = lo[2]; k <= hi[2]; ++k)
        for (int j = lo[1]; j <= hi[1]; ++j)
          for (int i = lo[0]; i <= hi[0]; ++i)
            cov[(size_t(k - nb.lo[2]) * ny + (j - nb.lo[1])) * nx + (i - nb.lo[0])] = 1;
    }
  // runs of each row, then rectangles grown over j (keyed by plane and
  // x-extent), then over k
  struct Rect {
    int lo[3], hi[3];
  };
  std::vector<Rect> plane_rects;
  std::vector<Rect> done;
  for (int k = nb.lo[2]; k <= nb.hi[2]; ++k) {
    std::map<std::pair<int, int>, Rect> open;
    for (int j = nb.lo[1]; j <= nb.hi[1]; ++j) {
      std::map<std::pair<int, int>, Rect> next;
      const unsigned char* row = &cov[(size_t(k - nb.lo[2]) * ny + (j - nb.lo[1])) * nx];
      for (int i = 0; i < nx;) {
        if (row[i]) {
          ++i;
          continue;
        }
        int e = i;
        while (e + 1 < nx && !row[e + 1]) ++e;
        const std::pair<int, int> key{nb.lo[0] + i, nb.lo[0] + e};
        auto it = open.find(key);
        Rect r;
        if (it != open.end()) {
          r = it->second;
          r.hi[1] = j;
          open.erase(it);
        } else {
          r = Rect{{key.first, j, k}, {key.second, j, k}};
        }
        next[key] = r;
        i = e + 1;
      }
      for (auto& kv : open) plane_rects.push_back(kv.second);
      open.swap(next);
    }
    for (auto& kv : open) plane_rects.push_back(kv.second);
  }
  // merge plane rectangles with equal x/y extents over consecutive k
  std::map<std::array<int, 4>, Rect> zopen;
  std::sort(plane_rects.begin(), plane_rects.end(), [](const Rect& a, const Rect& b) { return a.lo[2] < b.lo[2]; });
  for (const Rect& r : plane_rects) {
    const std::array<int, 4> key{r.lo[0], r.hi[0], r.lo[1], r.hi[1]};
    auto it = zopen.find(key);
    if (it != zopen.end() && it->second.hi[2] == r.lo[2] - 1) {
      it->second.hi[2] = r.hi[2];
    } else {
      if (it != zopen.end()) done.push_back(it->second);
      zopen[key] = r;
    }
  }
  for (auto& kv : zopen) done.push_back(kv.second);
  for (const Rect& r : done) {
    CopyTask t;
    t.dst = t.src = dst;
    for (int d = 0; d < 3; ++d) {
      t.lo[d] = r.lo[d];
      t.hi[d] = r.hi[d];
      t.sh[d] = 0;
    }
    t.start = 0;
    out.push_back(t);
  }
}

template <class T>
struct DevBuf {
  T* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

extern "C" {

const char* amr_last_error(void) { return g_err.c_str(); }

namespace {
// Device address of the no-allocation scratch array of each GPU, taken when
// the first context on that GPU is created, with that GPU current (each GPU
// holds its own copy of the array at its own address).
constexpr int SCRATCH_MAX_DEV = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[SCRATCH_MAX_DEV] = {};
}  // namespace

int amr_ctx_create(int device, void* stream, amr_ctx_t* out) {
  if (!out) return fail(AMR_ERR_ARG, "null output pointer");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(AMR_ERR_ARG, "device index out of range");
  HIPCHK(hipSetDevice(device));
  void* scr = nullptr;
  if (device < SCRATCH_MAX_DEV) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[device]) HIPCHK(noalloc_scratch_address(&g_scratch_dev[device]));
    scr = g_scratch_dev[device];
  }
  auto* c = new amr_ctx_s();
  c->device = device;
  c->scratch = scr;
  if (stream) {
    c->stream = reinterpret_cast<hipStream_t>(stream);
  } else {
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      return fail(AMR_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    c->own_stream = true;
  }
  *out = c;
  return AMR_OK;
}

namespace {
void release_ctx(amr_ctx_t ctx) {
  if (ctx->refs.fetch_sub(1) != 1) return;
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}
}  // namespace

int amr_ctx_destroy(amr_ctx_t ctx) {
  if (!ctx) return AMR_OK;
  release_ctx(ctx);
  return AMR_OK;
}

int amr_ctx_synchronize(amr_ctx_t ctx) {
  if (!ctx) return fail(AMR_ERR_ARG, "null context");
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return AMR_OK;
}

int amr_boxtab_create(amr_ctx_t ctx, const amr_geom_t* geom, int nbox, const int32_t* boxes, amr_boxtab_t* out) {
  if (!ctx || !geom || !boxes || !out || nbox <= 0) return fail(AMR_ERR_ARG, "invalid box table arguments");
  {
    int rc = check_geom(*geom);
    if (rc) return rc;
  }
  for (int b = 0; b < nbox; ++b)
    for (int d = 0; d < 3; ++d) {
      int lo = boxes[6 * b + d], hi = boxes[6 * b + 3 + d];
      if (hi < lo) return fail(AMR_ERR_ARG, "box " + std::to_string(b) + " is empty");
      if (lo < geom->dom_lo[d] || hi > geom->dom_hi[d])
        return fail(AMR_ERR_ARG, "box " + std::to_string(b) + " leaves the domain");
    }
  HIPCHK(hipSetDevice(ctx->device));
  auto* bt = new amr_boxtab_s();
  static std::atomic<uint64_t> next_serial{0};
  bt->serial = ++next_serial;
  bt->ctx = ctx;
  ctx->refs.fetch_add(1);
  bt->geom = *geom;
  bt->bt.nbox = nbox;
  bt->bt.h_boxes.resize(nbox);
  bt->geoms.assign(nbox, *geom);
  for (int b = 0; b < nbox; ++b)
    for (int d = 0; d < 3; ++d) {
      bt->bt.h_boxes[b].lo[d] = boxes[6 * b + d];
      bt->bt.h_boxes[b].hi[d] = boxes[6 * b + 3 + d];
    }
  int rc = finish_boxtab(bt);
  if (rc) {
    amr_boxtab_destroy(bt);
    return rc;
  }
  *out = bt;
  return AMR_OK;
}

int amr_batch_create(amr_ctx_t ctx, int nprob, const amr_geom_t* geoms, amr_boxtab_t* out) {
  if (!ctx || !geoms || !out || nprob <= 0) return fail(AMR_ERR_ARG, "invalid batch arguments");
  for (int b = 0; b < nprob; ++b) {
    int rc = check_geom(geoms[b]);
    if (rc) return rc;
  }
  HIPCHK(hipSetDevice(ctx->device));
  auto* bt = new amr_boxtab_s();
  bt->ctx = ctx;
  ctx->refs.fetch_add(1);
  bt->geom = geoms[0];
  bt->batch = true;
  bt->bt.nbox = nprob;
  bt->bt.h_boxes.resize(nprob);
  bt->geoms.assign(geoms, geoms + nprob);
  for (int b = 0; b < nprob; ++b)
    for (int d = 0; d < 3; ++d) {
      bt->bt.h_boxes[b].lo[d] = geoms[b].dom_lo[d];
      bt->bt.h_boxes[b].hi[d] = geoms[b].dom_hi[d];
    }
  int rc = finish_boxtab(bt);
  if (rc) {
    amr_boxtab_destroy(bt);
    return rc;
  }
  *out = bt;
  return AMR_OK;
}

int amr_batch_advance(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch, const amr_fab_t* state,
                      const double* dt) {
  if (!bt || !cls || !sch || !state || !dt) return fail(AMR_ERR_ARG, "null argument");
  int rc;
  if ((rc = check_scheme(*sch)) || (rc = check_closures(*cls))) return rc;
  int nb = bt->bt.nbox;
  bool any = false;
  for (int b = 0; b < nb; ++b) {
    if (std::isnan(dt[b]) || dt[b] < 0.0) return fail(AMR_ERR_ARG, "time steps must be >= 0");
    any = any || dt[b] > 0.0;
  }
  if (!any) return AMR_OK;
  hipStream_t st = bt->ctx->stream;
  HIPCHK(hipEventSynchronize(bt->dt_copied));  // staging buffer free again
  std::memcpy(bt->h_box_dt, dt, sizeof(double) * nb);
  HIPCHK(hipMemcpyAsync(bt->d_box_dt, bt->h_box_dt, sizeof(double) * nb, hipMemcpyHostToDevice, st));
  HIPCHK(hipEventRecord(bt->dt_copied, st));
  return after_step(bt, *cls, state, advance_impl(bt, *cls, *sch, state, 1.0, bt->d_box_dt));
}

int amr_batch_est_dt(amr_boxtab_t bt, const amr_closures_t* cls, const amr_fab_t* state, double cfl,
                     double* dt_out) {
  if (!bt || !cls || !state || !dt_out) return fail(AMR_ERR_ARG, "null argument");
  int rc = check_closures(*cls);
  if (rc) return rc;
  DescSet* ds = nullptr;
  if ((rc = get_desc(bt, state, 0, &ds))) return rc;
  Closures c = make_closures(*cls);
  hipStream_t st = bt->ctx->stream;
  int nb = bt->bt.nbox;
  HIPCHK(launch_estdt_boxes(bt->bt, ds->d, c, bt->d_partial, bt->d_result, st));
  HIPCHK(hipMemcpyAsync(dt_out, bt->d_result, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int b = 0; b < nb; ++b) {
    if (!(dt_out[b] > 0.0))
      return fail(AMR_ERR_STATE, "non-physical state in problem " + std::to_string(b) +
                                     ": CFL time step is not positive");
    dt_out[b] *= cfl;
  }
  return AMR_OK;
}

namespace {
int finish_boxtab(amr_boxtab_s* bt) {
  int nbox = bt->bt.nbox;
  for (int b = 0; b < nbox; ++b)
    for (int d = 0; d < 3; ++d) {
      bt->bt.h_boxes[b].dx[d] = bt->geoms[b].dx[d];
      bt->bt.h_boxes[b].dxi[d] = 1.0 / bt->geoms[b].dx[d];  // Geometry::InvCellSize
    }
  int rc = AMR_OK;
  hipError_t e = hipMalloc(&bt->bt.d_boxes, sizeof(BoxDev) * nbox);
  if (e == hipSuccess)
    e = hipMemcpy(bt->bt.d_boxes, bt->bt.h_boxes.data(), sizeof(BoxDev) * nbox, hipMemcpyHostToDevice);
  size_t npart = std::max<size_t>(5 * 1024, size_t(estdt_box_blocks(bt->bt)) * nbox);
  if (e == hipSuccess) e = hipMalloc(&bt->d_partial, sizeof(double) * npart);
  if (e == hipSuccess) e = hipMalloc(&bt->d_result, sizeof(double) * std::max(8, nbox));
  if (e == hipSuccess) e = hipHostMalloc(&bt->h_result, sizeof(double) * 8);
  if (e == hipSuccess) e = hipMalloc(&bt->d_box_dt, sizeof(double) * nbox);
  if (e == hipSuccess) e = hipHostMalloc(&bt->h_box_dt, sizeof(double) * nbox);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&bt->dt_copied, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc(&bt->d_clock, sizeof(Clock) * nbox);
  if (e == hipSuccess) e = hipMalloc(&bt->d_flag, sizeof(int) * 4);
  if (e == hipSuccess) e = hipMalloc(&bt->d_dt_levels, sizeof(double) * (MAX_DT_LEVELS + 1));
  if (e == hipSuccess) {
    std::vector<Clock> k(nbox, Clock{0.0, 0.0, 0, 1, 0, 0});
    e = hipMemcpy(bt->d_clock, k.data(), sizeof(Clock) * nbox, hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) e = prepare_block_maps(bt->bt);
  if (e != hipSuccess) return fail(AMR_ERR_HIP, std::string("box table allocation: ") + hipGetErrorString(e));
  rc = build_uncovered(bt);
  if (rc == AMR_OK) rc = build_cover(bt);
  if (rc == AMR_OK) rc = build_nbmap(bt);
  if (rc == AMR_OK) rc = build_xsides(bt);
  if (rc == AMR_OK) rc = build_srcmap(bt);
  if (rc == AMR_OK) rc = build_walls(bt);
  return rc;
}
}  // namespace

int amr_boxtab_destroy(amr_boxtab_t bt) {
  if (!bt) return AMR_OK;
  (void)hipStreamSynchronize(bt->ctx->stream);
  bt->graphs.clear();
  bt->tiles.clear();
  bt->descs.clear();
  if (bt->uw_mem) (void)hipFree(bt->uw_mem);
  for (void* p : bt->uw_tables) (void)hipFree(p);
  for (int w = 0; w < 2; ++w)
    if (bt->ws_mem[w]) (void)hipFree(bt->ws_mem[w]);
  if (bt->d_uncov) (void)hipFree(bt->d_uncov);
  if (bt->d_uncov_blk) (void)hipFree(bt->d_uncov_blk);
  if (bt->d_cover_bits) (void)hipFree(bt->d_cover_bits);
  if (bt->d_owner) (void)hipFree(bt->d_owner);
  if (bt->d_xside) (void)hipFree(bt->d_xside);
  if (bt->d_srcmap_bits) (void)hipFree(bt->d_srcmap_bits);
  if (bt->d_all_covered) (void)hipFree(bt->d_all_covered);
  if (bt->d_wall_tasks) (void)hipFree(bt->d_wall_tasks);
  if (bt->d_wall_blk) (void)hipFree(bt->d_wall_blk);
  bt->run.reset();
  bt->src.reset();
  bt->face_descs.clear();
  free_user_source(bt);
  free_fold(bt);
  bt->adj_g.reset();
  if (bt->adj_e_table) (void)hipFree(bt->adj_e_table);
  if (bt->adj_mem) (void)hipFree(bt->adj_mem);
  if (bt->adj_lq_table) (void)hipFree(bt->adj_lq_table);
  if (bt->adj_le_table) (void)hipFree(bt->adj_le_table);
  if (bt->adj_lmem) (void)hipFree(bt->adj_lmem);
  for (auto& m : bt->bt.block_maps) (void)hipFree(m.second.first);
  if (bt->bt.d_boxes) (void)hipFree(bt->bt.d_boxes);
  if (bt->d_partial) (void)hipFree(bt->d_partial);
  if (bt->d_result) (void)hipFree(bt->d_result);
  if (bt->h_result) (void)hipHostFree(bt->h_result);
  if (bt->d_box_dt) (void)hipFree(bt->d_box_dt);
  if (bt->d_clock) (void)hipFree(bt->d_clock);
  if (bt->d_diag) (void)hipFree(bt->d_diag);
  if (bt->d_flag) (void)hipFree(bt->d_flag);
  if (bt->d_dt_levels) (void)hipFree(bt->d_dt_levels);
  if (bt->h_box_dt) (void)hipHostFree(bt->h_box_dt);
  if (bt->dt_copied) (void)hipEventDestroy(bt->dt_copied);
  if (bt->aux) (void)hipStreamDestroy(bt->aux);
  amr_ctx_t ctx = bt->ctx;
  delete bt;
  release_ctx(ctx);
  return AMR_OK;
}

int amr_fill_boundary(amr_boxtab_t bt, const amr_fab_t* fabs) {
  if (!bt || !fabs) return fail(AMR_ERR_ARG, "null argument");
  DescSet* ds = nullptr;
  int rc = get_desc(bt, fabs, 0, &ds);
  if (rc) return rc;
  if ((rc = build_fill(bt, ds))) return rc;
  return enqueue_fill(bt, ds);
}

namespace {

int floor_div(int a, int r) { return a >= 0 ? a / r : -((-a + r - 1) / r); }

int check_two_levels(amr_boxtab_s* f, amr_boxtab_s* c, int r) {
  if (r != 2 && r != 4) return fail(AMR_ERR_ARG, "refinement ratio must be 2 or 4");
  if (f->batch || c->batch) return fail(AMR_ERR_ARG, "coarse-fine coupling takes level box tables, not batches");
  if (f->ctx != c->ctx) return fail(AMR_ERR_ARG, "fine and coarse box tables belong to different contexts");
  const amr_geom_t& gf = f->geoms[0];
  const amr_geom_t& gc = c->geoms[0];
  for (int d = 0; d < 3; ++d) {
    if (gf.dom_lo[d] != gc.dom_lo[d] * r || gf.dom_hi[d] != (gc.dom_hi[d] + 1) * r - 1)
      return fail(AMR_ERR_ARG, "fine domain is not the coarse domain refined by the ratio");
    if (gf.periodic[d] != gc.periodic[d] || gf.lo_bc[d] != gc.lo_bc[d] || gf.hi_bc[d] != gc.hi_bc[d])
      return fail(AMR_ERR_ARG, "fine and coarse levels differ in periodicity or boundary codes");
  }
  return AMR_OK;
}

CfTable* cf_table(amr_boxtab_s* f, amr_boxtab_s* c, int r, int ngrow) {
  const std::array<uint64_t, 3> key{c->serial, uint64_t(r), uint64_t(ngrow)};
  auto& t = f->cf[key];
  if (!t) {
    if (f->cf.size() > 16) {  // tables of coarse levels long gone: start over
      f->cf.clear();
      return cf_table(f, c, r, ngrow);
    }
    t = std::make_unique<CfTable>();
    t->crse_serial = c->serial;
    t->ratio = r;
    t->ngrow = ngrow;
  }
  return t.get();
}

// Interpolation tasks for fine regions `pieces` (task dst = fine box, region
// lo/hi, possibly outside the domain in periodic directions): each region is
// split by periodic shift and by the coarse box holding its coarse cells.  The
// seven-point coarse stencil of every cell must lie in the coarse level (valid
// cells, periodic images) or outside the domain across a physical boundary
// that is not ext_dir.
int cf_interp_tasks(amr_boxtab_s* F, amr_boxtab_s* C, int r, const std::vector<CopyTask>& pieces,
                    std::vector<InterpTask>& tasks, int64_t* ncell_out) {
  const amr_geom_t& g = F->geoms[0];
  const amr_geom_t& gc = C->geoms[0];
  int len[3], lenc[3];
  for (int d = 0; d < 3; ++d) {
    len[d] = g.dom_hi[d] - g.dom_lo[d] + 1;
    lenc[d] = gc.dom_hi[d] - gc.dom_lo[d] + 1;
  }
  // coverage of the coarse domain by the coarse boxes, one byte per cell
  std::vector<unsigned char> ccov(size_t(lenc[0]) * lenc[1] * lenc[2], 0);
  for (int cb = 0; cb < C->bt.nbox; ++cb) {
    const BoxDev& B = C->bt.h_boxes[cb];
    for (int k = B.lo[2]; k <= B.hi[2]; ++k)
      for (int j = B.lo[1]; j <= B.hi[1]; ++j)
        for (int i = B.lo[0]; i <= B.hi[0]; ++i)
          ccov[(size_t(k - gc.dom_lo[2]) * lenc[1] + (j - gc.dom_lo[1])) * lenc[0] + (i - gc.dom_lo[0])] = 1;
  }
  // every cell of S (clipped to the domain in non-periodic directions,
  // periodic images taken) is a coarse valid cell
  auto covered = [&](const IBox& S) {
    for (int k = S.lo[2]; k <= S.hi[2]; ++k)
      for (int j = S.lo[1]; j <= S.hi[1]; ++j)
        for (int i = S.lo[0]; i <= S.hi[0]; ++i) {
          int c[3] = {i, j, k};
          for (int d = 0; d < 3; ++d) c[d] = ((c[d] - gc.dom_lo[d]) % lenc[d] + lenc[d]) % lenc[d];
          if (!ccov[(size_t(c[2]) * lenc[1] + c[1]) * lenc[0] + c[0]]) return false;
        }
    return true;
  };
  int64_t ncell = 0;
  for (const CopyTask& p : pieces) {
    struct Part {
      IBox x;
      int sh[3];
    };
    std::vector<Part> parts{Part{IBox{{p.lo[0], p.lo[1], p.lo[2]}, {p.hi[0], p.hi[1], p.hi[2]}}, {0, 0, 0}}}, nx;
    for (int d = 0; d < 3; ++d) {
      if (!g.periodic[d]) continue;
      nx.clear();
      for (const Part& q : parts) {
        const int lo3[3] = {q.x.lo[d], std::max(q.x.lo[d], g.dom_lo[d]), std::max(q.x.lo[d], g.dom_hi[d] + 1)};
        const int hi3[3] = {std::min(q.x.hi[d], g.dom_lo[d] - 1), std::min(q.x.hi[d], g.dom_hi[d]), q.x.hi[d]};
        const int sh3[3] = {-len[d], 0, len[d]};
        for (int s = 0; s < 3; ++s) {
          if (lo3[s] > hi3[s]) continue;
          Part w = q;
          w.x.lo[d] = lo3[s];
          w.x.hi[d] = hi3[s];
          w.sh[d] = sh3[s];
          nx.push_back(w);
        }
      }
      parts.swap(nx);
    }
    for (const Part& q : parts) {
      IBox W;
      for (int d = 0; d < 3; ++d) {
        W.lo[d] = q.x.lo[d] - q.sh[d];
        W.hi[d] = q.x.hi[d] - q.sh[d];
      }
      const int64_t need = cells(W.lo, W.hi);
      int64_t got = 0;
      for (int cb = 0; cb < C->bt.nbox; ++cb) {
        const BoxDev& B = C->bt.h_boxes[cb];
        IBox Y;
        for (int d = 0; d < 3; ++d) {
          Y.lo[d] = std::max(W.lo[d], B.lo[d] * r);
          Y.hi[d] = std::min(W.hi[d], (B.hi[d] + 1) * r - 1);
        }
        if (ibox_empty(Y)) continue;
        // the coarse stencil: coarse cells of Y grown by one
        IBox S;
        for (int d = 0; d < 3; ++d) {
          const int clo = floor_div(Y.lo[d], r), chi = floor_div(Y.hi[d], r);
          if (!gc.periodic[d]) {
            const bool elo = gc.lo_bc[d] == 1 || gc.lo_bc[d] == 6, ehi = gc.hi_bc[d] == 1 || gc.hi_bc[d] == 6;
            if ((elo && clo == gc.dom_lo[d]) || (ehi && chi == gc.dom_hi[d]))
              return fail(AMR_ERR_ARG, "coarse-fine interpolation next to ext_dir (bc code 1 or 6) boundaries is not "
                                       "supported");
          }
          S.lo[d] = clo - 1;
          S.hi[d] = chi + 1;
          if (!gc.periodic[d]) {
            S.lo[d] = std::max(S.lo[d], gc.dom_lo[d]);
            S.hi[d] = std::min(S.hi[d], gc.dom_hi[d]);
          }
        }
        if (!covered(S))
          return fail(AMR_ERR_ARG, "coarse-fine interpolation: the coarse stencil leaves the coarse level (fine "
                                   "level not properly nested)");
        InterpTask tk;
        tk.dst = p.dst;
        tk.src = cb;
        for (int d = 0; d < 3; ++d) {
          tk.lo[d] = floor_div(Y.lo[d], r);
          tk.hi[d] = floor_div(Y.hi[d], r);
          tk.flo[d] = Y.lo[d];
          tk.fhi[d] = Y.hi[d];
          tk.sh[d] = q.sh[d];
        }
        tk.start = ncell;
        ncell += cells(tk.lo, tk.hi);
        got += cells(Y.lo, Y.hi);
        tasks.push_back(tk);
      }
      if (got != need)
        return fail(AMR_ERR_ARG, "coarse-fine interpolation: fine level not properly nested (fine cells over no "
                                 "coarse box)");
    }
  }
  *ncell_out = ncell;
  return AMR_OK;
}

// Fine ghost regions no fine box covers (inside the domain, periodic images
// unwrapped), as interpolation tasks.
int build_cf_interp(amr_boxtab_s* F, amr_boxtab_s* C, int r, CfTable* t) {
  if (t->interp_built) return AMR_OK;
  std::vector<InterpTask> tasks;
  int64_t ncell = 0, nun = 0;
  const std::vector<CopyTask> pieces = uncovered_pieces(F, t->ngrow, &nun);
  int rc = cf_interp_tasks(F, C, r, pieces, tasks, &ncell);
  if (rc) return rc;
  if (!tasks.empty()) {
    HIPCHK(hipMalloc(&t->d_interp, sizeof(InterpTask) * tasks.size()));
    HIPCHK(hipMemcpy(t->d_interp, tasks.data(), sizeof(InterpTask) * tasks.size(), hipMemcpyHostToDevice));
  }
  t->ninterp_cells = ncell;
  rc = upload_blocks(tasks, &t->d_interp_blk, &t->ninterp_blk);
  if (rc) return rc;
  t->interp_built = true;
  return AMR_OK;
}

int build_cf_avg(amr_boxtab_s* F, amr_boxtab_s* C, int r, CfTable* t) {
  if (t->avg_built) return AMR_OK;
  std::vector<CopyTask> tasks;
  int64_t ncell = 0;
  for (int fb = 0; fb < F->bt.nbox; ++fb) {
    const BoxDev& B = F->bt.h_boxes[fb];
    IBox X;
    for (int d = 0; d < 3; ++d) {
      if (((B.lo[d] % r) + r) % r != 0 || (((B.hi[d] + 1) % r) + r) % r != 0)
        return fail(AMR_ERR_ARG, "fine box " + std::to_string(fb) + " is not coarsenable by the ratio");
      X.lo[d] = floor_div(B.lo[d], r);
      X.hi[d] = floor_div(B.hi[d] + 1, r) - 1;
    }
    for (int cb = 0; cb < C->bt.nbox; ++cb) {
      const BoxDev& D = C->bt.h_boxes[cb];
      CopyTask tk;
      tk.dst = cb;
      tk.src = fb;
      for (int d = 0; d < 3; ++d) {
        tk.lo[d] = std::max(X.lo[d], D.lo[d]);
        tk.hi[d] = std::min(X.hi[d], D.hi[d]);
        tk.sh[d] = 0;
      }
      if (tk.lo[0] > tk.hi[0] || tk.lo[1] > tk.hi[1] || tk.lo[2] > tk.hi[2]) continue;
      tk.start = ncell;
      ncell += cells(tk.lo, tk.hi);
      tasks.push_back(tk);
    }
  }
  if (!tasks.empty()) {
    HIPCHK(hipMalloc(&t->d_avg, sizeof(CopyTask) * tasks.size()));
    HIPCHK(hipMemcpy(t->d_avg, tasks.data(), sizeof(CopyTask) * tasks.size(), hipMemcpyHostToDevice));
  }
  int rc = upload_blocks(tasks, &t->d_avg_blk, &t->navg_blk);
  if (rc) return rc;
  t->avg_built = true;
  return AMR_OK;
}

}  // namespace

namespace {
int fill_two_levels(amr_boxtab_t fine, amr_boxtab_t crse, int ratio, double a_old, const amr_fab_t* crse_old,
                    double b_new, const amr_fab_t* crse_new, int ngrow, const amr_fab_t* fine_state,
                    bool same_level) {
  if (!fine || !crse || !fine_state) return fail(AMR_ERR_ARG, "null argument");
  if (ngrow < 1 || ngrow > NGHOST) return fail(AMR_ERR_ARG, "ngrow must be 1, 2 or 3");
  if (!std::isfinite(a_old) || !std::isfinite(b_new)) return fail(AMR_ERR_ARG, "time weights must be finite");
  const bool use_old = !(a_old == 0.0 && b_new != 0.0), use_new = b_new != 0.0;
  if (use_old && !crse_old) return fail(AMR_ERR_ARG, "crse_old is required unless a_old == 0 and b_new != 0");
  if (use_new && !crse_new) return fail(AMR_ERR_ARG, "crse_new is required when b_new != 0");
  int rc;
  if ((rc = check_two_levels(fine, crse, ratio))) return rc;
  CfTable* t = cf_table(fine, crse, ratio, ngrow);
  if ((rc = build_cf_interp(fine, crse, ratio, t))) return rc;
  DescSet *dsf = nullptr, *d0 = nullptr, *d1 = nullptr;
  if ((rc = get_desc(fine, fine_state, NGHOST, &dsf))) return rc;
  if ((rc = build_fill(fine, dsf))) return rc;
  const int width = ngrow < NGHOST ? ngrow : 0;
  if (width && (rc = build_fill_narrow(fine, dsf, width))) return rc;
  const DescSet* tk = width ? dsf->narrow[width - 1].get() : dsf;
  if (use_old && (rc = get_desc(crse, crse_old, 1, &d0))) return rc;
  if (use_new && (rc = get_desc(crse, crse_new, 1, &d1))) return rc;
  hipStream_t st = fine->ctx->stream;
  if (fine->walls && !fine->has_problem)
    return fail(AMR_ERR_ARG, "bc code 6 walls need amr_set_problem (wall temperature) first");
  if (same_level) HIPCHK(launch_fill_copy(tk->d_copy, tk->ncopy, tk->d_copy_blk, tk->ncopy_blk, dsf->d, st));
  const bool blend = use_old && use_new;
  const Fab* c0 = use_old ? d0->d : d1->d;
  HIPCHK(launch_interp_cf(t->d_interp, t->d_interp_blk, t->ninterp_blk, c0, blend ? d1->d : c0, blend, a_old,
                          b_new, ratio, dsf->d, st));
  return enqueue_fill(fine, dsf, true, nullptr, width);
}
}  // namespace

int amr_fill_two_levels(amr_boxtab_t fine, amr_boxtab_t crse, int ratio, double a_old, const amr_fab_t* crse_old,
                        double b_new, const amr_fab_t* crse_new, int ngrow, const amr_fab_t* fine_state) {
  return fill_two_levels(fine, crse, ratio, a_old, crse_old, b_new, crse_new, ngrow, fine_state, true);
}

int amr_fill_two_levels_cf(amr_boxtab_t fine, amr_boxtab_t crse, int ratio, double a_old,
                           const amr_fab_t* crse_old, double b_new, const amr_fab_t* crse_new, int ngrow,
                           const amr_fab_t* fine_state) {
  // without an owner map the level may have physical boundaries, whose
  // ghost cells next to other boxes depend on the same-level copies; arrays
  // too far apart for the kernel's 32-bit offsets keep the copies too
  bool same = !(fine && fine->nbmap.on);
  if (!same && fine_state) {
    DescSet* ds = nullptr;
    int rc = get_desc(fine, fine_state, NGHOST, &ds);
    if (rc) return rc;
    desc_layout(ds);
    same = ds->span32 != 1;
  }
  return fill_two_levels(fine, crse, ratio, a_old, crse_old, b_new, crse_new, ngrow, fine_state, same);
}

int amr_average_down(amr_boxtab_t fine, amr_boxtab_t crse, int ratio, const amr_fab_t* fine_state,
                     const amr_fab_t* crse_state) {
  if (!fine || !crse || !fine_state || !crse_state) return fail(AMR_ERR_ARG, "null argument");
  int rc;
  if ((rc = check_two_levels(fine, crse, ratio))) return rc;
  CfTable* t = cf_table(fine, crse, ratio, 0);
  if ((rc = build_cf_avg(fine, crse, ratio, t))) return rc;
  DescSet *df = nullptr, *dc = nullptr;
  if ((rc = get_desc(fine, fine_state, 0, &df))) return rc;
  if ((rc = get_desc(crse, crse_state, 0, &dc))) return rc;
  HIPCHK(launch_avg_down(t->d_avg, t->d_avg_blk, t->navg_blk, df->d, dc->d, ratio, fine->ctx->stream));
  return AMR_OK;
}

int amr_tag_cells(amr_boxtab_t bt, const amr_tagger_t* tagger, const amr_fab_t* state, uint8_t* tags) {
  if (!bt || !tagger || !state || !tags) return fail(AMR_ERR_ARG, "null argument");
  if (bt->batch) return fail(AMR_ERR_ARG, "a batch of problems has no level to tag");
  const int kind = tagger->kind;
  if (kind != AMR_TAG_DENGRAD && kind != AMR_TAG_WALLBAND && kind != AMR_TAG_VORT)
    return fail(AMR_ERR_ARG, "unknown tagger kind " + std::to_string(kind));
  if (!std::isfinite(tagger->threshold) || !std::isfinite(tagger->h) || !std::isfinite(tagger->factor))
    return fail(AMR_ERR_ARG, "tagger parameters must be finite");
  int rc;
  DescSet* ds = nullptr;
  if ((rc = get_desc(bt, state, kind == AMR_TAG_WALLBAND ? 0 : 1, &ds))) return rc;
  TagArgs ta = tag_args(bt);
  ta.kind = kind;
  ta.thr = tagger->threshold;
  ta.h = tagger->h;
  ta.factor = tagger->factor;
  HIPCHK(launch_tag_cells(bt->bt, ds->d, ta, tags, bt->ctx->stream));
  return AMR_OK;
}

int amr_tag_blocks(amr_boxtab_t bt, const uint8_t* tags, int cb, uint8_t* out) {
  if (!bt || !tags || !out) return fail(AMR_ERR_ARG, "null argument");
  if (bt->batch) return fail(AMR_ERR_ARG, "a batch of problems has no level to tag");
  if (cb < 1) return fail(AMR_ERR_ARG, "block size must be positive");
  const TagArgs ta = tag_args(bt);
  for (int d = 0; d < 3; ++d)
    if (ta.n[d] % cb != 0) return fail(AMR_ERR_ARG, "domain extents must be multiples of the block size");
  HIPCHK(launch_tag_blocks(tags, ta, cb, out, bt->ctx->stream));
  return AMR_OK;
}

int amr_regrid_fill(amr_boxtab_t fine_new, amr_boxtab_t fine_old, const amr_fab_t* old_state, amr_boxtab_t crse,
                    int ratio, double a_old, const amr_fab_t* crse_old, double b_new, const amr_fab_t* crse_new,
                    const amr_fab_t* new_state) {
  if (!fine_new || !crse || !new_state) return fail(AMR_ERR_ARG, "null argument");
  if (fine_old && !old_state) return fail(AMR_ERR_ARG, "old_state is required with fine_old");
  if (!std::isfinite(a_old) || !std::isfinite(b_new)) return fail(AMR_ERR_ARG, "time weights must be finite");
  const bool use_old = !(a_old == 0.0 && b_new != 0.0), use_new = b_new != 0.0;
  if (use_old && !crse_old) return fail(AMR_ERR_ARG, "crse_old is required unless a_old == 0 and b_new != 0");
  if (use_new && !crse_new) return fail(AMR_ERR_ARG, "crse_new is required when b_new != 0");
  int rc;
  if ((rc = check_two_levels(fine_new, crse, ratio))) return rc;
  if (fine_old) {
    if (fine_old->batch || fine_old->ctx->device != fine_new->ctx->device)
      return fail(AMR_ERR_ARG, "the old level must be a level on the same device");
    const amr_geom_t &gn = fine_new->geoms[0], &go = fine_old->geoms[0];
    for (int d = 0; d < 3; ++d)
      if (gn.dom_lo[d] != go.dom_lo[d] || gn.dom_hi[d] != go.dom_hi[d])
        return fail(AMR_ERR_ARG, "the old and new levels have different domains");
  }
  DescSet *dn = nullptr, *dold = nullptr, *d0 = nullptr, *d1 = nullptr;
  if ((rc = get_desc(fine_new, new_state, 0, &dn))) return rc;
  if (fine_old && (rc = get_desc(fine_old, old_state, 0, &dold))) return rc;
  if (use_old && (rc = get_desc(crse, crse_old, 1, &d0))) return rc;
  if (use_new && (rc = get_desc(crse, crse_new, 1, &d1))) return rc;
  // copies from the old level's boxes, the rest from the coarse level
  std::vector<CopyTask> copies, pieces;
  int64_t ncopy = 0;
  for (int b = 0; b < fine_new->bt.nbox; ++b) {
    const BoxDev& N = fine_new->bt.h_boxes[b];
    if (fine_old)
      for (int o = 0; o < fine_old->bt.nbox; ++o) {
        const BoxDev& O = fine_old->bt.h_boxes[o];
        CopyTask t;
        bool empty = false;
        for (int d = 0; d < 3; ++d) {
          t.lo[d] = std::max(N.lo[d], O.lo[d]);
          t.hi[d] = std::min(N.hi[d], O.hi[d]);
          t.sh[d] = 0;
          empty = empty || t.lo[d] > t.hi[d];
        }
        if (empty) continue;
        t.dst = b;
        t.src = o;
        t.start = ncopy;
        ncopy += cells(t.lo, t.hi);
        copies.push_back(t);
      }
    uncovered_by(N, b, fine_old ? &fine_old->bt.h_boxes : nullptr, pieces);
  }
  std::vector<InterpTask> itasks;
  int64_t ninterp = 0;
  if ((rc = cf_interp_tasks(fine_new, crse, ratio, pieces, itasks, &ninterp))) return rc;
  DevBuf<CopyTask> dcopy;
  DevBuf<TaskBlock> dcopy_blk, dint_blk;
  DevBuf<InterpTask> dint;
  int ncopy_blk = 0, nint_blk = 0;
  if (!copies.empty()) {
    HIPCHK(hipMalloc(&dcopy.p, sizeof(CopyTask) * copies.size()));
    HIPCHK(hipMemcpy(dcopy.p, copies.data(), sizeof(CopyTask) * copies.size(), hipMemcpyHostToDevice));
  }
  if ((rc = upload_blocks(copies, &dcopy_blk.p, &ncopy_blk))) return rc;
  if (!itasks.empty()) {
    HIPCHK(hipMalloc(&dint.p, sizeof(InterpTask) * itasks.size()));
    HIPCHK(hipMemcpy(dint.p, itasks.data(), sizeof(InterpTask) * itasks.size(), hipMemcpyHostToDevice));
  }
  if ((rc = upload_blocks(itasks, &dint_blk.p, &nint_blk))) return rc;
  hipStream_t st = fine_new->ctx->stream;
  if (ncopy_blk) HIPCHK(launch_copy_regions(dcopy.p, dcopy_blk.p, ncopy_blk, dold->d, dn->d, nullptr, st));
  if (nint_blk) {
    const bool blend = use_old && use_new;
    const Fab* c0 = use_old ? d0->d : d1->d;
    HIPCHK(launch_interp_cf(dint.p, dint_blk.p, nint_blk, c0, blend ? d1->d : c0, blend, a_old, b_new, ratio, dn->d,
                            st));
  }
  // the task lists are freed on return
  HIPCHK(hipStreamSynchronize(st));
  return AMR_OK;
}

int amr_stage_coeffs(int order_rk, int stages_rk, double dt, int stage, int* nstages, int* mode, double* a,
                     double* b, double* c, double* d) {
  StageCoef sc;
  int n = 0;
  int rc = stage_coeffs(order_rk, stages_rk, dt, stage, &n, &sc);
  if (rc) return rc;
  if (nstages) *nstages = n;
  if (mode) *mode = sc.mode;
  if (a) *a = sc.a;
  if (b) *b = sc.b;
  if (c) *c = sc.c;
  if (d) *d = sc.d;
  return AMR_OK;
}

int amr_rk_stage(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch, int mode, double a, double b,
                 double c, double d, const amr_fab_t* s1, const amr_fab_t* sin, const amr_fab_t* out) {
  if (!bt || !cls || !sch || !sin || !out) return fail(AMR_ERR_ARG, "null argument");
  int rc;
  if ((rc = check_scheme(*sch)) || (rc = check_closures(*cls))) return rc;
  if (mode < 0 || mode > 2) return fail(AMR_ERR_ARG, "mode must be 0, 1 or 2");
  StageCoef sc{mode, a, b, c, d};
  bool need_s1 = mode != 2 && a != 0.0;
  if (need_s1 && !s1) return fail(AMR_ERR_ARG, "s1 is required when a != 0");
  DescSet *ds1 = nullptr, *dsin = nullptr, *dout = nullptr;
  if (need_s1 && (rc = get_desc(bt, s1, 0, &ds1))) return rc;
  if ((rc = get_desc(bt, sin, NGHOST, &dsin))) return rc;
  if ((rc = get_desc(bt, out, 0, &dout))) return rc;
  // out must not share memory with sin
  for (int i = 0; i < bt->bt.nbox; ++i)
    if (mapping_of(out[i]) == mapping_of(sin[i]))
      return fail(AMR_ERR_ARG, "out must not alias sin");
  StageArgs sa{sc, nullptr, 0, 0, 0};
  return do_stage(bt, *cls, *sch, sa, ds1, dsin, dout);
}

int amr_face_fluxes(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch, const amr_fab_t* state,
                    const amr_fab_t* flux_x, const amr_fab_t* flux_y, const amr_fab_t* flux_z) {
  if (!bt || !cls || !sch || !state || !flux_x || !flux_y || !flux_z) return fail(AMR_ERR_ARG, "null argument");
  int rc;
  if ((rc = check_scheme(*sch)) || (rc = check_closures(*cls))) return rc;
  DescSet* dsin = nullptr;
  if ((rc = get_desc(bt, state, NGHOST, &dsin))) return rc;
  const amr_fab_t* fl[3] = {flux_x, flux_y, flux_z};
  const Fab* faces[3] = {nullptr, nullptr, nullptr};
  const int nb = bt->bt.nbox;
  for (int d = 0; d < 3; ++d) {
    // face arrays: the box's faces normal to d, lo .. hi + e_d
    for (int b = 0; b < nb; ++b) {
      const amr_fab_t& f = fl[d][b];
      const BoxDev& B = bt->bt.h_boxes[b];
      for (int e = 0; e < 3; ++e)
        if (f.lo[e] != B.lo[e] || f.hi[e] != B.hi[e] + (e == d ? 1 : 0))
          return fail(AMR_ERR_ARG, "face fab " + std::to_string(b) + " of direction " + std::to_string(d) +
                                       " is not the box's face box (hi + 1 in that direction)");
      if (f.ng < 0 || !f.data) return fail(AMR_ERR_ARG, "face fab " + std::to_string(b) + " has no data");
      const int64_t nx = f.hi[0] - f.lo[0] + 1 + 2 * f.ng, ny = f.hi[1] - f.lo[1] + 1 + 2 * f.ng;
      const int64_t nz = f.hi[2] - f.lo[2] + 1 + 2 * f.ng;
      if (f.jstride < nx || f.kstride < f.jstride * ny || f.nstride < f.kstride * nz)
        return fail(AMR_ERR_ARG, "face fab " + std::to_string(b) + " strides do not cover its face box");
    }
    DescSet* ds = nullptr;
    for (auto& x : bt->face_descs)
      if (std::memcmp(x->host.data(), fl[d], sizeof(amr_fab_t) * nb) == 0) ds = x.get();
    if (!ds) {
      if (bt->face_descs.size() >= 12) {
        HIPCHK(hipStreamSynchronize(bt->ctx->stream));
        bt->face_descs.erase(bt->face_descs.begin());
      }
      std::unique_ptr<DescSet> n;
      if ((rc = new_desc(bt, fl[d], &n))) return rc;
      ds = n.get();
      bt->face_descs.push_back(std::move(n));
    }
    faces[d] = ds->d;
  }
  if ((rc = ensure_unfused(bt, *sch))) return rc;
  StageArgs sa{StageCoef{2, 0.0, 0.0, 0.0, 0.0}, nullptr, 0, 0, 0};
  HIPCHK(launch_unfused_stage(bt->bt, bt->uw, make_closures(*cls), *sch, sa, bt->wall_args, nullptr, dsin->d, nullptr,
                              bt->ctx->stream, faces));
  return AMR_OK;
}

int amr_rhs(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch, const amr_fab_t* state,
            const amr_fab_t* dsdt) {
  return amr_rk_stage(bt, cls, sch, 2, 0.0, 0.0, 0.0, 0.0, nullptr, state, dsdt);
}

int amr_advance(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch, const amr_fab_t* state,
                double time, double dt) {
  (void)time;
  if (!bt || !cls || !sch || !state) return fail(AMR_ERR_ARG, "null argument");
  int rc;
  if ((rc = check_scheme(*sch)) || (rc = check_closures(*cls))) return rc;
  return after_step(bt, *cls, state, advance_impl(bt, *cls, *sch, state, dt));
}

int amr_advance_graph(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch, const amr_fab_t* state,
                      double dt, int nsteps) {
  if (!bt || !cls || !sch || !state || nsteps <= 0) return fail(AMR_ERR_ARG, "invalid argument");
  int rc;
  if ((rc = check_scheme(*sch)) || (rc = check_closures(*cls))) return rc;
  int nb = bt->bt.nbox;
  // the tiling in force is part of the key: a graph records its launches
  const FusedConfig kcfg = choose_config(bt, *sch);
  std::vector<char> key(sizeof(amr_fab_t) * nb + sizeof(amr_closures_t) + sizeof(amr_scheme_t) + sizeof(double) +
                        sizeof(int) + sizeof(FusedConfig));
  char* p = key.data();
  std::memcpy(p + key.size() - sizeof(FusedConfig), &kcfg, sizeof(FusedConfig));
  std::memcpy(p, state, sizeof(amr_fab_t) * nb);
  p += sizeof(amr_fab_t) * nb;
  std::memcpy(p, cls, sizeof(amr_closures_t));
  p += sizeof(amr_closures_t);
  std::memcpy(p, sch, sizeof(amr_scheme_t));
  p += sizeof(amr_scheme_t);
  std::memcpy(p, &dt, sizeof(double));
  p += sizeof(double);
  std::memcpy(p, &nsteps, sizeof(int));
  GraphEntry* ge = nullptr;
  for (auto& g : bt->graphs)
    if (g->key == key) ge = g.get();
  hipStream_t st = bt->ctx->stream;
  if (!ge) {
    if ((rc = prepare_step_tables(bt, *sch, state))) return rc;
    HIPCHK(hipStreamSynchronize(st));
    auto g = std::make_unique<GraphEntry>();
    g->key = key;
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < nsteps; ++s) {
      rc = advance_impl(bt, *cls, *sch, state, dt);
      if (rc) {
        hipGraph_t junk = nullptr;
        (void)hipStreamEndCapture(st, &junk);
        if (junk) (void)hipGraphDestroy(junk);
        return rc;
      }
    }
    HIPCHK(hipStreamEndCapture(st, &g->graph));
    HIPCHK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    ge = g.get();
    if (bt->graphs.size() >= 8) bt->graphs.erase(bt->graphs.begin());
    bt->graphs.push_back(std::move(g));
  }
  HIPCHK(hipGraphLaunch(ge->exec, st));
  return AMR_OK;
}

int amr_clock_set(amr_boxtab_t bt, double time) {
  if (!bt) return fail(AMR_ERR_ARG, "null box table");
  if (std::isnan(time)) return fail(AMR_ERR_ARG, "time is NaN");
  const int n = bt->bt.nbox;
  std::vector<Clock> k(n, Clock{time, 0.0, 0, 1, 0, 0});
  hipStream_t st = bt->ctx->stream;
  HIPCHK(hipMemcpyAsync(bt->d_clock, k.data(), sizeof(Clock) * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  return AMR_OK;
}

int amr_clock_get(amr_boxtab_t bt, double* time, double* dt, int32_t* steps) {
  if (!bt) return fail(AMR_ERR_ARG, "null box table");
  const int np = nprob_of(bt);
  std::vector<Clock> k(np);
  hipStream_t st = bt->ctx->stream;
  HIPCHK(hipMemcpyAsync(k.data(), bt->d_clock, sizeof(Clock) * np, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  int bad = -1;
  for (int p = 0; p < np; ++p) {
    if (time) time[p] = k[p].t;
    if (dt) dt[p] = k[p].dt;
    if (steps) steps[p] = k[p].steps;
    if (k[p].bad && bad < 0) bad = p;
  }
  if (bad >= 0)
    return fail(AMR_ERR_STATE, "non-physical state in problem " + std::to_string(bad) +
                                   ": CFL time step not positive; its steps stopped there");
  return AMR_OK;
}

int amr_advance_cfl(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch, const amr_fab_t* state,
                    double cfl, double stop_time, int nsteps) {
  if (!bt || !cls || !sch || !state || nsteps < 0 || !(cfl > 0.0)) return fail(AMR_ERR_ARG, "invalid argument");
  int rc;
  if ((rc = check_scheme(*sch)) || (rc = check_closures(*cls))) return rc;
  for (int s = 0; s < nsteps; ++s)
    if ((rc = cfl_step(bt, *cls, *sch, state, cfl, stop_time))) return rc;
  return after_step(bt, *cls, state, AMR_OK);
}

int amr_advance_cfl_graph(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch,
                          const amr_fab_t* state, double cfl, double stop_time, int nsteps) {
  if (!bt || !cls || !sch || !state || nsteps <= 0 || !(cfl > 0.0)) return fail(AMR_ERR_ARG, "invalid argument");
  int rc;
  if ((rc = check_scheme(*sch)) || (rc = check_closures(*cls))) return rc;
  int nb = bt->bt.nbox;
  const FusedConfig kcfg = choose_config(bt, *sch);
  // key: a tag (CFL graph), the state, closures, scheme, cfl, stop time,
  // step count and the tiling in force
  const char tag = 'C';
  std::vector<char> key(1 + sizeof(amr_fab_t) * nb + sizeof(amr_closures_t) + sizeof(amr_scheme_t) +
                        2 * sizeof(double) + sizeof(int) + sizeof(FusedConfig));
  char* p = key.data();
  *p++ = tag;
  std::memcpy(p, state, sizeof(amr_fab_t) * nb);
  p += sizeof(amr_fab_t) * nb;
  std::memcpy(p, cls, sizeof(amr_closures_t));
  p += sizeof(amr_closures_t);
  std::memcpy(p, sch, sizeof(amr_scheme_t));
  p += sizeof(amr_scheme_t);
  std::memcpy(p, &cfl, sizeof(double));
  p += sizeof(double);
  std::memcpy(p, &stop_time, sizeof(double));
  p += sizeof(double);
  std::memcpy(p, &nsteps, sizeof(int));
  p += sizeof(int);
  std::memcpy(p, &kcfg, sizeof(FusedConfig));
  GraphEntry* ge = nullptr;
  for (auto& g : bt->graphs)
    if (g->key == key) ge = g.get();
  hipStream_t st = bt->ctx->stream;
  if (!ge) {
    if ((rc = prepare_step_tables(bt, *sch, state))) return rc;
    HIPCHK(hipStreamSynchronize(st));
    auto g = std::make_unique<GraphEntry>();
    g->key = key;
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < nsteps; ++s) {
      rc = cfl_step(bt, *cls, *sch, state, cfl, stop_time);
      if (rc) {
        hipGraph_t junk = nullptr;
        (void)hipStreamEndCapture(st, &junk);
        if (junk) (void)hipGraphDestroy(junk);
        return rc;
      }
    }
    HIPCHK(hipStreamEndCapture(st, &g->graph));
    HIPCHK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    ge = g.get();
    if (bt->graphs.size() >= 8) bt->graphs.erase(bt->graphs.begin());
    bt->graphs.push_back(std::move(g));
  }
  HIPCHK(hipGraphLaunch(ge->exec, st));
  return AMR_OK;
}

int amr_new_dt_levels(int nlev, const amr_boxtab_t* levels, const amr_closures_t* cls,
                      const amr_fab_t* const* states, const int32_t* n_cycle, double cfl, double cur_time,
                      double stop_time, int first, double* dt_level) {
  if (nlev <= 0 || nlev > MAX_DT_LEVELS || !levels || !cls || !states || !n_cycle || !dt_level || !(cfl > 0.0))
    return fail(AMR_ERR_ARG, "invalid argument (1 <= nlev <= 8, cfl > 0)");
  int rc = check_closures(*cls);
  if (rc) return rc;
  DtLevels a{};
  a.nlev = nlev;
  a.first = first ? 1 : 0;
  a.cfl = cfl;
  a.cur_time = cur_time;
  a.stop_time = stop_time;
  amr_boxtab_s* b0 = levels[0];
  if (!b0) return fail(AMR_ERR_ARG, "null box table");
  hipStream_t st0 = b0->ctx->stream;
  const Closures c = make_closures(*cls);
  for (int l = 0; l < nlev; ++l) {
    amr_boxtab_s* bt = levels[l];
    if (!bt || !states[l]) return fail(AMR_ERR_ARG, "null box table or state");
    if (n_cycle[l] <= 0) return fail(AMR_ERR_ARG, "n_cycle entries must be positive");
    if (bt->ctx->device != b0->ctx->device) return fail(AMR_ERR_ARG, "levels on different devices");
    DescSet* ds = nullptr;
    if ((rc = get_desc(bt, states[l], 0, &ds))) return rc;
    hipStream_t st = bt->ctx->stream;
    HIPCHK(launch_estdt(bt->bt, ds->d, c, bt->d_partial, bt->d_result, st));
    if (st != st0) HIPCHK(hipStreamSynchronize(st));
    a.est[l] = bt->d_result;
    a.ncycle[l] = n_cycle[l];
  }
  double h[MAX_DT_LEVELS + 1];
  for (int l = 0; l < nlev; ++l) h[l] = dt_level[l];
  HIPCHK(hipMemcpyAsync(b0->d_dt_levels, h, sizeof(double) * nlev, hipMemcpyHostToDevice, st0));
  HIPCHK(launch_dt_levels(a, b0->d_dt_levels, st0));
  HIPCHK(hipMemcpyAsync(h, b0->d_dt_levels, sizeof(double) * (nlev + 1), hipMemcpyDeviceToHost, st0));
  HIPCHK(hipStreamSynchronize(st0));
  if (!(h[nlev] > 0.0))
    return fail(AMR_ERR_STATE, "non-physical state: CFL time step is not positive (negative pressure or NaN)");
  for (int l = 0; l < nlev; ++l) dt_level[l] = h[l];
  return AMR_OK;
}

int amr_set_tiling(amr_ctx_t ctx, const amr_scheme_t* sch, int tx, int ty, int tz) {
  if (!ctx || !sch) return fail(AMR_ERR_ARG, "null argument");
  std::string key = scheme_key(*sch);
  if (tx == 0 && ty == 0 && tz == 0) {
    ctx->tuned.erase(key);
    return AMR_OK;
  }
  if (tx == -1) {
    ctx->tuned[key] = FusedConfig{-1, 0, 0};
    return AMR_OK;
  }
  if (tz <= 0) return fail(AMR_ERR_ARG, "tz must be positive");
  if (!fused_supported(*sch, tx, ty))
    return fail(AMR_ERR_ARG, "no fused kernel compiled for this scheme and tile shape");
  ctx->tuned[key] = FusedConfig{tx, ty, tz};
  return AMR_OK;
}

int amr_set_tiling_form(amr_ctx_t ctx, const amr_scheme_t* sch, int tx, int ty, int tz, int form) {
  if (!ctx || !sch) return fail(AMR_ERR_ARG, "null argument");
  if (form == 0) return amr_set_tiling(ctx, sch, tx, ty, tz);
  if (form != 1) return fail(AMR_ERR_ARG, "form must be 0 or 1");
  if (tz <= 0) return fail(AMR_ERR_ARG, "tz must be positive");
  if (!fused_self_supported(*sch, tx, ty))
    return fail(AMR_ERR_ARG, "the self-loading form takes 16x16 tiles of a viscous KEEP 2/4 scheme");
  FusedConfig f{tx, ty, tz};
  f.form = form;
  ctx->tuned[scheme_key(*sch)] = f;
  return AMR_OK;
}

int amr_set_box_tiling(amr_ctx_t ctx, const amr_scheme_t* sch, int tx, int ty, int tz) {
  if (!ctx || !sch) return fail(AMR_ERR_ARG, "null argument");
  if (tz <= 0) return fail(AMR_ERR_ARG, "tz must be positive");
  if (!fused_supported(*sch, tx, ty))
    return fail(AMR_ERR_ARG, "no fused kernel compiled for this scheme and tile shape");
  ctx->tuned[scheme_key(*sch)] = FusedConfig{tx, ty, tz, 1};
  return AMR_OK;
}

int amr_step_plan(amr_boxtab_t bt, const amr_scheme_t* sch, const amr_fab_t* state, int32_t info[6]) {
  if (!bt || !sch || !state || !info) return fail(AMR_ERR_ARG, "null argument");
  int rc;
  if ((rc = check_scheme(*sch))) return rc;
  if ((rc = ensure_ws(bt, state))) return rc;
  DescSet *dst = nullptr, *dw[2] = {nullptr, nullptr};
  if ((rc = get_desc(bt, state, NGHOST, &dst))) return rc;
  for (int w = 0; w < 2; ++w)
    if ((rc = get_desc(bt, bt->ws_fabs[w].data(), NGHOST, &dw[w]))) return rc;
  const bool level = use_level_tiles(bt, *sch, dst, dst, dw[0]);
  FusedConfig cfg;
  const bool fused = fused_path(bt, *sch, &cfg, level);
  info[0] = fused ? cfg.tx : -1;
  info[1] = fused ? cfg.ty : 0;
  info[2] = fused ? cfg.tz : 0;
  info[3] = fused ? (level ? 2 : (nb_direct(bt, *sch, level, dst) ? 3 : 1)) : 0;
  info[4] = 0;
  info[5] = fused ? cfg.form : 0;
  int nst = 0;
  StageCoef sc;
  if ((rc = stage_coeffs(sch->order_rk, sch->stages_rk, 1.0, 0, &nst, &sc))) return rc;
  if (fused && nst > 1 && bt->nuncov_cells > 0) {
    bool ok = false;
    if ((rc = uncovered_from_un(bt, *sch, dst, dw, &ok))) return rc;
    info[4] = ok ? 1 : 0;
  }
  return AMR_OK;
}

int amr_autotune(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch, const amr_fab_t* state,
                 int32_t tiling[3]) {
  if (!bt || !cls || !sch || !state || !tiling) return fail(AMR_ERR_ARG, "null argument");
  int rc;
  if ((rc = check_scheme(*sch)) || (rc = check_closures(*cls))) return rc;
  std::string key = scheme_key(*sch);
  bt->ctx->tuned.erase(key);  // tuning replaces an explicit setting
  if ((rc = ensure_ws(bt, state))) return rc;
  DescSet *din = nullptr, *dout = nullptr;
  if ((rc = get_desc(bt, state, NGHOST, &din))) return rc;
  if ((rc = get_desc(bt, bt->ws_fabs[0].data(), NGHOST, &dout))) return rc;
  // level tiles and box tiles both where level tiles apply
  const bool level = level_possible(bt, *sch, din, din, dout);
  int maxz = 0;
  for (auto& B : bt->bt.h_boxes) maxz = std::max(maxz, B.hi[2] - B.lo[2] + 1);
  const int maxz_level = level ? bt->cover.hi[2] - bt->cover.lo[2] + 1 : 0;  // level tiles run across boxes
  std::vector<FusedConfig> cand;
  const int shapes[3][2] = {{16, 16}, {32, 8}, {8, 8}};
  const int tzs[6] = {16, 24, 32, 48, 64, 128};
  for (int box_tiles = level ? 0 : 1; box_tiles <= 1; ++box_tiles)
    for (auto& sh : shapes)
      for (int tz : tzs) {
        FusedConfig c{sh[0], sh[1], tz, box_tiles};
        if (tz > 2 * (box_tiles ? maxz : maxz_level) && tz != 16) continue;
        if (!(fused_supported(*sch, c.tx, c.ty) && fused_fits(bt, c, !box_tiles))) continue;
        cand.push_back(c);
        // the self-loading form of the same tiling
        if (fused_self_supported(*sch, c.tx, c.ty)) {
          c.form = 1;
          cand.push_back(c);
        }
      }
  if (cand.empty()) {
    bt->tuned[key] = FusedConfig{-1, 0, 0};
    tiling[0] = -1;
    tiling[1] = tiling[2] = 0;
    return AMR_OK;
  }
  hipStream_t st = bt->ctx->stream;
  // the events are released and the previous choice restored on every exit
  // (a failed autotune leaves the tiling in force as it was)
  struct Guard {
    amr_boxtab_s* bt;
    std::string key;
    bool had;
    FusedConfig prev;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool done = false;
    ~Guard() {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      if (done) return;
      if (had) bt->tuned[key] = prev;
      else bt->tuned.erase(key);
    }
  };
  auto pit = bt->tuned.find(key);
  Guard guard{bt, key, pit != bt->tuned.end(), pit != bt->tuned.end() ? pit->second : FusedConfig{}};
  HIPCHK(hipEventCreate(&guard.e0));
  HIPCHK(hipEventCreate(&guard.e1));
  hipEvent_t e0 = guard.e0, e1 = guard.e1;
  float best = 1e30f;
  FusedConfig bestc = cand[0];
  // whole steps (fills and the uncovered-cell copy included) that write only
  // the workspaces; untimed steps first, so that the first candidates are
  // not timed while the clocks ramp up
  const double dt_probe = 1e-12;
  bt->tuned[key] = cand[0];
  for (int rep = 0; rep < 2; ++rep)
    if ((rc = advance_impl(bt, *cls, *sch, state, dt_probe, nullptr, true))) return rc;
  // best of three timed steps per candidate, then the three fastest again,
  // interleaved, so that a clock or thermal drift does not pick the winner
  auto time_cfg = [&](const FusedConfig& cfg, int reps, float* out) -> int {
    bt->tuned[key] = cfg;
    float tbest = 1e30f;
    for (int rep = 0; rep < reps + 1; ++rep) {
      HIPCHK(hipEventRecord(e0, st));
      int r = advance_impl(bt, *cls, *sch, state, dt_probe, nullptr, true);
      if (r) return r;
      HIPCHK(hipEventRecord(e1, st));
      HIPCHK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      if (rep > 0) tbest = std::min(tbest, ms);
    }
    *out = std::min(*out, tbest);
    return AMR_OK;
  };
  std::vector<float> t(cand.size(), 1e30f);
  for (size_t i = 0; i < cand.size(); ++i)
    if ((rc = time_cfg(cand[i], 3, &t[i]))) return rc;
  std::vector<size_t> order(cand.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return t[a] < t[b]; });
  order.resize(std::min<size_t>(3, order.size()));
  for (int round = 0; round < 2; ++round)
    for (size_t i : order)
      if ((rc = time_cfg(cand[i], 3, &t[i]))) return rc;
  for (size_t i : order)
    if (t[i] < best) {
      best = t[i];
      bestc = cand[i];
    }
  bt->tuned[key] = bestc;
  guard.done = true;
  tiling[0] = bestc.tx;
  tiling[1] = bestc.ty;
  tiling[2] = bestc.tz;
  return AMR_OK;
}

int amr_est_dt(amr_boxtab_t bt, const amr_closures_t* cls, const amr_fab_t* state, double cfl, double* dt_out) {
  if (!bt || !cls || !state || !dt_out) return fail(AMR_ERR_ARG, "null argument");
  int rc = check_closures(*cls);
  if (rc) return rc;
  DescSet* ds = nullptr;
  if ((rc = get_desc(bt, state, 0, &ds))) return rc;
  Closures c = make_closures(*cls);
  hipStream_t st = bt->ctx->stream;
  HIPCHK(launch_estdt(bt->bt, ds->d, c, bt->d_partial, bt->d_result, st));
  HIPCHK(hipMemcpyAsync(bt->h_result, bt->d_result, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  double dt = bt->h_result[0];
  if (!(dt > 0.0) || std::isnan(dt))
    return fail(AMR_ERR_STATE, "non-physical state: CFL time step is not positive (negative pressure or NaN)");
  *dt_out = dt * cfl;  // src/CNS.cpp:391
  return AMR_OK;
}

int amr_est_dt_noalloc(amr_boxtab_t bt, const amr_closures_t* cls, const amr_fab_t* state, double cfl,
                       double* dt_out) {
  if (!bt || !cls || !state || !dt_out) return fail(AMR_ERR_ARG, "null argument");
  int rc = check_closures(*cls);
  if (rc) return rc;
  const int nb = bt->bt.nbox;
  if (nb > NOALLOC_MAX_BOXES)
    return fail(AMR_ERR_ARG, "amr_est_dt_noalloc: " + std::to_string(nb) + " boxes, at most " +
                                 std::to_string(NOALLOC_MAX_BOXES));
  if ((rc = check_fabs(bt, state, 0))) return rc;
  void* scr = bt->ctx->scratch;
  if (!scr) return fail(AMR_ERR_ARG, "amr_est_dt_noalloc: no scratch array on this device");
  hipStream_t st = bt->ctx->stream;
  // descriptors into the scratch array through a fixed host buffer
  Fab chunk[128];
  for (int b0 = 0; b0 < nb; b0 += 128) {
    const int m = std::min(128, nb - b0);
    for (int i = 0; i < m; ++i) chunk[i] = make_fab(state[b0 + i]);
    HIPCHK(hipMemcpyAsync(static_cast<char*>(scr) + size_t(b0) * sizeof(Fab), chunk, sizeof(Fab) * m,
                          hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));  // before the buffer is refilled
  }
  double* res = nullptr;
  HIPCHK(launch_estdt_scratch(bt->bt, make_closures(*cls), scr, &res, st));
  double dt = 0.0;
  HIPCHK(hipMemcpyAsync(&dt, res, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (!(dt > 0.0) || std::isnan(dt))
    return fail(AMR_ERR_STATE, "non-physical state: CFL time step is not positive (negative pressure or NaN)");
  *dt_out = dt * cfl;  // src/CNS.cpp:391
  return AMR_OK;
}

int amr_check_state(amr_boxtab_t bt, const amr_closures_t* cls, const amr_fab_t* state, int32_t* bad_box) {
  if (!bt || !cls || !state) return fail(AMR_ERR_ARG, "null argument");
  int rc = check_closures(*cls);
  if (rc) return rc;
  return check_state_impl(bt, *cls, state, bad_box);
}

int amr_set_state_check(amr_boxtab_t bt, int on) {
  if (!bt) return fail(AMR_ERR_ARG, "null box table");
  bt->state_check = on != 0;
  return AMR_OK;
}

int amr_diagnostics(amr_boxtab_t bt, const amr_closures_t* cls, const amr_fab_t* state, double dt, double totals[5],
                    double prim_min[6], double prim_max[6], double wave_max[6], double cfl[3]) {
  if (!bt || !cls || !state) return fail(AMR_ERR_ARG, "null argument");
  int rc = check_closures(*cls);
  if (rc) return rc;
  DescSet* ds = nullptr;
  if ((rc = get_desc(bt, state, 0, &ds))) return rc;
  hipStream_t st = bt->ctx->stream;
  if (!bt->d_diag) HIPCHK(hipMalloc(&bt->d_diag, sizeof(double) * (21 * 1024 + 32)));
  double* res = bt->d_diag + 21 * 1024;
  HIPCHK(launch_diag(bt->bt, ds->d, make_closures(*cls), dt, bt->d_diag, res, st));
  double h[21];
  HIPCHK(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int m = 0; m < 6; ++m) {
    if (prim_min) prim_min[m] = h[m];
    if (prim_max) prim_max[m] = h[6 + m];
  }
  for (int d = 0; d < 3; ++d) {
    if (wave_max) {
      wave_max[2 * d] = h[12 + 3 * d];
      wave_max[2 * d + 1] = h[13 + 3 * d];
    }
    if (cfl) cfl[d] = h[14 + 3 * d];
  }
  if (totals) return amr_totals(bt, state, totals);
  return AMR_OK;
}

int amr_totals(amr_boxtab_t bt, const amr_fab_t* state, double out[5]) {
  if (!bt || !state || !out) return fail(AMR_ERR_ARG, "null argument");
  DescSet* ds = nullptr;
  int rc = get_desc(bt, state, 0, &ds);
  if (rc) return rc;
  hipStream_t st = bt->ctx->stream;
  HIPCHK(launch_totals(bt->bt, ds->d, bt->d_partial, bt->d_result, st));
  HIPCHK(hipMemcpyAsync(bt->h_result, bt->d_result, sizeof(double) * 5, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  double probsize = 1.0;
  for (int d = 0; d < 3; ++d) probsize *= (bt->geom.dom_hi[d] - bt->geom.dom_lo[d] + 1) * bt->geom.dx[d];
  for (int n = 0; n < 5; ++n) out[n] = bt->h_result[n] * probsize;  // src/CNS.cpp:541
  return AMR_OK;
}

int amr_set_source(amr_boxtab_t bt, const amr_fab_t* source) {
  if (!bt) return fail(AMR_ERR_ARG, "null box table");
  HIPCHK(hipStreamSynchronize(bt->ctx->stream));
  if (!source) {
    drop_graphs(bt);  // captured steps read the table being freed
    bt->src.reset();
    return AMR_OK;
  }
  int rc = check_fabs(bt, source, 0);
  if (rc) return rc;
  std::unique_ptr<DescSet> ds;
  if ((rc = new_desc(bt, source, &ds))) return rc;
  bt->src = std::move(ds);
  bt->graphs.clear();  // captured steps refer to the previous table
  return AMR_OK;
}

int amr_set_user_source(amr_boxtab_t bt, const char* code, const double* params, int nparams) {
  if (!bt) return fail(AMR_ERR_ARG, "null box table");
  if (nparams < 0 || (nparams > 0 && !params)) return fail(AMR_ERR_ARG, "params: nparams values required");
  drop_graphs(bt);  // captured stages launch the kernel being replaced
  HIPCHK(hipStreamSynchronize(bt->ctx->stream));
  free_user_source(bt);
  if (!code) return AMR_OK;
  HIPCHK(hipSetDevice(bt->ctx->device));
  std::string log;
  if (compile_user_source(code, bt->ctx->device, &bt->usrc_mod, &bt->usrc_fn, &log)) {
    free_user_source(bt);
    return fail(AMR_ERR_ARG, "user_source does not compile: " + log);
  }
  HIPCHK(hipMalloc(&bt->usrc_params, sizeof(double) * std::max(nparams, 1)));
  if (nparams) HIPCHK(hipMemcpy(bt->usrc_params, params, sizeof(double) * nparams, hipMemcpyHostToDevice));
  return AMR_OK;
}

int amr_set_problem(amr_boxtab_t bt, const amr_closures_t* cls, const amr_problem_t* problem) {
  if (!bt) return fail(AMR_ERR_ARG, "null box table");
  HIPCHK(hipStreamSynchronize(bt->ctx->stream));
  bt->graphs.clear();  // captured steps hold the previous parameters
  if (!problem) {
    bt->has_problem = false;
    bt->problem = amr_problem_t{};
    return AMR_OK;
  }
  if (!cls) return fail(AMR_ERR_ARG, "amr_set_problem needs the closures of the wall fill");
  int rc = check_closures(*cls);
  if (rc) return rc;
  if (bt->walls && !(problem->wall_T > 0.0)) return fail(AMR_ERR_ARG, "bc code 6 walls need a wall temperature > 0");
  bt->problem = *problem;
  bt->has_problem = true;
  bt->wall_fill = WallFill{problem->wall_T, cls->gamma, cls->Rspec};
  return AMR_OK;
}

int amr_vjp_stage(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch, int mode, double a, double b,
                  double c, double d, const amr_fab_t* sin, const amr_fab_t* lam_out, const amr_fab_t* lam_s1,
                  const amr_fab_t* lam_sin, int accumulate) {
  if (!bt || !cls || !sch || !sin || !lam_out || !lam_sin) return fail(AMR_ERR_ARG, "null argument");
  int rc;
  if ((rc = check_scheme(*sch)) || (rc = check_closures(*cls)) || (rc = check_adjoint(*sch)) ||
      (rc = check_adjoint_table(bt, *sch)))
    return rc;
  if (mode < 0 || mode > 2) return fail(AMR_ERR_ARG, "mode must be 0, 1 or 2");
  DescSet *ds = nullptr, *dlo = nullptr, *dls1 = nullptr, *dls = nullptr;
  if ((rc = get_desc(bt, sin, NGHOST, &ds))) return rc;
  if ((rc = get_desc(bt, lam_out, 0, &dlo))) return rc;
  if (lam_s1 && (rc = get_desc(bt, lam_s1, 0, &dls1))) return rc;
  if ((rc = get_desc(bt, lam_sin, 0, &dls))) return rc;
  const int nb = bt->bt.nbox;
  if (fabs_overlap(lam_sin, lam_out, nb) || (lam_s1 && fabs_overlap(lam_s1, lam_out, nb)))
    return fail(AMR_ERR_ARG, "lam_sin and lam_s1 must not share storage with lam_out");
  StageArgs sa{StageCoef{mode, a, b, c, d}, nullptr, 0, 0, 0};
  return adj_stage(bt, *cls, *sch, sa, ds, dlo, dls1, dls, accumulate != 0, 1);
}

int amr_vjp_rhs(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch, const amr_fab_t* state,
                const amr_fab_t* lam, const amr_fab_t* grad) {
  return amr_vjp_stage(bt, cls, sch, 2, 0.0, 0.0, 0.0, 0.0, state, lam, nullptr, grad, 0);
}

int amr_adjoint_run(amr_boxtab_t bt, const amr_closures_t* cls, const amr_scheme_t* sch, const amr_fab_t* state0,
                    double dt, int nsteps, int kcheck, int objective, const amr_fab_t* weights,
                    const amr_fab_t* grad, double* J_out) {
  if (!bt || !cls || !sch || !state0 || !grad) return fail(AMR_ERR_ARG, "null argument");
  int rc;
  if ((rc = check_scheme(*sch)) || (rc = check_closures(*cls)) || (rc = check_adjoint(*sch)) ||
      (rc = check_adjoint_table(bt, *sch)))
    return rc;
  if (nsteps < 1 || kcheck < 1) return fail(AMR_ERR_ARG, "nsteps and kcheck must be >= 1");
  if (objective != 0 && objective != 1) return fail(AMR_ERR_ARG, "objective must be 0 or 1");
  if (objective == 1 && !weights) return fail(AMR_ERR_ARG, "objective 1 needs weights");
  if (bt->batch) return fail(AMR_ERR_ARG, "adjoint runs take a level box table, not a batch");
  int nst = 0;
  StageCoef sc;
  if ((rc = stage_coeffs(sch->order_rk, sch->stages_rk, dt, 0, &nst, &sc))) return rc;
  DescSet *d0 = nullptr, *dg = nullptr, *dw = nullptr;
  if ((rc = get_desc(bt, state0, NGHOST, &d0))) return rc;
  if ((rc = get_desc(bt, grad, 0, &dg))) return rc;
  if (objective == 1 && (rc = get_desc(bt, weights, 0, &dw))) return rc;
  if ((rc = ensure_run(bt, state0, nsteps, kcheck, nst))) return rc;
  AdjRun* R = bt->run.get();
  hipStream_t st = bt->ctx->stream;
  const int k = kcheck;
  auto stage_args = [&](int s, StageArgs* sa) {
    StageCoef c2;
    int n2 = 0;
    int r = stage_coeffs(sch->order_rk, sch->stages_rk, dt, s, &n2, &c2);
    *sa = StageArgs{c2, nullptr, sch->order_rk, sch->stages_rk, s};
    return r;
  };
  // forward step n (global index) on stream sx; the last stage is skipped on
  // request
  auto run_step = [&](int n, bool skip_last, hipStream_t sx) -> int {
    const int seg = n / k, nl = n % k;
    for (int s = 0; s < nst; ++s) {
      DescSet* sin = R->sb(seg, nl, s);
      // stage inputs need their ghost cells (for the forward and the reverse
      // sweep) unless the level is one periodic array read through images
      const bool need_fill = !wrap_ok(bt, sin);
      if (skip_last && s == nst - 1) {
        // the stage itself is not needed, but the reverse sweep reads its
        // input with ghost cells
        if (need_fill) {
          int r = enqueue_fill(bt, sin, false, sx);
          if (r) return r;
        }
        break;
      }
      StageArgs sa;
      int r = stage_args(s, &sa);
      if (r) return r;
      DescSet* s1 = R->sb(seg, nl, 0);
      DescSet* out;
      if (s < nst - 1) out = R->sb(seg, nl, s + 1);
      else if (n + 1 == nsteps) out = R->uf();
      else if ((n + 1) % k == 0) out = R->ck(seg + 1);
      else out = R->sb(seg, nl + 1, 0);
      if (need_fill && (r = enqueue_fill(bt, sin, false, sx))) return r;
      if ((r = do_stage(bt, *cls, *sch, sa, s1, sin, out, nullptr, sx))) return r;
    }
    return AMR_OK;
  };
  HIPCHK(launch_copy_valid(bt->bt, d0->d, R->ck(0)->d, nullptr, st));
  for (auto& ds : R->states)
    if ((rc = copy_uncovered(bt, d0, ds.get()))) return rc;
  for (int n = 0; n < nsteps; ++n)
    if ((rc = run_step(n, false, st))) return rc;
  DescSet *lc = R->lam[0].get(), *lt = R->lam[1].get(), *la = R->lam[2].get();
  HIPCHK(launch_objective(bt->bt, objective, R->uf()->d, dw ? dw->d : nullptr, lc->d, bt->d_partial,
                          bt->d_result, st));
  HIPCHK(hipMemcpyAsync(bt->h_result, bt->d_result, sizeof(double), hipMemcpyDeviceToHost, st));
  // Reverse sweep, last segment first.  The forward sweep leaves the stage
  // inputs of the last two segments in place; an earlier segment is
  // recomputed from its checkpoint into its slot set, which the segment two
  // later has finished with.  Where the forward stages run the fused kernel
  // (no box-table workspace) the recompute of segment seg-1 runs on a second
  // stream while segment seg is reversed: the forward kernel of a small level
  // leaves most CUs to the reverse kernels.
  const bool overlap = R->nseg > 2 && fused_path(bt, *sch, nullptr);
  hipStream_t sr = st;
  if (overlap) {
    if (!bt->aux) HIPCHK(hipStreamCreateWithFlags(&bt->aux, hipStreamNonBlocking));
    sr = bt->aux;
    while (int(R->ev.size()) < 2 * R->nseg) {
      hipEvent_t e;
      HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      R->ev.push_back(e);
    }
  }
  auto ev_rev = [&](int seg) { return R->ev[seg]; };
  auto ev_rec = [&](int seg) { return R->ev[R->nseg + seg]; };
  for (int seg = R->nseg - 1; seg >= 0; --seg) {
    const int r = seg - 1;  // recomputed next, beside this segment's reverse sweep
    if (r >= 0 && r <= R->nseg - 3) {
      if (overlap) HIPCHK(hipStreamWaitEvent(sr, ev_rev(seg + 1), 0));
      const int n0 = r * k, len = std::min(k, nsteps - n0);
      for (int n = n0; n < n0 + len; ++n)
        if ((rc = run_step(n, n == n0 + len - 1, sr))) return rc;
      if (overlap) HIPCHK(hipEventRecord(ev_rec(r), sr));
    }
    if (overlap && seg <= R->nseg - 3) HIPCHK(hipStreamWaitEvent(st, ev_rec(seg), 0));
    const int len = std::min(k, nsteps - seg * k);
    for (int nl = len - 1; nl >= 0; --nl) {
      // la accumulates the adjoint of U^n: assigned by the last stage,
      // added to by the others (stage 0's input is U^n itself)
      for (int s = nst - 1; s >= 0; --s) {
        StageArgs sa;
        if ((rc = stage_args(s, &sa))) return rc;
        DescSet* sin = R->sb(seg, nl, s);
        if (s > 0) {
          if ((rc = adj_stage(bt, *cls, *sch, sa, sin, lc, la, lt, false, s == nst - 1 ? 2 : 1))) return rc;
          std::swap(lc, lt);
        } else {
          if ((rc = adj_stage(bt, *cls, *sch, sa, sin, lc, la, la, nst > 1, 1))) return rc;
        }
      }
      std::swap(lc, la);
    }
    if (overlap) HIPCHK(hipEventRecord(ev_rev(seg), st));
  }
  HIPCHK(launch_copy_valid(bt->bt, lc->d, dg->d, nullptr, st));
  HIPCHK(hipStreamSynchronize(st));
  if (J_out) *J_out = bt->h_result[0];
  return AMR_OK;
}


int amr_plot_vars(amr_boxtab_t bt, const amr_closures_t* cls, const amr_fab_t* state, int state_vars,
                  int derive_mask, const amr_fab_t* out) {
  if (!bt || !cls || !state || !out) return fail(AMR_ERR_ARG, "null argument");
  if (derive_mask < 0 || derive_mask > 31) return fail(AMR_ERR_ARG, "derive_mask uses bits 0..4 only");
  if (!state_vars && !derive_mask) return fail(AMR_ERR_ARG, "no plot variable selected");
  int rc;
  if ((rc = check_closures(*cls))) return rc;
  DescSet *ds = nullptr, *dout = nullptr;
  if ((rc = get_desc(bt, state, 0, &ds))) return rc;
  if ((rc = get_desc(bt, out, 0, &dout))) return rc;
  for (int b = 0; b < bt->bt.nbox; ++b)
    if (mapping_of(out[b]) == mapping_of(state[b])) return fail(AMR_ERR_ARG, "out must not alias state");
  HIPCHK(launch_plot_vars(bt->bt, ds->d, dout->d, state_vars ? 1 : 0, derive_mask, make_closures(*cls),
                          bt->ctx->stream));
  return AMR_OK;
}

int amr_write_plotfile(const char* dir, int nlev, const amr_boxtab_t* levels, const amr_fab_t* const* plot,
                       int ncomp, const char* const* names, double time, const int32_t* level_steps,
                       const int32_t* ref_ratio) {
  if (!dir || !levels || !plot || !names || !level_steps || nlev < 1 || ncomp < 1)
    return fail(AMR_ERR_ARG, "invalid argument");
  if (nlev > 1 && !ref_ratio) return fail(AMR_ERR_ARG, "ref_ratio is required with more than one level");
  for (int n = 0; n < ncomp; ++n)
    if (!names[n]) return fail(AMR_ERR_ARG, "null variable name");
  for (int l = 0; l < nlev; ++l) {
    if (!levels[l] || !plot[l]) return fail(AMR_ERR_ARG, "null level");
    if (levels[l]->batch) return fail(AMR_ERR_ARG, "plotfiles are written for level box tables, not batches");
    if (l > 0) {
      const amr_geom_t& gf = levels[l]->geoms[0];
      const amr_geom_t& gc = levels[l - 1]->geoms[0];
      const int r = ref_ratio[l - 1];
      for (int d = 0; d < 3; ++d)
        if (r < 1 || gf.dom_lo[d] != gc.dom_lo[d] * r || gf.dom_hi[d] != (gc.dom_hi[d] + 1) * r - 1)
          return fail(AMR_ERR_ARG, "level " + std::to_string(l) + " is not level " + std::to_string(l - 1) +
                                       " refined by ref_ratio");
    }
  }
  auto box_text = [](const int32_t* lo, const int32_t* hi) {
    std::ostringstream o;
    o << "((" << lo[0] << ',' << lo[1] << ',' << lo[2] << ") (" << hi[0] << ',' << hi[1] << ',' << hi[2]
      << ") (0,0,0))";
    return o.str();
  };
  auto make_dir = [](const std::string& d) {
    if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) return fail(AMR_ERR_ARG, "cannot create directory " + d);
    return int(AMR_OK);
  };
  const std::string root(dir);
  int rc;
  if ((rc = make_dir(root))) return rc;
  // per level: pack the valid cells on the device, copy them to the host,
  // write Level_l/Cell_D_00000 and Level_l/Cell_H (one data file per level)
  std::vector<std::vector<std::string>> level_boxes(nlev);
  for (int l = 0; l < nlev; ++l) {
    amr_boxtab_s* bt = levels[l];
    const int nb = bt->bt.nbox;
    DescSet* ds = nullptr;
    if ((rc = get_desc(bt, plot[l], 0, &ds))) return rc;
    std::vector<int64_t> off(nb + 1, 0);
    for (int b = 0; b < nb; ++b) {
      const BoxDev& B = bt->bt.h_boxes[b];
      off[b + 1] = off[b] + int64_t(B.hi[0] - B.lo[0] + 1) * (B.hi[1] - B.lo[1] + 1) * (B.hi[2] - B.lo[2] + 1) * ncomp;
    }
    int64_t* d_off = nullptr;
    double* d_buf = nullptr;
    std::vector<double> h(static_cast<size_t>(off[nb]));
    HIPCHK(hipMalloc(&d_off, sizeof(int64_t) * (nb + 1)));
    HIPCHK(hipMemcpy(d_off, off.data(), sizeof(int64_t) * (nb + 1), hipMemcpyHostToDevice));
    hipError_t e = hipMalloc(&d_buf, sizeof(double) * std::max<int64_t>(off[nb], 1));
    if (e == hipSuccess) e = launch_pack_valid(bt->bt, ds->d, ncomp, d_off, d_buf, bt->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(bt->ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(h.data(), d_buf, sizeof(double) * off[nb], hipMemcpyDeviceToHost);
    if (d_buf) (void)hipFree(d_buf);
    (void)hipFree(d_off);
    if (e != hipSuccess) return fail(AMR_ERR_HIP, std::string("plotfile copy: ") + hipGetErrorString(e));
    const std::string ldir = root + "/Level_" + std::to_string(l);
    if ((rc = make_dir(ldir))) return rc;
    std::ofstream fd(ldir + "/Cell_D_00000", std::ios::out | std::ios::trunc | std::ios::binary);
    if (!fd) return fail(AMR_ERR_ARG, "cannot write " + ldir + "/Cell_D_00000");
    std::vector<int64_t> fpos(nb);
    std::vector<double> mn(size_t(nb) * ncomp), mx(size_t(nb) * ncomp);
    for (int b = 0; b < nb; ++b) {
      const BoxDev& B = bt->bt.h_boxes[b];
      const std::string bs = box_text(B.lo, B.hi);
      level_boxes[l].push_back(bs);
      fpos[b] = int64_t(fd.tellp());
      fd << "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))" << bs << ' ' << ncomp << '\n';
      const int64_t nc = (off[b + 1] - off[b]) / ncomp;
      const double* p = h.data() + off[b];
      fd.write(reinterpret_cast<const char*>(p), std::streamsize(sizeof(double) * (off[b + 1] - off[b])));
      for (int n = 0; n < ncomp; ++n) {
        double lo = p[n * nc], hi = p[n * nc];
        for (int64_t t = 1; t < nc; ++t) {
          lo = std::min(lo, p[n * nc + t]);
          hi = std::max(hi, p[n * nc + t]);
        }
        mn[size_t(b) * ncomp + n] = lo;
        mx[size_t(b) * ncomp + n] = hi;
      }
    }
    fd.close();
    if (!fd) return fail(AMR_ERR_ARG, "write failed: " + ldir + "/Cell_D_00000");
    std::ofstream fh(ldir + "/Cell_H", std::ios::out | std::ios::trunc);
    fh.precision(17);
    fh << 1 << '\n' << 1 << '\n' << ncomp << '\n' << 0 << '\n';
    fh << '(' << nb << ' ' << 0 << '\n';
    for (const std::string& bs : level_boxes[l]) fh << bs << '\n';
    fh << ")\n";
    fh << nb << '\n';
    for (int b = 0; b < nb; ++b) fh << "FabOnDisk: Cell_D_00000 " << fpos[b] << '\n';
    fh << '\n' << nb << ',' << ncomp << '\n';
    for (int b = 0; b < nb; ++b) {
      for (int n = 0; n < ncomp; ++n) fh << mn[size_t(b) * ncomp + n] << ',';
      fh << '\n';
    }
    fh << '\n' << nb << ',' << ncomp << '\n';
    for (int b = 0; b < nb; ++b) {
      for (int n = 0; n < ncomp; ++n) fh << mx[size_t(b) * ncomp + n] << ',';
      fh << '\n';
    }
    fh.close();
    if (!fh) return fail(AMR_ERR_ARG, "write failed: " + ldir + "/Cell_H");
  }
  // the plotfile Header (CNS::writePlotFile, src/CNS.cpp:688-807)
  std::ofstream os(root + "/Header", std::ios::out | std::ios::trunc);
  os.precision(15);
  const amr_geom_t& g0 = levels[0]->geoms[0];
  os << "HyperCLaw-V1.1\n" << ncomp << '\n';
  for (int n = 0; n < ncomp; ++n) os << names[n] << '\n';
  os << 3 << '\n' << time << '\n' << nlev - 1 << '\n';
  for (int d = 0; d < 3; ++d) os << g0.prob_lo[d] << ' ';
  os << '\n';
  for (int d = 0; d < 3; ++d) os << g0.prob_hi[d] << ' ';
  os << '\n';
  for (int l = 0; l + 1 < nlev; ++l) os << ref_ratio[l] << ' ';
  os << '\n';
  for (int l = 0; l < nlev; ++l) os << box_text(levels[l]->geoms[0].dom_lo, levels[l]->geoms[0].dom_hi) << ' ';
  os << '\n';
  for (int l = 0; l < nlev; ++l) os << level_steps[l] << ' ';
  os << '\n';
  for (int l = 0; l < nlev; ++l) {
    for (int d = 0; d < 3; ++d) os << levels[l]->geoms[0].dx[d] << ' ';
    os << '\n';
  }
  os << 0 << '\n' << 0 << '\n';
  for (int l = 0; l < nlev; ++l) {
    amr_boxtab_s* bt = levels[l];
    const amr_geom_t& g = bt->geoms[0];
    os << l << ' ' << bt->bt.nbox << ' ' << time << '\n' << level_steps[l] << '\n';
    for (int b = 0; b < bt->bt.nbox; ++b) {
      const BoxDev& B = bt->bt.h_boxes[b];
      for (int d = 0; d < 3; ++d)
        os << g.prob_lo[d] + B.lo[d] * g.dx[d] << ' ' << g.prob_lo[d] + (B.hi[d] + 1) * g.dx[d] << '\n';
    }
    os << "Level_" << l << "/Cell\n";
  }
  os.close();
  if (!os) return fail(AMR_ERR_ARG, "write failed: " + root + "/Header");
  return AMR_OK;
}

}  // extern "C"
