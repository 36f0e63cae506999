// Ghost fill (single-level FillPatch) and valid-cell reductions (CFL dt,
// conserved totals).
#include <algorithm>

#include "kernels.h"

namespace amr {

namespace {

constexpr int FBLK = TASK_BLK;

// FillBoundary + periodic images: copy valid cells of `src` into ghost
// cells of `dst` (AMReX FillBoundary semantics restated).
// One lane per ghost cell of the concatenated task list (packed block map:
// a workgroup's 256 cells may span several small tasks; each lane steps
// forward from the task of the workgroup's first cell).
__global__ void __launch_bounds__(FBLK) k_fill_copy(const CopyTask* tasks, int ntask, const TaskBlock* blocks,
                                                    const Fab* fabs) {
  const TaskBlock tb = blocks[blockIdx.x];
  const int64_t g = int64_t(tb.first) + threadIdx.x;
  int t = tb.task;
  while (t + 1 < ntask && tasks[t + 1].start <= g) ++t;
  const CopyTask& T = tasks[t];
  const int nx = T.hi[0] - T.lo[0] + 1, ny = T.hi[1] - T.lo[1] + 1, nz = T.hi[2] - T.lo[2] + 1;
  const int64_t r = g - T.start;
  if (r >= int64_t(nx) * ny * nz) return;  // past the last task
  const int r2 = int(r / nx);
  int c[3] = {T.lo[0] + int(r - int64_t(r2) * nx), T.lo[1] + r2 % ny, T.lo[2] + r2 / ny};
  const Fab d = fabs[T.dst], s = fabs[T.src];
#pragma unroll
  for (int n = 0; n < NCONS; ++n) d(c[0], c[1], c[2], n) = s(c[0] - T.sh[0], c[1] - T.sh[1], c[2] - T.sh[2], n);
}

// bc code 6 ghost cell: CnsFillExtDir (src/set/bcs.cpp:17-87) calling the
// channel's bcnormal (exm/Channel/prob.h:157-187), an isothermal no-slip
// wall with dp/dn = 0.  s_int is the first interior cell of the column, and
// every ghost layer uses it; dratio = ghost distance / first-cell distance
// from the wall, formed as the reference forms it.
__device__ __forceinline__ void wall_ghost(const BcTask& T, const Fab& f, const int c[3], const BoxDev& bx,
                                           const WallFill& wf) {
  int s[3] = {c[0], c[1], c[2]};
  s[T.dir] = T.bnd;
  double si[5];
#pragma unroll
  for (int n = 0; n < NCONS; ++n) si[n] = f(s[0], s[1], s[2], n);
  const double dxd = bx.dx[T.dir];
  const double x = T.plo + static_cast<double>(c[T.dir] + 0.5) * dxd;
  double di, de;
  if (T.side == 0) {
    di = static_cast<double>(T.bnd + 0.5) * dxd - T.plo;
    de = T.plo - x;
  } else {
    di = T.phi - static_cast<double>(T.bnd + 0.5) * dxd;
    de = x - T.phi;
  }
  const double dratio = de / di;
  const double qu = -si[UMX] / si[URHO];
  const double qv = -si[UMY] / si[URHO];
  const double qw = -si[UMZ] / si[URHO];
  const double eint_int = (si[UET] - 0.5 * (si[UMX] * si[UMX] + si[UMY] * si[UMY] + si[UMZ] * si[UMZ]) / si[URHO]) / si[URHO];
  const double p_int = (wf.gamma - 1.0) * si[URHO] * eint_int;
  const double T_int = p_int / (wf.Rspec * si[URHO]);
  const double Tx = wf.Tw + dratio * (wf.Tw - T_int);
  const double qT = Tx < 50.0 ? 50.0 : Tx;
  const double qr = p_int / (wf.Rspec * qT);
  const double ekin = 0.5 * (qu * qu + qv * qv + qw * qw);
  const double eint = p_int / (qr * (wf.gamma - 1.0));
  f(c[0], c[1], c[2], URHO) = qr;
  f(c[0], c[1], c[2], UMX) = qr * qu;
  f(c[0], c[1], c[2], UMY) = qr * qv;
  f(c[0], c[1], c[2], UMZ) = qr * qw;
  f(c[0], c[1], c[2], UET) = qr * (eint + ekin);
}

// Physical boundaries (src/set/CNS_setup.cpp:18-31 BC map; AMReX foextrap
// and reflect_even/odd): ghost value from the boundary cell or its mirror.
__global__ void __launch_bounds__(FBLK) k_fill_bc(const BcTask* tasks, const TaskBlock* blocks, const Fab* fabs,
                                                  const BoxDev* boxes, WallFill wf) {
  const TaskBlock tb = blocks[blockIdx.x];
  const BcTask& T = tasks[tb.task];
  int c[3];
  if (!task_cell(T, tb, threadIdx.x, c)) return;
  const Fab f = fabs[T.box];
  if (T.type[0] == BC_WALL) {
    wall_ghost(T, f, c, boxes[T.box], wf);
    return;
  }
  const int cd = c[T.dir];
  for (int n = 0; n < NCONS; ++n) {
    int ty = T.type[n];
    if (ty == 0) continue;
    int s[3] = {c[0], c[1], c[2]};
    if (ty == 2) s[T.dir] = T.bnd;
    else s[T.dir] = T.side == 0 ? 2 * T.bnd - 1 - cd : 2 * T.bnd + 1 - cd;
    double v = f(s[0], s[1], s[2], n);
    f(c[0], c[1], c[2], n) = (ty == 4) ? -v : v;
  }
}

// Region copies between FAB sets: dst0[T.dst](c) = dst1[T.dst](c) = src[T.src](c - sh).
__global__ void __launch_bounds__(FBLK) k_copy_regions(const CopyTask* tasks, const TaskBlock* blocks,
                                                       const Fab* src, const Fab* dst0, const Fab* dst1) {
  const TaskBlock tb = blocks[blockIdx.x];
  const CopyTask& T = tasks[tb.task];
  int c[3];
  if (!task_cell(T, tb, threadIdx.x, c)) return;
  const Fab s = src[T.src], d0 = dst0[T.dst];
  double v[NCONS];
#pragma unroll
  for (int n = 0; n < NCONS; ++n) v[n] = s(c[0] - T.sh[0], c[1] - T.sh[1], c[2] - T.sh[2], n);
#pragma unroll
  for (int n = 0; n < NCONS; ++n) d0(c[0], c[1], c[2], n) = v[n];
  if (dst1 != nullptr) {
    const Fab d1 = dst1[T.dst];
#pragma unroll
    for (int n = 0; n < NCONS; ++n) d1(c[0], c[1], c[2], n) = v[n];
  }
}

__global__ void __launch_bounds__(FBLK) k_copy_valid(const BoxDev* boxes, const Fab* src, const Fab* dst,
                                                     const double* box_dt) {
  const int b = blockIdx.y;
  if (box_dt != nullptr && !(box_dt[b] > 0.0)) return;
  const BoxDev bx = boxes[b];
  int nx = bx.hi[0] - bx.lo[0] + 1, ny = bx.hi[1] - bx.lo[1] + 1, nz = bx.hi[2] - bx.lo[2] + 1;
  int64_t t = int64_t(blockIdx.x) * FBLK + threadIdx.x;
  if (t >= int64_t(nx) * ny * nz) return;
  int i = bx.lo[0] + int(t % nx);
  int64_t r = t / nx;
  int j = bx.lo[1] + int(r % ny);
  int k = bx.lo[2] + int(r / ny);
  const Fab s = src[b], d = dst[b];
#pragma unroll
  for (int n = 0; n < NCONS; ++n) d(i, j, k, n) = s(i, j, k, n);
}

// Plot variables of CNS::writePlotFile's plotMF: the five state components
// when `state_vars`, then the derived variables selected by `mask` in the
// order src/set/CNS_setup.cpp:156-175 registers them (bit 0 pressure, 1
// temperature, 2..4 x/y/z velocity), each as derpres, dertemp and dervel
// compute it (src/cls/CNS_K.h:76-129).
__global__ void __launch_bounds__(FBLK) k_plot_vars(const BoxDev* boxes, const Fab* src, const Fab* dst,
                                                    int state_vars, int mask, double gm1, double cv) {
  const int b = blockIdx.y;
  const BoxDev bx = boxes[b];
  int nx = bx.hi[0] - bx.lo[0] + 1, ny = bx.hi[1] - bx.lo[1] + 1, nz = bx.hi[2] - bx.lo[2] + 1;
  int64_t t = int64_t(blockIdx.x) * FBLK + threadIdx.x;
  if (t >= int64_t(nx) * ny * nz) return;
  int i = bx.lo[0] + int(t % nx);
  int64_t r = t / nx;
  int j = bx.lo[1] + int(r % ny);
  int k = bx.lo[2] + int(r / ny);
  const Fab s = src[b], d = dst[b];
  double u[NCONS];
#pragma unroll
  for (int n = 0; n < NCONS; ++n) u[n] = s(i, j, k, n);
  int c = 0;
  if (state_vars) {
#pragma unroll
    for (int n = 0; n < NCONS; ++n) d(i, j, k, c++) = u[n];
  }
  const double rhoinv = 1.0 / u[URHO];
  const double rhoeint = u[UET] - 0.5 * rhoinv * (u[UMX] * u[UMX] + u[UMY] * u[UMY] + u[UMZ] * u[UMZ]);
  if (mask & 1) d(i, j, k, c++) = gm1 * rhoeint;
  if (mask & 2) d(i, j, k, c++) = (rhoeint * rhoinv) / cv;
  if (mask & 4) d(i, j, k, c++) = u[UMX] / u[URHO];
  if (mask & 8) d(i, j, k, c++) = u[UMY] / u[URHO];
  if (mask & 16) d(i, j, k, c++) = u[UMZ] / u[URHO];
}

// Valid cells of every box packed into one buffer: box b's cells start at
// off[b], component-major, x fastest within a component (the order of a FAB
// on disk).
__global__ void __launch_bounds__(FBLK) k_pack_valid(const BoxDev* boxes, const Fab* src, int ncomp,
                                                     const int64_t* off, double* out) {
  const int b = blockIdx.y;
  const BoxDev bx = boxes[b];
  int nx = bx.hi[0] - bx.lo[0] + 1, ny = bx.hi[1] - bx.lo[1] + 1, nz = bx.hi[2] - bx.lo[2] + 1;
  const int64_t ncell = int64_t(nx) * ny * nz;
  int64_t t = int64_t(blockIdx.x) * FBLK + threadIdx.x;
  if (t >= ncell) return;
  int i = bx.lo[0] + int(t % nx);
  int64_t r = t / nx;
  int j = bx.lo[1] + int(r % ny);
  int k = bx.lo[2] + int(r / ny);
  const Fab s = src[b];
  for (int n = 0; n < ncomp; ++n) out[off[b] + n * ncell + t] = s(i, j, k, n);
}

constexpr int RBLK = 256;

// cns_estdt (src/cls/CNS_K.h:12-47; z momentum read from UMY, CNS_K.h:29):
// per-block minimum of the cell time steps.
__device__ __forceinline__ double cell_estdt(const Fab& u, int i, int j, int k, const Closures& c, double dx0,
                                             double dx1, double dx2) {
  double rho = u(i, j, k, URHO);
  double mx = u(i, j, k, UMX);
  double my = u(i, j, k, UMY);
  double mz = u(i, j, k, UMY);
  double rhoinv = 1.0 / dmax(rho, 1.0e-40);
  double vx = mx * rhoinv, vy = my * rhoinv, vz = mz * rhoinv;
  double ke = 0.5 * rho * (vx * vx + vy * vy + vz * vz);
  double rhoei = (u(i, j, k, UET) - ke);
  double p = c.gm1 * rhoei;
  double cs = sqrt(c.gamma * p * rhoinv);
  double dtx = dx0 / (fabs(vx) + cs);
  double dty = dx1 / (fabs(vy) + cs);
  double dtz = dx2 / (fabs(vz) + cs);
  return dmin(dtx, dmin(dty, dtz));
}

__device__ __forceinline__ double block_min(double v, double* sm) {
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int s = RBLK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = dmin(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  return sm[0];
}

// per-box minimum: blockIdx.y = box, partial[box * gridDim.x + blockIdx.x]
__global__ void __launch_bounds__(RBLK) k_estdt_boxes(const BoxDev* boxes, const Fab* st, Closures c,
                                                      double* partial, const Clock* clk, int batch, double stop) {
  __shared__ double sm[RBLK];
  const int b = blockIdx.y;
  // a problem whose clock has stopped takes step 0 whatever its estimate:
  // its cells are not read (uniform over the block)
  if (clk != nullptr && clock_done(clk[batch ? b : 0], stop)) {
    if (threadIdx.x == 0) partial[size_t(b) * gridDim.x + blockIdx.x] = 1.79769313486231570815e+308;
    return;
  }
  const BoxDev bx = boxes[b];
  const Fab u = st[b];
  double dt = 1.79769313486231570815e+308;
  int nx = bx.hi[0] - bx.lo[0] + 1, ny = bx.hi[1] - bx.lo[1] + 1, nz = bx.hi[2] - bx.lo[2] + 1;
  int64_t n = int64_t(nx) * ny * nz;
  for (int64_t t = int64_t(blockIdx.x) * RBLK + threadIdx.x; t < n; t += int64_t(gridDim.x) * RBLK) {
    int i = bx.lo[0] + int(t % nx);
    int64_t r = t / nx;
    int j = bx.lo[1] + int(r % ny);
    int k = bx.lo[2] + int(r / ny);
    dt = dmin(dt, cell_estdt(u, i, j, k, c, bx.dx[0], bx.dx[1], bx.dx[2]));
  }
  double m = block_min(dt, sm);
  if (threadIdx.x == 0) partial[size_t(b) * gridDim.x + blockIdx.x] = m;
}

__global__ void __launch_bounds__(RBLK) k_min_boxes(const double* partial, int nper, double* result) {
  __shared__ double sm[RBLK];
  double v = 1.79769313486231570815e+308;
  for (int t = threadIdx.x; t < nper; t += RBLK) v = dmin(v, partial[size_t(blockIdx.x) * nper + t]);
  double m = block_min(v, sm);
  if (threadIdx.x == 0) result[blockIdx.x] = m;
}

__global__ void __launch_bounds__(RBLK) k_estdt(const BoxDev* boxes, int nbox, const Fab* st,
                                                Closures c, double* partial) {
  __shared__ double sm[RBLK];
  double dt = 1.79769313486231570815e+308;
  for (int b = 0; b < nbox; ++b) {
    const BoxDev bx = boxes[b];
    const double dx0 = bx.dx[0], dx1 = bx.dx[1], dx2 = bx.dx[2];
    const Fab u = st[b];
    int nx = bx.hi[0] - bx.lo[0] + 1, ny = bx.hi[1] - bx.lo[1] + 1, nz = bx.hi[2] - bx.lo[2] + 1;
    int64_t n = int64_t(nx) * ny * nz;
    for (int64_t t = int64_t(blockIdx.x) * RBLK + threadIdx.x; t < n; t += int64_t(gridDim.x) * RBLK) {
      int i = bx.lo[0] + int(t % nx);
      int64_t r = t / nx;
      int j = bx.lo[1] + int(r % ny);
      int k = bx.lo[2] + int(r / ny);
      dt = dmin(dt, cell_estdt(u, i, j, k, c, dx0, dx1, dx2));
    }
  }
  sm[threadIdx.x] = dt;
  __syncthreads();
  for (int s = RBLK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = dmin(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sm[0];
}

__global__ void __launch_bounds__(RBLK) k_min_final(const double* partial, int n, double* result) {
  __shared__ double sm[RBLK];
  double v = 1.79769313486231570815e+308;
  for (int t = threadIdx.x; t < n; t += RBLK) v = dmin(v, partial[t]);
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int s = RBLK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = dmin(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) result[0] = sm[0];
}

// Conserved totals: per-block sums of the five components (fixed reduction
// tree, so the result is reproducible run to run).
__global__ void __launch_bounds__(RBLK) k_totals(const BoxDev* boxes, int nbox, const Fab* st,
                                                 double* partial) {
  __shared__ double sm[5][RBLK];
  double acc[5] = {0, 0, 0, 0, 0};
  for (int b = 0; b < nbox; ++b) {
    const BoxDev bx = boxes[b];
    const Fab u = st[b];
    int nx = bx.hi[0] - bx.lo[0] + 1, ny = bx.hi[1] - bx.lo[1] + 1, nz = bx.hi[2] - bx.lo[2] + 1;
    int64_t n = int64_t(nx) * ny * nz;
    for (int64_t t = int64_t(blockIdx.x) * RBLK + threadIdx.x; t < n; t += int64_t(gridDim.x) * RBLK) {
      int i = bx.lo[0] + int(t % nx);
      int64_t r = t / nx;
      int j = bx.lo[1] + int(r % ny);
      int k = bx.lo[2] + int(r / ny);
#pragma unroll
      for (int c = 0; c < 5; ++c) acc[c] += u(i, j, k, c);
    }
  }
#pragma unroll
  for (int c = 0; c < 5; ++c) sm[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int s = RBLK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
#pragma unroll
      for (int c = 0; c < 5; ++c) sm[c][threadIdx.x] += sm[c][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < 5; ++c) partial[5 * blockIdx.x + c] = sm[c][0];
}

__global__ void __launch_bounds__(RBLK) k_sum_final(const double* partial, int n, double* result) {
  __shared__ double sm[5][RBLK];
  double acc[5] = {0, 0, 0, 0, 0};
  for (int t = threadIdx.x; t < n; t += RBLK)
#pragma unroll
    for (int c = 0; c < 5; ++c) acc[c] += partial[5 * t + c];
#pragma unroll
  for (int c = 0; c < 5; ++c) sm[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int s = RBLK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
#pragma unroll
      for (int c = 0; c < 5; ++c) sm[c][threadIdx.x] += sm[c][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < 5; ++c) result[c] = sm[c][0];
}

}  // namespace

hipError_t launch_fill_copy(const CopyTask* tasks, int ntask, const TaskBlock* blocks, int nblk, const Fab* fabs,
                            hipStream_t st) {
  if (nblk == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_copy, dim3(unsigned(nblk)), dim3(FBLK), 0, st, tasks, ntask, blocks, fabs);
  return hipGetLastError();
}

hipError_t launch_fill_bc(const BcTask* tasks, const TaskBlock* blocks, int nblk, const Fab* fabs,
                          const BoxDev* boxes, const WallFill& wf, hipStream_t st) {
  if (nblk == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_bc, dim3(unsigned(nblk)), dim3(FBLK), 0, st, tasks, blocks, fabs, boxes, wf);
  return hipGetLastError();
}

hipError_t launch_copy_valid(const BoxTable& bt, const Fab* src, const Fab* dst, const double* box_dt,
                             hipStream_t st) {
  int64_t mx = 1;
  for (auto& b : bt.h_boxes)
    mx = std::max<int64_t>(mx, int64_t(b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) * (b.hi[2] - b.lo[2] + 1));
  hipLaunchKernelGGL(k_copy_valid, dim3(unsigned((mx + FBLK - 1) / FBLK), unsigned(bt.nbox)), dim3(FBLK), 0, st,
                     bt.d_boxes, src, dst, box_dt);
  return hipGetLastError();
}

int reduce_blocks(const BoxTable& bt) {
  int64_t n = 0;
  for (auto& b : bt.h_boxes)
    n += int64_t(b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) * (b.hi[2] - b.lo[2] + 1);
  int64_t blocks = (n / bt.nbox + RBLK - 1) / RBLK;
  if (blocks < 1) blocks = 1;
  if (blocks > 1024) blocks = 1024;
  return int(blocks);
}

hipError_t launch_estdt(const BoxTable& bt, const Fab* st, const Closures& c, double* partial, double* result,
                        hipStream_t s) {
  int nb = reduce_blocks(bt);
  hipLaunchKernelGGL(k_estdt, dim3(nb), dim3(RBLK), 0, s, bt.d_boxes, bt.nbox, st, c, partial);
  hipLaunchKernelGGL(k_min_final, dim3(1), dim3(RBLK), 0, s, partial, nb, result);
  return hipGetLastError();
}

// Scratch of the no-allocation CFL reduction: one copy per GPU, held until
// the process ends.  Host code never passes the array's name to a kernel
// (on the host it is a stand-in, not a device address): it takes the
// current device's address once through noalloc_scratch_address() and
// passes that.  Every call on a GPU shares its copy; nothing in it needs to
// start at zero (the descriptors are copied in, each partial slot and the
// result are written before they are read).
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

hipError_t noalloc_scratch_address(void** ptr) { return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch)); }

hipError_t launch_estdt_scratch(const BoxTable& bt, const Closures& c, void* scratch_dev, double** result_out,
                                hipStream_t s) {
  char* base = static_cast<char*>(scratch_dev);
  const Fab* st = reinterpret_cast<const Fab*>(base);
  double* partial = reinterpret_cast<double*>(base + NOALLOC_DESC_BYTES);
  double* result = partial + NOALLOC_MAX_BLOCKS;
  const int nb = reduce_blocks(bt);
  if (bt.nbox > NOALLOC_MAX_BOXES || nb > NOALLOC_MAX_BLOCKS) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_estdt, dim3(nb), dim3(RBLK), 0, s, bt.d_boxes, bt.nbox, st, c, partial);
  hipLaunchKernelGGL(k_min_final, dim3(1), dim3(RBLK), 0, s, partial, nb, result);
  *result_out = result;
  return hipGetLastError();
}

// The time-step rule of CNS::computeInitialDt / computeNewDt
// (src/CNS.cpp:282-377) for one problem: the estimate on the first step,
// else min(estimate, 1.1 * previous step) (std::min order), cut to
// stop_time - t when t + dt > stop_time - 0.001 dt (stop_time >= 0).  A
// problem within 1e-14 of stop_time takes no step (0); a non-positive or
// NaN minimum marks the problem bad and stops it.
__device__ double clock_step(Clock& k, double emin, double cfl, double stop) {
  if (clock_done(k, stop)) return 0.0;
  if (!(emin > 0.0)) {
    k.bad = 1;
    return 0.0;
  }
  const double est = emin * cfl;  // CNS::estTimeStep, src/CNS.cpp:391
  double dt0 = est;
  if (!k.first) {
    dt0 = dmin(est, 1.1 * k.dt);
    const double eps = 0.001 * dt0;
    if (stop >= 0.0 && (k.t + dt0) > (stop - eps)) dt0 = stop - k.t;
  }
  k.t = k.t + dt0;
  k.dt = dt0;
  k.steps += 1;
  k.first = 0;
  return dt0;
}

// est holds nper partial minima per box (k_estdt_boxes' blocks); a box's
// estimate is their minimum, exact in any order
__global__ void __launch_bounds__(RBLK) k_dt_policy(int nbox, int batch, const double* est, int nper, Clock* clk,
                                                    double cfl, double stop, double* box_dt) {
  if (batch) {
    const int p = blockIdx.x * RBLK + threadIdx.x;
    if (p < nbox) {
      double m = est[size_t(p) * nper];
      for (int t = 1; t < nper; ++t) m = dmin(m, est[size_t(p) * nper + t]);
      box_dt[p] = clock_step(clk[p], m, cfl, stop);
    }
    return;
  }
  // one problem: the minimum over its boxes (exact in any order)
  __shared__ double sm[RBLK];
  double v = 1.79769313486231570815e+308;
  for (int b = threadIdx.x; b < nbox * nper; b += RBLK) v = dmin(v, est[b]);
  const double m = block_min(v, sm);
  __shared__ double s_dt;
  if (threadIdx.x == 0) s_dt = clock_step(clk[0], m, cfl, stop);
  __syncthreads();
  for (int b = threadIdx.x; b < nbox; b += RBLK) box_dt[b] = s_dt;
}

hipError_t launch_dt_policy(int nbox, bool batch, const double* est, int nper, Clock* clk, double cfl,
                            double stop_time, double* box_dt, hipStream_t s) {
  const int grid = batch ? (nbox + RBLK - 1) / RBLK : 1;
  hipLaunchKernelGGL(k_dt_policy, dim3(grid), dim3(RBLK), 0, s, nbox, batch ? 1 : 0, est, nper, clk, cfl, stop_time,
                     box_dt);
  return hipGetLastError();
}

// computeNewDt over levels (src/CNS.cpp:337-376): per level the estimate,
// limited by 1.1 times its previous step (computeInitialDt on the first
// call: the estimate alone); dt0 = min over levels of nfactor_l * dt_l with
// nfactor_l the product of n_cycle[0..l]; the stop-time cut of dt0; dt_l =
// dt0 / nfactor_l.  dt_level[nlev] receives the steps, dt_level[nlev] the
// smallest minimum found (<= 0 or NaN: a non-physical state).
__global__ void k_dt_levels(DtLevels a, double* dt_level) {
  if (threadIdx.x != 0) return;
  double dt0 = 1.79769313486231570815e+308;
  double worst = dt0;
  int nfactor = 1;
  for (int l = 0; l < a.nlev; ++l) {
    const double m = *a.est[l];
    worst = (m > 0.0) ? dmin(worst, m) : m;
    double dmn = m * a.cfl;
    if (!a.first) dmn = dmin(dmn, 1.1 * dt_level[l]);
    nfactor *= a.ncycle[l];
    dt0 = dmin(dt0, nfactor * dmn);
  }
  if (!a.first) {
    const double eps = 0.001 * dt0;
    if (a.stop_time >= 0.0 && (a.cur_time + dt0) > (a.stop_time - eps)) dt0 = a.stop_time - a.cur_time;
  }
  nfactor = 1;
  for (int l = 0; l < a.nlev; ++l) {
    nfactor *= a.ncycle[l];
    dt_level[l] = dt0 / nfactor;
  }
  dt_level[a.nlev] = worst;
}

hipError_t launch_dt_levels(const DtLevels& a, double* dt_level, hipStream_t s) {
  hipLaunchKernelGGL(k_dt_levels, dim3(1), dim3(64), 0, s, a, dt_level);
  return hipGetLastError();
}

// printTotal's primitive extrema and pointCFL's wave speeds
// (src/CNS.cpp:544-547, src/cls/CNS_K.h:49-59) over the valid cells: per
// block 6 minima and 6 maxima of (rho, u, v, w, T, p) from cons2prim, then
// per direction max |u_d + c| and max |u_d - c| with c = sqrt(gamma R T),
// and max over cells of max(|u_d + c|, |u_d - c|) * dt / dx_d (the box's
// dx, so a batch's problems each use their own).  partial: 21 per block.
constexpr int NDIAG = 21;
__global__ void __launch_bounds__(RBLK) k_diag(const BoxDev* boxes, int nbox, const Fab* st, Closures c, double dt,
                                               double* partial) {
  __shared__ double sm[RBLK];
  double v[NDIAG];
  for (int q = 0; q < 6; ++q) {
    v[q] = 1.79769313486231570815e+308;
    v[6 + q] = -1.79769313486231570815e+308;
  }
  for (int q = 12; q < NDIAG; ++q) v[q] = 0.0;
  for (int b = 0; b < nbox; ++b) {
    const BoxDev bx = boxes[b];
    const Fab u = st[b];
    int nx = bx.hi[0] - bx.lo[0] + 1, ny = bx.hi[1] - bx.lo[1] + 1, nz = bx.hi[2] - bx.lo[2] + 1;
    int64_t n = int64_t(nx) * ny * nz;
    for (int64_t t = int64_t(blockIdx.x) * RBLK + threadIdx.x; t < n; t += int64_t(gridDim.x) * RBLK) {
      int i = bx.lo[0] + int(t % nx);
      int64_t r = t / nx;
      int j = bx.lo[1] + int(r % ny);
      int k = bx.lo[2] + int(r / ny);
      double q[NPRIM];
      cons2prim(u(i, j, k, 0), u(i, j, k, 1), u(i, j, k, 2), u(i, j, k, 3), u(i, j, k, 4), c, q);
      for (int m = 0; m < 6; ++m) {
        v[m] = dmin(v[m], q[m]);
        v[6 + m] = dmax(v[6 + m], q[m]);
      }
      const double sos = sqrt(c.gamma * c.Rspec * q[QT]);
      for (int d = 0; d < 3; ++d) {
        const double a = fabs(q[QU + d] + sos), bm = fabs(q[QU + d] - sos);
        v[12 + 3 * d] = dmax(v[12 + 3 * d], a);
        v[13 + 3 * d] = dmax(v[13 + 3 * d], bm);
        v[14 + 3 * d] = dmax(v[14 + 3 * d], dmax(a, bm) * dt / bx.dx[d]);
      }
    }
  }
  for (int q = 0; q < NDIAG; ++q) {
    sm[threadIdx.x] = v[q];
    __syncthreads();
    for (int s = RBLK / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s)
        sm[threadIdx.x] = q < 6 ? dmin(sm[threadIdx.x], sm[threadIdx.x + s]) : dmax(sm[threadIdx.x], sm[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[size_t(NDIAG) * blockIdx.x + q] = sm[0];
    __syncthreads();
  }
}

__global__ void __launch_bounds__(RBLK) k_diag_final(const double* partial, int n, double* result) {
  const int q = threadIdx.x;
  if (q >= NDIAG) return;
  double v = partial[q];
  for (int t = 1; t < n; ++t) v = q < 6 ? dmin(v, partial[size_t(NDIAG) * t + q]) : dmax(v, partial[size_t(NDIAG) * t + q]);
  result[q] = v;
}

hipError_t launch_diag(const BoxTable& bt, const Fab* st, const Closures& c, double dt, double* partial,
                       double* result, hipStream_t s) {
  const int nb = reduce_blocks(bt);
  hipLaunchKernelGGL(k_diag, dim3(nb), dim3(RBLK), 0, s, bt.d_boxes, bt.nbox, st, c, dt, partial);
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, s, partial, nb, result);
  return hipGetLastError();
}

// Non-physical states: flag[0] = lowest box with a valid cell whose
// component is NaN or infinite, whose density is <= 0 or whose pressure
// (gamma - 1)(E - |m|^2 / (2 rho)) is <= 0 (flag[0] = nbox when none).
__global__ void __launch_bounds__(RBLK) k_state_check(const BoxDev* boxes, const Fab* st, Closures c, int* flag) {
  const int b = blockIdx.y;
  const BoxDev bx = boxes[b];
  const Fab u = st[b];
  int nx = bx.hi[0] - bx.lo[0] + 1, ny = bx.hi[1] - bx.lo[1] + 1, nz = bx.hi[2] - bx.lo[2] + 1;
  int64_t n = int64_t(nx) * ny * nz;
  bool bad = false;
  for (int64_t t = int64_t(blockIdx.x) * RBLK + threadIdx.x; t < n && !bad; t += int64_t(gridDim.x) * RBLK) {
    int i = bx.lo[0] + int(t % nx);
    int64_t r = t / nx;
    int j = bx.lo[1] + int(r % ny);
    int k = bx.lo[2] + int(r / ny);
    double w[5];
    for (int m = 0; m < 5; ++m) {
      w[m] = u(i, j, k, m);
      if (!isfinite(w[m])) bad = true;
    }
    const double p = c.gm1 * (w[4] - 0.5 * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]) / w[0]);
    if (!(w[0] > 0.0) || !(p > 0.0)) bad = true;
  }
  if (bad) atomicMin(flag, b);
}

hipError_t launch_state_check(const BoxTable& bt, const Fab* st, const Closures& c, int* flag, hipStream_t s) {
  const int nper = estdt_box_blocks(bt);
  hipLaunchKernelGGL(k_state_check, dim3(nper, bt.nbox), dim3(RBLK), 0, s, bt.d_boxes, st, c, flag);
  return hipGetLastError();
}

int estdt_box_blocks(const BoxTable& bt) {
  int64_t mx = 1;
  for (auto& b : bt.h_boxes)
    mx = std::max<int64_t>(mx, int64_t(b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) * (b.hi[2] - b.lo[2] + 1));
  int64_t nb = (mx + RBLK * 8 - 1) / (RBLK * 8);
  return int(std::max<int64_t>(1, std::min<int64_t>(nb, 64)));
}

hipError_t launch_estdt_boxes(const BoxTable& bt, const Fab* st, const Closures& c, double* partial, double* result,
                              hipStream_t s, const Clock* clk, bool batch, double stop_time) {
  int nper = estdt_box_blocks(bt);
  hipLaunchKernelGGL(k_estdt_boxes, dim3(nper, bt.nbox), dim3(RBLK), 0, s, bt.d_boxes, st, c, partial, clk,
                     batch ? 1 : 0, stop_time);
  if (result) hipLaunchKernelGGL(k_min_boxes, dim3(bt.nbox), dim3(RBLK), 0, s, partial, nper, result);
  return hipGetLastError();
}

// Coarse-fine interpolation (conservative linear with one limiter factor per
// direction shared by the components): one lane per coarse cell, which reads
// its seven-point stencil once (from the FAB of the coarse box that owns it,
// ghost cells filled by the caller) and writes its children inside the
// task's fine region.
template <bool BLEND>
__global__ void __launch_bounds__(FBLK) k_interp_cf(const InterpTask* tasks, const TaskBlock* blocks, const Fab* c0,
                                                    const Fab* c1, double a, double b, int r, const Fab* fine) {
  const TaskBlock tb = blocks[blockIdx.x];
  const InterpTask& T = tasks[tb.task];
  int ic[3];
  if (!task_cell(T, tb, threadIdx.x, ic)) return;
  const Fab U0 = c0[T.src];
  const Fab U1 = BLEND ? c1[T.src] : U0;
  auto U = [&](int di, int dj, int dk, int n) -> double {
    const int i = ic[0] + di, j = ic[1] + dj, k = ic[2] + dk;
    if (BLEND) return a * U0(i, j, k, n) + b * U1(i, j, k, n);
    return U0(i, j, k, n);
  };
  double u0[NCONS], sl[3][NCONS];
  double sf[3] = {1.0, 1.0, 1.0};
#pragma unroll
  for (int n = 0; n < NCONS; ++n) {
    u0[n] = U(0, 0, 0, n);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int ex = d == 0, ey = d == 1, ez = d == 2;
      const double um = U(-ex, -ey, -ez, n), up = U(ex, ey, ez, n);
      const double dc = 0.5 * (up - um);
      const double df = 2.0 * (up - u0[n]);
      const double db = 2.0 * (u0[n] - um);
      double s = (df * db >= 0.0) ? dmin(fabs(df), fabs(db)) : 0.0;
      s = copysign(1.0, dc) * dmin(s, fabs(dc));
      if (dc != 0.0) sf[d] = dmin(sf[d], s / dc);
      sl[d][n] = dc;
    }
  }
#pragma unroll
  for (int n = 0; n < NCONS; ++n)
#pragma unroll
    for (int d = 0; d < 3; ++d) sl[d][n] = sl[d][n] * sf[d];
  const Fab F = fine[T.dst];
  int wlo[3], whi[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    wlo[d] = max(ic[d] * r, T.flo[d]);
    whi[d] = min(ic[d] * r + r - 1, T.fhi[d]);
  }
  const double rr = double(r);
  for (int wk = wlo[2]; wk <= whi[2]; ++wk) {
    const double oz = (double(wk - ic[2] * r) + 0.5) / rr - 0.5;
    for (int wj = wlo[1]; wj <= whi[1]; ++wj) {
      const double oy = (double(wj - ic[1] * r) + 0.5) / rr - 0.5;
      for (int wi = wlo[0]; wi <= whi[0]; ++wi) {
        const double ox = (double(wi - ic[0] * r) + 0.5) / rr - 0.5;
#pragma unroll
        for (int n = 0; n < NCONS; ++n)
          F(wi + T.sh[0], wj + T.sh[1], wk + T.sh[2], n) = u0[n] + ox * sl[0][n] + oy * sl[1][n] + oz * sl[2][n];
      }
    }
  }
}

// Average-down: one lane per coarse cell, children summed x fastest.
__global__ void __launch_bounds__(FBLK) k_avg_down(const CopyTask* tasks, const TaskBlock* blocks, const Fab* fine,
                                                   const Fab* crse, int r, double volfrac) {
  const TaskBlock tb = blocks[blockIdx.x];
  const CopyTask& T = tasks[tb.task];
  int c[3];
  if (!task_cell(T, tb, threadIdx.x, c)) return;
  const Fab F = fine[T.src], C = crse[T.dst];
  for (int n = 0; n < NCONS; ++n) {
    double s = 0.0;
    for (int kr = 0; kr < r; ++kr)
      for (int jr = 0; jr < r; ++jr)
        for (int ir = 0; ir < r; ++ir) s += F(c[0] * r + ir, c[1] * r + jr, c[2] * r + kr, n);
    C(c[0], c[1], c[2], n) = volfrac * s;
  }
}

hipError_t launch_plot_vars(const BoxTable& bt, const Fab* src, const Fab* dst, int state_vars, int mask,
                            const Closures& c, hipStream_t st) {
  int64_t mx = 0;
  for (auto& b : bt.h_boxes)
    mx = std::max<int64_t>(mx, int64_t(b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) * (b.hi[2] - b.lo[2] + 1));
  hipLaunchKernelGGL(k_plot_vars, dim3(unsigned((mx + FBLK - 1) / FBLK), unsigned(bt.nbox)), dim3(FBLK), 0, st,
                     bt.d_boxes, src, dst, state_vars, mask, c.gm1, c.cv);
  return hipGetLastError();
}

hipError_t launch_pack_valid(const BoxTable& bt, const Fab* src, int ncomp, const int64_t* off, double* out,
                             hipStream_t st) {
  int64_t mx = 0;
  for (auto& b : bt.h_boxes)
    mx = std::max<int64_t>(mx, int64_t(b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) * (b.hi[2] - b.lo[2] + 1));
  hipLaunchKernelGGL(k_pack_valid, dim3(unsigned((mx + FBLK - 1) / FBLK), unsigned(bt.nbox)), dim3(FBLK), 0, st,
                     bt.d_boxes, src, ncomp, off, out);
  return hipGetLastError();
}

hipError_t launch_interp_cf(const InterpTask* tasks, const TaskBlock* blocks, int nblk, const Fab* c0, const Fab* c1,
                            bool blend, double a, double b, int r, const Fab* fine, hipStream_t st) {
  if (nblk == 0) return hipSuccess;
  if (blend)
    hipLaunchKernelGGL(k_interp_cf<true>, dim3(unsigned(nblk)), dim3(FBLK), 0, st, tasks, blocks, c0, c1, a, b, r,
                       fine);
  else
    hipLaunchKernelGGL(k_interp_cf<false>, dim3(unsigned(nblk)), dim3(FBLK), 0, st, tasks, blocks, c0, c0, a, b, r,
                       fine);
  return hipGetLastError();
}

hipError_t launch_avg_down(const CopyTask* tasks, const TaskBlock* blocks, int nblk, const Fab* fine,
                           const Fab* crse, int r, hipStream_t st) {
  if (nblk == 0) return hipSuccess;
  const double volfrac = 1.0 / double(r * r * r);
  hipLaunchKernelGGL(k_avg_down, dim3(unsigned(nblk)), dim3(FBLK), 0, st, tasks, blocks, fine, crse, r, volfrac);
  return hipGetLastError();
}

hipError_t launch_copy_regions(const CopyTask* tasks, const TaskBlock* blocks, int nblk, const Fab* src,
                               const Fab* dst0, const Fab* dst1, hipStream_t st) {
  if (nblk == 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_regions, dim3(unsigned(nblk)), dim3(FBLK), 0, st, tasks, blocks, src, dst0, dst1);
  return hipGetLastError();
}

hipError_t launch_totals(const BoxTable& bt, const Fab* st, double* partial, double* result, hipStream_t s) {
  int nb = reduce_blocks(bt);
  hipLaunchKernelGGL(k_totals, dim3(nb), dim3(RBLK), 0, s, bt.d_boxes, bt.nbox, st, partial);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(RBLK), 0, s, partial, nb, result);
  return hipGetLastError();
}

}  // namespace amr
