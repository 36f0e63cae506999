// Error estimation for regridding: the per-cell taggers CNS::errorEst runs
// (src/CNS.cpp:469-507 calls the problem's user_tagging on every cell of
// every box of a FillPatch'ed state) into a byte map over the level's
// domain, and the reduction of that map to the clustering blocks.
#include "kernels.h"

namespace amr {

namespace {

constexpr int TBLK = 256;

__device__ __forceinline__ int64_t dom_index(const TagArgs& ta, int i, int j, int k) {
  return (int64_t(k - ta.dlo[2]) * ta.n[1] + (j - ta.dlo[1])) * ta.n[0] + (i - ta.dlo[0]);
}

// One lane per valid cell of box blockIdx.y (grid-stride in x).
__global__ void __launch_bounds__(TBLK) k_tag_cells(const BoxDev* boxes, const Fab* st, TagArgs ta,
                                                    unsigned char* tags) {
  const int b = blockIdx.y;
  const BoxDev bx = boxes[b];
  const Fab u = st[b];
  const int nx = bx.hi[0] - bx.lo[0] + 1, ny = bx.hi[1] - bx.lo[1] + 1, nz = bx.hi[2] - bx.lo[2] + 1;
  const int64_t n = int64_t(nx) * ny * nz;
  for (int64_t t = int64_t(blockIdx.x) * TBLK + threadIdx.x; t < n; t += int64_t(gridDim.x) * TBLK) {
    const int i = bx.lo[0] + int(t % nx);
    const int64_t r = t / nx;
    const int j = bx.lo[1] + int(r % ny);
    const int k = bx.lo[2] + int(r / ny);
    if (ta.kind == AMR_TAG_DENGRAD) {
      // exm/Sods/prob.h:187-197: |rho(i+1) - rho(i-1)| / rho(i) above the
      // threshold tags cells i .. i+3 (those past the domain are dropped)
      const double drhox = fabs(u(i + 1, j, k, URHO) - u(i - 1, j, k, URHO)) / u(i, j, k, URHO);
      if (drhox > ta.thr)
        for (int m = 0; m < 4; ++m)
          if (i + m <= ta.dlo[0] + ta.n[0] - 1) tags[dom_index(ta, i + m, j, k)] = 1;
    } else if (ta.kind == AMR_TAG_WALLBAND) {
      // exm/Channel/prob.h:199-209: y = (j + 1/2) dy (prob_lo is not added)
      // within factor*h of y = 0 or of prob_hi
      const double y = (j + 0.5) * ta.dx[1];
      if (y < ta.h * ta.factor || y > ta.prob_hi1 - ta.h * ta.factor) tags[dom_index(ta, i, j, k)] = 1;
    } else {
      // |curl u| by central differences of u = (rho u)/rho above the threshold
      auto vel = [&](int a, int bb, int c, int m) { return u(a, bb, c, m) / u(a, bb, c, URHO); };
      const double h0 = 0.5 / ta.dx[0], h1 = 0.5 / ta.dx[1], h2 = 0.5 / ta.dx[2];
      const double dwdy = (vel(i, j + 1, k, UMZ) - vel(i, j - 1, k, UMZ)) * h1;
      const double dvdz = (vel(i, j, k + 1, UMY) - vel(i, j, k - 1, UMY)) * h2;
      const double dudz = (vel(i, j, k + 1, UMX) - vel(i, j, k - 1, UMX)) * h2;
      const double dwdx = (vel(i + 1, j, k, UMZ) - vel(i - 1, j, k, UMZ)) * h0;
      const double dvdx = (vel(i + 1, j, k, UMY) - vel(i - 1, j, k, UMY)) * h0;
      const double dudy = (vel(i, j + 1, k, UMX) - vel(i, j - 1, k, UMX)) * h1;
      const double wx = dwdy - dvdz, wy = dudz - dwdx, wz = dvdx - dudy;
      const double mag = sqrt(wx * wx + wy * wy + wz * wz);
      if (mag > ta.thr) tags[dom_index(ta, i, j, k)] = 1;
    }
  }
}

__global__ void __launch_bounds__(TBLK) k_clear_bytes(unsigned char* p, int64_t n) {
  for (int64_t t = int64_t(blockIdx.x) * TBLK + threadIdx.x; t < n; t += int64_t(gridDim.x) * TBLK) p[t] = 0;
}

// One lane per block of cb^3 cells: 1 when any cell of the block is tagged.
__global__ void __launch_bounds__(TBLK) k_tag_blocks(const unsigned char* tags, TagArgs ta, int cb,
                                                     unsigned char* out) {
  const int nb0 = ta.n[0] / cb, nb1 = ta.n[1] / cb, nb2 = ta.n[2] / cb;
  const int64_t nbk = int64_t(nb0) * nb1 * nb2;
  for (int64_t t = int64_t(blockIdx.x) * TBLK + threadIdx.x; t < nbk; t += int64_t(gridDim.x) * TBLK) {
    const int bi = int(t % nb0);
    const int64_t r = t / nb0;
    const int bj = int(r % nb1), bk = int(r / nb1);
    unsigned char any = 0;
    for (int k = bk * cb; k < (bk + 1) * cb && !any; ++k)
      for (int j = bj * cb; j < (bj + 1) * cb && !any; ++j)
        for (int i = bi * cb; i < (bi + 1) * cb; ++i)
          if (tags[(int64_t(k) * ta.n[1] + j) * ta.n[0] + i]) {
            any = 1;
            break;
          }
    out[t] = any;
  }
}

int grid_for(int64_t n) { return int(std::max<int64_t>(1, std::min<int64_t>((n + TBLK - 1) / TBLK, 4096))); }

}  // namespace

hipError_t launch_tag_cells(const BoxTable& bt, const Fab* st, const TagArgs& ta, unsigned char* tags,
                            hipStream_t s) {
  const int64_t ndom = int64_t(ta.n[0]) * ta.n[1] * ta.n[2];
  hipLaunchKernelGGL(k_clear_bytes, dim3(grid_for(ndom)), dim3(TBLK), 0, s, tags, ndom);
  int64_t mx = 1;
  for (auto& b : bt.h_boxes)
    mx = std::max<int64_t>(mx, int64_t(b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) * (b.hi[2] - b.lo[2] + 1));
  const int nper = int(std::min<int64_t>((mx + TBLK - 1) / TBLK, 256));
  hipLaunchKernelGGL(k_tag_cells, dim3(nper, bt.nbox), dim3(TBLK), 0, s, bt.d_boxes, st, ta, tags);
  return hipGetLastError();
}

hipError_t launch_tag_blocks(const unsigned char* tags, const TagArgs& ta, int cb, unsigned char* out,
                             hipStream_t s) {
  const int64_t nbk = int64_t(ta.n[0] / cb) * (ta.n[1] / cb) * (ta.n[2] / cb);
  hipLaunchKernelGGL(k_tag_blocks, dim3(grid_for(nbk)), dim3(TBLK), 0, s, tags, ta, cb, out);
  return hipGetLastError();
}

}  // namespace amr
