// Launch regions of the pass-structured kernels: a region of box b is
// [lo + glo, hi + ghi] per direction; one launch covers the region of every
// box through a block map (block -> box, block within the box).
#pragma once

#include "kernels.h"

namespace amr {

constexpr int BLK = 256;

// Region of box b: [lo + glo, hi + ghi] per direction.
struct Region {
  int glo[3], ghi[3];
};

__device__ __forceinline__ bool region_cell(const BoxDev& bx, const Region& r, int64_t t, int& i,
                                            int& j, int& k) {
  int lo0 = bx.lo[0] + r.glo[0], lo1 = bx.lo[1] + r.glo[1], lo2 = bx.lo[2] + r.glo[2];
  int nx = bx.hi[0] + r.ghi[0] - lo0 + 1;
  int ny = bx.hi[1] + r.ghi[1] - lo1 + 1;
  int nz = bx.hi[2] + r.ghi[2] - lo2 + 1;
  if (t >= int64_t(nx) * ny * nz) return false;
  i = lo0 + int(t % nx);
  int64_t r2 = t / nx;
  j = lo1 + int(r2 % ny);
  k = lo2 + int(r2 / ny);
  return true;
}

#define REGION_LOOP_BEGIN                                                   \
  const BlockRef br = bmap[blockIdx.x];                                     \
  const int b = br.box;                                                     \
  if (box_dt != nullptr && !(box_dt[b] > 0.0)) return;                      \
  const BoxDev bx = boxes[b];                                               \
  int i, j, k;                                                              \
  if (!region_cell(bx, reg, int64_t(br.blk) * BLK + threadIdx.x, i, j, k)) return;


// Block map of region r over all boxes: block -> (box, block within box), so a
// launch over uneven boxes carries no empty blocks.
struct Launch {
  const BlockRef* map;
  dim3 grid;
};

inline std::array<int, 6> region_key(const Region& r) {
  return {r.glo[0], r.glo[1], r.glo[2], r.ghi[0], r.ghi[1], r.ghi[2]};
}

inline hipError_t build_map(const BoxTable& bt, const Region& r) {
  const auto key = region_key(r);
  if (bt.block_maps.count(key)) return hipSuccess;
  std::vector<BlockRef> h;
  for (int b = 0; b < bt.nbox; ++b) {
    int64_t n = 1;
    for (int d = 0; d < 3; ++d)
      n *= int64_t(bt.h_boxes[b].hi[d] + r.ghi[d] - bt.h_boxes[b].lo[d] - r.glo[d] + 1);
    const int nblk = int((n + BLK - 1) / BLK);
    for (int k = 0; k < nblk; ++k) h.push_back(BlockRef{b, k});
  }
  BlockRef* d = nullptr;
  hipError_t e = hipMalloc(&d, sizeof(BlockRef) * h.size());
  if (e != hipSuccess) return e;
  e = hipMemcpy(d, h.data(), sizeof(BlockRef) * h.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(d);
    return e;
  }
  bt.block_maps.emplace(key, std::make_pair(d, int(h.size())));
  return hipSuccess;
}

// Lookup only (no allocation), so launches can be recorded into a graph.
inline Launch grid_for(const BoxTable& bt, const Region& r) {
  auto it = bt.block_maps.find(region_key(r));
  if (it == bt.block_maps.end()) return Launch{nullptr, dim3(0)};
  return Launch{it->second.first, dim3(unsigned(it->second.second))};
}

inline Region grown(int g) { return Region{{-g, -g, -g}, {g, g, g}}; }
inline Region nodal() { return Region{{0, 0, 0}, {1, 1, 1}}; }
inline Region slope_region(int d) {
  Region r{{0, 0, 0}, {0, 0, 0}};
  r.glo[d] = -1;
  r.ghi[d] = 1;
  return r;
}
inline Region face_region(int d) {
  Region r{{0, 0, 0}, {0, 0, 0}};
  r.ghi[d] = 1;
  return r;
}


// cons2prim of `sin` into w.prims on region r of every box.
hipError_t launch_prims_region(const BoxTable& bt, const UnfusedWork& w, const Closures& c, const double* box_dt,
                               const Fab* sin, const Region& r, hipStream_t st);

}  // namespace amr
