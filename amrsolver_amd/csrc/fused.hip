// Fused KEEP (order 2/4/6) + viscous RHS and SSPRK stage update.
//
// One workgroup owns a TX x TY column tile of one box and marches up a
// run of planes (2.5D blocking).  Per plane it
//   1. converts the next plane's conserved state (tile + halo) to primitives
//      straight into LDS (cons2prim, CNS_hydro_K.h:15-34) -- primitives never
//      touch HBM;
//   2. computes cell-centred viscous fluxes (ViscousFluxes, CNS_hydro_K.h:108-172)
//      for the tile plus a one-cell ring, x/y parts into LDS, z part in
//      registers;
//   3. computes every x and y face flux of the plane once (KEEP,
//      CentralKEEP.h:40-145, plus the viscous face average,
//      CNS_hydro_K.h:254-261) into LDS;
//   4. computes the z face below the plane from a per-thread register queue
//      of its own column (order 6 also reads the column two planes up, a
//      load issued one plane ahead) and finalises the plane below: flux
//      difference
//      (CNS_advance.cpp:384-390) and stage combination (CNS::advance
//      LinComb/Saxpy/Xpay) written straight to the output.
// HBM traffic per cell and stage: the input state (plus tile halos, mostly
// served from L2), U^n when the stage needs it, and the output.
#include "kernels.h"
#include "fused_common.h"

namespace amr {

namespace {

using namespace fusedk;

template <int H, bool VISC, int TX, int TY>
struct Cfg {
  static constexpr int HX = (VISC && H < 2) ? 2 : H;  // xy halo
  static constexpr int KB = (VISC && H < 2) ? 2 : H;  // planes below k0 loaded
  // own-column queue c-1, c, c+1 (KEEP4 keeps its c-2 pair in a carry)
  static constexpr int NT = TX * TY;
  static constexpr int RX = TX + 2 * HX, RY = TY + 2 * HX;
  static constexpr int NREG = RX * RY;
  static constexpr int PL = NREG;  // LDS plane stride per primitive
  static constexpr int NLOAD = (NREG + NT - 1) / NT;
  static constexpr int NRING = 2 * TX + 2 * TY;
  // 16x16 viscous tiles (four waves) spread the work beyond the own cells
  // over the waves: the ring's viscous fluxes on wave 3 (one pass for both
  // directions), the KEEP part of the extra faces at i = TX and j = TY on
  // wave 0 before the viscous barrier (one 32-lane pass for both
  // directions), their viscous part on wave 1 after it
  static constexpr bool EARLY = VISC && NT == 256 && NRING == 64;
  static constexpr int VA = 64;  // first thread of the extra faces' viscous part
  // other shapes: threads of the extra x faces (i = TX) and y faces (j = TY),
  // waves of their own (the ring fluxes occupy wave 0) where the workgroup
  // has them
  static constexpr int XB = NT >= 128 ? 64 * ((NRING + 63) / 64) : (VISC ? NRING : 0);
  static constexpr int YB = NT >= XB + 128 ? XB + 64 : XB + TY;
  // first ring thread
  static constexpr int RB = EARLY ? NT - 64 : 0;
  // halo cells that are not ring cells (outer ring and the ring's corners)
  static constexpr int NOUT = NREG - NT - NRING;
};

// LDS cell of a thread's second load slot in the self-loading form (16x16
// viscous tiles): the ring threads load their ring cell, threads 64 ..
// 64 + NOUT - 1 (waves 1 and 2) the other halo cells (rows 0 and RY-1,
// columns 0 and RX-1, then the ring's four corners); -1 for threads with no
// second cell
template <class C, int TX>
__device__ __forceinline__ int self_halo_cell(int tid, int ring_idx) {
  constexpr int RX = C::RX, RY = C::RY;
  constexpr int OB = 64;  // first thread of the other halo cells
  static_assert(C::EARLY && C::RB + C::NRING == C::NT && OB + C::NOUT <= C::RB,
                "self-loading layout needs the ring on the last wave");
  if (tid >= C::RB) return ring_idx;
  if (tid < OB || tid >= OB + C::NOUT) return -1;
  const int o = tid - OB;
  int i, j;
  if (o < RX) { i = o; j = 0; }
  else if (o < 2 * RX) { i = o - RX; j = RY - 1; }
  else if (o < 2 * RX + RY - 2) { i = 0; j = o - 2 * RX + 1; }
  else if (o < 2 * RX + 2 * (RY - 2)) { i = RX - 1; j = o - 2 * RX - (RY - 2) + 1; }
  else {
    const int q = o - 2 * RX - 2 * (RY - 2);
    i = (q & 1) ? RX - 2 : 1;
    j = (q & 2) ? RY - 2 : 1;
  }
  return j * RX + i;
}

// SELF: the self-loading form (every cell a thread reads from the plane
// above its own is one it stored, so storing that plane needs no barrier;
// 16x16 viscous tiles).  SRC: the stage may carry a user_source array or
// a body force (the code for them is compiled out otherwise).  NB: tiles
// per box read every halo cell that another box of the level holds (its
// periodic image included) from that box's array, through an owner map of
// the level's blocks; halo cells no box holds come from the box's own ghost
// cells.  The stage inputs then need no same-level ghost copies.  The
// boxes' arrays lie within 2^31 elements of each other, so a slot's source
// is an element offset from the own array; NB 2: they share their strides
// (the registers of the plain form); NB 1: plane and component strides per
// slot too.
// At least two waves per SIMD for orders 2 and 4: the compiler would
// otherwise take registers past 256 (into AGPRs) rather than reuse them,
// halving the occupancy.  Order 6 holds more (six pairs per x/y face, two
// more z carries): one wave per SIMD with up to 512 registers rather than
// spills to scratch memory.
template <int H, bool VISC, int TX, int TY, bool UCS, bool SELF, bool SRC, int NB>
__global__ void __launch_bounds__(TX* TY) __attribute__((amdgpu_waves_per_eu(H == 3 ? 1 : 2))) k_fused(const TileDev* __restrict__ tiles, const BoxDev* __restrict__ boxes,
                                                     Closures c, double co0, double co1, double co2, StageArgs sa,
                                                     WrapArgs wa, Cover cov, UncovSrc us, NbMap nbm,
                                                     const Fab* __restrict__ s1f, const Fab* __restrict__ sinf,
                                                     const Fab* __restrict__ outf) {
  using C = Cfg<H, VISC, TX, TY>;
  static_assert(!NB || (!UCS && H <= 2), "neighbour reads take tiles per box of KEEP order 2 or 4");
  constexpr int RX = C::RX, HX = C::HX, PL = C::PL, NT = C::NT;
  static_assert(!SELF || (VISC && TX == 16 && TY == 16 && C::NLOAD == 2 && C::EARLY),
                "the self-loading form takes 16x16 viscous tiles");
  __shared__ double s_q[2][NPRIM * PL];
  __shared__ double s_px[VISC ? 4 : 1][TY][TX + 2];
  __shared__ double s_py[VISC ? 4 : 1][TY + 2][TX];
  __shared__ double s_fx[5][TY][TX + 1];
  __shared__ double s_fy[5][TY + 1][TX];
  __shared__ unsigned s_um[UCS ? C::NLOAD * NT : 1];
  // KEEP part of the extra faces, across barrier (c) (self-loading form:
  // s_fx/s_fy column TX / row TY may still be read for the plane below)
  __shared__ double s_fe[SELF ? 5 : 1][SELF ? TX + TY : 1];

  const TileDev T = tiles[blockIdx.x];
  if (T.nz <= 0 || !box_active(sa, T.box)) return;  // uniform: the whole workgroup leaves
  const StageCoef sc = box_stage(sa, T.box);
  const double dxi0 = boxes[T.box].dxi[0], dxi1 = boxes[T.box].dxi[1], dxi2 = boxes[T.box].dxi[2];
  const Fab S = sinf[T.box];
  const Fab O = outf[T.box];
  const bool use_s1 = sc.mode != 2 && sc.a != 0.0;
  // without s1 the epilogue still loads from sin (cache hits of the sv
  // loads): every iteration then issues the same loads in the same order,
  // and the compiler's load counters never wait on the plane prefetch
  const Fab S1 = use_s1 ? s1f[T.box] : S;
  const bool has_src = SRC && sa.src != nullptr;
  const Fab SR = has_src ? sa.src[T.box] : S;  // user_source FAB (valid cells)
  const int tid = threadIdx.x;
  const int ti = tid % TX, tj = tid / TX;
  const int gi = T.i0 + ti, gj = T.j0 + tj;
  // a tile may overhang its box (or, with level tiling, the boxes' bounding
  // box) in x and y: loads are clamped to the ghosted extent and only cells
  // inside are written; with level tiling only covered cells
  const int xhi = cov.on ? cov.hi[0] : boxes[T.box].hi[0], yhi = cov.on ? cov.hi[1] : boxes[T.box].hi[1];
  const int xcap = xhi + NGHOST, ycap = yhi + NGHOST;
  const bool own_in = gi <= xhi && gj <= yhi;
  // coverage bitmap (one all-ones word when every cell is covered)
  guint* cbits = (guint*)cov.bits;
  const int cplane = cov.nb[0] * cov.nb[1];
  const int ccol = own_in ? max((gj - cov.lo[1]) >> cov.sh, 0) * cov.nb[0] + max((gi - cov.lo[0]) >> cov.sh, 0) : 0;
  const int k0 = T.k0, nz = T.nz;
  const int kfirst = k0 - C::KB;
  const double co[3] = {co0, co1, co2};
  const double cv = c.cv;
  const int own = (tj + HX) * RX + (ti + HX);  // LDS index of the own cell
  // (a-b)*0.5*dxinv is evaluated as (a-b)*(0.5*dxinv): exact in IEEE
  // arithmetic for normal numbers (scaling by 0.5 is exact)
  const double hx = 0.5 * dxi0, hy = 0.5 * dxi1, hz = 0.5 * dxi2;

  // ring cell of this thread (viscous x-flux at i=-1,TX or y-flux at j=-1,TY)
  int ring_idx = 0, ring_kind = -1, ring_a = 0, ring_b = 0;
  if (VISC && tid >= C::RB && tid < C::RB + C::NRING) {
    int r = tid - C::RB;
    if (r < TY) { ring_kind = 0; ring_a = -1; ring_b = r; }
    else if (r < 2 * TY) { ring_kind = 0; ring_a = TX; ring_b = r - TY; }
    else if (r < 2 * TY + TX) { ring_kind = 1; ring_a = r - 2 * TY; ring_b = -1; }
    else { ring_kind = 1; ring_a = r - 2 * TY - TX; ring_b = TY; }
    ring_idx = (ring_b + HX) * RX + (ring_a + HX);
  }

  // LDS cell of each load slot (-1: loaded, not stored).  First form: slot
  // s holds cell tid + s*NT.  Self-loading form: slot 0 is the own cell,
  // slot 1 the cell self_halo_cell gives.
  int scell[C::NLOAD];
#pragma unroll
  for (int s = 0; s < C::NLOAD; ++s) {
    if constexpr (SELF) scell[s] = s == 0 ? own : self_halo_cell<C, TX>(tid, ring_idx);
    else scell[s] = tid + s * NT < C::NREG ? tid + s * NT : -1;
  }
  // cell a slot loads: lanes with nothing to store load a region cell
  auto load_cell = [&](int s) { return scell[s] >= 0 ? scell[s] : (SELF ? own : C::NREG - 1); };

  // per-lane element offsets of the loaded plane cells; plane bases are
  // wave-uniform
  const gdouble* sbase = gptr(S.p) + S.off;
  int loff[C::NLOAD];
#pragma unroll
  for (int s = 0; s < C::NLOAD; ++s) {
    const int cell = load_cell(s);
    int ii = cell % RX, jj = cell / RX;
    int x = T.i0 - HX + ii, y = T.j0 - HX + jj;
    if (wa.on) {
      x = x < wa.lo[0] ? x + wa.len[0] : (x >= wa.lo[0] + wa.len[0] ? x - wa.len[0] : x);
      y = y < wa.lo[1] ? y + wa.len[1] : (y >= wa.lo[1] + wa.len[1] ? y - wa.len[1] : y);
    } else {
      x = min(x, xcap);
      y = min(y, ycap);
    }
    loff[s] = x + y * int(S.js);
  }
  // NB: per slot, the element offset from sbase of its column in the array
  // that holds it (plane 0, the periodic shift of the current block folded
  // in) and, for NB 1, that array's plane and component strides; refreshed
  // when the plane enters another block of the owner map (uniform over the
  // tile)
  constexpr int NBS = NB == 1 ? C::NLOAD : 1;
  int nks[NBS], nns[NBS];
  int nzb = INT_MIN;
  auto nb_refresh = [&](int p) {
    const int zb = (p - nbm.dlo[2]) >> nbm.sh;
    if (zb == nzb) return;
    nzb = zb;
    int pw = p, zbw = zb;
    if (nbm.per[2]) {
      pw = p < nbm.dlo[2] ? p + nbm.len[2] : (p >= nbm.dlo[2] + nbm.len[2] ? p - nbm.len[2] : p);
      zbw = (pw - nbm.dlo[2]) >> nbm.sh;
    }
#pragma unroll
    for (int s = 0; s < C::NLOAD; ++s) {
      const int cell = load_cell(s);
      const int x = min(T.i0 - HX + cell % RX, xcap), y = min(T.j0 - HX + cell / RX, ycap);
      int xw = x, yw = y;
      if (nbm.per[0]) xw = x < nbm.dlo[0] ? x + nbm.len[0] : (x >= nbm.dlo[0] + nbm.len[0] ? x - nbm.len[0] : x);
      if (nbm.per[1]) yw = y < nbm.dlo[1] ? y + nbm.len[1] : (y >= nbm.dlo[1] + nbm.len[1] ? y - nbm.len[1] : y);
      const int ob = nbm.owner[(int64_t(zbw) * nbm.nb[1] + ((yw - nbm.dlo[1]) >> nbm.sh)) * nbm.nb[0] +
                               ((xw - nbm.dlo[0]) >> nbm.sh)];
      if (ob >= 0) {
        const Fab F = sinf[ob];
        loff[s] = int((F.p - S.p) + (F.off - S.off) + xw + int64_t(yw) * F.js + int64_t(pw - p) * F.ks);
        if constexpr (NB == 1) {
          nks[s] = int(F.ks);
          nns[s] = int(F.ns);
        }
      } else {
        loff[s] = x + y * int(S.js);
        if constexpr (NB == 1) {
          nks[s] = int(S.ks);
          nns[s] = int(S.ns);
        }
      }
    }
  };
  // ghost cells no box covers come from U^n (UCS): bit b of the lane's mask
  // for load cell s (in LDS, which has room where registers have not) is
  // set when the cell is such a ghost cell in plane block zb0 + b; at most
  // UNCOV_MAX_ZBLK blocks per tile
  const int zb0 = (kfirst - us.lo[2]) >> us.sh;
  const gdouble* u0base = sbase;
  unsigned um_own = 0u;  // the same mask for the own column (order 6)
  if (UCS) {
    const Fab U0 = us.u0[T.box];
    u0base = gptr(U0.p) + U0.off;
    const int zb1 = (k0 + nz + (H == 3 ? 2 : 1) - us.lo[2]) >> us.sh;
    guint* ubits = (guint*)us.bits;
#pragma unroll
    for (int s = 0; s < C::NLOAD; ++s) {
      const int cell = load_cell(s);
      const int bx = (min(T.i0 - HX + cell % RX, xcap) - us.lo[0]) >> us.sh;
      const int by = (min(T.j0 - HX + cell / RX, ycap) - us.lo[1]) >> us.sh;
      const bool inxy = bx >= 0 && bx < us.nb[0] && by >= 0 && by < us.nb[1];
      unsigned m = 0u;
      for (int zb = zb0; zb <= zb1; ++zb) {
        if (inxy && zb >= 0 && zb < us.nb[2]) {
          const int bit = (zb * us.nb[1] + by) * us.nb[0] + bx;
          if ((ubits[bit >> 5] >> (bit & 31)) & 1u) m |= 1u << (zb - zb0);
        }
      }
      s_um[s * NT + tid] = m;
    }
    if (H == 3) {
      const int bx = (min(gi, xhi) - us.lo[0]) >> us.sh, by = (min(gj, yhi) - us.lo[1]) >> us.sh;
      const bool inxy = bx >= 0 && bx < us.nb[0] && by >= 0 && by < us.nb[1];
      for (int zb = zb0; zb <= zb1; ++zb) {
        if (inxy && zb >= 0 && zb < us.nb[2]) {
          const int bit = (zb * us.nb[1] + by) * us.nb[0] + bx;
          if ((ubits[bit >> 5] >> (bit & 31)) & 1u) um_own |= 1u << (zb - zb0);
        }
      }
    }
  }
  // own column for the epilogue loads, clamped into the valid region (s1
  // may have no ghost cells): overhanging lanes load there and never store
  const int soff = min(gi, xhi) + min(gj, yhi) * int(S.js);
  const gdouble* s1base = gptr(S1.p) + S1.off;
  const int s1off = min(gi, xhi) + min(gj, yhi) * int(S1.js);
  double pre[C::NLOAD][5];
  auto load_plane = [&](int p) {
    if constexpr (NB == 1) {
      nb_refresh(p);
#pragma unroll
      for (int s = 0; s < C::NLOAD; ++s) {
        const int o = loff[s] + p * nks[s];
#pragma unroll
        for (int n = 0; n < 5; ++n) pre[s][n] = sbase[int64_t(o) + n * nns[s]];
      }
      return;
    }
    if constexpr (NB == 2) nb_refresh(p);  // then the plain form with the refreshed offsets
    if (!NB && wa.on) p = p < wa.lo[2] ? p + wa.len[2] : (p >= wa.lo[2] + wa.len[2] ? p - wa.len[2] : p);
    if (UCS) {
      const int b = ((p - us.lo[2]) >> us.sh) - zb0;
      const gdouble* pb = sbase + int64_t(p) * S.ks;
      const gdouble* pu = u0base + int64_t(p) * S.ks;
#pragma unroll
      for (int s = 0; s < C::NLOAD; ++s) {
        if ((s_um[s * NT + tid] >> b) & 1u) {
#pragma unroll
          for (int n = 0; n < 5; ++n) pre[s][n] = pu[n * S.ns + loff[s]];
        } else {
#pragma unroll
          for (int n = 0; n < 5; ++n) pre[s][n] = pb[n * S.ns + loff[s]];
        }
      }
    } else {
      const gdouble* pb = sbase + int64_t(p) * S.ks;
#pragma unroll
      for (int s = 0; s < C::NLOAD; ++s)
#pragma unroll
        for (int n = 0; n < 5; ++n) pre[s][n] = pb[n * S.ns + loff[s]];
    }
  };
  // order 6: the own column's conserved state two planes above an
  // iteration's plane c, loaded one iteration ahead into o2n
  double o2n[5], o2[5];
  auto load_own = [&](int p) {
    if (wa.on) p = p < wa.lo[2] ? p + wa.len[2] : (p >= wa.lo[2] + wa.len[2] ? p - wa.len[2] : p);
    const bool from_u0 = UCS && ((um_own >> (((p - us.lo[2]) >> us.sh) - zb0)) & 1u);
    const gdouble* pb = (from_u0 ? u0base : sbase) + int64_t(p) * S.ks;
#pragma unroll
    for (int n = 0; n < 5; ++n) o2n[n] = pb[n * S.ns + soff];
  };
  auto store_plane = [&](int p) {
    double* sq = s_q[p & 1];
#pragma unroll
    for (int s = 0; s < C::NLOAD; ++s) {
      const int cell = SELF ? scell[s] : tid + s * NT;
      if (SELF ? cell >= 0 : cell < C::NREG) {
        double q[6];
        cons2prim(pre[s][0], pre[s][1], pre[s][2], pre[s][3], pre[s][4], c, q);
#pragma unroll
        for (int n = 0; n < NPRIM; ++n) sq[n * PL + cell] = q[n];
      }
    }
  };

  // own column: plane c is snapshotted from LDS before the first barrier of
  // its iteration, plane c+1 is read from its LDS slot (overwritten only two
  // iterations later), plane c-1 is the previous snapshot; likewise u,v,w,T
  // of the ring cell
  double qm1[NPRIM], qs[NPRIM];
  double rm1[4], rs[4];
  double pfz_prev[4] = {0, 0, 0, 0}, pfz_cur[4] = {0, 0, 0, 0};
  double fz_prev[5] = {0, 0, 0, 0, 0};
  double zcarry[5] = {0, 0, 0, 0, 0};  // z pair (c+1, c-1) terms for the next plane
  // order 6: z pairs (c+2, c-1) of the previous plane and of the one before
  double z3p[5] = {0, 0, 0, 0, 0}, z3pp[5] = {0, 0, 0, 0, 0};
  double xy_prev[5] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int n = 0; n < NPRIM; ++n) qm1[n] = 0.0;
#pragma unroll
  for (int n = 0; n < 4; ++n) rm1[n] = 0.0;

  // prologue: plane kfirst into LDS
  load_plane(kfirst);
  store_plane(kfirst);
  load_plane(kfirst + 1);
  if (H == 3) load_own(kfirst + 2);
  __syncthreads();

  for (int cpl = kfirst; cpl <= k0 + nz; ++cpl) {
    const int pn = cpl + 1;
    // epilogue operands of the plane finalised in this iteration and its
    // coverage word, loaded by every lane on every iteration (the plane is
    // clamped to the tile's first one before the tile starts finalising)
    // and before the prefetch below: loads complete in issue order, so
    // waiting for these later never waits for the prefetch, which keeps a
    // whole iteration to arrive
    const bool fin = cpl >= k0 + 1 && own_in;
    const int ep = max(cpl - 1, k0);
    double sv[5], s1v[5];
    {
      const gdouble* pb = sbase + int64_t(ep) * S.ks;
#pragma unroll
      for (int n = 0; n < 5; ++n) sv[n] = pb[n * S.ns + soff];
      const gdouble* p1 = s1base + int64_t(ep) * S1.ks;
#pragma unroll
      for (int n = 0; n < 5; ++n) s1v[n] = p1[n * S1.ns + s1off];
    }
    const unsigned cword = cbits[(max((ep - cov.lo[2]) >> cov.sh, 0) * cplane + ccol) >> 5];
    // (a) next plane into LDS, prefetch the one after (the last iteration
    // has nothing left to prefetch and reloads plane k0+nz+1: every plane a
    // tile reads lies within two planes of its own)
    store_plane(pn);
    load_plane(min(pn + 1, k0 + nz + 1));
    if (H == 3) {
      // plane c+2 of the own column (issued one iteration ago, after that
      // iteration's prefetch), then plane c+3 for the next iteration
#pragma unroll
      for (int n = 0; n < 5; ++n) o2[n] = o2n[n];
      load_own(min(cpl + 3, k0 + nz + 2));
    }
    const double* sq = s_q[cpl & 1];        // plane c
    const double* sqn = s_q[(cpl + 1) & 1];  // plane c+1 (stored above)
#pragma unroll
    for (int n = 0; n < NPRIM; ++n) qs[n] = sq[n * PL + own];
    if (VISC && ring_kind >= 0)
#pragma unroll
      for (int n = 0; n < 4; ++n) rs[n] = sq[(QU + n) * PL + ring_idx];
    // (b) plane c+1 visible.  The self-loading form reads plane c+1 only at
    // cells the thread stored itself; it keeps the barrier while loading the
    // planes below the tile, whose iterations have no other one
    if (!SELF || cpl <= k0 - 1) __syncthreads();
    if (cpl < k0 - 1 || (!VISC && H == 1 && cpl < k0)) {  // uniform: only loading planes below the tile
      if (H == 3 && cpl == k0 - 2) {
        // (k0, k0-3): the first face's (c, c-3) pair
        double q2[NPRIM];
        cons2prim(o2[0], o2[1], o2[2], o2[3], o2[4], c, q2);
        zterms(q2, qm1, cv, z3p);
      }
#pragma unroll
      for (int n = 0; n < NPRIM; ++n) qm1[n] = qs[n];
#pragma unroll
      for (int n = 0; n < 4; ++n) rm1[n] = rs[n];
      continue;
    }
    // z pair (k0, k0-2) for the first face; order 6 also (k0+1, k0-2)
    if (H >= 2 && cpl == k0 - 1) {
      double qn[NPRIM];
#pragma unroll
      for (int n = 0; n < NPRIM; ++n) qn[n] = sqn[n * PL + own];
      zterms(qn, qm1, cv, zcarry);
      if (H == 3) {
#pragma unroll
        for (int n = 0; n < 5; ++n) z3pp[n] = z3p[n];
        double q2[NPRIM];
        cons2prim(o2[0], o2[1], o2[2], o2[3], o2[4], c, q2);
        zterms(q2, qm1, cv, z3p);
      }
      if (!VISC) {
#pragma unroll
        for (int n = 0; n < NPRIM; ++n) qm1[n] = qs[n];
        continue;
      }
    }

    const bool inner = cpl >= k0 && cpl < k0 + nz;  // x/y work on this plane
    if (VISC) {
      // own cell
      {
        VDer d;
        d.ux = qs[QU];
        d.uy = qs[QV];
        d.uz = qs[QW];
        d.T = qs[QT];
        const int e = own + 1, w = own - 1, nn = own + RX, s = own - RX;
        d.dudx = (sq[QU * PL + e] - sq[QU * PL + w]) * hx;
        d.dvdx = (sq[QV * PL + e] - sq[QV * PL + w]) * hx;
        d.dwdx = (sq[QW * PL + e] - sq[QW * PL + w]) * hx;
        d.dTdx = (sq[QT * PL + e] - sq[QT * PL + w]) * hx;
        d.dudy = (sq[QU * PL + nn] - sq[QU * PL + s]) * hy;
        d.dvdy = (sq[QV * PL + nn] - sq[QV * PL + s]) * hy;
        d.dwdy = (sq[QW * PL + nn] - sq[QW * PL + s]) * hy;
        d.dTdy = (sq[QT * PL + nn] - sq[QT * PL + s]) * hy;
        d.dudz = (sqn[QU * PL + own] - qm1[QU]) * hz;
        d.dvdz = (sqn[QV * PL + own] - qm1[QV]) * hz;
        d.dwdz = (sqn[QW * PL + own] - qm1[QW]) * hz;
        d.dTdz = (sqn[QT * PL + own] - qm1[QT]) * hz;
        double fx[4], fy[4];
        if (inner) {
          visc_flux<true, true, true>(d, c, fx, fy, pfz_cur);
#pragma unroll
          for (int n = 0; n < 4; ++n) {
            s_px[n][tj][ti + 1] = fx[n];
            s_py[n][tj + 1][ti] = fy[n];
          }
        } else {
          visc_flux<false, false, true>(d, c, fx, fy, pfz_cur);
        }
      }
      // ring cell
      if (inner && ring_kind >= 0) {
        VDer d;
        const int r = ring_idx;
        d.ux = rs[0];
        d.uy = rs[1];
        d.uz = rs[2];
        d.T = rs[3];
        const int e = r + 1, w = r - 1, nn = r + RX, s = r - RX;
        d.dudx = (sq[QU * PL + e] - sq[QU * PL + w]) * hx;
        d.dvdx = (sq[QV * PL + e] - sq[QV * PL + w]) * hx;
        d.dwdx = (sq[QW * PL + e] - sq[QW * PL + w]) * hx;
        d.dTdx = (sq[QT * PL + e] - sq[QT * PL + w]) * hx;
        d.dudy = (sq[QU * PL + nn] - sq[QU * PL + s]) * hy;
        d.dvdy = (sq[QV * PL + nn] - sq[QV * PL + s]) * hy;
        d.dwdy = (sq[QW * PL + nn] - sq[QW * PL + s]) * hy;
        d.dTdy = (sq[QT * PL + nn] - sq[QT * PL + s]) * hy;
        d.dudz = (sqn[QU * PL + r] - rm1[0]) * hz;
        d.dvdz = (sqn[QV * PL + r] - rm1[1]) * hz;
        d.dwdz = (sqn[QW * PL + r] - rm1[2]) * hz;
        d.dTdz = (sqn[QT * PL + r] - rm1[3]) * hz;
        // both directions in one pass (each component is computed as the
        // one-direction flux computes it), stored per lane
        double fx[4], fy[4], fz[4];
        visc_flux<true, true, false>(d, c, fx, fy, fz);
        const bool rx = ring_kind == 0;
        double* dst = rx ? &s_px[0][ring_b][ring_a + 1] : &s_py[0][ring_b + 1][ring_a];
        const int cs = rx ? TY * (TX + 2) : (TY + 2) * TX;
#pragma unroll
        for (int n = 0; n < 4; ++n) dst[n * cs] = rx ? fx[n] : fy[n];
      }
      // KEEP part of the faces at i = TX (threads 0 .. TY-1) and j = TY
      // (threads TY .. TY+TX-1) in one pass; their viscous part is added
      // after the barrier by other threads
      if (C::EARLY && inner && tid < TX + TY) {
        const bool isx = tid < TY;
        const int e = isx ? tid : tid - TY;
        double f[5];
        keep_face_xy<H, PL>(sq, isx ? (e + HX) * RX + (TX + HX) : (TY + HX) * RX + (e + HX), isx ? 1 : RX, isx, co,
                            cv, f);
        double* dst = isx ? &s_fx[0][e][TX] : &s_fy[0][TY][e];
        const int cs = isx ? TY * (TX + 1) : (TY + 1) * TX;
#pragma unroll
        for (int n = 0; n < 5; ++n) {
          if (SELF) s_fe[n][tid] = f[n];
          else dst[n * cs] = f[n];
        }
      }
      // (c) viscous x/y fluxes visible; below the tile this barrier instead
      // keeps the next store from overwriting the plane just read
      __syncthreads();
    }

    if (inner) {
      // x and y faces below the own cell
      double f[5];
      keep_face<H, 0, PL>(sq, own, 1, co, cv, f);
      if (VISC) {
        f[0] += 0.5 * (0.0 + 0.0);
#pragma unroll
        for (int n = 0; n < 4; ++n) f[n + 1] += 0.5 * (s_px[n][tj][ti] + s_px[n][tj][ti + 1]);
      }
#pragma unroll
      for (int n = 0; n < 5; ++n) s_fx[n][tj][ti] = f[n];
      keep_face<H, 1, PL>(sq, own, RX, co, cv, f);
      if (VISC) {
        f[0] += 0.5 * (0.0 + 0.0);
#pragma unroll
        for (int n = 0; n < 4; ++n) f[n + 1] += 0.5 * (s_py[n][tj][ti] + s_py[n][tj + 1][ti]);
      }
#pragma unroll
      for (int n = 0; n < 5; ++n) s_fy[n][tj][ti] = f[n];
      // the faces at i = TX and j = TY
      if (C::EARLY) {
        if (tid >= C::VA && tid < C::VA + TX + TY) {
          const int q = tid - C::VA;
          const bool isx = q < TY;
          const int e = isx ? q : q - TY;
          double* dst = isx ? &s_fx[0][e][TX] : &s_fy[0][TY][e];
          const int cs = isx ? TY * (TX + 1) : (TY + 1) * TX;
          // viscous cell fluxes on either side of the face
          const double* pv = isx ? &s_px[0][e][TX] : &s_py[0][TY][e];
          const int pcs = isx ? TY * (TX + 2) : (TY + 2) * TX, pst = isx ? 1 : TX;
#pragma unroll
          for (int n = 0; n < 5; ++n) f[n] = SELF ? s_fe[n][q] : dst[n * cs];
          f[0] += 0.5 * (0.0 + 0.0);
#pragma unroll
          for (int n = 0; n < 4; ++n) f[n + 1] += 0.5 * (pv[n * pcs] + pv[n * pcs + pst]);
#pragma unroll
          for (int n = 0; n < 5; ++n) dst[n * cs] = f[n];
        }
      } else if (tid >= C::XB && tid < C::XB + TY) {
        const int jj = tid - C::XB;
        keep_face<H, 0, PL>(sq, (jj + HX) * RX + (TX + HX), 1, co, cv, f);
        if (VISC) {
          f[0] += 0.5 * (0.0 + 0.0);
#pragma unroll
          for (int n = 0; n < 4; ++n) f[n + 1] += 0.5 * (s_px[n][jj][TX] + s_px[n][jj][TX + 1]);
        }
#pragma unroll
        for (int n = 0; n < 5; ++n) s_fx[n][jj][TX] = f[n];
      } else if (tid >= C::YB && tid < C::YB + TX) {
        const int ii = tid - C::YB;
        keep_face<H, 1, PL>(sq, (TY + HX) * RX + (ii + HX), RX, co, cv, f);
        if (VISC) {
          f[0] += 0.5 * (0.0 + 0.0);
#pragma unroll
          for (int n = 0; n < 4; ++n) f[n + 1] += 0.5 * (s_py[n][TY][ii] + s_py[n][TY + 1][ii]);
        }
#pragma unroll
        for (int n = 0; n < 5; ++n) s_fy[n][TY][ii] = f[n];
      }
      __syncthreads();  // (d) faces visible
    }

    double xy[5];
    if (inner) {
#pragma unroll
      for (int n = 0; n < 5; ++n)
        xy[n] = dxi0 * (s_fx[n][tj][ti] - s_fx[n][tj][ti + 1]) + dxi1 * (s_fy[n][tj][ti] - s_fy[n][tj + 1][ti]);
    }
    if (cpl >= k0) {
      // z face at cpl - 1/2 from the own column (planes c-1, c, c+1)
      double fz[5];
#pragma unroll
      for (int n = 0; n < 5; ++n) fz[n] = 0.0;
      double t[5];
      zterms(qs, qm1, cv, t);  // (c, c-1)
      keep_accum(co[0], t, fz);
      if (H >= 2) {
        // (c, c-2) is the previous plane's (c+1, c-1), carried in registers
        keep_accum(co[1], zcarry, fz);
        double qn[NPRIM];
#pragma unroll
        for (int n = 0; n < NPRIM; ++n) qn[n] = sqn[n * PL + own];
        zterms(qn, qm1, cv, zcarry);  // (c+1, c-1)
        keep_accum(co[1], zcarry, fz);
      }
      if (H == 3) {
        // (c, c-3) and (c+1, c-2) are the (c+2, c-1) pairs of the planes
        // two and one below
        keep_accum(co[2], z3pp, fz);
        keep_accum(co[2], z3p, fz);
#pragma unroll
        for (int n = 0; n < 5; ++n) z3pp[n] = z3p[n];
        double q2[NPRIM];
        cons2prim(o2[0], o2[1], o2[2], o2[3], o2[4], c, q2);
        zterms(q2, qm1, cv, z3p);  // (c+2, c-1)
        keep_accum(co[2], z3p, fz);
      }
      if (VISC) {
        fz[0] += 0.5 * (0.0 + 0.0);
#pragma unroll
        for (int n = 0; n < 4; ++n) fz[n + 1] += 0.5 * (pfz_prev[n] + pfz_cur[n]);
      }
      if (fin && ((cword >> ((((cpl - 1 - cov.lo[2]) >> cov.sh) * cplane + ccol) & 31)) & 1u)) {
        gdouble* po = gptr(O.p) + O.idx(gi, gj, cpl - 1);
        double L[5];
#pragma unroll
        for (int n = 0; n < 5; ++n) {
          L[n] = xy_prev[n] + dxi2 * (fz_prev[n] - fz[n]);
          if (has_src) L[n] = L[n] + SR(gi, gj, cpl - 1, n);  // user_source
        }
        if (SRC && sa.bfx != 0.0) {  // body-force user_source (exm/Channel/prob.h:189-193)
          L[UMX] = L[UMX] + sv[URHO] * sa.bfx;
          L[UET] = L[UET] + sv[UMX] * sa.bfx;
        }
        // s1 is read only where the stage uses it (it is not loaded otherwise)
        if (use_s1) {
#pragma unroll
          for (int n = 0; n < 5; ++n) po[n * O.ns] = stage_combine(sc, s1v[n], sv[n], L[n]);
        } else {
#pragma unroll
          for (int n = 0; n < 5; ++n) po[n * O.ns] = stage_combine(sc, 0.0, sv[n], L[n]);
        }
      }
#pragma unroll
      for (int n = 0; n < 5; ++n) fz_prev[n] = fz[n];
    }
    if (inner) {
#pragma unroll
      for (int n = 0; n < 5; ++n) xy_prev[n] = xy[n];
    }
    if (VISC) {
#pragma unroll
      for (int n = 0; n < 4; ++n) pfz_prev[n] = pfz_cur[n];
    }
#pragma unroll
    for (int n = 0; n < NPRIM; ++n) qm1[n] = qs[n];
#pragma unroll
    for (int n = 0; n < 4; ++n) rm1[n] = rs[n];
  }
}

#define AMR_FUSED_LAUNCH(UCS_, SELF_, SRC_, NB_)                                                              \
  hipLaunchKernelGGL((k_fused<H, VISC, TX, TY, UCS_, SELF_, SRC_, NB_>), dim3(ntiles), dim3(TX * TY), 0, st, tiles, \
                     boxes, c, co[0], co[1], co[2], sa, wa, cov, us, nbm, s1, sin, out)

template <int H, bool VISC, int TX, int TY, bool SELF>
void launch_t(const TileDev* tiles, int ntiles, const BoxDev* boxes, const Closures& c, const double co[3],
              const StageArgs& sa, const WrapArgs& wa, const Cover& cov, const UncovSrc& us, const NbMap& nbm,
              const Fab* s1, const Fab* sin, const Fab* out, hipStream_t st) {
  const bool src = sa.src != nullptr || sa.bfx != 0.0;
  if constexpr (H <= 2) {
    if (nbm.on) {
      if (nbm.ustride) {
        if (src) AMR_FUSED_LAUNCH(false, SELF, true, 2);
        else AMR_FUSED_LAUNCH(false, SELF, false, 2);
      } else {
        if (src) AMR_FUSED_LAUNCH(false, SELF, true, 1);
        else AMR_FUSED_LAUNCH(false, SELF, false, 1);
      }
      return;
    }
  }
  if (us.on) {
    if (src) AMR_FUSED_LAUNCH(true, SELF, true, 0);
    else AMR_FUSED_LAUNCH(true, SELF, false, 0);
  } else {
    if (src) AMR_FUSED_LAUNCH(false, SELF, true, 0);
    else AMR_FUSED_LAUNCH(false, SELF, false, 0);
  }
}
#undef AMR_FUSED_LAUNCH

template <int TX, int TY>
bool launch_shape(const amr_scheme_t& s, const TileDev* tiles, int ntiles, const BoxDev* boxes, const Closures& c,
                  const double co[3], const StageArgs& sa, const WrapArgs& wa, const Cover& cov,
                  const UncovSrc& us, const NbMap& nbm, const Fab* s1, const Fab* sin, const Fab* out,
                  hipStream_t st) {
  const bool v = s.rhs_visc != 0;
  if (s.order_keep == 6) {
    if (v) launch_t<3, true, TX, TY, false>(tiles, ntiles, boxes, c, co, sa, wa, cov, us, nbm, s1, sin, out, st);
    else launch_t<3, false, TX, TY, false>(tiles, ntiles, boxes, c, co, sa, wa, cov, us, nbm, s1, sin, out, st);
  } else if (s.order_keep == 4) {
    if (v) launch_t<2, true, TX, TY, false>(tiles, ntiles, boxes, c, co, sa, wa, cov, us, nbm, s1, sin, out, st);
    else launch_t<2, false, TX, TY, false>(tiles, ntiles, boxes, c, co, sa, wa, cov, us, nbm, s1, sin, out, st);
  } else if (s.order_keep == 2) {
    if (v) launch_t<1, true, TX, TY, false>(tiles, ntiles, boxes, c, co, sa, wa, cov, us, nbm, s1, sin, out, st);
    else launch_t<1, false, TX, TY, false>(tiles, ntiles, boxes, c, co, sa, wa, cov, us, nbm, s1, sin, out, st);
  } else {
    return false;
  }
  return true;
}

}  // namespace

bool fused_supported(const amr_scheme_t& s, int tx, int ty) {
  // any box shape: tiles overhanging a box are masked
  if (!s.rhs_euler || s.flux_euler != 1 || s.art_diss) return false;
  if (!(s.order_keep == 2 || s.order_keep == 4 || s.order_keep == 6)) return false;
  return (tx == 16 && ty == 16) || (tx == 8 && ty == 8) || (tx == 32 && ty == 8);
}

bool fused_self_supported(const amr_scheme_t& s, int tx, int ty) {
  return fused_supported(s, tx, ty) && s.order_keep <= 4 && s.rhs_visc && tx == 16 && ty == 16;
}

hipError_t launch_fused_stage(const amr_scheme_t& s, const FusedConfig& cfg, const WrapArgs& wa, const Cover& cov,
                              const UncovSrc& us, const TileDev* tiles, int ntiles, const BoxDev* boxes,
                              const Closures& c, const StageArgs& sa, const Fab* s1, const Fab* sin, const Fab* out,
                              hipStream_t st, const NbMap& nbm) {
  if (nbm.on && (us.on || s.order_keep > 4)) return hipErrorInvalidValue;
  double co[3];
  keep_coefficients(s.order_keep, co);
  bool ok = false;
  if (cfg.form == 1) {
    if (!fused_self_supported(s, cfg.tx, cfg.ty)) return hipErrorInvalidValue;
    if (s.order_keep == 4) launch_t<2, true, 16, 16, true>(tiles, ntiles, boxes, c, co, sa, wa, cov, us, nbm, s1, sin, out, st);
    else launch_t<1, true, 16, 16, true>(tiles, ntiles, boxes, c, co, sa, wa, cov, us, nbm, s1, sin, out, st);
    ok = true;
  } else if (cfg.tx == 16 && cfg.ty == 16)
    ok = launch_shape<16, 16>(s, tiles, ntiles, boxes, c, co, sa, wa, cov, us, nbm, s1, sin, out, st);
  else if (cfg.tx == 8 && cfg.ty == 8)
    ok = launch_shape<8, 8>(s, tiles, ntiles, boxes, c, co, sa, wa, cov, us, nbm, s1, sin, out, st);
  else if (cfg.tx == 32 && cfg.ty == 8)
    ok = launch_shape<32, 8>(s, tiles, ntiles, boxes, c, co, sa, wa, cov, us, nbm, s1, sin, out, st);
  if (!ok) return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace amr
