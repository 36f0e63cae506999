// Euler fluxes, one pass per direction: WENO5-JS with global Lax-Friedrichs
// splitting (flux_euler = 2) and MUSCL slopes with the two-shock Riemann
// solver (flux_euler = 0); with viscous terms each pass adds the face average
// of the cell-centred viscous fluxes of its direction (computed beforehand by
// a pass over the box grown by one cell) to its faces before differencing.
//
// The reference runs each scheme as full-box passes with every intermediate
// in HBM: cons2eulerflux_lambda over the grown box and one
// numericalflux_globallaxsplit pass per direction (High_resolution.h:83-319),
// or primitives over the grown box and, per direction, cns_slope and
// cns_riemann passes (Riemann.h:19-408), then the flux assembly
// (CNS_advance.cpp:375-391).  Here a workgroup owns a brick of 64 cells along
// x by 4 along a second axis and stages the pencil of cells its faces need
// (ghost cells along the pass direction only) in LDS:
// - WENO: conserved state, normal velocity, pressure and the direction's wave
//   speed per cell; each thread reconstructs one or two faces (the 6-cell
//   maximum of the wave speed, the split fluxes of the 6 cells, two WENO5-JS
//   evaluations per component) from LDS;
// - MUSCL: primitives per cell, then the characteristic slopes of the cells
//   next to a face, then one Riemann problem per face, all from LDS.
// With viscous terms the face flux is the Euler flux plus the average of the
// two cells' viscous fluxes (ViscousNumericalFluxes, CNS_hydro_K.h:254-261,
// read from HBM).  After a barrier each thread forms its cell's flux
// difference.  Pass x
// writes dxinv_x (F(i) - F(i+1)) to an accumulator, pass y adds its term,
// pass z adds its term, the source and the stage combination and writes the
// stage output, so the sum keeps the reference's order (x + y) + z.
#include "cns_point.h"
#include "kernels.h"

namespace amr {

namespace {

// One thread per cell of a WX x WB brick.  The faces beyond WX x WB (65 x 4
// in the x pass, 64 x 5 in the others) take a second round on the first
// threads, except where the kernels avoid it (repeated rows in the y and z
// passes; one (face, component) pair per thread in the WENO x pass);
// five-wave workgroups with one face per thread ran the 1D batch 18% slower.
constexpr int WX = 64, WB = 4, WNT = WX * WB;

// Brick geometry of pass D: loaded cells LX x LB, faces FX x FB (x fastest);
// the second axis is y for D = 0, 1 and z for D = 2.
template <int D>
struct WGeo {
  static constexpr int LX = D == 0 ? WX + 6 : WX;
  static constexpr int LB = D == 0 ? WB : WB + 6;
  static constexpr int NL = LX * LB;
  static constexpr int FX = D == 0 ? WX + 1 : WX;
  static constexpr int FB = D == 0 ? WB : WB + 1;
  static constexpr int NF = FX * FB;
  static constexpr int NV = 8;  // per cell: 5 conserved, normal velocity, pressure, wave speed
};

// (x, b, c) brick coordinates -> (i, j, k)
template <int D>
__device__ __forceinline__ void brick_ijk(int x, int b, int c, int& i, int& j, int& k) {
  i = x;
  if (D == 2) {
    j = c;
    k = b;
  } else {
    j = b;
    k = c;
  }
}

// The epilogue operands of the thread's cell, loaded when the kernel starts
// so that their latency hides behind the staging and the faces: the sum of
// the earlier passes (passes y and z) and, in the last pass, s1 and (PU) the
// stage input where the stage combination reads them (PU costs the WENO
// pass, at 122 VGPRs without it, a wave per SIMD).
template <int D, int LAST, bool PU>
struct EpiOps {
  bool on;
  int i, j, k;
  double a[5], s1[5], u[5];
};

template <int D, int LAST, bool PU>
__device__ __forceinline__ void epi_load(const WenoBrick& W, const BoxDev& bx, const StageArgs& sa,
                                         const Fab* __restrict__ sinf, const Fab* __restrict__ accf,
                                         const Fab* __restrict__ s1f, EpiOps<D, LAST, PU>& e) {
  constexpr int BA = D == 2 ? 2 : 1;
  const int tid = threadIdx.x;
  const int cx = tid % WX, cb = tid / WX;
  const int x = W.x0 + cx, b = W.b0 + cb;
  e.on = tid < WX * WB && x <= bx.hi[0] && b <= bx.hi[BA];
  if (!e.on) return;
  brick_ijk<D>(x, b, W.c, e.i, e.j, e.k);
  if (D != 0) {
    const Fab A = accf[W.box];
#pragma unroll
    for (int n = 0; n < 5; ++n) e.a[n] = A(e.i, e.j, e.k, n);
  }
  if (LAST) {
    const StageCoef sc = box_stage(sa, W.box);
    if (sc.mode != 2 && sc.a != 0.0) {
      const Fab F = s1f[W.box];
#pragma unroll
      for (int n = 0; n < 5; ++n) e.s1[n] = F(e.i, e.j, e.k, n);
    }
    if (PU && sc.mode != 2) {
      const Fab U = sinf[W.box];
#pragma unroll
      for (int n = 0; n < 5; ++n) e.u[n] = U(e.i, e.j, e.k, n);
    }
  }
}

// Flux difference of the thread's cell from the faces in sF ([5][NF], FX
// faces per row) and, in the last pass, source and stage combination.
template <int D, int LAST, int FX, int NF, bool PU>
__device__ __forceinline__ void dir_epilogue(const WenoBrick& W, const BoxDev& bx, const StageArgs& sa,
                                             const double* sF, const Fab& S, const Fab* __restrict__ accf,
                                             const EpiOps<D, LAST, PU>& e, const Fab* __restrict__ outf) {
  const int tid = threadIdx.x;
  // flux difference of the own cell (CNS_advance.cpp:384-390)
  if (!e.on) return;
  const int i = e.i, j = e.j, k = e.k;
  const int cx = tid % WX, cb = tid / WX;
  const int f0 = cb * FX + cx;
  const int f1 = D == 0 ? f0 + 1 : f0 + FX;
  const double dxi = bx.dxi[D];
  const Fab A = accf[W.box];
  if (!LAST) {
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      const double t = dxi * (sF[n * NF + f0] - sF[n * NF + f1]);
      A(i, j, k, n) = D == 0 ? t : e.a[n] + t;
    }
    return;
  }
  const StageCoef sc = box_stage(sa, W.box);
  const Fab O = outf[W.box];
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    double L = e.a[n] + dxi * (sF[n * NF + f0] - sF[n * NF + f1]);
    if (sa.src) L = L + sa.src[W.box](i, j, k, n);  // user_source, CNS_advance.cpp:397-407
    if (sa.bfx != 0.0 && n == UMX) L = L + S(i, j, k, URHO) * sa.bfx;
    if (sa.bfx != 0.0 && n == UET) L = L + S(i, j, k, UMX) * sa.bfx;
    double v;
    if (sc.mode == 2) {
      v = L;
    } else {
      const double a = (sc.a != 0.0) ? e.s1[n] : 0.0;
      v = stage_combine(sc, a, PU ? e.u[n] : S(i, j, k, n), L);
    }
    O(i, j, k, n) = v;
  }
}

// Faces 0 .. nf-1 of the thread's rounds into sF ([5][NF]); with repeated
// rows (rep) face row 0 also goes to row WB.
template <int FX, int NF, int NFT>
__device__ __forceinline__ void store_faces(const double (&F)[NFT][5], double* sF, int nf, bool rep) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int r = 0; r < NFT; ++r) {
    const int f = tid + r * WNT;
    if (f < nf)
#pragma unroll
      for (int n = 0; n < 5; ++n) sF[n * NF + f] = F[r][n];
  }
  if (rep && tid < FX)
#pragma unroll
    for (int n = 0; n < 5; ++n) sF[n * NF + WB * FX + tid] = F[0][n];
}

// The x ends of a brick's box as the x pass reads them with side tables:
// both entries and, for a neighbour side, that box's array, loaded once per
// brick (uniform scalar loads) instead of per cell.
struct XEnds {
  XSide s[2];
  Fab f[2];
};

template <int D>
__device__ __forceinline__ bool load_xends(const WenoBrick& W, const Fab* sinf, const XSide* xs, XEnds& e) {
  if (D != 0 || xs == nullptr) return false;
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    e.s[side] = xs[2 * W.box + side];
    if (e.s[side].kind == 1) e.f[side] = sinf[e.s[side].src];
  }
  return true;
}

// The conserved state of pencil cell (x, b) of pass D (x wrapped or clamped
// by the caller except, with side tables (xon), beyond the box's x ends in
// the x pass: there the neighbour's cell or the boundary rule of the side).
template <int D>
__device__ __forceinline__ void pencil_cell(const WenoBrick& W, const BoxDev& bx, const Fab& S, bool xon,
                                            const XEnds& xe, int x, int b, double u[5]) {
  int i, j, k;
  brick_ijk<D>(x, b, W.c, i, j, k);
  if (D == 0 && xon && (x < bx.lo[0] || x > bx.hi[0])) {
    const int side = x > bx.hi[0] ? 1 : 0;
    const XSide& X = xe.s[side];
    if (X.kind == 1) {
      const Fab& F = xe.f[side];
#pragma unroll
      for (int n = 0; n < 5; ++n) u[n] = F(i + X.shift, j, k, n);
      return;
    }
    if (X.kind == 2) {
      // k_fill_bc's rule (AMReX foextrap / reflect_even / reflect_odd)
#pragma unroll
      for (int n = 0; n < 5; ++n) {
        const int ty = X.type[n];
        const int sx = ty == 2 ? X.bnd : (side == 0 ? 2 * X.bnd - 1 - i : 2 * X.bnd + 1 - i);
        const double v = S(sx, j, k, n);
        u[n] = ty == 4 ? -v : v;
      }
      return;
    }
  }
#pragma unroll
  for (int n = 0; n < 5; ++n) u[n] = S(i, j, k, n);
}

// Viscous part of face f of a brick (f % FX along x, f / FX along the second
// axis; the face lies below its cell in direction D): F += 0.5 (Fv(left) +
// Fv(right)) per component (CNS_hydro_K.h:254-261, the reference's operand
// order).  Cells beyond the box (faces of overhanging lanes, whose fluxes
// are never used) are clamped into the viscous fluxes' grown-by-one extent.
template <int D, int FX>
__device__ __forceinline__ void add_visc_face(const WenoBrick& W, const BoxDev& bx, const Fab& P, int f,
                                              double F[5]) {
  constexpr int BA = D == 2 ? 2 : 1;
  const int fx = f % FX, fb = f / FX;
  int xr = min(W.x0 + fx, bx.hi[0] + 1), br = min(W.b0 + fb, bx.hi[BA] + 1);
  const int xl = D == 0 ? xr - 1 : xr, bl = D == 0 ? br : br - 1;
  int il, jl, kl, ir, jr, kr;
  brick_ijk<D>(xl, bl, W.c, il, jl, kl);
  brick_ijk<D>(xr, br, W.c, ir, jr, kr);
#pragma unroll
  for (int n = 0; n < 5; ++n) F[n] += 0.5 * (P(il, jl, kl, n) + P(ir, jr, kr, n));
}

template <int D, int LAST>
__global__ void __launch_bounds__(WNT) k_weno_dir(const WenoBrick* __restrict__ bricks, const BoxDev* __restrict__ boxes,
                                                  Closures c, StageArgs sa, const Fab* __restrict__ sinf,
                                                  const Fab* __restrict__ accf, const Fab* __restrict__ s1f,
                                                  const Fab* __restrict__ outf, const Fab* __restrict__ pvf,
                                                  const XSide* __restrict__ xs) {
  using G = WGeo<D>;
  __shared__ double sv[G::NV][G::NL];
  const WenoBrick W = bricks[blockIdx.x];
  if (!box_active(sa, W.box)) return;  // uniform: the whole workgroup leaves
  const BoxDev bx = boxes[W.box];
  const Fab S = sinf[W.box];
  const int tid = threadIdx.x;
  EpiOps<D, LAST, false> epi;
  epi_load<D, LAST, false>(W, bx, sa, sinf, accf, s1f, epi);
  constexpr int BA = D == 2 ? 2 : 1;  // axis of the brick's second coordinate
  const int xmax = bx.hi[0] + (D == 0 ? NGHOST : 0);
  const int bmax = bx.hi[BA] + (D == 0 ? 0 : NGHOST);

  // 1. stage the cells (High_resolution.h:83-121, the pass direction's part)
  const int dlo = bx.lo[D];
  XEnds xe;
  const bool xon = load_xends<D>(W, sinf, xs, xe);
  auto stage_cell = [&](const double uc[5], double v[G::NV]) {
    const double rho = uc[URHO], momx = uc[UMX], momy = uc[UMY];
    const double momz = uc[UMZ], rhoet = uc[UET];
    const double rhoinv = 1.0 / rho;
    const double ux = momx * rhoinv, uy = momy * rhoinv, uz = momz * rhoinv;
    const double rhoekin = 0.5 * rhoinv * (momx * momx + momy * momy + momz * momz);
    const double rhoeint = rhoet - rhoekin;
    const double P = c.gm1 * rhoeint;
    const double cs = sqrt(c.gamma * P * rhoinv);
    const double ud = D == 0 ? ux : (D == 1 ? uy : uz);
    v[0] = rho;
    v[1] = momx;
    v[2] = momy;
    v[3] = momz;
    v[4] = rhoet;
    v[5] = ud;
    v[6] = P;
    v[7] = fabs(dmax(dmax(ud + cs, ud - cs), ud));
  };
  // A box WB cells deep and periodic along D (the 1D problems' x4x4 boxes):
  // the pencil's rows repeat the box's WB rows and face row WB is face row
  // 0 (the same six cells in the same order), so each thread stages one cell
  // into every row holding it and the faces of rows 0 .. WB-1 are computed
  // once.  Uniform over the workgroup.
  const bool rep = D != 0 && W.wrap == WB && pvf == nullptr;
  if (rep) {
    const int lx = tid % G::LX, r = tid / G::LX;
    double uc[5], v[G::NV];
    pencil_cell<D>(W, bx, S, xon, xe, min(W.x0 + lx, xmax), dlo + r, uc);
    stage_cell(uc, v);
    for (int lb = (r + 3) % WB; lb < G::LB; lb += WB)
#pragma unroll
      for (int n = 0; n < G::NV; ++n) sv[n][lb * G::LX + lx] = v[n];
  } else {
    // every round's loads first, then the conversions
    constexpr int NR = (G::NL + WNT - 1) / WNT;
    double uc[NR][5];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int l = tid + r * WNT;
      if (l < G::NL) {
        const int lx = l % G::LX, lb = l / G::LX;
        int x = W.x0 + lx - (D == 0 ? 3 : 0);
        int b = W.b0 + lb - (D == 0 ? 0 : 3);
        if (D == 0 && W.wrap > 0) x = dlo + ((x - dlo) % W.wrap + W.wrap) % W.wrap;
        else x = min(x, xmax);
        if (D != 0 && W.wrap > 0) b = dlo + ((b - dlo) % W.wrap + W.wrap) % W.wrap;
        else b = min(b, bmax);
        pencil_cell<D>(W, bx, S, xon, xe, x, b, uc[r]);
      }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int l = tid + r * WNT;
      if (l < G::NL) {
        double v[G::NV];
        stage_cell(uc[r], v);
#pragma unroll
        for (int n = 0; n < G::NV; ++n) sv[n][l] = v[n];
      }
    }
  }
  __syncthreads();

  // 2. faces (numericalflux_globallaxsplit, High_resolution.h:162-246): face
  // f sits below stencil cell 3 of cells l0 .. l0+5 (stride st along D)
  constexpr int st = D == 0 ? 1 : G::LX;
  // wave-speed maximum over the stencil of the face at l0, and component n
  // of its flux
  auto face_me = [&](int l0) {
    double me = sv[7][l0];
#pragma unroll
    for (int o = 1; o < 6; ++o) me = dmax(me, sv[7][l0 + o * st]);
    return me;
  };
  auto face_comp = [&](int l0, double me, int n) {
    double df[6], Pf[6];
#pragma unroll
    for (int o = 0; o < 6; ++o) {
      const int l = l0 + o * st;
      const double ud = sv[5][l];
      double flux;
      if (n == URHO) flux = sv[1 + D][l];
      else if (n == UET) flux = (sv[4][l] + sv[6][l]) * ud;
      else if (n == 1 + D) flux = sv[n][l] * ud + sv[6][l];
      else flux = sv[n][l] * ud;
      df[o] = me * sv[n][l];
      Pf[o] = flux;
    }
    double flx = weno5js(0.5 * (Pf[0] + df[0]), 0.5 * (Pf[1] + df[1]), 0.5 * (Pf[2] + df[2]),
                         0.5 * (Pf[3] + df[3]), 0.5 * (Pf[4] + df[4]));
    flx += weno5js(0.5 * (Pf[5] - df[5]), 0.5 * (Pf[4] - df[4]), 0.5 * (Pf[3] - df[3]),
                   0.5 * (Pf[2] - df[2]), 0.5 * (Pf[1] - df[1]));
    return flx;
  };
  double* sF = &sv[0][0];  // [5][NF]
  if (D == 0 && pvf == nullptr) {
    // one (face, component) pair per thread and round, u = n * NF + f: the
    // 65th face of each row costs a fifth of a round, not a second round of
    // all five components on one wave
    constexpr int NU = 5 * G::NF;
    constexpr int NUT = (NU + WNT - 1) / WNT;
    double Fu[NUT];
#pragma unroll
    for (int r = 0; r < NUT; ++r) {
      const int u = tid + r * WNT;
      if (u < NU) {
        const int n = u / G::NF, f = u - n * G::NF;
        const int fx = f % G::FX, fb = f / G::FX;
        const int l0 = fb * G::LX + fx;
        Fu[r] = face_comp(l0, face_me(l0), n);
      }
    }
    __syncthreads();  // every stencil read done: the faces reuse the cell slots
#pragma unroll
    for (int r = 0; r < NUT; ++r) {
      const int u = tid + r * WNT;
      if (u < NU) sF[u] = Fu[r];
    }
    __syncthreads();
  } else {
    constexpr int NFT = (G::NF + WNT - 1) / WNT;
    const int nf = rep ? WB * G::FX : G::NF;
    double F[NFT][5];
#pragma unroll
    for (int r = 0; r < NFT; ++r) {
      const int f = tid + r * WNT;
      if (f < nf) {
        const int fx = f % G::FX, fb = f / G::FX;
        const int l0 = fb * G::LX + fx;
        const double me = face_me(l0);
#pragma unroll
        for (int n = 0; n < 5; ++n) F[r][n] = face_comp(l0, me, n);
        if (pvf) add_visc_face<D, G::FX>(W, bx, pvf[W.box], f, F[r]);
      }
    }
    __syncthreads();  // every stencil read done: the faces reuse the cell slots
    store_faces<G::FX, G::NF, NFT>(F, sF, nf, rep);
    __syncthreads();
  }

  dir_epilogue<D, LAST, G::FX, G::NF, false>(W, bx, sa, sF, S, accf, epi, outf);
}

// Brick geometry of the MUSCL pass D: faces f need cells f-2 .. f+1 (the
// slopes of cells f-1 and f; the z pass also reads T at f+1).
template <int D>
struct MGeo {
  static constexpr int LX = D == 0 ? WX + 4 : WX;
  static constexpr int LB = D == 0 ? WB : WB + 4;
  static constexpr int NL = LX * LB;
  static constexpr int FX = D == 0 ? WX + 1 : WX;
  static constexpr int FB = D == 0 ? WB : WB + 1;
  static constexpr int NF = FX * FB;
};

template <int D, int LAST>
__global__ void __launch_bounds__(WNT) k_muscl_dir(const WenoBrick* __restrict__ bricks,
                                                   const BoxDev* __restrict__ boxes, Closures c, StageArgs sa,
                                                   const Fab* __restrict__ sinf, const Fab* __restrict__ accf,
                                                   const Fab* __restrict__ s1f, const Fab* __restrict__ outf,
                                                   const Fab* __restrict__ pvf, const XSide* __restrict__ xs) {
  using G = MGeo<D>;
  __shared__ double sq[NPRIM][G::NL];
  __shared__ double sd[5][G::NL];
  const WenoBrick W = bricks[blockIdx.x];
  if (!box_active(sa, W.box)) return;  // uniform: the whole workgroup leaves
  const BoxDev bx = boxes[W.box];
  const Fab S = sinf[W.box];
  const int tid = threadIdx.x;
  EpiOps<D, LAST, true> epi;
  epi_load<D, LAST, true>(W, bx, sa, sinf, accf, s1f, epi);
  constexpr int BA = D == 2 ? 2 : 1;
  const int xmax = bx.hi[0] + (D == 0 ? NGHOST : 0);
  const int bmax = bx.hi[BA] + (D == 0 ? 0 : NGHOST);
  const int dlo = bx.lo[D];

  // 1. primitives of the pencil (cons2prim, CNS_hydro_K.h:15-34).  A box WB
  // cells deep and periodic along D: the pencil's rows repeat the box's WB
  // rows, so are the slopes of rows 1 .. LB-2, and face row WB is face row 0
  // (the same cells in the same order); each thread converts one cell and
  // computes one slope for every row holding them, the faces of rows
  // 0 .. WB-1 are computed once.  Uniform over the workgroup.
  XEnds xe;
  const bool xon = load_xends<D>(W, sinf, xs, xe);
  const bool rep = D != 0 && W.wrap == WB && pvf == nullptr;
  if (rep) {
    const int lx = tid % G::LX, r = tid / G::LX;
    double uc[5], q[NPRIM];
    pencil_cell<D>(W, bx, S, xon, xe, min(W.x0 + lx, xmax), dlo + r, uc);
    cons2prim(uc[0], uc[1], uc[2], uc[3], uc[4], c, q);
    for (int lb = (r + 2) % WB; lb < G::LB; lb += WB)
#pragma unroll
      for (int n = 0; n < NPRIM; ++n) sq[n][lb * G::LX + lx] = q[n];
  } else {
    // every round's loads first, then the conversions
    constexpr int NR = (G::NL + WNT - 1) / WNT;
    double uc[NR][5];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int l = tid + r * WNT;
      if (l < G::NL) {
        const int lx = l % G::LX, lb = l / G::LX;
        int x = W.x0 + lx - (D == 0 ? 2 : 0);
        int b = W.b0 + lb - (D == 0 ? 0 : 2);
        if (D == 0 && W.wrap > 0) x = dlo + ((x - dlo) % W.wrap + W.wrap) % W.wrap;
        else x = min(x, xmax);
        if (D != 0 && W.wrap > 0) b = dlo + ((b - dlo) % W.wrap + W.wrap) % W.wrap;
        else b = min(b, bmax);
        pencil_cell<D>(W, bx, S, xon, xe, x, b, uc[r]);
      }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int l = tid + r * WNT;
      if (l < G::NL) {
        double q[NPRIM];
        cons2prim(uc[r][0], uc[r][1], uc[r][2], uc[r][3], uc[r][4], c, q);
#pragma unroll
        for (int n = 0; n < NPRIM; ++n) sq[n][l] = q[n];
      }
    }
  }
  __syncthreads();

  // 2. slopes of the cells with both neighbours in the pencil (Riemann.h:19-120)
  constexpr int st = D == 0 ? 1 : G::LX;
  if (rep) {
    // row r + 2 (rows 2 .. 5, neighbours within 1 .. 6), stored in every row
    // of 1 .. LB-2 that repeats it
    const int lx = tid % G::LX, r = tid / G::LX;
    const int l = (r + 2) * G::LX + lx;
    auto Q = [&](int o, int comp) { return sq[comp][l + o * st]; };
    double d[5];
    muscl_slopes<D>(Q, c, d);
    for (int ld = (r + 2) % WB; ld <= G::LB - 2; ld += WB)
      if (ld >= 1)
#pragma unroll
        for (int n = 0; n < 5; ++n) sd[n][ld * G::LX + lx] = d[n];
  } else {
    for (int l = tid; l < G::NL; l += WNT) {
      const int ld = D == 0 ? l % G::LX : l / G::LX;
      if (ld < 1 || ld > (D == 0 ? G::LX : G::LB) - 2) continue;
      auto Q = [&](int o, int comp) { return sq[comp][l + o * st]; };
      double d[5];
      muscl_slopes<D>(Q, c, d);
#pragma unroll
      for (int n = 0; n < 5; ++n) sd[n][l] = d[n];
    }
  }
  __syncthreads();

  // 3. faces (Riemann.h:251-337): face f between pencil cells lL and lL + st
  constexpr int NFT = (G::NF + WNT - 1) / WNT;
  const int nf = rep ? WB * G::FX : G::NF;
  double F[NFT][5];
#pragma unroll
  for (int r = 0; r < NFT; ++r) {
    const int f = tid + r * WNT;
    if (f < nf) {
      const int fx = f % G::FX, fb = f / G::FX;
      const int lL = D == 0 ? fb * G::LX + fx + 1 : (fb + 1) * G::LX + fx;
      const int lR = lL + st;
      const double Tl = D == 2 ? sq[QT][lR + st] : sq[QT][lL];
      auto ql = [&](int n) { return sq[n][lL]; };
      auto dql = [&](int n) { return sd[n][lL]; };
      auto qr = [&](int n) { return sq[n][lR]; };
      auto dqr = [&](int n) { return sd[n][lR]; };
      muscl_face<D>(ql, dql, qr, dqr, Tl, c, F[r]);
      if (pvf) add_visc_face<D, G::FX>(W, bx, pvf[W.box], f, F[r]);
    }
  }
  __syncthreads();  // every pencil read done: the faces reuse the primitive slots
  double* sF = &sq[0][0];  // [5][NF]
  store_faces<G::FX, G::NF, NFT>(F, sF, nf, rep);
  __syncthreads();
  dir_epilogue<D, LAST, G::FX, G::NF, true>(W, bx, sa, sF, S, accf, epi, outf);
}

}  // namespace

int weno_brick_count(const BoxTable& bt, int d) {
  int64_t n = 0;
  for (int b = 0; b < bt.nbox; ++b) {
    const BoxDev& B = bt.h_boxes[b];
    const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1, nz = B.hi[2] - B.lo[2] + 1;
    const int nb = d == 2 ? nz : ny, nc = d == 2 ? ny : nz;
    n += int64_t((nx + WX - 1) / WX) * ((nb + WB - 1) / WB) * nc;
  }
  return int(n);
}

void weno_bricks(const BoxTable& bt, int d, const int* wrap, WenoBrick* out) {
  int64_t t = 0;
  for (int b = 0; b < bt.nbox; ++b) {
    const BoxDev& B = bt.h_boxes[b];
    const int ba = d == 2 ? 2 : 1, ca = d == 2 ? 1 : 2;
    for (int cc = B.lo[ca]; cc <= B.hi[ca]; ++cc)
      for (int b0 = B.lo[ba]; b0 <= B.hi[ba]; b0 += WB)
        for (int x0 = B.lo[0]; x0 <= B.hi[0]; x0 += WX) out[t++] = WenoBrick{b, x0, b0, cc, wrap[b]};
  }
}

hipError_t launch_dir_flux(int flux_euler, const WenoWork& ww, const BoxDev* boxes, const Closures& c,
                           const StageArgs& sa, const Fab* sin, const Fab* acc, const Fab* s1, const Fab* out,
                           hipStream_t st, const Fab* const* pv) {
  const Fab* pv0 = pv ? pv[0] : nullptr;
  const Fab* pv1 = pv ? pv[1] : nullptr;
  const Fab* pv2 = pv ? pv[2] : nullptr;
  if (flux_euler == 2) {
    if (ww.n[0] > 0)
      hipLaunchKernelGGL((k_weno_dir<0, 0>), dim3(unsigned(ww.n[0])), dim3(WNT), 0, st, ww.d[0], boxes, c, sa, sin,
                         acc, s1, out, pv0, ww.xside);
    if (ww.n[1] > 0)
      hipLaunchKernelGGL((k_weno_dir<1, 0>), dim3(unsigned(ww.n[1])), dim3(WNT), 0, st, ww.d[1], boxes, c, sa, sin,
                         acc, s1, out, pv1, ww.xside);
    if (ww.n[2] > 0)
      hipLaunchKernelGGL((k_weno_dir<2, 1>), dim3(unsigned(ww.n[2])), dim3(WNT), 0, st, ww.d[2], boxes, c, sa, sin,
                         acc, s1, out, pv2, ww.xside);
  } else if (flux_euler == 0) {
    if (ww.n[0] > 0)
      hipLaunchKernelGGL((k_muscl_dir<0, 0>), dim3(unsigned(ww.n[0])), dim3(WNT), 0, st, ww.d[0], boxes, c, sa, sin,
                         acc, s1, out, pv0, ww.xside);
    if (ww.n[1] > 0)
      hipLaunchKernelGGL((k_muscl_dir<1, 0>), dim3(unsigned(ww.n[1])), dim3(WNT), 0, st, ww.d[1], boxes, c, sa, sin,
                         acc, s1, out, pv1, ww.xside);
    if (ww.n[2] > 0)
      hipLaunchKernelGGL((k_muscl_dir<2, 1>), dim3(unsigned(ww.n[2])), dim3(WNT), 0, st, ww.d[2], boxes, c, sa, sin,
                         acc, s1, out, pv2, ww.xside);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace amr
