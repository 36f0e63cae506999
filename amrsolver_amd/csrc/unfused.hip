// Box-table kernels with the reference's pass structure: one launch per
// reference ParallelFor, covering every box of the level at once through a
// block map (block -> box, block within the box).  These serve every scheme (KEEP 2/4/6, WENO5,
// MUSCL+Riemann, viscous) and are the fallback where no fused stage kernel
// exists; intermediates live in box-table workspace arrays.
#include "cns_point.h"
#include "kernels.h"
#include "region.h"

namespace amr {

namespace {

// ---------------------------------------------------------------- prims
// CNS_advance.cpp:200-208: cons2prim on growntilebox(NGHOST)
__global__ void __launch_bounds__(BLK) k_prims(const BoxDev* boxes, const BlockRef* bmap, Region reg, const double* box_dt, const Fab* st,
                                               const Fab* pr, Closures c, WrapArgs wa) {
  REGION_LOOP_BEGIN
  const Fab u = st[b], q = pr[b];
  int ii = i, jj = j, kk = k;
  if (wa.on) {
    ii = ii < wa.lo[0] ? ii + wa.len[0] : (ii >= wa.lo[0] + wa.len[0] ? ii - wa.len[0] : ii);
    jj = jj < wa.lo[1] ? jj + wa.len[1] : (jj >= wa.lo[1] + wa.len[1] ? jj - wa.len[1] : jj);
    kk = kk < wa.lo[2] ? kk + wa.len[2] : (kk >= wa.lo[2] + wa.len[2] ? kk - wa.len[2] : kk);
  }
  double qq[6];
  cons2prim(u(ii, jj, kk, 0), u(ii, jj, kk, 1), u(ii, jj, kk, 2), u(ii, jj, kk, 3), u(ii, jj, kk, 4), c, qq);
#pragma unroll
  for (int n = 0; n < NPRIM; ++n) q(i, j, k, n) = qq[n];
}

// ---------------------------------------------------------------- KEEP
// CentralKEEP.h:40-145 on grownnodaltilebox(-1,0)
template <int H>
__global__ void __launch_bounds__(BLK) k_keep(const BoxDev* boxes, const BlockRef* bmap, Region reg, const double* box_dt, const Fab* pr,
                                              const Fab* nfx, const Fab* nfy, const Fab* nfz,
                                              Closures c, double c0, double c1, double c2) {
  REGION_LOOP_BEGIN
  const Fab q = pr[b];
  const double co[3] = {c0, c1, c2};
  double fx[5] = {0, 0, 0, 0, 0}, fy[5] = {0, 0, 0, 0, 0}, fz[5] = {0, 0, 0, 0, 0};
#define KEEP_DIR(F, NRM, DI, DJ, DK)                                                        \
  _Pragma("unroll") for (int l = 1; l <= H; ++l) {                                         \
    _Pragma("unroll") for (int m = 0; m <= l - 1; ++m) {                                   \
      int i1 = i + m * DI, j1 = j + m * DJ, k1 = k + m * DK;                               \
      int i2 = i + (m - l) * DI, j2 = j + (m - l) * DJ, k2 = k + (m - l) * DK;             \
      keep_pair<NRM>(co[l - 1], q(i1, j1, k1, QRHO), q(i1, j1, k1, QU), q(i1, j1, k1, QV),   \
                     q(i1, j1, k1, QW), c.cv * q(i1, j1, k1, QT), q(i1, j1, k1, QPRES),      \
                     q(i2, j2, k2, QRHO), q(i2, j2, k2, QU), q(i2, j2, k2, QV),              \
                     q(i2, j2, k2, QW), c.cv * q(i2, j2, k2, QT), q(i2, j2, k2, QPRES), F);  \
    }                                                                                       \
  }
  KEEP_DIR(fx, 0, 1, 0, 0)
  KEEP_DIR(fy, 1, 0, 1, 0)
  KEEP_DIR(fz, 2, 0, 0, 1)
#undef KEEP_DIR
  const Fab ox = nfx[b], oy = nfy[b], oz = nfz[b];
#pragma unroll
  for (int n = 0; n < NCONS; ++n) {
    ox(i, j, k, n) = fx[n];
    oy(i, j, k, n) = fy[n];
    oz(i, j, k, n) = fz[n];
  }
}

// ---------------------------------------------------------------- WENO5
// High_resolution.h:83-121 cons2eulerflux_lambda on growntilebox(NGHOST)
__global__ void __launch_bounds__(BLK) k_eulerflux(const BoxDev* boxes, const BlockRef* bmap, Region reg, const double* box_dt, const Fab* st,
                                                   const Fab* pfx, const Fab* pfy, const Fab* pfz,
                                                   const Fab* lam, Closures c) {
  REGION_LOOP_BEGIN
  const Fab u = st[b], fx = pfx[b], fy = pfy[b], fz = pfz[b], L = lam[b];
  double rhoinv = 1.0 / u(i, j, k, URHO);
  double momx = u(i, j, k, UMX), momy = u(i, j, k, UMY), momz = u(i, j, k, UMZ);
  double rhoet = u(i, j, k, UET);
  double ux = momx * rhoinv, uy = momy * rhoinv, uz = momz * rhoinv;
  double rhoekin = 0.5 * rhoinv * (momx * momx + momy * momy + momz * momz);
  double rhoeint = rhoet - rhoekin;
  double P = c.gm1 * rhoeint;
  fx(i, j, k, URHO) = momx;
  fx(i, j, k, UMX) = momx * ux + P;
  fx(i, j, k, UMY) = momy * ux;
  fx(i, j, k, UMZ) = momz * ux;
  fx(i, j, k, UET) = (rhoet + P) * ux;
  fy(i, j, k, URHO) = momy;
  fy(i, j, k, UMX) = momx * uy;
  fy(i, j, k, UMY) = momy * uy + P;
  fy(i, j, k, UMZ) = momz * uy;
  fy(i, j, k, UET) = (rhoet + P) * uy;
  fz(i, j, k, URHO) = momz;
  fz(i, j, k, UMX) = momx * uz;
  fz(i, j, k, UMY) = momy * uz;
  fz(i, j, k, UMZ) = momz * uz + P;
  fz(i, j, k, UET) = (rhoet + P) * uz;
  double cs = sqrt(c.gamma * P * rhoinv);
  L(i, j, k, 0) = fabs(dmax(dmax(ux + cs, ux - cs), ux));
  L(i, j, k, 1) = fabs(dmax(dmax(uy + cs, uy - cs), uy));
  L(i, j, k, 2) = fabs(dmax(dmax(uz + cs, uz - cs), uz));
}

// High_resolution.h:162-246 numericalflux_globallaxsplit for one direction
template <int D>
__device__ __forceinline__ double weno_face(int i, int j, int k, int n, const Fab& u, const Fab& pf,
                                            const Fab& lam) {
  constexpr int DI = D == 0, DJ = D == 1, DK = D == 2;
  double me = lam(i - 3 * DI, j - 3 * DJ, k - 3 * DK, D);
#pragma unroll
  for (int o = -2; o <= 2; ++o) me = dmax(me, lam(i + o * DI, j + o * DJ, k + o * DK, D));
  double df[6], P[6];
#pragma unroll
  for (int o = -3; o <= 2; ++o) {
    df[o + 3] = me * u(i + o * DI, j + o * DJ, k + o * DK, n);
    P[o + 3] = pf(i + o * DI, j + o * DJ, k + o * DK, n);
  }
  double flx = weno5js(0.5 * (P[0] + df[0]), 0.5 * (P[1] + df[1]), 0.5 * (P[2] + df[2]),
                       0.5 * (P[3] + df[3]), 0.5 * (P[4] + df[4]));
  flx += weno5js(0.5 * (P[5] - df[5]), 0.5 * (P[4] - df[4]), 0.5 * (P[3] - df[3]),
                 0.5 * (P[2] - df[2]), 0.5 * (P[1] - df[1]));
  return flx;
}

__global__ void __launch_bounds__(BLK) k_weno(const BoxDev* boxes, const BlockRef* bmap, Region reg, const double* box_dt, const Fab* st,
                                              const Fab* pfx, const Fab* pfy, const Fab* pfz,
                                              const Fab* lam, const Fab* nfx, const Fab* nfy,
                                              const Fab* nfz) {
  REGION_LOOP_BEGIN
  const Fab u = st[b], L = lam[b];
  const Fab ox = nfx[b], oy = nfy[b], oz = nfz[b];
  for (int n = 0; n < NCONS; ++n) {
    ox(i, j, k, n) = weno_face<0>(i, j, k, n, u, pfx[b], L);
    oy(i, j, k, n) = weno_face<1>(i, j, k, n, u, pfy[b], L);
    oz(i, j, k, n) = weno_face<2>(i, j, k, n, u, pfz[b], L);
  }
}

// ---------------------------------------------------------------- Riemann
// Riemann.h:19-120 cns_slope_{x,y,z} on grow(bx,dir,1)
template <int D>
__global__ void __launch_bounds__(BLK) k_slope(const BoxDev* boxes, const BlockRef* bmap, Region reg, const double* box_dt, const Fab* pr,
                                               const Fab* slope, Closures c) {
  REGION_LOOP_BEGIN
  constexpr int DI = D == 0, DJ = D == 1, DK = D == 2;
  const Fab q = pr[b], dq = slope[b];
  auto Q = [&](int o, int comp) { return q(i + o * DI, j + o * DJ, k + o * DK, comp); };
  double d[5];
  muscl_slopes<D>(Q, c, d);
#pragma unroll
  for (int n = 0; n < 5; ++n) dq(i, j, k, n) = d[n];
}

// Riemann.h:251-337 cns_riemann_{x,y,z} on surroundingNodes(bx,dir).  The z
// left state takes its sound speed from T at k+1 (Riemann.h:316).
template <int D>
__global__ void __launch_bounds__(BLK) k_riemann(const BoxDev* boxes, const BlockRef* bmap, Region reg, const double* box_dt, const Fab* pr,
                                                 const Fab* slope, const Fab* nf, Closures c) {
  REGION_LOOP_BEGIN
  constexpr int DI = D == 0, DJ = D == 1, DK = D == 2;
  const Fab q = pr[b], dq = slope[b], f = nf[b];
  const int il = i - DI, jl = j - DJ, kl = k - DK;
  const double Tl = D == 2 ? q(i, j, k + 1, QT) : q(il, jl, kl, QT);
  auto ql = [&](int n) { return q(il, jl, kl, n); };
  auto dql = [&](int n) { return dq(il, jl, kl, n); };
  auto qr = [&](int n) { return q(i, j, k, n); };
  auto dqr = [&](int n) { return dq(i, j, k, n); };
  double o[5];
  muscl_face<D>(ql, dql, qr, dqr, Tl, c, o);
#pragma unroll
  for (int n = 0; n < 5; ++n) f(i, j, k, n) = o[n];
}

// ---------------------------------------------------------------- JST
// CNS_hydro_K.h:38-44 disconSensor
__device__ __forceinline__ double discon_sensor(double pp, double pl, double pr) {
  double pjst = pr + 2.0 * pp + pl;
  double ptvd = fabs(pr - pp) + fabs(pp - pl);
  return fabs(2.0 * (pr - 2.0 * pp + pl) / (pjst + ptvd + 1.0e-40));
}

// ComputeSensorLambda (CNS_hydro_K.h:46-81) on the valid box; the rest of
// grow(bx,1) is set to 0, the value the reference's zeroed, never filled
// sensor and spectral-radius MultiFabs hold there (CNS_advance.cpp:237-257)
__global__ void __launch_bounds__(BLK) k_jst_sl(const BoxDev* boxes, const BlockRef* bmap, Region reg,
                                                const double* box_dt, const Fab* pr, const Fab* jst, Closures c) {
  REGION_LOOP_BEGIN
  const Fab q = pr[b], J = jst[b];
  const bool valid = i >= bx.lo[0] && i <= bx.hi[0] && j >= bx.lo[1] && j <= bx.hi[1] && k >= bx.lo[2] &&
                     k <= bx.hi[2];
  double v[6] = {0, 0, 0, 0, 0, 0};
  if (valid) {
    const double pp = q(i, j, k, QPRES);
    v[0] = discon_sensor(pp, q(i - 1, j, k, QPRES), q(i + 1, j, k, QPRES));
    v[1] = discon_sensor(pp, q(i, j - 1, k, QPRES), q(i, j + 1, k, QPRES));
    v[2] = discon_sensor(pp, q(i, j, k - 1, QPRES), q(i, j, k + 1, QPRES));
    const double cs = sqrt(c.gamma * q(i, j, k, QPRES) / q(i, j, k, QRHO));
    v[3] = fabs(q(i, j, k, QU)) + cs;
    v[4] = fabs(q(i, j, k, QV)) + cs;
    v[5] = fabs(q(i, j, k, QW)) + cs;
  }
#pragma unroll
  for (int n = 0; n < 6; ++n) J(i, j, k, n) = v[n];
}

// JSTflux (CNS_hydro_K.h:84-104) on grownnodaltilebox(-1,0): every direction
// reads component 0 of sensor and spectral radius, and the z damping stencil
// reads cons(k-1) three times, as the reference does
__global__ void __launch_bounds__(BLK) k_jst_flux(const BoxDev* boxes, const BlockRef* bmap, Region reg,
                                                  const double* box_dt, const Fab* st, const Fab* jst,
                                                  const Fab* nfx, const Fab* nfy, const Fab* nfz, Closures c) {
  REGION_LOOP_BEGIN
  const Fab u = st[b], J = jst[b], fx = nfx[b], fy = nfy[b], fz = nfz[b];
#pragma unroll
  for (int n = 0; n < NCONS; ++n) {
    const double uijk = u(i, j, k, n);
    double dw = uijk - u(i - 1, j, k, n);
    double sn = c.Cshock * dmax(J(i - 1, j, k, 0), J(i, j, k, 0));
    double rr = dmax(J(i - 1, j, k, 3), J(i, j, k, 3));
    double fd = u(i + 1, j, k, n) - 3.0 * u(i, j, k, n) + 3.0 * u(i - 1, j, k, n) - u(i - 2, j, k, n);
    fx(i, j, k, n) = fx(i, j, k, n) - (sn * dw - dmax(0.0, c.Cdamp - sn) * fd) * rr;
    dw = uijk - u(i, j - 1, k, n);
    sn = c.Cshock * dmax(J(i, j - 1, k, 0), J(i, j, k, 0));
    rr = dmax(J(i, j - 1, k, 3), J(i, j, k, 3));
    fd = u(i, j + 1, k, n) - 3.0 * u(i, j, k, n) + 3.0 * u(i, j - 1, k, n) - u(i, j - 2, k, n);
    fy(i, j, k, n) = fy(i, j, k, n) - (sn * dw - dmax(0.0, c.Cdamp - sn) * fd) * rr;
    dw = uijk - u(i, j, k - 1, n);
    sn = c.Cshock * dmax(J(i, j, k - 1, 0), J(i, j, k, 0));
    rr = dmax(J(i, j, k - 1, 3), J(i, j, k, 3));
    fd = u(i, j, k - 1, n) - 3.0 * u(i, j, k, n) + 3.0 * u(i, j, k - 1, n) - u(i, j, k - 1, n);
    fz(i, j, k, n) = fz(i, j, k, n) - (sn * dw - dmax(0.0, c.Cdamp - sn) * fd) * rr;
  }
}

// ---------------------------------------------------------------- viscous
// CNS_hydro_K.h:108-172 ViscousFluxes on growntilebox(1)
__global__ void __launch_bounds__(BLK) k_visc_pf(const BoxDev* boxes, const BlockRef* bmap, Region reg, const double* box_dt, const Fab* pr,
                                                 const Fab* pfx, const Fab* pfy, const Fab* pfz,
                                                 Closures c) {
  REGION_LOOP_BEGIN
  const double dx0 = bx.dxi[0], dx1 = bx.dxi[1], dx2 = bx.dxi[2];
  const Fab q = pr[b], fx = pfx[b], fy = pfy[b], fz = pfz[b];
  double ux = q(i, j, k, QU);
  double dudx = (q(i + 1, j, k, QU) - q(i - 1, j, k, QU)) * 0.5 * dx0;
  double dvdx = (q(i + 1, j, k, QV) - q(i - 1, j, k, QV)) * 0.5 * dx0;
  double dwdx = (q(i + 1, j, k, QW) - q(i - 1, j, k, QW)) * 0.5 * dx0;
  double dTdx = (q(i + 1, j, k, QT) - q(i - 1, j, k, QT)) * 0.5 * dx0;
  double uy = q(i, j, k, QV);
  double dudy = (q(i, j + 1, k, QU) - q(i, j - 1, k, QU)) * 0.5 * dx1;
  double dvdy = (q(i, j + 1, k, QV) - q(i, j - 1, k, QV)) * 0.5 * dx1;
  double dwdy = (q(i, j + 1, k, QW) - q(i, j - 1, k, QW)) * 0.5 * dx1;
  double dTdy = (q(i, j + 1, k, QT) - q(i, j - 1, k, QT)) * 0.5 * dx1;
  double uz = q(i, j, k, QW);
  double dudz = (q(i, j, k + 1, QU) - q(i, j, k - 1, QU)) * 0.5 * dx2;
  double dvdz = (q(i, j, k + 1, QV) - q(i, j, k - 1, QV)) * 0.5 * dx2;
  double dwdz = (q(i, j, k + 1, QW) - q(i, j, k - 1, QW)) * 0.5 * dx2;
  double dTdz = (q(i, j, k + 1, QT) - q(i, j, k - 1, QT)) * 0.5 * dx2;
  double div = dudx + dvdy + dwdz;
  double T = q(i, j, k, QT);
  double mu = visc(c, T);
  double lambda = cond(c, T, mu);
  double r1_3 = 1.0 / 3.0;
  double tauxx = 2.0 * mu * (dudx - r1_3 * div);
  double tauxy = mu * (dudy + dvdx);
  double tauxz = mu * (dudz + dwdx);
  double tauyy = 2.0 * mu * (dvdy - r1_3 * div);
  double tauyz = mu * (dvdz + dwdy);
  double tauzz = 2.0 * mu * (dwdz - r1_3 * div);
  fx(i, j, k, URHO) = 0.0;
  fx(i, j, k, UMX) = -tauxx;
  fx(i, j, k, UMY) = -tauxy;
  fx(i, j, k, UMZ) = -tauxz;
  fx(i, j, k, UET) = -lambda * dTdx - tauxx * ux - tauxy * uy - tauxz * uz;
  fy(i, j, k, URHO) = 0.0;
  fy(i, j, k, UMX) = -tauxy;
  fy(i, j, k, UMY) = -tauyy;
  fy(i, j, k, UMZ) = -tauyz;
  fy(i, j, k, UET) = -lambda * dTdy - tauxy * ux - tauyy * uy - tauyz * uz;
  fz(i, j, k, URHO) = 0.0;
  fz(i, j, k, UMX) = -tauxz;
  fz(i, j, k, UMY) = -tauyz;
  fz(i, j, k, UMZ) = -tauzz;
  fz(i, j, k, UET) = -lambda * dTdz - tauxz * ux - tauyz * uy - tauzz * uz;
}

// CNS_hydro_K.h:254-261 ViscousNumericalFluxes on grownnodaltilebox(-1,0)
__global__ void __launch_bounds__(BLK) k_visc_add(const BoxDev* boxes, const BlockRef* bmap, Region reg, const double* box_dt, const Fab* pfx,
                                                  const Fab* pfy, const Fab* pfz, const Fab* nfx,
                                                  const Fab* nfy, const Fab* nfz) {
  REGION_LOOP_BEGIN
  const Fab px = pfx[b], py = pfy[b], pz = pfz[b], ox = nfx[b], oy = nfy[b], oz = nfz[b];
#pragma unroll
  for (int n = 0; n < NCONS; ++n) {
    ox(i, j, k, n) += 0.5 * (px(i - 1, j, k, n) + px(i, j, k, n));
    oy(i, j, k, n) += 0.5 * (py(i, j - 1, k, n) + py(i, j, k, n));
    oz(i, j, k, n) += 0.5 * (pz(i, j, k - 1, n) + pz(i, j, k, n));
  }
}

// ---------------------------------------------------------------- walls
// Flux_2nd_Order_KEEP (src/rhs/CentralKEEP.h:233-286): next to a bc 6 wall
// the y fluxes of faces j = lo .. lo+h-1 (lower wall) or hi-h+1 .. hi (upper
// wall; the wall face hi+1 itself keeps its full-order flux, as in the
// reference) are recomputed with second-order KEEP (KEEPy, :147-186, one
// pair with coefficient 1).
__global__ void __launch_bounds__(BLK) k_keep_wall(const WallTask* tasks, const TaskBlock* blocks,
                                                   const BoxDev* boxes, const Fab* pr, const Fab* nfy, Closures c,
                                                   int h) {
  const TaskBlock tb = blocks[blockIdx.x];
  const WallTask& T = tasks[tb.task];
  int cc[3];
  if (!task_cell(T, tb, threadIdx.x, cc)) return;
  const int i = cc[0], j = cc[1], k = cc[2];
  const int wall_j = T.side == 0 ? boxes[T.box].lo[1] : boxes[T.box].hi[1];
  // lower wall: faces lo .. lo+h-1; upper wall: faces hi-h+1 .. hi (the task
  // region starts at the box's lowest face / ends at its highest cell)
  if (T.side == 0 ? j > wall_j + h - 1 : j < wall_j - h + 1) return;
  const Fab q = pr[T.box];
  double f[5] = {0, 0, 0, 0, 0};
  keep_pair<1>(1.0, q(i, j, k, QRHO), q(i, j, k, QU), q(i, j, k, QV), q(i, j, k, QW), c.cv * q(i, j, k, QT),
               q(i, j, k, QPRES), q(i, j - 1, k, QRHO), q(i, j - 1, k, QU), q(i, j - 1, k, QV), q(i, j - 1, k, QW),
               c.cv * q(i, j - 1, k, QT), q(i, j - 1, k, QPRES), f);
  const Fab oy = nfy[T.box];
#pragma unroll
  for (int n = 0; n < NCONS; ++n) oy(i, j, k, n) = f[n];
}

// ViscousWallFluxes (src/rhs/CNS_hydro_K.h:177-251) on the ghost row outside
// a bc 6 wall (src/tim/CNS_advance.cpp:303-327): the cell fluxes there are
// recomputed with one-sided second-order y derivatives, from the row and
// the two rows inward.
__global__ void __launch_bounds__(BLK) k_visc_wall(const WallTask* tasks, const TaskBlock* blocks,
                                                   const BoxDev* boxes, const Fab* pr, const Fab* pfx,
                                                   const Fab* pfy, const Fab* pfz, Closures c) {
  const TaskBlock tb = blocks[blockIdx.x];
  const WallTask& T = tasks[tb.task];
  int cc[3];
  if (!task_cell(T, tb, threadIdx.x, cc)) return;
  const int i = cc[0], j = cc[1], k = cc[2];
  const BoxDev bx = boxes[T.box];
  const double dx0 = bx.dxi[0], dx1 = bx.dxi[1], dx2 = bx.dxi[2];
  const Fab q = pr[T.box], fx = pfx[T.box], fy = pfy[T.box], fz = pfz[T.box];
  double ux = q(i, j, k, QU);
  double dudx = (q(i + 1, j, k, QU) - q(i - 1, j, k, QU)) * 0.5 * dx0;
  double dvdx = (q(i + 1, j, k, QV) - q(i - 1, j, k, QV)) * 0.5 * dx0;
  double dwdx = (q(i + 1, j, k, QW) - q(i - 1, j, k, QW)) * 0.5 * dx0;
  double dTdx = (q(i + 1, j, k, QT) - q(i - 1, j, k, QT)) * 0.5 * dx0;
  double uy = q(i, j, k, QV);
  double dudy, dvdy, dwdy, dTdy;
  if (T.side == 0) {
    dudy = (-q(i, j + 2, k, QU) + 4 * q(i, j + 1, k, QU) - 3 * q(i, j, k, QU)) * 0.5 * dx1;
    dvdy = (-q(i, j + 2, k, QV) + 4 * q(i, j + 1, k, QV) - 3 * q(i, j, k, QV)) * 0.5 * dx1;
    dwdy = (-q(i, j + 2, k, QW) + 4 * q(i, j + 1, k, QW) - 3 * q(i, j, k, QW)) * 0.5 * dx1;
    dTdy = (-q(i, j + 2, k, QT) + 4 * q(i, j + 1, k, QT) - 3 * q(i, j, k, QT)) * 0.5 * dx1;
  } else {
    dudy = (3 * q(i, j, k, QU) - 4 * q(i, j - 1, k, QU) + q(i, j - 2, k, QU)) * 0.5 * dx1;
    dvdy = (3 * q(i, j, k, QV) - 4 * q(i, j - 1, k, QV) + q(i, j - 2, k, QV)) * 0.5 * dx1;
    dwdy = (3 * q(i, j, k, QW) - 4 * q(i, j - 1, k, QW) + q(i, j - 2, k, QW)) * 0.5 * dx1;
    dTdy = (3 * q(i, j, k, QT) - 4 * q(i, j - 1, k, QT) + q(i, j - 2, k, QT)) * 0.5 * dx1;
  }
  double uz = q(i, j, k, QW);
  double dudz = (q(i, j, k + 1, QU) - q(i, j, k - 1, QU)) * 0.5 * dx2;
  double dvdz = (q(i, j, k + 1, QV) - q(i, j, k - 1, QV)) * 0.5 * dx2;
  double dwdz = (q(i, j, k + 1, QW) - q(i, j, k - 1, QW)) * 0.5 * dx2;
  double dTdz = (q(i, j, k + 1, QT) - q(i, j, k - 1, QT)) * 0.5 * dx2;
  double div = dudx + dvdy + dwdz;
  double T_ = q(i, j, k, QT);
  double mu = visc(c, T_);
  double lambda = cond(c, T_, mu);
  double r1_3 = 1.0 / 3.0;
  double tauxx = 2.0 * mu * (dudx - r1_3 * div);
  double tauxy = mu * (dudy + dvdx);
  double tauxz = mu * (dudz + dwdx);
  double tauyy = 2.0 * mu * (dvdy - r1_3 * div);
  double tauyz = mu * (dvdz + dwdy);
  double tauzz = 2.0 * mu * (dwdz - r1_3 * div);
  fx(i, j, k, URHO) = 0.0;
  fx(i, j, k, UMX) = -tauxx;
  fx(i, j, k, UMY) = -tauxy;
  fx(i, j, k, UMZ) = -tauxz;
  fx(i, j, k, UET) = -lambda * dTdx - tauxx * ux - tauxy * uy - tauxz * uz;
  fy(i, j, k, URHO) = 0.0;
  fy(i, j, k, UMX) = -tauxy;
  fy(i, j, k, UMY) = -tauyy;
  fy(i, j, k, UMZ) = -tauyz;
  fy(i, j, k, UET) = -lambda * dTdy - tauxy * ux - tauyy * uy - tauyz * uz;
  fz(i, j, k, URHO) = 0.0;
  fz(i, j, k, UMX) = -tauxz;
  fz(i, j, k, UMY) = -tauyz;
  fz(i, j, k, UMZ) = -tauzz;
  fz(i, j, k, UET) = -lambda * dTdz - tauxz * ux - tauyz * uy - tauzz * uz;
}

// ---------------------------------------------------------------- assembly
// CNS_advance.cpp:376-391 plus the stage combination of CNS::advance
__global__ void __launch_bounds__(BLK) k_assemble(const BoxDev* boxes, const BlockRef* bmap, Region reg, const double* box_dt, const Fab* nfx,
                                                  const Fab* nfy, const Fab* nfz, StageArgs sa,
                                                  const Fab* s1, const Fab* sin, const Fab* out) {
  REGION_LOOP_BEGIN
  const double dx0 = bx.dxi[0], dx1 = bx.dxi[1], dx2 = bx.dxi[2];
  const StageCoef sc = box_stage(sa, b);
  const Fab fx = nfx[b], fy = nfy[b], fz = nfz[b], o = out[b];
#pragma unroll
  for (int n = 0; n < NCONS; ++n) {
    double L = dx0 * (fx(i, j, k, n) - fx(i + 1, j, k, n)) + dx1 * (fy(i, j, k, n) - fy(i, j + 1, k, n)) +
               dx2 * (fz(i, j, k, n) - fz(i, j, k + 1, n));
    if (sa.src) L = L + sa.src[b](i, j, k, n);  // user_source, CNS_advance.cpp:397-407
    // body-force user_source (exm/Channel/prob.h:189-193)
    if (sa.bfx != 0.0 && n == UMX) L = L + sin[b](i, j, k, URHO) * sa.bfx;
    if (sa.bfx != 0.0 && n == UET) L = L + sin[b](i, j, k, UMX) * sa.bfx;
    double v;
    if (sc.mode == 2 || sa.rhs_only) {
      v = L;
    } else {
      double a = (sc.a != 0.0) ? s1[b](i, j, k, n) : 0.0;
      v = stage_combine(sc, a, sin[b](i, j, k, n), L);
    }
    o(i, j, k, n) = v;
  }
}

}  // namespace

hipError_t launch_prims(const BoxTable& bt, const UnfusedWork& w, const Closures& c, const double* box_dt,
                        const Fab* sin, const WrapArgs& wa, hipStream_t st) {
  const Region g3 = grown(NGHOST);
  Launch L = grid_for(bt, g3);
  if (!L.map) return hipErrorNotInitialized;
  hipLaunchKernelGGL(k_prims, L.grid, dim3(BLK), 0, st, bt.d_boxes, L.map, g3, box_dt, sin, w.prims, c, wa);
  return hipGetLastError();
}

hipError_t launch_prims_region(const BoxTable& bt, const UnfusedWork& w, const Closures& c, const double* box_dt,
                               const Fab* sin, const Region& r, hipStream_t st) {
  Launch L = grid_for(bt, r);
  if (!L.map) return hipErrorNotInitialized;
  hipLaunchKernelGGL(k_prims, L.grid, dim3(BLK), 0, st, bt.d_boxes, L.map, r, box_dt, sin, w.prims, c,
                     WrapArgs{0, {0, 0, 0}, {1, 1, 1}});
  return hipGetLastError();
}

hipError_t prepare_block_maps(const BoxTable& bt) {
  const Region fixed[4] = {grown(NGHOST), nodal(), grown(1), grown(0)};
  for (const Region& r : fixed) {
    hipError_t e = build_map(bt, r);
    if (e != hipSuccess) return e;
  }
  for (int d = 0; d < 3; ++d) {
    hipError_t e = build_map(bt, slope_region(d));
    if (e == hipSuccess) e = build_map(bt, face_region(d));
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Flux phase + assembly + stage combination, reference pass structure.
// out = the stage combination of s1, sin and the dSdt held in `rhs` (the
// combination k_assemble applies, after a pass that changed dSdt).
__global__ void __launch_bounds__(BLK) k_combine(const BoxDev* boxes, const BlockRef* bmap, Region reg,
                                                 const double* box_dt, StageArgs sa, const Fab* s1, const Fab* sin,
                                                 const Fab* rhs, const Fab* out) {
  REGION_LOOP_BEGIN
  const StageCoef sc = box_stage(sa, b);
  const Fab r = rhs[b], o = out[b];
#pragma unroll
  for (int n = 0; n < NCONS; ++n) {
    const double L = r(i, j, k, n);
    double v;
    if (sc.mode == 2) {
      v = L;
    } else {
      double a = (sc.a != 0.0) ? s1[b](i, j, k, n) : 0.0;
      v = stage_combine(sc, a, sin[b](i, j, k, n), L);
    }
    o(i, j, k, n) = v;
  }
}

// dst(i,j,k,n) = src(i,j,k,n) over region reg of every box (zero when src is
// null): the face fluxes of one direction out of the workspace.
__global__ void __launch_bounds__(BLK) k_copy_faces(const BoxDev* boxes, const BlockRef* bmap, Region reg,
                                                    const double* box_dt, const Fab* src, const Fab* dst) {
  REGION_LOOP_BEGIN
  const Fab o = dst[b];
  if (src) {
    const Fab f = src[b];
#pragma unroll
    for (int n = 0; n < NCONS; ++n) o(i, j, k, n) = f(i, j, k, n);
  } else {
#pragma unroll
    for (int n = 0; n < NCONS; ++n) o(i, j, k, n) = 0.0;
  }
}

hipError_t launch_combine(const BoxTable& bt, const StageArgs& sa, const Fab* s1, const Fab* sin, const Fab* rhs,
                          const Fab* out, hipStream_t st) {
  Launch L_ = grid_for(bt, grown(0));
  if (!L_.map) return hipErrorNotInitialized;
  hipLaunchKernelGGL(k_combine, L_.grid, dim3(BLK), 0, st, bt.d_boxes, L_.map, grown(0), sa.box_dt, sa, s1, sin,
                     rhs, out);
  return hipGetLastError();
}

hipError_t launch_unfused_stage(const BoxTable& bt, const UnfusedWork& w, const Closures& c,
                                const amr_scheme_t& s, const StageArgs& sa, const WallArgs& wa, const Fab* s1,
                                const Fab* sin, const Fab* out, hipStream_t st, const Fab* const* faces) {
  const BoxDev* B = bt.d_boxes;
  const double* bdt = sa.box_dt;
#define LAUNCH(KERNEL, REG, ...)                                                                    \
  do {                                                                                             \
    Launch L_ = grid_for(bt, REG);                                                                 \
    if (!L_.map) return hipErrorNotInitialized;                                                      \
    hipLaunchKernelGGL(KERNEL, L_.grid, dim3(BLK), 0, st, B, L_.map, REG, bdt, __VA_ARGS__);     \
  } while (0)
  const Region g3 = grown(NGHOST), nod = nodal(), g1 = grown(1), g0 = grown(0);
  // WENO or MUSCL: directional passes from the conserved state, each
  // forming what it needs of its pencil in LDS; with viscous terms the
  // cell-centred viscous fluxes are computed first (primitives on grow(bx,3),
  // ViscousFluxes on grow(bx,1), the wall rows' own fluxes next to bc 6
  // walls) and each pass adds their face averages
  if (!faces && !sa.rhs_only && s.rhs_euler && (s.flux_euler == 2 || s.flux_euler == 0) && !s.art_diss &&
      w.weno.d[0] != nullptr) {
    if (!s.rhs_visc) return launch_dir_flux(s.flux_euler, w.weno, B, c, sa, sin, w.nf[0], s1, out, st);
    if (w.pv[0] != nullptr) {
      LAUNCH(k_prims, g3, sin, w.prims, c, WrapArgs{0, {0, 0, 0}, {1, 1, 1}});
      LAUNCH(k_visc_pf, g1, w.prims, w.pv[0], w.pv[1], w.pv[2], c);
      if (wa.nvisc_blk > 0)
        hipLaunchKernelGGL(k_visc_wall, dim3(unsigned(wa.nvisc_blk)), dim3(BLK), 0, st, wa.visc, wa.visc_blk, B,
                           w.prims, w.pv[0], w.pv[1], w.pv[2], c);
      return launch_dir_flux(s.flux_euler, w.weno, B, c, sa, sin, w.nf[0], s1, out, st, w.pv);
    }
  }
  LAUNCH(k_prims, g3, sin, w.prims, c, WrapArgs{0, {0, 0, 0}, {1, 1, 1}});
  if (s.rhs_euler) {
    if (s.flux_euler == 1) {
      double co[3];
      keep_coefficients(s.order_keep, co);
      if (s.order_keep == 6)
        LAUNCH(k_keep<3>, nod, w.prims, w.nf[0], w.nf[1], w.nf[2], c, co[0], co[1], co[2]);
      else if (s.order_keep == 4)
        LAUNCH(k_keep<2>, nod, w.prims, w.nf[0], w.nf[1], w.nf[2], c, co[0], co[1], co[2]);
      else
        LAUNCH(k_keep<1>, nod, w.prims, w.nf[0], w.nf[1], w.nf[2], c, co[0], co[1], co[2]);
      // CNS_advance.cpp:226-229: order reduction next to bc 6 walls
      if (s.order_keep != 2 && wa.nkeep_blk > 0)
        hipLaunchKernelGGL(k_keep_wall, dim3(unsigned(wa.nkeep_blk)), dim3(BLK), 0, st, wa.keep, wa.keep_blk, B,
                           w.prims, w.nf[1], c, s.order_keep / 2);
    } else if (s.flux_euler == 2) {
      LAUNCH(k_eulerflux, g3, sin, w.pe[0], w.pe[1], w.pe[2], w.lam, c);
      LAUNCH(k_weno, nod, sin, w.pe[0], w.pe[1], w.pe[2], w.lam, w.nf[0], w.nf[1], w.nf[2]);
    } else {
      for (int d = 0; d < 3; ++d) {
        const Region rs = slope_region(d), rf = face_region(d);
        if (d == 0) {
          LAUNCH(k_slope<0>, rs, w.prims, w.slope, c);
          LAUNCH(k_riemann<0>, rf, w.prims, w.slope, w.nf[0], c);
        } else if (d == 1) {
          LAUNCH(k_slope<1>, rs, w.prims, w.slope, c);
          LAUNCH(k_riemann<1>, rf, w.prims, w.slope, w.nf[1], c);
        } else {
          LAUNCH(k_slope<2>, rs, w.prims, w.slope, c);
          LAUNCH(k_riemann<2>, rf, w.prims, w.slope, w.nf[2], c);
        }
      }
    }
  }
  // CNS_advance.cpp:235-265 JST dissipation subtracted from the Euler fluxes
  if (s.rhs_euler && s.art_diss) {
    LAUNCH(k_jst_sl, g1, w.prims, w.jst, c);
    LAUNCH(k_jst_flux, nod, sin, w.jst, w.nf[0], w.nf[1], w.nf[2], c);
  }
  if (s.rhs_visc) {
    LAUNCH(k_visc_pf, g1, w.prims, w.pv[0], w.pv[1], w.pv[2], c);
    if (wa.nvisc_blk > 0)
      hipLaunchKernelGGL(k_visc_wall, dim3(unsigned(wa.nvisc_blk)), dim3(BLK), 0, st, wa.visc, wa.visc_blk, B,
                         w.prims, w.pv[0], w.pv[1], w.pv[2], c);
    LAUNCH(k_visc_add, nod, w.pv[0], w.pv[1], w.pv[2], w.nf[0], w.nf[1], w.nf[2]);
  }
  if (faces) {  // the face fluxes themselves instead of their assembly
    for (int d = 0; d < 3; ++d) LAUNCH(k_copy_faces, face_region(d), s.rhs_euler ? w.nf[d] : nullptr, faces[d]);
    return hipGetLastError();
  }
  LAUNCH(k_assemble, g0, w.nf[0], w.nf[1], w.nf[2], sa, s1, sin, out);
#undef LAUNCH
  return hipGetLastError();
}

}  // namespace amr
