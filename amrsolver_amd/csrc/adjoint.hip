// Reverse mode of the KEEP + viscous right-hand side and of the SSPRK stage
// update: hand-written vector-Jacobian products over the box table.
//
// For a box b with ghost-filled input S (ghost width 3), the forward is
//   q = cons2prim(S)                       on grow(b,3)
//   F_d(f) = sum_l sum_m co_l P_d(q(f+m), q(f+m-l))    KEEP pairs, faces f
//          + 0.5 (pf_d(f-1) + pf_d(f))    viscous cell fluxes on grow(b,1)
//   dSdt(c) = sum_d dxi_d (F_d(c) - F_d(c+e_d))        valid cells
// (src/rhs/CNS_hydro_K.h:15-34,108-172,254-261; src/rhs/CentralKEEP.h:40-145;
// src/tim/CNS_advance.cpp:375-391).  Given lam = adjoint of dSdt on the valid
// cells, the kernels here compute G = (d dSdt / d S)^T lam on grow(b,3) by
// gathering, per cell, every term that reads that cell:
//   * the face adjoint is dxi_d (lam(f) - lam(f-1)), lam = 0 outside the box,
//     so the weight of pair (a, a-l) telescopes to co_l dxi_d (lam(a) - lam(a-l));
//   * viscous: a pre-pass stores per cell of grow(b,1) the adjoints of the
//     velocity/temperature differences and of the cell's own values;
//   * cons2prim is transposed per cell.
// The stage combination is transposed on the valid cells in the same pass;
// the ghost fill is then transposed (BC tasks in reverse direction order,
// then the copy tasks) by weighted accumulation into the stage adjoint at
// the source cells, in a fixed order and without atomics (one lane per
// boundary line; copy tasks in classes whose sources do not overlap), so
// that runs repeat bit for bit.
#include "cns_point.h"
#include "kernels.h"
#include "region.h"

namespace amr {

namespace {

constexpr int NE = 16;  // L_u, L_v, L_w, L_T, then (E_u, E_v, E_w, E_T) for x, y, z

__device__ __forceinline__ bool in_box(const BoxDev& bx, int g, int i, int j, int k) {
  return i >= bx.lo[0] - g && i <= bx.hi[0] + g && j >= bx.lo[1] - g && j <= bx.hi[1] + g &&
         k >= bx.lo[2] - g && k <= bx.hi[2] + g;
}

// Periodic index of direction d for a domain-wide array (identity when off).
template <bool WRAP>
__device__ __forceinline__ int wr(const WrapArgs& wa, int d, int x) {
  if (!WRAP) return x;
  return x < wa.lo[d] ? x + wa.len[d] : (x >= wa.lo[d] + wa.len[d] ? x - wa.len[d] : x);
}

// adjoint of dSdt at (i,j,k): zero outside the valid box
__device__ __forceinline__ void lam_at(const Fab& L, const BoxDev& bx, int i, int j, int k, double v[5]) {
  if (in_box(bx, 0, i, j, k)) {
#pragma unroll
    for (int n = 0; n < NCONS; ++n) v[n] = L(i, j, k, n);
  } else {
#pragma unroll
    for (int n = 0; n < NCONS; ++n) v[n] = 0.0;
  }
}

// d mu / dT and d k / dT of the closures (Sutherland: mu (1.5/T - 1/(S+T)))
__device__ __forceinline__ void transport_derivs(const Closures& c, double T, double mu, double k, double& dmu,
                                                 double& dk) {
  dmu = c.visc_law == 0 ? 0.0 : mu * (1.5 / T - 1.0 / (c.Tvisc_ref + T));
  if (c.cond_law == 0) dk = 0.0;
  else if (c.cond_law == 1) dk = k * (1.5 / T - 1.0 / (c.Tcond_ref + T));
  else dk = dmu * c.cp / c.Pr;
}

// Reverse of one KEEP pair in direction NRM (keep_pair): w[5] is the weight
// of the pair's five flux components; adds (dP/dq1)^T w to g1 (G1) and
// (dP/dq2)^T w to g2 (G2), q = (rho, u, v, w, T, p), ie = cv T.
// With A = (rho1+rho2)(un1+un2) the pair is
//   P.w = A Gs + 0.5 (p1+p2) w[1+NRM] + 0.5 w[4] (p1 un2 + p2 un1),
//   Gs = w0/4 + sum_t w[1+t] (ut1+ut2)/8 + w[4] (ke/4 + (ie1+ie2)/8).
template <int NRM, bool G1, bool G2>
__device__ __forceinline__ void keep_pair_vjp(const double w[5], const double q1[6], const double q2[6], double cv,
                                              double g1[6], double g2[6]) {
  const double R = q1[QRHO] + q2[QRHO];
  const double un1 = q1[QU + NRM], un2 = q2[QU + NRM];
  const double U = un1 + un2;
  const double A = R * U;
  const double ke = 0.5 * (q1[QU] * q2[QU] + q1[QV] * q2[QV] + q1[QW] * q2[QW]);
  const double we = w[UET];
  double Gs = 0.25 * w[URHO] + we * (0.25 * ke + 0.125 * cv * (q1[QT] + q2[QT]));
#pragma unroll
  for (int t = 0; t < 3; ++t) Gs += 0.125 * w[UMX + t] * (q1[QU + t] + q2[QU + t]);
  const double A8 = 0.125 * A;
  if (G1) {
    g1[QRHO] += U * Gs;
#pragma unroll
    for (int t = 0; t < 3; ++t) g1[QU + t] += A8 * (w[UMX + t] + we * q2[QU + t]);
    g1[QU + NRM] += R * Gs + 0.5 * we * q2[QPRES];
    g1[QT] += A8 * we * cv;
    g1[QPRES] += 0.5 * (w[UMX + NRM] + we * un2);
  }
  if (G2) {
    g2[QRHO] += U * Gs;
#pragma unroll
    for (int t = 0; t < 3; ++t) g2[QU + t] += A8 * (w[UMX + t] + we * q1[QU + t]);
    g2[QU + NRM] += R * Gs + 0.5 * we * q1[QPRES];
    g2[QT] += A8 * we * cv;
    g2[QPRES] += 0.5 * (w[UMX + NRM] + we * un1);
  }
}

__device__ __forceinline__ void load_q(const Fab& q, int i, int j, int k, double v[6]) {
#pragma unroll
  for (int n = 0; n < NPRIM; ++n) v[n] = q(i, j, k, n);
}

// KEEP gather of direction D into g for cell (i,j,k) with primitives qc.
// With WRAP (a periodic domain-wide level) every cell is valid: no range
// conditions, neighbour indices wrap periodically.  Each pair still sits
// behind a uniform branch on wa.on (always set when WRAP is launched): in one
// straight-line block the compiler issues every neighbour load of the three
// directions first and needs ~370 VGPRs (one wave per SIMD); with a block per
// pair it keeps ~110 (four waves).
template <int H, int D, bool WRAP = false>
__device__ __forceinline__ void adj_keep_dir(const Fab& q, const Fab& L, const BoxDev& bx, int i, int j, int k,
                                             const double qc[6], const double co[3], double dxi, double cv,
                                             double g[6], const WrapArgs& wa = WrapArgs{}) {
  const int ci[3] = {i, j, k};
  if (!WRAP) {
#pragma unroll
    for (int t = 0; t < 3; ++t)
      if (t != D && (ci[t] < bx.lo[t] || ci[t] > bx.hi[t])) return;  // every weight is zero
  }
  const int x = ci[D], lo = bx.lo[D], hi = bx.hi[D];
  const int di = D == 0, dj = D == 1, dk = D == 2;
  auto nb = [&](int o, int& ii, int& jj, int& kk) {
    ii = wr<WRAP>(wa, 0, i + o * di);
    jj = wr<WRAP>(wa, 1, j + o * dj);
    kk = wr<WRAP>(wa, 2, k + o * dk);
  };
  double lc[5];
  if (WRAP) {
#pragma unroll
    for (int n = 0; n < NCONS; ++n) lc[n] = L(i, j, k, n);
  } else {
    lam_at(L, bx, i, j, k, lc);
  }
#pragma unroll
  for (int l = 1; l <= H; ++l) {
    const double cw = co[l - 1] * dxi;
    if (WRAP ? wa.on != 0 : (x >= lo && x <= hi + l)) {  // cell 1 of the pair (c, c-l)
      double lm[5], w[5], qm[6];
      int ii, jj, kk;
      nb(-l, ii, jj, kk);
      if (WRAP) {
#pragma unroll
        for (int n = 0; n < NCONS; ++n) lm[n] = L(ii, jj, kk, n);
      } else {
        lam_at(L, bx, ii, jj, kk, lm);
      }
#pragma unroll
      for (int n = 0; n < NCONS; ++n) w[n] = cw * (lc[n] - lm[n]);
      load_q(q, ii, jj, kk, qm);
      keep_pair_vjp<D, true, false>(w, qc, qm, cv, g, nullptr);
    }
    if (WRAP ? wa.on != 0 : (x >= lo - l && x <= hi)) {  // cell 2 of the pair (c+l, c)
      double lp[5], w[5], qp[6];
      int ii, jj, kk;
      nb(l, ii, jj, kk);
      if (WRAP) {
#pragma unroll
        for (int n = 0; n < NCONS; ++n) lp[n] = L(ii, jj, kk, n);
      } else {
        lam_at(L, bx, ii, jj, kk, lp);
      }
#pragma unroll
      for (int n = 0; n < NCONS; ++n) w[n] = cw * (lp[n] - lc[n]);
      load_q(q, ii, jj, kk, qp);
      keep_pair_vjp<D, false, true>(w, qp, qc, cv, nullptr, g);
    }
  }
}

// Viscous pre-pass on grow(b,1): the adjoint of the cell fluxes pf_d(j) is
// 0.5 dxi_d (lam(j+e_d) - lam(j-e_d)); store the adjoints of the central
// differences (times their 0.5 dxi factor) and of the cell's own u, v, w, T.
template <bool WRAP>
__global__ void __launch_bounds__(BLK) k_adj_visc(const BoxDev* boxes, const BlockRef* bmap, Region reg,
                                                  const double* box_dt, const Fab* pr, const Fab* lam,
                                                  const Fab* ew, Closures c, WrapArgs wa) {
  REGION_LOOP_BEGIN
  const double dx0 = bx.dxi[0], dx1 = bx.dxi[1], dx2 = bx.dxi[2];
  const Fab q = pr[b], L = lam[b], E = ew[b];
  const int ip = wr<WRAP>(wa, 0, i + 1), im = wr<WRAP>(wa, 0, i - 1);
  const int jp = wr<WRAP>(wa, 1, j + 1), jm = wr<WRAP>(wa, 1, j - 1);
  const int kp = wr<WRAP>(wa, 2, k + 1), km = wr<WRAP>(wa, 2, k - 1);
  auto lget = [&](int ii, int jj, int kk, double v[5]) {
    if (WRAP) {
#pragma unroll
      for (int n = 0; n < NCONS; ++n) v[n] = L(ii, jj, kk, n);
    } else {
      lam_at(L, bx, ii, jj, kk, v);
    }
  };
  double a[5], bb[5], cc[5], lp[5], lm[5];
  lget(ip, j, k, lp);
  lget(im, j, k, lm);
#pragma unroll
  for (int n = 0; n < NCONS; ++n) a[n] = 0.5 * dx0 * (lp[n] - lm[n]);
  lget(i, jp, k, lp);
  lget(i, jm, k, lm);
#pragma unroll
  for (int n = 0; n < NCONS; ++n) bb[n] = 0.5 * dx1 * (lp[n] - lm[n]);
  lget(i, j, kp, lp);
  lget(i, j, km, lm);
#pragma unroll
  for (int n = 0; n < NCONS; ++n) cc[n] = 0.5 * dx2 * (lp[n] - lm[n]);
  // forward values at the cell (as in the viscous flux kernel)
  const double ux = q(i, j, k, QU), uy = q(i, j, k, QV), uz = q(i, j, k, QW), T = q(i, j, k, QT);
  const double dudx = (q(ip, j, k, QU) - q(im, j, k, QU)) * 0.5 * dx0;
  const double dvdx = (q(ip, j, k, QV) - q(im, j, k, QV)) * 0.5 * dx0;
  const double dwdx = (q(ip, j, k, QW) - q(im, j, k, QW)) * 0.5 * dx0;
  const double dTdx = (q(ip, j, k, QT) - q(im, j, k, QT)) * 0.5 * dx0;
  const double dudy = (q(i, jp, k, QU) - q(i, jm, k, QU)) * 0.5 * dx1;
  const double dvdy = (q(i, jp, k, QV) - q(i, jm, k, QV)) * 0.5 * dx1;
  const double dwdy = (q(i, jp, k, QW) - q(i, jm, k, QW)) * 0.5 * dx1;
  const double dTdy = (q(i, jp, k, QT) - q(i, jm, k, QT)) * 0.5 * dx1;
  const double dudz = (q(i, j, kp, QU) - q(i, j, km, QU)) * 0.5 * dx2;
  const double dvdz = (q(i, j, kp, QV) - q(i, j, km, QV)) * 0.5 * dx2;
  const double dwdz = (q(i, j, kp, QW) - q(i, j, km, QW)) * 0.5 * dx2;
  const double dTdz = (q(i, j, kp, QT) - q(i, j, km, QT)) * 0.5 * dx2;
  const double div = dudx + dvdy + dwdz;
  const double mu = visc(c, T);
  const double kk = cond(c, T, mu);
  const double r1_3 = 1.0 / 3.0;
  const double sxx = 2.0 * (dudx - r1_3 * div), syy = 2.0 * (dvdy - r1_3 * div), szz = 2.0 * (dwdz - r1_3 * div);
  const double sxy = dudy + dvdx, sxz = dudz + dwdx, syz = dvdz + dwdy;
  const double txx = mu * sxx, tyy = mu * syy, tzz = mu * szz, txy = mu * sxy, txz = mu * sxz, tyz = mu * syz;
  // adjoints of the stresses
  const double gxx = -a[UMX] - a[UET] * ux;
  const double gxy = -a[UMY] - a[UET] * uy - bb[UMX] - bb[UET] * ux;
  const double gxz = -a[UMZ] - a[UET] * uz - cc[UMX] - cc[UET] * ux;
  const double gyy = -bb[UMY] - bb[UET] * uy;
  const double gyz = -bb[UMZ] - bb[UET] * uz - cc[UMY] - cc[UET] * uy;
  const double gzz = -cc[UMZ] - cc[UET] * uz;
  const double s3 = (gxx + gyy + gzz) * r1_3;
  const double Dux = 2.0 * mu * (gxx - s3), Dvy = 2.0 * mu * (gyy - s3), Dwz = 2.0 * mu * (gzz - s3);
  const double Dxy = mu * gxy, Dxz = mu * gxz, Dyz = mu * gyz;
  // own-value adjoints: u, v, w through the work terms; T through mu and k
  const double Lu = -a[UET] * txx - bb[UET] * txy - cc[UET] * txz;
  const double Lv = -a[UET] * txy - bb[UET] * tyy - cc[UET] * tyz;
  const double Lw = -a[UET] * txz - bb[UET] * tyz - cc[UET] * tzz;
  const double gmu = gxx * sxx + gyy * syy + gzz * szz + gxy * sxy + gxz * sxz + gyz * syz;
  const double gk = -(a[UET] * dTdx + bb[UET] * dTdy + cc[UET] * dTdz);
  double dmu, dk;
  transport_derivs(c, T, mu, kk, dmu, dk);
  const double LT = gmu * dmu + gk * dk;
  const double hx = 0.5 * dx0, hy = 0.5 * dx1, hz = 0.5 * dx2;
  E(i, j, k, 0) = Lu;
  E(i, j, k, 1) = Lv;
  E(i, j, k, 2) = Lw;
  E(i, j, k, 3) = LT;
  E(i, j, k, 4) = hx * Dux;
  E(i, j, k, 5) = hx * Dxy;
  E(i, j, k, 6) = hx * Dxz;
  E(i, j, k, 7) = hx * (-kk * a[UET]);
  E(i, j, k, 8) = hy * Dxy;
  E(i, j, k, 9) = hy * Dvy;
  E(i, j, k, 10) = hy * Dyz;
  E(i, j, k, 11) = hy * (-kk * bb[UET]);
  E(i, j, k, 12) = hz * Dxz;
  E(i, j, k, 13) = hz * Dyz;
  E(i, j, k, 14) = hz * Dwz;
  E(i, j, k, 15) = hz * (-kk * cc[UET]);
}

// Reverse of the RHS plus the stage combination, one cell of grow(b,3) per
// thread (global-memory gathers; the neighbourhoods are cache resident).
// On valid cells the stage transpose is applied at once
//   mode 0: lam_sin (=|+=) b (c lam_out + d G),  mode 1: c lam_out + d G,
//   mode 2: G;   lam_s1 (+=|=) a lam_out
// where G = (d dSdt / d S)^T lam_out; on ghost cells G itself is stored for
// the transposed ghost fill.
// With WRAP the arrays are domain-wide and every cell of the region is valid.
template <int H, bool EULER, bool VISC, bool WRAP>
__global__ void __launch_bounds__(BLK) k_adj_cell(const BoxDev* boxes, const BlockRef* bmap, Region reg,
                                                  const double* box_dt, const Fab* pr, const Fab* lam,
                                                  const Fab* ew, const Fab* gw, Closures c, double co0,
                                                  double co1, double co2, StageArgs sa, const Fab* lam_s1,
                                                  const Fab* lam_sin, int accumulate, int s1_mode, WrapArgs wa) {
  REGION_LOOP_BEGIN
  const Fab q = pr[b], L = lam[b];
  double qc[6];
  load_q(q, i, j, k, qc);
  double g[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (EULER) {
    const double co[3] = {co0, co1, co2};
    adj_keep_dir<H, 0, WRAP>(q, L, bx, i, j, k, qc, co, bx.dxi[0], c.cv, g, wa);
    adj_keep_dir<H, 1, WRAP>(q, L, bx, i, j, k, qc, co, bx.dxi[1], c.cv, g, wa);
    adj_keep_dir<H, 2, WRAP>(q, L, bx, i, j, k, qc, co, bx.dxi[2], c.cv, g, wa);
  }
  if (VISC) {
    const Fab E = ew[b];
    if (WRAP || in_box(bx, 1, i, j, k)) {
#pragma unroll
      for (int m = 0; m < 4; ++m) g[QU + m] += E(i, j, k, m);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int di = d == 0, dj = d == 1, dk = d == 2;
      const int im = wr<WRAP>(wa, 0, i - di), jm = wr<WRAP>(wa, 1, j - dj), km = wr<WRAP>(wa, 2, k - dk);
      const int ip = wr<WRAP>(wa, 0, i + di), jp = wr<WRAP>(wa, 1, j + dj), kp = wr<WRAP>(wa, 2, k + dk);
      if (WRAP || in_box(bx, 1, im, jm, km)) {
#pragma unroll
        for (int m = 0; m < 4; ++m) g[QU + m] += E(im, jm, km, 4 + 4 * d + m);
      }
      if (WRAP || in_box(bx, 1, ip, jp, kp)) {
#pragma unroll
        for (int m = 0; m < 4; ++m) g[QU + m] -= E(ip, jp, kp, 4 + 4 * d + m);
      }
    }
  }
  // cons2prim transposed: T = p/(rho R), p = (gamma-1)(E - |m|^2/(2 rho)), v = m/rho
  const double rinv = 1.0 / qc[QRHO];
  const double ux = qc[QU], uy = qc[QV], uz = qc[QW];
  const double gp = g[QPRES] + g[QT] * rinv / c.Rspec;
  const double gr = g[QRHO] - g[QT] * qc[QT] * rinv;
  double G[5];
  G[URHO] = gr + c.gm1 * 0.5 * (ux * ux + uy * uy + uz * uz) * gp - (g[QU] * ux + g[QV] * uy + g[QW] * uz) * rinv;
  G[UMX] = g[QU] * rinv - c.gm1 * ux * gp;
  G[UMY] = g[QV] * rinv - c.gm1 * uy * gp;
  G[UMZ] = g[QW] * rinv - c.gm1 * uz * gp;
  G[UET] = c.gm1 * gp;
  if (!WRAP && !in_box(bx, 0, i, j, k)) {
    const Fab Gh = gw[b];
#pragma unroll
    for (int n = 0; n < NCONS; ++n) Gh(i, j, k, n) = G[n];
    return;
  }
  const StageCoef sc = box_stage(sa, b);
  const Fab Li = lam_sin[b];
  double lo[5];
#pragma unroll
  for (int n = 0; n < NCONS; ++n) lo[n] = L(i, j, k, n);
#pragma unroll
  for (int n = 0; n < NCONS; ++n) {
    double v;
    if (sc.mode == 0) v = sc.b * (sc.c * lo[n] + sc.d * G[n]);
    else if (sc.mode == 1) v = sc.c * lo[n] + sc.d * G[n];
    else v = G[n];
    if (accumulate) v = Li(i, j, k, n) + v;
    Li(i, j, k, n) = v;
  }
  if (s1_mode != 0 && lam_s1 != nullptr) {
    const Fab S = lam_s1[b];
    const double a = sc.mode == 2 ? 0.0 : sc.a;
#pragma unroll
    for (int n = 0; n < NCONS; ++n) {
      if (s1_mode == 2) S(i, j, k, n) = a * lo[n];
      else if (a != 0.0) S(i, j, k, n) = S(i, j, k, n) + a * lo[n];
    }
  }
}

constexpr int FB = 256;

template <class T>
__device__ __forceinline__ int find_task(const T* tasks, int ntask, int64_t t) {
  int lo = 0, hi = ntask - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].start <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Weight of G in lam_sin for a box's stage (b d, d or 1).
__device__ __forceinline__ double g_weight(const StageArgs& sa, int box) {
  const StageCoef sc = box_stage(sa, box);
  return sc.mode == 0 ? sc.b * sc.d : (sc.mode == 1 ? sc.d : 1.0);
}

// Transposed physical-boundary fill of one direction and side (z, y, x in
// turn, the reverse of the fill's order): one lane per line along the
// direction through the task's ghost layers, taken outermost first.  Each
// ghost cell's adjoint moves to its source cell (the boundary cell for
// foextrap, the mirror cell for reflections, negated for odd ones) -- into
// the stage adjoint, weighted, when that cell is valid, else into the ghost
// adjoint of another direction -- and is cleared.  A line's sources lie on
// the line inside the box, so no two lanes (or boxes) share one: plain adds
// in a fixed order, the same bits on every run.
__global__ void __launch_bounds__(FB) k_fold_bc_lines(const BcTask* tasks, int ntask, int64_t nlines,
                                                      const Fab* gfabs, const BoxDev* boxes, StageArgs sa,
                                                      const Fab* lam_sin) {
  int64_t t = int64_t(blockIdx.x) * FB + threadIdx.x;
  if (t >= nlines) return;
  const BcTask& T = tasks[find_task(tasks, ntask, t)];
  if (!box_active(sa, T.box)) return;
  const int d = T.dir, e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
  const int64_t r = t - T.start;
  const int n1 = T.hi[e1] - T.lo[e1] + 1;
  int c[3];
  c[e1] = T.lo[e1] + int(r % n1);
  c[e2] = T.lo[e2] + int(r / n1);
  const Fab f = gfabs[T.box], Li = lam_sin[T.box];
  const BoxDev bx = boxes[T.box];
  const double w = g_weight(sa, T.box);
  const int first = T.side == 0 ? T.lo[d] : T.hi[d], step = T.side == 0 ? 1 : -1;
  for (int cd = first; T.side == 0 ? cd <= T.hi[d] : cd >= T.lo[d]; cd += step) {
    c[d] = cd;
    for (int n = 0; n < NCONS; ++n) {
      const int ty = T.type[n];
      if (ty == 0) continue;
      int sc[3] = {c[0], c[1], c[2]};
      if (ty == 2) sc[d] = T.bnd;
      else sc[d] = T.side == 0 ? 2 * T.bnd - 1 - cd : 2 * T.bnd + 1 - cd;
      double v = f(c[0], c[1], c[2], n);
      f(c[0], c[1], c[2], n) = 0.0;
      if (ty == 4) v = -v;
      if (in_box(bx, 0, sc[0], sc[1], sc[2])) Li(sc[0], sc[1], sc[2], n) += w * v;
      else f(sc[0], sc[1], sc[2], n) += v;
    }
  }
}

// Transposed same-level / periodic copies of one colour class: the ghost
// adjoint of `dst` added, weighted, to lam_sin at the valid source cell of
// `src`.  Within a class no two tasks share a source cell, so the adds need
// no atomics; the classes run one after another in a fixed order.
__global__ void __launch_bounds__(FB) k_fold_copy_class(const CopyTask* tasks, int ntask, int64_t ncell,
                                                        const Fab* gfabs, StageArgs sa, const Fab* lam_sin) {
  int64_t t = int64_t(blockIdx.x) * FB + threadIdx.x;
  if (t >= ncell) return;
  const CopyTask& T = tasks[find_task(tasks, ntask, t)];
  if (!box_active(sa, T.dst)) return;
  int64_t r = t - T.start;
  int nx = T.hi[0] - T.lo[0] + 1, ny = T.hi[1] - T.lo[1] + 1;
  int i = T.lo[0] + int(r % nx);
  int64_t r2 = r / nx;
  int j = T.lo[1] + int(r2 % ny);
  int k = T.lo[2] + int(r2 / ny);
  const Fab d = gfabs[T.dst], s = lam_sin[T.src];
  const double w = g_weight(sa, T.src);
#pragma unroll
  for (int n = 0; n < NCONS; ++n) s(i - T.sh[0], j - T.sh[1], k - T.sh[2], n) += w * d(i, j, k, n);
}

constexpr int OB = 256;

// Objective over the valid cells and its gradient:
//   kind 0: J = sum |m|^2 / (2 rho) dV  (kinetic energy; dV per box)
//   kind 1: J = sum w . U
__global__ void __launch_bounds__(OB) k_objective(const BoxDev* boxes, int nbox, int kind, const Fab* st,
                                                  const Fab* wts, const Fab* lam, double* partial) {
  __shared__ double sm[OB];
  double acc = 0.0;
  for (int b = 0; b < nbox; ++b) {
    const BoxDev bx = boxes[b];
    const Fab u = st[b], L = lam[b];
    const double dV = bx.dx[0] * bx.dx[1] * bx.dx[2];
    int nx = bx.hi[0] - bx.lo[0] + 1, ny = bx.hi[1] - bx.lo[1] + 1, nz = bx.hi[2] - bx.lo[2] + 1;
    int64_t n = int64_t(nx) * ny * nz;
    for (int64_t t = int64_t(blockIdx.x) * OB + threadIdx.x; t < n; t += int64_t(gridDim.x) * OB) {
      int i = bx.lo[0] + int(t % nx);
      int64_t r = t / nx;
      int j = bx.lo[1] + int(r % ny);
      int k = bx.lo[2] + int(r / ny);
      if (kind == 0) {
        const double rho = u(i, j, k, URHO), mx = u(i, j, k, UMX), my = u(i, j, k, UMY), mz = u(i, j, k, UMZ);
        const double rinv = 1.0 / rho;
        const double m2 = mx * mx + my * my + mz * mz;
        acc += 0.5 * m2 * rinv * dV;
        L(i, j, k, URHO) = -0.5 * m2 * rinv * rinv * dV;
        L(i, j, k, UMX) = mx * rinv * dV;
        L(i, j, k, UMY) = my * rinv * dV;
        L(i, j, k, UMZ) = mz * rinv * dV;
        L(i, j, k, UET) = 0.0;
      } else {
        const Fab w = wts[b];
#pragma unroll
        for (int c = 0; c < NCONS; ++c) {
          const double wc = w(i, j, k, c);
          acc += wc * u(i, j, k, c);
          L(i, j, k, c) = wc;
        }
      }
    }
  }
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = OB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sm[0];
}

__global__ void __launch_bounds__(OB) k_sum1(const double* partial, int n, double* result) {
  __shared__ double sm[OB];
  double acc = 0.0;
  for (int t = threadIdx.x; t < n; t += OB) acc += partial[t];
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = OB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) result[0] = sm[0];
}

}  // namespace

// (a source added to dSdt does not depend on the state and drops out)
bool adjoint_supported(const amr_scheme_t& s) {
  if (s.rhs_euler && (s.flux_euler != 1 || s.art_diss)) return false;
  return s.rhs_euler || s.rhs_visc;
}

hipError_t launch_adj_stage(const BoxTable& bt, const UnfusedWork& w, const AdjWork& aw, const Closures& c,
                            const amr_scheme_t& s, const StageArgs& sa, const WrapArgs& wa, const Fab* sin, const Fab* lam_out, const Fab* lam_s1, const Fab* lam_sin,
                            bool accumulate, int s1_mode, const FoldTasks& ft, hipStream_t st) {
  const BoxDev* B = bt.d_boxes;
  double co[3] = {0.0, 0.0, 0.0};
  keep_coefficients(s.order_keep, co);
  const bool eu = s.rhs_euler != 0, vi = s.rhs_visc != 0;
  const int h = s.order_keep / 2;
  if (wa.on && aw.lq) {
    // periodic domain-wide level: the transpose of the periodic stencil,
    // gathered on the valid cells; no ghost outputs, no transposed fill
    const Region g0 = grown(0);
    Launch L0 = grid_for(bt, g0);
    if (!L0.map) return hipErrorNotInitialized;
    UnfusedWork lw = w;
    lw.prims = aw.lq;
    hipError_t e = launch_prims_region(bt, lw, c, sa.box_dt, sin, g0, st);
    if (e != hipSuccess) return e;
    if (vi)
      hipLaunchKernelGGL(k_adj_visc<true>, L0.grid, dim3(BLK), 0, st, B, L0.map, g0, sa.box_dt, aw.lq, lam_out,
                         aw.le, c, wa);
#define CELLW(HH, EE, VV)                                                                                      \
  hipLaunchKernelGGL((k_adj_cell<HH, EE, VV, true>), L0.grid, dim3(BLK), 0, st, B, L0.map, g0, sa.box_dt,     \
                     aw.lq, lam_out, aw.le, aw.g, c, co[0], co[1], co[2], sa, lam_s1, lam_sin, accumulate ? 1 : 0, \
                     s1_mode, wa)
    if (!eu) CELLW(1, false, true);
    else if (h == 3) { if (vi) CELLW(3, true, true); else CELLW(3, true, false); }
    else if (h == 2) { if (vi) CELLW(2, true, true); else CELLW(2, true, false); }
    else { if (vi) CELLW(1, true, true); else CELLW(1, true, false); }
#undef CELLW
    return hipGetLastError();
  }
  const WrapArgs nowrap{0, {0, 0, 0}, {1, 1, 1}};
  hipError_t e = launch_prims(bt, w, c, sa.box_dt, sin, wa, st);
  if (e != hipSuccess) return e;
  if (s.rhs_visc) {
    const Region g1 = grown(1);
    Launch L = grid_for(bt, g1);
    if (!L.map) return hipErrorNotInitialized;
    hipLaunchKernelGGL(k_adj_visc<false>, L.grid, dim3(BLK), 0, st, B, L.map, g1, sa.box_dt, w.prims, lam_out,
                       aw.e, c, nowrap);
  }
  const Region g3 = grown(NGHOST);
  Launch L3 = grid_for(bt, g3);
  if (!L3.map) return hipErrorNotInitialized;
#define CELL(HH, EE, VV)                                                                                     \
  hipLaunchKernelGGL((k_adj_cell<HH, EE, VV, false>), L3.grid, dim3(BLK), 0, st, B, L3.map, g3, sa.box_dt,   \
                     w.prims, lam_out, aw.e, aw.g, c, co[0], co[1], co[2], sa, lam_s1, lam_sin, accumulate ? 1 : 0, \
                     s1_mode, nowrap)
  if (!eu) CELL(1, false, true);
  else if (h == 3) { if (vi) CELL(3, true, true); else CELL(3, true, false); }
  else if (h == 2) { if (vi) CELL(2, true, true); else CELL(2, true, false); }
  else { if (vi) CELL(1, true, true); else CELL(1, true, false); }
#undef CELL
  for (int d = 2; d >= 0; --d)
    for (int side = 0; side < 2; ++side)
      if (ft.nlines[d][side] > 0)
        hipLaunchKernelGGL(k_fold_bc_lines, dim3(unsigned((ft.nlines[d][side] + FB - 1) / FB)), dim3(FB), 0, st,
                           ft.bc[d][side], ft.nbc[d][side], ft.nlines[d][side], aw.g, B, sa, lam_sin);
  for (int k = 0; k < ft.nclass; ++k)
    hipLaunchKernelGGL(k_fold_copy_class, dim3(unsigned((ft.class_cells[k] + FB - 1) / FB)), dim3(FB), 0, st,
                       ft.classes[k], ft.class_tasks[k], ft.class_cells[k], aw.g, sa, lam_sin);
  return hipGetLastError();
}

hipError_t launch_objective(const BoxTable& bt, int kind, const Fab* st, const Fab* wts, const Fab* lam,
                            double* partial, double* result, hipStream_t s) {
  const int nb = reduce_blocks(bt);
  hipLaunchKernelGGL(k_objective, dim3(nb), dim3(OB), 0, s, bt.d_boxes, bt.nbox, kind, st, wts, lam, partial);
  hipLaunchKernelGGL(k_sum1, dim3(1), dim3(OB), 0, s, partial, nb, result);
  return hipGetLastError();
}

}  // namespace amr
