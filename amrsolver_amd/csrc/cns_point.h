// Device point functions of the CNS right-hand side (FP64).
//
// Each function restates one reference point kernel in the reference's
// operation order; the build compiles device code with -ffp-contract=off so
// results round exactly as the reference's CPU build does.  max/min follow
// std::max/std::min semantics ((a<b)?b:a).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "amrsolver_hip.h"

namespace amr {

constexpr int URHO = 0, UMX = 1, UMY = 2, UMZ = 3, UET = 4, NCONS = 5;
constexpr int QRHO = 0, QU = 1, QV = 2, QW = 3, QT = 4, QPRES = 5, NPRIM = 6;
constexpr int NGHOST = 3;

__host__ __device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }
__host__ __device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }

// Array view with the AMReX Array4 index convention.  `off` folds the lower
// corner of the allocation into one offset so that index = off + i + j*js +
// k*ks + n*ns.
struct Fab {
  double* p;
  int64_t off, js, ks, ns;
  __host__ __device__ __forceinline__ double& operator()(int i, int j, int k, int n) const {
    return p[off + i + j * js + k * ks + n * ns];
  }
  __host__ __device__ __forceinline__ int64_t idx(int i, int j, int k) const {
    return off + i + j * js + k * ks;
  }
};

inline Fab make_fab(const amr_fab_t& f) {
  Fab r;
  r.p = f.data;
  r.js = f.jstride;
  r.ks = f.kstride;
  r.ns = f.nstride;
  r.off = -(int64_t(f.lo[0] - f.ng) + int64_t(f.lo[1] - f.ng) * f.jstride +
            int64_t(f.lo[2] - f.ng) * f.kstride);
  return r;
}

// Closures with the law switches resolved into plain fields.
struct Closures {
  double gamma, gm1, cv, cp, Rspec;
  int visc_law, cond_law;
  double visc_ref, Tvisc_ref, cond_ref, Tcond_ref, Pr;
  double Cshock, Cdamp;  // JST coefficients
};

inline Closures make_closures(const amr_closures_t& c) {
  Closures r;
  r.gamma = c.gamma;
  r.gm1 = c.gamma - 1.0;
  r.cv = c.cv;
  r.cp = c.cp;
  r.Rspec = c.Rspec;
  r.visc_law = c.visc_law;
  r.cond_law = c.cond_law;
  r.visc_ref = c.visc_ref;
  r.Tvisc_ref = c.Tvisc_ref;
  r.cond_ref = c.cond_ref;
  r.Tcond_ref = c.Tcond_ref;
  r.Pr = c.Pr;
  r.Cshock = c.Cshock;
  r.Cdamp = c.Cdamp;
  return r;
}

// src/cls/Closures.h:108-109, :118-119
__device__ __forceinline__ double visc(const Closures& c, double T) {
  if (c.visc_law == 0) return c.visc_ref;
  return c.visc_ref * T * sqrt(T) / (c.Tvisc_ref + T);
}
// src/cls/Closures.h:135-136, :161; law 2 = mu*cp/Pr (mu passed in)
__device__ __forceinline__ double cond(const Closures& c, double T, double mu) {
  if (c.cond_law == 0) return c.cond_ref;
  if (c.cond_law == 1) return c.cond_ref * T * sqrt(T) / (c.Tcond_ref + T);
  return mu * c.cp / c.Pr;
}

// src/rhs/CNS_hydro_K.h:15-34.  q[] = rho, u, v, w, T, p
__device__ __forceinline__ void cons2prim(const double u0, const double u1, const double u2,
                                          const double u3, const double u4, const Closures& c,
                                          double q[6]) {
  double rho = dmax(1e-40, u0);
  double rhoinv = 1.0 / rho;
  double ux = u1 * rhoinv;
  double uy = u2 * rhoinv;
  double uz = u3 * rhoinv;
  double rhoke = 0.5 * rho * (ux * ux + uy * uy + uz * uz);
  double rhoei = (u4 - rhoke);
  double p = c.gm1 * rhoei;
  q[QRHO] = rho;
  q[QU] = ux;
  q[QV] = uy;
  q[QW] = uz;
  q[QPRES] = p;
  q[QT] = p / (rho * c.Rspec);
}

// src/rhs/CentralKEEP.h:19-36
__device__ __forceinline__ double fDiv(double f, double fl) { return 0.5 * (f + fl); }
__device__ __forceinline__ double fgDiv(double f, double fl, double g, double gl) {
  return 0.5 * (f * gl + fl * g);
}
__device__ __forceinline__ double fgQuad(double f, double fl, double g, double gl) {
  return 0.25 * (f + fl) * (g + gl);
}
__device__ __forceinline__ double fghCubic(double f, double fl, double g, double gl, double h,
                                           double hl) {
  return 0.1250 * (f + fl) * (g + gl) * (h + hl);
}

// The five flux terms of one (l,m) pair of the KEEP flux in direction `nrm`
// (0,1,2) before the coefficient (CentralKEEP.h:58-84 and the y/z copies).
// Cell 1 = (rho1,u1,v1,w1,ie1,p1), cell 2 likewise.
template <int NRM>
__device__ __forceinline__ void keep_terms(double rho1, double ux1, double uy1, double uz1, double ie1,
                                           double p1, double rho2, double ux2, double uy2, double uz2,
                                           double ie2, double p2, double t[5]) {
  const double un1 = NRM == 0 ? ux1 : (NRM == 1 ? uy1 : uz1);
  const double un2 = NRM == 0 ? ux2 : (NRM == 1 ? uy2 : uz2);
  double massflx = fgQuad(rho1, rho2, un1, un2);
  double ke = 0.5 * (ux1 * ux2 + uy1 * uy2 + uz1 * uz2);
  t[URHO] = massflx;
  if (NRM == 0)
    t[UMX] = fghCubic(rho1, rho2, un1, un2, un1, un2) + fDiv(p1, p2);
  else
    t[UMX] = fghCubic(rho1, rho2, un1, un2, ux1, ux2);
  if (NRM == 1)
    t[UMY] = fghCubic(rho1, rho2, un1, un2, un1, un2) + fDiv(p1, p2);
  else
    t[UMY] = fghCubic(rho1, rho2, un1, un2, uy1, uy2);
  if (NRM == 2)
    t[UMZ] = fghCubic(rho1, rho2, un1, un2, un1, un2) + fDiv(p1, p2);
  else
    t[UMZ] = fghCubic(rho1, rho2, un1, un2, uz1, uz2);
  t[UET] = massflx * ke + fghCubic(rho1, rho2, un1, un2, ie1, ie2) + fgDiv(p1, p2, un1, un2);
}

// One pair accumulated into f[5] with coefficient co: f += co * terms, in
// the reference's component order.
__device__ __forceinline__ void keep_accum(double co, const double t[5], double f[5]) {
#pragma unroll
  for (int n = 0; n < 5; ++n) f[n] += co * t[n];
}

template <int NRM>
__device__ __forceinline__ void keep_pair(double co, double rho1, double ux1, double uy1,
                                          double uz1, double ie1, double p1, double rho2,
                                          double ux2, double uy2, double uz2, double ie2,
                                          double p2, double f[5]) {
  double t[5];
  keep_terms<NRM>(rho1, ux1, uy1, uz1, ie1, p1, rho2, ux2, uy2, uz2, ie2, p2, t);
  keep_accum(co, t, f);
}

// src/rhs/High_resolution.h:124-159 (beta1 with the 4th power of (s1-s3),
// eps = 1e-40; pow(x,2) written as x*x)
__device__ __forceinline__ double weno5js(double s0, double s1, double s2, double s3, double s4) {
  const double eps = 1e-40;
  double b0 = (13.0 / 12.0) * (s0 - 2.0 * s1 + s2) * (s0 - 2.0 * s1 + s2) +
              0.25 * (s0 - 4.0 * s1 + 3.0 * s2) * (s0 - 4.0 * s1 + 3.0 * s2);
  double b1 = (13.0 / 12.0) * (s1 - 2.0 * s2 + s3) * (s1 - 2.0 * s2 + s3) +
              0.25 * (s1 - s3) * (s1 - s3) * (s1 - s3) * (s1 - s3);
  double b2 = (13.0 / 12.0) * (s2 - 2.0 * s3 + s4) * (s2 - 2.0 * s3 + s4) +
              0.25 * (3.0 * s2 - 4.0 * s3 + s4) * (3.0 * s2 - 4.0 * s3 + s4);
  double t0 = eps + b0, t1 = eps + b1, t2 = eps + b2;
  double a0 = 0.1 / (t0 * t0);
  double a1 = 0.6 / (t1 * t1);
  double a2 = 0.3 / (t2 * t2);
  double sum = 1.0 / (a0 + a1 + a2);
  double v0 = 2.0 * s0 - 7.0 * s1 + 11.0 * s2;
  double v1 = -s1 + 5.0 * s2 + 2.0 * s3;
  double v2 = 2.0 * s2 + 5.0 * s3 - s4;
  return (1.0 / 6.0) * sum * (a0 * v0 + a1 * v1 + a2 * v2);
}

// src/rhs/Riemann.h:9-17
__device__ __forceinline__ double limiter(double dlft, double drgt) {
  double dcen = 0.5 * (dlft + drgt);
  double dsgn = copysign(1.0, dcen);
  double slop = 2.0 * dmin(fabs(dlft), fabs(drgt));
  double dlim = (dlft * drgt >= 0.0) ? slop : 0.0;
  return dsgn * dmin(dlim, fabs(dcen));
}

// src/rhs/Riemann.h:123-249; out = (rho, normal mom, t1 mom, t2 mom, energy) flux
__device__ __forceinline__ void riemann_prob(double gamma, double smallp, double rl, double ul,
                                             double pl, double ut1l, double ut2l, double rr,
                                             double ur, double pr, double ut1r, double ut2r,
                                             double out[5]) {
  const double weakwv = 1.e-3;
  const double small = 1.e-6;
  double clsql = gamma * pl * rl;
  double clsqr = gamma * pr * rr;
  double wl = sqrt(clsql);
  double wr = sqrt(clsqr);
  double cleft = wl / rl;
  double cright = wr / rr;
  double ccsmall = small * (cleft + cright);
  double pstar = (wl * pr + wr * pl - wr * wl * (ur - ul)) / (wl + wr);
  pstar = dmax(pstar, smallp);
  double pstnm1 = pstar;
  double wlsq = (0.5 * (gamma - 1.) * (pstar + pl) + pstar) * rl;
  double wrsq = (0.5 * (gamma - 1.) * (pstar + pr) + pstar) * rr;
  wl = sqrt(wlsq);
  wr = sqrt(wrsq);
  double ustarp = ul - (pstar - pl) / wl;
  double ustarm = ur + (pstar - pr) / wr;
  pstar = (wl * pr + wr * pl - wr * wl * (ur - ul)) / (wl + wr);
  pstar = dmax(pstar, smallp);
  double ustar = 0.0;
#pragma unroll
  for (int iter = 0; iter < 3; ++iter) {
    wlsq = (0.5 * (gamma - 1.) * (pstar + pl) + pstar) * rl;
    wrsq = (0.5 * (gamma - 1.) * (pstar + pr) + pstar) * rr;
    wl = 1. / sqrt(wlsq);
    wr = 1. / sqrt(wrsq);
    double ustnm1 = ustarm;
    double ustnp1 = ustarp;
    ustarm = ur - (pr - pstar) * wr;
    ustarp = ul + (pl - pstar) * wl;
    double dpditer = fabs(pstnm1 - pstar);
    double zp = fabs(ustarp - ustnp1);
    if (zp - weakwv * cleft < 0.0) zp = dpditer * wl;
    double zm = fabs(ustarm - ustnm1);
    if (zm - weakwv * cright < 0.0) zm = dpditer * wr;
    double zz = zp + zm;
    double denom = dpditer / dmax(zz, ccsmall);
    pstnm1 = pstar;
    pstar = pstar - denom * (ustarm - ustarp);
    pstar = dmax(pstar, smallp);
    ustar = 0.5 * (ustarm + ustarp);
  }
  double ro, uo, po, sgnm, utrans1, utrans2;
  if (ustar > 0.) {
    ro = rl; uo = ul; po = pl; sgnm = 1.; utrans1 = ut1l; utrans2 = ut2l;
  } else if (ustar < 0.) {
    ro = rr; uo = ur; po = pr; sgnm = -1.; utrans1 = ut1r; utrans2 = ut2r;
  } else {
    uo = 0.5 * (ur + ul);
    po = 0.5 * (pr + pl);
    ro = 2. * (rl * rr) / (rl + rr);
    sgnm = 1.;
    utrans1 = 0.5 * (ut1l + ut1r);
    utrans2 = 0.5 * (ut2l + ut2r);
  }
  double wosq = (0.5 * (gamma - 1.) * (pstar + po) + pstar) * ro;
  double co = sqrt(gamma * po / ro);
  double wo = sqrt(wosq);
  double dpjmp = pstar - po;
  double rstar = ro / (1. - ro * dpjmp / wosq);
  double cstar = sqrt(gamma * pstar / rstar);
  double spout = co - sgnm * uo;
  double spin = cstar - sgnm * uo;
  if (pstar >= po) {
    spin = wo / ro - sgnm * uo;
    spout = spin;
  }
  double ss = dmax(spout - spin, spout + spin);
  double frac = 0.5 * (1. + (spin + spout) / dmax(ss, ccsmall));
  double rgdnv, ugdnv, pgdnv;
  if (spout < 0.) {
    rgdnv = ro; ugdnv = uo; pgdnv = po;
  } else if (spin >= 0.) {
    rgdnv = rstar; ugdnv = ustar; pgdnv = pstar;
  } else {
    rgdnv = frac * rstar + (1. - frac) * ro;
    ugdnv = frac * ustar + (1. - frac) * uo;
    pgdnv = frac * pstar + (1. - frac) * po;
  }
  out[0] = rgdnv * ugdnv;
  out[1] = rgdnv * ugdnv * ugdnv + pgdnv;
  out[2] = rgdnv * ugdnv * utrans1;
  out[3] = rgdnv * ugdnv * utrans2;
  out[4] = ugdnv * (0.5 * rgdnv * (ugdnv * ugdnv + utrans1 * utrans1 + utrans2 * utrans2) +
                    pgdnv / (gamma - 1.) + pgdnv);
}

// Characteristic slopes of a cell in direction D (src/rhs/Riemann.h:19-120
// cns_slope_{x,y,z}): Q(o, comp) is primitive `comp` of the cell o steps
// along D (o = -1, 0, 1); the x slopes add 2e-40 to the sound speed, y and z
// 1e-40, as the reference does.
template <int D, class QF>
__device__ __forceinline__ void muscl_slopes(const QF& Q, const Closures& c, double dq[5]) {
  constexpr int qn = QU + D;
  constexpr int qt1 = D == 0 ? QV : QU;
  constexpr int qt2 = D == 2 ? QV : QW;
  const double add = D == 0 ? 2.e-40 : 1.e-40;
  double cspeed = sqrt(c.gamma * c.Rspec * Q(0, QT)) + add;
  double dlft = 0.5 * (Q(0, QPRES) - Q(-1, QPRES)) / cspeed - 0.5 * Q(0, QRHO) * (Q(0, qn) - Q(-1, qn));
  double drgt = 0.5 * (Q(1, QPRES) - Q(0, QPRES)) / cspeed - 0.5 * Q(0, QRHO) * (Q(1, qn) - Q(0, qn));
  dq[0] = limiter(dlft, drgt);
  double cs2 = cspeed * cspeed;
  dlft = (Q(0, QRHO) - Q(-1, QRHO)) - (Q(0, QPRES) - Q(-1, QPRES)) / cs2;
  drgt = (Q(1, QRHO) - Q(0, QRHO)) - (Q(1, QPRES) - Q(0, QPRES)) / cs2;
  dq[1] = limiter(dlft, drgt);
  dlft = 0.5 * (Q(0, QPRES) - Q(-1, QPRES)) / cspeed + 0.5 * Q(0, QRHO) * (Q(0, qn) - Q(-1, qn));
  drgt = 0.5 * (Q(1, QPRES) - Q(0, QPRES)) / cspeed + 0.5 * Q(0, QRHO) * (Q(1, qn) - Q(0, qn));
  dq[2] = limiter(dlft, drgt);
  dq[3] = limiter(Q(0, qt1) - Q(-1, qt1), Q(1, qt1) - Q(0, qt1));
  dq[4] = limiter(Q(0, qt2) - Q(-1, qt2), Q(1, qt2) - Q(0, qt2));
}

// Face flux in direction D from the left cell's primitives ql and slopes dql,
// the right cell's qr and dqr, and Tl, the temperature the left state's
// sound speed is taken from (the left cell's, except in z, where the
// reference reads T at k+1: Riemann.h:316); src/rhs/Riemann.h:251-337
// cns_riemann_{x,y,z}.  f[] is in conserved order (the transverse momentum
// components permuted back per direction).
template <int D, class QL, class DL, class QR, class DR>
__device__ __forceinline__ void muscl_face(const QL& ql, const DL& dql, const QR& qr, const DR& dqr, double Tl,
                                           const Closures& c, double f[5]) {
  constexpr int qn = QU + D;
  constexpr int qt1 = D == 0 ? QV : QU;
  constexpr int qt2 = D == 2 ? QV : QW;
  double cspeed = sqrt(c.gamma * c.Rspec * Tl) + 1.e-40;
  double rl = ql(QRHO) + 0.5 * ((dql(0) + dql(2)) / cspeed + dql(1));
  rl = dmax(rl, 1.0e-40);
  double ul = ql(qn) + 0.5 * ((dql(2) - dql(0)) / ql(QRHO));
  double pl = ql(QPRES) + 0.5 * (dql(0) + dql(2)) * cspeed;
  pl = dmax(pl, 1.0e-40);
  double ut1l = ql(qt1) + 0.5 * dql(3);
  double ut2l = ql(qt2) + 0.5 * dql(4);
  cspeed = sqrt(c.gamma * c.Rspec * qr(QT)) + 1.e-40;
  double rr = qr(QRHO) - 0.5 * ((dqr(0) + dqr(2)) / cspeed + dqr(1));
  rr = dmax(rr, 1.0e-40);
  double ur = qr(qn) - 0.5 * ((dqr(2) - dqr(0)) / qr(QRHO));
  double pr_ = qr(QPRES) - 0.5 * (dqr(0) + dqr(2)) * cspeed;
  pr_ = dmax(pr_, 1.0e-40);
  double ut1r = qr(qt1) - 0.5 * dqr(3);
  double ut2r = qr(qt2) - 0.5 * dqr(4);
  double o[5];
  riemann_prob(c.gamma, 1.0e-40, rl, ul, pl, ut1l, ut2l, rr, ur, pr_, ut1r, ut2r, o);
  constexpr int cn = UMX + D;
  constexpr int ct1 = D == 0 ? UMY : UMX;
  constexpr int ct2 = D == 2 ? UMY : UMZ;
  f[URHO] = o[0];
  f[cn] = o[1];
  f[ct1] = o[2];
  f[ct2] = o[3];
  f[UET] = o[4];
}

// Stage-combination epilogue (CNS::advance MultiFab ops, src/tim/CNS_advance.cpp:51-181).
struct StageCoef {
  int mode;  // 0: a*s1 + b*(c*sin + d*L); 1: (a*s1 + c*sin) + d*L; 2: L
  double a, b, c, d;
};

// Coefficients of stage `stage` of scheme (order_rk, stages_rk) for step
// size dt, computed with the reference's own expressions
// (src/tim/CNS_advance.cpp:51-181); false for an unknown scheme or stage.
__host__ __device__ inline bool stage_coef_of(int order_rk, int stages_rk, double dt, int stage, StageCoef* out,
                                              int* nstages) {
  StageCoef r{1, 0.0, 0.0, 1.0, dt};  // Sin + d*L
  int n = 0;
  if (order_rk == -2) {
    n = 2;
    if (stage == 1) r = StageCoef{1, 0.5, 0.0, 0.5, 0.5 * dt};
  } else if (order_rk == 0) {
    n = 1;
    r = StageCoef{2, 0, 0, 0, 0};
  } else if (order_rk == 1) {
    n = 1;
  } else if (order_rk == 2) {
    int m = stages_rk;
    if (m < 2) return false;
    n = m;
    if (stage < m - 1) r = StageCoef{1, 0.0, 0.0, 1.0, dt / double(m - 1)};
    else r = StageCoef{0, 1.0 / m, 1.0 / m, double(m - 1), dt};
  } else if (order_rk == 3 && stages_rk == 3) {
    n = 3;
    if (stage == 1) r = StageCoef{0, 3.0 / 4, 1.0 / 4, 1.0, dt};
    if (stage == 2) r = StageCoef{0, 1.0 / 3, 2.0 / 3, 1.0, dt};
  } else if (order_rk == 3 && stages_rk == 4) {
    n = 4;
    r = StageCoef{1, 0.0, 0.0, 1.0, dt / 2};
    if (stage == 2) r = StageCoef{1, 2.0 / 3, 0.0, 1.0 / 3, dt / 6};
  } else {
    return false;
  }
  if (stage < 0 || stage >= n) return false;
  *out = r;
  *nstages = n;
  return true;
}

// Per-launch stage arguments: one coefficient set for every box, or (for a
// batch of independent problems) coefficients from each box's own time step
// box_dt[b]; boxes with box_dt[b] <= 0 are inactive and left untouched.
struct StageArgs {
  StageCoef sc;
  const double* box_dt;
  int order_rk, stages_rk, stage;
  const Fab* src = nullptr;  // per-box source added to dSdt (rhs_source = 1)
  double bfx = 0.0;          // x body force of the source (0: none): rho*fx to rho*u, (rho*u)*fx to rho*E
  int rhs_only = 0;          // write dSdt itself (the stage combination follows in a later pass)
};

__device__ __forceinline__ bool box_active(const StageArgs& sa, int b) {
  return sa.box_dt == nullptr || sa.box_dt[b] > 0.0;
}

__device__ __forceinline__ StageCoef box_stage(const StageArgs& sa, int b) {
  if (sa.box_dt == nullptr) return sa.sc;
  StageCoef r;
  int n;
  stage_coef_of(sa.order_rk, sa.stages_rk, sa.box_dt[b], sa.stage, &r, &n);
  return r;
}

// With a == 0 the s1 term is left out entirely (s1 is then not read), which
// reproduces LinComb(dst, 1.0, x, dt, y) = x + dt*y bit for bit.
__device__ __forceinline__ double stage_combine(const StageCoef& sc, double s1, double sin,
                                                double L) {
  if (sc.mode == 0) {
    double t = sc.c * sin + sc.d * L;
    return sc.a == 0.0 ? sc.b * t : sc.a * s1 + sc.b * t;
  }
  if (sc.mode == 1) return (sc.a == 0.0 ? sc.c * sin : sc.a * s1 + sc.c * sin) + sc.d * L;
  return L;
}

}  // namespace amr
