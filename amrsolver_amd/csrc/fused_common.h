// Device helpers of the fused KEEP + viscous stage kernel: global-address
// views, the cell viscous flux, the KEEP face flux from an LDS plane and the
// z-direction pair terms.
#pragma once

#include "cns_point.h"

namespace amr {
namespace fusedk {

// Global-address-space views: loads and stores then compile to global_*
// instructions (flat ones would share the LDS wait counter and serialise
// the plane prefetch against every LDS wait).
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) const unsigned guint;
__device__ __forceinline__ gdouble* gptr(double* p) { return (gdouble*)p; }

struct VDer {
  double ux, uy, uz, T;
  double dudx, dvdx, dwdx, dTdx, dudy, dvdy, dwdy, dTdy, dudz, dvdz, dwdz, dTdz;
};

// ViscousFluxes (CNS_hydro_K.h:133-171), components UMX..UET (URHO is 0)
template <bool WX, bool WY, bool WZ>
__device__ __forceinline__ void visc_flux(const VDer& d, const Closures& c, double fx[4], double fy[4],
                                          double fz[4]) {
  double div = d.dudx + d.dvdy + d.dwdz;
  double mu = visc(c, d.T);
  double lambda = cond(c, d.T, mu);
  double r1_3 = 1.0 / 3.0;
  double tauxx = 2.0 * mu * (d.dudx - r1_3 * div);
  double tauxy = mu * (d.dudy + d.dvdx);
  double tauxz = mu * (d.dudz + d.dwdx);
  double tauyy = 2.0 * mu * (d.dvdy - r1_3 * div);
  double tauyz = mu * (d.dvdz + d.dwdy);
  double tauzz = 2.0 * mu * (d.dwdz - r1_3 * div);
  if (WX) {
    fx[0] = -tauxx;
    fx[1] = -tauxy;
    fx[2] = -tauxz;
    fx[3] = -lambda * d.dTdx - tauxx * d.ux - tauxy * d.uy - tauxz * d.uz;
  }
  if (WY) {
    fy[0] = -tauxy;
    fy[1] = -tauyy;
    fy[2] = -tauyz;
    fy[3] = -lambda * d.dTdy - tauxy * d.ux - tauyy * d.uy - tauyz * d.uz;
  }
  if (WZ) {
    fz[0] = -tauxz;
    fz[1] = -tauyz;
    fz[2] = -tauzz;
    fz[3] = -lambda * d.dTdz - tauxz * d.ux - tauyz * d.uy - tauzz * d.uz;
  }
}

// KEEP face flux in direction NRM at the face below LDS cell `idx`; cells
// idx + m*st, m = -H..H-1 (CentralKEEP.h:58-142 accumulation order).
template <int H, int NRM, int PL>
__device__ __forceinline__ void keep_face(const double* sq, int idx, int st, const double co[3], double cv,
                                          double f[5]) {
#pragma unroll
  for (int n = 0; n < 5; ++n) f[n] = 0.0;
#pragma unroll
  for (int l = 1; l <= H; ++l) {
#pragma unroll
    for (int m = 0; m <= l - 1; ++m) {
      const int a = idx + m * st, b = idx + (m - l) * st;
      keep_pair<NRM>(co[l - 1], sq[QRHO * PL + a], sq[QU * PL + a], sq[QV * PL + a], sq[QW * PL + a],
                     cv * sq[QT * PL + a], sq[QPRES * PL + a], sq[QRHO * PL + b], sq[QU * PL + b],
                     sq[QV * PL + b], sq[QW * PL + b], cv * sq[QT * PL + b], sq[QPRES * PL + b], f);
    }
  }
}

// KEEP pair terms with the normal direction (x or y) chosen per lane: the
// operations of keep_terms<0> / keep_terms<1> (CentralKEEP.h:19-36), the
// normal-momentum component selected between the cubic and the cubic plus
// the pressure average rather than computed on a branch, so that lanes of
// both directions share one instruction stream.
__device__ __forceinline__ void keep_terms_xy(bool isx, double rho1, double ux1, double uy1, double uz1,
                                              double ie1, double p1, double rho2, double ux2, double uy2,
                                              double uz2, double ie2, double p2, double t[5]) {
  const double un1 = isx ? ux1 : uy1;
  const double un2 = isx ? ux2 : uy2;
  const double massflx = fgQuad(rho1, rho2, un1, un2);
  const double ke = 0.5 * (ux1 * ux2 + uy1 * uy2 + uz1 * uz2);
  t[URHO] = massflx;
  const double cx = fghCubic(rho1, rho2, un1, un2, ux1, ux2);
  const double cy = fghCubic(rho1, rho2, un1, un2, uy1, uy2);
  const double pd = fDiv(p1, p2);
  const double cxp = cx + pd, cyp = cy + pd;
  t[UMX] = isx ? cxp : cx;
  t[UMY] = isx ? cy : cyp;
  t[UMZ] = fghCubic(rho1, rho2, un1, un2, uz1, uz2);
  t[UET] = massflx * ke + fghCubic(rho1, rho2, un1, un2, ie1, ie2) + fgDiv(p1, p2, un1, un2);
}

// keep_face with the direction chosen per lane (isx: x face, stride 1;
// else y face, stride RX)
template <int H, int PL>
__device__ __forceinline__ void keep_face_xy(const double* sq, int idx, int st, bool isx, const double co[3],
                                             double cv, double f[5]) {
#pragma unroll
  for (int n = 0; n < 5; ++n) f[n] = 0.0;
#pragma unroll
  for (int l = 1; l <= H; ++l) {
#pragma unroll
    for (int m = 0; m <= l - 1; ++m) {
      const int a = idx + m * st, b = idx + (m - l) * st;
      double t[5];
      keep_terms_xy(isx, sq[QRHO * PL + a], sq[QU * PL + a], sq[QV * PL + a], sq[QW * PL + a],
                    cv * sq[QT * PL + a], sq[QPRES * PL + a], sq[QRHO * PL + b], sq[QU * PL + b],
                    sq[QV * PL + b], sq[QW * PL + b], cv * sq[QT * PL + b], sq[QPRES * PL + b], t);
      keep_accum(co[l - 1], t, f);
    }
  }
}

// Terms of the z-direction KEEP pair (cell 1 = q1, cell 2 = q2).
__device__ __forceinline__ void zterms(const double q1[NPRIM], const double q2[NPRIM], double cv, double t[5]) {
  keep_terms<2>(q1[QRHO], q1[QU], q1[QV], q1[QW], cv * q1[QT], q1[QPRES], q2[QRHO], q2[QU], q2[QV], q2[QW],
                cv * q2[QT], q2[QPRES], t);
}

}  // namespace fusedk
}  // namespace amr
