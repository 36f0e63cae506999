/*
 * CPU oracle: restatement of amrsolver's per-box RHS and SSPRK stage update.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product library.  Built with -O3 -ffp-contract=off so that every operation
 * rounds as written, in the reference's operation order.
 *
 * The structure deliberately follows the reference: unfused full-box passes
 * over ghosted Fortran-order arrays, one pass per reference ParallelFor.
 * Each function cites the reference lines it restates (paths relative to the
 * reference repository root).
 *
 * Pinning status (what checks each function; tests/test_oracle_*.py):
 *   cons2prim           closed form + prims->cons->prims round trip
 *   KEEP 2/4/6          constant state -> 0, periodic telescoping, order of
 *                       accuracy 2/4/6 against analytic divergence,
 *                       momentum consistency for uniform velocity
 *   WENO5 candidates    exact on quadratic data (any convex weights), ENO
 *                       limits (one smooth sub-stencil selects that candidate)
 *   WENO5 weights       exact rational WENO5-JS (Jiang & Shu 1996): equal
 *                       to the textbook scheme where (s1-s3)^2 is 0 or 1,
 *                       to the quartic beta1 where s1-2s2+s3 = 0, and
 *                       eps = 1e-40, q = 2, d = {0.1,0.6,0.3} on data whose
 *                       smoothness indicators are of the size of eps
 *   Riemann             L=R -> exact Euler flux; mirror symmetry; weak waves
 *                       against an independent exact Riemann solver; Sod star
 *                       state (reference notebook values) within the
 *                       two-shock approximation error; symmetric collisions:
 *                       flux (0,p*,0,0,0) exactly and p* within the error
 *                       band of the third iterate against the exact
 *                       two-shock p* (excludes 2 and 4 iterations)
 *   Riemann z quirk     with given slopes the z face equals the x face of the
 *                       same data once the left cell's T is T(k+1), bit for
 *                       bit (riemann_face)
 *   viscous+Sutherland  linear fields -> exact stresses / heat flux;
 *                       Sutherland closed form
 *   SSPRK variants      linear ODE -> stability polynomials; conservation
 *   ghost fill          closed forms (periodic image, foextrap, reflections)
 *   estdt               closed form on uniform states (incl. UMY quirk)
 *   tangent (Dual)      the KEEP/viscous RHS, ghost fill and SSPRK sequence
 *                       are templates over the scalar type; the dual-number
 *                       instantiation's value part equals the double path
 *                       bit for bit and its derivative part matches central
 *                       differences of the double path
 */
#include "cns_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int URHO = 0, UMX = 1, UMY = 2, UMZ = 3, UET = 4, NCONS = 5;
constexpr int QRHO = 0, QU = 1, QV = 2, QW = 3, QT = 4, QPRES = 5, NPRIM = 6;
constexpr int NGHOST = 3;  // exm/*/prob.h: #define NGHOST 3

thread_local std::string g_err;

// problem parameters (orc_set_problem): exm/Channel/prob.h ProbParm::Tw and
// ::fx, and the gas constants its bcnormal reads from the closures
struct Problem {
  double Tw = 0.0, fx = 0.0, gamma = 1.4, Rspec = 287.0;
  bool set = false;
};
Problem g_prob;

int fail(const std::string& m) {
  g_err = m;
  return -1;
}

// Forward-mode dual number (value, derivative along one direction).  Every
// value operation is the double operation of the plain path, in the same
// order, so the value part of a Dual evaluation equals the double result.
struct Dual {
  double v = 0.0, d = 0.0;
  Dual() = default;
  Dual(double x) : v(x), d(0.0) {}  // NOLINT: constants enter as (x, 0)
  Dual(double x, double y) : v(x), d(y) {}
  Dual& operator+=(const Dual& o) {
    v += o.v;
    d += o.d;
    return *this;
  }
};
inline Dual operator+(const Dual& a, const Dual& b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(const Dual& a, const Dual& b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator-(const Dual& a) { return Dual(-a.v, -a.d); }
inline Dual operator*(const Dual& a, const Dual& b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(const Dual& a, const Dual& b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
inline Dual sqrt(const Dual& a) {
  const double r = std::sqrt(a.v);
  return Dual(r, a.d / (2.0 * r));
}
// std::max(lo, x) as the plain path evaluates it: x if lo < x, else lo
inline double clamp_low(double lo, double x) { return std::max(lo, x); }
inline Dual clamp_low(double lo, const Dual& x) { return lo < x.v ? x : Dual(lo); }
// max(a, b) as the reference's device code evaluates it for the wall
// temperature floor: b when a < b, else a
inline double max_of(double a, double b) { return a < b ? b : a; }
inline Dual max_of(const Dual& a, double b) { return a.v < b ? Dual(b) : a; }

// Operation-counting scalar: the plain double value plus a count of every
// FP64 operation applied to it (the reference arithmetic's flop count).
// Counting runs single-threaded.
struct FlopCounts {
  long add = 0, mul = 0, div = 0, sqrt = 0, cmp = 0;
};
FlopCounts g_flops;
struct Counted {
  double v = 0.0;
  Counted() = default;
  Counted(double x) : v(x) {}  // NOLINT: constants enter uncounted
  Counted& operator+=(const Counted& o) {
    ++g_flops.add;
    v += o.v;
    return *this;
  }
};
inline Counted operator+(const Counted& a, const Counted& b) { ++g_flops.add; return Counted(a.v + b.v); }
inline Counted operator-(const Counted& a, const Counted& b) { ++g_flops.add; return Counted(a.v - b.v); }
inline Counted operator-(const Counted& a) { return Counted(-a.v); }
inline Counted operator*(const Counted& a, const Counted& b) { ++g_flops.mul; return Counted(a.v * b.v); }
inline Counted operator/(const Counted& a, const Counted& b) { ++g_flops.div; return Counted(a.v / b.v); }
inline Counted sqrt(const Counted& a) { ++g_flops.sqrt; return Counted(std::sqrt(a.v)); }
inline Counted clamp_low(double lo, const Counted& x) {
  ++g_flops.cmp;
  return lo < x.v ? x : Counted(lo);
}
inline Counted max_of(const Counted& a, double b) {
  ++g_flops.cmp;
  return a.v < b ? Counted(b) : a;
}

// Comparisons, min/max (std::max/std::min semantics), magnitude and sign of
// the WENO and Riemann paths for plain doubles and for counted scalars
// (comparisons are counted; |x| and copysign are sign-bit operations).
inline bool operator<(const Counted& a, const Counted& b) { ++g_flops.cmp; return a.v < b.v; }
inline bool operator<(const Counted& a, double b) { ++g_flops.cmp; return a.v < b; }
inline bool operator>(const Counted& a, double b) { ++g_flops.cmp; return a.v > b; }
inline bool operator>=(const Counted& a, const Counted& b) { ++g_flops.cmp; return a.v >= b.v; }
inline bool operator>=(const Counted& a, double b) { ++g_flops.cmp; return a.v >= b; }
inline double gmax(double a, double b) { return std::max(a, b); }
inline double gmin(double a, double b) { return std::min(a, b); }
inline double gabs(double a) { return std::abs(a); }
inline double gsqrt(double a) { return std::sqrt(a); }
inline double gcopysign(double a, double b) { return std::copysign(a, b); }
inline Counted gmax(const Counted& a, const Counted& b) { return a < b ? b : a; }
inline Counted gmin(const Counted& a, const Counted& b) { return b < a ? b : a; }
inline Counted gabs(const Counted& a) { return Counted(std::abs(a.v)); }
inline Counted gsqrt(const Counted& a) { return sqrt(a); }
inline Counted gcopysign(double a, const Counted& b) { return Counted(std::copysign(a, b.v)); }

// Fortran-order view, component outermost (AMReX Array4 convention).
template <class R>
struct A4T {
  R* p = nullptr;
  int b[3] = {0, 0, 0};  // lower corner of the allocated box
  long js = 0, ks = 0, ns = 0;
  inline R& operator()(int i, int j, int k, int n) const {
    return p[(i - b[0]) + (j - b[1]) * js + (k - b[2]) * ks + n * ns];
  }
};
using A4 = A4T<double>;

template <class R>
A4T<R> view(R* p, const int lo[3], const int hi[3]) {
  A4T<R> a;
  a.p = p;
  for (int d = 0; d < 3; ++d) a.b[d] = lo[d];
  long nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
  a.js = nx;
  a.ks = nx * ny;
  a.ns = nx * ny * nz;
  return a;
}

long npts(const int lo[3], const int hi[3]) {
  return long(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
}

struct Box {
  int lo[3], hi[3];
};

Box grow(const Box& b, int g) {
  Box r = b;
  for (int d = 0; d < 3; ++d) {
    r.lo[d] -= g;
    r.hi[d] += g;
  }
  return r;
}
Box grow_dir(const Box& b, int dir, int g) {
  Box r = b;
  r.lo[dir] -= g;
  r.hi[dir] += g;
  return r;
}
Box nodal_all(const Box& b) {  // grownnodaltilebox(-1,0) / nodaltilebox()
  Box r = b;
  for (int d = 0; d < 3; ++d) r.hi[d] += 1;
  return r;
}
Box surrounding_nodes(const Box& b, int dir) {  // amrex::surroundingNodes(bx,dir)
  Box r = b;
  r.hi[dir] += 1;
  return r;
}

template <class F>
void for_box(const Box& b, F&& f) {
  for (int k = b.lo[2]; k <= b.hi[2]; ++k)
    for (int j = b.lo[1]; j <= b.hi[1]; ++j)
      for (int i = b.lo[0]; i <= b.hi[0]; ++i) f(i, j, k);
}

// Per-thread scratch arrays of the flux passes, kept between calls (fresh
// allocations of this size come from mmap, and their page faults on first
// touch are serial); `slot` distinguishes the arrays one call holds.
template <class R>
std::vector<R>& scratch(int slot) {
  thread_local std::vector<R> v[12];
  return v[slot];
}

// Threads for the point passes inside one box (set by a level whose boxes
// are fewer than its threads, e.g. a single 1D problem).  Used only for
// passes whose point function writes its own point alone, so the results do
// not depend on it; the tangent and operation-counting scalar types always
// run serially.
thread_local int g_inner_threads = 1;

// v = n zeros (threads as for_box_par)
template <class R>
void zero_fill(std::vector<R>& v, long n) {
  if (!std::is_same<R, double>::value || g_inner_threads <= 1) {
    v.assign(n, R(0.0));
    return;
  }
  v.resize(n);
  R* p = v.data();
#pragma omp parallel for schedule(static) num_threads(g_inner_threads)
  for (long i = 0; i < n; ++i) p[i] = R(0.0);
}

template <class R, class F>
void for_box_par(const Box& b, F&& f) {
  if (!std::is_same<R, double>::value || g_inner_threads <= 1) {
    for_box(b, f);
    return;
  }
  const int nj = b.hi[1] - b.lo[1] + 1, nk = b.hi[2] - b.lo[2] + 1;
#pragma omp parallel for schedule(static) num_threads(g_inner_threads)
  for (int r = 0; r < nj * nk; ++r) {
    const int j = b.lo[1] + r % nj, k = b.lo[2] + r / nj;
    for (int i = b.lo[0]; i <= b.hi[0]; ++i) f(i, j, k);
  }
}

// ---------------------------------------------------------------- closures
// src/cls/Closures.h:103-120 (viscosity), :130-166 (conductivity).  Law 2 is
// the constant-Prandtl closure k = mu*cp/Pr that Closures.h:145-153 sketches;
// it is the closure used for the TGV and channel workloads.
template <class R>
R visc(const orc_closures& c, R T) {
  using std::sqrt;
  if (c.visc_law == 0) return c.visc_ref;                        // :108-109
  return c.visc_ref * T * sqrt(T) / (c.Tvisc_ref + T);           // :118-119
}
template <class R>
R cond(const orc_closures& c, R T) {
  using std::sqrt;
  if (c.cond_law == 0) return c.cond_ref;                        // :135-136
  if (c.cond_law == 1) return c.cond_ref * T * sqrt(T) / (c.Tcond_ref + T);  // :161
  return visc(c, T) * c.cp / c.Pr;
}

// ---------------------------------------------------------------- cons2prim
// src/rhs/CNS_hydro_K.h:15-34
template <class R>
void cons2prim(int i, int j, int k, const A4T<R>& u, const A4T<R>& q, const orc_closures& c) {
  R rho = u(i, j, k, URHO);
  rho = clamp_low(1e-40, rho);
  R rhoinv = 1.0 / rho;
  R ux = u(i, j, k, UMX) * rhoinv;
  R uy = u(i, j, k, UMY) * rhoinv;
  R uz = u(i, j, k, UMZ) * rhoinv;
  R rhoke = 0.5 * rho * (ux * ux + uy * uy + uz * uz);
  R rhoei = (u(i, j, k, UET) - rhoke);
  R p = (c.gamma - 1.0) * rhoei;
  q(i, j, k, QRHO) = rho;
  q(i, j, k, QU) = ux;
  q(i, j, k, QV) = uy;
  q(i, j, k, QW) = uz;
  q(i, j, k, QPRES) = p;
  q(i, j, k, QT) = p / (rho * c.Rspec);
}

// ---------------------------------------------------------------- KEEP
// src/rhs/CentralKEEP.h:19-36 split forms
template <class R>
inline R fDiv(R f, R fl) { return 0.5 * (f + fl); }
template <class R>
inline R fgDiv(R f, R fl, R g, R gl) { return 0.5 * (f * gl + fl * g); }
template <class R>
inline R fgQuad(R f, R fl, R g, R gl) { return 0.25 * (f + fl) * (g + gl); }
template <class R>
inline R fghCubic(R f, R fl, R g, R gl, R h, R hl) {
  return 0.1250 * (f + fl) * (g + gl) * (h + hl);
}

// coefficients: src/CNS.cpp:243-251 (CPU build), src/rhs/CentralKEEP.h:196-202 (GPU build)
void keep_coeffs(int order, double c[3]) {
  if (order == 6) {
    c[0] = 6.0 / 4; c[1] = -6.0 / 20; c[2] = 2.0 / 60;
  } else if (order == 4) {
    c[0] = 4.0 / 3; c[1] = -2.0 / 12; c[2] = 0.0;
  } else {
    c[0] = 1.0; c[1] = 0.0; c[2] = 0.0;
  }
}

// src/rhs/CentralKEEP.h:40-145: fluxes at i-1/2, j-1/2, k-1/2 of node (i,j,k)
template <class R>
void keep_point(int i, int j, int k, int halfsten, const double co[3], const A4T<R>& q,
                const A4T<R>& fx, const A4T<R>& fy, const A4T<R>& fz, const orc_closures& c) {
  for (int n = 0; n < NCONS; ++n) {
    fx(i, j, k, n) = 0.0;
    fy(i, j, k, n) = 0.0;
    fz(i, j, k, n) = 0.0;
  }
  // one direction at a time; (di,dj,dk) is the unit vector, `nrm` the normal
  // velocity component; the accumulation order is l = 1..h, m = 0..l-1
  auto dir = [&](const A4T<R>& f, int di, int dj, int dk, int nrm) {
    for (int l = 1; l <= halfsten; ++l) {
      for (int m = 0; m <= l - 1; ++m) {
        int i1 = i + m * di, j1 = j + m * dj, k1 = k + m * dk;
        int i2 = i + (m - l) * di, j2 = j + (m - l) * dj, k2 = k + (m - l) * dk;
        R rho1 = q(i1, j1, k1, QRHO), ux1 = q(i1, j1, k1, QU), uy1 = q(i1, j1, k1, QV),
          uz1 = q(i1, j1, k1, QW), ie1 = c.cv * q(i1, j1, k1, QT), p1 = q(i1, j1, k1, QPRES);
        R rho2 = q(i2, j2, k2, QRHO), ux2 = q(i2, j2, k2, QU), uy2 = q(i2, j2, k2, QV),
          uz2 = q(i2, j2, k2, QW), ie2 = c.cv * q(i2, j2, k2, QT), p2 = q(i2, j2, k2, QPRES);
        R u1[3] = {ux1, uy1, uz1}, u2[3] = {ux2, uy2, uz2};
        R un1 = u1[nrm], un2 = u2[nrm];
        R massflx = fgQuad(rho1, rho2, un1, un2);
        R ke = 0.5 * (ux1 * ux2 + uy1 * uy2 + uz1 * uz2);
        f(i, j, k, URHO) += co[l - 1] * massflx;
        for (int a = 0; a < 3; ++a) {
          if (a == nrm)
            f(i, j, k, UMX + a) += co[l - 1] * (fghCubic(rho1, rho2, un1, un2, un1, un2) + fDiv(p1, p2));
          else
            f(i, j, k, UMX + a) += co[l - 1] * fghCubic(rho1, rho2, un1, un2, u1[a], u2[a]);
        }
        f(i, j, k, UET) += co[l - 1] * (massflx * ke + fghCubic(rho1, rho2, un1, un2, ie1, ie2) +
                                        fgDiv(p1, p2, un1, un2));
      }
    }
  };
  dir(fx, 1, 0, 0, 0);
  dir(fy, 0, 1, 0, 1);
  dir(fz, 0, 0, 1, 2);
}

// src/rhs/CentralKEEP.h:147-186 KEEPy: the y flux at j-1/2 of node (i,j,k)
template <class R>
void keepy_point(int i, int j, int k, int halfsten, const double co[3], const A4T<R>& q, const A4T<R>& fy,
                 const orc_closures& c) {
  R flx[NCONS];
  for (int n = 0; n < NCONS; ++n) flx[n] = 0.0;
  for (int l = 1; l <= halfsten; ++l) {
    for (int m = 0; m <= l - 1; ++m) {
      const int j1 = j + m, j2 = j + m - l;
      R rho1 = q(i, j1, k, QRHO), ux1 = q(i, j1, k, QU), uy1 = q(i, j1, k, QV), uz1 = q(i, j1, k, QW),
        ie1 = c.cv * q(i, j1, k, QT), p1 = q(i, j1, k, QPRES);
      R rho2 = q(i, j2, k, QRHO), ux2 = q(i, j2, k, QU), uy2 = q(i, j2, k, QV), uz2 = q(i, j2, k, QW),
        ie2 = c.cv * q(i, j2, k, QT), p2 = q(i, j2, k, QPRES);
      R massflx = fgQuad(rho1, rho2, uy1, uy2);
      R ke = 0.5 * (ux1 * ux2 + uy1 * uy2 + uz1 * uz2);
      flx[URHO] = flx[URHO] + co[l - 1] * massflx;
      flx[UMX] = flx[UMX] + co[l - 1] * fghCubic(rho1, rho2, uy1, uy2, ux1, ux2);
      flx[UMY] = flx[UMY] + co[l - 1] * (fghCubic(rho1, rho2, uy1, uy2, uy1, uy2) + fDiv(p1, p2));
      flx[UMZ] = flx[UMZ] + co[l - 1] * fghCubic(rho1, rho2, uy1, uy2, uz1, uz2);
      flx[UET] = flx[UET] + co[l - 1] * (massflx * ke + fghCubic(rho1, rho2, uy1, uy2, ie1, ie2) +
                                         fgDiv(p1, p2, uy1, uy2));
    }
  }
  for (int n = 0; n < NCONS; ++n) fy(i, j, k, n) = flx[n];
}

// ---------------------------------------------------------------- WENO5
// src/rhs/High_resolution.h:83-121
template <class R>
void cons2eulerflux_lambda(int i, int j, int k, const A4T<R>& u, const A4T<R>& fx, const A4T<R>& fy,
                           const A4T<R>& fz, const A4T<R>& lam, const orc_closures& c) {
  R rhoinv = 1.0 / u(i, j, k, URHO);
  R momx = u(i, j, k, UMX), momy = u(i, j, k, UMY), momz = u(i, j, k, UMZ);
  R rhoet = u(i, j, k, UET);
  R ux = momx * rhoinv, uy = momy * rhoinv, uz = momz * rhoinv;
  R rhoekin = 0.5 * rhoinv * (momx * momx + momy * momy + momz * momz);
  R rhoeint = rhoet - rhoekin;
  R P = (c.gamma - 1.0) * rhoeint;
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
  R cs = gsqrt(c.gamma * P * rhoinv);
  lam(i, j, k, 0) = gabs(gmax(gmax(ux + cs, ux - cs), ux));
  lam(i, j, k, 1) = gabs(gmax(gmax(uy + cs, uy - cs), uy));
  lam(i, j, k, 2) = gabs(gmax(gmax(uz + cs, uz - cs), uz));
}

// src/rhs/High_resolution.h:124-159 (beta[1] uses the 4th power of
// (s1-s3), eps = 1e-40, q = 2; pow(x,2) is written x*x)
template <class R>
R weno5js(const R s[5]) {
  const double eps = 1e-40;
  R beta[3], alpha[3], vl[3];
  beta[0] = (13.0 / 12.0) * (s[0] - 2.0 * s[1] + s[2]) * (s[0] - 2.0 * s[1] + s[2]) +
            0.25 * (s[0] - 4.0 * s[1] + 3.0 * s[2]) * (s[0] - 4.0 * s[1] + 3.0 * s[2]);
  beta[1] = (13.0 / 12.0) * (s[1] - 2.0 * s[2] + s[3]) * (s[1] - 2.0 * s[2] + s[3]) +
            0.25 * (s[1] - s[3]) * (s[1] - s[3]) * (s[1] - s[3]) * (s[1] - s[3]);
  beta[2] = (13.0 / 12.0) * (s[2] - 2.0 * s[3] + s[4]) * (s[2] - 2.0 * s[3] + s[4]) +
            0.25 * (3.0 * s[2] - 4.0 * s[3] + s[4]) * (3.0 * s[2] - 4.0 * s[3] + s[4]);
  R t0 = eps + beta[0], t1 = eps + beta[1], t2 = eps + beta[2];
  alpha[0] = 0.1 / (t0 * t0);
  alpha[1] = 0.6 / (t1 * t1);
  alpha[2] = 0.3 / (t2 * t2);
  R sum = 1.0 / (alpha[0] + alpha[1] + alpha[2]);
  vl[0] = 2.0 * s[0] - 7.0 * s[1] + 11.0 * s[2];
  vl[1] = -s[1] + 5.0 * s[2] + 2.0 * s[3];
  vl[2] = 2.0 * s[2] + 5.0 * s[3] - s[4];
  return (1.0 / 6.0) * sum * (alpha[0] * vl[0] + alpha[1] * vl[1] + alpha[2] * vl[2]);
}

// src/rhs/High_resolution.h:162-246; one direction (unit vector di,dj,dk,
// eigenvalue component `d`), component n
template <class R>
R weno_face(int i, int j, int k, int n, int di, int dj, int dk, int d, const A4T<R>& u, const A4T<R>& pf,
            const A4T<R>& lam) {
  auto at = [&](int o) { return std::array<int, 3>{i + o * di, j + o * dj, k + o * dk}; };
  R me = lam(i - 3 * di, j - 3 * dj, k - 3 * dk, d);
  for (int o = -2; o <= 2; ++o) {
    auto c = at(o);
    me = gmax(me, lam(c[0], c[1], c[2], d));
  }
  R df[6];
  for (int o = -3; o <= 2; ++o) {
    auto c = at(o);
    df[o + 3] = me * u(c[0], c[1], c[2], n);
  }
  auto P = [&](int o) {
    auto c = at(o);
    return pf(c[0], c[1], c[2], n);
  };
  R st[5];
  st[0] = 0.5 * (P(-3) + df[0]);
  st[1] = 0.5 * (P(-2) + df[1]);
  st[2] = 0.5 * (P(-1) + df[2]);
  st[3] = 0.5 * (P(0) + df[3]);
  st[4] = 0.5 * (P(1) + df[4]);
  R flx = weno5js(st);
  st[4] = 0.5 * (P(-2) - df[1]);
  st[3] = 0.5 * (P(-1) - df[2]);
  st[2] = 0.5 * (P(0) - df[3]);
  st[1] = 0.5 * (P(1) - df[4]);
  st[0] = 0.5 * (P(2) - df[5]);
  flx += weno5js(st);
  return flx;
}

// ---------------------------------------------------------------- Riemann
// src/rhs/Riemann.h:9-17
template <class R>
R limiter(R dlft, R drgt) {
  R dcen = 0.5 * (dlft + drgt);
  R dsgn = gcopysign(1.0, dcen);
  R slop = 2.0 * gmin(gabs(dlft), gabs(drgt));
  R dlim = (dlft * drgt >= 0.0) ? slop : R(0.0);
  return dsgn * gmin(dlim, gabs(dcen));
}

// src/rhs/Riemann.h:19-120.  dir: 0,1,2; the normal velocity is QU+dir and
// the two transverse velocities are (QV,QW) for x, (QU,QW) for y, (QU,QV)
// for z.  The x slope adds 2e-40 to the sound speed, y and z add 1e-40.
template <class R>
void slope_point(int i, int j, int k, int dir, const A4T<R>& dq, const A4T<R>& q, const orc_closures& c) {
  int di = dir == 0, dj = dir == 1, dk = dir == 2;
  int qn = QU + dir;
  int qt1 = dir == 0 ? QV : QU;
  int qt2 = dir == 2 ? QV : QW;
  double add = dir == 0 ? 2.e-40 : 1.e-40;
  auto Q = [&](int o, int comp) { return q(i + o * di, j + o * dj, k + o * dk, comp); };
  R cspeed = gsqrt(c.gamma * c.Rspec * Q(0, QT)) + add;
  R dlft = 0.5 * (Q(0, QPRES) - Q(-1, QPRES)) / cspeed - 0.5 * Q(0, QRHO) * (Q(0, qn) - Q(-1, qn));
  R drgt = 0.5 * (Q(1, QPRES) - Q(0, QPRES)) / cspeed - 0.5 * Q(0, QRHO) * (Q(1, qn) - Q(0, qn));
  R d0 = limiter(dlft, drgt);
  R cs2 = cspeed * cspeed;
  dlft = (Q(0, QRHO) - Q(-1, QRHO)) - (Q(0, QPRES) - Q(-1, QPRES)) / cs2;
  drgt = (Q(1, QRHO) - Q(0, QRHO)) - (Q(1, QPRES) - Q(0, QPRES)) / cs2;
  R d1 = limiter(dlft, drgt);
  dlft = 0.5 * (Q(0, QPRES) - Q(-1, QPRES)) / cspeed + 0.5 * Q(0, QRHO) * (Q(0, qn) - Q(-1, qn));
  drgt = 0.5 * (Q(1, QPRES) - Q(0, QPRES)) / cspeed + 0.5 * Q(0, QRHO) * (Q(1, qn) - Q(0, qn));
  R d2 = limiter(dlft, drgt);
  dlft = Q(0, qt1) - Q(-1, qt1);
  drgt = Q(1, qt1) - Q(0, qt1);
  R d3 = limiter(dlft, drgt);
  dlft = Q(0, qt2) - Q(-1, qt2);
  drgt = Q(1, qt2) - Q(0, qt2);
  R d4 = limiter(dlft, drgt);
  dq(i, j, k, 0) = d0;
  dq(i, j, k, 1) = d1;
  dq(i, j, k, 2) = d2;
  dq(i, j, k, 3) = d3;
  dq(i, j, k, 4) = d4;
}

// src/rhs/Riemann.h:123-249
template <class R>
void riemann_prob(double gamma, double smallp, R rl, R ul, R pl, R ut1l, R ut2l, R rr, R ur, R pr, R ut1r, R ut2r,
                  R& flxrho, R& flxu, R& flxut, R& flxutt, R& flxe) {
  const double weakwv = 1.e-3;
  const double small = 1.e-6;
  R clsql = gamma * pl * rl;
  R clsqr = gamma * pr * rr;
  R wl = gsqrt(clsql);
  R wr = gsqrt(clsqr);
  R cleft = wl / rl;
  R cright = wr / rr;
  R ccsmall = small * (cleft + cright);
  R pstar = (wl * pr + wr * pl - wr * wl * (ur - ul)) / (wl + wr);
  pstar = gmax(pstar, R(smallp));
  R pstnm1 = pstar;
  R wlsq = (0.5 * (gamma - 1.) * (pstar + pl) + pstar) * rl;
  R wrsq = (0.5 * (gamma - 1.) * (pstar + pr) + pstar) * rr;
  wl = gsqrt(wlsq);
  wr = gsqrt(wrsq);
  R ustarp = ul - (pstar - pl) / wl;
  R ustarm = ur + (pstar - pr) / wr;
  pstar = (wl * pr + wr * pl - wr * wl * (ur - ul)) / (wl + wr);
  pstar = gmax(pstar, R(smallp));
  R ustar = 0.0;
  for (int iter = 0; iter < 3; ++iter) {
    wlsq = (0.5 * (gamma - 1.) * (pstar + pl) + pstar) * rl;
    wrsq = (0.5 * (gamma - 1.) * (pstar + pr) + pstar) * rr;
    wl = 1. / gsqrt(wlsq);
    wr = 1. / gsqrt(wrsq);
    R ustnm1 = ustarm;
    R ustnp1 = ustarp;
    ustarm = ur - (pr - pstar) * wr;
    ustarp = ul + (pl - pstar) * wl;
    R dpditer = gabs(pstnm1 - pstar);
    R zp = gabs(ustarp - ustnp1);
    if (zp - weakwv * cleft < 0.0) zp = dpditer * wl;
    R zm = gabs(ustarm - ustnm1);
    if (zm - weakwv * cright < 0.0) zm = dpditer * wr;
    R zz = zp + zm;
    R denom = dpditer / gmax(zz, ccsmall);
    pstnm1 = pstar;
    pstar = pstar - denom * (ustarm - ustarp);
    pstar = gmax(pstar, R(smallp));
    ustar = 0.5 * (ustarm + ustarp);
  }
  R ro, uo, po, sgnm, utrans1, utrans2;
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
  R wosq = (0.5 * (gamma - 1.) * (pstar + po) + pstar) * ro;
  R co = gsqrt(gamma * po / ro);
  R wo = gsqrt(wosq);
  R dpjmp = pstar - po;
  R rstar = ro / (1. - ro * dpjmp / wosq);
  R cstar = gsqrt(gamma * pstar / rstar);
  R spout = co - sgnm * uo;
  R spin = cstar - sgnm * uo;
  if (pstar >= po) {
    spin = wo / ro - sgnm * uo;
    spout = spin;
  }
  R ss = gmax(spout - spin, spout + spin);
  R frac = 0.5 * (1. + (spin + spout) / gmax(ss, ccsmall));
  R rgdnv, ugdnv, pgdnv;
  if (spout < 0.) {
    rgdnv = ro; ugdnv = uo; pgdnv = po;
  } else if (spin >= 0.) {
    rgdnv = rstar; ugdnv = ustar; pgdnv = pstar;
  } else {
    rgdnv = frac * rstar + (1. - frac) * ro;
    ugdnv = frac * ustar + (1. - frac) * uo;
    pgdnv = frac * pstar + (1. - frac) * po;
  }
  flxrho = rgdnv * ugdnv;
  flxu = rgdnv * ugdnv * ugdnv + pgdnv;
  flxut = rgdnv * ugdnv * utrans1;
  flxutt = rgdnv * ugdnv * utrans2;
  flxe = ugdnv * (0.5 * rgdnv * (ugdnv * ugdnv + utrans1 * utrans1 + utrans2 * utrans2) +
                  pgdnv / (gamma - 1.) + pgdnv);
}

// src/rhs/Riemann.h:251-337.  Face at (i,j,k) - 1/2 e_dir.  The z direction
// takes the left-state sound speed from T at k+1 (Riemann.h:316).  The
// flux components are written normal-first and permuted back per direction
// (Riemann.h:278, :306, :336).
template <class R>
void riemann_point(int i, int j, int k, int dir, const A4T<R>& f, const A4T<R>& dq, const A4T<R>& q,
                   const orc_closures& c) {
  int di = dir == 0, dj = dir == 1, dk = dir == 2;
  int qn = QU + dir;
  int qt1 = dir == 0 ? QV : QU;
  int qt2 = dir == 2 ? QV : QW;
  int il = i - di, jl = j - dj, kl = k - dk;
  R Tl = dir == 2 ? q(i, j, k + 1, QT) : q(il, jl, kl, QT);
  R cspeed = gsqrt(c.gamma * c.Rspec * Tl) + 1.e-40;
  R rl = q(il, jl, kl, QRHO) + 0.5 * ((dq(il, jl, kl, 0) + dq(il, jl, kl, 2)) / cspeed + dq(il, jl, kl, 1));
  rl = gmax(rl, R(1.0e-40));
  R ul = q(il, jl, kl, qn) + 0.5 * ((dq(il, jl, kl, 2) - dq(il, jl, kl, 0)) / q(il, jl, kl, QRHO));
  R pl = q(il, jl, kl, QPRES) + 0.5 * (dq(il, jl, kl, 0) + dq(il, jl, kl, 2)) * cspeed;
  pl = gmax(pl, R(1.0e-40));
  R ut1l = q(il, jl, kl, qt1) + 0.5 * dq(il, jl, kl, 3);
  R ut2l = q(il, jl, kl, qt2) + 0.5 * dq(il, jl, kl, 4);

  cspeed = gsqrt(c.gamma * c.Rspec * q(i, j, k, QT)) + 1.e-40;
  R rr = q(i, j, k, QRHO) - 0.5 * ((dq(i, j, k, 0) + dq(i, j, k, 2)) / cspeed + dq(i, j, k, 1));
  rr = gmax(rr, R(1.0e-40));
  R ur = q(i, j, k, qn) - 0.5 * ((dq(i, j, k, 2) - dq(i, j, k, 0)) / q(i, j, k, QRHO));
  R pr = q(i, j, k, QPRES) - 0.5 * (dq(i, j, k, 0) + dq(i, j, k, 2)) * cspeed;
  pr = gmax(pr, R(1.0e-40));
  R ut1r = q(i, j, k, qt1) - 0.5 * dq(i, j, k, 3);
  R ut2r = q(i, j, k, qt2) - 0.5 * dq(i, j, k, 4);

  R frho, fu, fut, futt, fe;
  riemann_prob(c.gamma, 1.0e-40, rl, ul, pl, ut1l, ut2l, rr, ur, pr, ut1r, ut2r, frho, fu, fut,
               futt, fe);
  int cn = UMX + dir;
  int ct1 = dir == 0 ? UMY : UMX;
  int ct2 = dir == 2 ? UMY : UMZ;
  f(i, j, k, URHO) = frho;
  f(i, j, k, cn) = fu;
  f(i, j, k, ct1) = fut;
  f(i, j, k, ct2) = futt;
  f(i, j, k, UET) = fe;
}

// ---------------------------------------------------------------- viscous
// src/rhs/CNS_hydro_K.h:108-172 (cell-centred viscous fluxes)
template <class R>
void viscous_point(int i, int j, int k, const A4T<R>& q, const A4T<R>& fx, const A4T<R>& fy,
                   const A4T<R>& fz,
                   const double dxinv[3], const orc_closures& c) {
  R ux = q(i, j, k, QU);
  R dudx = (q(i + 1, j, k, QU) - q(i - 1, j, k, QU)) * 0.5 * dxinv[0];
  R dvdx = (q(i + 1, j, k, QV) - q(i - 1, j, k, QV)) * 0.5 * dxinv[0];
  R dwdx = (q(i + 1, j, k, QW) - q(i - 1, j, k, QW)) * 0.5 * dxinv[0];
  R dTdx = (q(i + 1, j, k, QT) - q(i - 1, j, k, QT)) * 0.5 * dxinv[0];
  R uy = q(i, j, k, QV);
  R dudy = (q(i, j + 1, k, QU) - q(i, j - 1, k, QU)) * 0.5 * dxinv[1];
  R dvdy = (q(i, j + 1, k, QV) - q(i, j - 1, k, QV)) * 0.5 * dxinv[1];
  R dwdy = (q(i, j + 1, k, QW) - q(i, j - 1, k, QW)) * 0.5 * dxinv[1];
  R dTdy = (q(i, j + 1, k, QT) - q(i, j - 1, k, QT)) * 0.5 * dxinv[1];
  R uz = q(i, j, k, QW);
  R dudz = (q(i, j, k + 1, QU) - q(i, j, k - 1, QU)) * 0.5 * dxinv[2];
  R dvdz = (q(i, j, k + 1, QV) - q(i, j, k - 1, QV)) * 0.5 * dxinv[2];
  R dwdz = (q(i, j, k + 1, QW) - q(i, j, k - 1, QW)) * 0.5 * dxinv[2];
  R dTdz = (q(i, j, k + 1, QT) - q(i, j, k - 1, QT)) * 0.5 * dxinv[2];
  R div = dudx + dvdy + dwdz;
  R mu = visc(c, q(i, j, k, QT));
  R lambda = cond(c, q(i, j, k, QT));
  double r1_3 = 1.0 / 3.0;
  R tauxx = 2.0 * mu * (dudx - r1_3 * div);
  R tauxy = mu * (dudy + dvdx);
  R tauxz = mu * (dudz + dwdx);
  R tauyy = 2.0 * mu * (dvdy - r1_3 * div);
  R tauyz = mu * (dvdz + dwdy);
  R tauzz = 2.0 * mu * (dwdz - r1_3 * div);
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

// src/rhs/CNS_hydro_K.h:177-251 ViscousWallFluxes: loc 0 (lower wall) and
// 1 (upper wall) take one-sided second-order y derivatives
template <class R>
void viscous_wall_point(int i, int j, int k, int loc, const A4T<R>& q, const A4T<R>& fx, const A4T<R>& fy,
                        const A4T<R>& fz, const double dxinv[3], const orc_closures& c) {
  R ux = q(i, j, k, QU);
  R dudx = (q(i + 1, j, k, QU) - q(i - 1, j, k, QU)) * 0.5 * dxinv[0];
  R dvdx = (q(i + 1, j, k, QV) - q(i - 1, j, k, QV)) * 0.5 * dxinv[0];
  R dwdx = (q(i + 1, j, k, QW) - q(i - 1, j, k, QW)) * 0.5 * dxinv[0];
  R dTdx = (q(i + 1, j, k, QT) - q(i - 1, j, k, QT)) * 0.5 * dxinv[0];
  R uy = q(i, j, k, QV);
  R dudy, dvdy, dwdy, dTdy;
  if (loc == 0) {
    dudy = (-q(i, j + 2, k, QU) + 4 * q(i, j + 1, k, QU) - 3 * q(i, j, k, QU)) * 0.5 * dxinv[1];
    dvdy = (-q(i, j + 2, k, QV) + 4 * q(i, j + 1, k, QV) - 3 * q(i, j, k, QV)) * 0.5 * dxinv[1];
    dwdy = (-q(i, j + 2, k, QW) + 4 * q(i, j + 1, k, QW) - 3 * q(i, j, k, QW)) * 0.5 * dxinv[1];
    dTdy = (-q(i, j + 2, k, QT) + 4 * q(i, j + 1, k, QT) - 3 * q(i, j, k, QT)) * 0.5 * dxinv[1];
  } else {
    dudy = (3 * q(i, j, k, QU) - 4 * q(i, j - 1, k, QU) + q(i, j - 2, k, QU)) * 0.5 * dxinv[1];
    dvdy = (3 * q(i, j, k, QV) - 4 * q(i, j - 1, k, QV) + q(i, j - 2, k, QV)) * 0.5 * dxinv[1];
    dwdy = (3 * q(i, j, k, QW) - 4 * q(i, j - 1, k, QW) + q(i, j - 2, k, QW)) * 0.5 * dxinv[1];
    dTdy = (3 * q(i, j, k, QT) - 4 * q(i, j - 1, k, QT) + q(i, j - 2, k, QT)) * 0.5 * dxinv[1];
  }
  R uz = q(i, j, k, QW);
  R dudz = (q(i, j, k + 1, QU) - q(i, j, k - 1, QU)) * 0.5 * dxinv[2];
  R dvdz = (q(i, j, k + 1, QV) - q(i, j, k - 1, QV)) * 0.5 * dxinv[2];
  R dwdz = (q(i, j, k + 1, QW) - q(i, j, k - 1, QW)) * 0.5 * dxinv[2];
  R dTdz = (q(i, j, k + 1, QT) - q(i, j, k - 1, QT)) * 0.5 * dxinv[2];
  R div = dudx + dvdy + dwdz;
  R mu = visc(c, q(i, j, k, QT));
  R lambda = cond(c, q(i, j, k, QT));
  double r1_3 = 1.0 / 3.0;
  R tauxx = 2.0 * mu * (dudx - r1_3 * div);
  R tauxy = mu * (dudy + dvdx);
  R tauxz = mu * (dudz + dwdx);
  R tauyy = 2.0 * mu * (dvdy - r1_3 * div);
  R tauyz = mu * (dvdz + dwdy);
  R tauzz = 2.0 * mu * (dwdz - r1_3 * div);
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

// Which y faces of a box sit on bc 6 walls (the box's own lower/upper end
// is the domain's and the code there is 6).
struct Walls {
  bool lo = false, hi = false;
};

// ---------------------------------------------------------------- box RHS
int check_args(const orc_scheme& s, int ng) {
  if (ng < NGHOST) return fail("state FAB needs at least 3 ghost cells");
  if (!(s.art_diss == 0 || s.art_diss == 1)) return fail("art_diss must be 0 or 1");
  if (s.rhs_euler && !(s.flux_euler >= 0 && s.flux_euler <= 2)) return fail("flux_euler must be 0, 1 or 2");
  if (s.rhs_euler && s.flux_euler == 1 && !(s.order_keep == 2 || s.order_keep == 4 || s.order_keep == 6))
    return fail("order_keep must be 2, 4 or 6");
  if (!s.rhs_euler && s.rhs_visc)
    return fail("rhs_euler=0 with rhs_visc=1 accumulates into an uninitialised flux array in the reference");
  return 0;
}

// ---------------------------------------------------------------- JST
// CNS_hydro_K.h:38-44 disconSensor
double discon_sensor(double pp, double pl, double pr) {
  double pjst = pr + 2.0 * pp + pl;
  double ptvd = std::abs(pr - pp) + std::abs(pp - pl);
  return std::abs(2.0 * (pr - 2.0 * pp + pl) / (pjst + ptvd + 1.0e-40));
}

// JST artificial dissipation of one box (art_diss = 1; CNS_advance.cpp:235-265):
// sensor and spectral radius (3 components each, NGHOST ghosts set to 0)
// computed on the valid box only (ComputeSensorLambda, CNS_hydro_K.h:46-81),
// so at a box face the cells outside read 0; then JSTflux
// (CNS_hydro_K.h:84-104) subtracts from every face flux of the all-nodal box,
// reading component 0 of sensor and spectral radius in every direction and,
// in z, cons(k-1) in place of cons(k+1) and cons(k-2).
void jst_box(const orc_closures& c, const Box& bx, const A4& u, const A4& q, const A4 nf[3]) {
  Box bg = grow(bx, NGHOST);
  std::vector<double> senb(npts(bg.lo, bg.hi) * 3, 0.0), lamb(npts(bg.lo, bg.hi) * 3, 0.0);
  A4 sen = view(senb.data(), bg.lo, bg.hi), lam = view(lamb.data(), bg.lo, bg.hi);
  for_box(bx, [&](int i, int j, int k) {
    double pp = q(i, j, k, QPRES);
    sen(i, j, k, 0) = discon_sensor(pp, q(i - 1, j, k, QPRES), q(i + 1, j, k, QPRES));
    sen(i, j, k, 1) = discon_sensor(pp, q(i, j - 1, k, QPRES), q(i, j + 1, k, QPRES));
    sen(i, j, k, 2) = discon_sensor(pp, q(i, j, k - 1, QPRES), q(i, j, k + 1, QPRES));
    double ux = q(i, j, k, QU), uy = q(i, j, k, QV), uz = q(i, j, k, QW);
    double cs = std::sqrt(c.gamma * q(i, j, k, QPRES) / q(i, j, k, QRHO));
    lam(i, j, k, 0) = std::abs(ux) + cs;
    lam(i, j, k, 1) = std::abs(uy) + cs;
    lam(i, j, k, 2) = std::abs(uz) + cs;
  });
  auto mx = [](double a, double b) { return a < b ? b : a; };  // amrex::max
  for (int n = 0; n < NCONS; ++n)
    for_box(nodal_all(bx), [&](int i, int j, int k) {
      double uijk = u(i, j, k, n);
      double dw = uijk - u(i - 1, j, k, n);
      double sn = c.Cshock * mx(sen(i - 1, j, k, 0), sen(i, j, k, 0));
      double rr = mx(lam(i - 1, j, k, 0), lam(i, j, k, 0));
      double fd = u(i + 1, j, k, n) - 3.0 * u(i, j, k, n) + 3.0 * u(i - 1, j, k, n) - u(i - 2, j, k, n);
      nf[0](i, j, k, n) -= (sn * dw - mx(0.0, c.Cdamp - sn) * fd) * rr;
      dw = uijk - u(i, j - 1, k, n);
      sn = c.Cshock * mx(sen(i, j - 1, k, 0), sen(i, j, k, 0));
      rr = mx(lam(i, j - 1, k, 0), lam(i, j, k, 0));
      fd = u(i, j + 1, k, n) - 3.0 * u(i, j, k, n) + 3.0 * u(i, j - 1, k, n) - u(i, j - 2, k, n);
      nf[1](i, j, k, n) -= (sn * dw - mx(0.0, c.Cdamp - sn) * fd) * rr;
      dw = uijk - u(i, j, k - 1, n);
      sn = c.Cshock * mx(sen(i, j, k - 1, 0), sen(i, j, k, 0));
      rr = mx(lam(i, j, k - 1, 0), lam(i, j, k, 0));
      fd = u(i, j, k - 1, n) - 3.0 * u(i, j, k, n) + 3.0 * u(i, j, k - 1, n) - u(i, j, k - 1, n);
      nf[2](i, j, k, n) -= (sn * dw - mx(0.0, c.Cdamp - sn) * fd) * rr;
    });
}

// Flux phase of CNS::compute_rhs (src/tim/CNS_advance.cpp:200-353) for one
// box: fills nf[3] (nodal FABs on [lo-ng, hi+1+ng]) in the same regions the
// reference writes.
template <class R>
int box_flux(const orc_closures& c, const orc_scheme& s, const Box& bx, int ng, const A4T<R>& u,
             const double dxinv[3], const A4T<R> nf[3], const Walls& walls = Walls()) {
  int rc = check_args(s, ng);
  if (rc) return rc;
  constexpr bool plain = std::is_same<R, double>::value;
  constexpr bool dual = std::is_same<R, Dual>::value;
  if (dual && s.rhs_euler && s.flux_euler != 1)
    return fail("the tangent (dual-number) path covers the KEEP and viscous right-hand sides");
  if (!plain && s.rhs_euler && s.art_diss)
    return fail("the tangent and counting paths do not cover JST dissipation");
  Box bg = grow(bx, NGHOST);
  std::vector<R>& qbuf = scratch<R>(3);
  zero_fill(qbuf, npts(bg.lo, bg.hi) * NPRIM);
  A4T<R> q = view(qbuf.data(), bg.lo, bg.hi);
  // CNS_advance.cpp:200-208 cons2prim on growntilebox(NGHOST)
  for_box_par<R>(bg, [&](int i, int j, int k) { cons2prim(i, j, k, u, q, c); });

  // CentralKEEP.h:233-286 Flux_2nd_Order_KEEP (called at CNS_advance.cpp:226-229):
  // y faces lo..lo+h-1 next to a lower wall and hi-h+1..hi next to an upper
  // wall redone with second-order KEEP (KEEPy with one pair, coefficient 1)
  auto wall_order_reduction = [&]() {
    if (!(s.flux_euler == 1 && s.order_keep != 2)) return;
    const int h = s.order_keep / 2;
    const double co2[3] = {1.0, 0.0, 0.0};
    Box bn = nodal_all(bx);
    for (int side = 0; side < 2; ++side) {
      if (side == 0 ? !walls.lo : !walls.hi) continue;
      Box bw = bn;
      if (side == 0) {
        bw.lo[1] = bx.lo[1];
        bw.hi[1] = bx.lo[1] + h - 1;
      } else {
        bw.lo[1] = bx.hi[1] - h + 1;
        bw.hi[1] = bx.hi[1];
      }
      for_box(bw, [&](int i, int j, int k) { keepy_point(i, j, k, 1, co2, q, nf[1], c); });
    }
  };
  if (s.rhs_euler) {
    if constexpr (dual) {
      double co[3];
      keep_coeffs(s.order_keep, co);
      int h = s.order_keep / 2;
      for_box(nodal_all(bx), [&](int i, int j, int k) { keep_point(i, j, k, h, co, q, nf[0], nf[1], nf[2], c); });
      wall_order_reduction();
    } else if (s.flux_euler == 2) {
      // High_resolution.h:279-319
      std::vector<R>& lb = scratch<R>(7);
      zero_fill(lb, npts(bg.lo, bg.hi) * 3);
      A4T<R> pf[3];
      for (int d = 0; d < 3; ++d) {
        std::vector<R>& pfb = scratch<R>(4 + d);
        zero_fill(pfb, npts(bg.lo, bg.hi) * NCONS);
        pf[d] = view(pfb.data(), bg.lo, bg.hi);
      }
      A4T<R> lam = view(lb.data(), bg.lo, bg.hi);
      for_box_par<R>(bg, [&](int i, int j, int k) { cons2eulerflux_lambda(i, j, k, u, pf[0], pf[1], pf[2], lam, c); });
      Box bn = nodal_all(bx);
      for (int n = 0; n < NCONS; ++n)
        for_box_par<R>(bn, [&](int i, int j, int k) {
          nf[0](i, j, k, n) = weno_face(i, j, k, n, 1, 0, 0, 0, u, pf[0], lam);
          nf[1](i, j, k, n) = weno_face(i, j, k, n, 0, 1, 0, 1, u, pf[1], lam);
          nf[2](i, j, k, n) = weno_face(i, j, k, n, 0, 0, 1, 2, u, pf[2], lam);
        });
    } else if (s.flux_euler == 1) {
      // CentralKEEP.h:189-229
      double co[3];
      keep_coeffs(s.order_keep, co);
      int h = s.order_keep / 2;
      for_box_par<R>(nodal_all(bx), [&](int i, int j, int k) { keep_point(i, j, k, h, co, q, nf[0], nf[1], nf[2], c); });
      wall_order_reduction();
    } else {
      // Riemann.h:339-408
      Box bg1 = grow(bx, 1);
      std::vector<R>& sb = scratch<R>(8);
      zero_fill(sb, npts(bg1.lo, bg1.hi) * NCONS);
      A4T<R> sl = view(sb.data(), bg1.lo, bg1.hi);
      for (int d = 0; d < 3; ++d) {
        for_box_par<R>(grow_dir(bx, d, 1), [&](int i, int j, int k) { slope_point(i, j, k, d, sl, q, c); });
        for_box_par<R>(surrounding_nodes(bx, d), [&](int i, int j, int k) { riemann_point(i, j, k, d, nf[d], sl, q, c); });
      }
    }
  }

  if constexpr (plain) {
    if (s.rhs_euler && s.art_diss) jst_box(c, bx, u, q, nf);
  }

  if (s.rhs_visc) {
    // CNS_advance.cpp:276-353 (no walls: Flux_2nd_Order_KEEP and
    // ViscousWallFluxes apply only to bc=6 y-walls)
    Box b1 = grow(bx, 1);
    A4T<R> pv[3];
    for (int d = 0; d < 3; ++d) {
      std::vector<R>& pb = scratch<R>(9 + d);
      zero_fill(pb, npts(b1.lo, b1.hi) * NCONS);
      pv[d] = view(pb.data(), b1.lo, b1.hi);
    }
    for_box_par<R>(b1, [&](int i, int j, int k) { viscous_point(i, j, k, q, pv[0], pv[1], pv[2], dxinv, c); });
    // CNS_advance.cpp:303-327: the ghost row outside a bc 6 wall takes
    // ViscousWallFluxes over growntilebox(1)'s x and z extent
    for (int side = 0; side < 2; ++side) {
      if (side == 0 ? !walls.lo : !walls.hi) continue;
      Box bw = b1;
      bw.lo[1] = bw.hi[1] = side == 0 ? bx.lo[1] - 1 : bx.hi[1] + 1;
      for_box(bw, [&](int i, int j, int k) { viscous_wall_point(i, j, k, side, q, pv[0], pv[1], pv[2], dxinv, c); });
    }
    // CNS_hydro_K.h:254-261 on grownnodaltilebox(-1,0)
    for (int n = 0; n < NCONS; ++n)
      for_box_par<R>(nodal_all(bx), [&](int i, int j, int k) {
        nf[0](i, j, k, n) += 0.5 * (pv[0](i - 1, j, k, n) + pv[0](i, j, k, n));
        nf[1](i, j, k, n) += 0.5 * (pv[1](i, j - 1, k, n) + pv[1](i, j, k, n));
        nf[2](i, j, k, n) += 0.5 * (pv[2](i, j, k - 1, n) + pv[2](i, j, k, n));
      });
  }
  return 0;
}

// CNS_advance.cpp:376-391 assembly
template <class R>
void box_assemble(const Box& bx, const A4T<R> nf[3], const double dxinv[3], const A4T<R>& dsdt) {
  for (int n = 0; n < NCONS; ++n)
    for_box_par<R>(bx, [&](int i, int j, int k) {
      dsdt(i, j, k, n) = dxinv[0] * (nf[0](i, j, k, n) - nf[0](i + 1, j, k, n)) +
                         dxinv[1] * (nf[1](i, j, k, n) - nf[1](i, j + 1, k, n)) +
                         dxinv[2] * (nf[2](i, j, k, n) - nf[2](i, j, k + 1, n));
    });
}

// src: source FAB on the valid box (NCONS comps) added after the assembly
// when rhs_source = 1 (the user_source hook, CNS_advance.cpp:397-407)
template <class R>
int box_rhs(const orc_closures& c, const orc_scheme& s, const Box& bx, int ng, const A4T<R>& u,
            const double dxinv[3], const A4T<R>& dsdt, const double* src, const Walls& walls = Walls()) {
  Box bn = grow(nodal_all(bx), NGHOST);
  A4T<R> nf[3];
  for (int d = 0; d < 3; ++d) {
    std::vector<R>& fb = scratch<R>(d);
    zero_fill(fb, npts(bn.lo, bn.hi) * NCONS);
    nf[d] = view(fb.data(), bn.lo, bn.hi);
  }
  int rc = box_flux(c, s, bx, ng, u, dxinv, nf, walls);
  if (rc) return rc;
  box_assemble(bx, nf, dxinv, dsdt);
  if (s.rhs_source) {
    const bool body = g_prob.set && g_prob.fx != 0.0;
    if (!src && !body) return fail("rhs_source=1 needs a source array or a body force");
    if (src) {
      A4 sv = view(const_cast<double*>(src), bx.lo, bx.hi);
      for (int n = 0; n < NCONS; ++n)
        for_box(bx, [&](int i, int j, int k) { dsdt(i, j, k, n) = dsdt(i, j, k, n) + sv(i, j, k, n); });
    }
    if (body) {  // exm/Channel/prob.h:189-193 user_source
      const double fx = g_prob.fx;
      for_box(bx, [&](int i, int j, int k) {
        dsdt(i, j, k, UMX) = dsdt(i, j, k, UMX) + u(i, j, k, URHO) * fx;
        dsdt(i, j, k, UET) = dsdt(i, j, k, UET) + u(i, j, k, UMX) * fx;
      });
    }
  }
  return 0;
}

// ---------------------------------------------------------------- level
struct Level {
  const orc_geom* g;
  int nbox;
  std::vector<Box> boxes;
  int ng;
};

// BC code -> AMReX BCType per component class (src/set/CNS_setup.cpp:18-31)
enum { INT_DIR = 0, EXT_DIR = 1, FOEXTRAP = 2, REFLECT_EVEN = 3, REFLECT_ODD = 4 };
const int scalar_bc[7] = {INT_DIR, EXT_DIR, FOEXTRAP, REFLECT_EVEN, REFLECT_EVEN, REFLECT_EVEN, EXT_DIR};
const int norm_vel_bc[7] = {INT_DIR, EXT_DIR, FOEXTRAP, REFLECT_ODD, REFLECT_ODD, REFLECT_ODD, EXT_DIR};
const int tang_vel_bc[7] = {INT_DIR, EXT_DIR, FOEXTRAP, REFLECT_EVEN, REFLECT_EVEN, REFLECT_ODD, EXT_DIR};

int bc_type(int code, int comp, int dir) {
  if (code < 0 || code > 6) return -1;
  if (comp == URHO || comp == UET) return scalar_bc[code];
  return (comp - UMX == dir) ? norm_vel_bc[code] : tang_vel_bc[code];
}

// ------------------------------------------------ coarse-fine coupling
// The reference takes these from AMReX (submodule lib/amrex,
// https://github.com/AMReX-Codes/amrex; the snapshot records no commit and the
// submodule is empty).  Restated from AMReX's published algorithms:
//  - FillPatch of a fine level (AmrLevel FillPatch -> FillPatchTwoLevels):
//    same-level copies and periodic images first; every ghost cell inside the
//    domain (periodic directions unbounded) that no fine box covers then takes
//    the coarse level's data at the fill time, interpolated; physical BCs last;
//  - the coarse data at time t (StateData::getData + FillPatchSingleLevel):
//    the new data if |t - t1| < teps, else the old data if |t - t0| < teps
//    (teps = 1e-3 (t1 - t0)), else a*old + b*new, a = (t1 - t)/(t1 - t0),
//    b = (t - t0)/(t1 - t0);
//  - CellConservativeLinear with linear limiting (lincc_interp, the
//    interpolater of State_Type, src/set/CNS_setup.cpp:112-114): per coarse
//    cell and direction d the central slope dc = (u+ - u-)/2 of every
//    component, the MC-limited slope s = sign(dc) min(|dc|, min(|2(u+ - u0)|,
//    |2(u0 - u-)|)) (0 where 2(u+ - u0) and 2(u0 - u-) differ in sign), one
//    factor f_d = min over components of s/dc (components with dc = 0 skip
//    it), slope f_d*dc; the fine cell gets u0 + x*sx + y*sy + z*sz with
//    offsets (i - r*ic + 1/2)/r - 1/2;
//  - average_down (CNS::avgDown, src/CNS.cpp:513-530): each coarse cell under
//    a fine box becomes (1/r^3) * the sum of its r^3 children, summed with i
//    fastest, then j, then k.
// Slopes next to ext_dir boundaries (one-sided in AMReX) are not restated:
// such stencils are rejected.
// Owner box of every cell of a level's domain (-1: no box), for lookups in
// constant time.
struct OwnerMap {
  int n[3] = {0, 0, 0}, lo[3] = {0, 0, 0};
  std::vector<int> own;
  void build(const Level& L) {
    const orc_geom& g = *L.g;
    for (int d = 0; d < 3; ++d) {
      lo[d] = g.dom_lo[d];
      n[d] = g.dom_hi[d] - g.dom_lo[d] + 1;
    }
    own.assign(size_t(n[0]) * n[1] * n[2], -1);
    for (int b = 0; b < L.nbox; ++b) {
      const Box& x = L.boxes[b];
      for (int k = x.lo[2]; k <= x.hi[2]; ++k)
        for (int j = x.lo[1]; j <= x.hi[1]; ++j)
          for (int i = x.lo[0]; i <= x.hi[0]; ++i)
            own[(size_t(k - lo[2]) * n[1] + (j - lo[1])) * n[0] + (i - lo[0])] = b;
    }
  }
  // c inside the domain
  int at(const int c[3]) const {
    return own[(size_t(c[2] - lo[2]) * n[1] + (c[1] - lo[1])) * n[0] + (c[0] - lo[0])];
  }
};

struct CoarseFine {
  const Level* C = nullptr;        // coarse level
  const double* const* u0 = nullptr;  // coarse FABs at t0, ghost cells filled
  const double* const* u1 = nullptr;  // coarse FABs at t1, ghost cells filled
  double a = 1.0, b = 0.0;          // weights of u0 and u1 at the fill time
  int ratio = 2;
  const OwnerMap* cmap = nullptr;   // owner boxes of the coarse level (set per fill)
};

void cf_weights(double t0, double t1, double t, double* a, double* b) {
  const double teps = (t1 - t0) * 1.0e-3;
  if (t1 == t0 || (t > t1 - teps && t < t1 + teps)) {
    *a = 0.0;
    *b = 1.0;
  } else if (t > t0 - teps && t < t0 + teps) {
    *a = 1.0;
    *b = 0.0;
  } else {
    *a = (t1 - t) / (t1 - t0);
    *b = (t - t0) / (t1 - t0);
  }
}

int floor_div(int a, int r) { return a >= 0 ? a / r : -((-a + r - 1) / r); }

bool in_box(const Box& s, const int c[3]) {
  return c[0] >= s.lo[0] && c[0] <= s.hi[0] && c[1] >= s.lo[1] && c[1] <= s.hi[1] && c[2] >= s.lo[2] &&
         c[2] <= s.hi[2];
}

// Box of `L` whose valid region holds cell c (periodic image taken), or -1;
// -2 if c lies outside the domain in a non-periodic direction.
int owner_box(const Level& L, const OwnerMap& m, const int c0[3]) {
  const orc_geom& g = *L.g;
  int c[3] = {c0[0], c0[1], c0[2]};
  for (int d = 0; d < 3; ++d) {
    const int len = g.dom_hi[d] - g.dom_lo[d] + 1;
    if (c[d] < g.dom_lo[d] || c[d] > g.dom_hi[d]) {
      if (!g.periodic[d]) return -2;
      c[d] = g.dom_lo[d] + ((c[d] - g.dom_lo[d]) % len + len) % len;
    }
  }
  return m.at(c);
}

int check_two_levels(const orc_geom& gf, const orc_geom& gc, int r) {
  if (r != 2 && r != 4) return fail("refinement ratio must be 2 or 4");
  for (int d = 0; d < 3; ++d) {
    if (gf.dom_lo[d] != gc.dom_lo[d] * r || gf.dom_hi[d] != (gc.dom_hi[d] + 1) * r - 1)
      return fail("fine domain is not the coarse domain refined");
    if (gf.periodic[d] != gc.periodic[d] || gf.lo_bc[d] != gc.lo_bc[d] || gf.hi_bc[d] != gc.hi_bc[d])
      return fail("fine and coarse levels differ in periodicity or boundary codes");
  }
  return 0;
}

// Interpolation of fine cell c (inside the domain, periodic image taken) from
// the coarse level.
int interp_cell(const CoarseFine& cf, const int c[3], double out[NCONS]) {
  const Level& C = *cf.C;
  const orc_geom& gc = *C.g;
  const int r = cf.ratio;
  int ic[3];
  for (int d = 0; d < 3; ++d) ic[d] = floor_div(c[d], r);
  const int cb = owner_box(C, *cf.cmap, ic);
  if (cb < 0) return fail("coarse-fine interpolation: fine level not properly nested (no coarse box holds a coarse cell)");
  // periodic image of ic inside box cb: the stencil is read from that FAB
  int icb[3];
  for (int d = 0; d < 3; ++d) {
    const int len = gc.dom_hi[d] - gc.dom_lo[d] + 1;
    icb[d] = ic[d];
    if (gc.periodic[d]) icb[d] = gc.dom_lo[d] + ((ic[d] - gc.dom_lo[d]) % len + len) % len;
  }
  for (int d = 0; d < 3; ++d) {
    for (int e = -1; e <= 1; e += 2) {
      int p[3] = {icb[0], icb[1], icb[2]};
      p[d] += e;
      if (owner_box(C, *cf.cmap, p) == -1)
        return fail("coarse-fine interpolation: the coarse stencil leaves the coarse level (fine level not properly nested)");
    }
    if (!gc.periodic[d]) {
      const bool elo = gc.lo_bc[d] == 1 || gc.lo_bc[d] == 6, ehi = gc.hi_bc[d] == 1 || gc.hi_bc[d] == 6;
      if ((elo && icb[d] == gc.dom_lo[d]) || (ehi && icb[d] == gc.dom_hi[d]))
        return fail("coarse-fine interpolation next to ext_dir boundaries is not supported");
    }
  }
  if (C.ng < 1) return fail("coarse FABs need at least one ghost cell");
  Box gb = grow(C.boxes[cb], C.ng);
  A4 v0, v1;
  if (cf.u0) v0 = view(const_cast<double*>(cf.u0[cb]), gb.lo, gb.hi);
  if (cf.u1) v1 = view(const_cast<double*>(cf.u1[cb]), gb.lo, gb.hi);
  auto U = [&](int di, int dj, int dk, int n) -> double {
    const int i = icb[0] + di, j = icb[1] + dj, k = icb[2] + dk;
    if (cf.b == 0.0) return v0(i, j, k, n);
    if (cf.a == 0.0) return v1(i, j, k, n);
    return cf.a * v0(i, j, k, n) + cf.b * v1(i, j, k, n);
  };
  static const int E[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  double u0[NCONS], sl[3][NCONS], sf[3] = {1.0, 1.0, 1.0};
  for (int n = 0; n < NCONS; ++n) {
    u0[n] = U(0, 0, 0, n);
    for (int d = 0; d < 3; ++d) {
      const double um = U(-E[d][0], -E[d][1], -E[d][2], n), up = U(E[d][0], E[d][1], E[d][2], n);
      const double dc = 0.5 * (up - um);
      const double df = 2.0 * (up - u0[n]);
      const double db = 2.0 * (u0[n] - um);
      double s = (df * db >= 0.0) ? std::min(std::abs(df), std::abs(db)) : 0.0;
      s = std::copysign(1.0, dc) * std::min(s, std::abs(dc));
      if (dc != 0.0) sf[d] = std::min(sf[d], s / dc);
      sl[d][n] = dc;
    }
  }
  double off[3];
  for (int d = 0; d < 3; ++d) off[d] = (double(c[d] - ic[d] * r) + 0.5) / double(r) - 0.5;
  for (int n = 0; n < NCONS; ++n) {
    const double sx = sl[0][n] * sf[0], sy = sl[1][n] * sf[1], sz = sl[2][n] * sf[2];
    out[n] = u0[n] + off[0] * sx + off[1] * sy + off[2] * sz;
  }
  return 0;
}

// Coarse-fine ghost cells of every fine box (after the same-level copies).
int interp_coarse_fine(const Level& L, std::vector<A4>& v, const CoarseFine& cf0) {
  const orc_geom& g = *L.g;
  OwnerMap fmap, cmap;
  fmap.build(L);
  cmap.build(*cf0.C);
  CoarseFine cf = cf0;
  cf.cmap = &cmap;
  int len[3];
  for (int d = 0; d < 3; ++d) len[d] = g.dom_hi[d] - g.dom_lo[d] + 1;
  int err = 0;
  std::string msg;
#pragma omp parallel for schedule(dynamic) reduction(| : err)
  for (int b = 0; b < L.nbox; ++b) {
    const Box& bx = L.boxes[b];
    Box gb = grow(bx, L.ng);
    for_box(gb, [&](int i, int j, int k) {
      if (err) return;
      const int iv[3] = {i, j, k};
      if (in_box(bx, iv)) return;
      int c[3] = {i, j, k};
      for (int d = 0; d < 3; ++d)
        if (c[d] < g.dom_lo[d] || c[d] > g.dom_hi[d]) {
          if (!g.periodic[d]) return;  // physical boundary
          c[d] += (c[d] < g.dom_lo[d]) ? len[d] : -len[d];
        }
      if (fmap.at(c) >= 0) return;  // covered: copied
      double val[NCONS];
      if (interp_cell(cf, c, val)) {
        err = 1;
#pragma omp critical(cf_error)
        if (msg.empty()) msg = g_err;  // the error text is per thread
        return;
      }
      for (int n = 0; n < NCONS; ++n) v[b](i, j, k, n) = val[n];
    });
  }
  return err ? fail(msg) : 0;
}

// Single-level FillPatch: valid data of every box is taken as given; ghost
// cells get (1) the valid data of the box covering their periodic image
// (AMReX FillBoundary), then (2) physical BCs per non-periodic direction in
// the order x, y, z (src/set/bcs.cpp + AMReX foextrap/reflect fills).
template <class R>
int fill_patch(const Level& L, R* const* fabs, const CoarseFine* cf = nullptr) {
  const orc_geom& g = *L.g;
  int ng = L.ng;
  std::vector<A4T<R>> v(L.nbox);
  for (int b = 0; b < L.nbox; ++b) {
    Box gb = grow(L.boxes[b], ng);
    v[b] = view(fabs[b], gb.lo, gb.hi);
  }
  int len[3];
  for (int d = 0; d < 3; ++d) len[d] = g.dom_hi[d] - g.dom_lo[d] + 1;
  OwnerMap owners;
  owners.build(L);
  // boxes in parallel: each writes only its own ghost cells, reading valid cells
#pragma omp parallel for schedule(dynamic)
  for (int b = 0; b < L.nbox; ++b) {
    const Box& bx = L.boxes[b];
    Box gb = grow(bx, ng);
    int hint = b;
    for_box(gb, [&](int i, int j, int k) {
      if (i >= bx.lo[0] && i <= bx.hi[0] && j >= bx.lo[1] && j <= bx.hi[1] && k >= bx.lo[2] && k <= bx.hi[2])
        return;
      int c[3] = {i, j, k};
      for (int d = 0; d < 3; ++d) {
        if (c[d] < g.dom_lo[d] || c[d] > g.dom_hi[d]) {
          if (!g.periodic[d]) return;  // physical boundary: second pass
          if (c[d] < g.dom_lo[d]) c[d] += len[d];
          else c[d] -= len[d];
        }
      }
      auto inside = [&](const Box& s) {
        return c[0] >= s.lo[0] && c[0] <= s.hi[0] && c[1] >= s.lo[1] && c[1] <= s.hi[1] && c[2] >= s.lo[2] &&
               c[2] <= s.hi[2];
      };
      int src = -1;
      if (inside(L.boxes[hint])) src = hint;
      else
        src = owners.at(c);
      if (src < 0) return;  // uncovered (coarse-fine): left untouched
      hint = src;
      for (int n = 0; n < NCONS; ++n) v[b](i, j, k, n) = v[src](c[0], c[1], c[2], n);
    });
  }
  if (cf) {
    if constexpr (std::is_same<R, double>::value) {
      if (interp_coarse_fine(L, v, *cf)) return -1;
    } else {
      return fail("coarse-fine interpolation is restated for double data only");
    }
  }
  for (int d = 0; d < 3; ++d) {
    if (g.periodic[d]) continue;
    for (int b = 0; b < L.nbox; ++b) {
      Box gb = grow(L.boxes[b], ng);
      // bc code 6: src/set/bcs.cpp:17-87 CnsFillExtDir with the channel's
      // bcnormal (exm/Channel/prob.h:157-187), all components at once
      for (int side = 0; side < 2; ++side) {
        if ((side == 0 ? g.lo_bc[d] : g.hi_bc[d]) != 6) continue;
        if (d != 1) return fail("bc code 6 walls are supported in y only");
        if (!g_prob.set) return fail("bc code 6 walls need orc_set_problem (wall temperature)");
        for_box(gb, [&](int i, int j, int k) {
          int iv[3] = {i, j, k};
          if (side == 0 ? !(iv[d] < g.dom_lo[d]) : !(iv[d] > g.dom_hi[d])) return;
          int firs[3] = {i, j, k};
          firs[d] = side == 0 ? g.dom_lo[d] : g.dom_hi[d];
          R si[NCONS];
          for (int n = 0; n < NCONS; ++n) si[n] = v[b](firs[0], firs[1], firs[2], n);
          const double x = g.prob_lo[d] + static_cast<double>(iv[d] + 0.5) * g.dx[d];
          double di, de;
          if (side == 0) {
            di = static_cast<double>(firs[d] + 0.5) * g.dx[d] - g.prob_lo[d];
            de = g.prob_lo[d] - x;
          } else {
            di = g.prob_hi[d] - static_cast<double>(firs[d] + 0.5) * g.dx[d];
            de = x - g.prob_hi[d];
          }
          const double dratio = de / di;
          R qu = -si[UMX] / si[URHO];
          R qv = -si[UMY] / si[URHO];
          R qw = -si[UMZ] / si[URHO];
          R eint_int = (si[UET] - 0.5 * (si[UMX] * si[UMX] + si[UMY] * si[UMY] + si[UMZ] * si[UMZ]) / si[URHO]) /
                       si[URHO];
          R p_int = (g_prob.gamma - 1.0) * si[URHO] * eint_int;
          R T_int = p_int / (g_prob.Rspec * si[URHO]);
          R qT = max_of(g_prob.Tw + dratio * (g_prob.Tw - T_int), 50.0);
          R qr = p_int / (g_prob.Rspec * qT);
          R ekin = 0.5 * (qu * qu + qv * qv + qw * qw);
          R eint = p_int / (qr * (g_prob.gamma - 1.0));
          v[b](i, j, k, URHO) = qr;
          v[b](i, j, k, UMX) = qr * qu;
          v[b](i, j, k, UMY) = qr * qv;
          v[b](i, j, k, UMZ) = qr * qw;
          v[b](i, j, k, UET) = qr * (eint + ekin);
        });
      }
      for (int n = 0; n < NCONS; ++n) {
        int tlo = g.lo_bc[d] == 6 ? INT_DIR : bc_type(g.lo_bc[d], n, d);
        int thi = g.hi_bc[d] == 6 ? INT_DIR : bc_type(g.hi_bc[d], n, d);
        if (tlo == EXT_DIR || thi == EXT_DIR) return fail("ext_dir boundaries need a problem bcnormal");
        if (tlo < 0 || thi < 0) return fail("invalid physical BC code");
        for_box(gb, [&](int i, int j, int k) {
          int c[3] = {i, j, k};
          int t = -1, src = 0;
          if (c[d] < g.dom_lo[d]) {
            t = tlo;
            src = (t == FOEXTRAP) ? g.dom_lo[d] : 2 * g.dom_lo[d] - 1 - c[d];
          } else if (c[d] > g.dom_hi[d]) {
            t = thi;
            src = (t == FOEXTRAP) ? g.dom_hi[d] : 2 * g.dom_hi[d] + 1 - c[d];
          } else
            return;
          if (t == INT_DIR) return;
          int s[3] = {i, j, k};
          s[d] = src;
          R val = v[b](s[0], s[1], s[2], n);
          v[b](i, j, k, n) = (t == REFLECT_ODD) ? -val : val;
        });
      }
    }
  }
  return 0;
}

// estdt: src/cls/CNS_K.h:12-47 (z momentum is read from UMY, CNS_K.h:29)
double box_estdt(const Box& bx, const A4& st, const double dx[3], const orc_closures& c) {
  double dt = std::numeric_limits<double>::max();
  for_box(bx, [&](int i, int j, int k) {
    double rho = st(i, j, k, URHO);
    double mx = st(i, j, k, UMX);
    double my = st(i, j, k, UMY);
    double mz = st(i, j, k, UMY);
    double rhoinv = 1.0 / std::max(rho, 1.0e-40);
    double vx = mx * rhoinv, vy = my * rhoinv, vz = mz * rhoinv;
    double ke = 0.5 * rho * (vx * vx + vy * vy + vz * vz);
    double rhoei = (st(i, j, k, UET) - ke);
    double p = (c.gamma - 1.0) * rhoei;
    double cs = std::sqrt(c.gamma * p * rhoinv);
    double dtx = dx[0] / (std::abs(vx) + cs);
    double dty = dx[1] / (std::abs(vy) + cs);
    double dtz = dx[2] / (std::abs(vz) + cs);
    dt = std::min(dt, std::min(dtx, std::min(dty, dtz)));
  });
  return dt;
}

// The time-integration sequence of CNS::advance (src/tim/CNS_advance.cpp:51-181)
// over an abstract set of operations.  Vectors: 0 = S1 (old), 1 = S2 (new),
// 2 = dSdt.  ops.rhs(src) evaluates dSdt = L(src) after FillPatch from src.
template <class Ops>
int advance_sequence(Ops& ops, int order_rk, int stages_rk, double dt) {
  enum { S1 = 0, S2 = 1, D = 2 };
  int rc = 0;
  if (order_rk == -2) {
    if ((rc = ops.rhs(S1))) return rc;
    ops.lincomb(S2, 1.0, S1, dt, D);
    if ((rc = ops.rhs(S2))) return rc;
    // LinComb(S2, 0.5, Sborder, 0.5, S1): Sborder holds S2 after FillPatch
    ops.lincomb(S2, 0.5, S2, 0.5, S1);
    ops.saxpy(S2, 0.5 * dt, D);
  } else if (order_rk == 0) {
    if ((rc = ops.rhs(S1))) return rc;
    ops.copy(S2, D);
  } else if (order_rk == 1) {
    if ((rc = ops.rhs(S1))) return rc;
    ops.lincomb(S2, 1.0, S1, dt, D);
  } else if (order_rk == 2) {
    int m = stages_rk;
    if (m < 2) return fail("SSPRK2 number of stages must be >= 2");
    ops.copy(S2, S1);
    for (int i = 1; i <= m - 1; ++i) {
      if ((rc = ops.rhs(S2))) return rc;
      ops.saxpy(S2, dt / double(m - 1), D);
    }
    if ((rc = ops.rhs(S2))) return rc;
    ops.lincomb(S2, double(m - 1), S2, dt, D);
    ops.lincomb(S2, 1.0 / m, S1, 1.0 / m, S2);
  } else if (order_rk == 3 && stages_rk == 3) {
    if ((rc = ops.rhs(S1))) return rc;
    ops.lincomb(S2, 1.0, S1, dt, D);
    if ((rc = ops.rhs(S2))) return rc;
    ops.xpay(D, dt, S2);
    ops.lincomb(S2, 3.0 / 4, S1, 1.0 / 4, D);
    if ((rc = ops.rhs(S2))) return rc;
    ops.xpay(D, dt, S2);
    ops.lincomb(S2, 1.0 / 3, S1, 2.0 / 3, D);
  } else if (order_rk == 3 && stages_rk == 4) {
    if ((rc = ops.rhs(S1))) return rc;
    ops.lincomb(S2, 1.0, S1, dt / 2, D);
    if ((rc = ops.rhs(S2))) return rc;
    ops.saxpy(S2, dt / 2, D);
    if ((rc = ops.rhs(S2))) return rc;
    ops.lincomb(S2, 2.0 / 3, S1, 1.0 / 3, S2);
    ops.saxpy(S2, dt / 6, D);
    if ((rc = ops.rhs(S2))) return rc;
    ops.saxpy(S2, dt / 2, D);
  } else {
    return fail("unsupported order_rk/stages_rk combination");
  }
  return 0;
}

// MultiFab::LinComb / Saxpy / Xpay / Copy on valid cells (AMReX semantics:
// LinComb dst = a*x + b*y; Saxpy dst += a*src; Xpay dst = src + a*dst)
template <class R>
struct LevelOps {
  const Level* L;
  const orc_closures* c;
  const orc_scheme* s;
  double dxinv[3];
  int nthreads;
  std::vector<std::vector<R>> vec[3];  // per vector, per box: valid data
  std::vector<Box> vbox;
  R* const* state;  // ghosted FABs (Sborder storage and final output)
  const double* const* srcs = nullptr;  // per-box source arrays (rhs_source)
  // coarse level of a fine level: coarse-fine ghost cells are interpolated
  // at each stage's FillPatch time tfill[stage] between its states at t0, t1
  const CoarseFine* cf = nullptr;
  std::vector<double> tfill;
  double t0 = 0.0, t1 = 0.0;
  int stage = 0;

  long nval(int b) const { return npts(vbox[b].lo, vbox[b].hi) * NCONS; }

  int rhs(int src) {
    int nb = L->nbox, ng = L->ng;
    // FillPatch(Sborder <- src): copy valid cells into the ghosted FABs
    for (int b = 0; b < nb; ++b) {
      Box gb = grow(vbox[b], ng);
      A4T<R> dst = view(state[b], gb.lo, gb.hi);
      A4T<R> sv = view(vec[src][b].data(), vbox[b].lo, vbox[b].hi);
      for (int n = 0; n < NCONS; ++n)
        for_box(vbox[b], [&](int i, int j, int k) { dst(i, j, k, n) = sv(i, j, k, n); });
    }
    int rc;
    if (cf) {
      CoarseFine w = *cf;
      if (stage >= int(tfill.size())) return fail("internal: more stages than FillPatch times");
      cf_weights(t0, t1, tfill[stage], &w.a, &w.b);
      rc = fill_patch(*L, state, &w);
    } else {
      rc = fill_patch(*L, state);
    }
    ++stage;
    if (rc) return rc;
    int err = 0;
    auto one = [&](int b) {
      Box gb = grow(vbox[b], ng);
      A4T<R> u = view(state[b], gb.lo, gb.hi);
      A4T<R> d = view(vec[2][b].data(), vbox[b].lo, vbox[b].hi);
      Walls w;
      const orc_geom& g = *L->g;
      w.lo = !g.periodic[1] && g.lo_bc[1] == 6 && vbox[b].lo[1] == g.dom_lo[1];
      w.hi = !g.periodic[1] && g.hi_bc[1] == 6 && vbox[b].hi[1] == g.dom_hi[1];
      return box_rhs(*c, *s, vbox[b], ng, u, dxinv, d, srcs ? srcs[b] : nullptr, w) ? 1 : 0;
    };
    if (nb < nthreads) {
      // fewer boxes than threads: boxes in turn, the threads inside each
      g_inner_threads = nthreads;
      for (int b = 0; b < nb; ++b) err |= one(b);
      g_inner_threads = 1;
    } else {
#pragma omp parallel for schedule(dynamic) num_threads(nthreads) reduction(| : err)
      for (int b = 0; b < nb; ++b) err |= one(b);
    }
    return err ? -1 : 0;
  }
  template <class F>
  void each(F&& f) {
    for (int b = 0; b < L->nbox; ++b) {
      long n = nval(b);
      for (long t = 0; t < n; ++t) f(b, t);
    }
  }
  void lincomb(int dst, double a, int x, double bb, int y) {
    each([&](int b, long t) { vec[dst][b][t] = a * vec[x][b][t] + bb * vec[y][b][t]; });
  }
  void saxpy(int dst, double a, int src) {
    each([&](int b, long t) { vec[dst][b][t] += a * vec[src][b][t]; });
  }
  void xpay(int dst, double a, int src) {
    each([&](int b, long t) { vec[dst][b][t] = vec[src][b][t] + a * vec[dst][b][t]; });
  }
  void copy(int dst, int src) {
    each([&](int b, long t) { vec[dst][b][t] = vec[src][b][t]; });
  }
};

struct LinOps {
  double lam;
  std::vector<double> vec[3];
  int rhs(int src) {
    for (size_t t = 0; t < vec[2].size(); ++t) vec[2][t] = lam * vec[src][t];
    return 0;
  }
  void lincomb(int dst, double a, int x, double b, int y) {
    for (size_t t = 0; t < vec[0].size(); ++t) vec[dst][t] = a * vec[x][t] + b * vec[y][t];
  }
  void saxpy(int dst, double a, int src) {
    for (size_t t = 0; t < vec[0].size(); ++t) vec[dst][t] += a * vec[src][t];
  }
  void xpay(int dst, double a, int src) {
    for (size_t t = 0; t < vec[0].size(); ++t) vec[dst][t] = vec[src][t] + a * vec[dst][t];
  }
  void copy(int dst, int src) { vec[dst] = vec[src]; }
};

Level make_level(const orc_geom* g, int nbox, const int* boxes, int ng) {
  Level L;
  L.g = g;
  L.nbox = nbox;
  L.ng = ng;
  L.boxes.resize(nbox);
  for (int b = 0; b < nbox; ++b)
    for (int d = 0; d < 3; ++d) {
      L.boxes[b].lo[d] = boxes[6 * b + d];
      L.boxes[b].hi[d] = boxes[6 * b + 3 + d];
    }
  return L;
}

// FillPatch times of the stages of CNS::advance, as its calls pass them
// (src/tim/CNS_advance.cpp:54,62,72,77,95,101,122,127,133,145,150,155,161).
int fill_times(int order_rk, int stages_rk, double time, double dt, std::vector<double>& t) {
  t.clear();
  if (order_rk == -2) {
    t = {time, time + dt};
  } else if (order_rk == 0 || order_rk == 1) {
    t = {time};
  } else if (order_rk == 2) {
    const int m = stages_rk;
    if (m < 2) return fail("SSPRK2 number of stages must be >= 2");
    for (int i = 1; i <= m - 1; ++i) t.push_back(time + dt * double(i - 1) / double(m - 1));
    t.push_back(time + dt);
  } else if (order_rk == 3 && stages_rk == 3) {
    t = {time, time + dt, time + dt / 2};
  } else if (order_rk == 3 && stages_rk == 4) {
    t = {time, time + dt / 2, time + dt, time + dt / 2};
  } else {
    return fail("unsupported order_rk/stages_rk combination");
  }
  return 0;
}

// One CNS::advance step of a level (valid data in, ghosted FABs as Sborder
// storage; the new state is written back into the valid region).  With `cf`
// the level is a fine level whose coarse level holds states at t0 and t1.
template <class R>
int advance_level(const Level& L, const orc_geom* g, const orc_closures* c, const orc_scheme* s, R* const* state,
                  const double* const* src, double dt, int nthreads, const CoarseFine* cf = nullptr,
                  double time = 0.0, double t0 = 0.0, double t1 = 0.0) {
  const int nbox = L.nbox, ng = L.ng;
  LevelOps<R> ops;
  ops.L = &L;
  ops.c = c;
  ops.s = s;
  ops.nthreads = nthreads > 0 ? nthreads : 1;
  for (int d = 0; d < 3; ++d) ops.dxinv[d] = 1.0 / g->dx[d];
  ops.vbox = L.boxes;
  ops.state = state;
  ops.srcs = src;
  if (cf) {
    ops.cf = cf;
    ops.t0 = t0;
    ops.t1 = t1;
    if (int rc = fill_times(s->order_rk, s->stages_rk, time, dt, ops.tfill)) return rc;
  }
  for (int v = 0; v < 3; ++v) ops.vec[v].resize(nbox);
  for (int b = 0; b < nbox; ++b) {
    long n = npts(L.boxes[b].lo, L.boxes[b].hi) * NCONS;
    for (int v = 0; v < 3; ++v) ops.vec[v][b].assign(n, R(0.0));
    Box gb = grow(L.boxes[b], ng);
    A4T<R> st = view(state[b], gb.lo, gb.hi);
    A4T<R> s1 = view(ops.vec[0][b].data(), L.boxes[b].lo, L.boxes[b].hi);
    for (int nn = 0; nn < NCONS; ++nn)
      for_box(L.boxes[b], [&](int i, int j, int k) { s1(i, j, k, nn) = st(i, j, k, nn); });
  }
  int rc = advance_sequence(ops, s->order_rk, s->stages_rk, dt);
  if (rc) return rc;
  for (int b = 0; b < nbox; ++b) {
    Box gb = grow(L.boxes[b], ng);
    A4T<R> st = view(state[b], gb.lo, gb.hi);
    A4T<R> s2 = view(ops.vec[1][b].data(), L.boxes[b].lo, L.boxes[b].hi);
    for (int nn = 0; nn < NCONS; ++nn)
      for_box(L.boxes[b], [&](int i, int j, int k) { st(i, j, k, nn) = s2(i, j, k, nn); });
  }
  return 0;
}

}  // namespace

// ====================================================================== C API
extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

void orc_set_problem(const orc_closures* c, double wall_T, double body_force_x) {
  g_prob.Tw = wall_T;
  g_prob.fx = body_force_x;
  g_prob.gamma = c ? c->gamma : 1.4;
  g_prob.Rspec = c ? c->Rspec : 287.0;
  g_prob.set = c != nullptr;
}

double orc_visc(const orc_closures* c, double T) { return visc(*c, T); }
double orc_cond(const orc_closures* c, double T) { return cond(*c, T); }

void orc_cons2prim_point(const orc_closures* c, const double u[5], double q[6]) {
  int lo[3] = {0, 0, 0};
  double ub[5], qb[6];
  std::memcpy(ub, u, sizeof ub);
  A4 U = view(ub, lo, lo), Q = view(qb, lo, lo);
  cons2prim(0, 0, 0, U, Q, *c);
  std::memcpy(q, qb, sizeof qb);
}

double orc_weno5js(const double s[5]) { return weno5js(s); }
double orc_limiter(double dlft, double drgt) { return limiter(dlft, drgt); }

void orc_riemann_prob(double gamma, double smallp, double rl, double ul, double pl, double ut1l,
                      double ut2l, double rr, double ur, double pr, double ut1r, double ut2r,
                      double out[5]) {
  riemann_prob(gamma, smallp, rl, ul, pl, ut1l, ut2l, rr, ur, pr, ut1r, ut2r, out[0], out[1], out[2],
               out[3], out[4]);
}

// One face of cns_riemann_{x,y,z} (riemann_point) from given primitives of
// the cells (face-1, face, face+1) along `dir` (q3[cell*6 + comp]) and given
// slopes of the two cells next to the face (dq2[cell*5 + comp], cells
// face-1 and face): the slopes are inputs here, not recomputed, so a test
// can vary one primitive field of one cell alone.  f[5] in the state's
// component order.
void orc_riemann_face(const orc_closures* c, int dir, const double q3[18], const double dq2[10], double f[5]) {
  double qb[18], db[15], fb[15];
  int loq[3] = {0, 0, 0};
  loq[dir] = -1;
  int hiq[3] = {0, 0, 0};
  hiq[dir] = 1;
  A4 Q = view(qb, loq, hiq), D = view(db, loq, hiq), F = view(fb, loq, hiq);
  for (int m = -1; m <= 1; ++m) {
    int ii = dir == 0 ? m : 0, jj = dir == 1 ? m : 0, kk = dir == 2 ? m : 0;
    for (int n = 0; n < 6; ++n) Q(ii, jj, kk, n) = q3[(m + 1) * 6 + n];
    for (int n = 0; n < 5; ++n) D(ii, jj, kk, n) = m <= 0 ? dq2[(m + 1) * 5 + n] : 0.0;
  }
  riemann_point(0, 0, 0, dir, F, D, Q, *c);
  for (int n = 0; n < 5; ++n) f[n] = F(0, 0, 0, n);
}

int orc_box_prims(const orc_closures* c, const int lo[3], const int hi[3], int ng, const double* state,
                  double* prims) {
  Box bx;
  for (int d = 0; d < 3; ++d) {
    bx.lo[d] = lo[d];
    bx.hi[d] = hi[d];
  }
  Box gb = grow(bx, ng);
  A4 u = view(const_cast<double*>(state), gb.lo, gb.hi);
  A4 q = view(prims, gb.lo, gb.hi);
  // every cell of the arrays the caller passed (the box grown by ng)
  for_box(gb, [&](int i, int j, int k) { cons2prim(i, j, k, u, q, *c); });
  return 0;
}

int orc_box_numflux(const orc_closures* c, const orc_scheme* s, const int lo[3], const int hi[3], int ng,
                    const double* state, const double dxinv[3], double* fx, double* fy, double* fz) {
  Box bx;
  for (int d = 0; d < 3; ++d) {
    bx.lo[d] = lo[d];
    bx.hi[d] = hi[d];
  }
  Box gb = grow(bx, ng);
  A4 u = view(const_cast<double*>(state), gb.lo, gb.hi);
  Box bn = grow(nodal_all(bx), ng);
  A4 nf[3] = {view(fx, bn.lo, bn.hi), view(fy, bn.lo, bn.hi), view(fz, bn.lo, bn.hi)};
  return box_flux(*c, *s, bx, ng, u, dxinv, nf);
}

int orc_box_rhs(const orc_closures* c, const orc_scheme* s, const int lo[3], const int hi[3], int ng,
                const double* state, const double dxinv[3], const double* src, double* dsdt) {
  Box bx;
  for (int d = 0; d < 3; ++d) {
    bx.lo[d] = lo[d];
    bx.hi[d] = hi[d];
  }
  Box gb = grow(bx, ng);
  A4 u = view(const_cast<double*>(state), gb.lo, gb.hi);
  A4 d = view(dsdt, bx.lo, bx.hi);
  return box_rhs(*c, *s, bx, ng, u, dxinv, d, src);
}

int orc_fill_patch(const orc_geom* g, int nbox, const int* boxes, int ng, double* const* fabs) {
  Level L = make_level(g, nbox, boxes, ng);
  return fill_patch(L, fabs);
}

int orc_advance(const orc_geom* g, const orc_closures* c, const orc_scheme* s, int nbox, const int* boxes,
                int ng, double* const* state, const double* const* src, double time, double dt, int nthreads) {
  (void)time;
  Level L = make_level(g, nbox, boxes, ng);
  return advance_level(L, g, c, s, state, src, dt, nthreads);
}

void orc_cf_weights(double t0, double t1, double t, double* a, double* b) { cf_weights(t0, t1, t, a, b); }

int orc_fill_two_levels(const orc_geom* gf, int nbf, const int* boxes_f, int ngf, double* const* fabs_f,
                        const orc_geom* gc, int nbc, const int* boxes_c, int ngc, const double* const* crse0,
                        const double* const* crse1, double a, double b, int ratio) {
  if (int rc = check_two_levels(*gf, *gc, ratio)) return rc;
  if ((b != 0.0 && !crse1) || (a != 0.0 && !crse0)) return fail("missing coarse data for nonzero weight");
  Level F = make_level(gf, nbf, boxes_f, ngf);
  Level C = make_level(gc, nbc, boxes_c, ngc);
  CoarseFine cf;
  cf.C = &C;
  cf.u0 = crse0;
  cf.u1 = crse1;
  cf.a = a;
  cf.b = b;
  cf.ratio = ratio;
  return fill_patch(F, fabs_f, &cf);
}

int orc_average_down(int nbf, const int* boxes_f, int ngf, const double* const* fabs_f, int nbc, const int* boxes_c,
                     int ngc, double* const* fabs_c, int ratio) {
  if (ratio != 2 && ratio != 4) return fail("refinement ratio must be 2 or 4");
  Level F = make_level(nullptr, nbf, boxes_f, ngf);
  Level C = make_level(nullptr, nbc, boxes_c, ngc);
  const int r = ratio;
  const double volfrac = 1.0 / double(r * r * r);
  for (const Box& bf : F.boxes)
    for (int d = 0; d < 3; ++d)
      if (bf.lo[d] % r != 0 || (bf.hi[d] + 1) % r != 0) return fail("fine boxes are not coarsenable by the ratio");
  for (int fb = 0; fb < F.nbox; ++fb) {
    const Box& bf = F.boxes[fb];
    Box gf = grow(bf, ngf);
    A4 fv = view(const_cast<double*>(fabs_f[fb]), gf.lo, gf.hi);
    for (int cb = 0; cb < C.nbox; ++cb) {
      const Box& bc = C.boxes[cb];
      Box x;
      bool empty = false;
      for (int d = 0; d < 3; ++d) {
        x.lo[d] = std::max(bf.lo[d] / r, bc.lo[d]);
        x.hi[d] = std::min((bf.hi[d] + 1) / r - 1, bc.hi[d]);
        empty = empty || x.lo[d] > x.hi[d];
      }
      if (empty) continue;
      Box gcb = grow(bc, ngc);
      A4 cv = view(fabs_c[cb], gcb.lo, gcb.hi);
      for (int n = 0; n < NCONS; ++n)
        for_box(x, [&](int i, int j, int k) {
          double c = 0.0;
          for (int kr = 0; kr < r; ++kr)
            for (int jr = 0; jr < r; ++jr)
              for (int ir = 0; ir < r; ++ir) c += fv(i * r + ir, j * r + jr, k * r + kr, n);
          cv(i, j, k, n) = volfrac * c;
        });
    }
  }
  return 0;
}

// Error estimation (CNS::errorEst, src/CNS.cpp:469-507: user_tagging on
// every valid cell of every box of the ghost-filled state) into a byte map
// over the domain, x fastest.
//  kind 1, exm/Sods/prob.h:187-197: drhox = |rho(i+1) - rho(i-1)| / rho(i);
//    drhox > thr tags i, i+1, i+2, i+3 (cells past the domain's hi end are
//    dropped; the reference writes them past its TagBox);
//  kind 2, exm/Channel/prob.h:199-209: y = (j + 0.5) dy; tagged where
//    y < h*factor or y > prob_hi[1] - h*factor;
//  kind 3: |curl u| > thr, u = (rho u)/rho, central differences
//    (u(+1) - u(-1)) * (0.5/dx) per derivative, the magnitude
//    sqrt(wx^2 + wy^2 + wz^2) with wx = dw/dy - dv/dz, wy = du/dz - dw/dx,
//    wz = dv/dx - du/dy (no reference tagger: the criterion of the refined
//    levels of the TGV workload).
int orc_tag_cells(const orc_geom* g, int nbox, const int* boxes, int ng, const double* const* state, int kind,
                  double thr, double h, double factor, unsigned char* tags) {
  if (kind < 1 || kind > 3) return fail("unknown tagger kind");
  if (kind != 2 && ng < 1) return fail("the tagger reads one ghost layer");
  Level L = make_level(g, nbox, boxes, ng);
  int n[3];
  for (int d = 0; d < 3; ++d) n[d] = g->dom_hi[d] - g->dom_lo[d] + 1;
  std::fill(tags, tags + size_t(n[0]) * n[1] * n[2], (unsigned char)0);
  auto at = [&](int i, int j, int k) -> unsigned char& {
    return tags[(size_t(k - g->dom_lo[2]) * n[1] + (j - g->dom_lo[1])) * n[0] + (i - g->dom_lo[0])];
  };
  for (int b = 0; b < L.nbox; ++b) {
    const Box& bx = L.boxes[b];
    Box gb = grow(bx, ng);
    A4 u = view(const_cast<double*>(state[b]), gb.lo, gb.hi);
    for_box(bx, [&](int i, int j, int k) {
      if (kind == 1) {
        const double drhox = std::abs(u(i + 1, j, k, URHO) - u(i - 1, j, k, URHO)) / u(i, j, k, URHO);
        if (drhox > thr)
          for (int m = 0; m < 4; ++m)
            if (i + m <= g->dom_hi[0]) at(i + m, j, k) = 1;
      } else if (kind == 2) {
        const double y = (j + 0.5) * g->dx[1];
        if (y < h * factor || y > g->prob_hi[1] - h * factor) at(i, j, k) = 1;
      } else {
        auto vel = [&](int a, int bb, int c, int m) { return u(a, bb, c, m) / u(a, bb, c, URHO); };
        const double h0 = 0.5 / g->dx[0], h1 = 0.5 / g->dx[1], h2 = 0.5 / g->dx[2];
        const double dwdy = (vel(i, j + 1, k, UMZ) - vel(i, j - 1, k, UMZ)) * h1;
        const double dvdz = (vel(i, j, k + 1, UMY) - vel(i, j, k - 1, UMY)) * h2;
        const double dudz = (vel(i, j, k + 1, UMX) - vel(i, j, k - 1, UMX)) * h2;
        const double dwdx = (vel(i + 1, j, k, UMZ) - vel(i - 1, j, k, UMZ)) * h0;
        const double dvdx = (vel(i + 1, j, k, UMY) - vel(i - 1, j, k, UMY)) * h0;
        const double dudy = (vel(i, j + 1, k, UMX) - vel(i, j - 1, k, UMX)) * h1;
        const double wx = dwdy - dvdz, wy = dudz - dwdx, wz = dvdx - dudy;
        if (std::sqrt(wx * wx + wy * wy + wz * wz) > thr) at(i, j, k) = 1;
      }
    });
  }
  return 0;
}

// CNS::init(AmrLevel& old) (src/CNS.cpp:155-167, FillPatch(old, S_new, 0,
// cur_time)) and, with no old level, CNS::init() (FillCoarsePatch,
// src/CNS.cpp:169-179): every valid cell of the new boxes takes the old
// level's value where an old box holds it, else the coarse level's data
// a*crse0 + b*crse1 interpolated as the coarse-fine ghost cells are.
int orc_regrid_fill(const orc_geom* gf, int nb_new, const int* boxes_new, int ng_new, double* const* fabs_new,
                    int nb_old, const int* boxes_old, int ng_old, const double* const* fabs_old,
                    const orc_geom* gc, int nbc, const int* boxes_c, int ngc, const double* const* crse0,
                    const double* const* crse1, double a, double b, int ratio) {
  if (int rc = check_two_levels(*gf, *gc, ratio)) return rc;
  if ((b != 0.0 && !crse1) || (a != 0.0 && !crse0)) return fail("missing coarse data for nonzero weight");
  Level N = make_level(gf, nb_new, boxes_new, ng_new);
  Level O = make_level(gf, nb_old, boxes_old, ng_old);
  Level C = make_level(gc, nbc, boxes_c, ngc);
  OwnerMap omap, cmap;
  omap.build(O);
  cmap.build(C);
  CoarseFine cf;
  cf.C = &C;
  cf.u0 = crse0;
  cf.u1 = crse1;
  cf.a = a;
  cf.b = b;
  cf.ratio = ratio;
  cf.cmap = &cmap;
  std::vector<A4> ov(O.nbox);
  for (int o = 0; o < O.nbox; ++o) {
    Box gb = grow(O.boxes[o], ng_old);
    ov[o] = view(const_cast<double*>(fabs_old[o]), gb.lo, gb.hi);
  }
  for (int nb = 0; nb < N.nbox; ++nb) {
    const Box& bx = N.boxes[nb];
    Box gb = grow(bx, ng_new);
    A4 v = view(fabs_new[nb], gb.lo, gb.hi);
    int err = 0;
    for_box(bx, [&](int i, int j, int k) {
      if (err) return;
      const int c[3] = {i, j, k};
      const int o = omap.at(c);
      if (o >= 0) {
        for (int n = 0; n < NCONS; ++n) v(i, j, k, n) = ov[o](i, j, k, n);
        return;
      }
      double val[NCONS];
      if (interp_cell(cf, c, val)) {
        err = 1;
        return;
      }
      for (int n = 0; n < NCONS; ++n) v(i, j, k, n) = val[n];
    });
    if (err) return -1;
  }
  return 0;
}

int orc_advance_two_levels(const orc_geom* g, const orc_closures* c, const orc_scheme* s, int nbox, const int* boxes,
                           int ng, double* const* state, const double* const* src, double time, double dt,
                           int nthreads, const orc_geom* gc, int nbc, const int* boxes_c, int ngc,
                           const double* const* crse0, const double* const* crse1, double t0, double t1, int ratio) {
  if (int rc = check_two_levels(*g, *gc, ratio)) return rc;
  if (!crse0 || !crse1) return fail("coarse states at t0 and t1 are required");
  Level L = make_level(g, nbox, boxes, ng);
  Level C = make_level(gc, nbc, boxes_c, ngc);
  CoarseFine cf;
  cf.C = &C;
  cf.u0 = crse0;
  cf.u1 = crse1;
  cf.ratio = ratio;
  return advance_level(L, g, c, s, state, src, dt, nthreads, &cf, time, t0, t1);
}

int orc_box_rhs_tangent(const orc_closures* c, const orc_scheme* s, const int lo[3], const int hi[3], int ng,
                        const double* state, const double* dstate, const double dxinv[3], const double* src,
                        double* dsdt, double* ddsdt) {
  Box bx;
  for (int d = 0; d < 3; ++d) {
    bx.lo[d] = lo[d];
    bx.hi[d] = hi[d];
  }
  Box gb = grow(bx, ng);
  const long ng_pts = npts(gb.lo, gb.hi) * NCONS, nv = npts(bx.lo, bx.hi) * NCONS;
  std::vector<Dual> ub(ng_pts), db(nv);
  for (long t = 0; t < ng_pts; ++t) ub[t] = Dual(state[t], dstate[t]);
  A4T<Dual> u = view(ub.data(), gb.lo, gb.hi);
  A4T<Dual> d = view(db.data(), bx.lo, bx.hi);
  int rc = box_rhs(*c, *s, bx, ng, u, dxinv, d, src);
  if (rc) return rc;
  for (long t = 0; t < nv; ++t) {
    dsdt[t] = db[t].v;
    ddsdt[t] = db[t].d;
  }
  return 0;
}

int orc_fill_patch_tangent(const orc_geom* g, int nbox, const int* boxes, int ng, double* const* fabs,
                           double* const* dfabs) {
  Level L = make_level(g, nbox, boxes, ng);
  std::vector<std::vector<Dual>> buf(nbox);
  std::vector<Dual*> ptr(nbox);
  for (int b = 0; b < nbox; ++b) {
    Box gb = grow(L.boxes[b], ng);
    long n = npts(gb.lo, gb.hi) * NCONS;
    buf[b].resize(n);
    for (long t = 0; t < n; ++t) buf[b][t] = Dual(fabs[b][t], dfabs[b][t]);
    ptr[b] = buf[b].data();
  }
  int rc = fill_patch(L, ptr.data());
  if (rc) return rc;
  for (int b = 0; b < nbox; ++b)
    for (size_t t = 0; t < buf[b].size(); ++t) {
      fabs[b][t] = buf[b][t].v;
      dfabs[b][t] = buf[b][t].d;
    }
  return 0;
}

int orc_advance_tangent(const orc_geom* g, const orc_closures* c, const orc_scheme* s, int nbox, const int* boxes,
                        int ng, double* const* state, double* const* dstate, const double* const* src, double time,
                        double dt, int nthreads) {
  (void)time;
  Level L = make_level(g, nbox, boxes, ng);
  std::vector<std::vector<Dual>> buf(nbox);
  std::vector<Dual*> ptr(nbox);
  for (int b = 0; b < nbox; ++b) {
    Box gb = grow(L.boxes[b], ng);
    long n = npts(gb.lo, gb.hi) * NCONS;
    buf[b].resize(n);
    for (long t = 0; t < n; ++t) buf[b][t] = Dual(state[b][t], dstate[b][t]);
    ptr[b] = buf[b].data();
  }
  int rc = advance_level(L, g, c, s, ptr.data(), src, dt, nthreads);
  if (rc) return rc;
  for (int b = 0; b < nbox; ++b)
    for (size_t t = 0; t < buf[b].size(); ++t) {
      state[b][t] = buf[b][t].v;
      dstate[b][t] = buf[b][t].d;
    }
  return 0;
}

int orc_advance_flops(const orc_geom* g, const orc_closures* c, const orc_scheme* s, int nbox, const int* boxes,
                      int ng, double* const* state, double dt, long counts[5]) {
  Level L = make_level(g, nbox, boxes, ng);
  std::vector<std::vector<Counted>> buf(nbox);
  std::vector<Counted*> ptr(nbox);
  for (int b = 0; b < nbox; ++b) {
    Box gb = grow(L.boxes[b], ng);
    long n = npts(gb.lo, gb.hi) * NCONS;
    buf[b].resize(n);
    for (long t = 0; t < n; ++t) buf[b][t] = Counted(state[b][t]);
    ptr[b] = buf[b].data();
  }
  g_flops = FlopCounts();
  int rc = advance_level(L, g, c, s, ptr.data(), nullptr, dt, 1);
  if (rc) return rc;
  counts[0] = g_flops.add;
  counts[1] = g_flops.mul;
  counts[2] = g_flops.div;
  counts[3] = g_flops.sqrt;
  counts[4] = g_flops.cmp;
  for (int b = 0; b < nbox; ++b)
    for (size_t t = 0; t < buf[b].size(); ++t) state[b][t] = buf[b][t].v;
  return 0;
}

double orc_est_dt(const orc_geom* g, const orc_closures* c, int nbox, const int* boxes, int ng,
                  double* const* state, double cfl) {
  Level L = make_level(g, nbox, boxes, ng);
  double dt = std::numeric_limits<double>::max();
  for (int b = 0; b < nbox; ++b) {
    Box gb = grow(L.boxes[b], ng);
    A4 st = view(state[b], gb.lo, gb.hi);
    dt = std::min(dt, box_estdt(L.boxes[b], st, g->dx, *c));
  }
  return dt * cfl;  // CNS.cpp:391
}

void orc_totals(const orc_geom* g, int nbox, const int* boxes, int ng, double* const* state, double out[5]) {
  Level L = make_level(g, nbox, boxes, ng);
  double probsize = 1.0;
  for (int d = 0; d < 3; ++d) probsize *= (g->dom_hi[d] - g->dom_lo[d] + 1) * g->dx[d];
  for (int n = 0; n < NCONS; ++n) {
    double s = 0.0;
    for (int b = 0; b < nbox; ++b) {
      Box gb = grow(L.boxes[b], ng);
      A4 st = view(state[b], gb.lo, gb.hi);
      for_box(L.boxes[b], [&](int i, int j, int k) { s += st(i, j, k, n); });
    }
    out[n] = s * probsize;  // CNS.cpp:541: sum * geom.ProbSize()
  }
}

int orc_rk_linear(int order_rk, int stages_rk, double lam, double dt, int n, double* u) {
  LinOps ops;
  ops.lam = lam;
  ops.vec[0].assign(u, u + n);
  ops.vec[1].assign(n, 0.0);
  ops.vec[2].assign(n, 0.0);
  int rc = advance_sequence(ops, order_rk, stages_rk, dt);
  if (rc) return rc;
  for (int t = 0; t < n; ++t) u[t] = ops.vec[1][t];
  return 0;
}

}  // extern "C"
