// cns_driver: a C++ host program over the C ABI alone (no Python, no torch)
// that runs the Taylor-Green vortex of exm/TGV on [0, 2 pi]^3 on one or more
// AMR levels, as amrsolver's main loop drives CNS: Amr::timeStep with
// subcycling (level l steps by dt / 2^l, each finer level takes two steps
// per step of the level below, then CNS::post_timestep averages it down,
// src/CNS.cpp:398-412), each level running the stage loop of CNS::advance
// with a FillPatch before every stage (src/tim/CNS_advance.cpp:54-161).
//
//   cns_driver [--n N] [--steps K] [--dt DT] [--stages 3|4] [--grids FILE]
//              [--init DIR] [--plot DIR] [--device D]
//
// --grids: lines "level lo0 lo1 lo2 hi0 hi1 hi2" (level 0 domain N^3,
// level l domain N*2^l); default one level of boxes of at most 64^3.
// --init: DIR/level_<l>.bin holds each level's valid cells over its whole
// domain as doubles in (component, z, y, x) order; default the TGV field.
// --plot: an AMReX plotfile of the final state (state and derived variables).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "amrsolver_hip.h"

namespace {

void check(int rc, const char* what) {
  if (rc != AMR_OK) {
    std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, amr_last_error());
    std::exit(1);
  }
}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    std::exit(1);
  }
}

// calorifically_perfect_gas_t (src/cls/Closures.h:72-82) with a Sutherland
// viscosity scaled so that the TGV set-up (Ma 0.1, Re 1600, T0 110.4, L 1)
// has rho0 = 1, and conductivity mu cp / Pr (Pr 0.71)
amr_closures_t tgv_closures() {
  amr_closures_t c{};
  c.gamma = 1.40;
  c.mw = 28.96e-3;
  c.Ru = 8.31451;
  c.cv = c.Ru / (c.mw * (c.gamma - 1.0));
  c.cp = c.gamma * c.Ru / (c.mw * (c.gamma - 1.0));
  c.Rspec = c.Ru / c.mw;
  c.Cshock = 1.0;
  c.Cdamp = 0.01;
  const double T0 = 110.4, mach = 0.1, Re = 1600.0, L = 1.0, rho0 = 1.0;
  const double cs = std::sqrt(c.gamma * c.Rspec * T0);
  const double v0 = mach * cs;
  const double mu0 = rho0 * L * v0 / Re;
  c.visc_law = 1;
  c.Tvisc_ref = 110.4;
  c.visc_ref = mu0 * (c.Tvisc_ref + T0) / (T0 * std::sqrt(T0));
  c.cond_law = 2;
  c.cond_ref = 0.0;
  c.Tcond_ref = 194.0;
  c.Pr = 0.71;
  return c;
}

struct Level {
  amr_geom_t geom{};
  std::vector<int32_t> boxes;  // lo[3], hi[3] per box
  amr_boxtab_t bt = nullptr;
  int nbox() const { return int(boxes.size() / 6); }
};

// one ghosted device array per box
struct State {
  std::vector<amr_fab_t> fabs;
  std::vector<double*> mem;
};

State alloc_state(const Level& L, int ncomp, int ng) {
  State s;
  for (int b = 0; b < L.nbox(); ++b) {
    amr_fab_t f{};
    int64_t n[3];
    for (int d = 0; d < 3; ++d) {
      f.lo[d] = L.boxes[6 * b + d];
      f.hi[d] = L.boxes[6 * b + 3 + d];
      n[d] = f.hi[d] - f.lo[d] + 1 + 2 * ng;
    }
    f.ng = ng;
    f.jstride = n[0];
    f.kstride = n[0] * n[1];
    f.nstride = n[0] * n[1] * n[2];
    double* p = nullptr;
    hip_check(hipMalloc(&p, sizeof(double) * f.nstride * ncomp), "hipMalloc");
    hip_check(hipMemset(p, 0, sizeof(double) * f.nstride * ncomp), "hipMemset");
    f.data = p;
    s.fabs.push_back(f);
    s.mem.push_back(p);
  }
  return s;
}

void free_state(State& s) {
  for (double* p : s.mem) (void)hipFree(p);
  s.mem.clear();
  s.fabs.clear();
}

// valid cells of every box from a (5, nz, ny, nx) domain array
void upload(const Level& L, State& s, const std::vector<double>& dom) {
  const int N0 = L.geom.dom_hi[0] + 1, N1 = L.geom.dom_hi[1] + 1, N2 = L.geom.dom_hi[2] + 1;
  for (int b = 0; b < L.nbox(); ++b) {
    const amr_fab_t& f = s.fabs[b];
    std::vector<double> h(size_t(f.nstride) * 5, 0.0);
    for (int n = 0; n < 5; ++n)
      for (int k = f.lo[2]; k <= f.hi[2]; ++k)
        for (int j = f.lo[1]; j <= f.hi[1]; ++j)
          for (int i = f.lo[0]; i <= f.hi[0]; ++i)
            h[size_t(n) * f.nstride + (k - f.lo[2] + f.ng) * f.kstride + (j - f.lo[1] + f.ng) * f.jstride +
              (i - f.lo[0] + f.ng)] = dom[((size_t(n) * N2 + k) * N1 + j) * N0 + i];
    hip_check(hipMemcpy(f.data, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice), "upload");
  }
}

// exm/TGV/prob.h:92-134 (omega = 1, not convecting) on the whole domain
std::vector<double> tgv_field(const amr_closures_t& c, const amr_geom_t& g) {
  const int N0 = g.dom_hi[0] + 1, N1 = g.dom_hi[1] + 1, N2 = g.dom_hi[2] + 1;
  const double T0 = 110.4, mach = 0.1, Re = 1600.0, L = 1.0;
  const double cs = std::sqrt(c.gamma * c.Rspec * T0), v0 = mach * cs;
  const double mu = c.visc_ref * T0 * std::sqrt(T0) / (c.Tvisc_ref + T0);
  const double rho0 = Re * mu / (L * v0), p0 = rho0 * c.Rspec * T0;
  std::vector<double> u(size_t(5) * N0 * N1 * N2);
  const size_t NN = size_t(N0) * N1 * N2;
  for (int k = 0; k < N2; ++k)
    for (int j = 0; j < N1; ++j)
      for (int i = 0; i < N0; ++i) {
        const double x = g.prob_lo[0] + (i + 0.5) * g.dx[0], y = g.prob_lo[1] + (j + 0.5) * g.dx[1],
                     z = g.prob_lo[2] + (k + 0.5) * g.dx[2];
        const double ux = v0 * std::sin(x / L) * std::cos(y / L) * std::cos(z / L);
        const double uy = -v0 * std::cos(x / L) * std::sin(y / L) * std::cos(z / L);
        const double p = p0 + rho0 * v0 * v0 / 16.0 * (std::cos(2.0 * x / L) + std::cos(2.0 * y / L)) *
                                  (std::cos(2.0 * z / L) + 2.0);
        const double rho = p / (c.Rspec * T0);
        const size_t t = (size_t(k) * N1 + j) * N0 + i;
        u[t] = rho;
        u[NN + t] = rho * ux;
        u[2 * NN + t] = rho * uy;
        u[3 * NN + t] = 0.0;
        u[4 * NN + t] = rho * (c.cv * T0 + 0.5 * (ux * ux + uy * uy));
      }
  return u;
}

// FillPatch times of CNS::advance's stages for SSPRK(3,3) and (3,4)
std::vector<double> fill_times(int stages, double time, double dt) {
  if (stages == 3) return {time, time + dt, time + dt / 2};
  return {time, time + dt / 2, time + dt, time + dt / 2};
}

// coarse-level weights at fill time t (AMReX StateData::getData rules)
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

struct Hierarchy {
  std::vector<Level> lev;
  std::vector<State> cur, old, tmp;
  std::vector<double> t0, t1;
  std::vector<int> ngrow;  // ghost layers each level's stages read (0: not yet known)
  amr_closures_t cls;
  amr_scheme_t sch;

  void fill(int l, const State& S, double t) {
    if (l == 0) {
      check(amr_fill_boundary(lev[0].bt, S.fabs.data()), "amr_fill_boundary");
      return;
    }
    if (ngrow[l] == 0) {  // the fused stage kernel reads two ghost layers, the others three
      int32_t info[6];
      check(amr_step_plan(lev[l].bt, &sch, S.fabs.data(), info), "amr_step_plan");
      ngrow[l] = info[3] != 0 ? 2 : 3;
    }
    double a, b;
    cf_weights(t0[l - 1], t1[l - 1], t, &a, &b);
    check(amr_fill_two_levels(lev[l].bt, lev[l - 1].bt, 2, a, old[l - 1].fabs.data(), b, cur[l - 1].fabs.data(),
                              ngrow[l], S.fabs.data()),
          "amr_fill_two_levels");
  }

  void step(int l, double time, double dt) {
    const std::vector<double> times = fill_times(sch.stages_rk, time, dt);
    State A = cur[l], B = old[l], C = tmp[l];
    const State* sin = &A;
    const State* out = nullptr;
    for (int s = 0; s < int(times.size()); ++s) {
      fill(l, *sin, times[s]);
      int nst, mode;
      double a, b, c, d;
      check(amr_stage_coeffs(sch.order_rk, sch.stages_rk, dt, s, &nst, &mode, &a, &b, &c, &d), "amr_stage_coeffs");
      out = (s % 2 == 0) ? &B : &C;
      check(amr_rk_stage(lev[l].bt, &cls, &sch, mode, a, b, c, d, A.fabs.data(), sin->fabs.data(), out->fabs.data()),
            "amr_rk_stage");
      sin = out;
    }
    old[l] = A;
    cur[l] = *out;
    tmp[l] = (out == &B) ? C : B;
    t0[l] = time;
    t1[l] = time + dt;
    if (l + 1 < int(lev.size())) {
      fill(l, cur[l], time + dt);
      const double dtf = dt / 2;
      for (int i = 1; i <= 2; ++i) step(l + 1, time + (i - 1) * dtf, dtf);
      check(amr_average_down(lev[l + 1].bt, lev[l].bt, 2, cur[l + 1].fabs.data(), cur[l].fabs.data()),
            "amr_average_down");
    }
  }
};

}  // namespace

int main(int argc, char** argv) {
  int n = 64, steps = 10, stages = 3, device = 0;
  double dt = -1.0;
  std::string grids, init, plot;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "missing value after %s\n", a.c_str());
        std::exit(2);
      }
      return argv[++i];
    };
    if (a == "--n") n = std::atoi(next().c_str());
    else if (a == "--steps") steps = std::atoi(next().c_str());
    else if (a == "--dt") dt = std::atof(next().c_str());
    else if (a == "--stages") stages = std::atoi(next().c_str());
    else if (a == "--grids") grids = next();
    else if (a == "--init") init = next();
    else if (a == "--plot") plot = next();
    else if (a == "--device") device = std::atoi(next().c_str());
    else {
      std::fprintf(stderr, "unknown argument %s\n", a.c_str());
      return 2;
    }
  }
  if (n < 8 || steps < 0 || (stages != 3 && stages != 4)) {
    std::fprintf(stderr, "need --n >= 8, --steps >= 0, --stages 3 or 4\n");
    return 2;
  }
  const double twopi = 2.0 * std::acos(-1.0);
  if (dt <= 0.0) dt = 5.0e-5 * 256.0 / n;

  Hierarchy H;
  H.cls = tgv_closures();
  H.sch = amr_scheme_t{};
  H.sch.rhs_euler = 1;
  H.sch.rhs_visc = 1;
  H.sch.flux_euler = 1;
  H.sch.order_keep = 4;
  H.sch.order_rk = 3;
  H.sch.stages_rk = stages;

  // levels and boxes
  std::vector<std::vector<int32_t>> lb;
  if (!grids.empty()) {
    std::ifstream f(grids);
    if (!f) {
      std::fprintf(stderr, "cannot read %s\n", grids.c_str());
      return 2;
    }
    int l;
    int32_t v[6];
    while (f >> l >> v[0] >> v[1] >> v[2] >> v[3] >> v[4] >> v[5]) {
      if (l < 0 || l > 8) {
        std::fprintf(stderr, "bad level %d in %s\n", l, grids.c_str());
        return 2;
      }
      if (int(lb.size()) <= l) lb.resize(l + 1);
      lb[l].insert(lb[l].end(), v, v + 6);
    }
  } else {
    lb.resize(1);
    const int m = n < 64 ? n : 64;
    for (int k = 0; k < n; k += m)
      for (int j = 0; j < n; j += m)
        for (int i = 0; i < n; i += m) {
          const int32_t v[6] = {i, j, k, std::min(i + m, n) - 1, std::min(j + m, n) - 1, std::min(k + m, n) - 1};
          lb[0].insert(lb[0].end(), v, v + 6);
        }
  }
  const int nlev = int(lb.size());
  for (int l = 0; l < nlev; ++l)
    if (lb[l].empty()) {
      std::fprintf(stderr, "level %d has no boxes\n", l);
      return 2;
    }

  hip_check(hipSetDevice(device), "hipSetDevice");
  hipStream_t stream;
  hip_check(hipStreamCreate(&stream), "hipStreamCreate");
  amr_ctx_t ctx = nullptr;
  check(amr_ctx_create(device, stream, &ctx), "amr_ctx_create");
  H.lev.resize(nlev);
  for (int l = 0; l < nlev; ++l) {
    Level& L = H.lev[l];
    const int N = n << l;
    for (int d = 0; d < 3; ++d) {
      L.geom.dom_lo[d] = 0;
      L.geom.dom_hi[d] = N - 1;
      L.geom.periodic[d] = 1;
      L.geom.prob_lo[d] = 0.0;
      L.geom.prob_hi[d] = twopi;
      L.geom.dx[d] = twopi / N;
    }
    L.boxes = lb[l];
    check(amr_boxtab_create(ctx, &L.geom, L.nbox(), L.boxes.data(), &L.bt), "amr_boxtab_create");
    H.cur.push_back(alloc_state(L, 5, 3));
    H.old.push_back(alloc_state(L, 5, 3));
    H.tmp.push_back(alloc_state(L, 5, 3));
    std::vector<double> dom;
    if (!init.empty()) {
      const std::string fn = init + "/level_" + std::to_string(l) + ".bin";
      std::ifstream f(fn, std::ios::binary);
      dom.resize(size_t(5) * N * N * N);
      if (!f.read(reinterpret_cast<char*>(dom.data()), std::streamsize(sizeof(double) * dom.size()))) {
        std::fprintf(stderr, "cannot read %s\n", fn.c_str());
        return 2;
      }
    } else {
      dom = tgv_field(H.cls, L.geom);
    }
    upload(L, H.cur[l], dom);
  }
  H.t0.assign(nlev, 0.0);
  H.t1.assign(nlev, 0.0);
  H.ngrow.assign(nlev, 0);

  double time = 0.0;
  hip_check(hipStreamSynchronize(stream), "sync");
  const auto c0 = std::chrono::steady_clock::now();
  for (int s = 0; s < steps; ++s) {
    H.step(0, time, dt);
    time += dt;
  }
  hip_check(hipStreamSynchronize(stream), "sync");
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
  int64_t cell_stages = 0;
  for (int l = 0; l < nlev; ++l) {
    int64_t cells = 0;
    for (int b = 0; b < H.lev[l].nbox(); ++b) {
      const int32_t* v = &H.lev[l].boxes[6 * b];
      cells += int64_t(v[3] - v[0] + 1) * (v[4] - v[1] + 1) * (v[5] - v[2] + 1);
    }
    cell_stages += cells * stages * (int64_t(1) << l) * steps;
  }
  double tot[5];
  check(amr_totals(H.lev[0].bt, H.cur[0].fabs.data(), tot), "amr_totals");
  std::printf("{\"levels\": %d, \"steps\": %d, \"dt\": %.17g, \"time\": %.17g, \"seconds\": %.6f, "
              "\"cell_stage_per_s\": %.6e, \"totals_level0\": [%.17g, %.17g, %.17g, %.17g, %.17g]}\n",
              nlev, steps, dt, time, secs, secs > 0 ? cell_stages / secs : 0.0, tot[0], tot[1], tot[2], tot[3],
              tot[4]);

  if (!plot.empty()) {
    const char* names[10] = {"density", "xmom", "ymom", "zmom", "energy", "pressure", "temperature",
                             "x_velocity", "y_velocity", "z_velocity"};
    std::vector<State> pv;
    std::vector<amr_boxtab_t> bts;
    std::vector<const amr_fab_t*> pf;
    std::vector<int32_t> lsteps, ratio(nlev > 1 ? nlev - 1 : 1, 2);
    for (int l = 0; l < nlev; ++l) {
      pv.push_back(alloc_state(H.lev[l], 10, 0));
      check(amr_plot_vars(H.lev[l].bt, &H.cls, H.cur[l].fabs.data(), 1, 31, pv[l].fabs.data()), "amr_plot_vars");
      bts.push_back(H.lev[l].bt);
      lsteps.push_back(steps << l);
    }
    for (int l = 0; l < nlev; ++l) pf.push_back(pv[l].fabs.data());
    check(amr_write_plotfile(plot.c_str(), nlev, bts.data(), pf.data(), 10, names, time, lsteps.data(),
                             ratio.data()),
          "amr_write_plotfile");
    for (State& s : pv) free_state(s);
  }
  for (int l = 0; l < nlev; ++l) {
    free_state(H.cur[l]);
    free_state(H.old[l]);
    free_state(H.tmp[l]);
    check(amr_boxtab_destroy(H.lev[l].bt), "amr_boxtab_destroy");
  }
  check(amr_ctx_destroy(ctx), "amr_ctx_destroy");
  hip_check(hipStreamDestroy(stream), "hipStreamDestroy");
  return 0;
}
