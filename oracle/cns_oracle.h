/*
 * CPU oracle for the per-box RHS + SSPRK hot path of amrsolver.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain restatement of the reference
 * algorithm (src/rhs, src/cls, src/tim/CNS_advance.cpp) used as the
 * checker for the HIP product path.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product library never links
 * or calls it.
 *
 * Layout of every array ("FAB"): Fortran order, i fastest, then j, then k,
 * component outermost (AMReX FArrayBox / Array4 convention).  A FAB covers
 * the box [lo-ng, hi+ng] in every direction.
 */
#ifndef CNS_ORACLE_H_
#define CNS_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double gamma, mw, Ru, cv, cp, Rspec, Cshock, Cdamp;
  int visc_law;          /* 0 constant, 1 Sutherland                     */
  double visc_ref, Tvisc_ref;
  int cond_law;          /* 0 constant, 1 Sutherland, 2 mu*cp/Pr        */
  double cond_ref, Tcond_ref, Pr;
} orc_closures;

typedef struct {
  int rhs_euler, rhs_visc, rhs_source;
  int flux_euler;        /* 0 MUSCL+Riemann, 1 KEEP, 2 WENO5-JS + LF     */
  int order_keep;        /* 2, 4 or 6                                    */
  int art_diss;          /* must be 0                                    */
  int order_rk, stages_rk;
} orc_scheme;

typedef struct {
  int dom_lo[3], dom_hi[3];
  int periodic[3];
  int lo_bc[3], hi_bc[3]; /* reference physical BC codes 0..6 (CNS_setup.cpp:18-31) */
  double dx[3];
  double prob_lo[3];
  double prob_hi[3];
} orc_geom;

/* Problem parameters (exm/Channel/prob.h ProbParm): bc 6 wall temperature
 * and x body force of the channel's user_source, with the closures its
 * bcnormal uses.  Process-wide; set before fills and steps that need them. */
void orc_set_problem(const orc_closures* c, double wall_T, double body_force_x);

/* ---- point functions ---- */
double orc_visc(const orc_closures* c, double T);
double orc_cond(const orc_closures* c, double T);
void   orc_cons2prim_point(const orc_closures* c, const double u[5], double q[6]);
double orc_weno5js(const double s[5]);
double orc_limiter(double dlft, double drgt);
void   orc_riemann_face(const orc_closures* c, int dir, const double q3[18], const double dq2[10], double f[5]);
void   orc_riemann_prob(double gamma, double smallp,
                        double rl, double ul, double pl, double ut1l, double ut2l,
                        double rr, double ur, double pr, double ut1r, double ut2r,
                        double out[5]);

/* ---- box level (state FAB grown by ng >= 3, dsdt FAB on the valid box) ---- */
int orc_box_prims(const orc_closures* c, const int lo[3], const int hi[3], int ng,
                  const double* state, double* prims);
/* numerical face fluxes (Euler + viscous as the scheme selects); each of
   fx, fy, fz is a FAB on the box [lo-ng, hi+1+ng] (nodal in every direction,
   ng ghosts).  This is the content of numflxmf after the flux phase of
   compute_rhs, before assembly. */
int orc_box_numflux(const orc_closures* c, const orc_scheme* s, const int lo[3],
                    const int hi[3], int ng, const double* state, const double dxinv[3],
                    double* fx, double* fy, double* fz);
/* src: NULL, or a source array on the valid box added to dSdt when
   rhs_source = 1 (the user_source hook) */
int orc_box_rhs(const orc_closures* c, const orc_scheme* s, const int lo[3],
                const int hi[3], int ng, const double* state, const double dxinv[3],
                const double* src, double* dsdt);

/* ---- level (list of boxes, one ghosted FAB per box) ---- */
int orc_fill_patch(const orc_geom* g, int nbox, const int* boxes, int ng,
                   double* const* fabs);
int orc_advance(const orc_geom* g, const orc_closures* c, const orc_scheme* s,
                int nbox, const int* boxes, int ng, double* const* state,
                const double* const* src, double time, double dt, int nthreads);
double orc_est_dt(const orc_geom* g, const orc_closures* c, int nbox,
                  const int* boxes, int ng, double* const* state, double cfl);
void orc_totals(const orc_geom* g, int nbox, const int* boxes, int ng,
                double* const* state, double out[5]);
/* ---- coarse-fine coupling (AMReX FillPatchTwoLevels with lincc_interp,
   average_down) ---- */
void orc_cf_weights(double t0, double t1, double t, double* a, double* b);
int orc_fill_two_levels(const orc_geom* gf, int nbf, const int* boxes_f, int ngf,
                        double* const* fabs_f, const orc_geom* gc, int nbc,
                        const int* boxes_c, int ngc, const double* const* crse0,
                        const double* const* crse1, double a, double b, int ratio);
int orc_average_down(int nbf, const int* boxes_f, int ngf, const double* const* fabs_f,
                     int nbc, const int* boxes_c, int ngc, double* const* fabs_c,
                     int ratio);
int orc_tag_cells(const orc_geom* g, int nbox, const int* boxes, int ng, const double* const* state, int kind,
                  double thr, double h, double factor, unsigned char* tags);
int orc_regrid_fill(const orc_geom* gf, int nb_new, const int* boxes_new, int ng_new, double* const* fabs_new,
                    int nb_old, const int* boxes_old, int ng_old, const double* const* fabs_old,
                    const orc_geom* gc, int nbc, const int* boxes_c, int ngc, const double* const* crse0,
                    const double* const* crse1, double a, double b, int ratio);
int orc_advance_two_levels(const orc_geom* g, const orc_closures* c, const orc_scheme* s,
                           int nbox, const int* boxes, int ng, double* const* state,
                           const double* const* src, double time, double dt, int nthreads,
                           const orc_geom* gc, int nbc, const int* boxes_c, int ngc,
                           const double* const* crse0, const double* const* crse1,
                           double t0, double t1, int ratio);
/* generic SSPRK driver on a linear test ODE dU/dt = lam*U (one scalar per
   entry), used to pin the stage coefficients against stability polynomials */
int orc_rk_linear(int order_rk, int stages_rk, double lam, double dt, int n,
                  double* u);

/* Tangent (forward-mode, dual numbers) of the KEEP/viscous path: the same
 * operations as the plain functions on (value, derivative) pairs.  dstate is
 * the direction (ghosted like state); dsdt/ddsdt receive L(state) and
 * dL(state)[dstate].  The fill and advance variants update state and
 * dstate in place. */
int orc_box_rhs_tangent(const orc_closures* c, const orc_scheme* s, const int lo[3],
                        const int hi[3], int ng, const double* state, const double* dstate,
                        const double dxinv[3], const double* src, double* dsdt, double* ddsdt);
int orc_fill_patch_tangent(const orc_geom* g, int nbox, const int* boxes, int ng,
                           double* const* fabs, double* const* dfabs);
int orc_advance_tangent(const orc_geom* g, const orc_closures* c, const orc_scheme* s,
                        int nbox, const int* boxes, int ng, double* const* state,
                        double* const* dstate, const double* const* src, double time,
                        double dt, int nthreads);
/* One step of the level (as orc_advance, one thread) on an operation-counting
 * scalar: counts[] = FP64 adds/subtracts, multiplies, divides, square roots
 * and comparisons of the reference arithmetic. */
int orc_advance_flops(const orc_geom* g, const orc_closures* c, const orc_scheme* s,
                      int nbox, const int* boxes, int ng, double* const* state, double dt,
                      long counts[5]);
const char* orc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
