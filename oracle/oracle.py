"""ctypes binding of the CPU oracle (oracle/cns_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "_build", "libcns_oracle.so")


class OClosures(ctypes.Structure):
    _fields_ = [
        ("gamma", ctypes.c_double), ("mw", ctypes.c_double), ("Ru", ctypes.c_double),
        ("cv", ctypes.c_double), ("cp", ctypes.c_double), ("Rspec", ctypes.c_double),
        ("Cshock", ctypes.c_double), ("Cdamp", ctypes.c_double),
        ("visc_law", ctypes.c_int), ("visc_ref", ctypes.c_double), ("Tvisc_ref", ctypes.c_double),
        ("cond_law", ctypes.c_int), ("cond_ref", ctypes.c_double), ("Tcond_ref", ctypes.c_double),
        ("Pr", ctypes.c_double),
    ]


class OScheme(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "rhs_euler", "rhs_visc", "rhs_source", "flux_euler", "order_keep", "art_diss",
        "order_rk", "stages_rk")]


class OGeom(ctypes.Structure):
    _fields_ = [
        ("dom_lo", ctypes.c_int * 3), ("dom_hi", ctypes.c_int * 3), ("periodic", ctypes.c_int * 3),
        ("lo_bc", ctypes.c_int * 3), ("hi_bc", ctypes.c_int * 3),
        ("dx", ctypes.c_double * 3), ("prob_lo", ctypes.c_double * 3), ("prob_hi", ctypes.c_double * 3),
    ]


def _copy_struct(dst_cls, src):
    d = dst_cls()
    for name, _ in dst_cls._fields_:
        v = getattr(src, name)
        if isinstance(v, ctypes.Array):
            arr = getattr(d, name)
            for i in range(len(v)):
                arr[i] = v[i]
        else:
            setattr(d, name, v)
    return d


def closures(c):
    return _copy_struct(OClosures, c)


def scheme(s):
    return _copy_struct(OScheme, s)


def geom(g):
    return _copy_struct(OGeom, g)


_lib = None
_DP = ctypes.POINTER(ctypes.c_double)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load(variant=None):
    """The oracle library; variant "x86v3" is the same code built for
    x86-64-v3 hosts (CPU-baseline timing only)."""
    global _lib
    if variant is None and _lib is not None:
        return _lib
    path = LIBPATH if variant is None else os.path.join(HERE, "_build", f"libcns_oracle_{variant}.so")
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    I3 = ctypes.c_int * 3
    lib.orc_last_error.restype = ctypes.c_char_p
    lib.orc_visc.restype = ctypes.c_double
    lib.orc_visc.argtypes = [ctypes.POINTER(OClosures), ctypes.c_double]
    lib.orc_cond.restype = ctypes.c_double
    lib.orc_cond.argtypes = [ctypes.POINTER(OClosures), ctypes.c_double]
    lib.orc_cons2prim_point.argtypes = [ctypes.POINTER(OClosures), _DP, _DP]
    lib.orc_weno5js.restype = ctypes.c_double
    lib.orc_weno5js.argtypes = [_DP]
    lib.orc_limiter.restype = ctypes.c_double
    lib.orc_limiter.argtypes = [ctypes.c_double, ctypes.c_double]
    lib.orc_riemann_prob.argtypes = [ctypes.c_double] * 12 + [_DP]
    lib.orc_riemann_face.restype = None
    lib.orc_riemann_face.argtypes = [ctypes.POINTER(OClosures), ctypes.c_int, _DP, _DP, _DP]
    lib.orc_box_prims.argtypes = [ctypes.POINTER(OClosures), I3, I3, ctypes.c_int, _DP, _DP]
    lib.orc_box_numflux.argtypes = [ctypes.POINTER(OClosures), ctypes.POINTER(OScheme), I3, I3, ctypes.c_int,
                                    _DP, _DP, _DP, _DP, _DP]
    lib.orc_box_rhs.argtypes = [ctypes.POINTER(OClosures), ctypes.POINTER(OScheme), I3, I3, ctypes.c_int,
                                _DP, _DP, _DP, _DP]
    PP = ctypes.POINTER(_DP)
    lib.orc_fill_patch.argtypes = [ctypes.POINTER(OGeom), ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                   ctypes.c_int, PP]
    lib.orc_advance.argtypes = [ctypes.POINTER(OGeom), ctypes.POINTER(OClosures), ctypes.POINTER(OScheme),
                                ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int, PP, PP,
                                ctypes.c_double, ctypes.c_double, ctypes.c_int]
    lib.orc_est_dt.restype = ctypes.c_double
    lib.orc_est_dt.argtypes = [ctypes.POINTER(OGeom), ctypes.POINTER(OClosures), ctypes.c_int,
                               ctypes.POINTER(ctypes.c_int), ctypes.c_int, PP, ctypes.c_double]
    lib.orc_totals.argtypes = [ctypes.POINTER(OGeom), ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                               ctypes.c_int, PP, _DP]
    lib.orc_set_problem.restype = None
    lib.orc_set_problem.argtypes = [ctypes.POINTER(OClosures), ctypes.c_double, ctypes.c_double]
    lib.orc_rk_linear.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_int, _DP]
    lib.orc_box_rhs_tangent.argtypes = [ctypes.POINTER(OClosures), ctypes.POINTER(OScheme), I3, I3, ctypes.c_int,
                                        _DP, _DP, _DP, _DP, _DP, _DP]
    lib.orc_fill_patch_tangent.argtypes = [ctypes.POINTER(OGeom), ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                           ctypes.c_int, PP, PP]
    lib.orc_advance_flops.argtypes = [ctypes.POINTER(OGeom), ctypes.POINTER(OClosures), ctypes.POINTER(OScheme),
                                      ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int, PP,
                                      ctypes.c_double, ctypes.POINTER(ctypes.c_long)]
    lib.orc_advance_tangent.argtypes = [ctypes.POINTER(OGeom), ctypes.POINTER(OClosures),
                                        ctypes.POINTER(OScheme), ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                        ctypes.c_int, PP, PP, PP, ctypes.c_double, ctypes.c_double, ctypes.c_int]
    IP = ctypes.POINTER(ctypes.c_int)
    lib.orc_cf_weights.restype = None
    lib.orc_cf_weights.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _DP, _DP]
    lib.orc_fill_two_levels.argtypes = [ctypes.POINTER(OGeom), ctypes.c_int, IP, ctypes.c_int, PP,
                                        ctypes.POINTER(OGeom), ctypes.c_int, IP, ctypes.c_int, PP, PP,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_int]
    lib.orc_average_down.argtypes = [ctypes.c_int, IP, ctypes.c_int, PP, ctypes.c_int, IP, ctypes.c_int, PP,
                                     ctypes.c_int]
    lib.orc_advance_two_levels.argtypes = [ctypes.POINTER(OGeom), ctypes.POINTER(OClosures),
                                           ctypes.POINTER(OScheme), ctypes.c_int, IP, ctypes.c_int, PP, PP,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                           ctypes.POINTER(OGeom), ctypes.c_int, IP, ctypes.c_int, PP, PP,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_int]
    lib.orc_tag_cells.argtypes = [ctypes.POINTER(OGeom), ctypes.c_int, IP, ctypes.c_int, PP, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.POINTER(ctypes.c_ubyte)]
    lib.orc_regrid_fill.argtypes = [ctypes.POINTER(OGeom), ctypes.c_int, IP, ctypes.c_int, PP,
                                    ctypes.c_int, IP, ctypes.c_int, PP,
                                    ctypes.POINTER(OGeom), ctypes.c_int, IP, ctypes.c_int, PP, PP,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_int]
    if variant is None:
        _lib = lib
    return lib


def _ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_DP)


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + load().orc_last_error().decode())


# ------------------------------------------------------------------ points


def visc(c, T):
    return load().orc_visc(ctypes.byref(closures(c)), T)


def cond(c, T):
    return load().orc_cond(ctypes.byref(closures(c)), T)


def cons2prim(c, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    q = np.zeros(6)
    load().orc_cons2prim_point(ctypes.byref(closures(c)), _ptr(u), _ptr(q))
    return q


def weno5js(s):
    s = np.ascontiguousarray(s, dtype=np.float64)
    return load().orc_weno5js(_ptr(s))


def limiter(a, b):
    return load().orc_limiter(a, b)


def riemann_face(c, d, q3, dq2):
    """Face flux of cns_riemann_{x,y,z} (d = 0, 1, 2) from primitives q3
    (3 cells along d: face-1, face, face+1; 6 components each) and the
    slopes dq2 of the two cells next to the face (5 components each)."""
    q3 = np.ascontiguousarray(q3, dtype=np.float64).reshape(18)
    dq2 = np.ascontiguousarray(dq2, dtype=np.float64).reshape(10)
    out = np.zeros(5)
    load().orc_riemann_face(ctypes.byref(closures(c)), d, _ptr(q3), _ptr(dq2), _ptr(out))
    return out


def riemann_prob(gamma, rl, ul, pl, ut1l, ut2l, rr, ur, pr, ut1r, ut2r, smallp=1.0e-40):
    out = np.zeros(5)
    load().orc_riemann_prob(gamma, smallp, rl, ul, pl, ut1l, ut2l, rr, ur, pr, ut1r, ut2r, _ptr(out))
    return out


# ------------------------------------------------------------------ boxes
# A box FAB is a numpy array (ncomp, nz+2ng, ny+2ng, nx+2ng) (i fastest).


def box_rhs(c, s, lo, hi, ng, state, dxinv, src=None):
    """dSdt of one box; src: (5, nz, ny, nx) source on the valid box, added
    when the scheme has rhs_source = 1."""
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    out = np.zeros((5, n[2], n[1], n[0]))
    I3 = ctypes.c_int * 3
    dx = np.ascontiguousarray(dxinv, dtype=np.float64)
    sp = None if src is None else np.ascontiguousarray(src, dtype=np.float64)
    _check(load().orc_box_rhs(ctypes.byref(closures(c)), ctypes.byref(scheme(s)), I3(*lo), I3(*hi), ng,
                              _ptr(np.ascontiguousarray(state)), _ptr(dx), None if sp is None else _ptr(sp),
                              _ptr(out)))
    return out


def box_rhs_tangent(c, s, lo, hi, ng, state, dstate, dxinv, src=None):
    """(dSdt, its derivative along dstate) of one box by dual numbers
    (KEEP and viscous right-hand sides)."""
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    out = np.zeros((5, n[2], n[1], n[0]))
    dout = np.zeros_like(out)
    I3 = ctypes.c_int * 3
    dx = np.ascontiguousarray(dxinv, dtype=np.float64)
    sp = None if src is None else np.ascontiguousarray(src, dtype=np.float64)
    _check(load().orc_box_rhs_tangent(ctypes.byref(closures(c)), ctypes.byref(scheme(s)), I3(*lo), I3(*hi), ng,
                                      _ptr(np.ascontiguousarray(state)), _ptr(np.ascontiguousarray(dstate)),
                                      _ptr(dx), None if sp is None else _ptr(sp), _ptr(out), _ptr(dout)))
    return out, dout


def box_numflux(c, s, lo, hi, ng, state, dxinv):
    n = [hi[d] - lo[d] + 2 + 2 * ng for d in range(3)]
    f = [np.zeros((5, n[2], n[1], n[0])) for _ in range(3)]
    I3 = ctypes.c_int * 3
    dx = np.ascontiguousarray(dxinv, dtype=np.float64)
    _check(load().orc_box_numflux(ctypes.byref(closures(c)), ctypes.byref(scheme(s)), I3(*lo), I3(*hi), ng,
                                  _ptr(np.ascontiguousarray(state)), _ptr(dx), _ptr(f[0]), _ptr(f[1]),
                                  _ptr(f[2])))
    return f


def box_prims(c, lo, hi, ng, state):
    q = np.zeros((6,) + state.shape[1:])
    I3 = ctypes.c_int * 3
    _check(load().orc_box_prims(ctypes.byref(closures(c)), I3(*lo), I3(*hi), ng,
                                _ptr(np.ascontiguousarray(state)), _ptr(q)))
    return q


# ------------------------------------------------------------------ level


class OracleLevel:
    """Boxes with one ghosted FAB each (AMReX MultiFab layout) on the CPU."""

    def __init__(self, g, boxes, ng=3, lib=None):
        self.lib = lib  # None: the default build
        self.g = geom(g)
        self.boxes = [(tuple(lo), tuple(hi)) for lo, hi in boxes]
        self.ng = ng
        self.n = [g.dom_hi[d] - g.dom_lo[d] + 1 for d in range(3)]
        self.barr = (ctypes.c_int * (6 * len(boxes)))()
        for b, (lo, hi) in enumerate(self.boxes):
            for d in range(3):
                self.barr[6 * b + d] = lo[d]
                self.barr[6 * b + 3 + d] = hi[d]
        self.fabs = []
        for lo, hi in self.boxes:
            m = [hi[d] - lo[d] + 1 + 2 * ng for d in range(3)]
            self.fabs.append(np.zeros((5, m[2], m[1], m[0])))
        # tangent FABs (direction of a forward-mode derivative), same layout
        self.tfabs = [np.zeros_like(f) for f in self.fabs]

    def _pp(self, fabs=None):
        fabs = self.fabs if fabs is None else fabs
        arr = (_DP * len(fabs))()
        for b, f in enumerate(fabs):
            arr[b] = _ptr(f)
        return arr

    def set_state(self, u, fabs=None):
        """u: (5, nz, ny, nx) domain array of valid cells."""
        g = self.ng
        for (lo, hi), f in zip(self.boxes, self.fabs if fabs is None else fabs):
            f[:, g:f.shape[1] - g, g:f.shape[2] - g, g:f.shape[3] - g] = \
                u[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]

    def get_state(self, fabs=None):
        g = self.ng
        out = np.zeros((5, self.n[2], self.n[1], self.n[0]))
        for (lo, hi), f in zip(self.boxes, self.fabs if fabs is None else fabs):
            out[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = \
                f[:, g:f.shape[1] - g, g:f.shape[2] - g, g:f.shape[3] - g]
        return out

    def fill_patch(self):
        _check(load().orc_fill_patch(ctypes.byref(self.g), len(self.boxes), self.barr, self.ng, self._pp()))

    def _src(self, source):
        """Per-box source arrays from a (5, nz, ny, nx) domain array."""
        if source is None:
            return None, None
        keep = [np.ascontiguousarray(source[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1])
                for lo, hi in self.boxes]
        arr = (_DP * len(keep))()
        for b, a in enumerate(keep):
            arr[b] = _ptr(a)
        return arr, keep

    def advance(self, c, s, dt, time=0.0, nthreads=1, source=None):
        sp, keep = self._src(source)
        _check((self.lib or load()).orc_advance(ctypes.byref(self.g), ctypes.byref(closures(c)),
                                                ctypes.byref(scheme(s)), len(self.boxes), self.barr, self.ng,
                                                self._pp(), sp, time, dt, nthreads))
        del keep

    # --- tangent (forward mode, dual numbers; KEEP and viscous schemes) ---
    def set_tangent(self, v):
        """Direction v on the valid cells; ghost cells of the tangent are
        zero (coarse-fine ghosts are constants) until a fill sets them."""
        for t in self.tfabs:
            t[...] = 0.0
        self.set_state(v, self.tfabs)

    def get_tangent(self):
        return self.get_state(self.tfabs)

    def fill_patch_tangent(self):
        _check(load().orc_fill_patch_tangent(ctypes.byref(self.g), len(self.boxes), self.barr, self.ng,
                                             self._pp(), self._pp(self.tfabs)))

    def advance_tangent(self, c, s, dt, time=0.0, nthreads=1, source=None):
        """One step of the state and of its tangent (the derivative of the
        step along the current tangent)."""
        sp, keep = self._src(source)
        _check(load().orc_advance_tangent(ctypes.byref(self.g), ctypes.byref(closures(c)),
                                          ctypes.byref(scheme(s)), len(self.boxes), self.barr, self.ng,
                                          self._pp(), self._pp(self.tfabs), sp, time, dt, nthreads))
        del keep

    def rhs_tangent(self, c, s, source=None):
        """(dSdt, its derivative along the tangent) after a FillPatch of
        both."""
        self.fill_patch_tangent()
        dx = [1.0 / self.g.dx[d] for d in range(3)]
        out = np.zeros((5, self.n[2], self.n[1], self.n[0]))
        dout = np.zeros_like(out)
        for (lo, hi), f, t in zip(self.boxes, self.fabs, self.tfabs):
            src = None if source is None else source[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
            sl = (slice(None), slice(lo[2], hi[2] + 1), slice(lo[1], hi[1] + 1), slice(lo[0], hi[0] + 1))
            out[sl], dout[sl] = box_rhs_tangent(c, s, lo, hi, self.ng, f, t, dx, src)
        return out, dout

    def advance_flops(self, c, s, dt):
        """One step on an operation-counting scalar (one thread); returns the
        counts {add, mul, div, sqrt, cmp} of FP64 operations."""
        cnt = (ctypes.c_long * 5)()
        _check(load().orc_advance_flops(ctypes.byref(self.g), ctypes.byref(closures(c)), ctypes.byref(scheme(s)),
                                        len(self.boxes), self.barr, self.ng, self._pp(), dt, cnt))
        return dict(zip(("add", "mul", "div", "sqrt", "cmp"), list(cnt)))

    def est_dt(self, c, cfl):
        return load().orc_est_dt(ctypes.byref(self.g), ctypes.byref(closures(c)), len(self.boxes), self.barr,
                                 self.ng, self._pp(), cfl)

    def totals(self):
        out = np.zeros(5)
        load().orc_totals(ctypes.byref(self.g), len(self.boxes), self.barr, self.ng, self._pp(), _ptr(out))
        return out

    def rhs(self, c, s, source=None):
        """dSdt of every box after a FillPatch of the current state."""
        self.fill_patch()
        dx = [1.0 / self.g.dx[d] for d in range(3)]
        out = np.zeros((5, self.n[2], self.n[1], self.n[0]))
        for (lo, hi), f in zip(self.boxes, self.fabs):
            src = None if source is None else source[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
            out[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = box_rhs(c, s, lo, hi, self.ng, f, dx, src)
        return out


    # --- coarse-fine coupling (this level is the fine one) ---
    def fill_two_levels(self, coarse, crse0, crse1, a, b, ratio=2):
        """FillPatch of a fine level: same-level copies, coarse-fine ghost
        cells interpolated from a*crse0 + b*crse1 (lists of ghosted coarse
        FABs, ghost cells filled), physical BCs."""
        _check(load().orc_fill_two_levels(ctypes.byref(self.g), len(self.boxes), self.barr, self.ng, self._pp(),
                                          ctypes.byref(coarse.g), len(coarse.boxes), coarse.barr, coarse.ng,
                                          None if crse0 is None else coarse._pp(crse0),
                                          None if crse1 is None else coarse._pp(crse1), a, b, ratio))

    def tag_cells(self, kind, threshold=0.0, h=0.0, factor=0.0):
        """CNS::errorEst on this level's ghost-filled FABs: a (nz, ny, nx)
        uint8 map over the domain (kinds as amr_tagger_t: 1 density
        gradient, 2 wall band, 3 vorticity magnitude)."""
        tags = np.zeros((self.n[2], self.n[1], self.n[0]), dtype=np.uint8)
        _check(load().orc_tag_cells(ctypes.byref(self.g), len(self.boxes), self.barr, self.ng, self._pp(), kind,
                                    threshold, h, factor, tags.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))))
        return tags

    def regridded(self, boxes, coarse, crse0, crse1, a, b, ratio=2):
        """A level on `boxes` whose valid cells are filled as CNS::init(old)
        fills a regridded level: this level's value where one of its boxes
        holds the cell, else interpolated from a*crse0 + b*crse1 (ghosted
        coarse FABs, ghost cells filled)."""
        new = OracleLevel(self.g, boxes, self.ng)
        _check(load().orc_regrid_fill(ctypes.byref(self.g), len(new.boxes), new.barr, new.ng, new._pp(),
                                      len(self.boxes), self.barr, self.ng, self._pp(),
                                      ctypes.byref(coarse.g), len(coarse.boxes), coarse.barr, coarse.ng,
                                      None if crse0 is None else coarse._pp(crse0),
                                      None if crse1 is None else coarse._pp(crse1), a, b, ratio))
        return new

    def average_down(self, coarse, ratio=2):
        """Coarse cells under this level's boxes := mean of their children."""
        _check(load().orc_average_down(len(self.boxes), self.barr, self.ng, self._pp(), len(coarse.boxes),
                                       coarse.barr, coarse.ng, coarse._pp(), ratio))

    def advance_two_levels(self, c, s, dt, time, coarse, crse0, crse1, t0, t1, ratio=2, nthreads=1, source=None):
        """One CNS::advance step of a fine level whose every FillPatch takes
        the coarse-fine ghost cells from the coarse states at t0 and t1."""
        sp, keep = self._src(source)
        _check(load().orc_advance_two_levels(ctypes.byref(self.g), ctypes.byref(closures(c)),
                                             ctypes.byref(scheme(s)), len(self.boxes), self.barr, self.ng,
                                             self._pp(), sp, time, dt, nthreads, ctypes.byref(coarse.g),
                                             len(coarse.boxes), coarse.barr, coarse.ng, coarse._pp(crse0),
                                             coarse._pp(crse1), t0, t1, ratio))
        del keep


def cf_weights(t0, t1, t):
    """Weights (a, b) of the coarse states at t0 and t1 at fill time t."""
    a, b = ctypes.c_double(), ctypes.c_double()
    load().orc_cf_weights(t0, t1, t, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


class OracleHierarchy:
    """Levels 0..L of one AMR hierarchy with refinement ratio r, advanced as
    AMReX's Amr::timeStep does with subcycling: level l steps by dt_l, then
    level l+1 takes r steps of dt_l / r at time + (i - 1) dt_{l+1}, then
    CNS::post_timestep averages level l+1 down onto level l (src/CNS.cpp:398-412,
    513-530; refluxing adds nothing there, its registers are never filled).
    Every FillPatch of a fine level interpolates its coarse-fine ghost cells
    from the coarse level's states at the start and end of the coarse step."""

    def __init__(self, levels, ratio=2):
        self.levels = [OracleLevel(g, boxes) for g, boxes in levels]
        self.ratio = ratio
        self.old = [None] * len(self.levels)    # ghosted FABs at the step's start
        self.tint = [(0.0, 0.0)] * len(self.levels)

    def set_state(self, us):
        for L, u in zip(self.levels, us):
            L.set_state(u)

    def get_state(self, lev):
        return self.levels[lev].get_state()

    def _fill(self, lev, t):
        L = self.levels[lev]
        if lev == 0:
            L.fill_patch()
            return
        C = self.levels[lev - 1]
        t0, t1 = self.tint[lev - 1]
        a, b = cf_weights(t0, t1, t)
        L.fill_two_levels(C, self.old[lev - 1], C.fabs, a, b, self.ratio)

    def _step(self, lev, c, s, time, dt, nthreads):
        L = self.levels[lev]
        self._fill(lev, time)
        self.old[lev] = [f.copy() for f in L.fabs]
        if lev == 0:
            L.advance(c, s, dt, time, nthreads=nthreads)
        else:
            C = self.levels[lev - 1]
            t0, t1 = self.tint[lev - 1]
            L.advance_two_levels(c, s, dt, time, C, self.old[lev - 1], C.fabs, t0, t1, self.ratio, nthreads)
        self.tint[lev] = (time, time + dt)
        if lev + 1 < len(self.levels):
            self._fill(lev, time + dt)
            dtf = dt / self.ratio
            for i in range(1, self.ratio + 1):
                self._step(lev + 1, c, s, time + (i - 1) * dtf, dtf, nthreads)
            self.levels[lev + 1].average_down(L, self.ratio)

    def advance(self, c, s, time, dt, nthreads=1):
        """One coarse step of dt (level l steps by dt / r^l)."""
        self._step(0, c, s, time, dt, nthreads)

    def regrid(self, lev, boxes):
        """Level lev+1 rebuilt on `boxes` at the time the hierarchy has
        reached (every level at the end of its last step): CNS::init(old)
        from the old level and level lev's new data."""
        C = self.levels[lev]
        C.fill_patch() if lev == 0 else self._fill(lev, self.tint[lev][1])
        self.levels[lev + 1] = self.levels[lev + 1].regridded(boxes, C, None, C.fabs, 0.0, 1.0, self.ratio)


def set_problem(c, wall_T=0.0, body_force_x=0.0):
    """Problem parameters (bc 6 wall temperature, x body force) and the
    closures of the wall fill, for every later oracle call; c None clears."""
    load().orc_set_problem(None if c is None else ctypes.byref(closures(c)), wall_T, body_force_x)


def rk_linear(order_rk, stages_rk, lam, dt, u):
    u = np.array(u, dtype=np.float64)
    _check(load().orc_rk_linear(order_rk, stages_rk, lam, dt, len(u), _ptr(u)))
    return u
