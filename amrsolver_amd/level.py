"""Host-side mirror of the reference's CNS level interface over the C ABI.

``Level`` plays the role of the CNS AmrLevel (src/CNS.h): it owns the box
table (BoxArray + post_regrid workspace) and offers ``advance`` (CNS::advance),
``compute_rhs`` (CNS::compute_rhs), ``fill_patch`` (FillPatch), ``est_time_step``
(CNS::estTimeStep) and ``totals`` (CNS::printTotal sums).  Device memory
comes from torch (plumbing only); all arithmetic runs in the HIP library.
"""
import ctypes

import numpy as np

from . import capi
from .capi import Fab, call

_torch = None


def _torch_mod():
    global _torch
    if _torch is None:
        import torch
        _torch = torch
    return _torch


class Context:
    """amr_ctx_t bound to torch's current stream on `device`."""

    def __init__(self, device=0, stream=None):
        torch = _torch_mod()
        if not torch.cuda.is_available():
            raise capi.AmrError("no GPU visible: the HIP path needs a device")
        torch.cuda.set_device(device)
        self.device = device
        if stream is None:
            stream = torch.cuda.current_stream(device)
            if stream.cuda_stream == 0:
                # one explicit stream shared by torch's copies and the library
                stream = torch.cuda.Stream(device=device)
                torch.cuda.set_stream(stream)
        self.stream = stream
        h = ctypes.c_void_p()
        call("amr_ctx_create", device, ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(h))
        self.h = h

    def synchronize(self):
        call("amr_ctx_synchronize", self.h)

    def close(self):
        if getattr(self, "h", None):
            capi.load().amr_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


LEVEL_PAD = 16  # x offset of cell 0 in a level-wide array's rows (elements)


class StateArrays:
    """Device arrays of one MultiFab: either one level-wide array (boxes are
    windows into it) or one array per box."""

    def __init__(self, level, ncomp=5, ng=3, layout="level"):
        torch = _torch_mod()
        self.level = level
        self.ng = ng
        self.layout = layout
        self.tensors = []
        nb = len(level.boxes)
        self.fabs = (Fab * nb)()
        dev = torch.device("cuda", level.ctx.device)
        if layout == "level":
            N = level.n
            # rows padded to a multiple of 16 elements with x = 0 at element
            # LEVEL_PAD of a row: the rows of 16-cell tiles start on 128-byte
            # lines (the allocation is 512-byte aligned)
            if ng > LEVEL_PAD:
                raise ValueError(f"ng {ng} > {LEVEL_PAD}")
            rowlen = -(-(LEVEL_PAD + N[0] + ng) // 16) * 16
            store = torch.zeros((ncomp, N[2] + 2 * ng, N[1] + 2 * ng, rowlen), dtype=torch.float64, device=dev)
            t = store[:, :, :, LEVEL_PAD - ng:LEVEL_PAD + N[0] + ng]  # (ncomp, nz+2g, ny+2g, nx+2g) view
            self.tensors.append(t)
            JS, KS, NS = store.stride(2), store.stride(1), store.stride(0)
            base = t.data_ptr()
            for b, (lo, hi) in enumerate(level.boxes):
                f = self.fabs[b]
                for d in range(3):
                    f.lo[d], f.hi[d] = lo[d], hi[d]
                f.ng = ng
                f.pad_ = 0
                off = lo[0] + lo[1] * JS + lo[2] * KS  # (lo-ng)+ng in every direction
                f.data = base + 8 * off
                f.jstride, f.kstride, f.nstride = JS, KS, NS
        elif layout == "boxes":
            # one compact ghosted array per box, carved from one allocation
            # as an AMReX arena hands out a MultiFab's FABs (each array
            # starts on a 512-byte boundary)
            shapes = []
            for lo, hi in level.boxes:
                n = [hi[d] - lo[d] + 1 for d in range(3)]
                shapes.append((ncomp, n[2] + 2 * ng, n[1] + 2 * ng, n[0] + 2 * ng))
            sizes = [int(np.prod(sh)) for sh in shapes]
            offs, total = [], 0
            for sz in sizes:
                offs.append(total)
                total += -(-sz // 64) * 64
            self.arena = torch.zeros(max(total, 1), dtype=torch.float64, device=dev)
            for b, (lo, hi) in enumerate(level.boxes):
                shape = shapes[b]
                t = self.arena[offs[b]:offs[b] + sizes[b]].view(shape)
                self.tensors.append(t)
                f = self.fabs[b]
                for d in range(3):
                    f.lo[d], f.hi[d] = lo[d], hi[d]
                f.ng = ng
                f.pad_ = 0
                f.data = t.data_ptr()
                f.jstride = shape[3]
                f.kstride = shape[3] * shape[2]
                f.nstride = shape[3] * shape[2] * shape[1]
        else:
            raise ValueError(layout)

    def upload(self, u):
        """u: (ncomp, nz, ny, nx) valid-cell array of the whole domain."""
        torch = _torch_mod()
        g = self.ng
        if self.layout == "level":
            t = self.tensors[0]
            t[:, g:t.shape[1] - g, g:t.shape[2] - g, g:t.shape[3] - g] = torch.from_numpy(np.ascontiguousarray(u)).to(t.device)
        else:
            for (lo, hi), t in zip(self.level.boxes, self.tensors):
                sub = u[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
                t[:, g:t.shape[1] - g, g:t.shape[2] - g, g:t.shape[3] - g] = torch.from_numpy(np.ascontiguousarray(sub)).to(t.device)
        torch.cuda.synchronize()

    def upload_tensor(self, u):
        """u: (ncomp, nz, ny, nx) device tensor of the whole domain's valid
        cells (no host round trip)."""
        g = self.ng
        if self.layout == "level":
            t = self.tensors[0]
            t[:, g:t.shape[1] - g, g:t.shape[2] - g, g:t.shape[3] - g] = u
        else:
            for (lo, hi), t in zip(self.level.boxes, self.tensors):
                t[:, g:t.shape[1] - g, g:t.shape[2] - g, g:t.shape[3] - g] = \
                    u[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]

    def download_tensor(self):
        """Valid cells of every box as one (ncomp, nz, ny, nx) device tensor
        (zero where no box covers), after the context's work."""
        torch = _torch_mod()
        self.level.ctx.synchronize()
        g = self.ng
        N = self.level.n
        if self.layout == "level":
            t = self.tensors[0]
            return t[:, g:t.shape[1] - g, g:t.shape[2] - g, g:t.shape[3] - g].clone()
        out = torch.zeros((self.tensors[0].shape[0], N[2], N[1], N[0]), dtype=torch.float64,
                          device=self.tensors[0].device)
        for (lo, hi), t in zip(self.level.boxes, self.tensors):
            out[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = \
                t[:, g:t.shape[1] - g, g:t.shape[2] - g, g:t.shape[3] - g]
        return out

    def upload_ghosted(self, arr):
        """The whole level array including every ghost cell,
        (ncomp, N2+2g, N1+2g, N0+2g) (level layout only); arr may be a numpy
        array or a device tensor."""
        torch = _torch_mod()
        if self.layout != "level":
            raise ValueError("upload_ghosted needs the level layout")
        t = self.tensors[0]
        if tuple(arr.shape) != tuple(t.shape):
            raise ValueError(f"shape {tuple(arr.shape)} != {tuple(t.shape)}")
        if isinstance(arr, np.ndarray):
            arr = torch.from_numpy(np.ascontiguousarray(arr))
        t.copy_(arr)

    def upload_box(self, b, arr):
        """Whole ghosted array of box b, (ncomp, nz+2g, ny+2g, nx+2g)
        (separate-box layout only)."""
        torch = _torch_mod()
        if self.layout != "boxes":
            raise ValueError("upload_box needs the separate-box layout")
        t = self.tensors[b]
        if tuple(arr.shape) != tuple(t.shape):
            raise ValueError(f"box {b}: shape {arr.shape} != {tuple(t.shape)}")
        t.copy_(torch.from_numpy(np.ascontiguousarray(arr)))

    def download(self):
        """Valid cells of every box as one (ncomp, nz, ny, nx) numpy array."""
        torch = _torch_mod()
        self.level.ctx.synchronize()
        torch.cuda.synchronize()
        g = self.ng
        N = self.level.n
        if self.layout == "level":
            t = self.tensors[0]
            return t[:, g:t.shape[1] - g, g:t.shape[2] - g, g:t.shape[3] - g].cpu().numpy().copy()
        out = np.zeros((self.tensors[0].shape[0], N[2], N[1], N[0]))
        for (lo, hi), t in zip(self.level.boxes, self.tensors):
            out[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = \
                t[:, g:t.shape[1] - g, g:t.shape[2] - g, g:t.shape[3] - g].cpu().numpy()
        return out

    def download_box(self, b):
        """Whole ghosted array of box b, (ncomp, nz+2g, ny+2g, nx+2g)."""
        torch = _torch_mod()
        self.level.ctx.synchronize()
        torch.cuda.synchronize()
        g = self.ng
        lo, hi = self.level.boxes[b]
        if self.layout == "boxes":
            return self.tensors[b].cpu().numpy().copy()
        t = self.tensors[0]
        return t[:, lo[2]:hi[2] + 1 + 2 * g, lo[1]:hi[1] + 1 + 2 * g, lo[0]:hi[0] + 1 + 2 * g].cpu().numpy().copy()


class FaceArrays:
    """Face-flux arrays of a level (AMReX's face-centred numflxmf): per
    direction d one array per box over its faces normal to d (the box with
    hi + 1 in d)."""

    def __init__(self, level, ncomp=5):
        torch = _torch_mod()
        dev = torch.device("cuda", level.ctx.device)
        nb = len(level.boxes)
        self.tensors = [[], [], []]
        self.fabs = [(Fab * nb)() for _ in range(3)]
        for d in range(3):
            for b, (lo, hi) in enumerate(level.boxes):
                n = [hi[e] - lo[e] + 1 + (1 if e == d else 0) for e in range(3)]
                t = torch.zeros((ncomp, n[2], n[1], n[0]), dtype=torch.float64, device=dev)
                self.tensors[d].append(t)
                f = self.fabs[d][b]
                for e in range(3):
                    f.lo[e], f.hi[e] = lo[e], hi[e] + (1 if e == d else 0)
                f.ng = 0
                f.pad_ = 0
                f.data = t.data_ptr()
                f.jstride, f.kstride, f.nstride = n[0], n[0] * n[1], n[0] * n[1] * n[2]

    def download(self, d, b):
        """Faces of box b normal to d, (ncomp, nz, ny, nx) with one more in d."""
        return self.tensors[d][b].cpu().numpy().copy()


class Level:
    """One AMR level on one GPU (mirror of the CNS AmrLevel's hot path)."""

    def __init__(self, geom, boxes, closures, scheme, ctx=None, device=0):
        self.ctx = ctx if ctx is not None else Context(device)
        self.geom = geom
        self.boxes = [(tuple(lo), tuple(hi)) for lo, hi in boxes]
        self.n = [geom.dom_hi[d] - geom.dom_lo[d] + 1 for d in range(3)]
        self.closures = closures
        self.scheme = scheme
        arr = (ctypes.c_int32 * (6 * len(self.boxes)))()
        for b, (lo, hi) in enumerate(self.boxes):
            for d in range(3):
                arr[6 * b + d] = lo[d]
                arr[6 * b + 3 + d] = hi[d]
        h = ctypes.c_void_p()
        call("amr_boxtab_create", self.ctx.h, ctypes.byref(geom), len(self.boxes), arr, ctypes.byref(h))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            capi.load().amr_boxtab_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def new_state(self, layout="level", ng=3, ncomp=5):
        return StateArrays(self, ncomp=ncomp, ng=ng, layout=layout)

    # --- reference interface -------------------------------------------
    def fill_patch(self, S):
        call("amr_fill_boundary", self.h, S.fabs)

    def face_fluxes(self, Sborder, F):
        """numflxmf of compute_rhs for the ghost-filled state into the
        FaceArrays F."""
        call("amr_face_fluxes", self.h, ctypes.byref(self.closures), ctypes.byref(self.scheme), Sborder.fabs,
             F.fabs[0], F.fabs[1], F.fabs[2])

    def compute_rhs(self, Sborder, dSdt):
        call("amr_rhs", self.h, ctypes.byref(self.closures), ctypes.byref(self.scheme), Sborder.fabs, dSdt.fabs)

    def rk_stage(self, mode, a, b, c, d, S1, Sin, Sout):
        call("amr_rk_stage", self.h, ctypes.byref(self.closures), ctypes.byref(self.scheme), mode, a, b, c, d,
             S1.fabs if S1 is not None else None, Sin.fabs, Sout.fabs)

    def advance(self, S, time, dt):
        call("amr_advance", self.h, ctypes.byref(self.closures), ctypes.byref(self.scheme), S.fabs, time, dt)

    # --- coarse-fine coupling (this level is the fine one) ---------------
    def fill_two_levels(self, S, coarse, crse_old, crse_new, a, b, ratio=2, ngrow=3, same_level=True):
        """FillPatch of a fine level (its `ngrow` inner ghost layers):
        same-level copies, coarse-fine ghost cells interpolated from
        a*crse_old + b*crse_new (states of `coarse` with filled ghost cells),
        physical BCs.  same_level=False leaves out the same-level copies
        where the level's fused stages read the other boxes' cells directly
        (amr_fill_two_levels_cf)."""
        name = "amr_fill_two_levels" if same_level else "amr_fill_two_levels_cf"
        call(name, self.h, coarse.h, ratio, a, None if crse_old is None else crse_old.fabs, b,
             None if crse_new is None else crse_new.fabs, ngrow, S.fabs)

    def plot_vars(self, S, out, state_vars=True, derive=()):
        """Plot variables of S into out (valid cells): the state components
        if state_vars, then the derived variables named in `derive`, in the
        order CNS registers them (DERIVED_NAMES)."""
        call("amr_plot_vars", self.h, ctypes.byref(self.closures), S.fabs, int(bool(state_vars)),
             derive_mask(derive), out.fabs)

    def average_down(self, S, coarse, Sc, ratio=2):
        """CNS::avgDown: coarse cells of Sc under this level := mean of S."""
        call("amr_average_down", self.h, coarse.h, ratio, S.fabs, Sc.fabs)

    def tag_cells(self, S, tagger, out=None):
        """CNS::errorEst on the ghost-filled state S: a (nz, ny, nx) uint8
        device tensor over the domain, 1 where `tagger` (a capi.Tagger)
        tags the cell."""
        torch = _torch_mod()
        if out is None:
            out = torch.empty((self.n[2], self.n[1], self.n[0]), dtype=torch.uint8,
                              device=torch.device("cuda", self.ctx.device))
        call("amr_tag_cells", self.h, ctypes.byref(tagger), S.fabs, ctypes.c_void_p(out.data_ptr()))
        return out

    def tag_blocks(self, tags, cb):
        """The tag map reduced to blocks of cb^3 cells: a (nb0, nb1, nb2)
        numpy bool array indexed [bi, bj, bk]."""
        torch = _torch_mod()
        nb = [self.n[d] // cb for d in range(3)]
        out = torch.empty((nb[2], nb[1], nb[0]), dtype=torch.uint8, device=tags.device)
        call("amr_tag_blocks", self.h, ctypes.c_void_p(tags.data_ptr()), cb, ctypes.c_void_p(out.data_ptr()))
        self.ctx.synchronize()
        return out.cpu().numpy().astype(bool).transpose(2, 1, 0)

    def regrid_fill(self, S, old, S_old, coarse, crse_old, crse_new, a, b, ratio=2):
        """CNS::init(old) of this (regridded) level: S's valid cells from
        the old level (`old`, `S_old`; None: none) where it holds them, else
        from a*crse_old + b*crse_new of `coarse`, interpolated."""
        call("amr_regrid_fill", self.h, None if old is None else old.h, None if S_old is None else S_old.fabs,
             coarse.h, ratio, a, None if crse_old is None else crse_old.fabs, b,
             None if crse_new is None else crse_new.fabs, S.fabs)

    def advance_graph(self, S, dt, nsteps):
        call("amr_advance_graph", self.h, ctypes.byref(self.closures), ctypes.byref(self.scheme), S.fabs, dt,
             nsteps)

    def set_tiling(self, tx, ty=0, tz=0, form=0):
        """Force the fused-kernel tiling for this level's scheme (0,0,0 = auto,
        -1 = pass-structured kernels); form 1 picks the fused kernel's
        self-loading form (16x16 tiles, viscous KEEP 2/4)."""
        if form:
            call("amr_set_tiling_form", self.ctx.h, ctypes.byref(self.scheme), tx, ty, tz, form)
        else:
            call("amr_set_tiling", self.ctx.h, ctypes.byref(self.scheme), tx, ty, tz)

    def set_box_tiling(self, tx, ty, tz):
        """Force the fused-kernel tiling with tiles per box, also on a
        level-wide array."""
        call("amr_set_box_tiling", self.ctx.h, ctypes.byref(self.scheme), tx, ty, tz)

    def step_plan(self, S):
        """How a step on S runs: tile, tile kind ("passes", "boxes" or
        "level"), whether tiles per box read the cells other boxes hold from
        those boxes' arrays (no same-level ghost copies), whether the stages
        after the first read the ghost cells no box covers straight from U^n
        (no copy pass), and the fused kernel's form (0 first, 1
        self-loading)."""
        out = (ctypes.c_int32 * 6)()
        call("amr_step_plan", self.h, ctypes.byref(self.scheme), S.fabs, out)
        return {"tiling": (out[0], out[1], out[2]), "tiles": ("passes", "boxes", "level", "boxes")[out[3]],
                "direct_neighbours": out[3] == 3, "uncovered_from_un": bool(out[4]), "form": out[5]}

    def autotune(self, S):
        out = (ctypes.c_int32 * 3)()
        call("amr_autotune", self.h, ctypes.byref(self.closures), ctypes.byref(self.scheme), S.fabs, out)
        return tuple(out)

    def set_source(self, src):
        """Register the source arrays (a StateArrays, or None) added to dSdt
        for schemes with rhs_source = 1."""
        call("amr_set_source", self.h, src.fabs if src is not None else None)

    def set_user_source(self, code, params=()):
        """The problem's user_source as HIP source text (compiled at run time
        for this GPU), applied after the flux assembly for schemes with
        rhs_source = 1; None removes it."""
        arr = (ctypes.c_double * max(len(params), 1))(*params)
        call("amr_set_user_source", self.h, None if code is None else code.encode(), arr, len(params))

    def set_problem(self, wall_T=0.0, body_force_x=0.0, clear=False):
        """Problem parameters: the bc 6 wall temperature and the x body force
        of the channel's user_source (added when rhs_source = 1)."""
        if clear:
            call("amr_set_problem", self.h, ctypes.byref(self.closures), None)
            return
        p = capi.Problem(wall_T=wall_T, body_force_x=body_force_x)
        call("amr_set_problem", self.h, ctypes.byref(self.closures), ctypes.byref(p))

    # --- reverse mode ----------------------------------------------------
    def vjp_rhs(self, Sborder, lam, grad):
        """grad = (dL/dU)^T lam at the ghost-filled state Sborder."""
        call("amr_vjp_rhs", self.h, ctypes.byref(self.closures), ctypes.byref(self.scheme), Sborder.fabs,
             lam.fabs, grad.fabs)

    def vjp_stage(self, mode, a, b, c, d, Sin, lam_out, lam_s1, lam_sin, accumulate=False):
        call("amr_vjp_stage", self.h, ctypes.byref(self.closures), ctypes.byref(self.scheme), mode, a, b, c, d,
             Sin.fabs, lam_out.fabs, lam_s1.fabs if lam_s1 is not None else None, lam_sin.fabs,
             1 if accumulate else 0)

    def adjoint_run(self, S0, dt, nsteps, kcheck, grad, objective=0, weights=None):
        """Gradient of the objective after nsteps steps w.r.t. the valid
        cells of S0 (checkpoints every kcheck steps); returns J."""
        out = ctypes.c_double()
        call("amr_adjoint_run", self.h, ctypes.byref(self.closures), ctypes.byref(self.scheme), S0.fabs, dt,
             nsteps, kcheck, objective, weights.fabs if weights is not None else None, grad.fabs,
             ctypes.byref(out))
        return out.value

    # --- device-side time-step control ----------------------------------
    def clock_set(self, time=0.0):
        """Every problem's device clock to `time`; the next step takes the
        CFL estimate itself (computeInitialDt)."""
        call("amr_clock_set", self.h, time)

    def clock_get(self):
        """(times, last steps, step counts), one entry per problem."""
        n = len(self.boxes) if isinstance(self, Batch) else 1
        t = (ctypes.c_double * n)()
        d = (ctypes.c_double * n)()
        k = (ctypes.c_int32 * n)()
        call("amr_clock_get", self.h, t, d, k)
        return list(t), list(d), list(k)

    def advance_cfl(self, S, cfl, stop_time, nsteps=1):
        """nsteps steps whose sizes the device picks (CFL estimate, 1.1x
        growth cap, stop-time cut); no host synchronisation."""
        call("amr_advance_cfl", self.h, ctypes.byref(self.closures), ctypes.byref(self.scheme), S.fabs, cfl,
             stop_time, nsteps)

    def advance_cfl_graph(self, S, cfl, stop_time, nsteps):
        call("amr_advance_cfl_graph", self.h, ctypes.byref(self.closures), ctypes.byref(self.scheme), S.fabs,
             cfl, stop_time, nsteps)

    def est_time_step(self, S, cfl):
        out = ctypes.c_double()
        call("amr_est_dt", self.h, ctypes.byref(self.closures), S.fabs, cfl, ctypes.byref(out))
        return out.value

    def diagnostics(self, S, dt=0.0):
        """CNS::printTotal's figures: dict of totals, primitive min/max
        (rho, u, v, w, T, p), wave speeds max |u_d +- c| and CFL numbers."""
        D = ctypes.c_double
        tot, pmin, pmax, wav, cfl = (D * 5)(), (D * 6)(), (D * 6)(), (D * 6)(), (D * 3)()
        call("amr_diagnostics", self.h, ctypes.byref(self.closures), S.fabs, dt, tot, pmin, pmax, wav, cfl)
        return {"totals": list(tot), "prim_min": list(pmin), "prim_max": list(pmax),
                "wave_max": [(wav[2 * d], wav[2 * d + 1]) for d in range(3)], "cfl": list(cfl)}

    def check_state(self, S):
        """Lowest box with a non-physical cell, or -1 (raises on other errors)."""
        b = ctypes.c_int32(-1)
        rc = capi.load().amr_check_state(self.h, ctypes.byref(self.closures), S.fabs, ctypes.byref(b))
        if rc not in (0, capi.AMR_ERR_STATE):
            raise capi.AmrError(capi.load().amr_last_error().decode())
        return b.value

    def set_state_check(self, on=True):
        call("amr_set_state_check", self.h, 1 if on else 0)

    def totals(self, S):
        out = (ctypes.c_double * 5)()
        call("amr_totals", self.h, S.fabs, out)
        return list(out)

    def synchronize(self):
        self.ctx.synchronize()


def new_dt_levels(levels, states, n_cycle, cfl, cur_time, stop_time, dt_prev, first):
    """computeNewDt / computeInitialDt over a hierarchy, on the device:
    per-level steps (list) from the levels' states and their previous
    steps dt_prev."""
    n = len(levels)
    hs = (ctypes.c_void_p * n)(*[lv.h.value for lv in levels])
    ps = (ctypes.POINTER(Fab) * n)(*[ctypes.cast(S.fabs, ctypes.POINTER(Fab)) for S in states])
    nc = (ctypes.c_int32 * n)(*n_cycle)
    dt = (ctypes.c_double * n)(*[0.0 if d is None else d for d in dt_prev])
    call("amr_new_dt_levels", n, hs, ctypes.byref(levels[0].closures), ps, nc, cfl, cur_time, stop_time,
         1 if first else 0, dt)
    return list(dt)


def compute_new_dt(est, dt_prev, cur_time, stop_time, first):
    """CNS::computeInitialDt / computeNewDt dt policy (src/CNS.cpp:282-377)
    for one level: growth capped at 1.1x, clipped to stop_time with
    eps = 0.001 dt."""
    if first:
        return est
    dt0 = min(est, 1.1 * dt_prev)
    eps = 0.001 * dt0
    if stop_time >= 0.0 and (cur_time + dt0) > (stop_time - eps):
        dt0 = stop_time - cur_time
    return dt0


STATE_NAMES = ("density", "xmom", "ymom", "zmom", "energy")
DERIVED_NAMES = ("pressure", "temperature", "x_velocity", "y_velocity", "z_velocity")


def derive_mask(derive):
    m = 0
    for name in derive:
        if name not in DERIVED_NAMES:
            raise ValueError(f"unknown derived variable {name!r}")
        m |= 1 << DERIVED_NAMES.index(name)
    return m


def plot_names(state_vars=True, derive=()):
    """Variable names of a plotfile, in the order CNS::writePlotFile writes
    them: state components, then derived variables in registration order."""
    m = derive_mask(derive)
    return (list(STATE_NAMES) if state_vars else []) + [n for b, n in enumerate(DERIVED_NAMES) if m >> b & 1]


def write_plotfile(path, levels, states, time, level_steps, ref_ratio=2, state_vars=True, derive=()):
    """CNS::writePlotFile for levels 0..L (src/CNS.cpp:644-874): plot
    variables per level on the device, then the AMReX plotfile (Header,
    Level_l/Cell_H, Level_l/Cell_D_00000) under `path`."""
    names = plot_names(state_vars, derive)
    outs = []
    for lev, S in zip(levels, states):
        out = lev.new_state(layout=S.layout, ng=0, ncomp=len(names))
        lev.plot_vars(S, out, state_vars, derive)
        outs.append(out)
    n = len(levels)
    hs = (ctypes.c_void_p * n)(*[lev.h.value for lev in levels])
    fabs = (ctypes.POINTER(Fab) * n)(*[ctypes.cast(o.fabs, ctypes.POINTER(Fab)) for o in outs])
    cn = (ctypes.c_char_p * len(names))(*[x.encode() for x in names])
    steps = (ctypes.c_int32 * n)(*level_steps)
    rr = (ctypes.c_int32 * max(n - 1, 1))(*([ref_ratio] * max(n - 1, 1)))
    call("amr_write_plotfile", str(path).encode(), n, hs, fabs, len(names), cn, time, steps, rr)
    return names


def fill_times(order_rk, stages_rk, time, dt):
    """FillPatch times of the stages of CNS::advance, as its calls pass them
    (src/tim/CNS_advance.cpp:54-161)."""
    if order_rk == -2:
        return [time, time + dt]
    if order_rk in (0, 1):
        return [time]
    if order_rk == 2:
        m = stages_rk
        return [time + dt * float(i - 1) / float(m - 1) for i in range(1, m)] + [time + dt]
    if order_rk == 3 and stages_rk == 3:
        return [time, time + dt, time + dt / 2]
    if order_rk == 3 and stages_rk == 4:
        return [time, time + dt / 2, time + dt, time + dt / 2]
    raise ValueError("unsupported order_rk/stages_rk combination")


def cf_weights(t0, t1, t):
    """Weights (a, b) of a coarse level's states at t0 and t1 at fill time t:
    AMReX StateData::getData takes the new state within teps = 1e-3 (t1 - t0)
    of t1, else the old one within teps of t0, else FillPatchSingleLevel
    blends them with a = (t1 - t)/(t1 - t0), b = (t - t0)/(t1 - t0)."""
    teps = (t1 - t0) * 1.0e-3
    if t1 == t0 or (t1 - teps < t < t1 + teps):
        return 0.0, 1.0
    if t0 - teps < t < t0 + teps:
        return 1.0, 0.0
    return (t1 - t) / (t1 - t0), (t - t0) / (t1 - t0)


class Hierarchy:
    """Levels 0..L of an AMR hierarchy (refinement ratio r), advanced as
    AMReX's Amr::timeStep does with subcycling: level l steps by dt_l, then
    level l+1 takes r steps of dt_l / r at time + (i - 1) dt_{l+1}, then
    CNS::post_timestep averages level l+1 down onto level l (src/CNS.cpp:398-412,
    513-530; its Reflux adds nothing, the flux registers are never filled,
    src/tim/CNS_advance.cpp:356-371).  Each level runs the stage loop of
    CNS::advance with a FillPatch before every stage: amr_fill_boundary on
    level 0, amr_fill_two_levels above it, with the coarse level's states at
    the start and end of its step.  Per level three ghosted states: the step's
    U^n (kept as the coarse old state of the finer level), and two that take
    the stage outputs in turn."""

    def __init__(self, levels, ratio=2, layout="level"):
        self.levels = list(levels)
        self.ratio = ratio
        self.cur = [lev.new_state(layout=layout) for lev in self.levels]
        self.old = [lev.new_state(layout=layout) for lev in self.levels]
        self.tmp = [lev.new_state(layout=layout) for lev in self.levels]
        self.tint = [(0.0, 0.0)] * len(self.levels)
        self.ngrow = [None] * len(self.levels)
        self.direct = {}  # (level, id of a state): its stages read the other boxes' cells directly

    def upload(self, us):
        for S, u in zip(self.cur, us):
            S.upload(u)

    def download(self, lev):
        return self.cur[lev].download()

    def _fill(self, lev, S, t, stage=False):
        """Ghost cells of level lev's state S at time t; a stage input of a
        level whose fused stages read the other boxes' cells directly gets
        no same-level copies."""
        L = self.levels[lev]
        key = (lev, id(S))
        if key not in self.direct:
            # the fused stage kernel reads two ghost layers, the others three;
            # tiles per box may read the other boxes' cells themselves (decided
            # per stage input)
            plan = L.step_plan(S)
            self.ngrow[lev] = 2 if plan["tiles"] != "passes" else 3
            self.direct[key] = plan["direct_neighbours"]
        direct = stage and self.direct[key]
        if lev == 0:
            # the base level's ghost cells are same-level copies and physical
            # boundaries only (none where its stages read neighbours directly)
            if not direct:
                L.fill_patch(S)
            return
        t0, t1 = self.tint[lev - 1]
        a, b = cf_weights(t0, t1, t)
        L.fill_two_levels(S, self.levels[lev - 1], self.old[lev - 1], self.cur[lev - 1], a, b, self.ratio,
                          self.ngrow[lev], same_level=not direct)

    def _step(self, lev, time, dt):
        L = self.levels[lev]
        sch = L.scheme
        times = fill_times(sch.order_rk, sch.stages_rk, time, dt)
        A, B, C = self.cur[lev], self.old[lev], self.tmp[lev]
        sin = A
        for s, t in enumerate(times):
            self._fill(lev, sin, t, stage=True)
            _, mode, a, b, c, d = capi.stage_coeffs(sch.order_rk, sch.stages_rk, dt, s)
            out = B if s % 2 == 0 else C
            L.rk_stage(mode, a, b, c, d, A, sin, out)
            sin = out
        self.old[lev], self.cur[lev], self.tmp[lev] = A, sin, (C if sin is B else B)
        self.tint[lev] = (time, time + dt)
        if lev + 1 < len(self.levels):
            self._fill(lev, self.cur[lev], time + dt)
            dtf = dt / self.ratio
            for i in range(1, self.ratio + 1):
                self._step(lev + 1, time + (i - 1) * dtf, dtf)
            self.levels[lev + 1].average_down(self.cur[lev + 1], L, self.cur[lev], self.ratio)

    def advance(self, time, dt):
        """One coarse step of dt (level l steps by dt / r^l)."""
        self._step(0, time, dt)

    def regrid(self, lev, tagger, blocking=8, max_grid=64, eff=0.7, nest=3, autotune=True):
        """Amr::regrid of level lev+1 at the time the hierarchy has reached
        (every level at the end of its last step): CNS::errorEst tags on
        level lev's ghost-filled state (amr_tag_cells), reduced to blocks of
        blocking/ratio coarse cells, clustered into boxes of at most
        max_grid fine cells properly nested `nest` coarse cells inside level
        lev (problems.boxes_from_block_tags), and the level rebuilt on them
        by CNS::init(old) (amr_regrid_fill) with a new box table (autotuned).
        Returns the new boxes (the level is kept when they are unchanged)."""
        from . import problems as P
        C = self.levels[lev]
        S = self.cur[lev]
        if lev == 0:
            C.fill_patch(S)
        else:
            self._fill(lev, S, self.tint[lev][1])
        cb = blocking // self.ratio
        tags = C.tag_blocks(C.tag_cells(S, tagger), cb)
        boxes = P.boxes_from_block_tags(tags, C.boxes, C.n, cb, self.ratio, max_grid // blocking, eff, nest)
        old = self.levels[lev + 1]
        if [tuple(map(tuple, b)) for b in boxes] == [tuple(map(tuple, b)) for b in old.boxes]:
            return boxes
        if not boxes:
            raise ValueError("regrid: no tagged cells (removing a level is not supported)")
        new = Level(old.geom, boxes, old.closures, old.scheme, ctx=old.ctx)
        layout = self.cur[lev + 1].layout
        cur = new.new_state(layout=layout)
        new.regrid_fill(cur, old, self.cur[lev + 1], C, None, S, 0.0, 1.0, self.ratio)
        self.levels[lev + 1] = new
        self.cur[lev + 1] = cur
        self.old[lev + 1] = new.new_state(layout=layout)
        self.tmp[lev + 1] = new.new_state(layout=layout)
        self.ngrow[lev + 1] = None
        self.direct = {k: v for k, v in self.direct.items() if k[0] != lev + 1}
        old.close()
        if autotune:
            new.autotune(cur)
        return boxes


class Batch(Level):
    """Independent single-box problems advanced together (one launch per
    kernel for all of them), each with its own geometry and time step."""

    def __init__(self, geoms, closures, scheme, ctx=None, device=0):
        self.ctx = ctx if ctx is not None else Context(device)
        self.geoms = list(geoms)
        self.geom = self.geoms[0]
        self.boxes = [(tuple(g.dom_lo), tuple(g.dom_hi)) for g in self.geoms]
        self.n = None
        self.closures = closures
        self.scheme = scheme
        arr = (capi.Geom * len(self.geoms))(*self.geoms)
        h = ctypes.c_void_p()
        call("amr_batch_create", self.ctx.h, len(self.geoms), arr, ctypes.byref(h))
        self.h = h

    def new_state(self, layout="boxes", ng=3, ncomp=5):
        if layout != "boxes":
            raise ValueError("a batch holds one array per problem")
        return BatchArrays(self, ncomp=ncomp, ng=ng)

    def advance_dt(self, S, dts):
        arr = (ctypes.c_double * len(dts))(*dts)
        call("amr_batch_advance", self.h, ctypes.byref(self.closures), ctypes.byref(self.scheme), S.fabs, arr)

    def est_time_steps(self, S, cfl):
        out = (ctypes.c_double * len(self.boxes))()
        call("amr_batch_est_dt", self.h, ctypes.byref(self.closures), S.fabs, cfl, out)
        return list(out)


class BatchArrays(StateArrays):
    """One ghosted array per problem; states are given per problem."""

    def __init__(self, batch, ncomp=5, ng=3):
        super().__init__(batch, ncomp=ncomp, ng=ng, layout="boxes")

    def upload_problems(self, states):
        torch = _torch_mod()
        g = self.ng
        for u, t in zip(states, self.tensors):
            t[:, g:t.shape[1] - g, g:t.shape[2] - g, g:t.shape[3] - g] = torch.from_numpy(np.ascontiguousarray(u)).to(t.device)
        torch.cuda.synchronize()

    def download_problems(self):
        torch = _torch_mod()
        self.level.ctx.synchronize()
        torch.cuda.synchronize()
        g = self.ng
        return [t[:, g:t.shape[1] - g, g:t.shape[2] - g, g:t.shape[3] - g].cpu().numpy().copy() for t in self.tensors]
