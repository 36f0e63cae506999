"""Problem set-ups: closures, geometry, box decomposition and initial states.

Initial states are generated on the host (numpy) so that the device path and
the CPU oracle start from identical bits.
"""
import math

import numpy as np

from .capi import Closures, Scheme, Geom

# ---------------------------------------------------------------- closures


def perfect_gas(gamma=1.40, mw=28.96e-3, Ru=8.31451):
    """calorifically_perfect_gas_t (src/cls/Closures.h:72-82), same operation order."""
    cv = Ru / (mw * (gamma - 1.0))
    cp = gamma * Ru / (mw * (gamma - 1.0))
    Rspec = Ru / mw
    return dict(gamma=gamma, mw=mw, Ru=Ru, cv=cv, cp=cp, Rspec=Rspec, Cshock=1.0, Cdamp=0.01)


def closures_reference():
    """closures_derived_t<visc_suth_t, cond_suth_t, calorifically_perfect_gas_t>
    (exm/Sods/prob.h:76): Sutherland viscosity and conductivity."""
    g = perfect_gas()
    return Closures(visc_law=1, visc_ref=1.458e-6, Tvisc_ref=110.4,
                    cond_law=1, cond_ref=2.495e-3, Tcond_ref=194.0, Pr=0.71, **g)


def closures_scaled_sutherland(mu_at_T0, T0=110.4, Pr=0.71, Tref=110.4):
    """Sutherland viscosity scaled so that mu(T0) = mu_at_T0, with the
    constant-Prandtl conductivity k = mu*cp/Pr."""
    g = perfect_gas()
    visc_ref = mu_at_T0 * (Tref + T0) / (T0 * math.sqrt(T0))
    return Closures(visc_law=1, visc_ref=visc_ref, Tvisc_ref=Tref,
                    cond_law=2, cond_ref=0.0, Tcond_ref=194.0, Pr=Pr, **g)


def scheme(flux_euler=1, order_keep=4, visc=True, order_rk=3, stages_rk=3, euler=True, source=False, jst=False):
    return Scheme(rhs_euler=int(euler), rhs_visc=int(visc), rhs_source=int(source), flux_euler=flux_euler,
                  order_keep=order_keep, art_diss=int(jst), order_rk=order_rk, stages_rk=stages_rk)


def geom(n, prob_lo, prob_hi, periodic, lo_bc=(0, 0, 0), hi_bc=(0, 0, 0)):
    g = Geom()
    for d in range(3):
        g.dom_lo[d] = 0
        g.dom_hi[d] = n[d] - 1
        g.periodic[d] = int(periodic[d])
        g.lo_bc[d] = lo_bc[d]
        g.hi_bc[d] = hi_bc[d]
        g.dx[d] = (prob_hi[d] - prob_lo[d]) / n[d]
        g.prob_lo[d] = prob_lo[d]
        g.prob_hi[d] = prob_hi[d]
    return g


def decompose(n, max_grid):
    """Chop the domain [0,n) into boxes of at most max_grid cells per side."""
    boxes = []
    mg = [max_grid] * 3 if np.isscalar(max_grid) else list(max_grid)
    for k0 in range(0, n[2], mg[2]):
        for j0 in range(0, n[1], mg[1]):
            for i0 in range(0, n[0], mg[0]):
                lo = (i0, j0, k0)
                hi = (min(i0 + mg[0], n[0]) - 1, min(j0 + mg[1], n[1]) - 1, min(k0 + mg[2], n[2]) - 1)
                boxes.append((lo, hi))
    return boxes


def cell_centres(g, n):
    x = g.prob_lo[0] + (np.arange(n[0]) + 0.5) * g.dx[0]
    y = g.prob_lo[1] + (np.arange(n[1]) + 0.5) * g.dx[1]
    z = g.prob_lo[2] + (np.arange(n[2]) + 0.5) * g.dx[2]
    return np.meshgrid(x, y, z, indexing="ij")


def to_fab_order(u_ijk):
    """(5, nx, ny, nz) -> (5, nz, ny, nx) C-contiguous (i fastest in memory)."""
    return np.ascontiguousarray(np.transpose(u_ijk, (0, 3, 2, 1)))


# ---------------------------------------------------------------- Sod


def sod_state(cls, g, n, rho_l=1.0, p_l=1.0, u_l=0.0, rho_r=0.125, p_r=0.1, u_r=0.0, prob_hi_x=None):
    """exm/Sods/prob.h:98-121 on the whole domain, FAB order (5,nz,ny,nx)."""
    if prob_hi_x is None:
        prob_hi_x = g.prob_lo[0] + n[0] * g.dx[0]
    x = g.prob_lo[0] + (np.arange(n[0]) + 0.5) * g.dx[0]
    left = x < prob_hi_x / 2
    P = np.where(left, p_l, p_r)
    rho = np.where(left, rho_l, rho_r)
    ux = np.where(left, u_l, u_r)
    u = np.zeros((5, n[0]))
    u[0] = rho
    u[1] = rho * ux
    et = P / (cls.gamma - 1.0)
    u[4] = et + 0.5 * rho * ux * ux
    full = np.broadcast_to(u[:, :, None, None], (5, n[0], n[1], n[2]))
    return to_fab_order(full)


# ---------------------------------------------------------------- TGV


def tgv_closures(mach=0.1, Re=1600.0, L=1.0, T0=110.4, Pr=0.71, rho0=1.0):
    """Scaled Sutherland so that the reference TGV set-up gives rho0 = 1."""
    g = perfect_gas()
    cs = math.sqrt(g["gamma"] * g["Rspec"] * T0)
    v0 = mach * cs
    mu0 = rho0 * L * v0 / Re
    return closures_scaled_sutherland(mu0, T0=T0, Pr=Pr)


def tgv_state(cls, g, n, mach=0.1, Re=1600.0, L=1.0, T0=110.4):
    """exm/TGV/prob.h:92-134 (omega = 1, convecting = false)."""
    X, Y, Z = cell_centres(g, n)
    return _tgv_fields(cls, X, Y, Z, mach, Re, L, T0)


def box_centres(g, lo, hi, ng):
    """Cell centres of the ghosted box [lo-ng, hi+ng] (index arrays, ij order)."""
    ax = [g.prob_lo[d] + (np.arange(lo[d] - ng, hi[d] + ng + 1) + 0.5) * g.dx[d] for d in range(3)]
    return np.meshgrid(*ax, indexing="ij")


def tgv_state_box(cls, g, lo, hi, ng=3, mach=0.1, Re=1600.0, L=1.0, T0=110.4):
    """The TGV initial state on the ghosted box [lo-ng, hi+ng] (FAB order);
    ghost cells hold the analytic field (periodic in 2 pi)."""
    X, Y, Z = box_centres(g, lo, hi, ng)
    return _tgv_fields(cls, X, Y, Z, mach, Re, L, T0)


def _tgv_fields(cls, X, Y, Z, mach, Re, L, T0):
    cs = math.sqrt(cls.gamma * cls.Rspec * T0)
    v0 = mach * cs
    if cls.visc_law == 0:
        mu = cls.visc_ref
    else:
        mu = cls.visc_ref * T0 * math.sqrt(T0) / (cls.Tvisc_ref + T0)
    rho0 = Re * mu / (L * v0)
    p0 = rho0 * cls.Rspec * T0
    u0 = v0 * np.sin(X / L) * np.cos(Y / L) * np.cos(Z / L)
    u1 = -v0 * np.cos(X / L) * np.sin(Y / L) * np.cos(Z / L)
    u2 = np.zeros_like(u0)
    p = p0 + rho0 * v0 * v0 / 16.0 * (np.cos(2.0 * X / L) + np.cos(2.0 * Y / L)) * (np.cos(2.0 * Z / L) + 2.0)
    rho = p / (cls.Rspec * T0)
    eint = cls.cv * T0
    u = np.empty((5,) + X.shape)
    u[0] = rho
    u[1] = rho * u0
    u[2] = rho * u1
    u[3] = rho * u2
    u[4] = rho * (eint + 0.5 * (u0 * u0 + u1 * u1 + u2 * u2))
    return to_fab_order(u)


def tgv_vorticity_magnitude(g, n, mach=0.1, L=1.0, T0=110.4, cls=None, X=None, Y=None, Z=None):
    """Analytic |omega| of the TGV velocity field at t = 0 (at the cell
    centres of the domain, or at the given coordinates)."""
    if X is None:
        X, Y, Z = cell_centres(g, n)
    gam = 1.4 if cls is None else cls.gamma
    Rs = (8.31451 / 28.96e-3) if cls is None else cls.Rspec
    v0 = mach * math.sqrt(gam * Rs * T0)
    # u = v0 sin x cos y cos z, v = -v0 cos x sin y cos z, w = 0
    wx = -v0 * np.cos(X) * np.sin(Y) * np.sin(Z) / L  # dw/dy - dv/dz
    wy = -v0 * np.sin(X) * np.cos(Y) * np.sin(Z) / L  # du/dz - dw/dx
    wz = 2.0 * v0 * np.sin(X) * np.sin(Y) * np.cos(Z) / L  # dv/dx - du/dy
    return np.sqrt(wx * wx + wy * wy + wz * wz)


# ---------------------------------------------------------------- AMR hierarchy


def cluster_blocks(tags, max_blocks, eff=0.7, allowed=None):
    """Berger-Rigoutsos style clustering of a boolean block mask tags[i,j,k]
    into boxes (inclusive block bounds): shrink to the tags' bounding box;
    accept it when the tagged fraction is >= eff, no side exceeds max_blocks
    and (if given) every block of it is `allowed`; otherwise cut at a hole of
    the signature nearest the centre, else at the strongest inflection of the
    signature, else in the middle of the longest side, and recurse.  Tagged
    blocks must be allowed."""
    out = []
    stack = [((0, 0, 0), tuple(n - 1 for n in tags.shape))]
    while stack:
        lo, hi = stack.pop()
        sub = tags[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1]
        if not sub.any():
            continue
        nlo, nhi = [], []
        for d in range(3):
            others = tuple(a for a in range(3) if a != d)
            nz = np.nonzero(sub.any(axis=others))[0]
            nlo.append(lo[d] + int(nz[0]))
            nhi.append(lo[d] + int(nz[-1]))
        lo, hi = tuple(nlo), tuple(nhi)
        sub = tags[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1]
        size = [hi[d] - lo[d] + 1 for d in range(3)]
        inside = allowed is None or bool(allowed[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1].all())
        if sub.sum() >= eff * sub.size and max(size) <= max_blocks and inside:
            out.append((lo, hi))
            continue
        cut = None
        sigs = [sub.sum(axis=tuple(a for a in range(3) if a != d)) for d in range(3)]
        best = None
        for d in range(3):  # holes
            if size[d] < 2:
                continue
            for c in range(1, size[d]):
                if sigs[d][c - 1] == 0 or sigs[d][c] == 0:
                    dist = abs(c - size[d] / 2)
                    if best is None or dist < best[0]:
                        best = (dist, d, c)
        if best is None:  # inflection of the Laplacian of the signature
            for d in range(3):
                if size[d] < 4:
                    continue
                lap = sigs[d][:-2] - 2 * sigs[d][1:-1] + sigs[d][2:]
                for c in range(1, len(lap)):
                    if lap[c - 1] * lap[c] < 0:
                        jump = abs(lap[c] - lap[c - 1])
                        if best is None or jump > best[0]:
                            best = (jump, d, c + 1)
        if best is not None:
            _, d, c = best
        else:
            d = int(np.argmax(size))
            c = size[d] // 2
        if max(size) > max_blocks and size[d] <= max_blocks:
            d = int(np.argmax(size))
            c = size[d] // 2
        cut = lo[d] + c
        h1 = list(hi)
        h1[d] = cut - 1
        l2 = list(lo)
        l2[d] = cut
        stack.append((lo, tuple(h1)))
        stack.append((tuple(l2), hi))
    return sorted(out, key=lambda b: (b[0][2], b[0][1], b[0][0]))


def _intersect(a, b):
    lo = tuple(max(a[0][d], b[0][d]) for d in range(3))
    hi = tuple(min(a[1][d], b[1][d]) for d in range(3))
    return (lo, hi) if all(lo[d] <= hi[d] for d in range(3)) else None


def _erode_periodic(mask, w):
    """Cells of `mask` whose (2w+1)^3 neighbourhood (periodic) lies in it."""
    out = mask.copy()
    for d in range(3):
        m = out.copy()
        for s in range(1, w + 1):
            m &= np.roll(out, s, axis=d) & np.roll(out, -s, axis=d)
        out = m
    return out


def tgv_amr_hierarchy(n_base=128, nlev=2, ratio=2, pct=70.0, blocking=8, max_grid=64, base_grid=64, eff=0.7,
                      nest=0):
    """Levels of the batched-AMR TGV workload on [0, 2 pi]^3: a uniform base
    level, then `nlev` refined levels (ratio 2).  On each level the cells
    whose analytic |omega| (t = 0) exceeds the level's `pct` percentile are
    tagged; tags are clustered in blocks of blocking/ratio coarse cells into
    boxes of at most max_grid fine cells, refined, and clipped to the coarse
    level's boxes.  With nest > 0 every fine box also stays `nest` coarse
    cells inside the coarse level's union (tags outside are dropped, clusters
    cut until they fit), as coarse-fine interpolation needs for ghost cells
    of width 3 (nest = 3).  Returns [(geom, boxes)] per level."""
    levels = []
    n = (n_base,) * 3
    g = geom(n, (0, 0, 0), (2 * math.pi,) * 3, (1, 1, 1))
    boxes = decompose(n, base_grid)
    levels.append((g, boxes))
    cb = blocking // ratio          # block edge in coarse cells
    mb = max_grid // blocking       # max box edge in blocks
    for _ in range(nlev):
        g, boxes = levels[-1]
        n = [g.dom_hi[d] + 1 for d in range(3)]
        # |omega| on the level's cells, its percentile, tags by block
        vals = []
        for lo, hi in boxes:
            X, Y, Z = box_centres(g, lo, hi, 0)
            vals.append(tgv_vorticity_magnitude(g, n, X=X, Y=Y, Z=Z))
        thr = np.percentile(np.concatenate([v.ravel() for v in vals]), pct)
        nb = [n[d] // cb for d in range(3)]
        tags = np.zeros(nb, dtype=bool)
        for (lo, hi), v in zip(boxes, vals):
            t = v > thr
            sh = [(hi[d] - lo[d] + 1) // cb for d in range(3)]
            blk = t.reshape(sh[0], cb, sh[1], cb, sh[2], cb).any(axis=(1, 3, 5))
            tags[lo[0] // cb:lo[0] // cb + sh[0], lo[1] // cb:lo[1] // cb + sh[1],
                 lo[2] // cb:lo[2] // cb + sh[2]] |= blk
        fine = boxes_from_block_tags(tags, boxes, n, cb, ratio, mb, eff, nest)
        nf = [n[d] * ratio for d in range(3)]
        gf = geom(nf, (0, 0, 0), (2 * math.pi,) * 3, (1, 1, 1))
        levels.append((gf, fine))
    return levels


def boxes_from_block_tags(tags, boxes, n, cb, ratio=2, max_blocks=8, eff=0.7, nest=0):
    """Fine boxes from a coarse level's tags reduced to blocks of cb^3 coarse
    cells (tags[bi, bj, bk], bool): with nest > 0 tags outside the coarse
    union eroded by `nest` cells are dropped and clusters are cut until they
    fit inside it; clusters (cluster_blocks, at most max_blocks blocks a
    side, tagged fraction >= eff) are clipped to the coarse boxes and
    refined by `ratio`.  Sorted by (k, j, i) of their low corners."""
    nb = [n[d] // cb for d in range(3)]
    tags = np.array(tags, dtype=bool)
    allowed = None
    if nest > 0:
        cover = np.zeros(n, dtype=bool)
        for lo, hi in boxes:
            cover[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1] = True
        inner = _erode_periodic(cover, nest)
        allowed = inner.reshape(nb[0], cb, nb[1], cb, nb[2], cb).all(axis=(1, 3, 5))
        tags &= allowed
    clusters = cluster_blocks(tags, max_blocks, eff, allowed)
    fine = []
    for blo, bhi in clusters:
        clo = tuple(blo[d] * cb for d in range(3))
        chi = tuple((bhi[d] + 1) * cb - 1 for d in range(3))
        for B in boxes:
            x = _intersect((clo, chi), B)
            if x is not None:
                fine.append((tuple(x[0][d] * ratio for d in range(3)),
                             tuple((x[1][d] + 1) * ratio - 1 for d in range(3))))
    return sorted(fine, key=lambda b: (b[0][2], b[0][1], b[0][0]))


# ---------------------------------------------------------------- channel-like box


def channel_closures(Re=500.0, mach=0.2, Ub=1.0, rho0=1.0, H=1.0, Pr=0.71):
    """Scaled Sutherland closure of the periodic channel-like box: mu(T0) =
    rho0 Ub H / Re at the reference temperature T0 = p0 / (rho0 R) with
    p0 = rho0 c^2 / gamma, c = Ub / mach."""
    g = perfect_gas()
    c = Ub / mach
    p0 = rho0 * c * c / g["gamma"]
    T0 = p0 / (rho0 * g["Rspec"])
    return closures_scaled_sutherland(rho0 * Ub * H / Re, T0=T0, Pr=Pr), p0, T0


def channel_geom(n=(128, 64, 64)):
    """[0, 2 pi] x [-1, 1] x [0, pi], periodic in every direction."""
    return geom(n, (0.0, -1.0, 0.0), (2 * math.pi, 1.0, math.pi), (1, 1, 1))


def channel_state(cls, g, n, p0, seed=20261202, ampl=0.05):
    """rho = 1, u = 1.5 (1 - y^2) plus seeded U[-1,1] perturbations of
    amplitude `ampl` on u, v, w (numpy PCG64), uniform pressure p0."""
    rng = np.random.default_rng(seed)
    X, Y, Z = cell_centres(g, n)
    shape = X.shape
    vel = [1.5 * (1.0 - Y * Y) + ampl * rng.uniform(-1, 1, size=shape),
           ampl * rng.uniform(-1, 1, size=shape),
           ampl * rng.uniform(-1, 1, size=shape)]
    rho = np.ones(shape)
    u = np.empty((5,) + shape)
    u[0] = rho
    for d in range(3):
        u[1 + d] = rho * vel[d]
    u[4] = p0 / (cls.gamma - 1.0) + 0.5 * rho * (vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)
    return to_fab_order(u)


def kinetic_energy(u, g):
    """sum |m|^2 / (2 rho) dx dy dz over a (5, nz, ny, nx) state."""
    dV = g.dx[0] * g.dx[1] * g.dx[2]
    return float(np.sum(0.5 * (u[1] ** 2 + u[2] ** 2 + u[3] ** 2) / u[0]) * dV)


# ---------------------------------------------------------------- random smooth states


def smooth_random_state(cls, n, seed, ampl=0.05, rho0=1.0, T0=300.0, u0=10.0):
    """A smooth periodic-ish positive state for parity tests."""
    rng = np.random.default_rng(seed)
    shape = (n[0], n[1], n[2])
    I, J, K = np.meshgrid(*(np.arange(m) for m in n), indexing="ij")

    def field():
        f = np.zeros(shape)
        for _ in range(3):
            kx, ky, kz = rng.integers(1, 3, size=3)
            ph = rng.uniform(0, 2 * np.pi, size=3)
            f += np.sin(2 * np.pi * kx * I / n[0] + ph[0]) * np.cos(2 * np.pi * ky * J / n[1] + ph[1]) * \
                np.sin(2 * np.pi * kz * K / n[2] + ph[2])
        return f / 3.0 + 0.3 * rng.uniform(-1, 1, size=shape)

    rho = rho0 * (1.0 + ampl * field())
    T = T0 * (1.0 + ampl * field())
    vel = [u0 * field() for _ in range(3)]
    p = rho * cls.Rspec * T
    u = np.empty((5,) + shape)
    u[0] = rho
    for d in range(3):
        u[1 + d] = rho * vel[d]
    u[4] = p / (cls.gamma - 1.0) + 0.5 * rho * (vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)
    return to_fab_order(u)


# ---------------------------------------------------------------- SplitMix64


class SplitMix64:
    """SplitMix64 generator (Steele, Lea & Flood 2014); integer-exact in Python."""

    MASK = (1 << 64) - 1

    def __init__(self, seed):
        self.s = seed & self.MASK

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & self.MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, a, b):
        return a + (b - a) * ((self.next_u64() >> 11) * (1.0 / 9007199254740992.0))

    def integer(self, lo, hi):
        """Uniform integer in [lo, hi]."""
        return lo + self.next_u64() % (hi - lo + 1)


def riemann_batch_params(n_inst=512, seed=20261202, lmin=64, lmax=2048):
    """Instances of the batched 1D Riemann workload: length and L/R states."""
    g = SplitMix64(seed)
    inst = []
    for _ in range(n_inst):
        L = g.integer(lmin, lmax)
        rl, ul, pl = g.uniform(0.1, 1.0), g.uniform(-1.0, 1.0), g.uniform(0.1, 1.0)
        rr, ur, pr = g.uniform(0.1, 1.0), g.uniform(-1.0, 1.0), g.uniform(0.1, 1.0)
        inst.append(dict(L=int(L), rho_l=rl, u_l=ul, p_l=pl, rho_r=rr, u_r=ur, p_r=pr))
    return inst


def riemann_batch_problems(cls, inst, ny=4, nz=4):
    """Geometry and initial state of every instance of the batched 1D
    Riemann workload: L x 4 x 4 cells on [0,1] x [0,4/L]^2, split at x = 0.5,
    outflow in x, periodic in y and z."""
    geoms, states = [], []
    for p in inst:
        L = p["L"]
        n = (L, ny, nz)
        g = geom(n, (0.0, 0.0, 0.0), (1.0, ny / L, nz / L), (0, 1, 1), lo_bc=(2, 0, 0), hi_bc=(2, 0, 0))
        geoms.append(g)
        states.append(sod_state(cls, g, n, rho_l=p["rho_l"], p_l=p["p_l"], u_l=p["u_l"],
                                rho_r=p["rho_r"], p_r=p["p_r"], u_r=p["u_r"], prob_hi_x=1.0))
    return geoms, states


# --------------------------------------------------------------- wall channel
def wall_channel_parm():
    """ProbParm of exm/Channel/prob.h:71-98 (same formulas, same order)."""
    h, Re_tau, Re_b, Ma_b, u_tau, Tw, Rgas = 0.006845, 220.0, 3.0e3, 1.5, 35.0, 500.0, 287.0
    muw = 2.670533348e-5
    rho_w = Re_tau * muw / (u_tau * h)
    Pw = rho_w * Rgas * Tw
    aw = math.pow(1.4 * Rgas * Tw, 0.5)
    ub = Ma_b * aw
    rho_b = Re_b * muw / (ub * h)
    fx = rho_w * u_tau * u_tau / (h * rho_b)
    return {"h": h, "Tw": Tw, "Pw": Pw, "fx": fx, "ub": ub, "rho_b": rho_b}


def wall_channel(n=(96, 128, 64)):
    """The compressible channel of exm/Channel: [0, 0.08214] x [0, 0.0136871]
    x [0, 0.02738] (inputs:12-15), periodic in x and z, bc 6 isothermal
    no-slip walls in y (prob.h:47-48), reference Sutherland closures, KEEP4 +
    viscous + body-force source, SSPRK(3,3) (prob.h:49-56).  Returns
    (closures, geometry, scheme, ProbParm dict)."""
    cls = closures_reference()
    g = geom(n, (0.0, 0.0, 0.0), (0.08214, 0.0136871, 0.02738), (1, 0, 1), lo_bc=(0, 6, 0), hi_bc=(0, 6, 0))
    sch = scheme(flux_euler=1, order_keep=4, visc=True, order_rk=3, stages_rk=3, source=True)
    return cls, g, sch, wall_channel_parm()


def wall_channel_state(cls, g, n, parm):
    """prob_initdata of exm/Channel/prob.h:100-155 with the GPU build's
    rand_num = 0.5 (turbulent-profile mean flow plus one Fourier mode)."""
    lo = [g.prob_lo[d] for d in range(3)]
    ln = [g.prob_hi[d] - g.prob_lo[d] for d in range(3)]
    ax = [lo[d] + (np.arange(n[d]) + 0.5) * g.dx[d] for d in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    h = parm["h"]
    Y = np.where(Y > h, 2 * h - Y, Y)
    T = 677.04110513 + 255.46817 * np.power(Y, 0.5) - 176.38967315 * np.exp(-3081.95957181 * Y) * 1.025
    ux = 581.54524091 + 2687.84321585 * np.power(Y, 0.5) - 593.73493364 * np.exp(-2008.87100276 * Y)
    pi = 3.14159265358979323846
    rand_num, A = 0.5, 100 * 2.5
    nx, ny, nz = 1, 2, 3
    up = rand_num * A * np.cos(2 * pi * X * nx / ln[0]) * np.sin(pi * Y * ny / ln[1]) * np.cos(2 * pi * Z * nz / ln[2])
    vp = rand_num * -A * np.sin(2 * pi * X * nx / ln[0]) * np.cos(pi * Y * ny / ln[1]) * np.sin(2 * pi * Z * nz / ln[2])
    wp = rand_num * -A * np.sin(2 * pi * X * nx / ln[0]) * np.cos(pi * Y * ny / ln[1]) * np.sin(2 * pi * Z * nz / ln[2])
    rho = parm["Pw"] / (cls.Rspec * T)
    u = np.empty((5,) + X.shape)
    u[0] = rho
    u[1] = rho * (ux + up)
    u[2] = rho * vp
    u[3] = rho * wp
    u[4] = rho * cls.cv * T + 0.5 * rho * (ux * ux + up * up + vp * vp + wp * wp)
    return u
