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


def scheme(flux_euler=1, order_keep=4, visc=True, order_rk=3, stages_rk=3, euler=True):
    return Scheme(rhs_euler=int(euler), rhs_visc=int(visc), rhs_source=0, flux_euler=flux_euler,
                  order_keep=order_keep, art_diss=0, order_rk=order_rk, stages_rk=stages_rk)


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


def tgv_vorticity_magnitude(g, n, mach=0.1, L=1.0, T0=110.4, cls=None):
    """Analytic |omega| of the TGV velocity field at t = 0."""
    X, Y, Z = cell_centres(g, n)
    gam = 1.4 if cls is None else cls.gamma
    Rs = (8.31451 / 28.96e-3) if cls is None else cls.Rspec
    v0 = mach * math.sqrt(gam * Rs * T0)
    # u = v0 sin x cos y cos z, v = -v0 cos x sin y cos z, w = 0
    wx = -v0 * np.cos(X) * np.sin(Y) * np.sin(Z) / L  # dw/dy - dv/dz
    wy = -v0 * np.sin(X) * np.cos(Y) * np.sin(Z) / L  # du/dz - dw/dx
    wz = 2.0 * v0 * np.sin(X) * np.sin(Y) * np.cos(Z) / L  # dv/dx - du/dy
    return np.sqrt(wx * wx + wy * wy + wz * wz)


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
