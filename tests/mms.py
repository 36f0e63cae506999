"""Manufactured solution of exm/MMS_euler/prob.h (test aid): the state
(prob_initdata, :190-212) and the exact Euler flux divergence that its
user_source adds to dSdt (:219-283), restated with numpy.  With this source
dSdt + source is the truncation error of the discrete operator (the
quantity exm/MMS_euler/convergence_order.py:161-186 takes norms of)."""
import math

import numpy as np

from amrsolver_amd import problems as P

A = 2.0 * math.pi  # alpha = 2 n pi / Lambda with n = 1, Lambda = 1


def _fields(x, y, z):
    c, s = np.cos, np.sin
    f = c(A * x) * s(A * y) * s(A * z)                  # u = v = w
    fx = -A * s(A * x) * s(A * y) * s(A * z)
    fy = A * c(A * x) * c(A * y) * s(A * z)
    fz = A * c(A * x) * s(A * y) * c(A * z)
    g = 4.0 + c(A * x) * s(A * y) * c(A * z)           # p = T
    gx = -A * s(A * x) * s(A * y) * c(A * z)
    gy = A * c(A * x) * c(A * y) * c(A * z)
    gz = -A * c(A * x) * s(A * y) * s(A * z)
    return f, (fx, fy, fz), g, (gx, gy, gz)


def mms_geom(n):
    return P.geom((n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (1, 1, 1))


def mms_state(cls, n):
    g = mms_geom(n)
    X, Y, Z = P.cell_centres(g, (n, n, n))
    f, _, pT, _ = _fields(X, Y, Z)
    rho = pT / (cls.Rspec * pT)
    u = np.empty((5,) + X.shape)
    u[0] = rho
    u[1] = rho * f
    u[2] = rho * f
    u[3] = rho * f
    u[4] = rho * (cls.cv * pT + 0.5 * (3.0 * f * f))
    return g, P.to_fab_order(u)


def mms_source(cls, n):
    """The terms user_source adds, rho = 1/Rspec (p = T, so rho is constant)."""
    g = mms_geom(n)
    X, Y, Z = P.cell_centres(g, (n, n, n))
    f, df, pT, dg = _fields(X, Y, Z)
    u = v = w = f
    dudx, dudy, dudz = df
    dvdx, dvdy, dvdz = df
    dwdx, dwdy, dwdz = df
    p = T = pT
    dpdx, dpdy, dpdz = dg
    dTdx, dTdy, dTdz = dg
    rho = 1.0 / cls.Rspec
    s = np.empty((5,) + X.shape)
    s[0] = rho * (dudx + dvdy + dwdz)
    s[1] = rho * (2 * u * dudx + v * dudy + u * dvdy + w * dudz + u * dwdz) + dpdx
    s[2] = rho * (u * dvdx + v * dudx + 2 * v * dvdy + w * dvdz + v * dwdz) + dpdy
    s[3] = rho * (u * dwdx + w * dudx + v * dwdy + w * dvdy + 2 * w * dwdz) + dpdz
    et = cls.cv * T + 0.5 * (u * u + v * v + w * w)
    det = cls.cv * dTdx + (u * dudx + v * dvdx + w * dwdx)
    e = rho * (et * dudx + u * det + (1 / rho) * (u * dpdx + p * dudx))
    det = cls.cv * dTdy + (u * dudy + v * dvdy + w * dwdy)
    e += rho * (et * dvdy + v * det + (1 / rho) * (v * dpdy + p * dvdy))
    det = cls.cv * dTdz + (u * dudz + v * dvdz + w * dwdz)
    e += rho * (et * dwdz + w * det + (1 / rho) * (w * dpdz + p * dwdz))
    s[4] = e
    return P.to_fab_order(s)


def norms(err):
    """Per component L2 (root mean square) and L-infinity norms."""
    n3 = err[0].size
    return (np.sqrt(np.sum(err.reshape(5, -1) ** 2, axis=1) / n3), np.max(np.abs(err.reshape(5, -1)), axis=1))


def rates(errs, ns):
    return [np.log(errs[i + 1] / errs[i]) / np.log(ns[i] / ns[i + 1]) for i in range(len(ns) - 1)]
