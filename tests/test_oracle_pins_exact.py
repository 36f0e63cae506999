"""Exact pins of the oracle's WENO5-JS weights and two-shock Riemann solver.

The nonlinear WENO weights and the Riemann iteration have no reference
golden vectors (the reference ships none and cannot be built here).  They
are pinned by what the mathematics fixes:

* WENO5-JS (Jiang & Shu, J. Comput. Phys. 126, 1996: q = 2, linear weights
  d = {0.1, 0.6, 0.3}, eps = 1e-40 as src/rhs/High_resolution.h:127-149
  sets them) evaluated here in exact rational arithmetic.  The reference's
  beta1 uses 1/4 (s1 - s3)^4 where the paper has the square
  (High_resolution.h:137-140).  On data with (s1 - s3)^2 in {0, 1} the two
  coincide, so the oracle must equal the textbook scheme there; on data with
  s1 - 2 s2 + s3 = 0 the first term of beta1 vanishes and beta1 is exactly
  (s1 - s3)^4 / 4, so the result's closed form depends on the exponent.
* The two-shock solver (src/rhs/Riemann.h:123-249) on symmetric colliding
  streams: by symmetry u* = 0 exactly and the Godunov flux is (0, p*, 0, 0,
  0); both waves are shocks, so the exact p* is the two-shock value, solved
  to machine precision by tests/exact_riemann.py.  The solver's three fixed
  iterations leave a known error against it.
* The z direction's left-state sound speed (Riemann.h:316 reads T at k+1):
  with given slopes the z face equals the x face of the permuted data once
  the left cell's temperature is replaced by the k+1 cell's, bit for bit.
"""
import math
from fractions import Fraction as Fr

import numpy as np
import pytest

from amrsolver_amd import problems as P
from oracle import oracle as O
import exact_riemann as ER  # noqa: E402

G = 1.4
EPS = Fr(1e-40)  # the double the reference's constexpr 1e-40 rounds to
D_LIN = (Fr(1, 10), Fr(6, 10), Fr(3, 10))


# ------------------------------------------------------------------ WENO5-JS
def weno_exact(s, beta1_power=4, eps=EPS, d=D_LIN, q=2):
    """WENO5-JS in exact rational arithmetic (Jiang & Shu 1996, eqs. 2.9,
    3.2-3.4).  beta1_power = 2 is the paper's beta1, 4 the reference's."""
    s0, s1, s2, s3, s4 = (Fr(x) for x in s)
    b0 = Fr(13, 12) * (s0 - 2 * s1 + s2) ** 2 + Fr(1, 4) * (s0 - 4 * s1 + 3 * s2) ** 2
    b1 = Fr(13, 12) * (s1 - 2 * s2 + s3) ** 2 + Fr(1, 4) * (s1 - s3) ** beta1_power
    b2 = Fr(13, 12) * (s2 - 2 * s3 + s4) ** 2 + Fr(1, 4) * (3 * s2 - 4 * s3 + s4) ** 2
    a = [d[0] / (eps + b0) ** q, d[1] / (eps + b1) ** q, d[2] / (eps + b2) ** q]
    v = [(2 * s0 - 7 * s1 + 11 * s2) / 6, (-s1 + 5 * s2 + 2 * s3) / 6, (2 * s2 + 5 * s3 - s4) / 6]
    return sum(ai * vi for ai, vi in zip(a, v)) / sum(a), a, v


def weno_tol(s, **kw):
    """Rounding bound of the double evaluation: a few ulps of the weighted
    candidate magnitudes (the alphas and the weight sum are sums of positive
    terms; the candidates may cancel)."""
    _, a, v = weno_exact(s, **kw)
    mag = sum(ai * abs(vi) for ai, vi in zip(a, v)) / sum(a) + max(abs(Fr(x)) for x in s)
    return 64 * 2.0 ** -52 * float(mag)


def dyadic(rng, lo, hi, bits=6):
    """Random multiples of 2^-bits: sums and differences of a few of them
    are exact in doubles."""
    return rng.integers(int(lo * 2 ** bits), int(hi * 2 ** bits)) / 2.0 ** bits


def family_textbook(rng, n):
    """(s1 - s3)^2 in {0, 1}: the quartic and quadratic beta1 coincide."""
    out = []
    for k in range(n):
        s = [dyadic(rng, -4, 4) for _ in range(5)]
        s[3] = s[1] + (-1.0, 0.0, 1.0)[k % 3]
        out.append(s)
    return out


def family_quartic(rng, n):
    """s1 - 2 s2 + s3 = 0 and |s1 - s3| >= 2: beta1 = (s1 - s3)^4 / 4."""
    out = []
    while len(out) < n:
        s1 = float(rng.integers(-6, 7))
        s3 = s1 + 2.0 * float(rng.integers(1, 4)) * (1 if rng.random() < 0.5 else -1)
        s2 = 0.5 * (s1 + s3)
        s0, s4 = dyadic(rng, -6, 6), dyadic(rng, -6, 6)
        out.append([s0, s1, s2, s3, s4])
    return out


@pytest.mark.parametrize("seed", range(4))
def test_weno_equals_textbook_where_beta1_forms_coincide(seed):
    rng = np.random.default_rng(100 + seed)
    for s in family_textbook(rng, 30):
        exact, _, _ = weno_exact(s, beta1_power=2)
        assert exact == weno_exact(s, beta1_power=4)[0]
        got = O.weno5js(s)
        assert abs(got - float(exact)) <= weno_tol(s), (s, got, float(exact))


@pytest.mark.parametrize("seed", range(4))
def test_weno_beta1_is_the_quartic(seed):
    """beta1 = (s1 - s3)^4 / 4 exactly on this family: the oracle equals the
    scheme with the quartic, and the paper's square gives a result outside
    the rounding bound (so the exponent is what this test pins)."""
    rng = np.random.default_rng(200 + seed)
    separated = 0
    for s in family_quartic(rng, 30):
        quart, _, _ = weno_exact(s, beta1_power=4)
        quad, _, _ = weno_exact(s, beta1_power=2)
        got = O.weno5js(s)
        tol = weno_tol(s)
        assert abs(got - float(quart)) <= tol, (s, got, float(quart))
        if abs(float(quad) - float(quart)) > 100 * tol:
            separated += 1
    assert separated >= 20  # the family does tell the two exponents apart


def test_weno_epsilon_and_exponent():
    """Data of amplitude 1e-20 make the smoothness indicators ~1e-40, the
    size of eps: the weights then depend on eps = 1e-40 and on q = 2 at O(1).
    The oracle matches the exact evaluation; eps = 1e-6, eps = 1e-36, q = 1
    and permuted linear weights all miss by far more than the rounding."""
    rng = np.random.default_rng(7)
    for _ in range(40):
        s = [x * 1e-20 for x in rng.uniform(-1, 1, 5)]
        exact = float(weno_exact(s)[0])
        got = O.weno5js(s)
        tol = weno_tol(s)
        assert abs(got - exact) <= tol, (s, got, exact)
        wrong = [weno_exact(s, eps=Fr(1e-6))[0], weno_exact(s, eps=Fr(1e-36))[0], weno_exact(s, q=1)[0],
                 weno_exact(s, d=(D_LIN[2], D_LIN[1], D_LIN[0]))[0]]
        for w in wrong:
            assert abs(float(w) - exact) > 100 * tol


def test_weno_large_amplitude_and_constant():
    """Constant data: every beta is 0, the weights are the linear ones and
    the result is the constant to rounding (eps keeps 0/0 away); steep data
    scaled by 1e6 still matches the exact evaluation."""
    for c in (0.0, 1.0, -3.25, 1e6):
        assert float(weno_exact([c] * 5)[0]) == c
        assert abs(O.weno5js([c] * 5) - c) <= weno_tol([c] * 5)
    rng = np.random.default_rng(9)
    for _ in range(20):
        s = [x * 1e6 for x in rng.uniform(-1, 1, 5)]
        assert abs(O.weno5js(s) - float(weno_exact(s)[0])) <= weno_tol(s)


# ------------------------------------------------------------------ Riemann
# Error of the solver's p* against the exact two-shock value on symmetric
# collisions rho = p = 1, u = +-U.  The iteration (a secant iteration on the
# shock impedances, two initial guesses then 3 updates) converges to the
# exact value; its iterates 2, 3, 4 have relative errors of about
#   U = 2: 4.2e-3, 6.7e-5, 8.8e-8;   U = 3: 1.2e-2, 3.6e-4, 1.3e-6;
#   U = 5: 3.1e-2, 1.6e-3, 1.3e-5.
# The bands below contain the third iterate and exclude the second and the
# fourth by more than an order of magnitude.
COLLISION_BANDS = [(2.0, 2e-5, 3e-4), (3.0, 1e-4, 1.5e-3), (5.0, 5e-4, 6e-3)]


@pytest.mark.parametrize("U,lo,hi", COLLISION_BANDS)
def test_riemann_symmetric_collision(U, lo, hi):
    for rho, p in ((1.0, 1.0), (0.5, 2.0)):
        c = math.sqrt(G * p / rho)
        u = U * c / math.sqrt(G)  # same Mach number for every (rho, p)
        f = O.riemann_prob(G, rho, u, p, 0.3, -0.7, rho, -u, p, 0.3, -0.7)
        # u* = 0 by symmetry: no mass, transverse momentum or energy flux
        assert f[0] == 0.0 and f[2] == 0.0 and f[3] == 0.0 and f[4] == 0.0
        pe, ue = ER.star_state(rho, u, p, rho, -u, p, G, tol=1e-16)
        assert abs(ue) <= 1e-12 * c and pe > p  # two shocks
        err = abs(f[1] - pe) / pe
        if (rho, p) == (1.0, 1.0):
            assert lo < err < hi, f"U={U}: relative error of p* {err:.3e} outside ({lo:.0e}, {hi:.0e})"
        else:
            assert err < hi


def test_riemann_collision_converges_with_strength():
    """Weaker collisions leave smaller errors (the iteration starts from the
    acoustic guess, exact as U -> 0)."""
    errs = []
    for U in (0.01, 0.1, 1.0, 2.0, 5.0):
        f = O.riemann_prob(G, 1.0, U, 1.0, 0.0, 0.0, 1.0, -U, 1.0, 0.0, 0.0)
        pe, _ = ER.star_state(1.0, U, 1.0, 1.0, -U, 1.0, G, tol=1e-16)
        errs.append(abs(f[1] - pe) / pe)
    assert errs[0] < 1e-8 and all(a <= b for a, b in zip(errs, errs[1:]))


def _perm_face_data(q3, src_dir):
    """Primitives of a z problem (src_dir 2) laid out for the x face: normal
    velocity in u, then the transverse components in the order the z face
    function takes them (z: u, v; x: v, w; Riemann.h:262-336)."""
    q = np.array(q3, dtype=np.float64).reshape(3, 6).copy()
    out = q.copy()
    if src_dir == 2:
        out[:, 1], out[:, 2], out[:, 3] = q[:, 3], q[:, 1], q[:, 2]
    return out


def _perm_flux(f, src_dir):
    """Face flux of direction src_dir in x's component order (normal
    momentum, then the transverse ones as x writes them)."""
    if src_dir == 1:  # y writes (URHO, UMY, UMX, UMZ, UET)
        return np.array([f[0], f[2], f[1], f[3], f[4]])
    if src_dir == 2:  # z writes (URHO, UMZ, UMX, UMY, UET)
        return np.array([f[0], f[3], f[1], f[2], f[4]])
    return np.asarray(f)


def _random_face(rng):
    q3 = []
    for _ in range(3):
        rho = rng.uniform(0.3, 2.0)
        p = rng.uniform(0.2, 3.0)
        T = p / (rho * P.closures_reference().Rspec)
        q3 += [rho, *rng.uniform(-0.5, 0.5, 3), T, p]
    dq2 = list(rng.uniform(-0.05, 0.05, 10))
    return np.array(q3), np.array(dq2)


@pytest.mark.parametrize("seed", range(5))
def test_riemann_z_left_sound_speed_reads_k_plus_1(seed):
    cls = P.closures_reference()
    rng = np.random.default_rng(300 + seed)
    for _ in range(20):
        q3, dq2 = _random_face(rng)
        fz = _perm_flux(O.riemann_face(cls, 2, q3, dq2), 2)
        fy = _perm_flux(O.riemann_face(cls, 1, _y_data(q3), dq2), 1)
        qx = _perm_face_data(q3, 2)
        fx = O.riemann_face(cls, 0, qx.ravel(), dq2)
        # y has no quirk: the same problem along y equals x bit for bit
        assert np.array_equal(fy, fx)
        # z reads the left cell's temperature from the cell above the face
        qq = qx.copy()
        qq[0, 4] = qx[2, 4]
        assert np.array_equal(fz, O.riemann_face(cls, 0, qq.ravel(), dq2))
        # and only that: with T(k+1) != T(k-1) z differs from x
        if qx[2, 4] != qx[0, 4]:
            assert not np.array_equal(fz, fx)
        # with T(k+1) == T(k-1) the quirk is invisible
        qs = q3.reshape(3, 6).copy()
        qs[2, 4] = qs[0, 4]
        fz2 = _perm_flux(O.riemann_face(cls, 2, qs.ravel(), dq2), 2)
        assert np.array_equal(fz2, O.riemann_face(cls, 0, _perm_face_data(qs, 2).ravel(), dq2))


def _y_data(q3_as_z):
    """The z problem's data laid out for the y face: normal velocity in v,
    first transverse (z's u) in u, second (z's v) in w."""
    q = np.array(q3_as_z, dtype=np.float64).reshape(3, 6)
    out = q.copy()
    out[:, 2] = q[:, 3]  # normal
    out[:, 1] = q[:, 1]  # t1
    out[:, 3] = q[:, 2]  # t2
    return out.ravel()
