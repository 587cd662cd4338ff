"""One-time construction of the RBF surface operators of a cell type.

Restates, in host FP64 (numpy / LAPACK), the setup the reference runs once
per cell type with cuSOLVER:
  * interpolation matrix [Phi P; P^T 0] over the data sites
    (bases/fill.h:97-130) with Phi = r^7 (bases/phs.h, polyharmonic_spline<7>)
    of the spherical chord metric r = sqrt(2 (1 - x.y)) between the
    parameter-space points mapped to the unit sphere
    (bases/shapes/sphere.h:18-54);
  * evaluator and first-derivative matrices solve(lu, [Phi_s P_s]) of the
    sample sites (bases/operators.h:145-177, bases/fill.h:136-163), derivatives
    taken in the sample parameters (theta, phi) (bases/rbf.h:30-41);
  * parameter-space quadrature weights from the r^1 system with a constant
    augmentation, divided by the sphere's Jacobian cos(phi) / (4 pi)
    (bases/closed_surface.h:38-68, bases/shapes/sphere.h:101-115).
The polynomial augmentation is spherical_harmonics<5> for RBCs (36 terms,
include/rbc.h:17) and polynomials<0> (the constant) for platelets
(include/platelet.h:17).  The interpolant depends only on the span of the
augmentation, not on the basis chosen for it, so the harmonics of degree
<= 5 are represented here by the equal-span monomials x^a y^b z^c,
c <= 1, a + b + c <= degree, of the unit-sphere embedding.

The matrices are returned sample-major, shape (ns, nd + np), C order: the
(nd + np) x ns column-major layout of the reference (lde = nd + np), ready
for ibm_rbf_create.
"""
import math

import numpy as np
import scipy.linalg as sla


def unit_sphere(theta, phi):
    return np.stack([np.cos(theta) * np.cos(phi), np.sin(theta) * np.cos(phi), np.sin(phi)])


def _cosines(ts, ps, td, pd):
    """x_s . x_d on the unit sphere (spherical_metric, sphere.h:21-44)."""
    return (np.cos(ts[:, None] - td[None, :]) * np.cos(ps)[:, None] * np.cos(pd)[None, :]
            + np.sin(ps)[:, None] * np.sin(pd)[None, :])


def _chord(t):
    return np.sqrt(2.0 * np.maximum(1.0 - t, 0.0))


def _exponents(degree):
    out = []
    for c in (0, 1):
        for tot in range(c, degree + 1):
            for a in range(tot - c, -1, -1):
                out.append((a, tot - c - a, c))
    return out


def poly_terms(degree):
    """Number of augmentation terms: (degree + 1)^2."""
    return len(_exponents(degree))


# Parameter derivatives of the unit-sphere embedding (x, y, z) =
# (cos t cos p, sin t cos p, sin p): first derivatives [d/dt, d/dp] and
# second derivatives [tt, tp, pp] of each coordinate.
def _embedding_derivs(theta, phi):
    ct, st, cp, sp_ = np.cos(theta), np.sin(theta), np.cos(phi), np.sin(phi)
    x, y, z = ct * cp, st * cp, sp_
    first = {0: (-y, x, np.zeros_like(x)), 1: (-ct * sp_, -st * sp_, cp)}
    second = {(0, 0): (-x, -y, np.zeros_like(x)), (0, 1): (st * sp_, -ct * sp_, np.zeros_like(x)),
              (1, 1): (-x, -y, -z)}
    return (x, y, z), first, second


def _wrt_key(wrt):
    """None, 0 / 1 (first derivative in theta / phi) or a sorted pair."""
    if wrt is None or isinstance(wrt, int):
        return wrt
    return tuple(sorted(wrt))


def _poly(theta, phi, degree, wrt=None):
    """Monomials (np x n) of the unit-sphere embedding, or their first (wrt =
    0 / 1) or second (wrt = (0, 0), (0, 1), (1, 1)) parameter derivatives,
    by the chain rule through the embedding."""
    wrt = _wrt_key(wrt)
    X, first, second = _embedding_derivs(theta, phi)

    def pw(v, k):
        return np.ones_like(v) if k < 0 else (np.ones_like(v) if k == 0 else v ** k)

    rows = []
    for e in _exponents(degree):
        # partial derivatives of x^a y^b z^c in the embedding coordinates
        def dmono(idx):
            cnt = [0, 0, 0]
            for i in idx:
                cnt[i] += 1
            coef = 1.0
            val = np.ones_like(X[0])
            for i in range(3):
                k = e[i]
                for j in range(cnt[i]):
                    coef *= (k - j)
                if coef == 0:
                    return np.zeros_like(X[0])
                val = val * pw(X[i], k - cnt[i])
            return coef * val
        if wrt is None:
            rows.append(dmono(()))
        elif isinstance(wrt, int):
            d = first[wrt]
            rows.append(sum(dmono((i,)) * d[i] for i in range(3)))
        else:
            da, db = first[wrt[0]], first[wrt[1]]
            dab = second[wrt]
            t = sum(dmono((i,)) * dab[i] for i in range(3))
            for i in range(3):
                for j in range(3):
                    t = t + dmono((i, j)) * da[i] * db[j]
            rows.append(t)
    return np.array(rows)


def _phs7_block(ts, ps, td, pd, wrt=None):
    """Phi (ns x nd) of r^7 in the chord metric r = sqrt(2 (1 - t)),
    t = x_s . x_d, or its first / second derivative in the sample
    parameters (bases/rbf.h:29-48, bases/phs.h:19-35):
        d_a r^7     = 7 r^5 (r r_a),                r r_a = -t_a
        d_ab r^7    = 35 r^3 (r r_a)(r r_b) + 7 r^5 (-t_ab)."""
    wrt = _wrt_key(wrt)
    dth = ts[:, None] - td[None, :]
    cs, ss = np.cos(dth), np.sin(dth)
    cps, sps = np.cos(ps)[:, None], np.sin(ps)[:, None]
    cpd, spd = np.cos(pd)[None, :], np.sin(pd)[None, :]
    t = cs * cps * cpd + sps * spd
    r = _chord(t)
    if wrt is None:
        return r ** 7
    dt = {0: -ss * cps * cpd, 1: -cs * sps * cpd + cps * spd}
    if isinstance(wrt, int):
        return 7.0 * r ** 5 * (-dt[wrt])
    d2t = {(0, 0): -cs * cps * cpd, (0, 1): ss * sps * cpd, (1, 1): -t}
    return 35.0 * r ** 3 * dt[wrt[0]] * dt[wrt[1]] - 7.0 * r ** 5 * d2t[wrt]


FIRST = (0, 1)
SECOND = ((0, 0), (0, 1), (1, 1))


def _interpolation_lu(td, pd, degree):
    nd = td.size
    npoly = poly_terms(degree)
    A = np.zeros((nd + npoly, nd + npoly))
    A[:nd, :nd] = _phs7_block(td, pd, td, pd)
    P = _poly(td, pd, degree)
    A[nd:, :nd] = P
    A[:nd, nd:] = P.T
    return sla.lu_factor(A, check_finite=False)


def _solve_ops(lu, td, pd, ts, ps, degree, kinds):
    out = []
    nd = td.size
    for wrt in kinds:
        F = np.empty((nd + poly_terms(degree), ts.size))
        F[:nd] = _phs7_block(ts, ps, td, pd, wrt).T
        F[nd:] = _poly(ts, ps, degree, wrt)
        E = sla.lu_solve(lu, F, check_finite=False)
        out.append(np.ascontiguousarray(E.T))
    return tuple(out)


def operators(data, sample, degree, seconds=False):
    """(evaluator, d/dtheta, d/dphi) of a cell type, each (ns, nd + np); with
    seconds also (d2/dtheta2, d2/dtheta dphi, d2/dphi2)
    (bases/operators.h:145-190).

    data, sample: (theta, phi) parameter arrays of the data and sample sites.
    """
    td, pd = (np.asarray(a, dtype=np.float64) for a in data)
    ts, ps = (np.asarray(a, dtype=np.float64) for a in sample)
    lu = _interpolation_lu(td, pd, degree)
    kinds = (None,) + FIRST + (SECOND if seconds else ())
    return _solve_ops(lu, td, pd, ts, ps, degree, kinds)


def data_operators(data, degree):
    """The five derivative operators from the data sites to themselves
    (d/dtheta, d/dphi, d2/dtheta2, d2/dtheta dphi, d2/dphi2), each
    (nd, nd + np): the data geometry of forces/bending.h:142-158 (its
    evaluator is the identity, bases/operators.h:145-148)."""
    td, pd = (np.asarray(a, dtype=np.float64) for a in data)
    lu = _interpolation_lu(td, pd, degree)
    return _solve_ops(lu, td, pd, td, pd, degree, FIRST + SECOND)


def sphere_weights(theta, phi):
    """Parameter-space quadrature weights (bases/closed_surface.h:38-68 with
    the sphere's weight cos(phi) / (4 pi), bases/shapes/sphere.h:101-115):
    sum_i w_i |x_theta x x_phi|(i) integrates over a surface of that
    parameterisation."""
    th = np.asarray(theta, dtype=np.float64)
    ph = np.asarray(phi, dtype=np.float64)
    n = th.size
    A = np.zeros((n + 1, n + 1))
    A[:n, :n] = _chord(_cosines(th, ph, th, ph))
    A[n, :n] = 1.0
    A[:n, n] = 1.0
    rhs = np.zeros(n + 1)
    rhs[n] = 1.0
    c = sla.solve(A, rhs, check_finite=False, assume_a="sym")[:n]
    return c / (np.cos(ph) / (4 * math.pi))


class CellType:
    """Reference shape, sites and operators of one cell type.

    shape: callable (theta, phi) -> 3 x n reference positions.
    """

    def __init__(self, shape, nd, ns, degree, site_fn, membrane=False):
        """membrane: also build the second-derivative sample operators
        (Euu, Euv, Evv) and the data-to-data derivative operators
        (data_ops) the tension and bending models need."""
        self.nd, self.ns, self.degree = nd, ns, degree
        self.data_params = site_fn(nd)
        self.sample_params = site_fn(ns)
        self.shape = shape
        ops = operators(self.data_params, self.sample_params, degree, seconds=membrane)
        self.E, self.Et, self.Ep = ops[:3]
        self.Euu, self.Euv, self.Evv = ops[3:] if membrane else (None, None, None)
        self.seconds = ops[3:] if membrane else ()
        self.data_ops = data_operators(self.data_params, degree) if membrane else ()
        self.lde = nd + poly_terms(degree)
        self.w = sphere_weights(*self.sample_params)

    def data_reference(self):
        return self.shape(*self.data_params)

    def sample_reference(self):
        return self.shape(*self.sample_params)
