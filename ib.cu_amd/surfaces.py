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


def _poly(theta, phi, degree, wrt=None):
    """Monomials (np x n) of the unit-sphere embedding, or their derivative
    in theta (wrt=0) or phi (wrt=1)."""
    x, y, z = unit_sphere(theta, phi)
    if wrt is None:
        dx = dy = dz = None
    elif wrt == 0:
        dx, dy, dz = -y, x, np.zeros_like(x)
    else:
        dx, dy, dz = -np.cos(theta) * np.sin(phi), -np.sin(theta) * np.sin(phi), np.cos(phi)

    def pw(v, k):
        return np.ones_like(v) if k == 0 else v ** k

    rows = []
    for a, b, c in _exponents(degree):
        if wrt is None:
            rows.append(pw(x, a) * pw(y, b) * pw(z, c))
        else:
            t = np.zeros_like(x)
            if a:
                t = t + a * pw(x, a - 1) * dx * pw(y, b) * pw(z, c)
            if b:
                t = t + b * pw(x, a) * pw(y, b - 1) * dy * pw(z, c)
            if c:
                t = t + c * pw(x, a) * pw(y, b) * pw(z, c - 1) * dz
            rows.append(t)
    return np.array(rows)


def _phs7_block(ts, ps, td, pd, wrt=None):
    """Phi (ns x nd) of r^7, or its derivative in the sample's theta / phi:
    d r^7 = 7 r^5 (r dr) with r dr = -dt (bases/phs.h:27-38, bases/rbf.h:30-41)."""
    t = _cosines(ts, ps, td, pd)
    r = _chord(t)
    if wrt is None:
        return r ** 7
    if wrt == 0:
        dt = -np.sin(ts[:, None] - td[None, :]) * np.cos(ps)[:, None] * np.cos(pd)[None, :]
    else:
        dt = (-np.cos(ts[:, None] - td[None, :]) * np.sin(ps)[:, None] * np.cos(pd)[None, :]
              + np.cos(ps)[:, None] * np.sin(pd)[None, :])
    return 7.0 * r ** 5 * (-dt)


def operators(data, sample, degree):
    """(evaluator, d/dtheta, d/dphi) of a cell type: each (ns, nd + np).

    data, sample: (theta, phi) parameter arrays of the data and sample sites.
    """
    td, pd = (np.asarray(a, dtype=np.float64) for a in data)
    ts, ps = (np.asarray(a, dtype=np.float64) for a in sample)
    nd, ns = td.size, ts.size
    npoly = poly_terms(degree)
    A = np.zeros((nd + npoly, nd + npoly))
    A[:nd, :nd] = _phs7_block(td, pd, td, pd)
    P = _poly(td, pd, degree)
    A[nd:, :nd] = P
    A[:nd, nd:] = P.T
    lu = sla.lu_factor(A, check_finite=False)
    out = []
    for wrt in (None, 0, 1):
        F = np.empty((nd + npoly, ns))
        F[:nd] = _phs7_block(ts, ps, td, pd, wrt).T
        F[nd:] = _poly(ts, ps, degree, wrt)
        E = sla.lu_solve(lu, F, check_finite=False)
        out.append(np.ascontiguousarray(E.T))
    return tuple(out)


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

    def __init__(self, shape, nd, ns, degree, site_fn):
        self.nd, self.ns, self.degree = nd, ns, degree
        self.data_params = site_fn(nd)
        self.sample_params = site_fn(ns)
        self.shape = shape
        self.E, self.Et, self.Ep = operators(self.data_params, self.sample_params, degree)
        self.lde = nd + poly_terms(degree)
        self.w = sphere_weights(*self.sample_params)

    def data_reference(self):
        return self.shape(*self.data_params)

    def sample_reference(self):
        return self.shape(*self.sample_params)
