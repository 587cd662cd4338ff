"""Pins of the oracle's membrane force restatement (oracle/forces.py) by the
mathematics the reference's formulas come from, on analytic surfaces.

Surfaces are ellipsoids x(u, v) = (a cos u cos v, b sin u cos v, c sin v)
in the reference's (theta, phi) parameterisation; their derivatives come
from sympy.  Quadrature: trapezoid in u, Gauss-Legendre in v, so sums of
smooth integrands converge spectrally.

  * zero tension on an undeformed shape (i1 = 2, i2 = 1);
  * zero bending on spheres (H^2 = K and H constant);
  * first variations: for a deformation X(eps) = X + eps dX,
        sum_k f_k . dX_k = -dE/deps
    with E the strain energy over the reference area,
        Skalak:      W = shear/4 (I1^2 + 2 I1 - 2 I2) + bulk/4 I2^2,
                     I1 = i1 - 2, I2 = i2 - 1              (Skalak et al. 1973)
        neo-Hookean: W = shear/2 (i1/sqrt(i2) - 2) + bulk/2 (sqrt(i2) - 1)^2
    (the laws' (w1, w2) are 2 dW/d(i1, i2)), and for bending the Helfrich
    energy
        E = 4 kappa int (H - H0)^2 dA
    differentiated by central differences of the quadrature (with a
    preferred curvature field H0 carried by the material points, the
    reference's force is the normal part of that variation, so it is
    checked under normal displacements).  A dropped
    term, a wrong sign or factor in any of them breaks the balance.  (The
    reference's F = -4 kappa (Lap H + 2 H (H^2 - K)) n, forces/bending.h:177,
    is the variation of 4 kappa int H^2 dA: twice the textbook
    2 kappa int H^2 dA, i.e. its `modulus` is half the Helfrich k_c.)
"""
import math
import os

import numpy as np
import pytest

from conftest import ROOT, _load

sp = pytest.importorskip("sympy")


def _forces():
    return _load("orc_forces", os.path.join(ROOT, "oracle", "forces.py"))


_U, _V, _A, _B, _C = sp.symbols("u v a b c", real=True)
_X = sp.Matrix([_A * sp.cos(_U) * sp.cos(_V), _B * sp.sin(_U) * sp.cos(_V), _C * sp.sin(_V)])
_DERIVS = {
    "x": _X, "u": _X.diff(_U), "v": _X.diff(_V),
    "uu": _X.diff(_U, 2), "uv": _X.diff(_U).diff(_V), "vv": _X.diff(_V, 2),
}
_FUN = {k: sp.lambdify((_U, _V, _A, _B, _C), list(e), "numpy") for k, e in _DERIVS.items()}


def _eval(name, u, v, abc):
    out = _FUN[name](u, v, *abc)
    return np.array([np.broadcast_to(np.asarray(c, dtype=float), u.shape) for c in out])


def quadrature(nu=64, nv=48):
    gx, gw = np.polynomial.legendre.leggauss(nv)
    v = gx * math.pi / 2
    wv = gw * math.pi / 2
    u = 2 * math.pi * np.arange(nu) / nu
    U, V = np.meshgrid(u, v, indexing="ij")
    W = np.outer(np.full(nu, 2 * math.pi / nu), wv)
    return U.ravel(), V.ravel(), W.ravel()


def geometry(abc, u, v, w):
    t = np.stack([_eval("u", u, v, abc), _eval("v", u, v, abc)])
    tt = np.stack([_eval("uu", u, v, abc), _eval("uv", u, v, abc), _eval("vv", u, v, abc)])
    cr = np.cross(t[0], t[1], axis=0)
    det = np.sqrt((cr * cr).sum(0))
    return t, tt, cr / det, w * det


def invariants(geo, geo0):
    fo = _forces()
    e, f, g = [fo.dot(a, b) for a, b in ((geo[0][0], geo[0][0]), (geo[0][0], geo[0][1]), (geo[0][1], geo[0][1]))]
    oe, of, og = [fo.dot(a, b) for a, b in ((geo0[0][0], geo0[0][0]), (geo0[0][0], geo0[0][1]),
                                            (geo0[0][1], geo0[0][1]))]
    oi = oe * og - of * of
    return (e * og + oe * g - 2 * of * f) / oi, (e * g - f * f) / oi


def W_skalak(shear, bulk):
    def W(i1, i2):
        I1, I2 = i1 - 2, i2 - 1
        return shear / 4 * (I1 * I1 + 2 * I1 - 2 * I2) + bulk / 4 * I2 * I2
    return W


def W_neohookean(shear, bulk):
    def W(i1, i2):
        j = np.sqrt(i2)
        return shear / 2 * (i1 / j - 2) + bulk / 2 * (j - 1) ** 2
    return W


@pytest.mark.parametrize("law", ["skalak", "neohookean"])
def test_tension_vanishes_on_the_reference_shape(law):
    fo = _forces()
    u, v, w = quadrature(16, 12)
    geo = geometry((1.0, 0.8, 0.4), u, v, w)
    f = fo.tension(getattr(fo, law)(2.5, 40.0), geo[0], geo[1], geo[0], geo[1], geo[3])
    assert np.abs(f).max() < 1e-12


def test_bending_vanishes_on_spheres():
    fo = _forces()
    u, v, w = quadrature(16, 12)
    for R in (0.39, 1.0, 3.0):
        t, tt, n, s = geometry((R, R, R), u, v, w)
        H, K = fo.curvatures(t, tt, n)
        assert np.allclose(H, -1 / R, rtol=1e-13) and np.allclose(K, 1 / R ** 2, rtol=1e-12)
        hd = np.zeros((5, u.size))
        f = fo.bending(1.7, t, tt, n, s, hd)
        assert np.abs(f).max() < 1e-12 * 1.7 / R ** 2


@pytest.mark.parametrize("law,energy,shear,bulk", [("skalak", W_skalak, 2.5, 250.0),
                                                    ("neohookean", W_neohookean, 1.0, 10.0)])
@pytest.mark.parametrize("ref,cur", [((1.0, 1.0, 1.0), (1.1, 1.1, 0.9)), ((1.0, 1.0, 0.5), (0.9, 0.9, 0.6))])
def test_tension_is_the_variation_of_the_strain_energy(law, energy, shear, bulk, ref, cur):
    fo = _forces()
    u, v, w = quadrature()
    geo0 = geometry(ref, u, v, w)
    W = energy(shear, bulk)

    def E(abc):
        geo = geometry(abc, u, v, w)
        i1, i2 = invariants(geo, geo0)
        return (W(i1, i2) * geo0[3]).sum()

    geo = geometry(cur, u, v, w)
    f = fo.tension(getattr(fo, law)(shear, bulk), geo[0], geo[1], geo0[0], geo0[1], geo0[3])
    for k, dX in enumerate([_eval("x", u, v, (1, 1, 0)), _eval("x", u, v, (0, 0, 1))]):
        # deform a = b (k = 0) or c (k = 1)
        eps = 1e-5
        d = np.array([1.0, 1.0, 0.0]) if k == 0 else np.array([0.0, 0.0, 1.0])
        dE = (E(tuple(np.add(cur, eps * d))) - E(tuple(np.subtract(cur, eps * d)))) / (2 * eps)
        work = (f * dX).sum()
        assert abs(work + dE) < 1e-7 * max(abs(dE), 1.0), (k, work, dE)


def _mean_curvature_of(X):
    """Symbolic H of a parameterised surface X(u, v), with a > 0 and cos v > 0
    (the Abs a square root introduces is dropped)."""
    xu, xv = X.diff(_U), X.diff(_V)
    cr = xu.cross(xv)
    nrm = cr / sp.sqrt(cr.dot(cr))
    e, f_, g = xu.dot(xu), xu.dot(xv), xv.dot(xv)
    l, k, m = X.diff(_U, 2).dot(nrm), X.diff(_U).diff(_V).dot(nrm), X.diff(_V, 2).dot(nrm)
    return ((l * g + m * e - 2 * k * f_) / (2 * (e * g - f_ * f_))).replace(sp.Abs, lambda z: z), nrm


_H_REV, _ = _mean_curvature_of(_X.subs(_B, _A))
_H_REV = sp.simplify(_H_REV).replace(sp.Abs, lambda z: z)
_H_REV_FUN = sp.lambdify((_V, _A, _C), [_H_REV, _H_REV.diff(_V), _H_REV.diff(_V, 2)], "numpy")


def _h_of_revolution(abc, v):
    """H of the ellipsoid (a, a, c) and its parameter derivatives
    (hu, hv, huu, huv, hvv) = (0, H', 0, 0, H'')."""
    H, H1, H2 = (np.broadcast_to(np.asarray(q, dtype=float), v.shape) for q in _H_REV_FUN(v, abc[0], abc[2]))
    return H, np.array([0 * v, H1, 0 * v, 0 * v, H2])


@pytest.mark.parametrize("abc,ref", [((1.0, 1.0, 0.6), None), ((1.0, 1.0, 1.3), None),
                                     ((1.0, 1.0, 0.6), (1.2, 1.2, 1.2)), ((0.9, 0.9, 1.1), (0.5, 0.5, 0.5))])
def test_bending_is_the_variation_of_the_helfrich_energy(abc, ref):
    """Ellipsoids of revolution deformed in their semi-axes, with no
    preferred curvature or the constant one of a reference sphere."""
    fo = _forces()
    kappa = 0.8
    u, v, w = quadrature(48, 64)
    H0 = 0.0
    if ref is not None:
        geo0 = geometry(ref, u, v, w)
        H0 = -1.0 / ref[0]

    def E(abc_):
        t, tt, n, s = geometry(abc_, u, v, w)
        H, _ = fo.curvatures(t, tt, n)
        return 4 * kappa * ((H - H0) ** 2 * s).sum()

    t, tt, n, s = geometry(abc, u, v, w)
    _, hd = _h_of_revolution(abc, v)
    if ref is not None:
        f = fo.bending(kappa, t, tt, n, s, hd, geo0[0], geo0[1], geo0[2])
    else:
        f = fo.bending(kappa, t, tt, n, s, hd)
    for d in (np.array([1.0, 1.0, 0.0]), np.array([0.0, 0.0, 1.0])):
        eps = 1e-5
        dE = (E(tuple(np.add(abc, eps * d))) - E(tuple(np.subtract(abc, eps * d)))) / (2 * eps)
        work = (f * _eval("x", u, v, tuple(d))).sum()
        assert abs(work + dE) < 1e-7 * max(abs(dE), 1.0), (work, dE)


def test_bending_with_a_preferred_curvature_field_is_its_normal_variation():
    """A reference ellipsoid whose H0 varies over the surface, carried by the
    material points: the force is the variation under normal displacements
    X + eps psi n (psi = 1 + sin v / 2), which leave every point's
    parameters and so its H0 in place."""
    fo = _forces()
    kappa = 0.8
    u, v, w = quadrature(48, 64)
    ref, cur = (1.0, 1.0, 0.5), (1.1, 1.1, 0.8)
    Xc = _X.subs(_B, _A)
    _, nrm = _mean_curvature_of(Xc)
    eps_s = sp.Symbol("eps", real=True)
    psi = 1 + sp.sin(_V) / 2
    Xe = Xc + eps_s * psi * nrm
    He, _ = _mean_curvature_of(Xe)
    xu, xv = Xe.diff(_U), Xe.diff(_V)
    cr = xu.cross(xv)
    dA = sp.sqrt(cr.dot(cr))
    energy_density = sp.lambdify((_U, _V, _A, _C, eps_s), [He, dA], "numpy")
    H0, hd0 = _h_of_revolution(ref, v)
    geo0 = geometry(ref, u, v, w)

    def E(eps):
        He_, dA_ = (np.broadcast_to(np.asarray(q, dtype=float), v.shape)
                    for q in energy_density(u, v, cur[0], cur[2], eps))
        return 4 * kappa * ((He_ - H0) ** 2 * dA_ * w).sum()

    t, tt, n, s = geometry(cur, u, v, w)
    _, hd = _h_of_revolution(cur, v)
    f = fo.bending(kappa, t, tt, n, s, hd - hd0, geo0[0], geo0[1], geo0[2])
    eps = 1e-5
    dE = (E(eps) - E(-eps)) / (2 * eps)
    work = (f * (1 + np.sin(v) / 2) * n).sum()
    assert abs(work + dE) < 1e-7 * max(abs(dE), 1.0), (work, dE)
