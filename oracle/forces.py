"""Oracle (test infrastructure only): numpy restatement of the reference's
membrane force models, evaluated point by point from surface geometry.

Geometry at a surface point (bases/geometry.h:74-123, forces/load.h):
  t = (x_u, x_v)                   first parameter derivatives (u = theta, v = phi)
  tt = (x_uu, x_uv, x_vv)          second derivatives, in this order
                                   (bases/operators.h:168-190)
  n = x_u x x_v / |x_u x x_v|,  s = w |x_u x x_v|   (sigma, the patch area)
The reference geometry is one undeformed cell shared by every cell of the
type: the current point k reads reference point k % ns (forces/load.h:96-110).

  metric_terms   forces/tension.h:34-48 and bending.h:74-88 (`helper`):
                 e, f, g = x_u.x_u, x_u.x_v, x_v.x_v, their u / v
                 derivatives from the Christoffel symbols
                 Gamma_{ij,k} = x_ij . x_k (forces/load.h:293-333), the
                 determinant i = eg - f^2 and its derivatives
  tension        forces/tension.h:50-99 with a constitutive law
                 w(i1, i2) -> (w1, w2, w11, w12, w22):
                   skalak      forces/skalak.h:11-36
                   neohookean  forces/neohookean.h:10-34
  mean / gaussian curvature   forces/bending.h:48-72
  bending        forces/bending.h:142-223:
                 F = -4 kappa s (Lap_s(H - H0) + 2 (H - H0)(H^2 - K + H H0)) n,
                 H at the data sites, its parameter derivatives evaluated at
                 the sample sites by the RBF derivative operators, and the
                 Laplace-Beltrami operator written with the metric terms.

Pinned in tests/test_oracle_forces.py by the mathematics: zero tension on an
undeformed shape; zero bending on spheres of any radius; and the first
variation of the energies the laws derive from, -dE/de = sum_k f_k . dX_k,
on analytic ellipsoids of revolution with spectrally accurate quadrature
(Skalak's and the neo-Hookean strain energies over the reference area,
and 4 kappa int (H - H0)^2 dA for bending).
"""
import numpy as np


def dot(a, b):
    return (a * b).sum(0)


def metric_terms(t, tt):
    """(e, f, g, eu, fu, gu, ev, fv, gv, i, iu, iv) per point.
    t: (2, 3, P), tt: (3, 3, P)."""
    u, v = t
    uu, uv, vv = tt
    e, f, g = dot(u, u), dot(u, v), dot(v, v)
    eu = 2 * dot(uu, u)
    fu = dot(uu, v) + dot(uv, u)
    gu = 2 * dot(uv, v)
    ev = 2 * dot(uv, u)
    fv = dot(uv, v) + dot(vv, u)
    gv = 2 * dot(vv, v)
    i = e * g - f * f
    iu = eu * g + e * gu - 2 * f * fu
    iv = ev * g + e * gv - 2 * f * fv
    return e, f, g, eu, fu, gu, ev, fv, gv, i, iu, iv


def skalak(shear, bulk):
    def w(i1, i2):
        one = np.ones_like(i1)
        return shear * (i1 - 1), bulk * (i2 - 1) - shear, shear * one, 0 * one, bulk * one
    return w


def neohookean(shear, bulk):
    def w(i1, i2):
        j = np.sqrt(i2)
        w1 = shear / j
        w2 = bulk * (1 - 1 / j) - 0.5 * w1 * i1 / i2
        return w1, w2, 0 * w1, -0.5 * w1 / i2, 0.75 * w1 * i1 / (i2 * i2) + 0.5 * bulk / (i2 * j)
    return w


def _ref(a, P):
    """Reference quantity (last axis ns) broadcast to P = ns * nc points."""
    ns = a.shape[-1]
    return np.take(a, np.arange(P) % ns, axis=-1)


def tension(law, t, tt, t0, tt0, s0):
    """Force (3, P) of a tension law at current geometry (t, tt) against the
    reference (t0, tt0, s0) (forces/tension.h:50-99)."""
    P = t.shape[-1]
    t0, tt0, s0 = _ref(t0, P), _ref(tt0, P), _ref(s0, P)
    oe, of, og, oeu, ofu, ogu, oev, ofv, ogv, oi, oiu, oiv = metric_terms(t0, tt0)
    ce, cf, cg, ceu, cfu, cgu, cev, cfv, cgv, ci, ciu, civ = metric_terms(t, tt)
    i1 = (ce * og + oe * cg - 2 * of * cf) / oi
    i2 = ci / oi
    i1u = (ceu * og + ce * ogu + oeu * cg + oe * cgu - 2 * (ofu * cf + of * cfu) - i1 * oiu) / oi
    i1v = (cev * og + ce * ogv + oev * cg + oe * cgv - 2 * (ofv * cf + of * cfv) - i1 * oiv) / oi
    i2u = (ciu - i2 * oiu) / oi
    i2v = (civ - i2 * oiv) / oi
    w1, w2, w11, w12, w22 = law(i1, i2)
    w1u = w11 * i1u + w12 * i2u
    w1v = w11 * i1v + w12 * i2v
    w2u = w12 * i1u + w22 * i2u
    w2v = w12 * i1v + w22 * i2v
    u, v = t
    uu, uv, vv = tt
    A = w1 * og + w2 * cg
    B = w1 * of + w2 * cf
    C = w1 * oe + w2 * ce
    coef_u = (-A * oiu / (2 * oi) + B * oiv / (2 * oi)
              + (w1u * og + w1 * ogu + w2u * cg + w2 * cgu)
              - (w1v * of + w1 * ofv + w2v * cf + w2 * cfv))
    coef_v = (B * oiu / (2 * oi) - C * oiv / (2 * oi)
              - (w1u * of + w1 * ofu + w2u * cf + w2 * cfu)
              + (w1v * oe + w1 * oev + w2v * ce + w2 * cev))
    return s0 / oi * (A * uu - 2 * B * uv + C * vv + coef_u * u + coef_v * v)


def curvatures(t, tt, n):
    """(H, K): mean and Gaussian curvature (forces/bending.h:48-72)."""
    e, f, g = dot(t[0], t[0]), dot(t[0], t[1]), dot(t[1], t[1])
    l, k, m = dot(tt[0], n), dot(tt[1], n), dot(tt[2], n)
    den = e * g - f * f
    return (l * g + m * e - 2 * k * f) / (2 * den), (l * m - k * k) / den


def mean_curvature_sites(t, tt, n, t0=None, tt0=None, n0=None):
    """h at data sites: H of the current geometry, minus the reference's H
    when a preferred (reference) curvature is given (bending.h:150-158)."""
    h, _ = curvatures(t, tt, n)
    if t0 is not None:
        h0, _ = curvatures(_ref(t0, h.size), _ref(tt0, h.size), _ref(n0, h.size))
        h = h - h0
    return h


def bending(modulus, t, tt, n, s, hd, t0=None, tt0=None, n0=None):
    """Force (3, P) of the bending model at the sample sites: hd = (hu, hv,
    huu, huv, hvv), the parameter derivatives of h (forces/bending.h:160-223).
    With a reference geometry (t0, tt0, n0) the preferred curvature is the
    reference's."""
    e, f, g, eu, fu, gu, ev, fv, gv, i, iu, iv = metric_terms(t, tt)
    hu, hv, huu, huv, hvv = hd
    lh = ((((-iu * g + iv * f) / (2 * i) + gu - fv) / i * hu
           + ((iu * f - iv * e) / (2 * i) - fu + ev) / i * hv)
          + g / i * huu - 2 * f / i * huv + e / i * hvv)
    H, K = curvatures(t, tt, n)
    dh = H
    sh = H * H - K
    if t0 is not None:
        P = H.size
        H0, _ = curvatures(_ref(t0, P), _ref(tt0, P), _ref(n0, P))
        dh = dh - H0
        sh = sh + H * H0
    b = lh + 2 * dh * sh
    return -4 * s * modulus * b * n
