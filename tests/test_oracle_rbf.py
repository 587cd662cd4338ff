"""Pins the RBF operator construction (ib.cu_amd/surfaces.py) and the
oracle's evaluation / geometry (oracle/rbf.py) against what the mathematics
fixes: the interpolant reproduces its polynomial augmentation exactly (the
RBC map of include/rbc.h:19-35 is a degree-5 polynomial of the unit-sphere
embedding), derivative operators differentiate it exactly, evaluation at
the data sites is the identity, and the unit sphere has normal x, area
4 pi and volume 4 pi / 3."""
import math

import numpy as np
import pytest

from conftest import load_inputs, load_orc_rbf, load_surfaces

S = load_surfaces()
I = load_inputs()
O = load_orc_rbf()


def _rbc_derivs(theta, phi):
    h = 1e-5
    d_t = (I.rbc_shape(theta + h, phi) - I.rbc_shape(theta - h, phi)) / (2 * h)
    d_p = (I.rbc_shape(theta, phi + h) - I.rbc_shape(theta, phi - h)) / (2 * h)
    return d_t, d_p


@pytest.fixture(scope="module")
def rbc_small():
    dp = I.bauer_params(300)
    sp = I.bauer_params(700)
    return dp, sp, S.operators(dp, sp, 5)


def test_poly_terms():
    assert [S.poly_terms(d) for d in range(6)] == [1, 4, 9, 16, 25, 36]


def test_evaluator_reproduces_constants_and_rbc_shape(rbc_small):
    dp, sp, (E, Et, Ep) = rbc_small
    nd = dp[0].size
    assert np.abs(E[:, :nd].sum(1) - 1).max() < 1e-13
    assert np.abs(Et[:, :nd].sum(1)).max() < 1e-12 and np.abs(Ep[:, :nd].sum(1)).max() < 1e-12
    X = I.rbc_shape(*dp)                                  # 3 x nd
    Y = O.evaluate(E, nd, X)                              # 3 x ns
    exact = I.rbc_shape(*sp)
    assert np.abs(Y - exact).max() < 1e-12 * np.abs(exact).max()
    d_t, d_p = _rbc_derivs(*sp)
    assert np.abs(O.evaluate(Et, nd, X) - d_t).max() < 1e-8 * np.abs(d_t).max()
    assert np.abs(O.evaluate(Ep, nd, X) - d_p).max() < 1e-8 * np.abs(d_p).max()


def test_data_to_data_is_identity():
    dp = I.bauer_params(150)
    E, _, _ = S.operators(dp, dp, 5)
    # Rounding amplified by the conditioning of the r^7 interpolation matrix.
    assert np.abs(E[:, :150] - np.eye(150)).max() < 1e-8


def test_platelet_constant_augmentation():
    dp, sp = I.bauer_params(120), I.bauer_params(400)
    E, Et, Ep = S.operators(dp, sp, 0)
    assert E.shape == (400, 121)
    assert np.abs(E[:, :120].sum(1) - 1).max() < 1e-13


def test_unit_sphere_geometry_closed_forms(rbc_small):
    dp, sp, (E, Et, Ep) = rbc_small
    nd, ns = dp[0].size, sp[0].size
    w = S.sphere_weights(*sp)
    assert abs((w * np.cos(sp[1])).sum() - 4 * math.pi) < 1e-12
    X = np.stack([S.unit_sphere(*dp)] * 2, axis=1).reshape(6, nd)     # 2 cells, rows comp*nc + cell
    y, t1, t2 = (O.evaluate(M, nd, X).reshape(3, 2, ns) for M in (E, Et, Ep))
    normal, sigma, area, volume = O.geometry(y, t1, t2, w)
    u = S.unit_sphere(*sp)
    assert np.abs(normal - u[:, None, :]).max() < 1e-10
    assert np.abs(sigma - w * np.cos(sp[1])).max() < 1e-10
    assert np.allclose(area, 4 * math.pi, rtol=1e-10)
    assert np.allclose(volume, 4 * math.pi / 3, rtol=1e-10)


def test_long_double_evaluation_is_exact_on_integers():
    rng = np.random.default_rng(0)
    Et = rng.integers(-5, 6, (37, 20)).astype(np.float64)
    X = rng.integers(-5, 6, (4, 17)).astype(np.float64)
    assert np.array_equal(O.evaluate(Et, 17, X), X @ Et[:, :17].T)


def _rbc_symbolic_derivs():
    """Exact parameter derivatives of the RBC map (sympy), keyed like the
    operators: 0, 1, (0, 0), (0, 1), (1, 1)."""
    sp = pytest.importorskip("sympy")
    t, p = sp.symbols("t p", real=True)
    x, y, z0 = sp.cos(t) * sp.cos(p), sp.sin(t) * sp.cos(p), sp.sin(p)
    r2 = x * x + y * y
    z = sp.Rational(1, 2) * z0 * (sp.Rational(21, 100) + 2 * r2 - sp.Rational(112, 100) * r2 * r2)
    X = sp.Matrix([I.RBC_RADIUS * x, I.RBC_RADIUS * z, I.RBC_RADIUS * y])
    keys = {0: X.diff(t), 1: X.diff(p), (0, 0): X.diff(t, 2), (0, 1): X.diff(t).diff(p), (1, 1): X.diff(p, 2)}
    return {k: sp.lambdify((t, p), list(v), "numpy") for k, v in keys.items()}


def _call(f, th, ph):
    return np.array([np.broadcast_to(np.asarray(c, dtype=float), th.shape) for c in f(th, ph)])


def test_second_derivative_operators_differentiate_the_rbc_shape_exactly():
    """The RBC map is a degree-5 polynomial of the unit-sphere embedding, so
    the interpolant reproduces it and every derivative operator returns its
    exact parameter derivatives (bases/operators.h:168-190)."""
    dp = I.bauer_params(300)
    sp_ = I.bauer_params(700)
    ops = S.operators(dp, sp_, 5, seconds=True)
    assert len(ops) == 6
    X = I.rbc_shape(*dp)
    exact = _rbc_symbolic_derivs()
    for op, key in zip(ops[1:], [0, 1, (0, 0), (0, 1), (1, 1)]):
        want = _call(exact[key], *sp_)
        got = O.evaluate(op, 300, X)
        assert np.abs(got - want).max() < 1e-9 * np.abs(want).max(), key
    # data-to-data derivative operators (the data geometry of bending)
    dops = S.data_operators(dp, 5)
    for op, key in zip(dops, [0, 1, (0, 0), (0, 1), (1, 1)]):
        want = _call(exact[key], *dp)
        got = O.evaluate(op, 300, X)
        assert np.abs(got - want).max() < 1e-8 * np.abs(want).max(), key
    # derivatives annihilate constants
    for op in ops[1:] + dops:
        assert np.abs(op[:, :300].sum(1)).max() < 1e-9
