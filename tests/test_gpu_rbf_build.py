"""Device construction of the RBF surface operators (ibm_rbf_build) and of
the sphere quadrature weights (ibm_sphere_weights).

Pins, as for the host construction (tests/test_oracle_rbf.py): the RBC map of
include/rbc.h is a degree-5 polynomial of the unit-sphere embedding, so the
interpolant reproduces it and every operator returns its exact value and
parameter derivatives (sympy); the evaluator reproduces constants and the
derivative operators annihilate them; the weights integrate the unit sphere
(sum w cos(phi) = 4 pi).  Against the host construction
(ib.cu_amd/surfaces.py, LAPACK) the operators and weights agree to the
rounding the interpolation matrix's conditioning allows (1e-7 relative;
measured 5e-9 to 2e-8).
"""
import math

import numpy as np
import pytest

from conftest import load_inputs, load_surfaces, load_orc_rbf

pytestmark = pytest.mark.gpu

I = load_inputs()
S = load_surfaces()
O = load_orc_rbf()


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _exact_rbc():
    sp = pytest.importorskip("sympy")
    t, p = sp.symbols("t p", real=True)
    x, y, z0 = sp.cos(t) * sp.cos(p), sp.sin(t) * sp.cos(p), sp.sin(p)
    r2 = x * x + y * y
    z = sp.Rational(1, 2) * z0 * (sp.Rational(21, 100) + 2 * r2 - sp.Rational(112, 100) * r2 * r2)
    X = sp.Matrix([I.RBC_RADIUS * x, I.RBC_RADIUS * z, I.RBC_RADIUS * y])
    exprs = [X, X.diff(t), X.diff(p), X.diff(t, 2), X.diff(t).diff(p), X.diff(p, 2)]
    return [sp.lambdify((t, p), list(e), "numpy") for e in exprs]


def test_device_operators_reproduce_the_rbc_map(ibm, ctx):
    nd, ns = 300, 700
    dp, sp_ = I.bauer_params(nd), I.bauer_params(ns)
    kinds = ["value", "t", "p", "tt", "tp", "pp"]
    ops = ibm.rbf_build(ctx, _t(np.stack(dp)), _t(np.stack(sp_)), 5, kinds)
    ctx.synchronize()
    ops = [o.cpu().numpy() for o in ops]
    X = I.rbc_shape(*dp)
    exact = _exact_rbc()
    for op, f in zip(ops, exact):
        want = np.array([np.broadcast_to(np.asarray(c, dtype=float), sp_[0].shape) for c in f(*sp_)])
        got = O.evaluate(op, nd, X)
        assert np.abs(got - want).max() < 1e-9 * np.abs(want).max()
    assert np.abs(ops[0][:, :nd].sum(1) - 1).max() < 1e-12
    for op in ops[1:]:
        assert np.abs(op[:, :nd].sum(1)).max() < 1e-9
    host = S.operators(dp, sp_, 5, seconds=True)
    for a, b in zip(ops, host):
        err = np.abs(a - b).max() / np.abs(b).max()
        print("device vs host operator", err)
        assert err < 1e-7


def test_device_weights_integrate_the_sphere(ibm, ctx):
    for n in (400, 1552):
        th, ph = I.bauer_params(n)
        w = ibm.sphere_weights(ctx, _t(np.stack([th, ph]))).cpu().numpy()
        assert abs((w * np.cos(ph)).sum() - 4 * math.pi) < 1e-9
        host = S.sphere_weights(th, ph)
        err = np.abs(w - host).max() / np.abs(host).max()
        print("device vs host weights", n, err)
        assert err < 1e-7


def test_platelet_operators_and_errors(ibm, ctx):
    nd, ns = 120, 400
    dp, sp_ = I.bauer_params(nd), I.bauer_params(ns)
    (E,) = ibm.rbf_build(ctx, _t(np.stack(dp)), _t(np.stack(sp_)), 0, ["value"])
    E = E.cpu().numpy()
    assert E.shape == (ns, nd + 1)
    assert np.abs(E[:, :nd].sum(1) - 1).max() < 1e-12
    with pytest.raises(ibm.IbmError):
        ibm.rbf_build(ctx, _t(np.stack(dp)), _t(np.stack(sp_)), 9, ["value"])
