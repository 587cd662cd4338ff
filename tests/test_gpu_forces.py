"""GPU parity of the membrane force models (forces.hip) against the oracle's
restatement (oracle/forces.py, pinned in tests/test_oracle_forces.py), and of
timesteps whose cells carry those models against the oracle's step.

The force passes take the same FP64 geometry on both sides (the RBF
operators evaluated on the GPU for the library, in long double for the
oracle), so the comparison covers the operators, the reference table, the
mean curvature, its derivatives and the tension / bending formulas together.
Bars: forces within 1e-12 relative (max norm) of the oracle (measured
2.3e-13 Skalak, 1.4e-13 neo-Hookean); three FP64 steps within 1e-10 of the
oracle step (measured <= 7.6e-12).
"""
import os
from types import SimpleNamespace

import numpy as np
import pytest

from conftest import ROOT, _load, load_orc_rbf
import step_cases as C

pytestmark = pytest.mark.gpu

fd = _load("orc_fd", os.path.join(ROOT, "oracle", "fd.py"))
step_mod = _load("orc_step", os.path.join(ROOT, "oracle", "step.py"))
fo = _load("orc_forces", os.path.join(ROOT, "oracle", "forces.py"))
O = load_orc_rbf()
REF = 16


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def mtypes():
    return C.membrane_cell_types()


def _ops(ibm, ctx, T):
    op = ibm.RbfOperator(ctx, "f64", [_t(m) for m in (T["E"], T["Et"], T["Ep"], *T["seconds"])], T["nd"])
    dop = ibm.RbfOperator(ctx, "f64", [_t(m) for m in T["data_ops"]], T["nd"])
    return op, dop


def _model(ibm, m):
    return ibm.membrane(m["tension"], m["shear"], m["bulk"], m["bending"], m["preferred"])


def test_membrane_forces_match_oracle(ibm, ctx, mtypes):
    for T in mtypes:
        nd, ns, nc = T["nd"], T["ns"], T["nc"]
        op, dop = _ops(ibm, ctx, T)
        X = _t(T["X"])
        ref = _t(T["reference"])
        w = _t(T["w"])
        (Y,) = ibm.rbf_eval(ctx, [op], [X])                  # (6, 3 nc, ns)
        (R,) = ibm.rbf_eval(ctx, [op], [ref])                # (6, 3, ns)
        table = ibm.membrane_reference(ctx, ns, R[1:].contiguous(), w)
        normal, sigma, _, _ = ibm.surface_geometry(ctx, ns, nc, Y[0], Y[1], Y[2], w)
        (Yd,) = ibm.rbf_eval(ctx, [dop], [X])                # (5, 3 nc, nd)
        h0 = None
        if T["membrane"]["preferred"]:
            (Rd,) = ibm.rbf_eval(ctx, [dop], [ref])
            h0 = ibm.mean_curvature(ctx, nd, 1, Rd)
        h = ibm.mean_curvature(ctx, nd, nc, Yd, h0)
        (hd,) = ibm.rbf_eval(ctx, [op], [h])                 # (6, nc, ns)
        f = ibm.membrane_forces(ctx, _model(ibm, T["membrane"]), ns, nc, Y[1:].contiguous(), normal, sigma,
                                hd[1:].contiguous(), table)
        ctx.synchronize()
        holder = SimpleNamespace(rbf=O, forces_mod=fo)
        y = O.evaluate(T["E"], nd, T["X"].reshape(3 * nc, nd)).reshape(3, nc, ns)
        want = step_mod.Step.forces(holder, T, T["X"], (y,))
        err = _rel(f.cpu().numpy().reshape(3, nc, ns), want)
        print("membrane forces", T["membrane"]["tension"], "relative error", err, "max |f|", np.abs(want).max())
        assert err < 1e-12
        # beta accumulates into existing contents
        f2 = ibm.membrane_forces(ctx, _model(ibm, T["membrane"]), ns, nc, Y[1:].contiguous(), normal, sigma,
                                 hd[1:].contiguous(), table, out=f.clone(), beta=1.0)
        assert np.allclose(f2.cpu().numpy(), 2 * f.cpu().numpy(), rtol=1e-15, atol=0)


def _solver(ibm, ctx, types, walls, precision=64, tol=1e-12):
    keep = []
    ct = []
    for T in types:
        op, dop = _ops(ibm, ctx, T)
        keep += [op, dop]
        ct.append({"ops": op, "cells": T["nc"], "weights": _t(T["w"]), "membrane": _model(ibm, T["membrane"]),
                   "data_ops": dop, "reference": _t(T["reference"])})
    s = ibm.Solver(ctx, C.domain(ibm, walls), REF, C.DT, C.NU, density=C.RHO, tolerance=tol, precision=precision,
                   max_iterations=60, cell_types=ct)
    st = s.zero_state([_t(T["X"]) for T in types])
    return s, st, keep


@pytest.mark.parametrize("walls", [False, True])
def test_membrane_steps_match_oracle(ibm, orc, ctx, mtypes, walls):
    s, st, keep = _solver(ibm, ctx, mtypes, walls)
    ost = step_mod.Step(orc, fd, O, C.domain_spec(walls), REF, C.DT, C.NU, C.RHO, "cosine", mtypes, forces=fo)
    u = [np.zeros(st.u[i].numel()) for i in range(3)]
    p = np.zeros(st.p.numel())
    xs = [T["X"].copy() for T in mtypes]
    for k in range(3):
        s.step(st)
        u, p, xs, geo = ost.step(u, p, xs)
        info = s.info()
        print("step", k, info)
        assert max(info["residual"][:3]) <= 1e-12 and info["residual"][3] <= 1e-11
    errs = [_rel(st.u[i].cpu().numpy(), u[i]) for i in range(3)]
    errs.append(_rel(st.p.cpu().numpy(), p))
    errs += [_rel(st.xs[j].cpu().numpy(), xs[j]) for j in range(len(mtypes))]
    print("relative errors u, v, w, p, x...", errs)
    assert max(errs) < 1e-10
    s.close()


def test_membrane_model_errors(ibm, ctx, mtypes):
    T = mtypes[0]
    op3 = ibm.RbfOperator(ctx, "f64", [_t(m) for m in (T["E"], T["Et"], T["Ep"])], T["nd"])
    ct = [{"ops": op3, "cells": T["nc"], "weights": _t(T["w"]), "membrane": _model(ibm, T["membrane"]),
           "reference": _t(T["reference"])}]
    with pytest.raises(ibm.IbmError) as e:
        ibm.Solver(ctx, C.domain(ibm, True), REF, C.DT, C.NU, density=C.RHO, cell_types=ct)
    assert e.value.code == ibm.ERR_ARG
    bad = ibm.membrane("skalak", 1.0, 1.0)
    bad.tension = 7
    op, dop = _ops(ibm, ctx, T)
    ct = [{"ops": op, "cells": T["nc"], "weights": _t(T["w"]), "membrane": bad, "reference": _t(T["reference"])}]
    with pytest.raises(ibm.IbmError) as e:
        ibm.Solver(ctx, C.domain(ibm, True), REF, C.DT, C.NU, density=C.RHO, cell_types=ct)
    assert e.value.code == ibm.ERR_ARG
