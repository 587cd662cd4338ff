"""GPU parity of the immersed-boundary timestep (ibm_solver_*) against the
oracle's composition of the same scheme (oracle/step.py), plus the
boundary-data operator against its assembled CSR form.

Bars: after three FP64 steps, velocity, pressure and data-site positions
within 1e-10 relative (max norm) of the oracle (solves to rtol 1e-12 on both
sides; sparse direct solves in the oracle; measured <= 7.6e-12); the mixed
form (FP32 V-cycles inside FP64 defect correction) within 1e-12 of the FP64
form (measured 6.4e-14); the sheared channel within 1e-10 of the oracle; a
replayed HIP graph bitwise equal to the same fixed-count step run eagerly.
Channel domains run both solve paths: the direct FFT x tridiagonal solves
(the default there) and forced multigrid."""
import os

import numpy as np
import pytest

from conftest import ROOT, _load, load_orc_rbf
import step_cases as C

pytestmark = pytest.mark.gpu

fd = _load("orc_fd", os.path.join(ROOT, "oracle", "fd.py"))
step_mod = _load("orc_step", os.path.join(ROOT, "oracle", "step.py"))
O = load_orc_rbf()
REF = 16


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def types():
    return C.cell_types()


def _solver(ibm, ctx, types, walls, precision=64, rbf="f64", tol=1e-12, pseudo_pressure=False, multigrid=False):
    ops = [ibm.RbfOperator(ctx, rbf, [_t(T["E"]), _t(T["Et"]), _t(T["Ep"])], T["nd"]) for T in types]
    ct = [{"ops": op, "cells": T["nc"], "weights": _t(T["w"]), "tether": _t(T["ref"].reshape(-1)),
           "stiffness": T["k"]} for op, T in zip(ops, types)]
    s = ibm.Solver(ctx, C.domain(ibm, walls), REF, C.DT, C.NU, density=C.RHO, tolerance=tol, precision=precision,
                   max_iterations=60, cell_types=ct, pseudo_pressure=pseudo_pressure, multigrid=multigrid)
    st = s.zero_state([_t(T["X"]) for T in types])
    return s, st, ops


def _oracle(orc, types, walls, pseudo_pressure=False):
    return step_mod.Step(orc, fd, O, C.domain_spec(walls), REF, C.DT, C.NU, C.RHO, "cosine", types,
                         pseudo_pressure=pseudo_pressure)


# Channel domains (walls) are solved directly (FFT x tridiagonal) unless
# multigrid is forced; the periodic box always uses multigrid.
@pytest.mark.parametrize("walls,pseudo,mg", [(False, False, False), (True, False, False), (True, True, False),
                                             (True, False, True), (True, True, True)])
def test_three_steps_match_oracle(ibm, orc, ctx, types, walls, pseudo, mg):
    s, st, ops = _solver(ibm, ctx, types, walls, pseudo_pressure=pseudo, multigrid=mg)
    ost = _oracle(orc, types, walls, pseudo_pressure=pseudo)
    u = [np.zeros(st.u[i].numel()) for i in range(3)]
    p = np.zeros(st.p.numel())
    xs = [T["X"].copy() for T in types]
    for k in range(3):
        s.step(st)
        u, p, xs, geo = ost.step(u, p, xs)
        info = s.info()
        print("step", k, info)
        # Helmholtz solves to rtol; the projection may stop at its absolute
        # floor tol * 2 r ||u*|| / dt, a little above rtol ||b||.
        assert max(info["residual"][:3]) <= 1e-12 and info["residual"][3] <= 1e-11
        umax = max(np.abs(c).max() for c in u)
        assert info["divergence"] <= 1e-9 * umax * 20
        # the step's fused check (max |D u|, nothing stored) equals the
        # maximum of the stored divergence field bitwise
        g, _ = s.grids()
        assert info["divergence"] == float(ibm.divergence(ctx, g, st.u).abs().max())
        if walls and not mg:
            # direct solves: the residuals are independent FP64 stencil
            # evaluations at rounding level, not iteration tolerances
            assert max(info["residual"][:3]) <= 1e-14 and info["residual"][3] <= 1e-12
    errs = [_rel(st.u[i].cpu().numpy(), u[i]) for i in range(3)]
    errs.append(_rel(st.p.cpu().numpy(), p))
    errs += [_rel(st.xs[j].cpu().numpy(), xs[j]) for j in range(len(types))]
    print("relative errors u, v, w, p, x...", errs)
    assert max(errs) < 1e-10
    # Geometry of the last step against the oracle's.
    for j, (y, t1, t2) in enumerate(geo):
        c = s.cells(j)
        assert _rel(c["positions"].cpu().numpy(), y) < 1e-12
        rn, rs, ra, rv = O.geometry(y, t1, t2, types[j]["w"])
        assert _rel(c["area"].cpu().numpy(), ra) < 1e-12
        assert _rel(c["volume"].cpu().numpy(), rv) < 1e-12
    s.close()


def test_mixed_precision_matches_fp64(ibm, ctx, types):
    a, sa, _ = _solver(ibm, ctx, types, True, precision=64, tol=1e-11, multigrid=True)
    b, sb, _ = _solver(ibm, ctx, types, True, precision=32, tol=1e-11, multigrid=True)
    for _ in range(3):
        a.step(sa)
        b.step(sb)
    errs = [_rel(sb.u[i].cpu().numpy(), sa.u[i].cpu().numpy()) for i in range(3)]
    errs.append(_rel(sb.p.cpu().numpy(), sa.p.cpu().numpy()))
    errs += [_rel(x.cpu().numpy(), y.cpu().numpy()) for x, y in zip(sb.xs, sa.xs)]
    print("mixed vs FP64 relative errors", errs, b.info())
    assert max(errs) < 1e-12
    a.close()
    b.close()


@pytest.mark.parametrize("mg", [True, False])
def test_graph_replay_matches_fixed_steps(ibm, ctx, types, mg):
    """A replayed HIP graph equals the same fixed-count step run eagerly by a
    second, identically calibrated solver (each solver keeps its own
    warm-start state), bitwise, with unrelated eager work between replays."""
    import torch
    side = torch.cuda.Stream()                   # capture needs a created stream
    gctx = ibm.Context(0, stream=side)
    s, st, _ = _solver(ibm, gctx, types, True, precision=32, tol=1e-10, multigrid=mg)
    e, eager, _ = _solver(ibm, gctx, types, True, precision=32, tol=1e-10, multigrid=mg)
    torch.cuda.synchronize()
    s.step(st)                                   # calibrates the V-cycle counts
    e.step(eager)
    torch.cuda.synchronize()
    for i in range(3):
        assert torch.equal(st.u[i], eager.u[i])
    s.capture(st)
    for k in range(3):
        s.replay(st)
        torch.cuda.synchronize()
        e.step_fixed(eager)
        torch.cuda.synchronize()
        # unrelated eager work between replays
        junk = torch.randn(1 << 20, device="cuda").sum()
        torch.cuda.synchronize()
        for i in range(3):
            assert torch.equal(st.u[i], eager.u[i]), (k, i)
        assert torch.equal(st.p, eager.p)
        for x, y in zip(st.xs, eager.xs):
            assert torch.equal(x, y)
        info = s.info()
        assert max(info["residual"]) < 1e-9, info
    assert abs(st.t - eager.t) == 0
    s.close()
    e.close()
    gctx.close()


def test_fixed_step_geometry_matches_adaptive(ibm, ctx, types):
    """A fixed-count step evaluates the tangents and the surface geometry on
    a second stream; from the same positions they equal, bitwise, those of
    an adaptive step, which runs everything on the caller's stream."""
    import torch
    a, sa, _ = _solver(ibm, ctx, types, True, precision=32, tol=1e-10)
    b, sb, _ = _solver(ibm, ctx, types, True, precision=32, tol=1e-10)
    a.step(sa)
    b.step(sb)
    torch.cuda.synchronize()
    for x, y in zip(sa.xs, sb.xs):
        assert torch.equal(x, y)
    a.step_fixed(sa)
    b.step(sb)
    torch.cuda.synchronize()
    for j in range(len(types)):
        ca, cb = a.cells(j), b.cells(j)
        for key in ("positions", "forces", "normals", "sigma", "area", "volume"):
            assert torch.equal(ca[key], cb[key]), (j, key)
        for t in range(2):
            assert torch.equal(ca["tangents"][t], cb["tangents"][t]), (j, t)
    a.close()
    b.close()


@pytest.mark.parametrize("order", [0, 2])
def test_boundary_apply_matches_csr(ibm, orc, ctx, order):
    dom = C.domain(ibm, True)
    spec = C.domain_spec(True)
    rng = np.random.default_rng(5)
    for comp, align in enumerate(step_mod.MAC):
        g = ibm.grid_make(dom, REF, align)
        axes = [orc.grid_axes(spec, REF, align)[j] for j in range(3)]
        for side in (0, 1):
            B = fd.boundary_op(axes, 1, side == 0, order)
            vals = rng.standard_normal(B.shape[1])
            y0 = rng.standard_normal(g.size)
            y = ibm.boundary_apply(ctx, g, 1, side, _t(y0.copy()), values=_t(vals), alpha=0.5, order=order)
            assert _rel(y.cpu().numpy(), y0 + 0.5 * (B @ vals)) < 1e-14
            y = ibm.boundary_apply(ctx, g, 1, side, _t(y0.copy()), value=2.0, order=order)
            assert _rel(y.cpu().numpy(), y0 + B @ np.full(B.shape[1], 2.0)) < 1e-14
        # periodic view: no columns, nothing added
        y = ibm.boundary_apply(ctx, g, 0, 1, _t(np.zeros(g.size)), value=1.0, order=order)
        assert float(y.abs().max()) == 0


def test_sheared_channel_matches_oracle(ibm, orc, ctx):
    """Moving top wall (the boundary data of test/ib-whole.cc:129-152), no
    cells: three steps against the oracle."""
    import torch
    s = ibm.Solver(ctx, C.domain(ibm, True), REF, C.DT, C.NU, density=C.RHO, tolerance=1e-12, precision=32)
    g, pg = s.grids()
    st = s.zero_state([])
    wall = 0.1
    ub = [torch.zeros(g[i].size, dtype=torch.float64, device="cuda") for i in range(3)]
    ibm.boundary_apply(ctx, g[0], 1, 1, ub[0], value=wall, order=2)
    ost = _oracle(orc, [], True)
    ubn = [c.cpu().numpy() for c in ub]
    u = [np.zeros(g[i].size) for i in range(3)]
    p = np.zeros(pg.size)
    for _ in range(3):
        s.step(st, boundary=ub)
        u, p, _, _ = ost.step(u, p, [], boundary=ubn)
    err = _rel(st.u[0].cpu().numpy(), u[0])
    print("sheared channel relative error", err)
    assert err < 1e-10
    assert float(st.u[0].max()) > 0
    s.close()


def test_tracker_step_sizes(ibm, orc, ctx, types):
    """ins::tracker on the device forces: the step sizes follow the oracle's
    restatement (oracle/tracker.py) fed with the same forces; too_small is
    IBM_ERR_DT_TOO_SMALL."""
    tracker = _load("orc_tracker", os.path.join(ROOT, "oracle", "tracker.py"))
    f0 = []
    for T in types:
        y = O.evaluate(T["E"], T["nd"], T["X"].reshape(3 * T["nc"], T["nd"]))
        f0.append(-T["k"] * (y.reshape(3, T["nc"], -1) - T["ref"]))
    fmax = np.sqrt(sum(float(np.abs(f).max()) ** 2 for f in f0))
    fscale = (4 * 2 ** 2.5 * C.DT) ** 2 * (fmax + 1e-3)          # k_max = 2^2.5 k_min at the start
    ops = [ibm.RbfOperator(ctx, "f64", [_t(T["E"]), _t(T["Et"]), _t(T["Ep"])], T["nd"]) for T in types]
    ct = [{"ops": op, "cells": T["nc"], "weights": _t(T["w"]), "tether": _t(T["ref"].reshape(-1)),
           "stiffness": T["k"]} for op, T in zip(ops, types)]
    s = ibm.Solver(ctx, C.domain(ibm, True), REF, C.DT, C.NU, density=C.RHO, tolerance=1e-10, precision=32,
                   cell_types=ct, fscale=fscale)
    st = s.zero_state([_t(T["X"]) for T in types])
    ref = tracker.Tracker(C.DT, fscale)
    t = 0.0
    for _ in range(3):
        s.step(st)
        forces = [s.cells(j)["forces"].cpu().numpy() for j in range(len(types))]
        k, e = ref(forces)
        info = s.info()
        assert info["exponent"] == e and info["timestep"] == k, (info, k, e)
        t += k
    assert e == 2
    assert abs(st.t - t) <= 1e-15 * t
    s.close()
    tiny = ibm.Solver(ctx, C.domain(ibm, True), REF, C.DT, C.NU, density=C.RHO, tolerance=1e-10, precision=32,
                      cell_types=ct, fscale=fscale * 1e-4)
    st = tiny.zero_state([_t(T["X"]) for T in types])
    with pytest.raises(ibm.IbmError) as err:
        tiny.step(st)
    assert err.value.code == 5
    tiny.close()


@pytest.mark.parametrize("order", [0, 2])
def test_ins_boundary_apply_matches_csr(ibm, orc, ctx, order):
    """ibm_ins_boundary_apply = the oracle's ins::boundary (ins/boundary.h:16-56)
    on every MAC component of a walled domain, for every excluded axis."""
    dom = C.domain(ibm, True)
    spec = C.domain_spec(True)
    rng = np.random.default_rng(9)
    for align in step_mod.MAC + [step_mod.CENTRE]:
        g = ibm.grid_make(dom, REF, align)
        axes = [orc.grid_axes(spec, REF, align)[j] for j in range(3)]
        for dim in (-1, 0, 1, 2):
            B = fd.ins_boundary(axes, dim, order)
            x = rng.standard_normal(g.size)
            y0 = rng.standard_normal(g.size)
            for alpha, beta in ((0.5, 1.0), (-2.0, 0.25), (1.0, 0.0)):
                y = ibm.ins_boundary_apply(ctx, g, dim, _t(x), _t(y0.copy()), alpha=alpha, beta=beta, order=order)
                ref = alpha * (B @ x) + beta * y0
                assert _rel(y.cpu().numpy(), ref) < 1e-15, (align, dim, alpha, beta)
