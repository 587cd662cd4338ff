"""The C4 timestep at its full size, as the benchmark runs it: 200 RBCs
(2210 data / 8832 sample sites, Skalak tension + bending) and 400 platelets
(605 / 1552, neo-Hookean tension + bending with preferred curvature), the
moduli of test/ib-whole.cc:208-215, in a 32 x 64 x 32 um channel on a
128 x 256 x 128 MAC grid with y walls and the top wall moving at 1e-3 m/s;
cells start 1% inflated against their undeformed shapes.

A whole oracle step at this size takes the CPU many minutes, so the step is
checked through what holds at any size:
* the membrane forces of every 50th cell of each type (cells are
  independent) against the oracle's forces (oracle/step.py, oracle/forces.py;
  operator products in long double) at the same data sites: at 1% strain the
  forces are small differences of the geometry terms, so the bar is four
  times the difference between the oracle with long-double and with FP64
  products (its own FP64 rounding), and at most 1e-11 where that is smaller;
* the plain FP64 step's solves meet their tolerance (FP64 residuals) and its
  velocity is divergence-free to rounding;
* the mixed form (FP32 V-cycles in FP64 defect correction) gives the FP64
  form's u, p and X within 1e-10 relative after two steps.
"""
import os
import types as pytypes

import numpy as np
import pytest

from conftest import ROOT, _load, load_inputs, load_surfaces, load_orc_rbf

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def c4():
    bench = _load("ibm_bench", os.path.join(ROOT, "bench.py"))
    L, types = bench.c4_cells(load_inputs(), load_surfaces())
    return bench, L, types


def test_c4_forces_solves_and_mixed_form(ibm, ctx, c4):
    bench, L, types = c4
    s64, st64, ub, keep64 = bench.make_solver(ibm, ctx, L, types, 64, tol=1e-12)
    s64.step(st64, boundary=ub)
    info = s64.info()
    print("C4 FP64 step", info)
    assert max(info["residual"][:3]) <= 1e-12 and info["residual"][3] <= 1e-11
    g, _ = s64.grids()
    umax = max(float(a.abs().max()) for a in st64.u)
    res = g[0].axis[0].resolution                             # cells per unit length
    assert umax > 0 and info["divergence"] <= 1e-9 * umax * res

    step_mod = _load("orc_step", os.path.join(ROOT, "oracle", "step.py"))
    fo = _load("orc_forces", os.path.join(ROOT, "oracle", "forces.py"))
    O = load_orc_rbf()
    O64 = pytypes.SimpleNamespace(evaluate=lambda M, K, Z: O.evaluate(M, K, Z, dtype=np.float64))
    orc_view = pytypes.SimpleNamespace(rbf=O, forces_mod=fo)
    for k, T in enumerate(types):
        ct = T["ct"]
        f = s64.cells(k)["forces"].cpu().numpy()                   # (3, nc, ns)
        sub = np.arange(0, T["nc"], 50)
        X = np.ascontiguousarray(T["X"][:, sub, :]).reshape(-1)
        OT = {"E": ct.E, "Et": ct.Et, "Ep": ct.Ep, "seconds": ct.seconds, "data_ops": ct.data_ops, "nd": ct.nd,
              "ns": ct.ns, "nc": len(sub), "w": ct.w, "k": 0.0, "reference": T["reference"].reshape(-1),
              "membrane": dict(zip(("tension", "shear", "bulk", "bending", "preferred"), T["membrane"]))}
        orc_view.types = [OT]
        orc_view.rbf = O                                      # products in long double, rounded once
        geo = step_mod.Step.geometry(orc_view, [X])[0]
        ref = step_mod.Step.forces(orc_view, OT, X, geo)
        orc_view.rbf = O64                                    # the same with FP64 (BLAS) products
        ref64 = step_mod.Step.forces(orc_view, OT, X, step_mod.Step.geometry(orc_view, [X])[0])
        err, spread = _rel(f[:, sub], ref), _rel(ref64, ref)
        print(f"C4 type {k}: forces of {len(sub)} cells vs oracle, relative error {err:.2e} "
              f"(the oracle's own FP64-product spread {spread:.2e})")
        # Forces at 1% strain are small differences of the geometry terms;
        # the GPU must stay within a few times the rounding of an FP64
        # evaluation of the same formulas.
        assert err < max(1e-11, 4 * spread)

    s32, st32, ub32, keep32 = bench.make_solver(ibm, ctx, L, types, 32, tol=1e-12)
    s32.step(st32, boundary=ub32)
    s64.step(st64, boundary=ub)
    s32.step(st32, boundary=ub32)
    errs = [_rel(st32.u[i].cpu().numpy(), st64.u[i].cpu().numpy()) for i in range(3)]
    errs.append(_rel(st32.p.cpu().numpy(), st64.p.cpu().numpy()))
    errs += [_rel(a.cpu().numpy(), b.cpu().numpy()) for a, b in zip(st32.xs, st64.xs)]
    print("C4 mixed vs FP64 after two steps: u, v, w, p, X per type", errs)
    assert max(errs) < 1e-10
    s32.close()
    s64.close()
