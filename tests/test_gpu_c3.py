"""The C3 workload's two heavy operators at their real sizes.

* RBF evaluation of 64 RBCs: the 8832 x 2210 evaluator applied to the
  2210 x 192 data-site coordinates (N = 3 x 64) in FP64, FP32 and split
  BF16, against the long-double product on a column subset.  Bars in units
  of (|E| |X|)_ij as in tests/test_gpu_rbf.py: 1e-13, 4e-6, 4e-5; the
  measured errors and rates are printed.
* The 128^3 pressure Poisson (cell centres of a 32 um box, periodic x, z,
  Neumann walls in y; singular, mean-free) solved with FP32 V-cycles in FP64
  defect correction to 1e-10, checked by an independent FP64 residual.
"""
import math

import numpy as np
import pytest

from conftest import load_inputs, load_orc_rbf, load_surfaces

pytestmark = pytest.mark.gpu

I = load_inputs()
S = load_surfaces()
O = load_orc_rbf()
UM = 0.1


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.fixture(scope="module")
def c3_rbf():
    dp, sp = I.bauer_params(2210), I.bauer_params(8832)
    E = S.operators(dp, sp, 5)[0]
    L = (32 * UM,) * 3
    rbcs, _ = I.cell_poses(*L, 64, 0, 3)
    X = np.stack([I.posed(I.rbc_shape(*dp), q) for q in rbcs], axis=1)      # (3, 64, 2210)
    return E, np.ascontiguousarray(X.reshape(3 * 64, 2210))


@pytest.mark.parametrize("precision,bar", [("f64", 1e-13), ("f32", 4e-6), ("bf16x3", 4e-5)])
def test_c3_rbf_evaluation(ibm, ctx, c3_rbf, precision, bar):
    import torch
    E, X = c3_rbf
    K = 2210
    op = ibm.RbfOperator(ctx, precision, [_t(E)], K)
    Xd = _t(X)
    (Y,) = ibm.rbf_eval(ctx, [op], [Xd])
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        ibm.rbf_eval(ctx, [op], [Xd], ys=[Y])
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    Y = Y.cpu().numpy()[0]                                     # (192, 8832)
    cols = np.arange(0, 192, 16)
    ref = O.evaluate(E, K, X[cols])
    scale = np.abs(X[cols]) @ np.abs(E[:, :K]).T
    err = (np.abs(Y[cols] - ref) / np.maximum(scale, 1e-300)).max()
    rel = np.abs(Y[cols] - ref).max() / np.abs(ref).max()
    flops = 2.0 * 8832 * 192 * K
    print(f"C3 RBF {precision}: scaled error {err:.2e}, relative {rel:.2e}, {ms:.3f} ms, "
          f"{flops / ms / 1e9:.1f} TFLOP/s")
    assert err < bar
    op.close()


def test_c3_pressure_poisson_mixed(ibm, ctx):
    import torch
    dom = [ibm.periodic(32 * UM), ibm.neumann(32 * UM), ibm.periodic(32 * UM)]
    pg = ibm.centered_grid(dom, 128)
    assert [a.points for a in pg.axis] == [128, 128, 128]
    rng = I.SplitMix64(33)
    b = _t(rng.normal(pg.size))
    b -= b.mean()
    mg = ibm.Multigrid(ctx, pg, (0.0, 1.0), precision=32)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    x, it, rr = mg.solve(b, rtol=1e-10, maxit=60)
    e1.record()
    torch.cuda.synchronize()
    r = b - ibm.op_apply(ctx, pg, (0.0, 1.0), x)
    ctx.synchronize()
    rel = float(r.norm() / b.norm())
    print(f"C3 pressure Poisson 128^3 mixed: {it} cycles, reported {rr:.2e}, FP64 residual {rel:.2e}, "
          f"{e0.elapsed_time(e1):.1f} ms")
    assert rel <= 1e-10 and abs(float(x.mean())) <= 1e-12 * float(x.abs().max())
    mg.close()
