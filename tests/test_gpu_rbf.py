"""GPU parity of the RBF surface evaluation (ibm_rbf_eval) and surface
geometry (ibm_surface_geometry) against the oracle (oracle/rbf.py).

Bars, elementwise against the long-double product, in units of
(|E| |X|)_ij = sum_k |E_ki| |X_kj| (the scale of the rounding of any
summation order):
  FP64 matrix cores        1e-13
  FP32                      4e-6
  split BF16 (3 products)   4e-5
Geometry from FP64 evaluations: 1e-12 relative.  The measured errors of the
reduced-precision paths on the real RBC evaluator are printed."""
import math

import numpy as np
import pytest

from conftest import load_inputs, load_orc_rbf, load_surfaces

pytestmark = pytest.mark.gpu

S = load_surfaces()
I = load_inputs()
O = load_orc_rbf()

BARS = {"f64": 1e-13, "f32": 4e-6, "bf16x3": 4e-5}


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _scaled_err(Y, Et, K, X):
    ref = O.evaluate(Et, K, X)
    scale = np.abs(X) @ np.abs(Et[:, :K]).T
    return (np.abs(Y - ref) / np.maximum(scale, 1e-300)).max()


@pytest.fixture(scope="module")
def rbc_ops():
    dp = I.bauer_params(400)
    sp = I.bauer_params(1500)
    return dp, sp, S.operators(dp, sp, 5)


def _cells(dp, nc, seed):
    rng = np.random.default_rng(seed)
    X = I.rbc_shape(*dp)
    cells = []
    for c in range(nc):
        ang = rng.uniform(0, 2 * math.pi)
        p = I.rotate(X, ang, rng.standard_normal(3)) + rng.uniform(0, 3, 3)[:, None]
        cells.append(p)
    return np.stack(cells, axis=1).reshape(3 * nc, -1)    # rows comp*nc + cell


@pytest.mark.parametrize("precision", ["f64", "f32", "bf16x3"])
def test_rbf_eval_matches_oracle(ibm, ctx, rbc_ops, precision):
    dp, sp, mats = rbc_ops
    nd, ns = dp[0].size, sp[0].size
    X = _cells(dp, 45, 1)                                  # N = 135: two column tiles, one partial
    op = ibm.RbfOperator(ctx, precision, [_t(m) for m in mats], nd)
    (Y,) = ibm.rbf_eval(ctx, [op], [_t(X)])
    Y = Y.cpu().numpy()
    errs = [_scaled_err(Y[j], mats[j], nd, X) for j in range(3)]
    print(precision, "scaled errors", errs)
    assert max(errs) < BARS[precision]
    op.close()


@pytest.mark.parametrize("precision", ["f64", "bf16x3"])
def test_rbf_odd_shapes_and_batch(ibm, ctx, precision):
    rng = np.random.default_rng(2)
    shapes = [(300, 77, 5), (129, 64, 1), (40, 1, 130)]            # (M, K, N): padding in every dim
    ops, xs, mats_all = [], [], []
    for M, K, N in shapes:
        lde = K + 3
        mats = [rng.standard_normal((M, lde)) for _ in range(2)]
        mats_all.append(mats)
        ops.append(ibm.RbfOperator(ctx, precision, [_t(m) for m in mats], K))
        xs.append(rng.standard_normal((N, K)))
    Ys = ibm.rbf_eval(ctx, ops, [_t(x) for x in xs])
    for (M, K, N), mats, x, Y, op in zip(shapes, mats_all, xs, Ys, ops):
        Y = Y.cpu().numpy()
        for j in range(2):
            assert _scaled_err(Y[j], mats[j], K, x) < BARS[precision]
        # The same product alone in its own launch is bitwise identical.
        (Y1,) = ibm.rbf_eval(ctx, [op], [_t(x)])
        assert np.array_equal(Y1.cpu().numpy(), Y)
    for op in ops:
        op.close()


def test_rbf_errors(ibm, ctx):
    import torch
    m = torch.zeros((10, 4), dtype=torch.float64, device="cuda")
    with pytest.raises(ibm.IbmError):
        ibm.RbfOperator(ctx, 8, [m], 4)
    with pytest.raises(ibm.IbmError):
        ibm.RbfOperator(ctx, "f64", [m], 5)                       # K > lde


def test_surface_geometry_matches_oracle_and_sphere(ibm, ctx, rbc_ops):
    dp, sp, mats = rbc_ops
    nd, ns = dp[0].size, sp[0].size
    w = S.sphere_weights(*sp)
    nc = 3
    radii = np.array([1.0, 2.0, 0.5])
    unit = S.unit_sphere(*dp)                                         # 3 x nd
    X = (unit[:, None, :] * radii[None, :, None]).reshape(3 * nc, nd)  # rows comp*nc + cell
    op = ibm.RbfOperator(ctx, "f64", [_t(m) for m in mats], nd)
    (Y,) = ibm.rbf_eval(ctx, [op], [_t(X)])
    y, t1, t2 = (Y[j].reshape(3, nc, ns).contiguous() for j in range(3))
    normal, sigma, area, volume = ibm.surface_geometry(ctx, ns, nc, y, t1, t2, _t(w))
    Yc = Y.cpu().numpy()
    rn, rs, ra, rv = O.geometry(*(Yc[j].reshape(3, nc, ns) for j in range(3)), w)
    assert np.abs(normal.cpu().numpy() - rn).max() < 1e-12
    assert np.abs(sigma.cpu().numpy() - rs).max() < 1e-12 * np.abs(rs).max()
    assert np.allclose(area.cpu().numpy(), ra, rtol=1e-12)
    assert np.allclose(volume.cpu().numpy(), rv, rtol=1e-12)
    assert np.allclose(area.cpu().numpy(), 4 * math.pi * radii ** 2, rtol=1e-9)
    assert np.allclose(volume.cpu().numpy(), 4 * math.pi / 3 * radii ** 3, rtol=1e-9)
    op.close()


@pytest.mark.parametrize("shape", [(8832, 2210, 600), (1552, 605, 1200), (300, 77, 5)],
                         ids=["rbc-c4", "platelet-c4", "small"])
def test_f64_kernel_variants_agree_bitwise(ibm, ctx, shape):
    """ibm_rbf_gemm_variant: the 4-wave (64 x 64) and the double-buffered
    8-wave (64 x 32) FP64 kernels and the timed choice give bitwise the same
    products, at the
    C4 shapes (200 RBCs, 400 platelets: N = 3 x cells) and a padded one."""
    import torch
    M, K, N = shape
    rng = np.random.default_rng(M + K)
    mats = [rng.standard_normal((M, K)) for _ in range(2)]
    x = rng.standard_normal((N, K))
    op = ibm.RbfOperator(ctx, "f64", [_t(m) for m in mats], K)
    xd = _t(x)
    out, ms = {}, {}
    try:
        for wn in (2, 4, 0):
            ibm.rbf_gemm_variant(wn)
            (Y,) = ibm.rbf_eval(ctx, [op], [xd])
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(5):
                ibm.rbf_eval(ctx, [op], [xd], ys=[Y])
            e1.record()
            torch.cuda.synchronize()
            out[wn], ms[wn] = Y.cpu().numpy(), e0.elapsed_time(e1) / 5
    finally:
        ibm.rbf_gemm_variant(0)
    flops = 2.0 * M * N * K * 2
    print(f"FP64 {M}x{K}x{N} x2: 4 waves {ms[2]:.3f} ms ({flops / ms[2] / 1e9:.1f} TFLOP/s), "
          f"8 waves {ms[4]:.3f} ms ({flops / ms[4] / 1e9:.1f} TFLOP/s), timed choice {ms[0]:.3f} ms")
    assert np.array_equal(out[2], out[4]) and np.array_equal(out[2], out[0])
    cols = np.arange(0, N, max(1, N // 8))
    for j in range(2):
        assert _scaled_err(out[4][j][cols], mats[j], K, x[cols]) < BARS["f64"]
    with pytest.raises(ibm.IbmError):
        ibm.rbf_gemm_variant(3)
    op.close()
