"""Block-target autotuning of the multigrid's z-marching kernels
(ibm_mg_autotune / ibm_mg_shape).

The tuner only moves the z-chunk boundaries of the smoothing, residual +
restriction and prolongation + residual kernels (the tile stays the one the
grid size selects), so a solver created with tuning must give bitwise the
same iterates as one created with the built-in targets.
Checked on the grids the timestep solves on: the 128^3 periodic FP32 mixed
Helmholtz, the C4 channel's pressure grid (128 x 256 x 128 cell centres,
Neumann walls in y, singular) and an FP64 solve on a walled 64^3 grid; the
shapes chosen and the time of a fixed number of cycles either way are
printed.
"""
import numpy as np
import pytest

from conftest import load_inputs

pytestmark = pytest.mark.gpu

UM = 0.1


def _fixed(ibm, ctx, grid, op, precision, b, iters):
    import torch
    mg = ibm.Multigrid(ctx, grid, op, precision=precision)
    mg.solve_fixed(b, iters)                                  # warm-up
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    x, rr = mg.solve_fixed(b, iters)
    e1.record()
    torch.cuda.synchronize()
    shapes = mg.shapes()
    mg.close()
    return x, float(rr), e0.elapsed_time(e1), shapes


CASES = [
    ("128^3 periodic Helmholtz FP32", lambda ibm: ibm.centered_grid([ibm.periodic(32 * UM)] * 3, 128),
     (1.0, -0.05), 32, 4),
    ("C4 pressure 128x256x128 FP32", lambda ibm: ibm.centered_grid(
        [ibm.periodic(32 * UM), ibm.neumann(64 * UM), ibm.periodic(32 * UM)], 128), (0.0, 1.0), 32, 6),
    ("64^3 walled Helmholtz FP64", lambda ibm: ibm.centered_grid([ibm.dirichlet(16 * UM)] * 3, 64),
     (1.0, -0.02), 64, 4),
]


@pytest.mark.parametrize("name,make,op,precision,iters", CASES, ids=[c[0] for c in CASES])
def test_tuned_shapes_give_identical_iterates(ibm, ctx, name, make, op, precision, iters):
    import torch
    grid = make(ibm)
    rng = load_inputs().SplitMix64(71)
    b = torch.from_numpy(rng.normal(grid.size)).cuda()
    if op[0] == 0.0:
        b -= b.mean()
    try:
        ibm.mg_autotune(False, clear=True)
        x0, r0, ms0, shapes0 = _fixed(ibm, ctx, grid, op, precision, b, iters)
        ibm.mg_autotune(True, clear=True)
        x1, r1, ms1, shapes1 = _fixed(ibm, ctx, grid, op, precision, b, iters)
    finally:
        ibm.mg_autotune(True)
    assert all(t == 0 for row in shapes0 for t, _ in row)
    assert [[w for _, w in row] for row in shapes0] == [[w for _, w in row] for row in shapes1]
    print(f"{name}: {iters} cycles, built-in targets {ms0:.3f} ms, tuned {ms1:.3f} ms, relres {r1:.2e}")
    print("  (target, wide) per level for smoothing / restriction / prolongation:", shapes1)
    assert any(t in (256, 512, 1024, 2048) for t, _ in shapes1[0])
    assert r0 == r1
    assert torch.equal(x0, x1)
