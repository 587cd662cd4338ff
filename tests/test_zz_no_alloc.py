"""ibm_spread_sets_noalloc (the spread with the library's fixed per-device
workspace instead of allocated scratch) against ibm_spread_sets.

Both run the same binning and brick kernels, whose results do not depend on
scheduling, so the bar is bitwise equality.  Two calls in a row, the second
with fewer markers on a smaller grid, show that the shared workspace keeps no
state that leaks from one call into the next.  Oversized inputs are refused
on the host before any work is queued.
"""
import numpy as np
import pytest

from conftest import load_inputs

UM = 0.1

pytestmark = pytest.mark.gpu


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _two_sets(inputs, n0, n1, L, seed):
    x0 = inputs.random_markers(n0, L, seed)
    x1 = inputs.random_markers(n1, L, seed + 1)
    rng = inputs.SplitMix64(seed + 2)
    return [x0, x1], [rng.normal(3 * n0), rng.normal(3 * n1)]


def _compare(ibm, ctx, grids, xs, fs, delta="cosine"):
    import torch
    xs_d = [_t(x) for x in xs]
    fs_d = [_t(f) for f in fs]
    ref = ibm.spread_sets(ctx, grids, delta, xs_d, fs_d)
    # poison the outputs first: every point must be written
    out = [torch.full_like(r, float("nan")) for r in ref]
    ibm.spread_sets_noalloc(ctx, grids, delta, xs_d, fs_d, out)
    ctx.synchronize()
    for c in range(3):
        a = out[c].cpu().numpy()
        b = ref[c].cpu().numpy()
        assert np.array_equal(a.view(np.int64), b.view(np.int64)), f"component {c} differs"
        assert np.abs(b).max() > 0


def test_noalloc_matches_allocating_spread_twice(ibm, ctx):
    inputs = load_inputs()
    # call 1: 64^3 cells over 16 um with walls in y, two marker sets (60,000 markers)
    L1 = (16 * UM, 16 * UM, 16 * UM)
    dom1 = [ibm.periodic(L1[0]), ibm.dirichlet(L1[1]), ibm.periodic(L1[2])]
    g1 = ibm.mac_grids(dom1, 64)
    xs, fs = _two_sets(inputs, 40000, 20000, L1, 31)
    _compare(ibm, ctx, g1, xs, fs)
    # call 2, right after: the C1 sphere (2562 markers) on the periodic 32^3 grid
    cfg = inputs.config_c1()
    g2 = ibm.mac_grids([ibm.periodic(16 * UM)] * 3, cfg["refinement"])
    _compare(ibm, ctx, g2, [cfg["x"]], [cfg["f"]])
    # and the larger input once more, now after a smaller one
    _compare(ibm, ctx, g1, xs, fs, delta="peskin")


def test_noalloc_refuses_oversized_inputs(ibm, ctx):
    import torch
    inputs = load_inputs()
    L = (16 * UM,) * 3
    grids = ibm.mac_grids([ibm.periodic(L[0])] * 3, 32)
    n = (1 << 18) + 1
    x = _t(inputs.random_markers(n, L, 5))
    f = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
    out = [torch.zeros(grids[c].size, dtype=torch.float64, device="cuda") for c in range(3)]
    with pytest.raises(ibm.IbmError) as e:
        ibm.spread_sets_noalloc(ctx, grids, "cosine", [x], [f], out)
    assert e.value.code == ibm.ERR_SIZE
    # a grid beyond the workspace's bins: 32 x 32 x 48 um has a 16 um base
    # unit, so refinement 64 gives 128 x 128 x 192 cells (3,145,728 bins);
    # the outputs are full size all the same
    Lb = (32 * UM, 32 * UM, 48 * UM)
    big = ibm.mac_grids([ibm.periodic(a) for a in Lb], 64)
    assert [a.cells for a in big[0].axis] == [128, 128, 192]
    small_x = _t(inputs.random_markers(16, Lb, 6))
    small_f = torch.zeros(48, dtype=torch.float64, device="cuda")
    out = [torch.zeros(big[c].size, dtype=torch.float64, device="cuda") for c in range(3)]
    with pytest.raises(ibm.IbmError) as e:
        ibm.spread_sets_noalloc(ctx, big, "cosine", [small_x], [small_f], out)
    assert e.value.code == ibm.ERR_SIZE
    ctx.synchronize()
