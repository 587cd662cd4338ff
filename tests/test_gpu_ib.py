"""GPU parity of spreading and interpolation against the CPU oracle.

Bar: 1e-12 relative (max-norm) to the oracle's restatement of the reference
algorithm (FP64 everywhere; summation order differs), bitwise run-to-run
reproducibility, and size-independent properties at full sizes.
"""
import numpy as np
import pytest

from conftest import load_inputs

UM = 0.1
RTOL = 1e-12

pytestmark = pytest.mark.gpu


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _spec(dom):
    return [(d.length, d.solid, tuple(d.lower), tuple(d.upper)) for d in dom]


def _relerr(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / scale


def _c1(ibm):
    inputs = load_inputs()
    cfg = inputs.config_c1()
    dom = [ibm.periodic(16 * UM)] * 3
    return cfg, dom


@pytest.mark.parametrize("delta", ["cosine", "peskin", "bspline3"])
def test_spread_c1_matches_oracle(ibm, orc, ctx, delta):
    cfg, dom = _c1(ibm)
    grids = ibm.mac_grids(dom, cfg["refinement"])
    fe = ibm.spread(ctx, grids, delta, _t(cfg["x"]), _t(cfg["f"]))
    ctx.synchronize()
    ref = orc.spread(orc.mac_axes(_spec(dom), cfg["refinement"]), delta, cfg["x"], cfg["f"])
    for c in range(3):
        assert _relerr(fe[c].cpu().numpy(), ref[c]) < RTOL


@pytest.mark.parametrize("delta", ["cosine", "peskin", "bspline3"])
def test_interp_c1_matches_oracle(ibm, orc, ctx, delta):
    cfg, dom = _c1(ibm)
    inputs = load_inputs()
    grids = ibm.mac_grids(dom, cfg["refinement"])
    rng = inputs.SplitMix64(21)
    ue = [rng.normal(grids[c].size) for c in range(3)]
    ul = ibm.interpolate(ctx, grids, delta, _t(cfg["x"]), [_t(u) for u in ue])
    ctx.synchronize()
    ref = orc.interpolate(orc.mac_axes(_spec(dom), cfg["refinement"]), delta, cfg["x"], ue)
    assert _relerr(ul.cpu().numpy(), ref) < RTOL


@pytest.mark.parametrize("walls", ["dirichlet", "neumann"])
def test_spread_interp_walls_match_oracle(ibm, orc, ctx, walls):
    inputs = load_inputs()
    wall = ibm.dirichlet if walls == "dirichlet" else ibm.neumann
    dom = [ibm.periodic(2.0), wall(1.5), ibm.periodic(1.0)]     # unit 0.5 -> 40 x 30 x 20 cells at ref 10
    grids = ibm.mac_grids(dom, 10)
    assert grids[0].points == (40, 30, 20)
    n = 3000
    # markers anywhere in the box, including the first and last half cell
    x = inputs.random_markers(n, (2.0, 1.5, 1.0), seed=31)
    f = inputs.SplitMix64(32).normal(3 * n)
    axes = orc.mac_axes(_spec(dom), 10)
    fe = ibm.spread(ctx, grids, "cosine", _t(x), _t(f))
    ref = orc.spread(axes, "cosine", x, f)
    for c in range(3):
        assert _relerr(fe[c].cpu().numpy(), ref[c]) < RTOL
    ue = [inputs.SplitMix64(40 + c).normal(grids[c].size) for c in range(3)]
    ul = ibm.interpolate(ctx, grids, "cosine", _t(x), [_t(u) for u in ue])
    assert _relerr(ul.cpu().numpy(), orc.interpolate(axes, "cosine", x, ue)) < RTOL


def test_markers_outside_walls_follow_reference_clamping(ibm, orc, ctx):
    # Markers beyond a solid wall: the reference clamps the sort key and keeps
    # the unclamped offset; spread drops, interpolation clamps.
    dom = [ibm.periodic(0.8), ibm.dirichlet(0.8), ibm.periodic(0.8)]
    grids = ibm.mac_grids(dom, 8)
    x = np.array([0.1, 0.4, 0.7, 0.25, -0.03, 0.83, 0.5, 0.2, 0.6])
    f = np.array([1.0, -2.0, 0.5, 0.3, 1.0, 2.0, -1.0, 0.25, 4.0])
    axes = orc.mac_axes(_spec(dom), 8)
    fe = ibm.spread(ctx, grids, "peskin", _t(x), _t(f))
    ref = orc.spread(axes, "peskin", x, f)
    for c in range(3):
        assert _relerr(fe[c].cpu().numpy(), ref[c]) < RTOL
    ue = [np.arange(grids[c].size, dtype=np.float64) for c in range(3)]
    ul = ibm.interpolate(ctx, grids, "peskin", _t(x), [_t(u) for u in ue])
    assert _relerr(ul.cpu().numpy(), orc.interpolate(axes, "peskin", x, ue)) < RTOL


def test_markers_on_cell_faces(ibm, orc, ctx):
    # Markers exactly on grid lines of every staggering (h = 1/8 exactly).
    dom = [ibm.periodic(1.0)] * 3
    grids = ibm.mac_grids(dom, 8)
    vals = np.array([0.0, 1 / 16, 1 / 8, 0.5, 15 / 16, 1.0 - 1 / 8])
    g = np.stack(np.meshgrid(vals, vals, vals, indexing="ij")).reshape(3, -1)
    x = np.ascontiguousarray(g.reshape(-1))
    n = g.shape[1]
    f = np.linspace(-1, 1, 3 * n)
    axes = orc.mac_axes(_spec(dom), 8)
    fe = ibm.spread(ctx, grids, "cosine", _t(x), _t(f))
    ref = orc.spread(axes, "cosine", x, f)
    for c in range(3):
        assert _relerr(fe[c].cpu().numpy(), ref[c]) < RTOL


def test_empty_and_sets(ibm, orc, ctx):
    import torch
    inputs = load_inputs()
    cfg, dom = _c1(ibm)
    grids = ibm.mac_grids(dom, cfg["refinement"])
    # n = 0: the grid is written with zeros
    out = [torch.full((grids[c].size,), 7.0, dtype=torch.float64, device="cuda") for c in range(3)]
    ibm.spread(ctx, grids, "cosine", _t(np.zeros(0)), _t(np.zeros(0)), out=out)
    for c in range(3):
        assert float(out[c].abs().max()) == 0.0
    # two uneven sets == one concatenated set, bitwise
    x, f, n = cfg["x"], cfg["f"], cfg["n"]
    k = 1000
    xa = np.concatenate([x[j * n:j * n + k] for j in range(3)])
    xb = np.concatenate([x[j * n + k:(j + 1) * n] for j in range(3)])
    fa = np.concatenate([f[j * n:j * n + k] for j in range(3)])
    fb = np.concatenate([f[j * n + k:(j + 1) * n] for j in range(3)])
    one = ibm.spread(ctx, grids, "cosine", _t(x), _t(f))
    two = ibm.spread_sets(ctx, grids, "cosine", [_t(xa), _t(xb)], [_t(fa), _t(fb)])
    for c in range(3):
        assert torch.equal(one[c], two[c])
    ue = [_t(inputs.SplitMix64(5 + c).normal(grids[c].size)) for c in range(3)]
    ul = ibm.interpolate(ctx, grids, "cosine", _t(x), ue).cpu().numpy()
    ua, ub = ibm.interpolate_sets(ctx, grids, "cosine", [_t(xa), _t(xb)], ue)
    ua, ub = ua.cpu().numpy(), ub.cpu().numpy()
    for j in range(3):
        assert np.array_equal(ul[j * n:j * n + k], ua[j * k:(j + 1) * k])
        assert np.array_equal(ul[j * n + k:(j + 1) * n], ub[j * (n - k):(j + 1) * (n - k)])


def test_c4_full_size_properties(ibm, ctx):
    """C4 at full size: bitwise reproducibility, conservation and adjointness
    (markers stay clear of the walls, so both hold to rounding)."""
    import torch
    inputs = load_inputs()
    cfg = inputs.config_c4()
    L = cfg["L"]
    dom = [ibm.periodic(L[0]), ibm.dirichlet(L[1]), ibm.periodic(L[2])]
    grids = ibm.mac_grids(dom, cfg["refinement"])
    assert grids[1].points == (128, 255, 128)
    xs = [_t(cfg["x_rbc"]), _t(cfg["x_plt"])]
    fs = [_t(cfg["f_rbc"]), _t(cfg["f_plt"])]
    a = ibm.spread_sets(ctx, grids, "cosine", xs, fs)
    b = ibm.spread_sets(ctx, grids, "cosine", xs, fs)
    for c in range(3):
        assert torch.equal(a[c], b[c])
    h3 = (1.0 / grids[0].resolution) ** 3
    nr, npl = cfg["n_rbc"], cfg["n_plt"]
    for c in range(3):
        total = cfg["f_rbc"][c * nr:(c + 1) * nr].sum() + cfg["f_plt"][c * npl:(c + 1) * npl].sum()
        assert abs(float(a[c].sum()) * h3 - total) < 1e-9 * np.abs(cfg["f_rbc"]).sum()
    ue = [torch.randn(grids[c].size, dtype=torch.float64, device="cuda",
                      generator=torch.Generator(device="cuda").manual_seed(c)) for c in range(3)]
    ur, up = ibm.interpolate_sets(ctx, grids, "cosine", xs, ue)
    lhs = sum(float(a[c] @ ue[c]) for c in range(3)) * h3
    rhs = float(fs[0] @ ur) + float(fs[1] @ up)
    assert abs(lhs - rhs) < 1e-10 * abs(rhs)


def test_bad_arguments_raise(ibm, ctx):
    cfg, dom = _c1(ibm)
    grids = ibm.mac_grids(dom, cfg["refinement"])
    x = _t(cfg["x"])
    with pytest.raises(ibm.IbmError) as e:
        ibm.spread(ctx, grids, 7, x, x)
    assert e.value.code == ibm.ERR_ARG
    tiny = ibm.mac_grids([ibm.periodic(1.0)] * 3, 2)     # 2 cells per periodic axis
    with pytest.raises(ibm.IbmError) as e:
        ibm.spread(ctx, tiny, "cosine", x, x)
    assert e.value.code == ibm.ERR_BAD_GRID
