"""Many independent boxes in one launch per operator (ibm_spread_boxes,
ibm_interpolate_boxes) and the tile spread they use.

Bars: each box's spread and interpolation within 1e-12 relative (max norm)
of the oracle run on that box alone (orc.spread / orc.interpolate, the
restatement of ib/novel.h and ib/interpolate.h); the tile spread of the C4
channel within 1e-12 of the brick spread (ibm_spread_sets); at full C5 size
(64 boxes, 335,388 markers) per-box conservation sum(fe) h^3 = sum(f) and
adjointness <S f, u> h^3 = <f, J u> to 1e-12 relative, and bitwise equal
repeated runs.
"""
import numpy as np
import pytest

from conftest import load_inputs

UM = 0.1
pytestmark = pytest.mark.gpu


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _spec(dom):
    return [(d.length, d.solid, tuple(d.lower), tuple(d.upper)) for d in dom]


@pytest.fixture(scope="module")
def c5():
    return load_inputs().config_c5()


def _box(cfg, b):
    o0, o1 = cfg["offsets"][b], cfg["offsets"][b + 1]
    n = cfg["n"]
    x = cfg["x"].reshape(3, n)[:, o0:o1]
    f = cfg["f"].reshape(3, n)[:, o0:o1]
    return np.ascontiguousarray(x.reshape(-1)), np.ascontiguousarray(f.reshape(-1))


@pytest.mark.parametrize("delta", ["cosine", "peskin"])
def test_boxes_match_oracle_per_box(ibm, orc, ctx, c5, delta):
    dom = [ibm.periodic(c5["L"])] * 3
    grids = ibm.mac_grids(dom, c5["refinement"])
    axes = orc.mac_axes(_spec(dom), c5["refinement"])
    nb = 6
    off = c5["offsets"][:nb + 1]
    n = off[-1]
    x = np.ascontiguousarray(c5["x"].reshape(3, -1)[:, :n].reshape(-1))
    f = np.ascontiguousarray(c5["f"].reshape(3, -1)[:, :n].reshape(-1))
    fe = ibm.spread_boxes(ctx, grids, delta, off, _t(x), _t(f))
    rng = load_inputs().SplitMix64(55)
    ue = [rng.normal(nb * grids[c].size) for c in range(3)]
    ul = ibm.interpolate_boxes(ctx, grids, delta, off, _t(x), [_t(u) for u in ue])
    ctx.synchronize()
    fe = [a.cpu().numpy() for a in fe]
    ul = ul.cpu().numpy().reshape(3, n)
    for b in range(nb):
        xb, fb = _box({"offsets": off, "n": n, "x": x, "f": f}, b)
        ref = orc.spread(axes, delta, xb, fb)
        for c in range(3):
            sz = grids[c].size
            assert _rel(fe[c][b * sz:(b + 1) * sz], ref[c]) < 1e-12, (b, c)
        ub = [ue[c][b * grids[c].size:(b + 1) * grids[c].size] for c in range(3)]
        refl = orc.interpolate(axes, delta, xb, ub).reshape(3, -1)
        assert _rel(ul[:, off[b]:off[b + 1]], refl) < 1e-12, b


@pytest.mark.parametrize("cells", [(8, 16, 8), (16, 8, 24)])
def test_tile_spread_on_one_and_two_brick_axes(ibm, orc, ctx, cells):
    """Periodic x and z of one, two or three bricks (8, 16, 24 cells; the
    tiles wrap onto themselves or their neighbour) with walls in y, markers
    spread over the whole box including the wall layers."""
    h = 0.5 * UM
    dom = [ibm.periodic(cells[0] * h), ibm.dirichlet(cells[1] * h), ibm.periodic(cells[2] * h)]
    grids = ibm.mac_grids(dom, 8)                             # 8 cells per 4 um (the lengths' GCD)
    assert [grids[0].axis[j].cells for j in range(3)] == list(cells)
    axes = orc.mac_axes(_spec(dom), 8)
    rng = load_inputs().SplitMix64(sum(cells))
    n = 3000
    x = np.concatenate([rng.uniform(n) * cells[j] * h for j in range(3)])
    f = rng.normal(3 * n)
    for delta in ("cosine", "peskin"):
        fe = ibm.spread_boxes(ctx, grids, delta, [0, n], _t(x), _t(f))
        ctx.synchronize()
        ref = orc.spread(axes, delta, x, f)
        for c in range(3):
            assert _rel(fe[c].cpu().numpy(), ref[c]) < 1e-12, (delta, c)


def test_tile_spread_matches_brick_spread_on_the_c4_channel(ibm, ctx):
    """One box: the 128 x 256 x 128 channel with y walls and all C4 cells."""
    inputs = load_inputs()
    cfg = inputs.config_c4()
    L = cfg["L"]
    dom = [ibm.periodic(L[0]), ibm.dirichlet(L[1]), ibm.periodic(L[2])]
    grids = ibm.mac_grids(dom, cfg["refinement"])
    xr, xp = _t(cfg["x_rbc"]), _t(cfg["x_plt"])
    fr, fp = _t(cfg["f_rbc"]), _t(cfg["f_plt"])
    ref = ibm.spread_sets(ctx, grids, "cosine", [xr, xp], [fr, fp])
    x = _t(np.concatenate([cfg["x_rbc"].reshape(3, -1), cfg["x_plt"].reshape(3, -1)], axis=1).reshape(-1))
    f = _t(np.concatenate([cfg["f_rbc"].reshape(3, -1), cfg["f_plt"].reshape(3, -1)], axis=1).reshape(-1))
    n = cfg["n_rbc"] + cfg["n_plt"]
    tile = ibm.spread_boxes(ctx, grids, "cosine", [0, n], x, f)
    ctx.synchronize()
    for c in range(3):
        err = _rel(tile[c].cpu().numpy(), ref[c].cpu().numpy())
        print("tile vs brick spread, component", c, err)
        assert err < 1e-12
    import torch
    ms = {}
    for name, call in (("brick", lambda: ibm.spread_sets(ctx, grids, "cosine", [xr, xp], [fr, fp])),
                       ("tile", lambda: ibm.spread_boxes(ctx, grids, "cosine", [0, n], x, f))):
        call()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            call()
        e1.record()
        torch.cuda.synchronize()
        ms[name] = e0.elapsed_time(e1) / 5
    print(f"C4 spread of {n} markers: brick {ms['brick']:.3f} ms, tile {ms['tile']:.3f} ms (with allocation)")


def test_c5_full_size_properties(ibm, ctx, c5):
    import torch
    dom = [ibm.periodic(c5["L"])] * 3
    grids = ibm.mac_grids(dom, c5["refinement"])
    off = c5["offsets"]
    nb = c5["nboxes"]
    x, f = _t(c5["x"]), _t(c5["f"])
    fe = ibm.spread_boxes(ctx, grids, "cosine", off, x, f)
    fe2 = ibm.spread_boxes(ctx, grids, "cosine", off, x, f)
    rng = load_inputs().SplitMix64(77)
    ue = [_t(rng.normal(nb * grids[c].size)) for c in range(3)]
    ul = ibm.interpolate_boxes(ctx, grids, "cosine", off, x, ue)
    ctx.synchronize()
    h3 = (c5["L"] / c5["refinement"]) ** 3
    n = c5["n"]
    fh = c5["f"].reshape(3, n)
    ulh = ul.cpu().numpy().reshape(3, n)
    for c in range(3):
        a, b = fe[c].cpu().numpy(), fe2[c].cpu().numpy()
        assert np.array_equal(a.view(np.int64), b.view(np.int64))
        sz = grids[c].size
        u = ue[c].cpu().numpy()
        for bx in range(nb):
            s = a[bx * sz:(bx + 1) * sz]
            fb = fh[c, off[bx]:off[bx + 1]]
            assert abs(s.sum() * h3 - fb.sum()) <= 1e-12 * np.abs(fb).sum()
            lhs = (s * u[bx * sz:(bx + 1) * sz]).sum() * h3
            rhs = (fb * ulh[c, off[bx]:off[bx + 1]]).sum()
            assert abs(lhs - rhs) <= 1e-12 * np.abs(fb).sum() * np.abs(u).max()
    # timing of the batched operators (reported, not asserted)
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    e0.record()
    for _ in range(5):
        ibm.spread_boxes(ctx, grids, "cosine", off, x, f, out=fe)
    e1.record()
    for _ in range(5):
        ibm.interpolate_boxes(ctx, grids, "cosine", off, x, ue, out=ul)
    e2.record()
    torch.cuda.synchronize()
    print(f"C5: {n} markers in {nb} boxes; spread {e0.elapsed_time(e1) / 5:.3f} ms, "
          f"interpolate {e1.elapsed_time(e2) / 5:.3f} ms")


def test_boxes_argument_errors(ibm, ctx, c5):
    import torch
    dom = [ibm.periodic(c5["L"])] * 3
    grids = ibm.mac_grids(dom, c5["refinement"])
    x = _t(c5["x"][: 3 * 10])
    f = torch.zeros(30, dtype=torch.float64, device="cuda")
    out = [torch.zeros(grids[c].size, dtype=torch.float64, device="cuda") for c in range(3)]
    with pytest.raises(ibm.IbmError) as e:
        ibm.spread_boxes(ctx, grids, "cosine", [0, 6, 5, 10], x, f)     # decreasing offsets
    assert e.value.code == ibm.ERR_SIZE
    # 20 cells per periodic axis is not a multiple of the brick edge
    g20 = ibm.mac_grids([ibm.periodic(2.0)] * 3, 20)
    assert g20[0].axis[0].cells == 20
    with pytest.raises(ibm.IbmError) as e:
        ibm.spread_boxes(ctx, g20, "cosine", [0, 10], _t(np.full(30, 0.5)), f,
                         out=[torch.zeros(g20[c].size, dtype=torch.float64, device="cuda") for c in range(3)])
    assert e.value.code == ibm.ERR_BAD_GRID
    del out


def test_boxes_helmholtz_matches_direct_solves(ibm, orc, ctx, c5):
    """(I - nu dt L) x = b on every MAC component of 6 boxes at once, each
    against a sparse direct solve of the oracle's assembled operator
    (oracle/fd.py), backward Euler with nu = 1e-6 m^2/s, dt = 0.1 us."""
    import os
    import scipy.sparse.linalg as spla
    from conftest import ROOT, _load
    fd = _load("orc_fd", os.path.join(ROOT, "oracle", "fd.py"))
    dom = [ibm.periodic(c5["L"])] * 3
    grids = ibm.mac_grids(dom, c5["refinement"])
    lam = 10.0 * 1e-4
    nb = 6
    rng = load_inputs().SplitMix64(91)
    MAC = [[0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]]
    for c in range(3):
        n = grids[c].size
        b = rng.normal(nb * n)
        mg = ibm.BoxesMultigrid(ctx, grids[c], nb, (1.0, -lam))
        x, it, rel = mg.solve(_t(b), rtol=1e-12)
        ctx.synchronize()
        print("component", c, "cycles", it, "worst relative residual", rel)
        assert rel <= 1e-12 and it <= 20
        ax = [orc.grid_axes(_spec(dom), c5["refinement"], MAC[c])[j] for j in range(3)]
        A = fd.operator(ax, 1.0, -lam).tocsc()
        lu = spla.splu(A)
        xh = x.cpu().numpy()
        for bx in range(nb):
            ref = lu.solve(b[bx * n:(bx + 1) * n])
            assert _rel(xh[bx * n:(bx + 1) * n], ref) < 1e-10, (c, bx)
        mg.close()


def test_c5_helmholtz_full_size(ibm, ctx, c5):
    """All 64 boxes: the spread forces as right-hand sides, every box's FP64
    residual to 1e-10, in one launch per operation for all boxes."""
    import torch
    dom = [ibm.periodic(c5["L"])] * 3
    grids = ibm.mac_grids(dom, c5["refinement"])
    nb = c5["nboxes"]
    fe = ibm.spread_boxes(ctx, grids, "cosine", c5["offsets"], _t(c5["x"]), _t(c5["f"]))
    lam = 10.0 * 1e-4
    dt_rho = 1e-4 / 1e-3
    for c in range(3):
        mg = ibm.BoxesMultigrid(ctx, grids[c], nb, (1.0, -lam))
        b = fe[c] * dt_rho
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        x, it, rel = mg.solve(b, rtol=1e-10)
        e1.record()
        torch.cuda.synchronize()
        print(f"C5 Helmholtz component {c}: {nb} boxes, {it} cycles, worst relative residual {rel:.2e}, "
              f"{e0.elapsed_time(e1):.3f} ms")
        assert rel <= 1e-10
        # an independent FP64 residual through the single-grid operator on two boxes
        n = grids[c].size
        for bx in (0, nb - 1):
            xb = x[bx * n:(bx + 1) * n].contiguous()
            r = ibm.op_apply(ctx, grids[c], (1.0, -lam), xb)
            ctx.synchronize()
            bb = b[bx * n:(bx + 1) * n]
            assert float((bb - r).norm() / bb.norm()) <= 1e-10
        mg.close()


def test_boxes_helmholtz_errors(ibm, ctx):
    walls = ibm.mac_grids([ibm.periodic(1.6), ibm.dirichlet(1.6), ibm.periodic(1.6)], 16)
    with pytest.raises(ibm.IbmError) as e:
        ibm.BoxesMultigrid(ctx, walls[0], 4, (1.0, -1e-3))
    assert e.value.code == ibm.ERR_BAD_GRID
    per = ibm.mac_grids([ibm.periodic(1.6)] * 3, 16)
    with pytest.raises(ibm.IbmError) as e:
        ibm.BoxesMultigrid(ctx, per[0], 4, (0.0, 1.0))
    assert e.value.code == ibm.ERR_ARG
