"""GPU parity of the staggered-grid operators and solvers against the oracle
(oracle/fd.py, CSR restatement of the reference's assembled operators).

Bars: operator application 1e-13 relative (FP64) / 1e-6 (FP32); Gershgorin
bounds 1e-13; Chebyshev smoothing 1e-12; one FP64 V-cycle 1e-11 against the
oracle's V-cycle with the same (x-fastest) transfers; solves to rtol 1e-10
against a sparse direct solve (error bound 1e-8 relative, both precisions).
"""
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from conftest import ROOT, _load

pytestmark = pytest.mark.gpu

fd = _load("orc_fd", os.path.join(ROOT, "oracle", "fd.py"))


def _t(a, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return t if dtype is None else t.to(dtype)


def _spec(dom):
    return [(d.length, d.solid, tuple(d.lower), tuple(d.upper)) for d in dom]


def _axes(orc, dom, ref, align):
    a = orc.grid_axes(_spec(dom), ref, align)
    return [a[j] for j in range(3)]


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


DOMAINS = {
    "periodic": lambda ibm: [ibm.periodic(1.0), ibm.periodic(0.75), ibm.periodic(1.25)],    # 8k x 6k x 10k
    "walls": lambda ibm: [ibm.periodic(1.0), ibm.dirichlet(0.75), ibm.periodic(1.25)],
    "neumann": lambda ibm: [ibm.periodic(1.0), ibm.neumann(0.75), ibm.periodic(1.25)],
}
ALIGNS = {"u": [0.0, 0.5, 0.5], "v": [0.5, 0.0, 0.5], "w": [0.5, 0.5, 0.0], "p": [0.5, 0.5, 0.5]}


@pytest.mark.parametrize("dom_kind", ["periodic", "walls", "neumann"])
@pytest.mark.parametrize("comp", ["u", "v", "p"])
def test_operators_match_assembled_csr(ibm, orc, ctx, dom_kind, comp):
    import torch
    dom = DOMAINS[dom_kind](ibm)
    ref = 8                          # unit 0.25 -> 32 x 24 x 40 cells
    g = ibm.grid_make(dom, ref, ALIGNS[comp])
    ax = _axes(orc, dom, ref, ALIGNS[comp])
    rng = np.random.default_rng(3)
    x = rng.standard_normal(g.size)
    y0 = rng.standard_normal(g.size)
    lam = 0.4 / g.resolution ** 2
    for op in (ibm.LAPLACIAN, ibm.IDENTITY, ibm.helmholtz(lam)):
        A = fd.operator(ax, *op)
        ref_y = 0.5 * (A @ x) - 2.0 * y0
        y = ibm.op_apply(ctx, g, op, _t(x), alpha=0.5, beta=-2.0, y=_t(y0.copy()))
        assert _rel(y.cpu().numpy(), ref_y) < 1e-13
        y32 = ibm.op_apply(ctx, g, op, _t(x, torch.float32), alpha=0.5, beta=-2.0,
                           y=_t(y0.copy(), torch.float32))
        assert _rel(y32.double().cpu().numpy(), ref_y) < 2e-6
        lo, hi = ibm.gershgorin(g, op)
        rlo, rhi = fd.gershgorin(A)
        scale = max(abs(rlo), abs(rhi))
        assert abs(lo - rlo) < 1e-13 * scale and abs(hi - rhi) < 1e-13 * scale


def test_chebyshev_apply_matches_oracle(ibm, orc, ctx):
    dom = DOMAINS["walls"](ibm)
    g = ibm.grid_make(dom, 8, ALIGNS["v"])
    ax = _axes(orc, dom, 8, ALIGNS["v"])
    lam = 0.4 / g.resolution ** 2
    op = ibm.helmholtz(lam)
    A = fd.operator(ax, *op)
    a, b = fd.smoother_interval(ax, A)
    assert np.allclose(ibm.chebyshev_weights(a, b), fd.chebw4(a, b), rtol=1e-15, atol=0)
    r = np.random.default_rng(4).standard_normal(g.size)
    out = ibm.chebyshev_apply(ctx, g, op, a, b, _t(r))
    assert _rel(out.cpu().numpy(), fd.chebyshev_apply(A, a, b, r)) < 1e-12


@pytest.mark.parametrize("dom_kind,ref,comp", [
    ("flat", 4, "w"),          # periodic z of 4 planes: shorter than the march's 12-plane lead
    ("flat_walls", 4, "v"),
    ("walls", 4, "w"),
    ("all_walls", 4, "p"),
    ("periodic", 8, "u"),
    ("all_walls", 8, "v"),
    ("walls", 24, "u"),        # 829K points: the wide tile
])
def test_chebyshev_apply_shapes(ibm, orc, ctx, dom_kind, ref, comp):
    """The one-pass smoother on every wall pattern, on z axes shorter than
    its pipeline and on grids large enough for the wide tile."""
    doms = dict(DOMAINS,
                all_walls=lambda ibm: [ibm.dirichlet(1.0), ibm.dirichlet(0.75), ibm.neumann(1.25)],
                flat=lambda ibm: [ibm.periodic(1.0), ibm.periodic(0.75), ibm.periodic(0.25)],
                flat_walls=lambda ibm: [ibm.periodic(1.0), ibm.dirichlet(0.75), ibm.periodic(0.25)])
    dom = doms[dom_kind](ibm)
    g = ibm.grid_make(dom, ref, ALIGNS[comp])
    ax = _axes(orc, dom, ref, ALIGNS[comp])
    op = ibm.helmholtz(0.4 / g.resolution ** 2)
    A = fd.operator(ax, *op)
    a, b = fd.smoother_interval(ax, A)
    r = np.random.default_rng(ref).standard_normal(g.size)
    out = ibm.chebyshev_apply(ctx, g, op, a, b, _t(r))
    assert _rel(out.cpu().numpy(), fd.chebyshev_apply(A, a, b, r)) < 1e-12


def _mg_case(ibm, orc, comp, dom_kind, ref=8):
    dom = DOMAINS[dom_kind](ibm)
    g = ibm.grid_make(dom, ref, ALIGNS[comp])
    af = lambda r: _axes(orc, dom, r, ALIGNS[comp])
    return dom, g, af


@pytest.mark.parametrize("comp", ["u", "v", "w"])
def test_vcycle_fp64_matches_oracle(ibm, orc, ctx, comp):
    dom, g, af = _mg_case(ibm, orc, comp, "walls")
    lam = 0.4 / g.resolution ** 2
    op = ibm.helmholtz(lam)
    mg = ibm.Multigrid(ctx, g, op, precision=64)
    om = fd.Multigrid(af, 8, *op, order="x-fastest")
    assert len(mg.levels()) == len(om.levels)
    b = np.random.default_rng(5).standard_normal(g.size)
    x = mg.vcycle(_t(b)).cpu().numpy()
    assert _rel(x, om.vcycle(b)) < 1e-11
    mg.close()


@pytest.mark.parametrize("zlen,ref,comp", [(0.0625, 4, "v"), (0.125, 4, "w"), (0.0625, 8, "w")])
def test_vcycle_fp64_thin_z_matches_oracle(ibm, orc, ctx, zlen, ref, comp):
    """V-cycles whose fused transfer kernels see coarse z axes of two planes
    (64 x 47 x 4 over 32 x 23 x 2, ...): periodic wraps by more than one
    period, and a coarse-plane ring longer than the coarse axis."""
    dom = [ibm.periodic(1.0), ibm.dirichlet(0.75), ibm.periodic(zlen)]
    g = ibm.grid_make(dom, ref, ALIGNS[comp])
    op = ibm.helmholtz(0.4 / g.resolution ** 2)
    mg = ibm.Multigrid(ctx, g, op, precision=64)
    om = fd.Multigrid(lambda r: _axes(orc, dom, r, ALIGNS[comp]), ref, *op, order="x-fastest")
    assert len(mg.levels()) == len(om.levels) >= 2
    b = np.random.default_rng(9).standard_normal(g.size)
    x = mg.vcycle(_t(b)).cpu().numpy()
    assert _rel(x, om.vcycle(b)) < 1e-11
    mg.close()


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("dom_kind", ["periodic", "walls"])
def test_helmholtz_mg_solve(ibm, orc, ctx, precision, dom_kind):
    dom, g, af = _mg_case(ibm, orc, "v", dom_kind)
    lam = 0.4 / g.resolution ** 2
    op = ibm.helmholtz(lam)
    A = fd.operator(af(8), *op)
    b = np.random.default_rng(6).standard_normal(g.size)
    exact = spla.spsolve(A.tocsc(), b)
    mg = ibm.Multigrid(ctx, g, op, precision=precision)
    x, it, rr = mg.solve(_t(b), rtol=1e-10, maxit=60)
    assert rr <= 1e-10
    assert np.linalg.norm(x.cpu().numpy() - exact) <= 1e-8 * np.linalg.norm(exact)
    mg.close()


@pytest.mark.parametrize("precision", [64, 32])
def test_pressure_poisson_mg_solve(ibm, orc, ctx, precision):
    # Singular Neumann/periodic Laplacian on the cell centres: mean-free solution.
    dom, g, af = _mg_case(ibm, orc, "p", "neumann")
    L = fd.laplacian(af(8))
    rng = np.random.default_rng(7)
    b = rng.standard_normal(g.size)
    b -= b.mean()
    mg = ibm.Multigrid(ctx, g, ibm.LAPLACIAN, precision=precision)
    x, it, rr = mg.solve(_t(b), rtol=1e-10, maxit=80)
    x = x.cpu().numpy()
    assert abs(x.mean()) < 1e-12 * np.abs(x).max()
    assert np.linalg.norm(L @ x - b) <= 2e-10 * np.linalg.norm(b)
    mg.close()


def test_chebpcg_matches_direct(ibm, orc, ctx):
    dom, g, af = _mg_case(ibm, orc, "u", "walls")
    lam = 0.4 / g.resolution ** 2
    op = ibm.helmholtz(lam)
    A = fd.operator(af(8), *op)
    b = np.random.default_rng(8).standard_normal(g.size)
    exact = spla.spsolve(A.tocsc(), b)
    x, it, rr = ibm.chebpcg_solve(ctx, g, op, _t(b), rtol=1e-12)
    assert np.linalg.norm(x.cpu().numpy() - exact) <= 1e-10 * np.linalg.norm(exact)


def test_c3_helmholtz_mixed_precision_full_size(ibm, ctx):
    """128^3 Helmholtz (C3 grid, y walls) in FP32 multigrid with FP64 defect
    correction: converges to rtol 1e-10, checked by an independent FP64
    residual."""
    import torch
    um = 0.1
    dom = [ibm.periodic(32 * um), ibm.dirichlet(32 * um), ibm.periodic(32 * um)]
    g = ibm.grid_make(dom, 128, ALIGNS["u"])
    assert g.points == (128, 128, 128)
    nu, dt = 10.0, 1e-4                    # 1e-6 m^2/s, 1e-7 s in library units
    op = ibm.helmholtz(nu * dt)
    b = torch.randn(g.size, dtype=torch.float64, device="cuda", generator=torch.Generator(device="cuda").manual_seed(1))
    mg = ibm.Multigrid(ctx, g, op, precision=32)
    x, it, rr = mg.solve(b, rtol=1e-10, maxit=50)
    r = ibm.op_apply(ctx, g, op, x, alpha=-1.0, beta=1.0, y=b.clone())
    assert float(r.norm()) <= 1.01e-10 * float(b.norm())
    assert it <= 20
    mg.close()


def test_fixed_and_floor_solves(ibm, orc, ctx):
    """ibm_mg_solve_fixed reproduces the adaptive solve with the same cycle
    count bitwise; ibm_mg_solve_atol accepts a right-hand side that is pure
    rounding noise (the adaptive relative test alone cannot)."""
    import torch
    dom, g, af = _mg_case(ibm, orc, "p", "neumann")
    mg = ibm.Multigrid(ctx, g, ibm.LAPLACIAN, precision=32)
    rng = np.random.default_rng(9)
    b = rng.standard_normal(g.size)
    b -= b.mean()
    x, it, rr = mg.solve(_t(b), rtol=1e-10, maxit=80)
    xf, rrf = mg.solve_fixed(_t(b), it)
    assert torch.equal(x, xf)
    assert abs(float(rrf) - rr) <= 1e-12 * rr + 1e-300
    noise = _t(1e-17 * rng.standard_normal(g.size))
    with pytest.raises(ibm.IbmError):
        mg.solve(noise, rtol=1e-12, maxit=5)
    x, it, rr = mg.solve(noise, rtol=1e-12, maxit=5, atol=1e-12)
    assert it <= 5
    mg.close()
