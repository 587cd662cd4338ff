"""GPU parity of the Navier-Stokes building blocks (fd::differential,
fd::average, ins::divergence, ins::gradient, ins::advection) against the
oracle's assembled CSR operators (oracle/fd.py) on periodic, Dirichlet-wall
and Neumann-wall domains.  Bar: 1e-13 relative (FP64, two-point stencils and
one product; only the summation order differs)."""
import os

import numpy as np
import pytest

from conftest import ROOT, _load

pytestmark = pytest.mark.gpu

fd = _load("orc_fd", os.path.join(ROOT, "oracle", "fd.py"))

MAC = [[0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]]
CENTRE = [0.5, 0.5, 0.5]
DOMAINS = {
    "periodic": lambda ibm: [ibm.periodic(1.0), ibm.periodic(0.75), ibm.periodic(1.25)],
    "walls": lambda ibm: [ibm.periodic(1.0), ibm.dirichlet(0.75), ibm.periodic(1.25)],
    "neumann": lambda ibm: [ibm.neumann(1.0), ibm.dirichlet(0.75), ibm.periodic(1.25)],
}


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _axes_of(orc, dom, ref):
    spec = [(d.length, d.solid, tuple(d.lower), tuple(d.upper)) for d in dom]

    def f(align):
        a = orc.grid_axes(spec, ref, align)
        return [a[j] for j in range(3)]
    return f


@pytest.mark.parametrize("dom_kind", ["periodic", "walls", "neumann"])
def test_differential_and_average_match_csr(ibm, orc, ctx, dom_kind):
    dom = DOMAINS[dom_kind](ibm)
    af = _axes_of(orc, dom, 8)
    rng = np.random.default_rng(11)
    for align in MAC + [CENTRE]:
        g = ibm.grid_make(dom, 8, align)
        ax = af(align)
        x = rng.standard_normal(g.size)
        for view in range(3):
            dst = ibm.grid_shift(g, view)
            y0 = rng.standard_normal(dst.size)
            for fn, ofn in ((ibm.differential, fd.differential), (ibm.average, fd.average)):
                M = ofn(ax, view)
                assert M.shape == (dst.size, g.size)
                y = fn(ctx, g, view, _t(x), alpha=0.75, beta=-1.5, y=_t(y0.copy()))
                assert _rel(y.cpu().numpy(), 0.75 * (M @ x) - 1.5 * y0) < 1e-13


@pytest.mark.parametrize("dom_kind", ["periodic", "walls", "neumann"])
def test_divergence_gradient_advection_match_oracle(ibm, orc, ctx, dom_kind):
    dom = DOMAINS[dom_kind](ibm)
    af = _axes_of(orc, dom, 8)
    grids = ibm.mac_grids(dom, 8)
    rng = np.random.default_rng(12)
    u = [rng.standard_normal(grids[i].size) for i in range(3)]
    ut = [_t(c) for c in u]

    div = ibm.divergence(ctx, grids, ut).cpu().numpy()
    assert _rel(div, fd.divergence(af, MAC, u)) < 1e-13

    pg = ibm.centered_grid(dom, 8)
    p = rng.standard_normal(pg.size)
    base = [rng.standard_normal(grids[i].size) for i in range(3)]
    out = ibm.gradient(ctx, pg, _t(p), out=[_t(b.copy()) for b in base], alpha=-0.5, beta=1.0)
    for i, gi in enumerate(fd.gradient(af, CENTRE, p)):
        assert _rel(out[i].cpu().numpy(), base[i] - 0.5 * gi) < 1e-13

    H = ibm.advection(ctx, grids, ut)
    for i, hi in enumerate(fd.advection(af, MAC, u)):
        assert _rel(H[i].cpu().numpy(), hi) < 1e-13


def test_advection_full_size_vanishes_for_shear_flow(ibm, ctx):
    """C4 grid (128 x 256 x 128, y walls): u = (sin(pi y / Ly), 0, 0) is an
    exact steady shear, so the advective term is zero to rounding."""
    import torch
    um = 0.1
    dom = [ibm.periodic(32 * um), ibm.dirichlet(64 * um), ibm.periodic(32 * um)]
    grids = ibm.mac_grids(dom, 128)
    gu = grids[0]
    nx, ny, nz = gu.points
    y = (torch.arange(ny, dtype=torch.float64, device="cuda") + gu.axis[1].shift) / gu.axis[1].resolution
    u = torch.sin(np.pi * y / (64 * um)).view(1, ny, 1).expand(nz, ny, nx).contiguous().view(-1)
    zeros = [torch.zeros(grids[i].size, dtype=torch.float64, device="cuda") for i in (1, 2)]
    H = ibm.advection(ctx, grids, [u] + zeros)
    assert max(float(h.abs().max()) for h in H) < 1e-12
    div = ibm.divergence(ctx, grids, [u] + zeros)
    assert float(div.abs().max()) < 1e-12
