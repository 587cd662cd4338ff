"""Direct FFT x tridiagonal solves on channel grids (ibm_fft_*).

Pins: on small channel grids the solution equals the sparse LU solution of
the oracle's CSR operator (oracle/fd.py, the kron-assembled restatement of
fd/laplacian.h and fd/identity.h) to 1e-12 relative, for the Helmholtz
operator on each MAC component (Dirichlet walls, moving-wall ghost rows)
and for the pressure Poisson (Neumann walls, singular: mean-free
solutions).  At the C4 size (128 x 256 x 128) the FP64 residual
||b - A x|| / ||b|| computed by the independent stencil (ibm_op_apply) is at
rounding level (1e-13) and the solutions agree with FP64 multigrid solves
to the multigrid's tolerance times the conditioning.
"""
import os

import numpy as np
import pytest

from conftest import ROOT, _load, load_inputs

pytestmark = pytest.mark.gpu

UM = 0.1
MAC = [[0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]]
CENTRE = [0.5, 0.5, 0.5]


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _spec(dom):
    return [(d.length, d.solid, tuple(d.lower), tuple(d.upper)) for d in dom]


def _channel(ibm, cells, h=0.5 * UM, walls="dirichlet"):
    wall = ibm.dirichlet if walls == "dirichlet" else ibm.neumann
    return [ibm.periodic(cells[0] * h), wall(cells[1] * h), ibm.periodic(cells[2] * h)]


@pytest.mark.parametrize("cells", [(8, 12, 16), (16, 8, 4)])
def test_fft_matches_sparse_lu(ibm, orc, ctx, cells):
    import scipy.sparse.linalg as spla
    fd = _load("orc_fd", os.path.join(ROOT, "oracle", "fd.py"))
    rng = load_inputs().SplitMix64(sum(cells))
    dom = _channel(ibm, cells)
    ref = int(np.gcd.reduce(cells))                         # cells per GCD of the lengths
    grids = ibm.mac_grids(dom, ref)
    assert [grids[0].axis[j].cells for j in range(3)] == list(cells)
    lam = 0.37
    for c in range(3):
        ax = [orc.grid_axes(_spec(dom), ref, MAC[c])[j] for j in range(3)]
        A = fd.operator(ax, 1.0, -lam).tocsc()
        b = rng.normal(A.shape[0])
        want = spla.splu(A).solve(b)
        solver = ibm.FftSolver(ctx, grids[c])
        got = solver.solve(_t(b), (1.0, -lam)).cpu().numpy()
        err = np.abs(got - want).max() / np.abs(want).max()
        print(f"{cells} Helmholtz component {c}: FFT vs sparse LU {err:.2e}")
        assert err < 1e-12
        solver.close()
    pdom = _channel(ibm, cells, walls="neumann")
    pg = ibm.centered_grid(pdom, ref)
    ax = [orc.grid_axes(_spec(pdom), ref, CENTRE)[j] for j in range(3)]
    L = fd.operator(ax, 0.0, 1.0).tolil()
    b = rng.normal(L.shape[0])
    b -= b.mean()
    L[0, :] = 0
    L[0, 0] = 1.0
    d = b.copy()
    d[0] = 0.0
    want = spla.splu(L.tocsc()).solve(d)
    want -= want.mean()
    solver = ibm.FftSolver(ctx, pg)
    got = solver.solve(_t(b), (0.0, 1.0)).cpu().numpy()
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"{cells} pressure Poisson: FFT vs sparse LU {err:.2e}, mean {got.mean():.1e}")
    assert err < 1e-12 and abs(got.mean()) < 1e-13 * np.abs(got).max()
    solver.close()


def test_fft_c4_channel_residuals_and_multigrid(ibm, ctx):
    import torch
    L = (32 * UM, 64 * UM, 32 * UM)
    dom = [ibm.periodic(L[0]), ibm.dirichlet(L[1]), ibm.periodic(L[2])]
    grids = ibm.mac_grids(dom, 128)
    rng = load_inputs().SplitMix64(404)
    lam = 0.5 * 1e-4 * 10.0                                  # dt nu / 2 of the benchmark
    for c in range(3):
        g = grids[c]
        b = _t(rng.normal(g.size))
        solver = ibm.FftSolver(ctx, g)
        x = solver.solve(b, (1.0, -lam))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            solver.solve(b, (1.0, -lam), x=x)
        e1.record()
        torch.cuda.synchronize()
        r = b - ibm.op_apply(ctx, g, (1.0, -lam), x)
        rel = float(r.norm() / b.norm())
        mg = ibm.Multigrid(ctx, g, (1.0, -lam), precision=64)
        xm, it, rr = mg.solve(b, rtol=1e-13, maxit=60)
        diff = float((x - xm).abs().max() / x.abs().max())
        print(f"C4 Helmholtz component {c}: FFT {e0.elapsed_time(e1) / 5:.3f} ms, residual {rel:.2e}; "
              f"multigrid ({it} cycles to {rr:.1e}) differs by {diff:.2e}")
        assert rel < 1e-13 and diff < 1e-11
        mg.close()
        solver.close()
    pdom = [ibm.periodic(L[0]), ibm.neumann(L[1]), ibm.periodic(L[2])]
    pg = ibm.centered_grid(pdom, 128)
    b = _t(rng.normal(pg.size))
    b -= b.mean()
    solver = ibm.FftSolver(ctx, pg)
    x = solver.solve(b, (0.0, 1.0))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        solver.solve(b, (0.0, 1.0), x=x)
    e1.record()
    torch.cuda.synchronize()
    r = b - ibm.op_apply(ctx, pg, (0.0, 1.0), x)
    rel = float(r.norm() / b.norm())
    mg = ibm.Multigrid(ctx, pg, (0.0, 1.0), precision=64)
    xm, it, rr = mg.solve(b, rtol=1e-12, maxit=80)
    diff = float((x - xm).abs().max() / x.abs().max())
    print(f"C4 pressure Poisson: FFT {e0.elapsed_time(e1) / 5:.3f} ms, residual {rel:.2e}, "
          f"mean {float(x.mean()):.1e}; multigrid ({it} cycles to {rr:.1e}) differs by {diff:.2e}")
    assert rel < 1e-12 and diff < 1e-7 and abs(float(x.mean())) < 1e-13 * float(x.abs().max())
    mg.close()
    solver.close()


def test_fft_aliasing_and_errors(ibm, ctx):
    dom = _channel(ibm, (16, 8, 16))
    g = ibm.mac_grids(dom, 4)[0]
    rng = load_inputs().SplitMix64(9)
    b = _t(rng.normal(g.size))
    solver = ibm.FftSolver(ctx, g)
    x = solver.solve(b, (1.0, -0.2))
    y = b.clone()
    solver.solve(y, (1.0, -0.2), x=y)                      # in place
    assert bool((x == y).all())
    solver.close()
    bad = [ibm.mac_grids(_channel(ibm, (12, 8, 16)), 4)[0],               # 12 points: not a power of two
           ibm.mac_grids([ibm.periodic(4 * UM)] * 3, 4)[0],                # no walled axis
           ibm.mac_grids([ibm.dirichlet(4 * UM), ibm.dirichlet(4 * UM), ibm.periodic(4 * UM)], 4)[0]]
    for grid in bad:
        with pytest.raises(ibm.IbmError) as e:
            ibm.FftSolver(ctx, grid)
        assert e.value.code == ibm.ERR_BAD_GRID
