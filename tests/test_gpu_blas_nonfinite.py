"""BLAS-1 entry points (linalg/blas.h:63-98,140-171) against numpy, and the
handling of non-finite inputs: a NaN or infinity must surface as an error
(solvers) or as NaN in the result (spread, reductions), never as a finite
number.

Bars: dot within 1e-13 relative of a compensated host sum and bitwise
reproducible from call to call; axpby within 1 ulp-scale (1e-15 relative)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("n", [0, 1, 1000, 1 << 20, 3_000_017])
def test_dot_matches_host(ibm, ctx, n):
    rng = np.random.default_rng(n)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    a = ibm.dot(ctx, _t(x), _t(y))
    ref = math.fsum(x * y)
    scale = max(float(np.abs(x * y).sum()), 1e-300)
    assert abs(a - ref) <= 1e-13 * scale, (a, ref)
    assert ibm.dot(ctx, _t(x), _t(y)) == a          # fixed summation order


@pytest.mark.parametrize("n", [1, 4097, 1 << 20])
def test_axpby_matches_host(ibm, ctx, n):
    rng = np.random.default_rng(7)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    out = ibm.axpby(ctx, 0.75, _t(x), -1.5, _t(y.copy()))
    ctx.synchronize()
    ref = 0.75 * x - 1.5 * y
    assert np.abs(out.cpu().numpy() - ref).max() <= 1e-15 * np.abs(ref).max()


def _grid(ibm):
    dom = [ibm.periodic(0.4), ibm.dirichlet(0.4), ibm.periodic(0.4)]
    return dom, ibm.grid_make(dom, 40, [0.5, 0.5, 0.5])


@pytest.mark.parametrize("precision", [64, 32])
def test_multigrid_nan_rhs_is_not_converged(ibm, ctx, precision):
    import torch
    _, g = _grid(ibm)
    mg = ibm.Multigrid(ctx, g, ibm.helmholtz(0.4 / g.resolution ** 2), precision)
    b = torch.randn(g.size, dtype=torch.float64, device="cuda")
    b[g.size // 3] = float("nan")
    with pytest.raises(ibm.IbmError) as err:
        mg.solve(b, rtol=1e-10, maxit=20)
    assert err.value.code == ibm.ERR_NOT_CONVERGED
    b[g.size // 3] = float("inf")
    with pytest.raises(ibm.IbmError) as err:
        mg.solve(b, rtol=1e-10, maxit=20, warm=True, x=torch.zeros_like(b))
    assert err.value.code == ibm.ERR_NOT_CONVERGED
    # a fixed-count solve reports a NaN relative residual on the device
    b[g.size // 3] = float("nan")
    _, rr = mg.solve_fixed(b, 3)
    assert math.isnan(float(rr))
    mg.close()


def test_chebpcg_nan_rhs_is_not_converged(ibm, ctx):
    import torch
    _, g = _grid(ibm)
    b = torch.randn(g.size, dtype=torch.float64, device="cuda")
    b[5] = float("nan")
    with pytest.raises(ibm.IbmError) as err:
        ibm.chebpcg_solve(ctx, g, ibm.helmholtz(0.4 / g.resolution ** 2), b, rtol=1e-10, maxit=50)
    assert err.value.code == ibm.ERR_NOT_CONVERGED


@pytest.mark.parametrize("bad", [float("nan"), float("inf")])
def test_spread_nonfinite_force_poisons_grid(ibm, ctx, bad):
    dom = [ibm.periodic(0.4)] * 3
    grids = ibm.mac_grids(dom, 40)
    rng = np.random.default_rng(3)
    n = 500
    x = rng.uniform(0, 0.4, (3, n))
    f = rng.standard_normal((3, n))
    good = ibm.spread(ctx, grids, "cosine", _t(x.reshape(-1)), _t(f.reshape(-1)))
    assert all(bool(c.isfinite().all()) for c in good)
    f[1, 17] = bad
    out = ibm.spread(ctx, grids, "cosine", _t(x.reshape(-1)), _t(f.reshape(-1)))
    for c in out:
        assert bool(c.isnan().all())


def test_solver_nan_velocity_is_an_error(ibm, ctx):
    import torch
    dom = [ibm.periodic(0.4), ibm.dirichlet(0.4), ibm.periodic(0.4)]
    s = ibm.Solver(ctx, dom, 40, 1e-4, 10.0, density=1e-3, tolerance=1e-10, precision=32)
    st = s.zero_state([])
    st.u[0][123] = float("nan")
    with pytest.raises(ibm.IbmError) as err:
        s.step(st)
    assert err.value.code == ibm.ERR_NOT_CONVERGED
    s.close()
