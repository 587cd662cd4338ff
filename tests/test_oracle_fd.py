"""Pins the oracle's restatement of the staggered-grid operators and solvers
(oracle/fd.py).  The reference ships no expected outputs, so each function is
pinned by what the mathematics fixes:

  laplacian     periodic / Dirichlet / Neumann spectra in closed form;
                quadratics reproduced exactly with the second-order identity
                (the test/diffusion.cc steady state: L u = -I2 1)
  identity      Dirichlet/Neumann corner values (3/4, 1)
  differential  div(grad) equals the cell-centred Laplacian (periodic and
                Neumann walls); D applied to a linear field is its slope
  gershgorin    closed-form intervals of the 7-point operators
  chebyshev     1 - z p(z) = T4(s(z)) / T4(s(0)) (algo/chebyshev.h:9-41)
  multigrid     V-cycle contracts the error; solve == sparse direct solve;
                reference transfer order == x-fastest order on centred grids
  chebpcg       == sparse direct solve
"""
import math

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from conftest import ROOT, _load
import os

fd = _load("orc_fd", os.path.join(ROOT, "oracle", "fd.py"))


def P(L):
    return (L, 0, (0.0, 0.0), (0.0, 0.0))


def D(L):
    return (L, 1, (1.0, 0.0), (1.0, 0.0))


def N(L):
    return (L, 1, (0.0, 1.0), (0.0, 1.0))


def axes(orc, spec, ref, align):
    a = orc.grid_axes(spec, ref, align)
    return [a[j] for j in range(3)]


def coords(ax, j):
    return (np.arange(ax[j].points) + ax[j].shift) / ax[j].resolution


def field(ax, fn):
    x, y, z = np.meshgrid(coords(ax, 0), coords(ax, 1), coords(ax, 2), indexing="ij")
    return np.ascontiguousarray(fn(x, y, z).transpose(2, 1, 0).reshape(-1))   # x fastest


def test_periodic_laplacian_spectrum(orc):
    ax = axes(orc, [P(1.0), P(0.75), P(1.25)], 2, [0.5, 0.0, 0.5])    # unit 0.25: 8 x 6 x 10
    L = fd.laplacian(ax)
    n = [a.points for a in ax]
    r = [a.resolution for a in ax]
    for k in [(1, 2, 3), (0, 1, 4), (3, 0, 0)]:
        u = field(ax, lambda x, y, z: np.cos(2 * np.pi * k[0] * x / 1.0) * np.cos(2 * np.pi * k[1] * y / 0.75)
                  * np.cos(2 * np.pi * k[2] * z / 1.25))
        lam = sum(-4 * r[d] ** 2 * math.sin(math.pi * k[d] / n[d]) ** 2 for d in range(3))
        assert np.abs(L @ u - lam * u).max() < 1e-11 * abs(lam)


@pytest.mark.parametrize("kind", ["dirichlet", "neumann"])
def test_walled_laplacian_spectrum(orc, kind):
    wall = D if kind == "dirichlet" else N
    ax = axes(orc, [P(0.5), wall(1.0), P(0.5)], 10, [0.5, 0.5, 0.5])  # unit 0.5 -> 10 x 20 x 10, y cell-centred
    L = fd.laplacian(ax)
    n = ax[1].points
    r = ax[1].resolution
    for k in (1, 3, 7):
        if kind == "dirichlet":
            u = field(ax, lambda x, y, z: np.sin(np.pi * k * y / 1.0) + 0 * x)
        else:
            u = field(ax, lambda x, y, z: np.cos(np.pi * k * y / 1.0) + 0 * x)
        lam = -4 * r * r * math.sin(math.pi * k / (2 * n)) ** 2
        assert np.abs(L @ u - lam * u).max() < 1e-10 * abs(lam)


def test_quadratic_steady_state_with_corrected_identity(orc):
    # test/diffusion.cc:54-100: u0 = y (L - y) / 2 is the steady state of
    # u_t = nu (L u + 1): the scheme satisfies L u0 = -I2 1 exactly.
    for align in ([0.5, 0.5, 0.5], [0.5, 0.0, 0.5]):      # cell-centred and wall-aligned y
        ax = axes(orc, [P(0.5), D(1.0), P(0.5)], 10, align)
        L = fd.laplacian(ax)
        I2 = fd.identity(ax, 2)
        u = field(ax, lambda x, y, z: y * (1.0 - y) / 2 + 0 * x)
        one = np.ones(u.size)
        assert np.abs(L @ u + I2 @ one).max() < 1e-9
    # cell-centred Dirichlet rows carry the 3/4 correction
    ax = axes(orc, [P(0.5), D(1.0), P(0.5)], 10, [0.5, 0.5, 0.5])
    d = fd.identity(ax, 2).diagonal().reshape(ax[2].points, ax[1].points, ax[0].points)
    assert np.allclose(d[:, 0, :], 0.75) and np.allclose(d[:, -1, :], 0.75) and np.allclose(d[:, 1:-1, :], 1.0)


@pytest.mark.parametrize("walls", [False, True])
def test_div_grad_is_cell_centred_laplacian(orc, walls):
    # Divergence on the no-slip (Dirichlet) velocity grids composed with the
    # gradient from the cell centres is the cell-centred Laplacian with
    # homogeneous Neumann walls: the pressure-Poisson operator.
    vel = [P(0.5), D(1.0) if walls else P(1.0), P(0.75)]
    pre = [P(0.5), N(1.0) if walls else P(1.0), P(0.75)]
    ref = 5
    cent = axes(orc, vel, ref, [0.5, 0.5, 0.5])
    Lc = fd.laplacian(axes(orc, pre, ref, [0.5, 0.5, 0.5]))
    total = None
    for i in range(3):
        al = [0.5, 0.5, 0.5]
        al[i] = 0.0
        face = axes(orc, vel, ref, al)
        G = fd.differential(cent, i)           # centred -> face i
        Dv = fd.differential(face, i)          # face i -> centred
        assert G.shape == (np.prod([a.points for a in face]), np.prod([a.points for a in cent]))
        term = Dv @ G
        total = term if total is None else total + term
    assert abs(total - Lc).max() < 1e-9 * abs(Lc).max()


def test_differential_of_linear_field(orc):
    ax = axes(orc, [P(0.5), D(1.0), P(0.5)], 10, [0.5, 0.0, 0.5])   # v grid (walls in y)
    Dy = fd.differential(ax, 1)
    u = field(ax, lambda x, y, z: 3.0 * y + 0 * x)
    out = Dy @ u
    # interior rows: exact slope; the two wall rows see the homogeneous wall value
    o = out.reshape(ax[2].points, ax[1].points + 1, ax[0].points)
    assert np.allclose(o[:, 1:-1, :], 3.0)


def test_gershgorin_closed_forms(orc):
    ax = axes(orc, [P(1.0)] * 3, 8, [0.5] * 3)
    r2 = ax[0].resolution ** 2
    lo, hi = fd.gershgorin(fd.laplacian(ax))
    assert abs(lo + 12 * r2) < 1e-9 * r2 and abs(hi) < 1e-9 * r2
    lam = 0.3 / r2
    lo, hi = fd.gershgorin(fd.operator(ax, 1.0, -lam))
    assert abs(lo - 1) < 1e-12 and abs(hi - (1 + 12 * lam * r2)) < 1e-12


@pytest.mark.parametrize("ab", [(1.0, 5.8), (-1.5, -12.0), (0.3, 2.0)])
def test_chebyshev_weights_are_shifted_t4(ab):
    a, b = ab
    w = fd.chebw4(a, b)

    def p(z):       # Horner of algo/chebyshev.h:155-162 in scalar form
        s = w[0]
        for i in range(1, 4):
            s = w[i] - z * s
        return s

    def t4(x):
        return 8 * x ** 4 - 8 * x ** 2 + 1

    s = lambda z: -1 + 2 * (z - a) / (b - a)
    for z in np.linspace(min(a, b), max(a, b), 9):
        assert abs((1 - z * p(z)) - t4(s(z)) / t4(s(0))) < 1e-12


def _axes_of(orc, spec, align):
    return lambda ref: axes(orc, spec, ref, align)


def test_multigrid_vcycle_contracts_and_solves(orc):
    spec = [P(2.0), D(2.0), P(2.0)]
    af = _axes_of(orc, spec, [0.5, 0.0, 0.5])                # v component, walls in y
    mg = fd.Multigrid(af, 16, 1.0, -0.4 / (af(16)[0].resolution ** 2))
    assert len(mg.levels) >= 2
    A = mg.levels[0]["A"]
    rng = np.random.default_rng(0)
    b = rng.standard_normal(A.shape[0])
    exact = spla.spsolve(A.tocsc(), b)
    e = exact - mg.vcycle(b)
    assert np.linalg.norm(e) < 0.2 * np.linalg.norm(exact)
    x, it = mg.solve(b, 1e-10 * np.linalg.norm(b))
    assert np.linalg.norm(x - exact) < 1e-9 * np.linalg.norm(exact)


def test_transfer_orders_agree_on_centred_grid(orc):
    spec = [P(2.0), N(2.0), P(2.0)]
    af = _axes_of(orc, spec, [0.5, 0.5, 0.5])
    f, c = af(16), af(8)
    for kind in ("restriction", "interpolation"):
        a = fd.transfer(f, c, kind, "reference")
        b = fd.transfer(f, c, kind, "x-fastest")
        assert abs(a - b).max() < 1e-15


def test_chebpcg_matches_direct(orc):
    ax = axes(orc, [P(1.0), D(1.0), P(1.0)], 16, [0.0, 0.5, 0.5])
    lam = 0.4 / ax[0].resolution ** 2
    A = fd.operator(ax, 1.0, -lam)
    b = np.random.default_rng(1).standard_normal(A.shape[0])
    x, it = fd.chebpcg(A, b, 1e-12)
    exact = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - exact) < 1e-10 * np.linalg.norm(exact)
