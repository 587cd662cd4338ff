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
  transfers     interp1 reproduces linear fields on interior rows of periodic,
                Dirichlet and Neumann axes, face-aligned and centred; Dirichlet
                wall rows reproduce fields vanishing at that wall, Neumann wall
                rows reproduce constants; restrict1 = interp1^T / 2 except the
                reference's Neumann face-aligned wall fold
  smoother      interval [a + (b - a) prop, b] with prop and the Gershgorin
                interval in closed form (cube: 1/8; 32 x 64 x 32 box: 1/16)
  boundary_op   Couette identity L u_lin + B_lo g_lo + B_hi g_hi = 0 for linear
                wall-to-wall profiles on centred and face-aligned walled axes
  transfer order  on staggered grids the reference's kron order is not a
                linear interpolation along x and z, the per-axis order is;
                both V-cycles contract alike and stay within 1e-3 of each other
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


# ----------------------------------------------------------------------
# Multigrid transfers (solvers/mg/interpolation.h:24-120)
# ----------------------------------------------------------------------
def _axis(orc, spec, ref, align):
    return orc.grid_axes([spec] * 3, ref, [align] * 3)[0]


def _coords1(a):
    return (np.arange(a.points) + a.shift) / a.resolution


@pytest.mark.parametrize("kind", ["periodic", "dirichlet", "neumann"])
@pytest.mark.parametrize("align", [0.0, 0.5])
def test_interpolation_reproduces_linear_fields(orc, kind, align):
    spec = {"periodic": P(1.0), "dirichlet": D(1.0), "neumann": N(1.0)}[kind]
    f, c = _axis(orc, spec, 16, align), _axis(orc, spec, 8, align)
    Pm = fd.interp1(f, c).toarray()
    xf, xc = _coords1(f), _coords1(c)
    # interior rows: the 4-tap weights 1 - |dx|/2 interpolate linearly
    err = Pm @ (2 * xc + 1) - (2 * xf + 1)
    assert np.abs(err[1:-1]).max() < 1e-13
    if kind == "dirichlet":
        # the ghost value -u0 (centred) / the dropped wall point (face) is the
        # homogeneous wall: fields vanishing at a wall are exact on its row
        assert abs((Pm @ xc - xf)[0]) < 1e-13
        assert abs((Pm @ (1 - xc) - (1 - xf))[-1]) < 1e-13
        assert abs((Pm @ np.ones(c.points))[0] - 1) > 0.1     # constants are not
    else:
        # periodic wrap and Neumann ghost (= u0) keep constants on every row
        assert np.abs(Pm @ np.ones(c.points) - 1).max() < 1e-13
    if kind == "periodic":
        # the wrap rows see the coarse field's other end
        assert abs(err[-1]) > 0.1


@pytest.mark.parametrize("kind", ["periodic", "dirichlet", "neumann"])
@pytest.mark.parametrize("align", [0.0, 0.5])
def test_restriction_is_half_interpolation_transpose(orc, kind, align):
    spec = {"periodic": P(1.0), "dirichlet": D(1.0), "neumann": N(1.0)}[kind]
    f, c = _axis(orc, spec, 16, align), _axis(orc, spec, 8, align)
    Pm = fd.interp1(f, c).toarray()
    R = fd.restrict1(f, c).toarray()
    d = R - Pm.T / 2
    if kind == "neumann" and align == 0.0:
        # The reference folds the wall ghost into its interpolation with
        # weights[3] = 1/2 but into its restriction with weights[0] / 2 = 0
        # (solvers/mg/interpolation.h:79-80 vs 116-118): exactly the two
        # corner entries differ, by 1/4.
        bad = np.argwhere(np.abs(d) > 1e-15)
        assert sorted(map(tuple, bad)) == [(0, 0), (c.points - 1, f.points - 1)]
        assert np.allclose(np.abs(d[tuple(bad.T)]), 0.25)
    else:
        assert np.abs(d).max() < 1e-15


def test_smoother_interval_closed_form(orc):
    # cube, periodic: Gershgorin of I - lam L is [1, 1 + 12 lam r^2], prop 1/8
    ax = axes(orc, [P(1.0)] * 3, 8, [0.5] * 3)
    r2 = ax[0].resolution ** 2
    lam = 0.4 / r2
    a, b = fd.smoother_interval(ax, fd.operator(ax, 1.0, -lam))
    lo, hi = 1.0, 1 + 12 * lam * r2
    assert abs(a - (lo + (hi - lo) / 8)) < 1e-12 and abs(b - hi) < 1e-12
    # 32 x 64 x 32 box (the C4 shape at 1/4 scale), y no-slip walls, cell
    # centred in y: wall rows of I~ - lam L carry 3/4 + 6 lam r^2 on the
    # diagonal and 4 lam r^2 off it, so lo = min(1, 3/4 + 2 lam r^2);
    # prop = (16/32) (16/64) (16/32) = 1/16.
    for lamr2 in (0.1, 0.4):
        ax = axes(orc, [P(0.32), D(0.64), P(0.32)], 25, [0.0, 0.5, 0.5])
        r2 = ax[1].resolution ** 2
        lam = lamr2 / r2
        A = fd.operator(ax, 1.0, -lam)
        a, b = fd.smoother_interval(ax, A)
        lo, hi = min(1.0, 0.75 + 2 * lamr2), 1 + 12 * lamr2
        assert abs(b - hi) < 1e-12
        assert abs(a - (lo + (hi - lo) / 16)) < 1e-12, (a, lo + (hi - lo) / 16)
    # the Laplacian's interval [-12 r^2, 0] has |a| > |b|: the ends swap to
    # (0, -12 r^2), so the smoothing interval is [-12 r^2 / 8, -12 r^2]
    ax = axes(orc, [P(1.0)] * 3, 8, [0.5] * 3)
    r2 = ax[0].resolution ** 2
    a, b = fd.smoother_interval(ax, fd.laplacian(ax))
    assert abs(b + 12 * r2) < 1e-9 * r2 and abs(a + 12 * r2 / 8) < 1e-9 * r2


@pytest.mark.parametrize("align", [0.5, 0.0])
def test_boundary_operators_couette_identity(orc, align):
    # A linear profile between walls with values g_lo, g_hi is a steady state:
    # L u + B_lo g_lo + B_hi g_hi = 0 on every row (fd/boundary_ops.h:29-71
    # supplies the ghost-cell contribution A gamma r^2 the Laplacian's wall
    # rows leave out).  y is cell-centred (u, w grids) or face-aligned with
    # the wall points dropped (v grid).
    ax = axes(orc, [P(0.5), D(1.0), P(0.5)], 10, [0.5, align, 0.5])
    L = fd.laplacian(ax)
    g_lo, g_hi = 0.3, 1.7
    u = field(ax, lambda x, y, z: g_lo + (g_hi - g_lo) * y / 1.0 + 0 * x)
    Blo = fd.boundary_op(ax, 1, True, 2)
    Bhi = fd.boundary_op(ax, 1, False, 2)
    nface = ax[0].points * ax[2].points
    assert Blo.shape == (u.size, nface)
    res = L @ u + Blo @ np.full(nface, g_lo) + Bhi @ np.full(nface, g_hi)
    assert np.abs(res).max() < 1e-12 * ax[1].resolution ** 2
    # without the boundary data the wall rows are off by A gamma r^2
    assert np.abs(L @ u).max() > 0.1 * ax[1].resolution ** 2
    # order 0 is the same operator without the r^2 factor
    B0 = fd.boundary_op(ax, 1, False, 0)
    assert abs(B0 * ax[1].resolution ** 2 - Bhi).max() < 1e-9


def test_reference_transfer_order_on_staggered_grid(orc):
    """The reference folds its 1-D transfers as kron(kron(Px, Py), Pz), which
    applies Pz along the fastest (x) index of the grid vectors; on the u
    grid (x face-aligned, z centred) that is no longer a linear interpolation
    along x or z.  The matrix-free kernels apply each axis's own transfer
    (the x-fastest fold).  Both V-cycles are convergent and agree closely."""
    spec = [P(2.0), D(2.0), P(2.0)]
    align = [0.0, 0.5, 0.5]
    af = _axes_of(orc, spec, align)
    f, c = af(16), af(8)
    n = [a.points for a in f]
    for order, exact in (("reference", False), ("x-fastest", True)):
        Pm = fd.transfer(f, c, "interpolation", order)
        for fn in (lambda x, y, z: x + 0 * y, lambda x, y, z: z + 0 * x):
            e = (Pm @ field(c, fn) - field(f, fn)).reshape(n[2], n[1], n[0])[2:-2, 2:-2, 2:-2]
            assert (np.abs(e).max() < 1e-13) == exact
    lam = 0.4 / f[0].resolution ** 2
    b = np.random.default_rng(0).standard_normal(np.prod(n))
    A = fd.operator(f, 1.0, -lam)
    exact = spla.spsolve(A.tocsc(), b)
    v = {}
    for order in ("reference", "x-fastest"):
        mg = fd.Multigrid(af, 16, 1.0, -lam, order)
        v[order] = mg.vcycle(b)
        assert np.linalg.norm(exact - v[order]) < 2e-3 * np.linalg.norm(exact)
    # measured 1.6e-4
    assert np.linalg.norm(v["reference"] - v["x-fastest"]) < 1e-3 * np.linalg.norm(v["x-fastest"])
