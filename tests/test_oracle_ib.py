"""Pins the CPU oracle's restatement of the reference's IB coupling.

The reference ships no golden vectors or known-answer tests for this path
(its test/*.cc programs print values without asserting anything), so each
oracle function is pinned by what the mathematics fixes:

  delta kernels   closed forms; partition of unity; symmetry; support;
                  Peskin's moment conditions (first moment 0, sum of squares
                  3/8); B-spline linear reproduction
  grid geometry   textbook ghost-cell coefficients for Dirichlet/Neumann
                  walls; counts of fd::discretization; the product's own
                  independent host implementation
  spread          long-double direct evaluation of the defining sum;
                  conservation sum(fe) h^3 = sum(f); exact translation by one
                  mesh width; the test/grid.cc scenario
  interpolate     long-double direct evaluation; adjointness with spread;
                  constants reproduced exactly (partition of unity)
"""
import math

import numpy as np
import pytest

UM = 0.1


def _periodic(L):
    return (L, 0, (0.0, 0.0), (0.0, 0.0))


def _dirichlet(L):
    return (L, 1, (1.0, 0.0), (1.0, 0.0))


def _neumann(L):
    return (L, 1, (0.0, 1.0), (0.0, 1.0))


# ----------------------------------------------------------------------
# Delta kernels
# ----------------------------------------------------------------------
@pytest.mark.parametrize("delta", ["cosine", "peskin", "bspline3"])
def test_delta_partition_symmetry_support(orc, delta):
    for d in np.linspace(-1, 0, 41)[1:]:
        vals = [orc.phi(delta, d + s) for s in (-1, 0, 1, 2)]
        assert abs(sum(vals) - 1.0) < 1e-15
        assert all(v >= -1e-16 for v in vals)
    for r in np.linspace(0, 2, 37):
        assert abs(orc.phi(delta, r) - orc.phi(delta, -r)) < 1e-16
    assert abs(orc.phi(delta, 2.0)) < 1e-15
    assert abs(orc.phi(delta, -2.0)) < 1e-15


def test_delta_cosine_closed_form(orc):
    assert orc.phi("cosine", 0.0) == 0.5
    assert abs(orc.phi("cosine", 1.0) - 0.25) < 1e-16
    for r in np.linspace(-2, 2, 17):
        assert abs(orc.phi("cosine", r) - 0.25 * (1 + math.cos(math.pi * r / 2))) < 1e-16


def test_delta_peskin_moments(orc):
    # Peskin (2002): even/odd sums 1/2, first moment 0, sum of squares 3/8.
    for d in np.linspace(-1, 0, 33)[1:]:
        rs = [d + s for s in (-1, 0, 1, 2)]
        v = [orc.phi("peskin", r) for r in rs]
        assert abs(v[0] + v[2] - 0.5) < 1e-15
        assert abs(v[1] + v[3] - 0.5) < 1e-15
        assert abs(sum(r * w for r, w in zip(rs, v))) < 1e-15
        assert abs(sum(w * w for w in v) - 3.0 / 8.0) < 1e-15


def test_delta_bspline_closed_form_and_linear_reproduction(orc):
    def b3(r):
        s = abs(r)
        if s < 1:
            return 2.0 / 3.0 - s * s + s ** 3 / 2
        if s < 2:
            return (2 - s) ** 3 / 6
        return 0.0

    for r in np.linspace(-2.5, 2.5, 101):
        assert abs(orc.phi("bspline3", r) - b3(r)) < 1e-15
    for d in np.linspace(-1, 0, 33)[1:]:
        rs = [d + s for s in (-1, 0, 1, 2)]
        assert abs(sum(r * orc.phi("bspline3", r) for r in rs)) < 1e-15


# ----------------------------------------------------------------------
# Grid geometry
# ----------------------------------------------------------------------
def test_grid_counts_and_ghost_coefficients(orc):
    # 16 um periodic x, 12 um Dirichlet y, 16 um Neumann z; fd::mac{20}
    # over the 4 um base unit -> resolution 5 per unit, h = 0.2 um... in library units
    spec = [_periodic(16 * UM), _dirichlet(12 * UM), _neumann(16 * UM)]
    axes = orc.mac_axes(spec, 20)
    unit = 4 * UM
    r = 20 / unit
    h = 1 / r
    for c in range(3):
        for j in range(3):
            a = axes[3 * c + j]
            assert abs(a.resolution - r) < 1e-12 * r
            cells = int(spec[j][0] * r + 0.5)
            assert a.cells == cells
            face = (c == j)
            assert a.alignment == (0.0 if face else 0.5)
            solid = spec[j][1]
            assert a.points == cells - (1 if (face and solid) else 0)
            assert a.shift == (0.0 if face else 0.5) + (1 if (face and solid) else 0)
    # Dirichlet, cell-centred: ghost u_-1 = 2 g - u_0; identity correction -1/4
    a = axes[3 * 0 + 1]
    assert np.allclose(list(a.boundary), [2.0, 2.0])
    assert np.allclose(list(a.interior), [-1.0, -1.0])
    assert np.allclose(list(a.laplacian), [-0.25, -0.25])
    # Dirichlet, on the wall (v grid): ghost is the wall value itself
    a = axes[3 * 1 + 1]
    assert np.allclose(list(a.boundary), [1.0, 1.0])
    assert np.allclose(list(a.interior), [0.0, 0.0])
    assert np.allclose(list(a.laplacian), [0.0, 0.0])
    # Neumann, cell-centred: u_-1 = u_0 -/+ h g; no correction
    a = axes[3 * 0 + 2]
    assert np.allclose(list(a.boundary), [-h, h])
    assert np.allclose(list(a.interior), [1.0, 1.0])
    assert np.allclose(list(a.laplacian), [0.0, 0.0])


def test_domain_unit_gcd(orc):
    # base unit of (16, 12, 16) um is 4 um; of (32, 64, 32) um is 32 um
    axes = orc.grid_axes([_periodic(16 * UM), _dirichlet(12 * UM), _periodic(16 * UM)], 1, [0.5] * 3)
    assert abs(1 / axes[0].resolution - 4 * UM) < 1e-15
    axes = orc.grid_axes([_periodic(32 * UM), _dirichlet(64 * UM), _periodic(32 * UM)], 128, [0.5] * 3)
    assert [axes[j].cells for j in range(3)] == [128, 256, 128]


# ----------------------------------------------------------------------
# Spread / interpolate
# ----------------------------------------------------------------------
SMALL_SPECS = {
    "periodic": [_periodic(0.8), _periodic(1.0), _periodic(1.2)],      # 8 x 10 x 12 cells
    "walls": [_periodic(0.8), _dirichlet(1.0), _periodic(1.2)],
}


def _small(orc, inputs, kind, n=60, seed=11, margin_cells=0.0):
    spec = SMALL_SPECS[kind]
    axes = orc.mac_axes(spec, 2)          # unit 0.2 -> resolution 10
    L = [s[0] for s in spec]
    h = 1 / axes[0].resolution
    x = inputs.random_markers(n, L, seed, margin=margin_cells * h)
    rng = inputs.SplitMix64(seed + 7)
    f = rng.normal(3 * n)
    return axes, x, f, h


@pytest.mark.parametrize("kind", ["periodic", "walls"])
@pytest.mark.parametrize("delta", ["cosine", "peskin", "bspline3"])
def test_spread_matches_direct_sum(orc, inputs, kind, delta):
    axes, x, f, h = _small(orc, inputs, kind)
    a = orc.spread(axes, delta, x, f)
    b = orc.spread(axes, delta, x, f, direct=True)
    for c in range(3):
        scale = np.abs(b[c]).max()
        assert np.abs(a[c] - b[c]).max() <= 1e-13 * scale


@pytest.mark.parametrize("kind", ["periodic", "walls"])
@pytest.mark.parametrize("delta", ["cosine", "peskin", "bspline3"])
def test_interp_matches_direct_sum(orc, inputs, kind, delta):
    axes, x, f, h = _small(orc, inputs, kind)
    rng = inputs.SplitMix64(5)
    ue = [rng.normal(s) for s in orc.sizes(axes)]
    a = orc.interpolate(axes, delta, x, ue)
    b = orc.interpolate(axes, delta, x, ue, direct=True)
    assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()


@pytest.mark.parametrize("delta", ["cosine", "peskin"])
def test_spread_conservation_and_adjointness(orc, inputs, delta):
    axes, x, f, h = _small(orc, inputs, "periodic", n=200)
    fe = orc.spread(axes, delta, x, f)
    rng = inputs.SplitMix64(9)
    ue = [rng.normal(s) for s in orc.sizes(axes)]
    ul = orc.interpolate(axes, delta, x, ue)
    n = x.size // 3
    h3 = h ** 3
    lhs = sum(float(fe[c] @ ue[c]) for c in range(3)) * h3
    rhs = float(f @ ul)
    assert abs(lhs - rhs) <= 1e-13 * max(abs(lhs), 1.0)
    for c in range(3):
        assert abs(fe[c].sum() * h3 - f[c * n:(c + 1) * n].sum()) <= 1e-12


def test_spread_translation_by_one_cell(orc):
    # h = 1/8 exactly (length 1, refinement 8); markers on a 1/1024 lattice
    spec = [_periodic(1.0)] * 3
    axes = orc.mac_axes(spec, 8)
    h = 1.0 / 8
    rng = np.random.default_rng(3)
    n = 40
    x = (rng.integers(0, 1024, size=3 * n) / 1024.0).astype(np.float64)
    f = rng.standard_normal(3 * n)
    base = orc.spread(axes, "cosine", x, f)
    for j in range(3):
        xs = x.copy()
        xs[j * n:(j + 1) * n] = np.mod(xs[j * n:(j + 1) * n] + h, 1.0)
        moved = orc.spread(axes, "cosine", xs, f)
        for c in range(3):
            rolled = np.roll(base[c].reshape(8, 8, 8), 1, axis=2 - j).reshape(-1)
            assert np.abs(moved[c] - rolled).max() <= 1e-13 * np.abs(base[c]).max()


def test_grid_cc_scenario(orc):
    # test/grid.cc:23-63: two markers at (5,5,5) um in a periodic 10 um box,
    # fd::mac{16}; f = 1.  The spread integral equals the total force and the
    # interpolated unit field is exactly 1.
    spec = [_periodic(10 * UM)] * 3
    axes = orc.mac_axes(spec, 16)
    h = 1 / axes[0].resolution
    x = np.full(6, 5 * UM)
    f = np.ones(6)
    fe = orc.spread(axes, "cosine", x, f)
    for c in range(3):
        assert abs(fe[c].sum() * h ** 3 - 2.0) < 1e-13
    ue = [np.ones(s) for s in orc.sizes(axes)]
    ul = orc.interpolate(axes, "cosine", x, ue)
    assert np.abs(ul - 1.0).max() < 1e-15


def test_walls_drop_in_spread_and_clamp_in_interp(orc):
    # A marker half a cell from the lower y wall: spread drops stencil points
    # outside the v grid, interpolation clamps them onto the first row.
    spec = [_periodic(0.8), _dirichlet(0.8), _periodic(0.8)]
    axes = orc.mac_axes(spec, 8)            # unit 0.8 -> resolution 10, h = 0.1
    h = 0.1
    x = np.array([0.4, 0.5 * h, 0.4])
    f = np.array([0.0, 1.0, 0.0])
    fe = orc.spread(axes, "cosine", x, f)
    direct = orc.spread(axes, "cosine", x, f, direct=True)
    assert np.abs(fe[1] - direct[1]).max() < 1e-13 * np.abs(direct[1]).max()
    assert fe[1].sum() * h ** 3 < 1.0 - 1e-3          # mass lost through the wall
    ue = [np.ones(s) for s in orc.sizes(axes)]
    ul = orc.interpolate(axes, "cosine", x, ue)
    assert np.abs(ul - 1.0).max() < 1e-15             # clamped stencil still sums to 1
