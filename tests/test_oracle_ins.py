"""Pins the oracle's ins::advection / divergence / gradient compositions.

  advection   a hand-assembled flux form for u = (f(x), 0, 0):
              H_x[i] = ((u[i] + u[i+1])^2 - (u[i-1] + u[i])^2) / (4h)
              (the test/advection.cc style check); zero for a plane shear
              flow u = (g(y), 0, 0) and for constant flow
  divergence  of a gradient = the Neumann cell-centred Laplacian
              (tests/test_oracle_fd.py) and zero for constant fields
  ins::boundary  closed-form diagonals (Dirichlet centred 2, face-aligned 1,
              Neumann -h; corners of two walled axes sum), the excluded axis
              contributes nothing, and B x equals fd::lower/upper_boundary
              applied to x's wall rows (the Couette-pinned operators of
              tests/test_oracle_fd.py)
  step        with pseudo-pressure wall data (w = u + k grad phi) the
              extrapolated tangential wall slip of a sheared cell-laden
              channel falls against plain wall data
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import ROOT, _load

fd = _load("orc_fd", os.path.join(ROOT, "oracle", "fd.py"))

MAC = [[0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]]


def _axes_of(orc, spec, ref):
    def f(align):
        a = orc.grid_axes(spec, ref, align)
        return [a[j] for j in range(3)]
    return f


def _field(ax, fn):
    c = [(np.arange(a.points) + a.shift) / a.resolution for a in ax]
    x, y, z = np.meshgrid(*c, indexing="ij")
    return np.ascontiguousarray(fn(x, y, z).transpose(2, 1, 0).reshape(-1))


def test_advection_flux_form_1d(orc):
    spec = [(1.0, 0, (0, 0), (0, 0))] * 3
    af = _axes_of(orc, spec, 8)                  # 8^3 periodic, h = 1/8
    ax = [af(a) for a in MAC]
    u = [_field(ax[0], lambda x, y, z: np.sin(2 * np.pi * x) + 0.3 * np.cos(4 * np.pi * x) + 0 * y),
         np.zeros(512), np.zeros(512)]
    H = fd.advection(af, MAC, u)
    U = u[0].reshape(8, 8, 8)                    # [z, y, x]
    h = 1.0 / 8
    up = np.roll(U, -1, axis=2)
    um = np.roll(U, 1, axis=2)
    hand = ((U + up) ** 2 - (um + U) ** 2) / (4 * h)
    assert np.abs(H[0].reshape(8, 8, 8) - hand).max() < 1e-12 * np.abs(hand).max()
    assert np.abs(H[1]).max() == 0 and np.abs(H[2]).max() == 0


def test_advection_vanishes_for_shear_and_constant_flow(orc):
    spec = [(1.0, 0, (0, 0), (0, 0)), (1.0, 1, (1, 0), (1, 0)), (1.0, 0, (0, 0), (0, 0))]
    af = _axes_of(orc, spec, 8)
    ax = [af(a) for a in MAC]
    u = [_field(ax[0], lambda x, y, z: np.sin(np.pi * y) + 0 * x), np.zeros(ax[1][0].points * ax[1][1].points * ax[1][2].points),
         np.zeros(512)]
    H = fd.advection(af, MAC, u)
    assert max(np.abs(h).max() for h in H) < 1e-12
    spec = [(1.0, 0, (0, 0), (0, 0))] * 3
    af = _axes_of(orc, spec, 8)
    u = [np.full(512, 0.7), np.full(512, -0.2), np.full(512, 1.3)]
    H = fd.advection(af, MAC, u)
    assert max(np.abs(h).max() for h in H) < 1e-12
    assert np.abs(fd.divergence(af, MAC, u)).max() < 1e-12


# ----------------------------------------------------------------------
# ins::boundary (ins/boundary.h:16-56)
# ----------------------------------------------------------------------
def _wall_rows(ax, axis):
    n = [a.points for a in ax]
    idx = np.zeros((n[2], n[1], n[0]), dtype=bool)
    sl_lo = [slice(None)] * 3
    sl_hi = [slice(None)] * 3
    sl_lo[2 - axis] = 0
    sl_hi[2 - axis] = n[axis] - 1
    return idx, tuple(sl_lo), tuple(sl_hi)


@pytest.mark.parametrize("kind,align", [("dirichlet", 0.5), ("dirichlet", 0.0), ("neumann", 0.5)])
def test_ins_boundary_closed_forms(orc, kind, align):
    wall = (1.0, 1, (1.0, 0.0), (1.0, 0.0)) if kind == "dirichlet" else (1.0, 1, (0.0, 1.0), (0.0, 1.0))
    spec = [(0.5, 0, (0, 0), (0, 0)), wall, (0.5, 0, (0, 0), (0, 0))]
    ax = _axes_of(orc, spec, 10)([0.5, align, 0.5])
    h = 1.0 / ax[1].resolution
    # ghost u_-1 = A gamma + B u_0: Dirichlet centred A = 2 (u_-1 = 2 g - u_0),
    # Dirichlet on the dropped wall point A = 1 (u_-1 = g), Neumann centred
    # A = -h at the lower wall and +h at the upper one (outward derivative)
    lo, hi = {("dirichlet", 0.5): (2.0, 2.0), ("dirichlet", 0.0): (1.0, 1.0),
              ("neumann", 0.5): (-h, h)}[(kind, align)]
    for order in (0, 2):
        B = fd.ins_boundary(ax, 0, order)
        d = B.diagonal().reshape(ax[2].points, ax[1].points, ax[0].points)
        assert abs(B - sp.diags(B.diagonal())).max() == 0          # diagonal
        sc = ax[1].resolution ** 2 if order == 2 else 1.0
        assert np.allclose(d[:, 0, :], lo * sc, rtol=1e-14) and np.allclose(d[:, -1, :], hi * sc, rtol=1e-14)
        assert not d[:, 1:-1, :].any()
    # the excluded axis (comp == dim) contributes nothing
    assert fd.ins_boundary(ax, 1, 0).nnz == 0


def test_ins_boundary_corners_sum_two_walled_axes(orc):
    wall = (1.0, 1, (1.0, 0.0), (1.0, 0.0))
    spec = [wall, wall, (0.5, 0, (0, 0), (0, 0))]
    ax = _axes_of(orc, spec, 8)([0.5, 0.5, 0.5])
    d = fd.ins_boundary(ax, 2, 0).diagonal().reshape(ax[2].points, ax[1].points, ax[0].points)
    assert np.allclose(d[:, 0, 0], 4.0) and np.allclose(d[:, -1, -1], 4.0)     # x and y walls: 2 + 2
    assert np.allclose(d[:, 0, 3], 2.0) and np.allclose(d[:, 3, 0], 2.0)
    # excluding x leaves only the y walls
    d = fd.ins_boundary(ax, 0, 0).diagonal().reshape(ax[2].points, ax[1].points, ax[0].points)
    assert np.allclose(d[:, 0, 0], 2.0) and np.allclose(d[:, 3, 0], 0.0)


@pytest.mark.parametrize("align", [[0.0, 0.5, 0.5], [0.5, 0.5, 0.0]])
@pytest.mark.parametrize("order", [0, 2])
def test_ins_boundary_is_boundary_data_of_the_wall_rows(orc, align, order):
    """B x = B_lo (x on the first y row) + B_hi (x on the last y row), with
    B_lo / B_hi the fd::lower/upper_boundary operators (one walled axis, so
    their corrected identities on the other axes are plain identities)."""
    wall = (1.0, 1, (1.0, 0.0), (1.0, 0.0))
    spec = [(0.5, 0, (0, 0), (0, 0)), wall, (0.75, 0, (0, 0), (0, 0))]
    ax = _axes_of(orc, spec, 8)(align)
    n = [a.points for a in ax]
    x = np.random.default_rng(4).standard_normal(np.prod(n))
    X = x.reshape(n[2], n[1], n[0])
    lo = fd.boundary_op(ax, 1, True, order) @ X[:, 0, :].reshape(-1)
    hi = fd.boundary_op(ax, 1, False, order) @ X[:, -1, :].reshape(-1)
    dim = 0 if align[0] == 0.0 else 2
    assert np.abs(fd.ins_boundary(ax, dim, order) @ x - (lo + hi)).max() < 1e-12 * np.abs(lo + hi).max()


def test_pseudo_pressure_walls_reduce_slip(orc):
    import step_cases as C
    from conftest import load_orc_rbf
    step_mod = _load("orc_step", os.path.join(ROOT, "oracle", "step.py"))
    O = load_orc_rbf()
    types = C.cell_types()
    slip = {}
    for pp in (False, True):
        st = step_mod.Step(orc, fd, O, C.domain_spec(True), 16, C.DT, C.NU, C.RHO, "cosine", types,
                           pseudo_pressure=pp)
        u = [np.zeros(st.I[i].shape[0]) for i in range(3)]
        p = np.zeros(np.prod([a.points for a in st.paxes]))
        xs = [T["X"].copy() for T in types]
        out = []
        for _ in range(4):
            u, p, xs, _ = st.step(u, p, xs)
            ax = st.uaxes[0]
            U = u[0].reshape(ax[2].points, ax[1].points, ax[0].points)
            # second-order extrapolation to the walls (cell-centred y: walls
            # half a cell below the first and above the last row)
            out.append(max(np.abs(1.5 * U[:, 0, :] - 0.5 * U[:, 1, :]).max(),
                           np.abs(1.5 * U[:, -1, :] - 0.5 * U[:, -2, :]).max()))
        slip[pp] = out
    # identical first step (p = 0), then less slip (measured 0.0118 vs 0.0156 at step 4)
    assert slip[True][0] == slip[False][0]
    for k in (1, 2, 3):
        assert slip[True][k] < 0.85 * slip[False][k], slip
