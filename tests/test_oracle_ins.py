"""Pins the oracle's ins::advection / divergence / gradient compositions.

  advection   a hand-assembled flux form for u = (f(x), 0, 0):
              H_x[i] = ((u[i] + u[i+1])^2 - (u[i-1] + u[i])^2) / (4h)
              (the test/advection.cc style check); zero for a plane shear
              flow u = (g(y), 0, 0) and for constant flow
  divergence  of a gradient = the Neumann cell-centred Laplacian
              (tests/test_oracle_fd.py) and zero for constant fields
"""
import os

import numpy as np

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
