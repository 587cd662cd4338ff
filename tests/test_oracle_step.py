"""Pins the oracle's timestep composition (oracle/step.py) by what the
scheme fixes exactly: the Crank-Nicolson amplification of a decaying shear
mode, momentum balance on a periodic box (advection, diffusion and the
pressure gradient conserve momentum; spreading conserves the total force),
a divergence-free result after projection, and rest staying at rest."""
import math
import os

import numpy as np
import pytest

from conftest import ROOT, _load, load_orc, load_orc_rbf
import step_cases as C

fd = _load("orc_fd", os.path.join(ROOT, "oracle", "fd.py"))
step_mod = _load("orc_step", os.path.join(ROOT, "oracle", "step.py"))


def _step(walls, types, ref=16):
    return step_mod.Step(load_orc(), fd, load_orc_rbf(), C.domain_spec(walls), ref, C.DT, C.NU, C.RHO, "cosine",
                         types)


def _zeros(st):
    u = [np.zeros(st.I[i].shape[0]) for i in range(3)]
    return u, np.zeros(st.paxes[0].points * st.paxes[1].points * st.paxes[2].points)


def test_shear_mode_decays_with_crank_nicolson_factor():
    st = _step(False, [])
    u, p = _zeros(st)
    ax = st.uaxes[2]                                     # w on (0.5, 0.5, 0) alignment
    n = ax[0].points
    h = 1.0 / ax[0].resolution
    x = (np.arange(n) + ax[0].shift) * h
    mode = np.sin(2 * math.pi * x / 0.8)
    u[2] = np.tile(mode, ax[1].points * ax[2].points)
    un, phi, _, _ = st.step(u, p, [])
    kappa = 4 / h ** 2 * math.sin(math.pi * h / 0.8) ** 2
    lam = C.DT * C.NU / 2
    factor = (1 - lam * kappa) / (1 + lam * kappa)
    assert np.abs(un[2] - factor * u[2]).max() < 1e-12
    assert np.abs(un[0]).max() < 1e-14 and np.abs(un[1]).max() < 1e-14
    assert np.abs(phi).max() < 1e-10


def test_momentum_balance_and_divergence_free_periodic():
    types = C.cell_types()
    st = _step(False, types)
    u, p = _zeros(st)
    xs = [T["X"] for T in types]
    un, phi, xn, geo = st.step(u, p, xs)
    h3 = 1.0 / (st.uaxes[0][0].resolution ** 3)
    total_f = np.zeros(3)
    for T, (y, _, _) in zip(types, geo):
        total_f += (-T["k"] * (y - T["ref"])).sum(axis=(1, 2))
    for i in range(3):
        assert abs(un[i].sum() * h3 - C.DT * total_f[i] / C.RHO) < 1e-9 * np.abs(C.DT * total_f / C.RHO).max()
    div = fd.divergence(st.axes_of, step_mod.MAC, un)
    umax = max(np.abs(c).max() for c in un)
    assert np.abs(div).max() < 1e-9 * umax * st.uaxes[0][0].resolution
    assert umax > 0


def test_walls_divergence_free_and_rest_is_steady():
    types = C.cell_types()
    st = _step(True, types)
    u, p = _zeros(st)
    xs = [T["X"] for T in types]
    un, phi, xn, _ = st.step(u, p, xs)
    div = fd.divergence(st.axes_of, step_mod.MAC, un)
    umax = max(np.abs(c).max() for c in un)
    assert np.abs(div).max() < 1e-9 * umax * st.uaxes[0][0].resolution
    for T in types:
        T["k"] = 0.0
    st0 = _step(True, types)
    u0, p0 = _zeros(st0)
    un0, phi0, xn0, _ = st0.step(u0, p0, xs)
    assert max(np.abs(c).max() for c in un0) == 0
    assert all(np.array_equal(a, b) for a, b in zip(xn0, xs))
