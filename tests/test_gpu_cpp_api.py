"""The reference-shaped C++ API (include/ibm.hpp) driven by a compiled
program (tests/cpp/api_check.cpp): ib::spread / ib::interpolate on C1 match
the oracle to 1e-12; solvers::multigrid and ins::diffusion results satisfy
their equations (assembled CSR operators of the oracle)."""
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from conftest import ROOT, _load, load_inputs

pytestmark = pytest.mark.gpu

fd = _load("orc_fd", os.path.join(ROOT, "oracle", "fd.py"))
BIN = os.path.join(ROOT, "tests", "cpp", "api_check")


def test_cpp_api_c1(ibm, orc, tmp_path):
    cfg = load_inputs().config_c1()
    cfg["x"].tofile(tmp_path / "x.bin")
    cfg["f"].tofile(tmp_path / "f.bin")
    r = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    dom = [ibm.periodic(16 * ibm.um)] * 3
    spec = [(d.length, d.solid, tuple(d.lower), tuple(d.upper)) for d in dom]
    axes = orc.mac_axes(spec, 32)
    ref = orc.spread(axes, "cosine", cfg["x"], cfg["f"])
    refl = orc.interpolate(axes, "cosine", cfg["x"], ref)
    for c in range(3):
        fe = np.fromfile(tmp_path / f"fe{c}.bin")
        assert np.abs(fe - ref[c]).max() <= 1e-12 * np.abs(ref[c]).max()
    ul = np.fromfile(tmp_path / "ul.bin")
    assert np.abs(ul - refl).max() <= 1e-12 * np.abs(refl).max()

    ax0 = [orc.grid_axes(spec, 32, [0.0, 0.5, 0.5])[j] for j in range(3)]
    h = 1.0 / ax0[0].resolution
    A = fd.operator(ax0, 1.0, -0.4 * h * h)
    xs = np.fromfile(tmp_path / "helm.bin")
    assert np.linalg.norm(A @ xs - ref[0]) <= 1.01e-10 * np.linalg.norm(ref[0])

    dt, nu = 1e-4, 10.0
    L = fd.laplacian(ax0)
    H = fd.operator(ax0, 1.0, -dt * nu / 2)
    du = spla.spsolve(H.tocsc(), dt * nu * (L @ ref[0]))
    un = np.fromfile(tmp_path / "diff.bin")
    assert np.abs(un - (ref[0] + du)).max() <= 1e-9 * np.abs(ref[0]).max()
