"""The C-ABI library loads, exports every function include/ibm.h declares, and
its host-only grid construction agrees with the oracle (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

UM = 0.1


def declared_functions():
    text = open(os.path.join(ROOT, "include", "ibm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+\**(ibm_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(ibm):
    L = ctypes.CDLL(ibm.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header(ibm):
    assert set(declared_functions()) <= set(ibm.exported_symbols()) | {"ibm_grid_size"}


@pytest.mark.parametrize("walls", [False, True])
def test_grid_make_matches_oracle(ibm, orc, walls):
    dom = [ibm.periodic(16 * UM), ibm.dirichlet(12 * UM) if walls else ibm.periodic(12 * UM),
           ibm.neumann(16 * UM) if walls else ibm.periodic(16 * UM)]
    spec = [(d.length, d.solid, tuple(d.lower), tuple(d.upper)) for d in dom]
    grids = ibm.mac_grids(dom, 20)
    axes = orc.mac_axes(spec, 20)
    fields = ["length", "resolution", "alignment", "shift", "cells", "points", "solid"]
    for c in range(3):
        for j in range(3):
            a, b = grids[c].axis[j], axes[3 * c + j]
            for name in fields:
                assert getattr(a, name) == getattr(b, name), (c, j, name)
            for name in ["boundary", "interior", "laplacian"]:
                assert list(getattr(a, name)) == list(getattr(b, name)), (c, j, name)


def test_grid_errors(ibm):
    with pytest.raises(ibm.IbmError) as e:
        ibm.mac_grids([ibm.periodic(-1.0)] * 3, 8)
    assert e.value.code == ibm.ERR_BAD_GRID
    with pytest.raises(ibm.IbmError) as e:
        ibm.mac_grids([ibm.periodic(1.0)] * 3, 0)
    assert e.value.code == ibm.ERR_ARG


def test_reference_gcd_collapse_is_rejected(ibm):
    # gcd(gcd(1.6, 1.2), 0.8) collapses to ~1e-15 under the reference's
    # floating Euclid (algo/gcd.h:9-23); the grid would need ~1e15 cells.
    assert ibm.domain_unit([ibm.periodic(1.6), ibm.periodic(1.2), ibm.periodic(0.8)]) < 1e-12
    with pytest.raises(ibm.IbmError) as e:
        ibm.mac_grids([ibm.periodic(1.6), ibm.periodic(1.2), ibm.periodic(0.8)], 10)
    assert e.value.code == ibm.ERR_BAD_GRID


def test_grid_coarsen_halves_refinement(ibm):
    dom = [ibm.periodic(32 * UM), ibm.dirichlet(64 * UM), ibm.periodic(32 * UM)]
    g = ibm.centered_grid(dom, 128)
    assert g.points == (128, 256, 128)
    c = ibm.grid_coarsen(g)
    assert c.points == (64, 128, 64)
    assert c.refinement == 64
    assert list(c.axis[1].interior) == list(g.axis[1].interior)


def test_cpp_header_is_plain_cpp():
    """include/ibm.hpp compiles as host C++17 with no HIP headers."""
    import subprocess
    src = '#include "ibm.hpp"\nint main() { fd::mac m{32}; return m.refinement() == 32 ? 0 : 1; }\n'
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), "-x", "c++", "-"],
                       input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_abi_version_matches_header(ibm):
    text = open(os.path.join(ROOT, "include", "ibm.h")).read()
    declared = int(re.search(r"#define IBM_ABI_VERSION (\d+)", text).group(1))
    assert ibm.lib().ibm_abi_version() == declared == ibm.ABI_VERSION
