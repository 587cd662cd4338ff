"""The oracle's restatement of the reference's linalg binary format
(oracle/linalg_io.py), pinned by the reference's own data files and by the
format's fixed magic words.

The reference ships its parameter-site and quadrature-weight files in this
format (data/*.bin); they are read where the reference tree is present and
the tests skip elsewhere.  The weights file pins the reader's value order and
type through a property of the quadrature it holds: for the unit sphere,
sum_i w_i |x_theta x x_phi|(i) = sum_i w_i cos(phi_i) = 4 pi at the 15000
Bauer-spiral sites the file was made for (bases/shapes/sphere.h:80-115).
"""
import math
import os
import struct

import numpy as np
import pytest

from conftest import ROOT, _load, load_inputs

REF_DATA = "/root/reference/data"


def _oracle():
    return _load("orc_linalg_io", os.path.join(ROOT, "oracle", "linalg_io.py"))


def _fixture(name):
    path = os.path.join(REF_DATA, name)
    if not os.path.exists(path):
        pytest.skip("reference data files not present")
    with open(path, "rb") as f:
        return f.read()


def test_magic_words_spell_the_reference_tags():
    io = _oracle()
    # "ve"/"ma" container, "d" dense, "d" double: linalg/io.h:67-85, stored little-endian
    assert io.dense_bytes(np.zeros(3), False)[:4] == b"ddve"
    assert io.dense_bytes(np.zeros((2, 3)), True)[:4] == b"ddma"
    assert struct.pack("<i", io.magic(False, single=True)) == b"sdve"
    assert struct.pack("<i", io.magic(True, sparse=True)) == b"dsma"


def test_weights_fixture_reads_as_the_sphere_quadrature():
    io = _oracle()
    buf = _fixture("sphere.w.15000.bin")
    w, end = io.read_dense(buf, 0, False)
    assert end == len(buf) and w.size == 15000
    th, ph = load_inputs().bauer_params(15000)
    assert abs((w * np.cos(ph)).sum() - 4 * math.pi) < 1e-9
    assert io.dense_bytes(w, False) == buf


@pytest.mark.parametrize("name,n", [("rbc.2210.bin", 2210), ("plt.605.bin", 605), ("rbc.8832.bin", 8832)])
def test_site_fixtures_read_as_parameter_pairs(name, n):
    io = _oracle()
    buf = _fixture(name)
    sites, end = io.read_dense(buf, 0, True)
    assert end == len(buf) and sites.shape == (n, 2)
    theta, phi = sites[:, 0], sites[:, 1]
    # column-major: all longitudes, then all latitudes
    assert np.all(np.abs(theta) <= math.pi) and np.all(np.abs(phi) <= math.pi / 2)
    assert theta.max() - theta.min() > 6.0 and phi.max() - phi.min() > 2.8
    assert io.dense_bytes(sites, True) == buf


def test_state_records_round_trip_and_drop_a_truncated_tail():
    io = _oracle()
    rng = np.random.default_rng(0)
    recs = []
    buf = b""
    for k in range(3):
        u = [rng.standard_normal(n) for n in (12, 10, 12)]
        p = rng.standard_normal(8)
        xs = [rng.standard_normal((5, 6)), rng.standard_normal((4, 3))]
        recs.append((0.5 * k, u, p, xs))
        buf += io.state_bytes(0.5 * k, u, p, xs)
    got = io.read_states(buf + buf[: len(buf) // 5], 3, 2)
    assert len(got) == 3
    for (t, u, p, xs), (t2, u2, p2, xs2) in zip(recs, got):
        assert t == t2
        assert all(np.array_equal(a, b) for a, b in zip(u, u2))
        assert np.array_equal(p, p2)
        assert all(np.array_equal(a, b) for a, b in zip(xs, xs2))
    with pytest.raises(ValueError):
        io.read_dense(buf, 8, True)       # a vector where a matrix is expected
