"""Checkpoint / resume in the reference's linalg binary format through the
C ABI (ibm_state_write / ibm_state_read / ibm_io_*), checked against the
oracle's independent reader and writer (oracle/linalg_io.py): files the
library writes parse bit for bit, files the oracle writes load bit for bit,
the last complete record is found when a run was cut off mid-record
(util/kaboose.cc), and shape or type mismatches are refused."""
import os

import numpy as np
import pytest

from conftest import ROOT, _load

pytestmark = pytest.mark.gpu


def _oracle():
    return _load("orc_linalg_io", os.path.join(ROOT, "oracle", "linalg_io.py"))


def _rand_state(ibm, rng, u_sizes, p_size, shapes, t):
    import torch
    u = [torch.from_numpy(rng.standard_normal(n)).cuda() for n in u_sizes]
    p = torch.from_numpy(rng.standard_normal(p_size)).cuda()
    xs = [torch.from_numpy(rng.standard_normal(r * c)).cuda() for r, c in shapes]
    return ibm.State(u, p, xs, t)


def test_state_files_match_the_oracle_format(ibm, ctx, tmp_path):
    io = _oracle()
    rng = np.random.default_rng(3)
    u_sizes, p_size, shapes = (64, 60, 64), 64, [(7, 6), (3, 9)]
    L = ibm.state_layout(u_sizes, p_size, shapes)
    path = str(tmp_path / "states.bin")
    states = [_rand_state(ibm, rng, u_sizes, p_size, shapes, 0.25 * k) for k in range(3)]
    for k, st in enumerate(states):
        ibm.state_write(ctx, path, st, L, append=k > 0)
    buf = open(path, "rb").read()
    recs = io.read_states(buf, 3, 2)
    assert len(recs) == 3
    for st, (t, u, p, xs) in zip(states, recs):
        assert t == st.t
        for a, b in zip(st.u, u):
            assert np.array_equal(a.cpu().numpy(), b)
        assert np.array_equal(st.p.cpu().numpy(), p)
        for a, b, (r, c) in zip(st.xs, xs, shapes):
            # column-major rows x cols on the device, as bases::matrix stores it
            assert np.array_equal(a.cpu().numpy().reshape(c, r).T, b)
    # the oracle's bytes are the library's bytes
    ref = b"".join(io.state_bytes(st.t, [a.cpu().numpy() for a in st.u], st.p.cpu().numpy(),
                                  [a.cpu().numpy().reshape(c, r).T for a, (r, c) in zip(st.xs, shapes)])
                   for st in states)
    assert ref == buf


def test_resume_reads_the_last_complete_record(ibm, ctx, tmp_path):
    import torch
    io = _oracle()
    rng = np.random.default_rng(4)
    u_sizes, p_size, shapes = (40, 36, 40), 40, [(5, 12)]
    L = ibm.state_layout(u_sizes, p_size, shapes)
    recs = []
    for k in range(4):
        u = [rng.standard_normal(n) for n in u_sizes]
        p = rng.standard_normal(p_size)
        x = rng.standard_normal(shapes[0])
        recs.append((1.5 * k, u, p, [x]))
    buf = b"".join(io.state_bytes(*r) for r in recs)
    path = str(tmp_path / "run.bin")
    with open(path, "wb") as f:
        f.write(buf + io.state_bytes(*recs[0])[:700])     # a run killed while writing a fifth record
    st = _rand_state(ibm, rng, u_sizes, p_size, shapes, -1.0)
    n = ibm.state_read(ctx, path, st, L, record=-1)
    assert n == 4
    t, u, p, xs = recs[3]
    assert st.t == t
    for a, b in zip(st.u, u):
        assert np.array_equal(a.cpu().numpy(), b)
    assert np.array_equal(st.p.cpu().numpy(), p)
    assert np.array_equal(st.xs[0].cpu().numpy(), xs[0].T.reshape(-1))
    assert ibm.state_read(ctx, path, st, L, record=1) == 4
    assert st.t == recs[1][0]
    # a record beyond the file, a wrong shape, a wrong object type
    with pytest.raises(ibm.IbmError) as e:
        ibm.state_read(ctx, path, st, L, record=4)
    assert e.value.code == ibm.ERR_SIZE
    bad = ibm.state_layout((41, 36, 40), p_size, shapes)
    with pytest.raises(ibm.IbmError) as e:
        ibm.state_read(ctx, path, st, bad, record=0)
    assert e.value.code == ibm.ERR_SIZE
    v = torch.zeros(40, dtype=torch.float64, device="cuda")
    with pytest.raises(ibm.IbmError) as e:
        ibm.io_read_dense(ctx, path, 8, v, True, 40, 1)     # a vector read as a matrix
    assert e.value.code == ibm.ERR_ARG


def test_dense_objects_round_trip(ibm, ctx, tmp_path):
    import torch
    io = _oracle()
    path = str(tmp_path / "dense.bin")
    a = torch.arange(12, dtype=torch.float64, device="cuda")
    ibm.io_write_dense(ctx, path, a, False, append=False)
    ibm.io_write_dense(ctx, path, a, True, rows=3, cols=4)
    assert ibm.io_peek_dense(ctx, path, 0) == (io.magic(False), 12, 1)
    off = ibm.io_read_dense(ctx, path, 0, torch.empty_like(a), False, 12)
    assert ibm.io_peek_dense(ctx, path, off) == (io.magic(True), 3, 4)
    b = torch.empty_like(a)
    assert ibm.io_read_dense(ctx, path, off, b, True, 3, 4) == os.path.getsize(path)
    assert torch.equal(a, b)
    m, _ = io.read_dense(open(path, "rb").read(), off, True)
    assert np.array_equal(m, np.arange(12.0).reshape(4, 3).T)
