"""Oracle (test infrastructure only): numpy restatement of the reference's
linear-algebra binary format, used to check the library's checkpoint files.

linalg/io.h:67-99  magic = container | layout | value type, as one int32:
                   vector 0x65760000, matrix 0x616d0000; dense 0x6400,
                   sparse 0x7300; double 0x64, float 0x73
linalg/io.h:421-430 dense vector: magic, {rows, 1, rows}, values
linalg/io.h:445-454 dense matrix: magic, {rows, cols, rows*cols}, values
                   (column-major)
linalg/io.h:414-419 a scalar (a state's time) is its raw bytes
ins/state.h:48-68  ins::state record order t, u..., p; util/kaboose.cc:21-31
                   extends it with one matrix per cell type.
Host byte order is little-endian on every machine this runs on.
"""
import struct

import numpy as np

VECTOR = 0x65760000
MATRIX = 0x616D0000
DENSE = 0x00006400
SPARSE = 0x00007300
DOUBLE = 0x00000064
SINGLE = 0x00000073


def magic(is_matrix, sparse=False, single=False):
    return (MATRIX if is_matrix else VECTOR) | (SPARSE if sparse else DENSE) | (SINGLE if single else DOUBLE)


def dense_bytes(a, is_matrix):
    """Bytes of a dense double vector (1-D a) or matrix (2-D a, written
    column-major)."""
    a = np.asarray(a, dtype=np.float64)
    if is_matrix:
        rows, cols = a.shape
        vals = np.asfortranarray(a).ravel(order="F")
    else:
        rows, cols = a.size, 1
        vals = a.ravel()
    return struct.pack("<4i", magic(is_matrix), rows, cols, rows * cols) + vals.astype("<f8").tobytes()


def read_dense(buf, offset, is_matrix):
    """(array, next offset) of the dense double object at `offset`."""
    m, rows, cols, n = struct.unpack_from("<4i", buf, offset)
    if m != magic(is_matrix):
        raise ValueError(f"magic 0x{m:08x} is not a dense double {'matrix' if is_matrix else 'vector'}")
    if n != rows * cols or (not is_matrix and cols != 1):
        raise ValueError("inconsistent header")
    start = offset + 16
    vals = np.frombuffer(buf, dtype="<f8", count=n, offset=start).copy()
    if is_matrix:
        vals = vals.reshape((cols, rows)).T
    return vals, start + 8 * n


def state_bytes(t, u, p, xs):
    """One ib::state record: t, u[...], p, x[...]."""
    out = struct.pack("<d", t)
    for v in u:
        out += dense_bytes(v, False)
    out += dense_bytes(p, False)
    for x in xs:
        out += dense_bytes(x, True)
    return out


def read_states(buf, dims, ntypes):
    """Every complete record of a state file, as (t, [u], p, [x]) tuples;
    a truncated final record is dropped (as kaboose stops at end of file)."""
    out = []
    off = 0
    while off + 8 <= len(buf):
        try:
            (t,) = struct.unpack_from("<d", buf, off)
            o = off + 8
            u = []
            for _ in range(dims):
                v, o = read_dense(buf, o, False)
                u.append(v)
            p, o = read_dense(buf, o, False)
            xs = []
            for _ in range(ntypes):
                x, o = read_dense(buf, o, True)
                xs.append(x)
        except (struct.error, ValueError):
            break
        if o > len(buf):
            break
        out.append((t, u, p, xs))
        off = o
    return out
