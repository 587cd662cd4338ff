"""ctypes binding of the CPU oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY:
imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
as the checker; the product never routes through it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class orc_dim(ctypes.Structure):
    _fields_ = [("length", ctypes.c_double), ("solid", ctypes.c_int), ("reserved", ctypes.c_int),
                ("lower", ctypes.c_double * 2), ("upper", ctypes.c_double * 2)]


class orc_axis(ctypes.Structure):
    _fields_ = [("length", ctypes.c_double), ("resolution", ctypes.c_double),
                ("alignment", ctypes.c_double), ("shift", ctypes.c_double),
                ("cells", ctypes.c_int), ("points", ctypes.c_int), ("solid", ctypes.c_int),
                ("reserved", ctypes.c_int),
                ("boundary", ctypes.c_double * 2), ("interior", ctypes.c_double * 2),
                ("laplacian", ctypes.c_double * 2), ("robin", (ctypes.c_double * 2) * 2)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        D = ctypes.c_double
        L.orc_grid_axes.argtypes = [P(orc_dim), ctypes.c_int, ctypes.c_int, P(D), P(orc_axis)]
        L.orc_phi.restype = D
        L.orc_phi.argtypes = [ctypes.c_int, D]
        for name in ("orc_spread", "orc_spread_direct"):
            getattr(L, name).argtypes = [P(orc_axis), ctypes.c_int, ctypes.c_int, P(D), P(D), P(D), P(D), P(D)]
        for name in ("orc_interp", "orc_interp_direct"):
            getattr(L, name).argtypes = [P(orc_axis), ctypes.c_int, ctypes.c_int, P(D), P(D), P(D), P(D), P(D)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def dims(spec):
    """spec: list of (length, solid, lower(alpha,beta), upper(alpha,beta))"""
    arr = (orc_dim * len(spec))()
    for i, (length, solid, lower, upper) in enumerate(spec):
        arr[i].length = length
        arr[i].solid = solid
        arr[i].lower[0], arr[i].lower[1] = lower
        arr[i].upper[0], arr[i].upper[1] = upper
    return arr


def grid_axes(spec, refinement, alignment):
    d = dims(spec)
    out = (orc_axis * len(spec))()
    al = (ctypes.c_double * 3)(*alignment)
    lib().orc_grid_axes(d, len(spec), refinement, al, out)
    return out


def mac_axes(spec, refinement):
    """[3 components x 3 axes] array for fd::mac{refinement} (fd/cell.h:194-196)."""
    out = (orc_axis * 9)()
    for c in range(3):
        al = [0.0 if j == c else 0.5 for j in range(3)]
        ax = grid_axes(spec, refinement, al)
        for j in range(3):
            out[3 * c + j] = ax[j]
    return out


def sizes(axes):
    return [axes[3 * c].points * axes[3 * c + 1].points * axes[3 * c + 2].points for c in range(3)]


DELTA = {"cosine": 0, "peskin": 1, "bspline3": 2}


def spread(axes, delta, x, f, direct=False):
    x = np.ascontiguousarray(x, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    n = x.size // 3
    out = [np.zeros(s) for s in sizes(axes)]
    fn = lib().orc_spread_direct if direct else lib().orc_spread
    fn(axes, DELTA.get(delta, delta), n, _dp(x), _dp(f), _dp(out[0]), _dp(out[1]), _dp(out[2]))
    return out


def interpolate(axes, delta, x, ue, direct=False):
    x = np.ascontiguousarray(x, dtype=np.float64)
    ue = [np.ascontiguousarray(u, dtype=np.float64) for u in ue]
    n = x.size // 3
    out = np.zeros(3 * n)
    fn = lib().orc_interp_direct if direct else lib().orc_interp
    fn(axes, DELTA.get(delta, delta), n, _dp(x), _dp(ue[0]), _dp(ue[1]), _dp(ue[2]), _dp(out))
    return out


def phi(delta, r):
    return lib().orc_phi(DELTA.get(delta, delta), r)
