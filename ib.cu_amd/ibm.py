"""Python host binding of libibm.so (ctypes).

Mirrors the reference's operator interface for the hot path so the tests and
the benchmark read like the reference's programs:

    grids = mac_grids(domain, refinement)              # fd::grid per component
    fe = spread(ctx, grids, delta, x, f)               # ib::spread (ib/novel.h:236-244)
    ul = interpolate(ctx, grids, delta, x, fe)         # ib::interpolate (ib/interpolate.h:124-131)

Device memory comes from torch tensors (plumbing only); every computation
runs in libibm.so.  Importing this module does not need a GPU; calling into
the library does.  There is no fallback: a missing library raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libibm.so")

# ----------------------------------------------------------------------
# Units (units.h:16-18,221-276): 1 length unit = 10 um, 1 mass unit = 1 ug,
# 1 time unit = 1 ms.
# ----------------------------------------------------------------------
um = 0.1
cm = 1000.0
m = 1e5
s = 1000.0
ms = 1.0
ns = 1e-6
g = 1e6
cP = 0.01
mL = cm ** 3

DELTA = {"cosine": 0, "peskin": 1, "bspline3": 2}


class ibm_dimension(ctypes.Structure):
    _fields_ = [("length", ctypes.c_double), ("solid", ctypes.c_int), ("reserved", ctypes.c_int),
                ("lower", ctypes.c_double * 2), ("upper", ctypes.c_double * 2)]


class ibm_axis(ctypes.Structure):
    _fields_ = [("length", ctypes.c_double), ("resolution", ctypes.c_double),
                ("alignment", ctypes.c_double), ("shift", ctypes.c_double),
                ("cells", ctypes.c_int), ("points", ctypes.c_int), ("solid", ctypes.c_int),
                ("reserved", ctypes.c_int),
                ("boundary", ctypes.c_double * 2), ("interior", ctypes.c_double * 2),
                ("laplacian", ctypes.c_double * 2), ("robin", (ctypes.c_double * 2) * 2)]


class ibm_grid(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int), ("refinement", ctypes.c_int), ("unit", ctypes.c_double),
                ("axis", ibm_axis * 3)]

    @property
    def points(self):
        return tuple(self.axis[i].points for i in range(self.dims))

    @property
    def size(self):
        n = 1
        for p in self.points:
            n *= p
        return n

    @property
    def resolution(self):
        return self.axis[0].resolution


class IbmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ibm status {code}: {msg}")
        self.code = code


OK, ERR_HIP, ERR_ARG, ERR_SIZE, ERR_NOT_CONVERGED, ERR_DT_TOO_SMALL, ERR_BAD_GRID = range(7)

_lib = None

# (name, restype, argtypes) of every exported C entry point used from Python.
_VP = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_PD = ctypes.POINTER(ctypes.c_double)
ABI_VERSION = 4          # IBM_ABI_VERSION of include/ibm.h this binding matches

_SIGS = {
    "ibm_abi_version": (_I, []),
    "ibm_domain_unit": (_I, [ctypes.POINTER(ibm_dimension), _I, _PD]),
    "ibm_grid_make": (_I, [ctypes.POINTER(ibm_dimension), _I, _I, _PD, ctypes.POINTER(ibm_grid)]),
    "ibm_mac_grids": (_I, [ctypes.POINTER(ibm_dimension), _I, _I, ctypes.POINTER(ibm_grid)]),
    "ibm_centered_grid": (_I, [ctypes.POINTER(ibm_dimension), _I, _I, ctypes.POINTER(ibm_grid)]),
    "ibm_grid_coarsen": (_I, [ctypes.POINTER(ibm_grid), ctypes.POINTER(ibm_grid)]),
    "ibm_grid_size": (ctypes.c_int64, [ctypes.POINTER(ibm_grid)]),
    "ibm_create": (_I, [ctypes.POINTER(_VP), _I]),
    "ibm_destroy": (_I, [_VP]),
    "ibm_set_stream": (_I, [_VP, _VP]),
    "ibm_synchronize": (_I, [_VP]),
    "ibm_last_error": (ctypes.c_char_p, [_VP]),
    "ibm_profile_enable": (_I, [_VP, _I]),
    "ibm_profile_read": (_I, [_VP, ctypes.c_char_p, _PD, ctypes.POINTER(_I)]),
    "ibm_profile_work": (_I, [_VP, ctypes.c_char_p, _PD]),
    "ibm_spread": (_I, [_VP, ctypes.POINTER(ibm_grid), _I, _I, _VP, _VP, ctypes.POINTER(_VP)]),
    "ibm_spread_sets": (_I, [_VP, ctypes.POINTER(ibm_grid), _I, _I, ctypes.POINTER(_I),
                             ctypes.POINTER(_VP), ctypes.POINTER(_VP), ctypes.POINTER(_VP)]),
    "ibm_spread_sets_noalloc": (_I, [_VP, ctypes.POINTER(ibm_grid), _I, _I, ctypes.POINTER(_I),
                                     ctypes.POINTER(_VP), ctypes.POINTER(_VP), ctypes.POINTER(_VP)]),
    "ibm_spread_boxes": (_I, [_VP, ctypes.POINTER(ibm_grid), _I, _I, ctypes.POINTER(_I), _VP, _VP,
                              ctypes.POINTER(_VP)]),
    "ibm_interpolate_boxes": (_I, [_VP, ctypes.POINTER(ibm_grid), _I, _I, ctypes.POINTER(_I), _VP,
                                   ctypes.POINTER(_VP), _VP]),
    "ibm_interpolate": (_I, [_VP, ctypes.POINTER(ibm_grid), _I, _I, _VP, ctypes.POINTER(_VP), _VP]),
    "ibm_interpolate_sets": (_I, [_VP, ctypes.POINTER(ibm_grid), _I, _I, ctypes.POINTER(_I),
                                  ctypes.POINTER(_VP), ctypes.POINTER(_VP), ctypes.POINTER(_VP)]),
    "ibm_op_apply": (_I, [_VP, ctypes.POINTER(ibm_grid), _D, _D, _D, _VP, _D, _VP]),
    "ibm_op_apply_f32": (_I, [_VP, ctypes.POINTER(ibm_grid), _D, _D, ctypes.c_float, _VP, ctypes.c_float, _VP]),
    "ibm_gershgorin": (_I, [ctypes.POINTER(ibm_grid), _D, _D, _PD, _PD]),
    "ibm_chebyshev_weights": (_I, [_D, _D, _PD]),
    "ibm_chebyshev_apply": (_I, [_VP, ctypes.POINTER(ibm_grid), _D, _D, _D, _D, _VP, _VP, _VP]),
    "ibm_dot": (_I, [_VP, ctypes.c_int64, _VP, _VP, _PD]),
    "ibm_mg_create": (_I, [_VP, ctypes.POINTER(ibm_grid), _D, _D, _I, ctypes.POINTER(_VP)]),
    "ibm_mg_destroy": (_I, [_VP]),
    "ibm_mg_levels": (_I, [_VP, ctypes.POINTER(_I), ctypes.POINTER(ibm_grid), _I]),
    "ibm_mg_vcycle": (_I, [_VP, _VP, _VP, _VP]),
    "ibm_fft_create": (_I, [_VP, ctypes.POINTER(ibm_grid), ctypes.POINTER(_VP)]),
    "ibm_fft_destroy": (_I, [_VP]),
    "ibm_fft_solve": (_I, [_VP, _VP, _D, _D, _VP, _VP]),
    "ibm_mg_autotune": (_I, [_I, _I]),
    "ibm_mg_shape": (_I, [_VP, _I, _I, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "ibm_mg_solve": (_I, [_VP, _VP, _VP, _VP, _D, _I, ctypes.POINTER(_I), _PD]),
    "ibm_chebpcg_solve": (_I, [_VP, ctypes.POINTER(ibm_grid), _D, _D, _VP, _VP, _D, _I, _VP,
                               ctypes.POINTER(_I), _PD]),
    "ibm_grid_shift": (_I, [ctypes.POINTER(ibm_grid), _I, ctypes.POINTER(ibm_grid)]),
    "ibm_differential": (_I, [_VP, ctypes.POINTER(ibm_grid), _I, _D, _VP, _D, _VP]),
    "ibm_average": (_I, [_VP, ctypes.POINTER(ibm_grid), _I, _D, _VP, _D, _VP]),
    "ibm_divergence": (_I, [_VP, ctypes.POINTER(ibm_grid), ctypes.POINTER(_VP), _VP]),
    "ibm_gradient": (_I, [_VP, ctypes.POINTER(ibm_grid), _D, _VP, _D, ctypes.POINTER(_VP)]),
    "ibm_advection": (_I, [_VP, ctypes.POINTER(ibm_grid), ctypes.POINTER(_VP), ctypes.POINTER(_VP)]),
    "ibm_rbf_create": (_I, [_VP, _I, _I, _I, _I, ctypes.POINTER(_VP), _I, ctypes.POINTER(_VP)]),
    "ibm_rbf_destroy": (_I, [_VP]),
    "ibm_rbf_gemm_variant": (_I, [_I]),
    "ibm_rbf_eval": (_I, [_VP, _I, ctypes.POINTER(_VP), ctypes.POINTER(_I), ctypes.POINTER(_VP),
                          ctypes.POINTER(_VP)]),
    "ibm_rbf_build": (_I, [_VP, _I, _VP, _I, _VP, _I, _I, ctypes.POINTER(_I), ctypes.POINTER(_VP)]),
    "ibm_sphere_weights": (_I, [_VP, _I, _VP, _VP]),
    "ibm_surface_geometry": (_I, [_VP, _I, _I, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "ibm_axpby": (_I, [_VP, ctypes.c_int64, _D, _VP, _D, _VP]),
    "ibm_boundary_apply": (_I, [_VP, ctypes.POINTER(ibm_grid), _I, _I, _I, _D, _VP, _D, _VP]),
    "ibm_mg_solve_fixed": (_I, [_VP, _VP, _VP, _VP, _I, _I, _VP]),
    "ibm_ins_boundary_apply": (_I, [_VP, ctypes.POINTER(ibm_grid), _I, _I, _D, _VP, _D, _VP]),
    "ibm_boxes_mg_create": (_I, [_VP, ctypes.POINTER(ibm_grid), _I, _D, _D, ctypes.POINTER(_VP)]),
    "ibm_boxes_mg_destroy": (_I, [_VP]),
    "ibm_boxes_mg_solve": (_I, [_VP, _VP, _VP, _VP, _D, _I, ctypes.POINTER(_I), _PD]),
    "ibm_mg_solve_atol": (_I, [_VP, _VP, _VP, _VP, _D, _D, _I, _I, ctypes.POINTER(_I), _PD]),
    "ibm_solver_create": (_I, [_VP, ctypes.POINTER(ibm_dimension), _I, _I, _VP, _I, _VP, ctypes.POINTER(_VP)]),
    "ibm_solver_destroy": (_I, [_VP]),
    "ibm_solver_step": (_I, [_VP, _VP, _VP, _VP]),
    "ibm_solver_step_fixed": (_I, [_VP, _VP, _VP, _VP]),
    "ibm_solver_capture": (_I, [_VP, _VP, _VP, _VP]),
    "ibm_solver_replay": (_I, [_VP, _VP, _VP]),
    "ibm_solver_info": (_I, [_VP, _VP, _VP]),
    "ibm_solver_cells": (_I, [_VP, _I, _VP]),
    "ibm_solver_grids": (_I, [_VP, ctypes.POINTER(ibm_grid), ctypes.POINTER(ibm_grid)]),
    "ibm_solver_layout": (_I, [_VP, _VP]),
    "ibm_membrane_reference": (_I, [_VP, _I, _VP, _VP, _VP]),
    "ibm_mean_curvature": (_I, [_VP, _I, _I, _VP, _VP, _VP]),
    "ibm_membrane_forces": (_I, [_VP, _VP, _I, _I, _VP, _VP, _VP, _VP, _VP, _D, _VP]),
    "ibm_io_write_dense": (_I, [_VP, ctypes.c_char_p, _I, _I, ctypes.c_longlong, ctypes.c_longlong, _VP]),
    "ibm_io_read_dense": (_I, [_VP, ctypes.c_char_p, ctypes.c_longlong, _I, ctypes.c_longlong, ctypes.c_longlong,
                               _VP, ctypes.POINTER(ctypes.c_longlong)]),
    "ibm_io_peek_dense": (_I, [_VP, ctypes.c_char_p, ctypes.c_longlong, ctypes.POINTER(_I), ctypes.POINTER(_I),
                               ctypes.POINTER(_I)]),
    "ibm_state_write": (_I, [_VP, ctypes.c_char_p, _I, _VP, _VP]),
    "ibm_state_read": (_I, [_VP, ctypes.c_char_p, _I, _VP, _VP, ctypes.POINTER(_I)]),
}


def lib():
    """Loads libibm.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        if L.ibm_abi_version() != ABI_VERSION:
            raise RuntimeError(f"{LIB_PATH} has ABI version {L.ibm_abi_version()}, this binding expects "
                               f"{ABI_VERSION}: rebuild with __graft_entry__.build()")
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def _check(code, ctx=None):
    if code != OK:
        if ctx is None:
            raise IbmError(code, "grid construction failed")
        msg = lib().ibm_last_error(ctx.handle)
        raise IbmError(code, msg.decode() if msg else "")


# ----------------------------------------------------------------------
# Domains and grids (fd::dimension, fd::domain, fd::grid)
# ----------------------------------------------------------------------
def periodic(length):
    """fd::dimension{length, fd::boundary::periodic}"""
    d = ibm_dimension()
    d.length = length
    d.solid = 0
    return d


def robin(length, lower=(1.0, 0.0), upper=None):
    """fd::dimension{length, robin lower[, upper]}; default Dirichlet (1, 0)."""
    upper = lower if upper is None else upper
    d = ibm_dimension()
    d.length = length
    d.solid = 1
    d.lower[0], d.lower[1] = lower
    d.upper[0], d.upper[1] = upper
    return d


def dirichlet(length):
    return robin(length, (1.0, 0.0))


def neumann(length):
    return robin(length, (0.0, 1.0))


def _dims(domain):
    arr = (ibm_dimension * len(domain))(*domain)
    return arr, len(domain)


def domain_unit(domain):
    arr, n = _dims(domain)
    out = ctypes.c_double()
    _check(lib().ibm_domain_unit(arr, n, ctypes.byref(out)))
    return out.value


def grid_make(domain, refinement, alignment):
    arr, n = _dims(domain)
    al = (ctypes.c_double * 3)(*alignment)
    out = ibm_grid()
    _check(lib().ibm_grid_make(arr, n, refinement, al, ctypes.byref(out)))
    return out


def mac_grids(domain, refinement):
    arr, n = _dims(domain)
    out = (ibm_grid * n)()
    _check(lib().ibm_mac_grids(arr, n, refinement, out))
    return out


def centered_grid(domain, refinement):
    arr, n = _dims(domain)
    out = ibm_grid()
    _check(lib().ibm_centered_grid(arr, n, refinement, ctypes.byref(out)))
    return out


def grid_coarsen(grid):
    out = ibm_grid()
    _check(lib().ibm_grid_coarsen(ctypes.byref(grid), ctypes.byref(out)))
    return out


# ----------------------------------------------------------------------
# Context
# ----------------------------------------------------------------------
class Context:
    """Owns an ibm_ctx bound to a device; work is enqueued on `stream`
    (a torch.cuda.Stream or raw hipStream_t int; default torch's current)."""

    def __init__(self, device=0, stream=None):
        h = ctypes.c_void_p()
        code = lib().ibm_create(ctypes.byref(h), device)
        if code != OK:
            raise IbmError(code, "ibm_create failed")
        self.handle = h
        self.set_stream(stream)

    def set_stream(self, stream=None):
        if stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        elif hasattr(stream, "cuda_stream"):
            stream = stream.cuda_stream
        _check(lib().ibm_set_stream(self.handle, ctypes.c_void_p(stream)), self)

    def synchronize(self):
        _check(lib().ibm_synchronize(self.handle), self)

    def profile(self, on=True):
        _check(lib().ibm_profile_enable(self.handle, 1 if on else 0), self)

    def profile_read(self, name):
        """(total milliseconds, launches) recorded for a kernel name."""
        t = ctypes.c_double()
        k = ctypes.c_int()
        _check(lib().ibm_profile_read(self.handle, name.encode(), ctypes.byref(t), ctypes.byref(k)), self)
        return t.value, k.value

    def profile_work(self, name):
        """Summed algorithmic bytes (flops for rbf_gemm) of a kernel name's recorded launches."""
        w = ctypes.c_double()
        _check(lib().ibm_profile_work(self.handle, name.encode(), ctypes.byref(w)), self)
        return w.value

    def close(self):
        if self.handle:
            lib().ibm_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _ptrs(ts):
    return (ctypes.c_void_p * len(ts))(*[t.data_ptr() for t in ts])


def _check_markers(x, n):
    import torch
    if x.dtype != torch.float64 or not x.is_contiguous() or x.numel() != 3 * n:
        raise ValueError("marker arrays must be contiguous float64 with 3*n entries")


# ----------------------------------------------------------------------
# ib::spread / ib::interpolate
# ----------------------------------------------------------------------
def spread(ctx, grids, delta, x, f, out=None):
    """Spreads SoA forces f (3*n) at SoA positions x (3*n) onto the three
    component grids; returns (or fills) three float64 tensors."""
    import torch
    n = x.numel() // 3
    _check_markers(x, n)
    _check_markers(f, n)
    if out is None:
        out = [torch.empty(grids[c].size, dtype=torch.float64, device=x.device) for c in range(3)]
    code = lib().ibm_spread(ctx.handle, grids, DELTA.get(delta, delta), n, _ptr(x), _ptr(f), _ptrs(out))
    _check(code, ctx)
    return out


def spread_sets(ctx, grids, delta, xs, fs, out=None):
    import torch
    ns = [x.numel() // 3 for x in xs]
    for x, f, n in zip(xs, fs, ns):
        _check_markers(x, n)
        _check_markers(f, n)
    if out is None:
        out = [torch.empty(grids[c].size, dtype=torch.float64, device=xs[0].device) for c in range(3)]
    code = lib().ibm_spread_sets(ctx.handle, grids, DELTA.get(delta, delta), len(xs),
                                 (ctypes.c_int * len(ns))(*ns), _ptrs(xs), _ptrs(fs), _ptrs(out))
    _check(code, ctx)
    return out


def spread_sets_noalloc(ctx, grids, delta, xs, fs, out):
    """ibm_spread_sets_noalloc: spread_sets into caller-given outputs with the
    library's fixed per-device workspace (no device allocation)."""
    ns = [x.numel() // 3 for x in xs]
    for x, f, n in zip(xs, fs, ns):
        _check_markers(x, n)
        _check_markers(f, n)
    code = lib().ibm_spread_sets_noalloc(ctx.handle, grids, DELTA.get(delta, delta), len(xs),
                                         (ctypes.c_int * len(ns))(*ns), _ptrs(xs), _ptrs(fs), _ptrs(out))
    _check(code, ctx)
    return out


def spread_boxes(ctx, grids, delta, offsets, x, f, out=None):
    """ibm_spread_boxes: markers of box b are x[offsets[b]:offsets[b+1]]
    (SoA over all markers); returns three tensors of nboxes * size(grid_c)."""
    import torch
    nb = len(offsets) - 1
    n = int(offsets[-1])
    _check_markers(x, n)
    _check_markers(f, n)
    if out is None:
        out = [torch.empty(nb * grids[c].size, dtype=torch.float64, device=x.device) for c in range(3)]
    off = (ctypes.c_int * (nb + 1))(*[int(o) for o in offsets])
    _check(lib().ibm_spread_boxes(ctx.handle, grids, DELTA.get(delta, delta), nb, off, _ptr(x), _ptr(f),
                                  _ptrs(out)), ctx)
    return out


def interpolate_boxes(ctx, grids, delta, offsets, x, ue, out=None):
    import torch
    nb = len(offsets) - 1
    n = int(offsets[-1])
    _check_markers(x, n)
    if out is None:
        out = torch.empty(3 * n, dtype=torch.float64, device=x.device)
    off = (ctypes.c_int * (nb + 1))(*[int(o) for o in offsets])
    _check(lib().ibm_interpolate_boxes(ctx.handle, grids, DELTA.get(delta, delta), nb, off, _ptr(x), _ptrs(ue),
                                       _ptr(out)), ctx)
    return out


def interpolate(ctx, grids, delta, x, ue, out=None):
    import torch
    n = x.numel() // 3
    _check_markers(x, n)
    if out is None:
        out = torch.empty(3 * n, dtype=torch.float64, device=x.device)
    code = lib().ibm_interpolate(ctx.handle, grids, DELTA.get(delta, delta), n, _ptr(x), _ptrs(ue), _ptr(out))
    _check(code, ctx)
    return out


def interpolate_sets(ctx, grids, delta, xs, ue, outs=None):
    import torch
    ns = [x.numel() // 3 for x in xs]
    for x, n in zip(xs, ns):
        _check_markers(x, n)
    if outs is None:
        outs = [torch.empty(3 * n, dtype=torch.float64, device=xs[0].device) for n in ns]
    code = lib().ibm_interpolate_sets(ctx.handle, grids, DELTA.get(delta, delta), len(xs),
                                      (ctypes.c_int * len(ns))(*ns), _ptrs(xs), _ptrs(ue), _ptrs(outs))
    _check(code, ctx)
    return outs


# ----------------------------------------------------------------------
# Operators and solvers (fd::laplacian, fd::identity, ins::diffusion's
# Helmholtz operator, algo::chebyshev, solvers::multigrid, solvers::chebpcg)
# ----------------------------------------------------------------------
LAPLACIAN = (0.0, 1.0)
IDENTITY = (1.0, 0.0)


def helmholtz(lam):
    """I2 - lambda L (ins/diffusion.h:21-35)."""
    return (1.0, -lam)


def op_apply(ctx, grid, op, x, alpha=1.0, beta=0.0, y=None):
    import torch
    if y is None:
        y = torch.zeros_like(x)
    sigma, tau = op
    fn = lib().ibm_op_apply if x.dtype == torch.float64 else lib().ibm_op_apply_f32
    _check(fn(ctx.handle, ctypes.byref(grid), sigma, tau, alpha, _ptr(x), beta, _ptr(y)), ctx)
    return y


def gershgorin(grid, op):
    lo, hi = ctypes.c_double(), ctypes.c_double()
    _check(lib().ibm_gershgorin(ctypes.byref(grid), op[0], op[1], ctypes.byref(lo), ctypes.byref(hi)))
    return lo.value, hi.value


def chebyshev_weights(a, b):
    w = (ctypes.c_double * 4)()
    _check(lib().ibm_chebyshev_weights(a, b, w))
    return list(w)


def chebyshev_apply(ctx, grid, op, a, b, r):
    import torch
    out = torch.empty_like(r)
    work = torch.empty(2 * r.numel(), dtype=torch.float64, device=r.device)
    _check(lib().ibm_chebyshev_apply(ctx.handle, ctypes.byref(grid), op[0], op[1], a, b, _ptr(r), _ptr(out),
                                     _ptr(work)), ctx)
    return out


def dot(ctx, x, y):
    """linalg::dot (linalg/blas.h:140-155): FP64, fixed summation order."""
    out = ctypes.c_double()
    _check(lib().ibm_dot(ctx.handle, x.numel(), _ptr(x), _ptr(y), ctypes.byref(out)), ctx)
    return out.value


def axpby(ctx, a, x, b, y):
    """y = a x + b y in place (linalg::axpy / scal, linalg/blas.h:63-98)."""
    _check(lib().ibm_axpby(ctx.handle, y.numel(), a, _ptr(x), b, _ptr(y)), ctx)
    return y


class FftSolver:
    """Direct solver of sigma I~ + tau L on channel grids (ibm_fft_*)."""

    def __init__(self, ctx, grid):
        self.ctx = ctx
        h = ctypes.c_void_p()
        _check(lib().ibm_fft_create(ctx.handle, ctypes.byref(grid), ctypes.byref(h)), ctx)
        self.handle = h

    def solve(self, b, op, x=None):
        import torch
        if x is None:
            x = torch.empty_like(b)
        _check(lib().ibm_fft_solve(self.ctx.handle, self.handle, op[0], op[1], _ptr(b), _ptr(x)), self.ctx)
        return x

    def close(self):
        if self.handle:
            lib().ibm_fft_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mg_autotune(enable=True, clear=False):
    """Switches the multigrid block-target autotuning (ibm_mg_autotune);
    clear=True forgets the targets recorded so far."""
    _check(lib().ibm_mg_autotune(1 if enable else 0, 1 if clear else 0))


class Multigrid:
    """solvers::multigrid(grid, tol, op, smoother) over A = sigma*I2 + tau*L."""

    def __init__(self, ctx, grid, op, precision=64):
        self.ctx = ctx
        h = ctypes.c_void_p()
        _check(lib().ibm_mg_create(ctx.handle, ctypes.byref(grid), op[0], op[1], precision, ctypes.byref(h)), ctx)
        self.handle = h

    def levels(self):
        n = ctypes.c_int()
        grids = (ibm_grid * 32)()
        _check(lib().ibm_mg_levels(self.handle, ctypes.byref(n), grids, 32))
        return [grids[i] for i in range(n.value)]

    def shapes(self):
        """Per level, the launch shape (block target, wide tile) of the
        smoothing, residual + restriction and prolongation + residual kernels
        (target 0: none recorded, the built-in target is used)."""
        out = []
        for level in range(len(self.levels())):
            row = []
            for kind in range(3):
                t, w = ctypes.c_int(), ctypes.c_int()
                _check(lib().ibm_mg_shape(self.handle, level, kind, ctypes.byref(t), ctypes.byref(w)))
                row.append((t.value, w.value))
            out.append(row)
        return out

    def vcycle(self, b):
        import torch
        x = torch.empty_like(b)
        _check(lib().ibm_mg_vcycle(self.ctx.handle, self.handle, _ptr(b), _ptr(x)), self.ctx)
        return x

    def solve(self, b, rtol=1e-10, maxit=100, x=None, atol=0.0, warm=False):
        """warm=True starts from the given x."""
        import torch
        if x is None:
            x = torch.zeros_like(b)
        it = ctypes.c_int()
        rr = ctypes.c_double()
        if atol > 0 or warm:
            _check(lib().ibm_mg_solve_atol(self.ctx.handle, self.handle, _ptr(b), _ptr(x), rtol, atol, maxit,
                                           1 if warm else 0, ctypes.byref(it), ctypes.byref(rr)), self.ctx)
        else:
            _check(lib().ibm_mg_solve(self.ctx.handle, self.handle, _ptr(b), _ptr(x), rtol, maxit,
                                      ctypes.byref(it), ctypes.byref(rr)), self.ctx)
        return x, it.value, rr.value

    def solve_fixed(self, b, iters, x=None, warm=False):
        """Exactly `iters` cycles, no host synchronisation; returns (x, relres
        as a one-element device tensor)."""
        import torch
        if x is None:
            x = torch.empty_like(b)
        rr = torch.zeros(1, dtype=torch.float64, device=b.device)
        _check(lib().ibm_mg_solve_fixed(self.ctx.handle, self.handle, _ptr(b), _ptr(x), iters, 1 if warm else 0,
                                        _ptr(rr)), self.ctx)
        return x, rr

    def close(self):
        if self.handle:
            lib().ibm_mg_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BoxesMultigrid:
    """ibm_boxes_mg: Helmholtz solves on nbatch copies of one periodic grid."""

    def __init__(self, ctx, grid, nbatch, op):
        sigma, tau = op
        self.ctx, self.grid, self.nbatch = ctx, grid, nbatch
        h = _VP()
        _check(lib().ibm_boxes_mg_create(ctx.handle, ctypes.byref(grid), nbatch, sigma, tau, ctypes.byref(h)), ctx)
        self.handle = h

    def solve(self, b, x=None, rtol=1e-10, maxit=60):
        import torch
        if x is None:
            x = torch.zeros_like(b)
        it, rel = _I(0), _D(0)
        _check(lib().ibm_boxes_mg_solve(self.ctx.handle, self.handle, _ptr(b), _ptr(x), rtol, maxit,
                                        ctypes.byref(it), ctypes.byref(rel)), self.ctx)
        return x, it.value, rel.value

    def close(self):
        if self.handle:
            lib().ibm_boxes_mg_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def chebpcg_solve(ctx, grid, op, b, rtol=1e-10, maxit=1000):
    import torch
    x = torch.empty_like(b)
    work = torch.empty(6 * b.numel(), dtype=torch.float64, device=b.device)
    it = ctypes.c_int()
    rr = ctypes.c_double()
    _check(lib().ibm_chebpcg_solve(ctx.handle, ctypes.byref(grid), op[0], op[1], _ptr(b), _ptr(x), rtol, maxit,
                                   _ptr(work), ctypes.byref(it), ctypes.byref(rr)), ctx)
    return x, it.value, rr.value


# ----------------------------------------------------------------------
# ins::divergence / ins::gradient / ins::advection, fd::differential / average
# ----------------------------------------------------------------------
def grid_shift(grid, view):
    out = ibm_grid()
    _check(lib().ibm_grid_shift(ctypes.byref(grid), view, ctypes.byref(out)))
    return out


def differential(ctx, grid, view, x, alpha=1.0, beta=0.0, y=None):
    import torch
    if y is None:
        y = torch.zeros(grid_shift(grid, view).size, dtype=torch.float64, device=x.device)
    _check(lib().ibm_differential(ctx.handle, ctypes.byref(grid), view, alpha, _ptr(x), beta, _ptr(y)), ctx)
    return y


def average(ctx, grid, view, x, alpha=1.0, beta=0.0, y=None):
    import torch
    if y is None:
        y = torch.zeros(grid_shift(grid, view).size, dtype=torch.float64, device=x.device)
    _check(lib().ibm_average(ctx.handle, ctypes.byref(grid), view, alpha, _ptr(x), beta, _ptr(y)), ctx)
    return y


def divergence(ctx, grids, u, out=None):
    import torch
    if out is None:
        out = torch.empty(grid_shift(grids[0], 0).size, dtype=torch.float64, device=u[0].device)
    _check(lib().ibm_divergence(ctx.handle, grids, _ptrs(u), _ptr(out)), ctx)
    return out


def gradient(ctx, pgrid, p, out=None, alpha=1.0, beta=0.0):
    import torch
    if out is None:
        out = [torch.zeros(grid_shift(pgrid, i).size, dtype=torch.float64, device=p.device) for i in range(3)]
    _check(lib().ibm_gradient(ctx.handle, ctypes.byref(pgrid), alpha, _ptr(p), beta, _ptrs(out)), ctx)
    return out


def advection(ctx, grids, u, out=None):
    import torch
    if out is None:
        out = [torch.empty(grids[i].size, dtype=torch.float64, device=u[0].device) for i in range(3)]
    _check(lib().ibm_advection(ctx.handle, grids, _ptrs(u), _ptrs(out)), ctx)
    return out


# ----------------------------------------------------------------------
# RBF surface operators
# ----------------------------------------------------------------------
RBF_PRECISION = {"f64": 64, "f32": 32, "bf16x3": 16, 64: 64, 32: 32, 16: 16}


class RbfOperator:
    """Packed operators of one cell type (ibm_rbf_create).  `mats` are
    float64 CUDA tensors of shape (M, lde) (sample-major: the reference's
    lde x M column-major matrix), of which the first K columns are used."""

    def __init__(self, ctx, precision, mats, K):
        import torch
        mats = list(mats)
        M, lde = mats[0].shape
        for m in mats:
            if m.shape != (M, lde) or not m.is_contiguous() or m.dtype != torch.float64:
                raise ValueError("operators must be contiguous float64 (M, lde) tensors of one shape")
        self.M, self.K, self.nops = M, K, len(mats)
        self.precision = RBF_PRECISION.get(precision, precision)
        h = _VP()
        _check(lib().ibm_rbf_create(ctx.handle, self.precision, M, K, len(mats), _ptrs(mats), lde,
                                    ctypes.byref(h)), ctx)
        self.handle = h

    def close(self):
        if self.handle:
            lib().ibm_rbf_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rbf_gemm_variant(waves_n=0):
    """FP64 evaluation kernel: 2 or 4 waves along the columns of a tile, or
    0 to time both per shape (ibm_rbf_gemm_variant)."""
    _check(lib().ibm_rbf_gemm_variant(int(waves_n)))


def rbf_eval(ctx, ops, xs, ys=None):
    """Y_b = [E_j^T X_b]_j for each (operator set, X) pair in one launch.
    X_b: float64 CUDA tensor holding K x N column-major (any shape with
    K*N elements, e.g. (N, K) row-major).  Returns tensors (nops, N, M)."""
    import torch
    ns = []
    for op, x in zip(ops, xs):
        if x.numel() % op.K or not x.is_contiguous() or x.dtype != torch.float64:
            raise ValueError("X must be contiguous float64 with a multiple of K elements")
        ns.append(x.numel() // op.K)
    if ys is None:
        ys = [torch.empty((op.nops, n, op.M), dtype=torch.float64, device=xs[0].device) for op, n in zip(ops, ns)]
    handles = (_VP * len(ops))(*[op.handle.value for op in ops])
    _check(lib().ibm_rbf_eval(ctx.handle, len(ops), handles, (_I * len(ns))(*ns), _ptrs(xs), _ptrs(ys)), ctx)
    return ys


RBF_KINDS = {"value": 0, "t": 1, "p": 2, "tt": 3, "tp": 4, "pp": 5}


def rbf_build(ctx, data, sample, degree, kinds=("value", "t", "p")):
    """ibm_rbf_build: operators (ns, nd + np) (sample-major, the reference's
    (nd + np) x ns column-major layout) built on the device from (theta, phi)
    parameter tensors data (2, nd) and sample (2, ns)."""
    import torch
    nd, ns = data.shape[1], sample.shape[1]
    npoly = (degree + 1) ** 2
    outs = [torch.empty((ns, nd + npoly), dtype=torch.float64, device=data.device) for _ in kinds]
    kk = (_I * len(kinds))(*[RBF_KINDS.get(k, k) for k in kinds])
    _check(lib().ibm_rbf_build(ctx.handle, nd, _ptr(data.contiguous()), ns, _ptr(sample.contiguous()), degree,
                               len(kinds), kk, _ptrs(outs)), ctx)
    return outs


def sphere_weights(ctx, params):
    import torch
    n = params.shape[1]
    w = torch.empty(n, dtype=torch.float64, device=params.device)
    _check(lib().ibm_sphere_weights(ctx.handle, n, _ptr(params.contiguous()), _ptr(w)), ctx)
    return w


def surface_geometry(ctx, ns, nc, y, t1, t2, w):
    """Normals (3, nc, ns), sigma (nc, ns), area (nc), volume (nc) in FP64."""
    import torch
    dev = y.device
    normal = torch.empty((3, nc, ns), dtype=torch.float64, device=dev)
    sigma = torch.empty((nc, ns), dtype=torch.float64, device=dev)
    area = torch.empty(nc, dtype=torch.float64, device=dev)
    volume = torch.empty(nc, dtype=torch.float64, device=dev)
    _check(lib().ibm_surface_geometry(ctx.handle, ns, nc, _ptr(y), _ptr(t1), _ptr(t2), _ptr(w), _ptr(normal),
                                      _ptr(sigma), _ptr(area), _ptr(volume)), ctx)
    return normal, sigma, area, volume


def ins_boundary_apply(ctx, grid, dim, x, y, alpha=1.0, beta=1.0, order=0):
    """y = alpha * ins::boundary(grid, dim) x + beta * y (order 2: scaled by
    resolution^2 per axis)."""
    _check(lib().ibm_ins_boundary_apply(ctx.handle, ctypes.byref(grid), dim, order, alpha, _ptr(x), beta,
                                        _ptr(y)), ctx)
    return y


def boundary_apply(ctx, grid, view, side, y, value=0.0, values=None, alpha=1.0, order=2):
    """y += alpha * fd::{lower,upper}_boundary(grid, view, order) g (side 0 / 1)."""
    _check(lib().ibm_boundary_apply(ctx.handle, ctypes.byref(grid), view, side, order, alpha,
                                    _ptr(values) if values is not None else None, value, _ptr(y)), ctx)
    return y


# ----------------------------------------------------------------------
# Immersed-boundary timestep
# ----------------------------------------------------------------------
MAX_CELL_TYPES = 8
PSEUDO_PRESSURE = 1       # IBM_SOLVER_PSEUDO_PRESSURE
MULTIGRID = 2             # IBM_SOLVER_MULTIGRID


class ibm_solver_params(ctypes.Structure):
    _fields_ = [("timestep", _D), ("diffusivity", _D), ("density", _D), ("tolerance", _D),
                ("delta", _I), ("precision", _I), ("max_iterations", _I), ("flags", _I), ("fscale", _D)]


TENSION = {"none": 0, "skalak": 1, "neohookean": 2}


class ibm_membrane(ctypes.Structure):
    _fields_ = [("tension", _I), ("preferred", _I), ("shear", _D), ("bulk", _D), ("bending", _D)]


def membrane(tension="none", shear=0.0, bulk=0.0, bending=0.0, preferred=False):
    """ibm_membrane: a tension law ("skalak", "neohookean") with its moduli
    and a bending modulus (forces/skalak.h, neohookean.h, bending.h)."""
    return ibm_membrane(TENSION.get(tension, tension), int(bool(preferred)), shear, bulk, bending)


class ibm_cell_type(ctypes.Structure):
    _fields_ = [("ops", _VP), ("data_sites", _I), ("sample_sites", _I), ("cells", _I), ("reserved", _I),
                ("weights", _VP), ("tether", _VP), ("stiffness", _D), ("membrane", ibm_membrane),
                ("data_ops", _VP), ("reference", _VP)]


class ibm_state(ctypes.Structure):
    _fields_ = [("t", _D), ("u", _VP * 3), ("p", _VP), ("x", _VP * MAX_CELL_TYPES)]


class ibm_state_layout(ctypes.Structure):
    _fields_ = [("dims", _I), ("ntypes", _I), ("u_size", ctypes.c_longlong * 3), ("p_size", ctypes.c_longlong),
                ("rows", _I * MAX_CELL_TYPES), ("cols", _I * MAX_CELL_TYPES)]


def state_layout(u_sizes, p_size, shapes):
    """ibm_state_layout of velocity / pressure sizes and per-type (rows, cols)."""
    L = ibm_state_layout()
    L.dims = len(u_sizes)
    L.ntypes = len(shapes)
    for i, n in enumerate(u_sizes):
        L.u_size[i] = n
    L.p_size = p_size
    for k, (r, c) in enumerate(shapes):
        L.rows[k] = r
        L.cols[k] = c
    return L


def membrane_reference(ctx, ns, geometry, weights, out=None):
    """ibm_membrane_reference: the 14 x ns table of an undeformed shape."""
    import torch
    if out is None:
        out = torch.empty(14 * ns, dtype=torch.float64, device=weights.device)
    _check(lib().ibm_membrane_reference(ctx.handle, ns, _ptr(geometry), _ptr(weights), _ptr(out)), ctx)
    return out


def mean_curvature(ctx, nd, nc, geometry, h0=None, out=None):
    import torch
    if out is None:
        out = torch.empty(nd * nc, dtype=torch.float64, device=geometry.device)
    _check(lib().ibm_mean_curvature(ctx.handle, nd, nc, _ptr(geometry), _ptr(h0) if h0 is not None else None,
                                    _ptr(out)), ctx)
    return out


def membrane_forces(ctx, model, ns, nc, geometry, normal, sigma, hd, table, out=None, beta=0.0):
    import torch
    if out is None:
        out = torch.zeros(3 * ns * nc, dtype=torch.float64, device=geometry.device)
    _check(lib().ibm_membrane_forces(ctx.handle, ctypes.byref(model), ns, nc, _ptr(geometry),
                                     _ptr(normal) if normal is not None else None,
                                     _ptr(sigma) if sigma is not None else None,
                                     _ptr(hd) if hd is not None else None, _ptr(table), beta, _ptr(out)), ctx)
    return out


def state_write(ctx, path, st, layout, append=True):
    """Appends (or writes) one ib::state record in the linalg binary format."""
    _check(lib().ibm_state_write(ctx.handle, os.fsencode(path), int(append), ctypes.byref(st.c),
                                 ctypes.byref(layout)), ctx)


def state_read(ctx, path, st, layout, record=-1):
    """Reads record `record` (-1: the last complete one) into st's buffers;
    returns the number of complete records in the file."""
    n = _I(0)
    _check(lib().ibm_state_read(ctx.handle, os.fsencode(path), record, ctypes.byref(st.c) if st is not None else None,
                                ctypes.byref(layout), ctypes.byref(n)), ctx)
    return n.value


def io_write_dense(ctx, path, a, is_matrix, rows=None, cols=None, append=True):
    rows = a.numel() if rows is None else rows
    cols = 1 if cols is None else cols
    _check(lib().ibm_io_write_dense(ctx.handle, os.fsencode(path), int(append), int(is_matrix), rows, cols,
                                    _ptr(a)), ctx)


def io_read_dense(ctx, path, offset, out, is_matrix, rows, cols=1):
    nxt = ctypes.c_longlong(0)
    _check(lib().ibm_io_read_dense(ctx.handle, os.fsencode(path), offset, int(is_matrix), rows, cols, _ptr(out),
                                   ctypes.byref(nxt)), ctx)
    return nxt.value


def io_peek_dense(ctx, path, offset):
    m, r, c = _I(0), _I(0), _I(0)
    _check(lib().ibm_io_peek_dense(ctx.handle, os.fsencode(path), offset, ctypes.byref(m), ctypes.byref(r),
                                   ctypes.byref(c)), ctx)
    return m.value & 0xffffffff, r.value, c.value


class ibm_step_info(ctypes.Structure):
    _fields_ = [("iterations", _I * 4), ("residual", _D * 4), ("divergence", _D), ("timestep", _D),
                ("exponent", _I), ("reserved", _I)]


class ibm_cell_view(ctypes.Structure):
    _fields_ = [("positions", _VP), ("tangents", _VP * 2), ("forces", _VP), ("normals", _VP), ("sigma", _VP),
                ("area", _VP), ("volume", _VP), ("velocity", _VP)]


_hip = None


def _copy_device(dst, src_ptr):
    """Device-to-device copy of dst.numel() elements from a raw pointer."""
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.restype = _I
        _hip.hipMemcpy.argtypes = [_VP, _VP, ctypes.c_size_t, _I]
    import torch
    torch.cuda.synchronize()
    rc = _hip.hipMemcpy(ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(src_ptr),
                        dst.numel() * dst.element_size(), 3)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed ({rc})")
    return dst


class State:
    """ib::state: time, MAC velocity, centred pressure and per-type data sites
    (torch float64 CUDA tensors owned here)."""

    def __init__(self, u, p, xs, t=0.0):
        self.u, self.p, self.xs = list(u), p, list(xs)
        self.c = ibm_state()
        self.c.t = t
        for i in range(3):
            self.c.u[i] = self.u[i].data_ptr()
        self.c.p = p.data_ptr()
        for k, x in enumerate(self.xs):
            self.c.x[k] = x.data_ptr()

    @property
    def t(self):
        return self.c.t

    def clone(self):
        return State([a.clone() for a in self.u], self.p.clone(), [x.clone() for x in self.xs], self.c.t)


class Solver:
    """The immersed-boundary timestep (ibm_solver_*).

    cell_types: list of dicts with keys ops (RbfOperator with 3 operators:
    position, d/dtheta, d/dphi, or 6 with the second derivatives), cells,
    weights (CUDA tensor, ns), and the forces: tether (CUDA tensor,
    ns x 3 cells) with stiffness, and / or membrane (ibm.membrane(...)) with
    reference (CUDA tensor, the undeformed data sites nd x 3) and, for
    bending, data_ops (RbfOperator, 5 data-to-data derivative operators)."""

    def __init__(self, ctx, domain, refinement, timestep, diffusivity, density=1.0, tolerance=1e-10,
                 delta="cosine", precision=64, max_iterations=100, cell_types=(), fscale=0.0,
                 pseudo_pressure=False, multigrid=False):
        """With fscale > 0, `timestep` is the tracker's smallest step k_min;
        pseudo_pressure selects the u* = g + dt grad p wall condition;
        multigrid forces V-cycle solves on channel grids, which are
        otherwise solved directly (FFT x tridiagonal)."""
        self.ctx = ctx
        flags = (PSEUDO_PRESSURE if pseudo_pressure else 0) | (MULTIGRID if multigrid else 0)
        par = ibm_solver_params(timestep, diffusivity, density, tolerance, DELTA.get(delta, delta), precision,
                                max_iterations, flags, fscale)
        n = len(cell_types)
        arr = (ibm_cell_type * max(n, 1))()
        self._keep = []
        for k, c in enumerate(cell_types):
            op = c["ops"]
            tether = c.get("tether")
            mem = c.get("membrane") or ibm_membrane()
            dops = c.get("data_ops")
            ref = c.get("reference")
            arr[k] = ibm_cell_type(op.handle.value, op.K, op.M, c["cells"], 0, c["weights"].data_ptr(),
                                   tether.data_ptr() if tether is not None else None, c.get("stiffness", 0.0),
                                   mem, dops.handle.value if dops is not None else None,
                                   ref.data_ptr() if ref is not None else None)
            self._keep.append(c)
        dims, nd = _dims(domain)
        h = _VP()
        _check(lib().ibm_solver_create(ctx.handle, dims, nd, refinement, ctypes.byref(par), n, arr,
                                       ctypes.byref(h)), ctx)
        self.handle = h
        self.types = list(cell_types)

    def grids(self):
        g = (ibm_grid * 3)()
        pg = ibm_grid()
        _check(lib().ibm_solver_grids(self.handle, g, ctypes.byref(pg)))
        return g, pg

    def zero_state(self, xs):
        import torch
        g, pg = self.grids()
        u = [torch.zeros(g[i].size, dtype=torch.float64, device="cuda") for i in range(3)]
        p = torch.zeros(pg.size, dtype=torch.float64, device="cuda")
        return State(u, p, xs)

    def _bnd(self, boundary):
        return None if boundary is None else _ptrs(boundary)

    def step(self, st, boundary=None):
        _check(lib().ibm_solver_step(self.ctx.handle, self.handle, ctypes.byref(st.c), self._bnd(boundary)), self.ctx)

    def step_fixed(self, st, boundary=None):
        _check(lib().ibm_solver_step_fixed(self.ctx.handle, self.handle, ctypes.byref(st.c), self._bnd(boundary)),
               self.ctx)

    def capture(self, st, boundary=None):
        _check(lib().ibm_solver_capture(self.ctx.handle, self.handle, ctypes.byref(st.c), self._bnd(boundary)),
               self.ctx)

    def replay(self, st):
        _check(lib().ibm_solver_replay(self.ctx.handle, self.handle, ctypes.byref(st.c)), self.ctx)

    def layout(self):
        """ibm_state_layout of this solver's states (checkpoint / resume)."""
        L = ibm_state_layout()
        _check(lib().ibm_solver_layout(self.handle, ctypes.byref(L)), self.ctx)
        return L

    def info(self):
        i = ibm_step_info()
        _check(lib().ibm_solver_info(self.ctx.handle, self.handle, ctypes.byref(i)), self.ctx)
        return {"iterations": list(i.iterations), "residual": list(i.residual), "divergence": i.divergence,
                "timestep": i.timestep, "exponent": i.exponent}

    def cells(self, k):
        """Copies of a cell type's positions (3, nc, ns), tangents, forces,
        normals, sigma (nc, ns), area, volume (nc) and data-site velocity."""
        import torch
        v = ibm_cell_view()
        _check(lib().ibm_solver_cells(self.handle, k, ctypes.byref(v)))
        c = self.types[k]
        op = c["ops"]
        nc, ns, nd = c["cells"], op.M, op.K

        def get(ptr, shape):
            return _copy_device(torch.empty(shape, dtype=torch.float64, device="cuda"), ptr)

        return {"positions": get(v.positions, (3, nc, ns)), "tangents": [get(v.tangents[j], (3, nc, ns)) for j in range(2)],
                "forces": get(v.forces, (3, nc, ns)), "normals": get(v.normals, (3, nc, ns)),
                "sigma": get(v.sigma, (nc, ns)), "area": get(v.area, (nc,)), "volume": get(v.volume, (nc,)),
                "velocity": get(v.velocity, (3, nc, nd))}

    def close(self):
        if self.handle:
            lib().ibm_solver_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
