"""CPU oracle for the staggered-grid operators and solvers.  TEST
INFRASTRUCTURE ONLY (used by tests/, smoke() and bench.py's cpu_baseline leg
as the checker; the product never routes through it).

Restates, with scipy.sparse CSR matrices in FP64:
  1-D Laplacian          fd/laplacian.h:18-49 (3-point rows + boundary entries)
  1-D identity           fd/identity.h:25-51 (second-order corner corrections)
  boundary entries       fd/discretization.h:16-45,94-108 (single_entry, interior)
  ins::boundary          ins/boundary.h:16-56 (pseudo-pressure wall operator)
  d-D assembly           fd/combine.h:10-49 (Kronecker sum/product, x fastest)
  1-D differential/avg   fd/differential.h:44-82, fd/average.h:44-81
  Gershgorin             algo/gershgorin.h:14-50
  Chebyshev polynomial   algo/chebyshev.h:120-169 (chebw<4>, Horner)
  smoother interval      solvers/mg/chebyshev.h:14-40
  transfers              solvers/mg/interpolation.h:19-152 (kron order x slowest)
  multigrid              solvers/multigrid.h:15-212
  PCG / chebpcg          solvers/pcg.h:11-44, solvers/chebpcg.h, preconditioners/chebyshev.h

Grid axes come from the C oracle (orc.grid_axes), which restates
fd/discretization.h.  Pinning of every function is listed in
tests/test_oracle_fd.py.
"""
import math

import numpy as np
import scipy.sparse as sp


def _modulo(x, s):
    return x - s * math.floor(x / s)


# ----------------------------------------------------------------------
# 1-D pieces
# ----------------------------------------------------------------------
def single_entry(rows, cols, scale, row_end, col_end, value):
    """fd/discretization.h:18-45: one nonzero at a corner, or nothing."""
    v = scale * value
    if not rows or not cols or not v:
        return sp.csr_matrix((rows, cols))
    r = rows - 1 if row_end else 0
    c = cols - 1 if col_end else 0
    return sp.csr_matrix(([v], ([r], [c])), shape=(rows, cols))


def interior_entry(axis, rows, cols, scale, lower):
    """discretization::interior(rows, cols, scale, tag) (fd/discretization.h:94-100,129-134)."""
    if not axis.solid:
        return single_entry(rows, cols, scale, not lower, lower, 1.0)
    return single_entry(rows, cols, scale, not lower, not lower, axis.interior[0 if lower else 1])


def lap1(axis):
    """fd/laplacian.h:18-49"""
    rows = axis.points
    if not rows:
        return sp.csr_matrix((0, 0))
    scale = axis.resolution * axis.resolution
    main = sp.diags([np.ones(rows - 1), -2 * np.ones(rows), np.ones(rows - 1)], [-1, 0, 1],
                    shape=(rows, rows), format="csr") * scale
    m = main + interior_entry(axis, rows, rows, scale, True) + interior_entry(axis, rows, rows, scale, False)
    return sp.csr_matrix(m)


def id1(axis, order):
    """fd/identity.h:25-51; order 2 adds the laplacian coefficients at the ends."""
    rows = axis.points
    d = np.ones(rows)
    if order == 2 and axis.solid and rows:
        d[0] += axis.laplacian[0]
        d[rows - 1] += axis.laplacian[1]
    return sp.diags(d, 0, format="csr")


def boundary1(axis, lower):
    """fd::__1::boundary (fd/boundary_ops.h:16-27): points x solid, one entry
    (the ghost-cell boundary coefficient) in the first / last row."""
    rows = axis.points
    if not axis.solid:
        return sp.csr_matrix((rows, 0))
    return single_entry(rows, 1, 1.0, not lower, not lower, axis.boundary[0 if lower else 1])


def boundary_op(axes, view, lower, order):
    """fd::lower_boundary / fd::upper_boundary (fd/boundary_ops.h:29-71)."""
    mats = []
    for j, a in enumerate(axes):
        mats.append(a.resolution ** order * boundary1(a, lower) if j == view else id1(a, order))
    return kron_product(mats)


def ins_boundary(axes, dim, order=0):
    """ins::boundary(grid, dim) (ins/boundary.h:16-56): per axis the
    discretization's boundary entries at (0, 0) and (n-1, n-1)
    (fd/discretization.h:102-108, 138-141) with scale `solid`, none on the
    axis `dim` (comp == dim) or on periodic axes; folded over the reversed
    components as kron(lb, ri) + kron(li, rb) with the zeroth-order identity,
    so x is fastest.  order 2 scales each axis's entries by resolution^2."""
    parts = []
    for j, a in enumerate(axes):
        rows = a.points
        if j == dim or not a.solid:
            b = sp.csr_matrix((rows, rows))
        else:
            b = (single_entry(rows, rows, 1.0, False, False, a.boundary[0])
                 + single_entry(rows, rows, 1.0, True, True, a.boundary[1]))
            if order == 2:
                b = b * a.resolution ** 2
        parts.append((sp.csr_matrix(b), sp.identity(rows, format="csr")))
    seq = list(reversed(parts))
    lb, li = seq[0]
    for rb, ri in seq[1:]:
        lb, li = sp.kron(lb, ri, format="csr") + sp.kron(li, rb, format="csr"), sp.kron(li, ri, format="csr")
    return sp.csr_matrix(lb)


def shifted_axis(axis):
    """The discretization with alignment shifted by 1/2 (fd/differential.h:49-52)."""
    class A:
        pass
    a = A()
    al = _modulo(axis.shift + 0.5, 1.0)
    on_boundary = al == 0
    a.solid = axis.solid
    a.resolution = axis.resolution
    a.alignment = al
    a.shift = al + (1 if (axis.solid and on_boundary) else 0)
    a.cells = axis.cells
    a.points = a.cells - (1 if (on_boundary and axis.solid) else 0)
    a.interior = list(axis.interior)  # the source component's coefficients are used (:78-79)
    return a


def _two_point(axis, value_fn, lo_scale, hi_scale):
    """Common structure of fd::differential / fd::average (1-D)."""
    src_shift = axis.shift
    dst = shifted_axis(axis)
    dst_shift = dst.shift
    skip_one = 1 if src_shift > dst_shift else 0
    skip_end = _modulo(1 - src_shift, 1) < _modulo(1 - dst_shift, 1)
    cols = axis.points
    rows = dst.points
    nnz = rows + cols - 1
    r_idx, c_idx, vals = [], [], []
    starts = [0] * (rows + 1)
    for tid in range(nnz):
        loc = (tid + skip_one) % 2
        row = (tid + skip_one) // 2
        col = loc + row - skip_one
        r_idx.append(row)
        c_idx.append(col)
        vals.append(value_fn(loc))
    m = sp.csr_matrix((vals, (r_idx, c_idx)), shape=(rows, cols))
    if skip_one:
        m = m + interior_entry(axis, rows, cols, lo_scale, True)
    if skip_end:
        m = m + interior_entry(axis, rows, cols, hi_scale, False)
    return sp.csr_matrix(m), dst


def diff1(axis):
    """fd/differential.h:44-82"""
    s = axis.resolution
    return _two_point(axis, lambda loc: s * (-1 + 2 * loc), -s, s)


def avg1(axis):
    """fd/average.h:44-81"""
    return _two_point(axis, lambda loc: 0.5, 0.5, 0.5)


# ----------------------------------------------------------------------
# d-D assembly, x fastest (fd/combine.h:35-49: fold over the reversed axes)
# ----------------------------------------------------------------------
def kron_product(mats):
    """combine(grid, k) with Kronecker products: mats given x first."""
    out = None
    for m in reversed(mats):
        out = m if out is None else sp.kron(out, m, format="csr")
    return sp.csr_matrix(out)


def laplacian(axes):
    """fd::laplacian(grid) (fd/laplacian.h:53-68): Kronecker sum with I2."""
    op, ident = None, None
    for a in reversed(axes):
        L, I = lap1(a), id1(a, 2)
        if op is None:
            op, ident = L, I
        else:
            op = sp.kron(op, I, format="csr") + sp.kron(ident, L, format="csr")
            ident = sp.kron(ident, I, format="csr")
    return sp.csr_matrix(op)


def identity(axes, order=2):
    """fd::identity(grid, order) (fd/identity.h:55-66)."""
    return kron_product([id1(a, order) for a in axes])


def operator(axes, sigma, tau):
    """sigma * I2 + tau * L (e.g. Helmholtz I2 - lambda L, ins/diffusion.h:33)."""
    return sp.csr_matrix(sigma * identity(axes, 2) + tau * laplacian(axes))


def differential(axes, view):
    """fd::differential(grid, view) (fd/differential.h:86-103)."""
    mats = []
    for j, a in enumerate(axes):
        mats.append(diff1(a)[0] if j == view else id1(a, 1))
    return kron_product(mats)


def average(axes, view):
    """fd::average(grid, view) (fd/average.h:85-102)."""
    mats = []
    for j, a in enumerate(axes):
        mats.append(avg1(a)[0] if j == view else id1(a, 0))
    return kron_product(mats)


# ----------------------------------------------------------------------
# Gershgorin, Chebyshev
# ----------------------------------------------------------------------
def gershgorin(A):
    """algo/gershgorin.h:14-50: per row, the diagonal entry and the sum of the
    magnitudes of the others (CSR rows hold no duplicate columns)."""
    A = sp.csr_matrix(A)
    A.sum_duplicates()
    diag = A.diagonal()
    off = np.asarray(abs(A).sum(axis=1)).ravel() - np.abs(diag)
    return float((diag - off).min()), float((diag + off).max())


def _poly(c, a, b):
    n = len(c)
    return sum(c[i] * a ** i * b ** (n - 1 - i) for i in range(n))


def chebw4(a, b):
    """chebw<4> (algo/chebyshev.h:120-139)."""
    q = _poly([1, 28, 70, 28, 1], a, b)
    return [_poly([128], a, b) / q, _poly([256, 256], a, b) / q,
            _poly([160, 448, 160], a, b) / q, _poly([32, 224, 224, 32], a, b) / q]


def chebyshev_apply(A, a, b, r):
    """algo::chebyshev::polynomial (algo/chebyshev.h:149-163)."""
    w = chebw4(a, b)
    s = w[0] * r
    for i in range(1, 4):
        t = w[i] * r - A @ s
        s = t
    return s


def smoother_interval(axes, A):
    """solvers/mg/chebyshev.h:14-37"""
    lengths = [a.length for a in axes]
    smallest = min(lengths) / 2
    prop = 1.0
    for L in reversed(lengths):
        prop = (smallest / L) * prop
    a, b = gershgorin(A)
    if abs(a) > abs(b):
        a, b = b, a
    return a + (b - a) * prop, b


# ----------------------------------------------------------------------
# Multigrid
# ----------------------------------------------------------------------
def _weights(s):
    g = lambda dx: 1 - abs(dx) / 2
    return [g(-1 - s), g(-s), g(1 - s), g(2 - s)]


def interp1(fine, coarse):
    """interpolation(component, interpolation_tag) (solvers/mg/interpolation.h:43-82)."""
    w = _weights(fine.shift)
    rows, cols = fine.points, coarse.points
    correction = rows - 2 * cols
    nnz = 4 * cols - 2 + correction
    swizzle = [3, 1, 2, 0]
    shifts = [-1, 0, 0, 1]
    starts = [0] * (rows + 1)
    for r in range(rows):
        starts[r] = min(1 + 2 * (r - 1), nnz - 1) if r else 0
    starts[rows] = nnz
    data, ind = [0.0] * nnz, [0] * nnz
    for tid in range(nnz):
        col = (tid + 1) // 4
        shift = (tid + 1) % 4
        swiz = swizzle[shift] if col + shifts[shift] < cols else shift
        data[tid] = w[swiz]
        ind[tid] = min(cols - 1, col + shifts[shift])
    m = sp.csr_matrix((data, ind, starts), shape=(rows, cols))
    m.sum_duplicates()
    m = m + interior_entry(fine, rows, cols, w[correction], False) + interior_entry(fine, rows, cols, w[3], True)
    return sp.csr_matrix(m)


def restrict1(fine, coarse):
    """interpolation(component, restriction_tag) (solvers/mg/interpolation.h:84-120)."""
    w = _weights(fine.shift)
    cols, rows = fine.points, coarse.points
    correction = cols - 2 * rows
    nnz = 4 * rows - 2 + correction
    starts = [0] * (rows + 1)
    for r in range(rows):
        starts[r] = 4 * r - 1 if r else 0
    starts[rows] = nnz
    data, ind = [0.0] * nnz, [0] * nnz
    for tid in range(nnz):
        shift = (tid + 1) % 4
        row = (tid + 1) // 4
        ind[tid] = 2 * row + shift - 1
        data[tid] = w[shift] / 2
    m = sp.csr_matrix((data, ind, starts), shape=(rows, cols))
    m = m + interior_entry(fine, rows, cols, w[0] / 2, True)
    if not correction:
        m = m + interior_entry(fine, rows, cols, w[3] / 2, False)
    return sp.csr_matrix(m)


def transfer(fine_axes, coarse_axes, kind, order="reference"):
    """mg::interpolation / mg::restriction (solvers/mg/interpolation.h:124-152).
    order="reference": kron(kron(Px, Py), Pz) as the reference folds (x
    slowest); order="x-fastest": kron(kron(Pz, Py), Px), matching the grid
    vectors' layout."""
    f = interp1 if kind == "interpolation" else restrict1
    mats = [f(a, b) for a, b in zip(fine_axes, coarse_axes)]
    seq = mats if order == "reference" else list(reversed(mats))
    out = seq[0]
    for m in seq[1:]:
        out = sp.kron(out, m, format="csr")
    return sp.csr_matrix(out)


def refined(axes):
    """solvers/multigrid.h:153-172"""
    for a in axes:
        c = a.cells
        if c & 1:
            return False
        if (not a.solid and c <= 2) or (a.solid and c <= 4):
            return False
    return True


class Multigrid:
    """solvers::multigrid over A = sigma*I2 + tau*L with Chebyshev smoothing.
    `axes_of(refinement)` returns the grid axes at a refinement."""

    def __init__(self, axes_of, refinement, sigma, tau, order="reference"):
        self.levels = []
        ref = refinement
        while True:
            axes = axes_of(ref)
            A = operator(axes, sigma, tau)
            a, b = smoother_interval(axes, A)
            lvl = {"axes": axes, "A": A, "ab": (a, b)}
            self.levels.append(lvl)
            if refined(axes) and ref >= 2:
                coarse = axes_of(ref >> 1)
                lvl["R"] = transfer(axes, coarse, "restriction", order)
                lvl["P"] = transfer(axes, coarse, "interpolation", order)
                ref >>= 1
            else:
                break

    def smooth(self, l, b):
        lv = self.levels[l]
        return chebyshev_apply(lv["A"], *lv["ab"], b)

    def vcycle(self, b, l=0):
        """solvers/multigrid.h:120-128 and the direct solver :46-63"""
        lv = self.levels[l]
        x = self.smooth(l, b)
        if l + 1 == len(self.levels):
            return x
        r = b - lv["A"] @ x
        rc = lv["R"] @ r
        ec = self.vcycle(rc, l + 1)
        x = x + lv["P"] @ ec
        return self.smooth(l, b - lv["A"] @ x) + x

    def nested_iteration(self, b, l=0):
        """solvers/multigrid.h:110-117"""
        lv = self.levels[l]
        if l + 1 == len(self.levels):
            return self.smooth(l, b)
        x = lv["P"] @ self.nested_iteration(lv["R"] @ b, l + 1)
        r = b - lv["A"] @ x
        return x + self.vcycle(r, l)

    def solve(self, b, tol, maxit=200):
        """iterate() (solvers/multigrid.h:88-104) with an absolute tolerance."""
        A = self.levels[0]["A"]
        x = self.nested_iteration(b)
        r = b - A @ x
        it = 0
        while np.linalg.norm(r) > tol and it < maxit:
            e = self.vcycle(r)
            x = x + e
            r = r - A @ e
            it += 1
        return x, it


def pcg(precondition, A, r, tol, maxit=10000):
    """solvers/pcg.h:11-44 (stopping on the search direction norm)."""
    n = r.shape[0]
    d = precondition(r)
    delta_new = float(r @ d)
    x = np.zeros(n)
    count = 0
    while np.linalg.norm(d) > tol and count < maxit:
        q = A @ d
        nu = float(d @ q)
        if abs(nu) < tol * tol:
            break
        alpha = delta_new / nu
        x = x + alpha * d
        r = r - alpha * q
        q = precondition(r)
        delta_old = delta_new
        delta_new = float(r @ q)
        beta = delta_new / delta_old
        q = q + beta * d
        d = q
        count += 1
    return x, count


def chebpcg(A, b, tol):
    """solvers::chebpcg: PCG with the Chebyshev preconditioner over the
    Gershgorin interval, arguments swapped as preconditioners/chebyshev.h:11-12."""
    lo, hi = gershgorin(A)
    return pcg(lambda r: chebyshev_apply(A, hi, lo, r), A, b, tol)


# ----------------------------------------------------------------------
# ins::divergence / ins::gradient / ins::advection (ins/differential.h,
# ins/advection.h), composed from the assembled 1-D operators.
# `axes_of(align)` returns the axes of the grid with those alignments.
# ----------------------------------------------------------------------
def _shift_align(align, view, axes):
    a = list(align)
    s = axes[view].shift + 0.5
    a[view] = s - math.floor(s)
    return a


def divergence(axes_of, aligns, u):
    """sum_i D_i u_i (ins/differential.h:46-83)."""
    out = None
    for i in range(3):
        ax = axes_of(aligns[i])
        t = differential(ax, i) @ u[i]
        out = t if out is None else out + t
    return out


def gradient(axes_of, align, p):
    """(D_i p)_i (ins/differential.h:88-112)."""
    ax = axes_of(align)
    return [differential(ax, i) @ p for i in range(3)]


def advection(axes_of, aligns, u):
    """H^i = sum_j D_j[(A_i u^j) o (A_j u^i)] (ins/advection.h:58-125)."""
    H = []
    for i in range(3):
        h = None
        for j in range(3):
            axj = axes_of(aligns[j])
            axi = axes_of(aligns[i])
            a = average(axj, i) @ u[j]
            b = average(axi, j) @ u[i]
            prod = a * b
            g = axes_of(_shift_align(aligns[j], i, axj))
            t = differential(g, j) @ prod
            h = t if h is None else h + t
        H.append(h)
    return H
