"""CPU restatement of the RBF surface evaluation and surface geometry.

Test infrastructure only: the tests and the benchmark's CPU baseline use it
as the checker; nothing in the library calls it.

  evaluate   bases::impl::slice * matrix (bases/operators.h:95-109):
             Y = E[0:K, :]^T X, here in long double (80-bit) so that it is
             a reference for FP64 matrix-core results.
  geometry   bases::geometry::compute (bases/geometry.h:95-119):
             n = t1 x t2 / |t1 x t2|, sigma = w |t1 x t2|; plus the cell
             area sum(sigma) and volume sum(x . n sigma) / 3.

Pinned by (tests/test_oracle_rbf.py): the evaluator reproducing constants
and the degree-5 RBC shape exactly; derivative operators reproducing the
shape's parameter derivatives; data-to-data evaluation = identity; the
unit sphere's normals, area 4 pi and volume 4 pi / 3.
"""
import numpy as np


def evaluate(Et, K, X, dtype=np.longdouble):
    """Et: (M, lde) sample-major operator; X: (N, K) = K x N column-major.
    Returns (N, M) = M x N column-major."""
    A = np.asarray(Et)[:, :K].astype(dtype)
    B = np.asarray(X).astype(dtype)
    return (B @ A.T).astype(np.float64)


def geometry(y, t1, t2, w):
    """y, t1, t2: (3, nc, ns); w: (ns).  Returns normal (3, nc, ns),
    sigma (nc, ns), area (nc), volume (nc)."""
    n = np.cross(t1, t2, axis=0)
    det = np.sqrt((n * n).sum(0))
    normal = n / det
    sigma = w[None, :] * det
    area = sigma.sum(1)
    volume = (w[None, :] * (y * n).sum(0)).sum(1) / 3.0
    return normal, sigma, area, volume
