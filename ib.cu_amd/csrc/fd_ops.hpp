// Matrix-free staggered-grid operators shared by the stencil kernels, the
// multigrid hierarchy and the host-side Gershgorin bounds.
//
// The reference assembles every operator as a CSR matrix from 1-D pieces with
// Kronecker sums/products, x fastest (fd/combine.h:35-49):
//   laplacian(grid) = Lz(x)Iz'... = Lz (x) Iy (x) Ix + Iz (x) Ly (x) Ix + Iz (x) Iy (x) Lx
//   (fd/laplacian.h:53-68), with I the second-order corrected identity
//   (fd/identity.h:25-51) and L the 3-point Laplacian whose wall rows carry
//   the ghost-cell coefficient (fd/laplacian.h:18-49, fd/discretization.h:94-108).
// Because every I factor is diagonal, row (i, j, k) of the assembled matrix is
//   sigma*Ix_i*Iy_j*Iz_k*x + tau*(Iy_j*Iz_k*(Lx x) + Ix_i*Iz_k*(Ly x) + Ix_i*Iy_j*(Lz x)),
// which the kernels evaluate directly from the grid vector.
#pragma once
#include "ibm.h"

namespace ibm {

// One axis of an operator, in the element type of the kernel.
template <typename T>
struct op_axis {
    int n;            // points
    int periodic;
    T r2;             // resolution^2
    T lap_lo, lap_hi; // ghost coefficient folded into the first/last diagonal of L (times r2)
    T id_lo, id_hi;   // identity correction on the first/last row (second order)
};

template <typename T>
struct op3 {
    op_axis<T> a[3];
    T sigma, tau;     // A = sigma * I + tau * L
};

// Multigrid transfer along one axis, fine -> coarse (restriction) and coarse
// -> fine (interpolation), solvers/mg/interpolation.h:24-120: 4-tap weights
// 1 - |dx|/2 at dx in {-1-s, -s, 1-s, 2-s} with s the fine grid's shift.
template <typename T>
struct transfer_axis {
    int nf, nc;
    int periodic;
    T w[4];
    T int_lo, int_hi; // fine grid ghost coefficients (interior)
};

template <typename T>
struct transfer3 {
    transfer_axis<T> a[3];
};

inline double transfer_weight(double dx) { return 1 - (dx < 0 ? -dx : dx) / 2; }

template <typename T>
op_axis<T> make_op_axis(const ibm_axis& a)
{
    op_axis<T> o;
    o.n = a.points;
    o.periodic = !a.solid;
    o.r2 = (T) (a.resolution * a.resolution);
    o.lap_lo = a.solid ? (T) a.interior[0] : (T) 0;
    o.lap_hi = a.solid ? (T) a.interior[1] : (T) 0;
    o.id_lo = a.solid ? (T) a.laplacian[0] : (T) 0;
    o.id_hi = a.solid ? (T) a.laplacian[1] : (T) 0;
    return o;
}

template <typename T>
op3<T> make_op(const ibm_grid& g, double sigma, double tau)
{
    op3<T> o;
    for (int j = 0; j < 3; ++j) o.a[j] = make_op_axis<T>(g.axis[j]);
    o.sigma = (T) sigma;
    o.tau = (T) tau;
    return o;
}

template <typename T>
transfer3<T> make_transfer(const ibm_grid& fine, const ibm_grid& coarse)
{
    transfer3<T> t;
    for (int j = 0; j < 3; ++j) {
        const ibm_axis& f = fine.axis[j];
        transfer_axis<T>& x = t.a[j];
        x.nf = f.points;
        x.nc = coarse.axis[j].points;
        x.periodic = !f.solid;
        double s = f.shift;
        x.w[0] = (T) transfer_weight(-1 - s);
        x.w[1] = (T) transfer_weight(-s);
        x.w[2] = (T) transfer_weight(1 - s);
        x.w[3] = (T) transfer_weight(2 - s);
        x.int_lo = f.solid ? (T) f.interior[0] : (T) 0;
        x.int_hi = f.solid ? (T) f.interior[1] : (T) 0;
    }
    return t;
}

// Two-point operator along one axis ("view") from a grid to the grid whose
// alignment in that axis is shifted by 1/2: the first difference
// fd::differential (values -r, +r) and the average fd::average (1/2, 1/2)
// (fd/differential.h:44-82, fd/average.h:44-81).  Row r of the 1-D matrix is
//   c0 * x[r - s1] + c1 * x[r + 1 - s1],  s1 = (src shift > dst shift),
// where an index one past either end is the ghost value: periodic wrap, or
// the source grid's ghost coefficient times the end point (the matrices'
// skip_one / skip_end boundary entries).  Other axes: identity.
template <typename T>
struct twopoint_op {
    int sn[3], dn[3];
    int view;
    int s1;
    int periodic;
    T c0, c1;
    T int_lo, int_hi;
};

// Grid of `src` with the alignment of axis `view` shifted by 1/2.
ibm_status shifted_grid(const ibm_grid& src, int view, ibm_grid* dst);

// Gershgorin interval of A = sigma*I + tau*L on grid g (algo/gershgorin.h:14-50):
// rows differ only by which end of each axis they sit on, so the extrema are
// taken over the at most 27 row classes.
void gershgorin_bounds(const ibm_grid& g, double sigma, double tau, double* lo, double* hi);

} // namespace ibm
