// Navier-Stokes building blocks on the staggered grid: fd::differential /
// fd::average along one axis, ins::divergence, ins::gradient and
// ins::advection, all matrix-free (two-point kernels in fd_ops.hip).
//
// Restates:
//   fd::differential, fd::average          fd/differential.h:44-103, fd/average.h:44-102
//   fd::shift::directionally               fd/cell.h:96-104,165-184
//   ins::divergence / ins::gradient        ins/differential.h:46-115
//   ins::average / ins::advection          ins/advection.h:14-125
#include <hip/hip_runtime.h>
#include <cmath>
#include "context.hpp"
#include "fd_kernels.hpp"

namespace ibm {

ibm_status make_twopoint(const ibm_grid& src, int view, int kind, twopoint_op<double>* op, ibm_grid* dst)
{
    if (view < 0 || view >= src.dims) return IBM_ERR_ARG;
    ibm_status st = shifted_grid(src, view, dst);
    if (st != IBM_OK) return st;
    const ibm_axis& a = src.axis[view];
    for (int d = 0; d < 3; ++d) {
        op->sn[d] = src.axis[d].points;
        op->dn[d] = dst->axis[d].points;
    }
    op->view = view;
    op->s1 = a.shift > dst->axis[view].shift ? 1 : 0;
    op->periodic = !a.solid;
    double r = a.resolution;
    op->c0 = kind == 0 ? -r : 0.5;
    op->c1 = kind == 0 ? r : 0.5;
    op->int_lo = a.solid ? a.interior[0] : 0.0;
    op->int_hi = a.solid ? a.interior[1] : 0.0;
    return IBM_OK;
}

// ins::boundary(grid, dim) (ins/boundary.h:16-56): per axis, the
// discretization's boundary entries at (0, 0) and (n-1, n-1)
// (fd/discretization.h:102-108,138-141; zero on periodic axes) scaled by
// `solid` = 1, kron the zeroth-order identity of the other axes, summed over
// the axes; the axis `dim` contributes nothing (its comp == dim test).  order
// 2 scales each axis's entries by its resolution^2, the factor the ghost
// value carries in the Laplacian's wall rows.
ins_boundary_op make_ins_boundary(const ibm_grid& g, int dim, int order)
{
    ins_boundary_op B;
    for (int d = 0; d < 3; ++d) {
        const ibm_axis& a = g.axis[d];
        B.n[d] = a.points;
        const double scale = order == 2 ? a.resolution * a.resolution : 1.0;
        const bool on = a.solid && d != dim;
        B.lo[d] = on ? scale * a.boundary[0] : 0.0;
        B.hi[d] = on ? scale * a.boundary[1] : 0.0;
    }
    return B;
}

static bool same_points(const ibm_grid& a, const ibm_grid& b)
{
    for (int d = 0; d < 3; ++d)
        if (a.axis[d].points != b.axis[d].points) return false;
    return true;
}

static ibm_status check3(ibm_ctx* ctx, const ibm_grid* g, int n)
{
    if (!g) return fail(ctx, IBM_ERR_ARG, "null grid");
    for (int i = 0; i < n; ++i)
        if (g[i].dims != 3) return fail(ctx, IBM_ERR_ARG, "3-dimensional grids required");
    return IBM_OK;
}

// ins::divergence (ins/differential.h:46-83): out = sum_i D_i u_i on the
// grid reached by shifting u_i's grid along axis i (the cell centres for MAC),
// the three terms in one pass.
ibm_status divergence_of(ibm_ctx* ctx, const ibm_grid* u_grids, const double* const* u, double scale, double* out,
                         double* partial, double* max_out)
{
    twopoint_op<double> ops[3];
    ibm_grid first;
    for (int i = 0; i < 3; ++i) {
        if (!u[i]) return fail(ctx, IBM_ERR_ARG, "null component");
        ibm_grid dst;
        IBM_TRY(make_twopoint(u_grids[i], i, 0, &ops[i], &dst));
        if (i == 0) first = dst;
        else if (!same_points(first, dst)) return fail(ctx, IBM_ERR_SIZE, "divergence terms on different grids");
    }
    return launch_divergence(ctx, ops, scale, u, out, partial, max_out);
}

} // namespace ibm

using namespace ibm;

extern "C" {

ibm_status ibm_grid_shift(const ibm_grid* src, int view, ibm_grid* dst)
{
    if (!src || !dst || view < 0 || view >= src->dims) return IBM_ERR_ARG;
    return shifted_grid(*src, view, dst);
}

static ibm_status twopoint_apply(ibm_ctx* ctx, const ibm_grid* src, int view, int kind, double alpha,
                                 const double* x, double beta, double* y)
{
    if (!ctx) return IBM_ERR_ARG;
    IBM_TRY(check3(ctx, src, 1));
    if (!x || !y) return fail(ctx, IBM_ERR_ARG, "null vector");
    if (x == y) return fail(ctx, IBM_ERR_ARG, "x and y must not alias");
    twopoint_op<double> op;
    ibm_grid dst;
    if (make_twopoint(*src, view, kind, &op, &dst) != IBM_OK) return fail(ctx, IBM_ERR_ARG, "bad view %d", view);
    return launch_twopoint<double>(ctx, op, alpha, x, beta, y);
}

ibm_status ibm_differential(ibm_ctx* ctx, const ibm_grid* src, int view, double alpha, const double* x,
                            double beta, double* y)
{
    return twopoint_apply(ctx, src, view, 0, alpha, x, beta, y);
}

ibm_status ibm_average(ibm_ctx* ctx, const ibm_grid* src, int view, double alpha, const double* x, double beta,
                       double* y)
{
    return twopoint_apply(ctx, src, view, 1, alpha, x, beta, y);
}

// fd::lower_boundary / fd::upper_boundary (fd/boundary_ops.h:21-71) times
// face data: y += alpha * B g.  B is (points x 1 along `view`, scaled by
// resolution^order) kron the (order-corrected) identity of the other axes;
// g holds one value per face point (other axes, x fastest) or, when null,
// the constant `value` (the algo::fill vectors of test/ib-whole.cc:146-149).
ibm_status ibm_boundary_apply(ibm_ctx* ctx, const ibm_grid* g, int view, int side, int order, double alpha,
                              const double* values, double value, double* y)
{
    if (!ctx) return IBM_ERR_ARG;
    IBM_TRY(check3(ctx, g, 1));
    if (view < 0 || view > 2 || (side != 0 && side != 1) || (order != 0 && order != 2) || !y)
        return fail(ctx, IBM_ERR_ARG, "bad boundary arguments (view %d side %d order %d)", view, side, order);
    const ibm_axis& a = g->axis[view];
    if (!a.solid) return IBM_OK;            // periodic: the boundary operator has no columns
    boundary_op B;
    for (int d = 0; d < 3; ++d) {
        const ibm_axis& e = g->axis[d];
        B.n[d] = e.points;
        bool corr = order == 2 && e.solid;
        B.id_lo[d] = corr ? 1.0 + e.laplacian[0] : 1.0;
        B.id_hi[d] = corr ? 1.0 + e.laplacian[1] : 1.0;
    }
    B.view = view;
    B.row = side ? a.points - 1 : 0;
    B.scale = std::pow(a.resolution, order) * a.boundary[side];
    return launch_boundary(ctx, B, alpha, values, value, y);
}

ibm_status ibm_ins_boundary_apply(ibm_ctx* ctx, const ibm_grid* g, int dim, int order, double alpha, const double* x,
                                  double beta, double* y)
{
    if (!ctx) return IBM_ERR_ARG;
    IBM_TRY(check3(ctx, g, 1));
    if (dim < -1 || dim > 2 || (order != 0 && order != 2))
        return fail(ctx, IBM_ERR_ARG, "bad ins::boundary arguments (dim %d order %d)", dim, order);
    if (!x || !y) return fail(ctx, IBM_ERR_ARG, "null vector");
    return launch_ins_boundary(ctx, make_ins_boundary(*g, dim, order), alpha, x, beta, y);
}

ibm_status ibm_divergence(ibm_ctx* ctx, const ibm_grid* u_grids, const double* const* u, double* out)
{
    if (!ctx) return IBM_ERR_ARG;
    IBM_TRY(check3(ctx, u_grids, 3));
    if (!u || !out) return fail(ctx, IBM_ERR_ARG, "null vector");
    return divergence_of(ctx, u_grids, u, 1.0, out, nullptr, nullptr);
}

// ins::gradient (ins/differential.h:88-112): out_i = D_i p, on p's grid
// shifted along axis i.
ibm_status ibm_gradient(ibm_ctx* ctx, const ibm_grid* p_grid, double alpha, const double* p, double beta,
                        double* const* out)
{
    if (!ctx) return IBM_ERR_ARG;
    IBM_TRY(check3(ctx, p_grid, 1));
    if (!p || !out) return fail(ctx, IBM_ERR_ARG, "null vector");
    for (int i = 0; i < 3; ++i) {
        if (!out[i]) return fail(ctx, IBM_ERR_ARG, "null component");
        twopoint_op<double> op;
        ibm_grid dst;
        IBM_TRY(make_twopoint(*p_grid, i, 0, &op, &dst));
        IBM_TRY(launch_twopoint<double>(ctx, op, alpha, p, beta, out[i]));
    }
    return IBM_OK;
}

// ins::advection (ins/advection.h:58-125):
//   H^i = sum_j D_j [ (A_i u^j) o (A_j u^i) ],
// A_i averaging along axis i, D_j the difference along j on the grid of
// component j shifted along i (fd::shift::directionally).
ibm_status ibm_advection(ibm_ctx* ctx, const ibm_grid* u_grids, const double* const* u, double* const* H)
{
    if (!ctx) return IBM_ERR_ARG;
    IBM_TRY(check3(ctx, u_grids, 3));
    if (!u || !H) return fail(ctx, IBM_ERR_ARG, "null vector");
    for (int i = 0; i < 3; ++i)
        if (!u[i] || !H[i]) return fail(ctx, IBM_ERR_ARG, "null component");
    ibm_grid ga[3][3];
    twopoint_op<double> avg_ij[3][3];      // A_i applied to u^j
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) IBM_TRY(make_twopoint(u_grids[j], i, 1, &avg_ij[i][j], &ga[i][j]));
    for (int i = 0; i < 3; ++i) {
        advection_op<double> Op;
        for (int j = 0; j < 3; ++j) {
            const ibm_grid& g = ga[i][j];                        // grid of A_i u^j
            if (!same_points(g, ga[j][i])) return fail(ctx, IBM_ERR_SIZE, "advection products on different grids");
            ibm_grid back;
            IBM_TRY(make_twopoint(g, j, 0, &Op.D[j], &back));
            if (!same_points(back, u_grids[i])) return fail(ctx, IBM_ERR_SIZE, "advection result off the component grid");
            Op.A[j] = avg_ij[i][j];
            Op.B[j] = avg_ij[j][i];
        }
        IBM_TRY(launch_advection<double>(ctx, Op, u, i, H[i]));
    }
    return IBM_OK;
}

} // extern "C"
