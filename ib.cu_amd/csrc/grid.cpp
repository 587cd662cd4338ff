// Host-side grid construction.
//
// Restates fd::domain's base unit (fd/domain.h:30-32 with algo/gcd.h:9-34),
// fd::grid's resolution (fd/grid.h:39-46), fd::discretization's counts and
// ghost-cell coefficients (fd/discretization.h:47-78,158-164) and the MAC
// staggering rules (fd/cell.h:10-17,80-114,194-196).
#include <cfloat>
#include <cmath>
#include <utility>
#include "geometry.hpp"

namespace {

using ibm::ref_floor;
using ibm::ref_modulo;

// Euclid's algorithm for floating-point values (algo/gcd.h:9-23).
double gcd_impl(double a, double b)
{
    double eps = DBL_EPSILON;
    while ((b < 0 ? -b : b) > eps) {
        if (a < b) std::swap(a, b);
        double c = a - ref_floor(a / b) * b;
        a = b;
        b = c;
        eps *= 2;
    }
    return a;
}

// algo::gcd for floating point (algo/gcd.h:27-34).
double gcd(double a, double b)
{
    if (a < b) std::swap(a, b);
    return gcd_impl(1.0, b / a) * a;
}

// Ghost-cell coefficients for alpha u + beta du = gamma (fd/discretization.h:68-77).
struct coeffs { double boundary, interior, laplacian; };

coeffs robin_coefficients(const double p[2], double h, double s_in)
{
    double s = 1 - ref_modulo(1 - s_in, 1);
    double denom = p[1] - p[0] * s * h;
    coeffs c;
    c.boundary = -h / denom;
    c.interior = (p[1] + p[0] * h * (1 - s)) / denom;
    c.laplacian = (-(p[1] * (1 - 2 * s) - p[0] * s * h * (1 - s))) / (2 * denom);
    return c;
}

ibm_status make_axis(const ibm_dimension& d, double alignment, double resolution, ibm_axis* a)
{
    if (!(d.length >= 0)) return IBM_ERR_BAD_GRID;    // fd/dimension.h:29-30
    double al = ref_modulo(alignment, 1.0);          // fd/cell.h:15-16
    bool on_boundary = !al;                          // fd/cell.h:11
    a->length = d.length;
    a->resolution = resolution;
    a->alignment = al;
    a->shift = al + (d.solid && on_boundary);
    double cells = d.length * resolution + 0.5;
    // The base unit is the reference's floating GCD (algo/gcd.h), which can
    // collapse to ~1e-15 for lengths whose ratio is not representable; such a
    // grid would need ~1e15 cells and is rejected.
    if (!(cells >= 1.0 && cells < 2147483647.0)) return IBM_ERR_BAD_GRID;
    a->cells = (int) cells;
    a->points = a->cells - (on_boundary && d.solid);
    a->solid = d.solid ? 1 : 0;
    a->reserved = 0;
    if (d.solid) {
        coeffs lo = robin_coefficients(d.lower, 1 / resolution, al);
        coeffs up = robin_coefficients(d.upper, -1 / resolution, 1 - al);
        a->boundary[0] = lo.boundary;   a->boundary[1] = up.boundary;
        a->interior[0] = lo.interior;   a->interior[1] = up.interior;
        a->laplacian[0] = lo.laplacian; a->laplacian[1] = up.laplacian;
    } else {
        for (int i = 0; i < 2; ++i) a->boundary[i] = a->interior[i] = a->laplacian[i] = 0;
    }
    a->robin[0][0] = d.lower[0]; a->robin[0][1] = d.lower[1];
    a->robin[1][0] = d.upper[0]; a->robin[1][1] = d.upper[1];
    if (a->cells < 1 || a->points < 0) return IBM_ERR_BAD_GRID;
    return IBM_OK;
}

} // namespace

extern "C" {

ibm_status ibm_domain_unit(const ibm_dimension* dims, int ndims, double* unit)
{
    if (!dims || !unit || ndims < 1 || ndims > 3) return IBM_ERR_ARG;
    double u = dims[0].length;
    for (int i = 1; i < ndims; ++i) u = gcd(u, dims[i].length);
    *unit = u;
    return IBM_OK;
}

ibm_status ibm_grid_make(const ibm_dimension* dims, int ndims, int refinement,
                         const double* alignment, ibm_grid* out)
{
    if (!dims || !alignment || !out || ndims < 1 || ndims > 3 || refinement < 1) return IBM_ERR_ARG;
    for (int i = 0; i < ndims; ++i)
        if (!(dims[i].length > 0)) return IBM_ERR_BAD_GRID;
    double unit;
    ibm_status st = ibm_domain_unit(dims, ndims, &unit);
    if (st != IBM_OK) return st;
    ibm_grid g = {};
    g.dims = ndims;
    g.refinement = refinement;
    g.unit = unit;
    double res = refinement / unit;                  // fd/grid.h:43-44
    for (int i = 0; i < ndims; ++i) {
        st = make_axis(dims[i], alignment[i], res, &g.axis[i]);
        if (st != IBM_OK) return st;
    }
    for (int i = ndims; i < 3; ++i) {
        // Unused trailing axes: one point, so products over axes stay valid.
        ibm_axis& a = g.axis[i];
        a = ibm_axis{};
        a.length = 1; a.resolution = 1; a.alignment = 0.5; a.shift = 0.5;
        a.cells = 1; a.points = 1; a.solid = 0;
    }
    *out = g;
    return IBM_OK;
}

ibm_status ibm_mac_grids(const ibm_dimension* dims, int ndims, int refinement, ibm_grid* out)
{
    if (!out) return IBM_ERR_ARG;
    for (int c = 0; c < ndims; ++c) {
        double al[3];
        for (int j = 0; j < 3; ++j) al[j] = (j == c) ? 0.0 : 0.5;
        ibm_status st = ibm_grid_make(dims, ndims, refinement, al, &out[c]);
        if (st != IBM_OK) return st;
    }
    return IBM_OK;
}

ibm_status ibm_centered_grid(const ibm_dimension* dims, int ndims, int refinement, ibm_grid* out)
{
    double al[3] = {0.5, 0.5, 0.5};
    return ibm_grid_make(dims, ndims, refinement, al, out);
}

ibm_status ibm_grid_coarsen(const ibm_grid* fine, ibm_grid* coarse)
{
    if (!fine || !coarse) return IBM_ERR_ARG;
    int ref = fine->refinement >> 1;                 // solvers/multigrid.h:189
    if (ref < 1) return IBM_ERR_BAD_GRID;
    ibm_dimension dims[3];
    double al[3];
    for (int i = 0; i < fine->dims; ++i) {
        const ibm_axis& a = fine->axis[i];
        dims[i].length = a.length;
        dims[i].solid = a.solid;
        dims[i].reserved = 0;
        dims[i].lower[0] = a.robin[0][0]; dims[i].lower[1] = a.robin[0][1];
        dims[i].upper[0] = a.robin[1][0]; dims[i].upper[1] = a.robin[1][1];
        al[i] = a.alignment;
    }
    return ibm_grid_make(dims, fine->dims, ref, al, coarse);
}

int64_t ibm_grid_size(const ibm_grid* g)
{
    if (!g) return 0;
    int64_t n = 1;
    for (int i = 0; i < g->dims; ++i) n *= g->axis[i].points;
    return n;
}

} // extern "C"
