// CPU oracle for the immersed-boundary hot path.  TEST INFRASTRUCTURE ONLY:
// this library is loaded by tests/, by __graft_entry__.smoke() and by the
// cpu_baseline leg of bench.py, always as the checker, never as the product.
//
// It restates the reference's algorithms operation by operation in FP64:
//   grid geometry     fd/discretization.h:47-179, fd/grid.h:39-46,
//                     algo/gcd.h:9-34, util/math.h:158-168
//   delta kernels     ib/cosine.h:9-16, ib/peskin.h:6-14, ib/bspline.h:13-27
//   stencil pattern   ib/delta.h:113-147 (every_other_pattern, meshwidths 4)
//   index policy      ib/indexing.h:10-122, ib/sweep.h:38-125
//   spread            ib/novel.h:58-168 (sort by key, 8 sweeps of 8 values,
//                     segmented sums, 8 plane buffers, plane combine)
//   interpolate       ib/interpolate.h:49-76 (64-point gather in pattern order)
// plus a long-double direct evaluation of the spread/interp sums from their
// definition, used to pin the restatement.
//
// Pinning of each function is listed in tests/test_oracle_ib.py.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <utility>
#include <vector>

extern "C" {

typedef struct orc_dim {
    double length;
    int solid;
    int reserved;
    double lower[2];
    double upper[2];
} orc_dim;

// Same field order as the product's ibm_axis so tests can compare them.
typedef struct orc_axis {
    double length, resolution, alignment, shift;
    int cells, points, solid, reserved;
    double boundary[2], interior[2], laplacian[2];
    double robin[2][2];
} orc_axis;

}

namespace {

// util/math.h:158-168
double r_floor(double x) { return (double) (long) (x - (x < 0 && (-x != (double) (long) -x))); }
double r_modulo(double x, double s) { return x - s * r_floor(x / s); }

// algo/gcd.h:9-34
double r_gcd(double a, double b)
{
    if (a < b) std::swap(a, b);
    double aa = 1.0, bb = b / a;
    double eps = DBL_EPSILON;
    while (std::fabs(bb) > eps) {
        if (aa < bb) std::swap(aa, bb);
        double c = aa - r_floor(aa / bb) * bb;
        aa = bb;
        bb = c;
        eps *= 2;
    }
    return aa * a;
}

struct axis {
    double length, res, align, shift;
    int cells, points;
    bool solid;
    double units(double x) const          // fd/discretization.h:143-149
    {
        double s = shift, r = res;
        double y = x - s / r;
        if (!solid) y = r_modulo(y, length);
        return y * r + s;
    }
    int index(int i) const { return solid ? i : (i + cells) % cells; }   // :151-156
};

axis from_orc(const orc_axis& a)
{
    axis x;
    x.length = a.length; x.res = a.resolution; x.align = a.alignment; x.shift = a.shift;
    x.cells = a.cells; x.points = a.points; x.solid = a.solid != 0;
    return x;
}

// ib/cosine.h, ib/peskin.h, ib/bspline.h
double phi_cosine(double r) { return 0.25 * (1 + std::cos(M_PI_2 * r)); }
double phi_peskin(double r)
{
    double s = std::fabs(r);
    return s < 1 ? (3 - 2 * s + std::sqrt(1 + 4 * s - 4 * s * s)) / 8
                 : (5 - 2 * s - std::sqrt(-7 + 12 * s - 4 * s * s)) / 8;
}
double phi_bspline(int n, double r)
{
    if (n == 0) return (-0.5 <= r && r < 0.5) ? 1.0 : 0.0;
    double b = (n + 1.0) / (2.0 * n);
    double a = r / n;
    return (b + a) * phi_bspline(n - 1, r + 0.5) + (b - a) * phi_bspline(n - 1, r - 0.5);
}
double phi(int kind, double r)
{
    if (kind == 0) return phi_cosine(r);
    if (kind == 1) return phi_peskin(r);
    return phi_bspline(3, r);
}

// every_other_pattern<3, 4> (ib/delta.h:131-146): offset + i -> shift vector.
void pattern(int n, int s[3])
{
    int x = n / 8, y = n % 8;
    for (int j = 0; j < 3; ++j) {
        s[j] = x % 2 + 2 * (y % 2) - 1;
        x /= 2;
        y /= 2;
    }
}

struct grid3 { axis a[3]; };

// ib::sweep::sort (ib/sweep.h:64-77) with the clamped sort index.
int sort_key(const grid3& g, const double* p)
{
    int index = 0, weight = 1;
    for (int j = 0; j < 3; ++j) {
        const axis& c = g.a[j];
        int solid = c.solid;
        int i = (int) r_floor(c.units(p[j]) - c.shift);
        int jj = c.index(i);
        int w = c.points + solid;
        int k = jj + solid;
        k = k < 0 ? 0 : (k >= w ? w - 1 : k);
        index += weight * k;        // combine, x fastest (ib/indexing.h:27-33)
        weight *= w;
    }
    return index;
}

// ib::sweep::difference (ib/sweep.h:49-62)
void difference(const grid3& g, const double* p, double d[3])
{
    for (int j = 0; j < 3; ++j) {
        double diff = g.a[j].units(p[j]) - g.a[j].shift;
        d[j] = r_floor(diff) - diff;
    }
}

// sorter::decompose (ib/indexing.h:103-115)
void decompose(const grid3& g, int key, int dec[3])
{
    for (int j = 0; j < 3; ++j) {
        int solid = g.a[j].solid;
        int w = g.a[j].points + solid;
        dec[j] = key % w - solid;
        key /= w;
    }
}

// ib::sweep::grid with errors_negative (spread) or clamped (interp).
int grid_index(const grid3& g, const int dec[3], const int s[3], bool clamp)
{
    int index = 0, weight = 1;
    for (int j = 0; j < 3; ++j) {
        const axis& c = g.a[j];
        int jj = c.index(dec[j] + s[j]);
        if (clamp) {
            jj = jj < 0 ? 0 : (jj >= c.points ? c.points - 1 : jj);
        } else if (jj < 0 || jj >= c.points || index < 0) {
            index = -1;
        }
        if (index >= 0) index += weight * jj;
        weight *= c.points;
    }
    return index;
}

double weight_of(int kind, const double d[3], const int s[3])
{
    // (v * ... * 1) over the dimensions (ib/sweep.h:107-109)
    double v0 = phi(kind, d[0] + s[0]), v1 = phi(kind, d[1] + s[1]), v2 = phi(kind, d[2] + s[2]);
    return v0 * (v1 * (v2 * 1));
}

long double phi_ld(int kind, long double r)
{
    if (kind == 0) {
        if (fabsl(r) >= 2) return 0;       // the kernel is only used on its 4-point support
        return 0.25L * (1 + cosl(3.141592653589793238462643383279502884L / 2 * r));
    }
    if (kind == 1) {
        long double s = fabsl(r);
        if (s >= 2) return 0;
        return s < 1 ? (3 - 2 * s + sqrtl(1 + 4 * s - 4 * s * s)) / 8
                     : (5 - 2 * s - sqrtl(-7 + 12 * s - 4 * s * s)) / 8;
    }
    long double s = fabsl(r);
    if (s >= 2) return 0;
    if (s < 1) return 2.0L / 3 - s * s + s * s * s / 2;
    long double t = 2 - s;
    return t * t * t / 6;
}

} // namespace

extern "C" {

// Grid axes as fd::grid builds them: resolution = refinement / GCD(lengths).
int orc_grid_axes(const orc_dim* dims, int ndims, int refinement, const double* alignment, orc_axis* out)
{
    double unit = dims[0].length;
    for (int i = 1; i < ndims; ++i) unit = r_gcd(unit, dims[i].length);
    double res = refinement / unit;
    for (int i = 0; i < ndims; ++i) {
        const orc_dim& d = dims[i];
        orc_axis& a = out[i];
        std::memset(&a, 0, sizeof(a));
        double al = r_modulo(alignment[i], 1.0);
        bool on_boundary = !al;
        a.length = d.length;
        a.resolution = res;
        a.alignment = al;
        a.shift = al + (d.solid && on_boundary);
        a.cells = (int) (d.length * res + 0.5);
        a.points = a.cells - (on_boundary && d.solid);
        a.solid = d.solid ? 1 : 0;
        if (d.solid) {
            const double* ps[2] = {d.lower, d.upper};
            double hs[2] = {1 / res, -1 / res};
            double ss[2] = {al, 1 - al};
            for (int e = 0; e < 2; ++e) {
                const double* p = ps[e];
                double h = hs[e];
                double s = 1 - r_modulo(1 - ss[e], 1);
                double den = p[1] - p[0] * s * h;
                a.boundary[e] = -h / den;
                a.interior[e] = (p[1] + p[0] * h * (1 - s)) / den;
                a.laplacian[e] = (-(p[1] * (1 - 2 * s) - p[0] * s * h * (1 - s))) / (2 * den);
            }
        }
        a.robin[0][0] = d.lower[0]; a.robin[0][1] = d.lower[1];
        a.robin[1][0] = d.upper[0]; a.robin[1][1] = d.upper[1];
    }
    return 0;
}

// ib::novel::spread for one marker set; `grids` is [3 components][3 axes].
void orc_spread(const orc_axis* grids, int kind, int n, const double* x, const double* f,
                double* fe0, double* fe1, double* fe2)
{
    double* fe[3] = {fe0, fe1, fe2};
    grid3 g[3];
    size_t size[3], maxsize = 0;
    for (int c = 0; c < 3; ++c) {
        size[c] = 1;
        for (int j = 0; j < 3; ++j) {
            g[c].a[j] = from_orc(grids[3 * c + j]);
            size[c] *= g[c].a[j].points;
        }
        maxsize = std::max(maxsize, size[c]);
    }
    std::vector<double> buf(8 * maxsize, 0.0);     // 8 plane buffers (ib/novel.h:180-191)
    std::vector<int> keys(n), perm(n), skeys(n), seg;
    std::vector<double> dd(3 * (size_t) n), fv(n);
    for (int c = 0; c < 3; ++c) {
        const grid3& gc = g[c];
        double res = gc.a[0].res * (gc.a[1].res * (gc.a[2].res * 1));
        // index(): keys and identity permutation, then sort by key
#pragma omp parallel for schedule(static)
        for (int t = 0; t < n; ++t) {
            double p[3] = {x[t], x[n + t], x[2 * n + t]};
            keys[t] = sort_key(gc, p);
            perm[t] = t;
        }
        std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return keys[a] < keys[b]; });
        seg.clear();
        for (int t = 0; t < n; ++t) {
            skeys[t] = keys[perm[t]];
            if (t == 0 || skeys[t] != skeys[t - 1]) seg.push_back(t);
        }
        seg.push_back(n);
        // sorted offsets and forces of this component
#pragma omp parallel for schedule(static)
        for (int t = 0; t < n; ++t) {
            int j = perm[t];
            double p[3] = {x[j], x[n + j], x[2 * n + j]};
            difference(gc, p, &dd[3 * (size_t) t]);
            fv[t] = f[(size_t) c * n + j];
        }
        // reduce(): 8 sweeps of 8 values each, reduce_by_key over equal
        // consecutive keys, then the plane writes.  Plane i only ever
        // receives value i of each sweep, so the planes are independent and
        // each keeps the sequential summation order of the reference.
        const int nseg = (int) seg.size() - 1;
#pragma omp parallel for schedule(static, 1)
        for (int i = 0; i < 8; ++i) {
            double* plane = &buf[(size_t) i * size[c]];
            for (int sweep = 0; sweep < 8; ++sweep) {
                int s[3];
                pattern(8 * sweep + i, s);
                for (int q = 0; q < nseg; ++q) {
                    int t = seg[q], e = seg[q + 1];
                    double r = weight_of(kind, &dd[3 * (size_t) t], s) * fv[t];
                    for (int u = t + 1; u < e; ++u) r += weight_of(kind, &dd[3 * (size_t) u], s) * fv[u];
                    int dec[3];
                    decompose(gc, skeys[t], dec);
                    int k = grid_index(gc, dec, s, false);
                    if (k >= 0) plane[k] += res * r;
                }
            }
        }
        // combine planes (ib/novel.h:153-162)
#pragma omp parallel for schedule(static)
        for (size_t k = 0; k < size[c]; ++k) {
            double acc = 0.0;
            for (int i = 0; i < 8; ++i) {
                acc += buf[(size_t) i * size[c] + k];
                buf[(size_t) i * size[c] + k] = 0.0;
            }
            fe[c][k] = acc;
        }
    }
}

// ib::interpolate::accumulate for each component (ib/interpolate.h:49-76).
void orc_interp(const orc_axis* grids, int kind, int n, const double* x,
                const double* u0, const double* u1, const double* u2, double* ul)
{
    const double* ue[3] = {u0, u1, u2};
    for (int c = 0; c < 3; ++c) {
        grid3 g;
        for (int j = 0; j < 3; ++j) g.a[j] = from_orc(grids[3 * c + j]);
#pragma omp parallel for schedule(static)
        for (int t = 0; t < n; ++t) {
            double p[3] = {x[t], x[n + t], x[2 * n + t]};
            int key = sort_key(g, p);
            int dec[3];
            decompose(g, key, dec);
            double d[3];
            difference(g, p, d);
            double v = 0.0;
            for (int i = 0; i < 64; ++i) {
                int s[3];
                pattern(i, s);
                int k = grid_index(g, dec, s, true);
                double w = weight_of(kind, d, s);
                if (k >= 0) v += w * ue[c][k];
            }
            ul[(size_t) c * n + t] = v;
        }
    }
}

// Direct long-double evaluation of the spread sum from its definition:
// fe_c(g) = prod_j r_j * sum_m f_c(m) prod_j phi(r_j X_j(g) - r_j x_j(m)),
// with grid point coordinates X_j(g) = (g_j + shift_j) / r_j
// (test/ib-whole.cc:104-116) and periodic images.  Markers must lie inside
// the box; solid axes drop stencil points outside [0, points).
void orc_spread_direct(const orc_axis* grids, int kind, int n, const double* x, const double* f,
                       double* fe0, double* fe1, double* fe2)
{
    double* fe[3] = {fe0, fe1, fe2};
    for (int c = 0; c < 3; ++c) {
        const orc_axis* a = grids + 3 * c;
        size_t size = (size_t) a[0].points * a[1].points * a[2].points;
        std::vector<long double> acc(size, 0.0L);
        long double res = (long double) a[0].resolution * a[1].resolution * a[2].resolution;
        for (int m = 0; m < n; ++m) {
            int lo[3];
            long double w[3][5];
            int idx[3][5];
            for (int j = 0; j < 3; ++j) {
                long double u = (long double) x[(size_t) j * n + m] * a[j].resolution;   // x / h
                int g0 = (int) floorl(u - a[j].shift) - 2;
                lo[j] = g0;
                for (int o = 0; o < 5; ++o) {
                    int gj = g0 + o;
                    long double r = (gj + (long double) a[j].shift) - u;
                    w[j][o] = phi_ld(kind, r);
                    int k = gj;
                    if (!a[j].solid) k = ((k % a[j].cells) + a[j].cells) % a[j].cells;
                    idx[j][o] = (k >= 0 && k < a[j].points) ? k : -1;
                }
            }
            (void) lo;
            long double fv = f[(size_t) c * n + m];
            for (int oz = 0; oz < 5; ++oz)
                for (int oy = 0; oy < 5; ++oy)
                    for (int ox = 0; ox < 5; ++ox) {
                        if (idx[0][ox] < 0 || idx[1][oy] < 0 || idx[2][oz] < 0) continue;
                        long double ww = w[0][ox] * w[1][oy] * w[2][oz];
                        if (ww == 0) continue;
                        size_t k = idx[0][ox] + (size_t) a[0].points * (idx[1][oy] + (size_t) a[1].points * idx[2][oz]);
                        acc[k] += res * fv * ww;
                    }
        }
        for (size_t k = 0; k < size; ++k) fe[c][k] = (double) acc[k];
    }
}

// Direct long-double interpolation: U_c(m) = sum_g u_c(g) prod_j phi(r X_j(g) - r x_j(m)).
void orc_interp_direct(const orc_axis* grids, int kind, int n, const double* x,
                       const double* u0, const double* u1, const double* u2, double* ul)
{
    const double* ue[3] = {u0, u1, u2};
    for (int c = 0; c < 3; ++c) {
        const orc_axis* a = grids + 3 * c;
        for (int m = 0; m < n; ++m) {
            long double w[3][5];
            int idx[3][5];
            for (int j = 0; j < 3; ++j) {
                long double u = (long double) x[(size_t) j * n + m] * a[j].resolution;
                int g0 = (int) floorl(u - a[j].shift) - 2;
                for (int o = 0; o < 5; ++o) {
                    int gj = g0 + o;
                    long double r = (gj + (long double) a[j].shift) - u;
                    w[j][o] = phi_ld(kind, r);
                    int k = gj;
                    if (!a[j].solid) k = ((k % a[j].cells) + a[j].cells) % a[j].cells;
                    else k = k < 0 ? 0 : (k >= a[j].points ? a[j].points - 1 : k);
                    idx[j][o] = k;
                }
            }
            long double v = 0;
            for (int oz = 0; oz < 5; ++oz)
                for (int oy = 0; oy < 5; ++oy)
                    for (int ox = 0; ox < 5; ++ox) {
                        long double ww = w[0][ox] * w[1][oy] * w[2][oz];
                        if (ww == 0) continue;
                        size_t k = idx[0][ox] + (size_t) a[0].points * (idx[1][oy] + (size_t) a[1].points * idx[2][oz]);
                        v += ww * ue[c][k];
                    }
            ul[(size_t) c * n + m] = (double) v;
        }
    }
}

// Scalar delta evaluation for the closed-form checks.
double orc_phi(int kind, double r) { return phi(kind, r); }

} // extern "C"
