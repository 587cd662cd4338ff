// Construction of the RBF surface operators of a cell type on the device.
//
// Restates the setup the reference runs once per cell type:
//   * the interpolation matrix [Phi P^T; P 0] over the data sites
//     (bases/fill.h:97-130): Phi = r^7 (bases/phs.h:13-46,
//     polyharmonic_spline<7>) of the chord distance r = sqrt(2 (1 - t)),
//     t = x_s . x_d, between the parameter-space points mapped to the unit
//     sphere (bases/shapes/sphere.h:18-54), P the polynomial augmentation;
//   * its LU factorisation and the operators solve(lu, [Phi_s P_s]^T) of the
//     sample sites (bases/operators.h:145-190, solvers/lu.h:14-48), for the
//     evaluator and the first and second parameter derivatives, with the
//     chain rule of bases/rbf.h:29-48:
//         d_a r^7  = 7 r^5 (r r_a),                     r r_a = -t_a
//         d_ab r^7 = 35 r^3 (r r_a)(r r_b) + 7 r^5 (-t_ab);
//   * the parameter-space quadrature weights of closed_surface::weights
//     (bases/closed_surface.h:38-68): the r^1 system with a constant
//     augmentation and right-hand side (0, 1), divided by the sphere's
//     Jacobian cos(phi) / (4 pi) (bases/shapes/sphere.h:101-115).
// The augmentation spans the spherical harmonics of degree <= d
// (spherical_harmonics<5> for RBCs, include/rbc.h:17; polynomials<0> for
// platelets, include/platelet.h:17) through the equal-span monomials
// x^a y^b z^c, c <= 1, a + b + c <= d, of the unit-sphere embedding (the
// interpolant depends on the span only).
//
// The fills are elementwise kernels; the dense LU and the solves are
// rocSOLVER's getrf / getrs (the reference uses cuSOLVER's).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <cmath>
#include <vector>
#include "context.hpp"

namespace ibm {
namespace {

constexpr int MAX_DEGREE = 8;
constexpr int MAX_TERMS = (MAX_DEGREE + 1) * (MAX_DEGREE + 1);

struct poly_exps {
    int count;
    unsigned char e[MAX_TERMS][3];
};

// The monomial exponents in the order of ib.cu_amd/surfaces.py: c = 0 then
// 1; total degree rising; a falling.
poly_exps exponents(int degree)
{
    poly_exps p{};
    for (int c = 0; c <= 1; ++c)
        for (int tot = c; tot <= degree; ++tot)
            for (int a = tot - c; a >= 0; --a) {
                p.e[p.count][0] = (unsigned char) a;
                p.e[p.count][1] = (unsigned char) (tot - c - a);
                p.e[p.count][2] = (unsigned char) c;
                ++p.count;
            }
    return p;
}

__device__ inline double ipow(double v, int k)
{
    double r = 1.0;
    for (int i = 0; i < k; ++i) r *= v;
    return r;
}

// d^n/dv ... of x^a y^b z^c for the coordinate counts cnt[3].
__device__ inline double dmono(const unsigned char* e, const double X[3], int c0, int c1, int c2)
{
    const int cnt[3] = {c0, c1, c2};
    double coef = 1.0, val = 1.0;
    for (int i = 0; i < 3; ++i) {
        int k = e[i];
        for (int j = 0; j < cnt[i]; ++j) coef *= (double) (k - j);
        if (coef == 0.0) return 0.0;
        val *= ipow(X[i], k - cnt[i]);
    }
    return coef * val;
}

// Embedding x = (cos t cos p, sin t cos p, sin p) and its parameter
// derivatives: d[0] = d/dt, d[1] = d/dp; dd[0] = tt, dd[1] = tp, dd[2] = pp.
struct embed {
    double X[3], d[2][3], dd[3][3];
};

__device__ inline embed embedding(double t, double p)
{
    embed m;
    double ct = cos(t), st = sin(t), cp = cos(p), sp = sin(p);
    m.X[0] = ct * cp; m.X[1] = st * cp; m.X[2] = sp;
    m.d[0][0] = -m.X[1]; m.d[0][1] = m.X[0]; m.d[0][2] = 0.0;
    m.d[1][0] = -ct * sp; m.d[1][1] = -st * sp; m.d[1][2] = cp;
    m.dd[0][0] = -m.X[0]; m.dd[0][1] = -m.X[1]; m.dd[0][2] = 0.0;
    m.dd[1][0] = st * sp; m.dd[1][1] = -ct * sp; m.dd[1][2] = 0.0;
    m.dd[2][0] = -m.X[0]; m.dd[2][1] = -m.X[1]; m.dd[2][2] = -m.X[2];
    return m;
}

// Derivative kinds: 0 value, 1 d/dt, 2 d/dp, 3 tt, 4 tp, 5 pp.
__device__ inline void kind_axes(int kind, int& a, int& b)
{
    a = kind == 1 || kind == 3 || kind == 4 ? 0 : 1;
    b = kind == 3 ? 0 : 1;
}

__device__ inline double poly_value(const unsigned char* e, const embed& m, int kind)
{
    if (kind == 0) return dmono(e, m.X, 0, 0, 0);
    if (kind <= 2) {
        const double* d = m.d[kind - 1];
        return dmono(e, m.X, 1, 0, 0) * d[0] + dmono(e, m.X, 0, 1, 0) * d[1] + dmono(e, m.X, 0, 0, 1) * d[2];
    }
    int a, b;
    kind_axes(kind, a, b);
    const double* dab = m.dd[kind - 3];
    double v = dmono(e, m.X, 1, 0, 0) * dab[0] + dmono(e, m.X, 0, 1, 0) * dab[1] + dmono(e, m.X, 0, 0, 1) * dab[2];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            int c[3] = {0, 0, 0};
            c[i] += 1;
            c[j] += 1;
            v += dmono(e, m.X, c[0], c[1], c[2]) * m.d[a][i] * m.d[b][j];
        }
    return v;
}

// r^7 of the chord distance between sample (ts, ps) and data (td, pd), or
// its derivative `kind` in the sample parameters.
__device__ inline double phs7(double ts, double ps, double td, double pd, int kind, int power)
{
    double dth = ts - td;
    double cs = cos(dth), ss = sin(dth);
    double cps = cos(ps), sps = sin(ps), cpd = cos(pd), spd = sin(pd);
    double t = cs * cps * cpd + sps * spd;
    double r = sqrt(2.0 * fmax(1.0 - t, 0.0));
    if (power == 1) return r;                  // the r^1 kernel of the quadrature weights
    double r2 = r * r, r3 = r2 * r, r5 = r3 * r2;
    if (kind == 0) return r5 * r2;
    double dt[2] = {-ss * cps * cpd, -cs * sps * cpd + cps * spd};
    if (kind <= 2) return 7.0 * r5 * (-dt[kind - 1]);
    double d2t = kind == 3 ? -cs * cps * cpd : (kind == 4 ? ss * sps * cpd : -t);
    int a, b;
    kind_axes(kind, a, b);
    return 35.0 * r3 * dt[a] * dt[b] - 7.0 * r5 * d2t;
}

// A = [Phi P^T; P 0] (column-major, lda = nd + np) of the data sites.
__global__ void fill_interp_kernel(int nd, const double* __restrict__ prm, poly_exps P, int power,
                                   double* __restrict__ A)
{
    const int n = nd + P.count;
    const size_t total = (size_t) n * n;
    for (size_t g = (size_t) blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t) gridDim.x * blockDim.x) {
        const int i = (int) (g % n), j = (int) (g / n);      // row i, column j
        double v = 0.0;
        if (i < nd && j < nd) {
            v = phs7(prm[i], prm[nd + i], prm[j], prm[nd + j], 0, power);
        } else if (i < nd && j >= nd) {
            embed m = embedding(prm[i], prm[nd + i]);
            v = poly_value(P.e[j - nd], m, 0);
        } else if (i >= nd && j < nd) {
            embed m = embedding(prm[j], prm[nd + j]);
            v = poly_value(P.e[i - nd], m, 0);
        }
        A[g] = v;
    }
}

// F = [Phi_s; P_s] (column-major (nd + np) x ns): column s holds the kernel
// `kind` between sample s and every data site, then the monomials at s.
__global__ void fill_rhs_kernel(int nd, const double* __restrict__ dprm, int ns, const double* __restrict__ sprm,
                                poly_exps P, int kind, double* __restrict__ F)
{
    const int n = nd + P.count;
    const size_t total = (size_t) n * ns;
    for (size_t g = (size_t) blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t) gridDim.x * blockDim.x) {
        const int i = (int) (g % n), s = (int) (g / n);
        double v;
        if (i < nd) {
            v = phs7(sprm[s], sprm[ns + s], dprm[i], dprm[nd + i], kind, 7);
        } else {
            embed m = embedding(sprm[s], sprm[ns + s]);
            v = poly_value(P.e[i - nd], m, kind);
        }
        F[g] = v;
    }
}

// weights w_i = c_i / (cos(phi_i) / (4 pi))
__global__ void sphere_scale_kernel(int n, const double* __restrict__ prm, const double* __restrict__ c,
                                    double* __restrict__ w)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    w[i] = c[i] / (cos(prm[n + i]) / (4.0 * M_PI));
}

__global__ void unit_rhs_kernel(int n, double* __restrict__ b)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    b[i] = i == n ? 1.0 : 0.0;
}

int grid_of(size_t total) { return (int) std::min<size_t>((total + 255) / 256, 65536); }

struct solver_handle {
    rocblas_handle h = nullptr;
    ~solver_handle() { if (h) rocblas_destroy_handle(h); }
};

ibm_status rb(ibm_ctx* ctx, rocblas_status s, const char* what)
{
    if (s == rocblas_status_success) return IBM_OK;
    return fail(ctx, IBM_ERR_HIP, "%s: rocBLAS / rocSOLVER status %d", what, (int) s);
}

// LU of A (n x n, device) in place; pivots and info in `piv` (n + 1 ints).
ibm_status lu_factor(ibm_ctx* ctx, rocblas_handle h, int n, double* A, int* piv)
{
    IBM_TRY(rb(ctx, rocsolver_dgetrf(h, n, n, A, n, piv, piv + n), "getrf"));
    int info = 0;
    IBM_HIP(ctx, hipMemcpyAsync(&info, piv + n, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    IBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (info != 0) return fail(ctx, IBM_ERR_ARG, "singular interpolation matrix (getrf info %d): repeated sites?", info);
    return IBM_OK;
}

} // namespace
} // namespace ibm

extern "C" {

ibm_status ibm_rbf_build(ibm_ctx* ctx, int nd, const double* data, int ns, const double* sample, int degree,
                         int nkinds, const int* kinds, double* const* E)
{
    using namespace ibm;
    if (!ctx) return IBM_ERR_ARG;
    if (nd <= 0 || ns <= 0 || !data || !sample || !kinds || !E || nkinds <= 0)
        return fail(ctx, IBM_ERR_ARG, "bad RBF construction arguments");
    if (degree < 0 || degree > MAX_DEGREE) return fail(ctx, IBM_ERR_ARG, "augmentation degree must be 0..%d", MAX_DEGREE);
    for (int k = 0; k < nkinds; ++k)
        if (kinds[k] < IBM_RBF_VALUE || kinds[k] > IBM_RBF_D_PP || !E[k])
            return fail(ctx, IBM_ERR_ARG, "bad operator kind or output %d", k);
    const poly_exps P = exponents(degree);
    const int n = nd + P.count;
    double* A = nullptr;
    int* piv = nullptr;
    IBM_HIP(ctx, hipMalloc(&A, sizeof(double) * (size_t) n * n));
    struct freer {
        void* p;
        ~freer() { if (p) (void) hipFree(p); }
    } fa{A};
    IBM_HIP(ctx, hipMalloc(&piv, sizeof(int) * (size_t) (n + 1)));
    freer fp{piv};
    solver_handle sh;
    IBM_TRY(rb(ctx, rocblas_create_handle(&sh.h), "rocblas_create_handle"));
    IBM_TRY(rb(ctx, rocblas_set_stream(sh.h, ctx->stream), "rocblas_set_stream"));
    fill_interp_kernel<<<grid_of((size_t) n * n), 256, 0, ctx->stream>>>(nd, data, P, 7, A);
    IBM_LAUNCH_CHECK(ctx);
    IBM_TRY(lu_factor(ctx, sh.h, n, A, piv));
    for (int k = 0; k < nkinds; ++k) {
        fill_rhs_kernel<<<grid_of((size_t) n * ns), 256, 0, ctx->stream>>>(nd, data, ns, sample, P, kinds[k], E[k]);
        IBM_LAUNCH_CHECK(ctx);
        IBM_TRY(rb(ctx, rocsolver_dgetrs(sh.h, rocblas_operation_none, n, ns, A, n, piv, E[k], n), "getrs"));
    }
    IBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return IBM_OK;
}

ibm_status ibm_sphere_weights(ibm_ctx* ctx, int n, const double* params, double* w)
{
    using namespace ibm;
    if (!ctx) return IBM_ERR_ARG;
    if (n <= 0 || !params || !w) return fail(ctx, IBM_ERR_ARG, "bad quadrature arguments");
    poly_exps P{};
    P.count = 1;                                  // the constant, polynomials<0>
    const int m = n + 1;
    double *A = nullptr, *b = nullptr;
    int* piv = nullptr;
    IBM_HIP(ctx, hipMalloc(&A, sizeof(double) * (size_t) m * m));
    struct freer {
        void* p;
        ~freer() { if (p) (void) hipFree(p); }
    } fa{A};
    IBM_HIP(ctx, hipMalloc(&b, sizeof(double) * (size_t) m));
    freer fb{b};
    IBM_HIP(ctx, hipMalloc(&piv, sizeof(int) * (size_t) (m + 1)));
    freer fp{piv};
    solver_handle sh;
    IBM_TRY(rb(ctx, rocblas_create_handle(&sh.h), "rocblas_create_handle"));
    IBM_TRY(rb(ctx, rocblas_set_stream(sh.h, ctx->stream), "rocblas_set_stream"));
    fill_interp_kernel<<<grid_of((size_t) m * m), 256, 0, ctx->stream>>>(n, params, P, 1, A);
    IBM_LAUNCH_CHECK(ctx);
    IBM_TRY(lu_factor(ctx, sh.h, m, A, piv));
    unit_rhs_kernel<<<(m + 255) / 256, 256, 0, ctx->stream>>>(n, b);
    IBM_LAUNCH_CHECK(ctx);
    IBM_TRY(rb(ctx, rocsolver_dgetrs(sh.h, rocblas_operation_none, m, 1, A, m, piv, b, m), "getrs"));
    sphere_scale_kernel<<<(n + 255) / 256, 256, 0, ctx->stream>>>(n, params, b, w);
    IBM_LAUNCH_CHECK(ctx);
    IBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return IBM_OK;
}

} // extern "C"
