// Chebyshev smoothing, geometric multigrid and Chebyshev-preconditioned CG on
// the matrix-free operators A = sigma * I + tau * L.
//
// Restates:
//   algo::chebyshev + chebw<4>            algo/chebyshev.h:120-169
//   algo::gershgorin                      algo/gershgorin.h:14-50
//   solvers::mg::chebyshev (interval)     solvers/mg/chebyshev.h:14-40
//   solvers::multigrid (levels, V-cycle,  solvers/multigrid.h:88-195
//     nested iteration, direct solver)
//   solvers::pcg / chebpcg                solvers/pcg.h:11-44, solvers/chebpcg.h:9-23,
//                                         solvers/preconditioners/chebyshev.h:9-26
// Design choices (the reference does not pin them):
//   * transfers act per physical axis (x fastest, like the grid vectors); the
//     reference's kron(kron(Px, Py), Pz) applies Pz along the fastest index,
//     which coincides with this only when the x and z axes share a staggering;
//   * stopping rules are relative: ||r|| <= rtol ||b|| (the reference stops on
//     absolute norms, solvers/multigrid.h:96, solvers/pcg.h:21);
//   * mixed precision: FP32 V-cycles inside an FP64 defect correction
//     x += V(b - A x), with the residual, its norm and the update in FP64.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <atomic>
#include <cmath>
#include <vector>
#include "context.hpp"
#include "fd_kernels.hpp"

namespace ibm {

// chebw<4> (algo/chebyshev.h:120-139): weights p_i(a, b) / q(a, b) with
// polynomial(c...)(a, b) = sum_i c_i a^i b^(n-1-i).
static double poly(const std::vector<double>& c, double a, double b)
{
    double v = 0;
    int n = (int) c.size();
    for (int i = 0; i < n; ++i) {
        double ai = 1, bi = 1;
        for (int k = 0; k < i; ++k) ai *= a;
        for (int k = 0; k < n - 1 - i; ++k) bi *= b;
        v += c[i] * ai * bi;
    }
    return v;
}

void chebyshev_weights(double a, double b, double w[4])
{
    double q = poly({1, 28, 70, 28, 1}, a, b);
    w[0] = poly({128}, a, b) / q;
    w[1] = poly({256, 256}, a, b) / q;
    w[2] = poly({160, 448, 160}, a, b) / q;
    w[3] = poly({32, 224, 224, 32}, a, b) / q;
}

// solvers/mg/chebyshev.h:14-37: smoothing interval [a + (b - a) * prop, b]
// with prop = prod_d (min_length / 2) / length_d, after ordering |a| <= |b|.
static void smoother_interval(const ibm_grid& g, double sigma, double tau, double* a, double* b)
{
    double lo, hi;
    gershgorin_bounds(g, sigma, tau, &lo, &hi);
    if (std::fabs(lo) > std::fabs(hi)) std::swap(lo, hi);
    double smallest = INFINITY;
    for (int d = 0; d < g.dims; ++d) smallest = std::fmin(smallest, g.axis[d].length);
    smallest /= 2;
    double prop = 1.0;
    for (int d = g.dims - 1; d >= 0; --d) prop = (smallest / g.axis[d].length) * prop;
    *a = lo + (hi - lo) * prop;
    *b = hi;
}

// solvers/multigrid.h:153-172: an axis can be coarsened if it has an even
// number of cells, more than 2 if periodic and more than 4 if walled.
static bool refinable(const ibm_grid& g)
{
    for (int d = 0; d < g.dims; ++d) {
        int c = g.axis[d].cells;
        if (c & 1) return false;
        if (!g.axis[d].solid ? c <= 2 : c <= 4) return false;
    }
    return g.refinement >= 2;
}

template <typename T>
struct level_bufs {
    T *x = nullptr, *b = nullptr, *r = nullptr, *s1 = nullptr, *s2 = nullptr;
};

struct mg_level {
    ibm_grid g;
    size_t n;
    double w[4];
    op3<double> A64;
    op3<float> A32;
    transfer3<double> T64;   // this level <-> next coarser
    transfer3<float> T32;
};

} // namespace ibm

struct ibm_mg {
    int precision = 64;
    double sigma = 0, tau = 0;
    bool singular = false;
    std::vector<ibm::mg_level> lv;
    std::vector<ibm::level_bufs<double>> b64;
    std::vector<ibm::level_bufs<float>> b32;
    double* r = nullptr;        // outer FP64 residual
    double* rhs = nullptr;      // projected right-hand side (singular operators)
    double* partial = nullptr;  // reduction scratch
    double* scalars = nullptr;  // device scalars
    std::vector<void*> owned;
    size_t coarse_start = 0;    // first level handled by the one-workgroup coarse kernel (lv.size(): none)
    ibm::coarse_levels<double> coarse64{};
    ibm::coarse_levels<float> coarse32{};
};

namespace ibm {

template <typename T>
static op3<T> level_op(const mg_level& l);
template <>
op3<double> level_op<double>(const mg_level& l) { return l.A64; }
template <>
op3<float> level_op<float>(const mg_level& l) { return l.A32; }
template <typename T>
static transfer3<T> level_tr(const mg_level& l);
template <>
transfer3<double> level_tr<double>(const mg_level& l) { return l.T64; }
template <>
transfer3<float> level_tr<float>(const mg_level& l) { return l.T32; }
template <typename T>
static std::vector<level_bufs<T>>& bufs(ibm_mg* m);
template <>
std::vector<level_bufs<double>>& bufs<double>(ibm_mg* m) { return m->b64; }
template <>
std::vector<level_bufs<float>>& bufs<float>(ibm_mg* m) { return m->b32; }

// Smoothing in one z-marching kernel (launch_cheb_fused) or as three stencil
// passes.
constexpr bool FUSED_SMOOTHING = true;
// Residual and restriction of the down-leg in one kernel
// (launch_residual_restrict) or as a stencil pass and a restriction.
constexpr bool FUSED_RESTRICT = true;
// Prolongation and residual of the up-leg in one kernel
// (launch_prolong_residual) or as a prolongation and a stencil pass.
constexpr bool FUSED_PROLONG = true;

// out = p(A) v, the degree-4 Chebyshev polynomial by Horner's rule
// (algo/chebyshev.h:149-163): s0 = w0 v, s_k = w_k v - A s_{k-1}.
// With add = true, out += p(A) v instead (out must not alias v, s1, s2).
// With acc (add mode only), the final out + p(A) v is added to the FP64
// vector acc instead of being stored in out.
template <typename T>
static ibm_status cheb_apply(ibm_ctx* ctx, const op3<T>& A, const double w[4], const T* v, T* s1, T* s2,
                             T* out, bool add, double* acc = nullptr)
{
    if (FUSED_SMOOTHING && cheb_fused_fits(A.a[0].n, A.a[1].n))
        return launch_cheb_fused<T>(ctx, A, w, v, out, add, acc);
    IBM_TRY(launch_stencil<T>(ctx, A, (T) -w[0], v, 0, nullptr, (T) w[1], v, s1));
    IBM_TRY(launch_stencil<T>(ctx, A, (T) -1, s1, 0, nullptr, (T) w[2], v, s2));
    if (add && acc) return launch_stencil_acc<T>(ctx, A, (T) -1, s2, (T) 1, out, (T) w[3], v, acc);
    if (add) return launch_stencil<T>(ctx, A, (T) -1, s2, (T) 1, out, (T) w[3], v, out);
    return launch_stencil<T>(ctx, A, (T) -1, s2, 0, nullptr, (T) w[3], v, out);
}

// solvers/multigrid.h:120-128 (iterative levels) and :46-63 (coarsest):
//   x = S b; r = b - A x; x += P vcycle(R r); x += S(b - A x)
template <typename T>
static const coarse_levels<T>& coarse_desc(const ibm_mg* m);
template <>
const coarse_levels<double>& coarse_desc<double>(const ibm_mg* m) { return m->coarse64; }
template <>
const coarse_levels<float>& coarse_desc<float>(const ibm_mg* m) { return m->coarse32; }

// The levels from coarse_start down, for one precision.
template <typename T>
static void fill_coarse(ibm_mg* m, coarse_levels<T>& C)
{
    C.count = (int) (m->lv.size() - m->coarse_start);
    C.lds_bytes = 0;
    for (int i = 0; i < C.count; ++i) {
        const mg_level& L = m->lv[m->coarse_start + i];
        C.n[i] = (int) L.n;
        C.off[i] = (int) (C.lds_bytes / sizeof(T));
        C.A[i] = level_op<T>(L);
        if (m->coarse_start + i + 1 < m->lv.size()) C.R[i] = level_tr<T>(L);
        for (int k = 0; k < 4; ++k) C.w[i][k] = (T) L.w[k];
        C.lds_bytes += COARSE_VECTORS * sizeof(T) * L.n;
    }
    level_bufs<T>& B0 = bufs<T>(m)[m->coarse_start];
    C.b0 = B0.b;
    C.x0 = B0.x;
}

// With acc, the level's final correction is added to the FP64 vector acc
// (used on level 0 when fused(), see below) instead of being stored in x.
template <typename T>
static ibm_status vcycle(ibm_ctx* ctx, ibm_mg* m, size_t l, double* acc = nullptr)
{
    if (l == m->coarse_start) return launch_coarse_vcycle<T>(ctx, coarse_desc<T>(m));
    const mg_level& L = m->lv[l];
    level_bufs<T>& B = bufs<T>(m)[l];
    op3<T> A = level_op<T>(L);
    IBM_TRY(cheb_apply<T>(ctx, A, L.w, B.b, B.s1, B.s2, B.x, false));
    if (l + 1 == m->lv.size()) return IBM_OK;
    level_bufs<T>& C = bufs<T>(m)[l + 1];
    if (FUSED_RESTRICT) {
        IBM_TRY(launch_residual_restrict<T>(ctx, A, level_tr<T>(L), B.x, B.b, C.b));
    } else {
        IBM_TRY(launch_stencil<T>(ctx, A, (T) -1, B.x, 0, nullptr, (T) 1, B.b, B.r));
        IBM_TRY(launch_restrict<T>(ctx, level_tr<T>(L), B.r, C.b));
    }
    IBM_TRY(vcycle<T>(ctx, m, l + 1));
    if (FUSED_PROLONG && FUSED_SMOOTHING && cheb_fused_fits(A.a[0].n, A.a[1].n)) {
        // xn = x + P e (into s1) and r = b - A xn in one pass; x = xn + S r
        IBM_TRY(launch_prolong_residual<T>(ctx, A, level_tr<T>(L), C.x, B.x, B.b, B.s1, B.r));
        return launch_cheb_fused<T>(ctx, A, L.w, B.r, B.x, true, acc, B.s1);
    }
    IBM_TRY(launch_prolong_add<T>(ctx, level_tr<T>(L), C.x, B.x));
    IBM_TRY(launch_stencil<T>(ctx, A, (T) -1, B.x, 0, nullptr, (T) 1, B.b, B.r));
    return cheb_apply<T>(ctx, A, L.w, B.r, B.s1, B.s2, B.x, true, acc);
}

// FP32 V-cycles whose level 0 runs the separate kernels (two levels or more,
// the coarse end not starting at level 0) fold the defect-correction update
// x += e into the last smoothing pass and produce the next FP32 right-hand
// side in the FP64 residual pass.
template <typename T>
static bool fused(const ibm_mg* m)
{
    return sizeof(T) == 4 && m->lv.size() > 1 && m->coarse_start > 0;
}

// One defect-correction iteration: x += V(r), r = b - A x; with fused(),
// B0.b already holds r in FP32 and is refreshed by the residual pass, which
// skips the FP64 copy of r unless keep_r.
template <typename T>
static ibm_status correct(ibm_ctx* ctx, ibm_mg* m, const double* b, double* x, bool keep_r = true)
{
    const mg_level& L0 = m->lv[0];
    const size_t n = L0.n;
    level_bufs<T>& B0 = bufs<T>(m)[0];
    if (fused<T>(m)) {
        IBM_TRY(vcycle<T>(ctx, m, 0, x));
        return launch_stencil_split(ctx, L0.A64, -1.0, x, 0.0, nullptr, 1.0, b, keep_r ? m->r : nullptr,
                                    (float*) B0.b);
    }
    IBM_TRY(launch_convert<T, double>(ctx, n, m->r, B0.b));
    IBM_TRY(vcycle<T>(ctx, m, 0));
    IBM_TRY(launch_axpby<double, T>(ctx, n, 1.0, B0.x, 1.0, x));
    return launch_stencil<double>(ctx, L0.A64, -1.0, x, 0.0, nullptr, 1.0, b, m->r);
}

static ibm_status read_scalar(ibm_ctx* ctx, const double* d, double* h)
{
    IBM_HIP(ctx, hipMemcpyAsync(h, d, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    IBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return IBM_OK;
}

static ibm_status norm2(ibm_ctx* ctx, ibm_mg* m, size_t n, const double* x, double* out)
{
    IBM_TRY(launch_dot<double, double>(ctx, n, x, x, m->partial, m->scalars));
    double v;
    IBM_TRY(read_scalar(ctx, m->scalars, &v));
    *out = std::sqrt(v);
    return IBM_OK;
}

// Initial residual: r = b from x = 0, or r = b - A x from the caller's x
// (warm start; on a singular operator x is made mean-free first).
static ibm_status start(ibm_ctx* ctx, ibm_mg* m, const double* b, double* x, bool warm)
{
    const mg_level& L0 = m->lv[0];
    const size_t n = L0.n;
    if (!warm) {
        IBM_TRY(launch_fill<double>(ctx, n, 0.0, x));
        return launch_convert<double, double>(ctx, n, b, m->r);
    }
    if (m->singular) IBM_TRY(launch_remove_mean<double>(ctx, n, x, m->partial, m->scalars + 1));
    return launch_stencil<double>(ctx, L0.A64, -1.0, x, 0.0, nullptr, 1.0, b, m->r);
}

template <typename T>
static ibm_status solve_loop(ibm_ctx* ctx, ibm_mg* m, const double* b, double* x, double rtol, double atol,
                             int maxit, bool warm, int* iters, double* relres)
{
    const mg_level& L0 = m->lv[0];
    const size_t n = L0.n;
    level_bufs<T>& B0 = bufs<T>(m)[0];
    if (m->singular) {
        // Solutions of a singular (periodic / Neumann) Laplacian are defined
        // up to a constant: solve with the mean-free part of b, return the
        // mean-free solution.
        IBM_TRY(launch_convert<double, double>(ctx, n, b, m->rhs));
        IBM_TRY(launch_remove_mean<double>(ctx, n, m->rhs, m->partial, m->scalars + 1));
        b = m->rhs;
    }
    double nb;
    IBM_TRY(norm2(ctx, m, n, b, &nb));
    IBM_TRY(start(ctx, m, b, x, warm));
    int it = 0;
    double nr = nb;
    if (warm) IBM_TRY(norm2(ctx, m, n, m->r, &nr));
    // A NaN or infinity in b, x or the operator makes the norms non-finite:
    // no cycle can fix that, and the comparisons below are written so that a
    // NaN fails them (nr <= stop is false for a NaN).
    if (!std::isfinite(nb) || !std::isfinite(nr)) {
        if (iters) *iters = 0;
        if (relres) *relres = NAN;
        return fail(ctx, IBM_ERR_NOT_CONVERGED, "multigrid: non-finite right-hand side or initial residual");
    }
    const double stop = std::max(rtol * nb, atol);
    if (fused<T>(m)) IBM_TRY(launch_convert<T, double>(ctx, n, m->r, B0.b));
    while (!(nr <= stop) && it < maxit) {
        IBM_TRY(correct<T>(ctx, m, b, x));
        ++it;
        IBM_TRY(norm2(ctx, m, n, m->r, &nr));
        if (!std::isfinite(nr)) break;
    }
    // A 1 = 0 on the singular operators, so the constant component of x does
    // not change the residuals above: it is projected out once, here.
    if (m->singular) IBM_TRY(launch_remove_mean<double>(ctx, n, x, m->partial, m->scalars + 1));
    if (iters) *iters = it;
    if (relres) *relres = nb > 0 ? nr / nb : (std::isfinite(nr) ? 0.0 : NAN);
    if (!(nr <= stop))
        return fail(ctx, IBM_ERR_NOT_CONVERGED, "multigrid: relative residual %.3e after %d cycles", nr / nb, it);
    return IBM_OK;
}

// The same defect correction with a fixed number of V-cycles and no host
// synchronisation (so it can be recorded in a HIP graph); the final
// relative residual ||b - A x|| / ||b|| is left in device memory.
template <typename T>
static ibm_status solve_fixed(ibm_ctx* ctx, ibm_mg* m, const double* b, double* x, int iters, bool warm,
                              double* relres_dev)
{
    const mg_level& L0 = m->lv[0];
    const size_t n = L0.n;
    level_bufs<T>& B0 = bufs<T>(m)[0];
    if (m->singular) {
        IBM_TRY(launch_convert<double, double>(ctx, n, b, m->rhs));
        IBM_TRY(launch_remove_mean<double>(ctx, n, m->rhs, m->partial, m->scalars + 1));
        b = m->rhs;
    }
    IBM_TRY(start(ctx, m, b, x, warm));
    if (fused<T>(m) && iters > 0) IBM_TRY(launch_convert<T, double>(ctx, n, m->r, B0.b));
    // Only the last residual is read (for relres_dev) in FP64.
    for (int it = 0; it < iters; ++it) IBM_TRY(correct<T>(ctx, m, b, x, relres_dev && it == iters - 1));
    if (m->singular) IBM_TRY(launch_remove_mean<double>(ctx, n, x, m->partial, m->scalars + 1));
    if (relres_dev) {
        IBM_TRY(launch_dot<double, double>(ctx, n, b, b, m->partial, m->scalars + 2));
        IBM_TRY(launch_dot<double, double>(ctx, n, m->r, m->r, m->partial, m->scalars + 3));
        IBM_TRY(launch_sqrt_ratio(ctx, m->scalars + 3, m->scalars + 2, relres_dev));
    }
    return IBM_OK;
}

// Several independent fixed-count solves (each on its own context, so on
// its own stream), issued cycle by cycle in turn: cycle k of every solve
// before cycle k + 1 of any.  Each solve runs exactly the kernels of
// solve_fixed; only the issue order changes.  With events, the streams also
// join into ctxs[0]'s and fork again after every cycle, so that a graph
// captured from them consists of one short parallel section per cycle: a
// graph executor that submits each parallel branch whole before the next
// then still keeps every stream supplied.
template <typename T>
static ibm_status solve_fixed_multi(int count, ibm_ctx* const* ctxs, ibm_mg* const* ms, const double* const* bs,
                                    double* const* xs, const int* iters, bool warm, double* const* relres_dev,
                                    hipEvent_t fork_ev, const hipEvent_t* join_ev)
{
    auto rejoin = [&]() -> ibm_status {
        if (!fork_ev || count < 2) return IBM_OK;
        for (int k = 1; k < count; ++k) {
            IBM_HIP(ctxs[0], hipEventRecord(join_ev[k - 1], ctxs[k]->stream));
            IBM_HIP(ctxs[0], hipStreamWaitEvent(ctxs[0]->stream, join_ev[k - 1], 0));
        }
        IBM_HIP(ctxs[0], hipEventRecord(fork_ev, ctxs[0]->stream));
        for (int k = 1; k < count; ++k) IBM_HIP(ctxs[0], hipStreamWaitEvent(ctxs[k]->stream, fork_ev, 0));
        return IBM_OK;
    };
    std::vector<const double*> b(bs, bs + count);
    int most = 0;
    for (int k = 0; k < count; ++k) {
        ibm_mg* m = ms[k];
        const size_t n = m->lv[0].n;
        if (m->singular) {
            IBM_TRY(launch_convert<double, double>(ctxs[k], n, b[k], m->rhs));
            IBM_TRY(launch_remove_mean<double>(ctxs[k], n, m->rhs, m->partial, m->scalars + 1));
            b[k] = m->rhs;
        }
        IBM_TRY(start(ctxs[k], m, b[k], xs[k], warm));
        if (fused<T>(m) && iters[k] > 0) IBM_TRY(launch_convert<T, double>(ctxs[k], n, m->r, bufs<T>(m)[0].b));
        most = std::max(most, iters[k]);
    }
    for (int it = 0; it < most; ++it) {
        for (int k = 0; k < count; ++k)
            if (it < iters[k])
                IBM_TRY(correct<T>(ctxs[k], ms[k], b[k], xs[k], relres_dev[k] && it == iters[k] - 1));
        if (it + 1 < most) IBM_TRY(rejoin());
    }
    for (int k = 0; k < count; ++k) {
        ibm_mg* m = ms[k];
        const size_t n = m->lv[0].n;
        if (m->singular) IBM_TRY(launch_remove_mean<double>(ctxs[k], n, xs[k], m->partial, m->scalars + 1));
        if (relres_dev[k]) {
            IBM_TRY(launch_dot<double, double>(ctxs[k], n, b[k], b[k], m->partial, m->scalars + 2));
            IBM_TRY(launch_dot<double, double>(ctxs[k], n, m->r, m->r, m->partial, m->scalars + 3));
            IBM_TRY(launch_sqrt_ratio(ctxs[k], m->scalars + 3, m->scalars + 2, relres_dev[k]));
        }
    }
    return IBM_OK;
}

ibm_status mg_solve_fixed_multi(int count, ibm_ctx* const* ctxs, ibm_mg* const* ms, const double* const* b,
                                double* const* x, const int* iters, bool warm, double* const* relres,
                                hipEvent_t fork_ev, const hipEvent_t* join_ev)
{
    if (fork_ev && (!join_ev || count < 2)) return IBM_ERR_ARG;
    for (int k = 1; fork_ev && k < count; ++k)
        if (ctxs[k] == ctxs[0] || !join_ev[k - 1]) return fail(ctxs[0], IBM_ERR_ARG, "fixed multi-solve: bad fork");
    if (count < 1 || !ctxs || !ms || !b || !x || !iters || !relres) return IBM_ERR_ARG;
    for (int k = 0; k < count; ++k) {
        if (!ctxs[k] || !ms[k] || !b[k] || !x[k] || b[k] == x[k] || iters[k] < 0 ||
            ms[k]->precision != ms[0]->precision)
            return fail(ctxs[0], IBM_ERR_ARG, "fixed multi-solve: bad entry %d", k);
    }
    if (ms[0]->precision == 64)
        return solve_fixed_multi<double>(count, ctxs, ms, b, x, iters, warm, relres, fork_ev, join_ev);
    return solve_fixed_multi<float>(count, ctxs, ms, b, x, iters, warm, relres, fork_ev, join_ev);
}

// Block-target autotuning of the z-marching kernels.  For every level above
// the coarse end whose grid the process has not timed yet, each kernel
// (smoothing, residual + restriction, prolongation + residual) runs on the
// level's zeroed buffers with z chunkings aimed at 256 ... 2048 blocks, and
// the fastest target is recorded for the grid.  The tile stays the one the
// grid size selects, so every point is computed by the same code whatever
// the target: tuning changes timings, never results.
static std::atomic<bool> autotune_on{true};
constexpr int TUNE_TARGETS[] = {256, 512, 1024, 2048};
constexpr int TUNE_REPS = 4;
constexpr float TUNE_MARGIN = 0.97f;   // a target must beat the default's time by 3%

static int per_of(const ibm_grid& g)
{
    return (g.axis[0].solid ? 0 : 1) | (g.axis[1].solid ? 0 : 2) | (g.axis[2].solid ? 0 : 4);
}

template <typename T>
static ibm_status tune_run(ibm_ctx* ctx, ibm_mg* m, size_t l, int kind)
{
    const mg_level& L = m->lv[l];
    level_bufs<T>& B = bufs<T>(m)[l];
    const op3<T> A = level_op<T>(L);
    if (kind == TUNE_CHEB) return launch_cheb_fused<T>(ctx, A, L.w, B.b, B.x, false, nullptr);
    level_bufs<T>& C = bufs<T>(m)[l + 1];
    if (kind == TUNE_RR) return launch_residual_restrict<T>(ctx, A, level_tr<T>(L), B.x, B.b, C.b);
    return launch_prolong_residual<T>(ctx, A, level_tr<T>(L), C.x, B.x, B.b, B.s1, B.r);
}

template <typename T>
static ibm_status tune_levels(ibm_ctx* ctx, ibm_mg* m)
{
    if (!autotune_on.load()) return IBM_OK;
    const int elem = (int) sizeof(T);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) hipEventDestroy(e0);
        return fail(ctx, IBM_ERR_HIP, "multigrid tuning: event creation failed");
    }
    const bool profiling = ctx->profiling;
    ctx->profiling = false;
    ibm_status st = IBM_OK;
    for (size_t l = 0; st == IBM_OK && l < m->coarse_start; ++l) {
        const ibm_grid& g = m->lv[l].g;
        const int nx = g.axis[0].points, ny = g.axis[1].points, nz = g.axis[2].points;
        if (!cheb_fused_fits(nx, ny)) continue;
        const int per = per_of(g);
        bool zeroed = false;
        for (int kind = TUNE_CHEB; st == IBM_OK && kind <= TUNE_PR; ++kind) {
            if (kind != TUNE_CHEB && l + 1 == m->lv.size()) continue;
            if (target_tuned(kind, elem, nx, ny, nz, per)) continue;
            if (!zeroed) {
                for (size_t k = l; st == IBM_OK && k <= std::min(l + 1, m->lv.size() - 1); ++k) {
                    const level_bufs<T>& B = bufs<T>(m)[k];
                    for (T* p : {B.x, B.b, B.r, B.s1, B.s2})
                        if (hipMemsetAsync(p, 0, m->lv[k].n * sizeof(T), ctx->stream) != hipSuccess)
                            st = fail(ctx, IBM_ERR_HIP, "multigrid tuning: clearing level %zu failed", k);
                }
                zeroed = true;
            }
            const int fallback = kind == TUNE_CHEB ? MARCH_DEFAULT_TARGET : TRANSFER_DEFAULT_TARGET;
            float ms[sizeof(TUNE_TARGETS) / sizeof(int)] = {};
            for (size_t c = 0; st == IBM_OK && c < sizeof(TUNE_TARGETS) / sizeof(int); ++c) {
                set_tuned_target(kind, elem, nx, ny, nz, per, TUNE_TARGETS[c]);
                if ((st = tune_run<T>(ctx, m, l, kind)) != IBM_OK) break;   // warm-up
                hipEventRecord(e0, ctx->stream);
                for (int r = 0; st == IBM_OK && r < TUNE_REPS; ++r) st = tune_run<T>(ctx, m, l, kind);
                hipEventRecord(e1, ctx->stream);
                if (st != IBM_OK) break;
                if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms[c], e0, e1) != hipSuccess)
                    st = fail(ctx, IBM_ERR_HIP, "multigrid tuning: timing failed");
            }
            // The default stays unless another target beats it by the margin;
            // then the fastest target wins.
            int best = fallback;
            float bar = 0;
            for (size_t c = 0; c < sizeof(TUNE_TARGETS) / sizeof(int); ++c)
                if (TUNE_TARGETS[c] == fallback) bar = TUNE_MARGIN * ms[c];
            for (size_t c = 0; c < sizeof(TUNE_TARGETS) / sizeof(int); ++c)
                if (ms[c] < bar) {
                    best = TUNE_TARGETS[c];
                    bar = ms[c];
                }
            set_tuned_target(kind, elem, nx, ny, nz, per, st == IBM_OK ? best : 0);
        }
    }
    ctx->profiling = profiling;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return st;
}

template <typename T>
static ibm_status alloc(ibm_ctx* ctx, ibm_mg* m, size_t n, T** p)
{
    void* v = nullptr;
    IBM_HIP(ctx, hipMalloc(&v, sizeof(T) * (n ? n : 1)));
    m->owned.push_back(v);
    *p = (T*) v;
    return IBM_OK;
}

static ibm_status check_grid(ibm_ctx* ctx, const ibm_grid* g)
{
    if (!g) return fail(ctx, IBM_ERR_ARG, "null grid");
    if (g->dims != 3) return fail(ctx, IBM_ERR_ARG, "operators support 3-dimensional grids (got %d)", g->dims);
    for (int d = 0; d < 3; ++d)
        if (g->axis[d].points < 1) return fail(ctx, IBM_ERR_BAD_GRID, "empty axis");
    return IBM_OK;
}

} // namespace ibm

using namespace ibm;

extern "C" {

ibm_status ibm_gershgorin(const ibm_grid* g, double sigma, double tau, double* lo, double* hi)
{
    if (!g || !lo || !hi) return IBM_ERR_ARG;
    gershgorin_bounds(*g, sigma, tau, lo, hi);
    return IBM_OK;
}

ibm_status ibm_chebyshev_weights(double a, double b, double* w)
{
    if (!w) return IBM_ERR_ARG;
    chebyshev_weights(a, b, w);
    return IBM_OK;
}

ibm_status ibm_op_apply(ibm_ctx* ctx, const ibm_grid* g, double sigma, double tau, double alpha,
                        const double* x, double beta, double* y)
{
    if (!ctx) return IBM_ERR_ARG;
    IBM_TRY(check_grid(ctx, g));
    if (!x || !y) return fail(ctx, IBM_ERR_ARG, "null vector");
    if (x == y) return fail(ctx, IBM_ERR_ARG, "x and y must not alias");
    return launch_stencil<double>(ctx, make_op<double>(*g, sigma, tau), alpha, x, beta, beta != 0 ? y : nullptr,
                                  0.0, nullptr, y);
}

ibm_status ibm_op_apply_f32(ibm_ctx* ctx, const ibm_grid* g, double sigma, double tau, float alpha,
                            const float* x, float beta, float* y)
{
    if (!ctx) return IBM_ERR_ARG;
    IBM_TRY(check_grid(ctx, g));
    if (!x || !y) return fail(ctx, IBM_ERR_ARG, "null vector");
    if (x == y) return fail(ctx, IBM_ERR_ARG, "x and y must not alias");
    return launch_stencil<float>(ctx, make_op<float>(*g, sigma, tau), alpha, x, beta, beta != 0 ? y : nullptr,
                                 0.0f, nullptr, y);
}

ibm_status ibm_chebyshev_apply(ibm_ctx* ctx, const ibm_grid* g, double sigma, double tau, double a, double b,
                               const double* r, double* out, double* work /* 2 * size */)
{
    if (!ctx) return IBM_ERR_ARG;
    IBM_TRY(check_grid(ctx, g));
    if (!r || !out || !work) return fail(ctx, IBM_ERR_ARG, "null vector");
    double w[4];
    chebyshev_weights(a, b, w);
    size_t n = (size_t) ibm_grid_size(g);
    return cheb_apply<double>(ctx, make_op<double>(*g, sigma, tau), w, r, work, work + n, out, false);
}

ibm_status ibm_mg_create(ibm_ctx* ctx, const ibm_grid* g, double sigma, double tau, int precision, ibm_mg** out)
{
    if (!ctx || !out) return IBM_ERR_ARG;
    *out = nullptr;
    IBM_TRY(check_grid(ctx, g));
    if (precision != 64 && precision != 32) return fail(ctx, IBM_ERR_ARG, "precision must be 64 or 32");
    auto* m = new ibm_mg;
    m->precision = precision;
    m->sigma = sigma;
    m->tau = tau;
    bool singular = sigma == 0;
    for (int d = 0; d < 3; ++d)
        if (g->axis[d].solid && (g->axis[d].interior[0] != 1.0 || g->axis[d].interior[1] != 1.0)) singular = false;
    m->singular = singular;
    ibm_grid cur = *g;
    ibm_status st = IBM_OK;
    while (true) {
        mg_level L;
        L.g = cur;
        L.n = (size_t) ibm_grid_size(&cur);
        double a, b;
        smoother_interval(cur, sigma, tau, &a, &b);
        chebyshev_weights(a, b, L.w);
        L.A64 = make_op<double>(cur, sigma, tau);
        L.A32 = make_op<float>(cur, sigma, tau);
        bool more = refinable(cur);
        ibm_grid next;
        if (more && ibm_grid_coarsen(&cur, &next) == IBM_OK) {
            L.T64 = make_transfer<double>(cur, next);
            L.T32 = make_transfer<float>(cur, next);
        } else {
            more = false;
        }
        m->lv.push_back(L);
        if (!more) break;
        cur = next;
    }
    for (auto& L : m->lv) {
        if (precision == 64) {
            level_bufs<double> B;
            if ((st = alloc(ctx, m, L.n, &B.x)) || (st = alloc(ctx, m, L.n, &B.b)) || (st = alloc(ctx, m, L.n, &B.r)) ||
                (st = alloc(ctx, m, L.n, &B.s1)) || (st = alloc(ctx, m, L.n, &B.s2)))
                break;
            m->b64.push_back(B);
        } else {
            level_bufs<float> B;
            if ((st = alloc(ctx, m, L.n, &B.x)) || (st = alloc(ctx, m, L.n, &B.b)) || (st = alloc(ctx, m, L.n, &B.r)) ||
                (st = alloc(ctx, m, L.n, &B.s1)) || (st = alloc(ctx, m, L.n, &B.s2)))
                break;
            m->b32.push_back(B);
        }
    }
    // The coarsest levels whose vectors fit in one workgroup's LDS (at most
    // MAX_COARSE_LEVELS of them) run in the one-workgroup kernel.
    m->coarse_start = m->lv.size();
    {
        const size_t elem = precision == 64 ? sizeof(double) : sizeof(float);
        const size_t limit = coarse_lds_limit(ctx->device);
        size_t bytes = 0;
        for (size_t l = m->lv.size(); l-- > 0;) {
            bytes += COARSE_VECTORS * elem * m->lv[l].n;
            if (bytes > limit || m->lv[l].n > COARSE_MAX_POINTS || m->lv.size() - l > (size_t) MAX_COARSE_LEVELS) break;
            m->coarse_start = l;
        }
    }
    if (st == IBM_OK && m->coarse_start < m->lv.size()) {
        if (precision == 64) {
            fill_coarse<double>(m, m->coarse64);
            st = prepare_coarse_vcycle<double>(ctx, m->coarse64.lds_bytes);
        } else {
            fill_coarse<float>(m, m->coarse32);
            st = prepare_coarse_vcycle<float>(ctx, m->coarse32.lds_bytes);
        }
    }
    if (st == IBM_OK) st = precision == 64 ? tune_levels<double>(ctx, m) : tune_levels<float>(ctx, m);
    if (st == IBM_OK) st = alloc(ctx, m, m->lv[0].n, &m->r);
    if (st == IBM_OK) st = alloc(ctx, m, m->lv[0].n, &m->rhs);
    if (st == IBM_OK) st = alloc(ctx, m, REDUCE_PARTS, &m->partial);
    if (st == IBM_OK) st = alloc(ctx, m, 8, &m->scalars);
    if (st != IBM_OK) {
        for (void* p : m->owned) hipFree(p);
        delete m;
        return st;
    }
    *out = m;
    return IBM_OK;
}

ibm_status ibm_mg_destroy(ibm_mg* m)
{
    if (!m) return IBM_OK;
    hipDeviceSynchronize();
    for (void* p : m->owned) hipFree(p);
    delete m;
    return IBM_OK;
}

ibm_status ibm_mg_levels(const ibm_mg* m, int* levels, ibm_grid* grids, int max_grids)
{
    if (!m || !levels) return IBM_ERR_ARG;
    *levels = (int) m->lv.size();
    for (int i = 0; grids && i < max_grids && i < (int) m->lv.size(); ++i) grids[i] = m->lv[i].g;
    return IBM_OK;
}

ibm_status ibm_mg_autotune(int enable, int clear)
{
    if (clear) clear_tuned_targets();
    autotune_on.store(enable != 0);
    return IBM_OK;
}

ibm_status ibm_mg_shape(const ibm_mg* m, int level, int kind, int* target, int* wide)
{
    if (!m || !target || !wide || kind < TUNE_CHEB || kind > TUNE_PR) return IBM_ERR_ARG;
    if (level < 0 || (size_t) level >= m->lv.size()) return IBM_ERR_ARG;
    const ibm_grid& g = m->lv[level].g;
    *target = tuned_target(kind, m->precision == 64 ? 8 : 4, g.axis[0].points, g.axis[1].points, g.axis[2].points,
                           per_of(g));
    *wide = march_wide(m->lv[level].n) ? 1 : 0;
    return IBM_OK;
}

ibm_status ibm_mg_vcycle(ibm_ctx* ctx, ibm_mg* m, const double* b, double* x)
{
    if (!ctx || !m) return IBM_ERR_ARG;
    if (!b || !x) return fail(ctx, IBM_ERR_ARG, "null vector");
    size_t n = m->lv[0].n;
    if (m->precision == 64) {
        IBM_TRY(launch_convert<double, double>(ctx, n, b, m->b64[0].b));
        IBM_TRY(vcycle<double>(ctx, m, 0));
        return launch_convert<double, double>(ctx, n, m->b64[0].x, x);
    }
    IBM_TRY(launch_convert<float, double>(ctx, n, b, m->b32[0].b));
    IBM_TRY(vcycle<float>(ctx, m, 0));
    return launch_convert<double, float>(ctx, n, m->b32[0].x, x);
}

ibm_status ibm_mg_solve(ibm_ctx* ctx, ibm_mg* m, const double* b, double* x, double rtol, int maxit, int* iters,
                        double* relres)
{
    if (!ctx || !m) return IBM_ERR_ARG;
    if (!b || !x) return fail(ctx, IBM_ERR_ARG, "null vector");
    if (b == x) return fail(ctx, IBM_ERR_ARG, "b and x must not alias");
    if (m->precision == 64) return solve_loop<double>(ctx, m, b, x, rtol, 0.0, maxit, false, iters, relres);
    return solve_loop<float>(ctx, m, b, x, rtol, 0.0, maxit, false, iters, relres);
}

ibm_status ibm_mg_solve_atol(ibm_ctx* ctx, ibm_mg* m, const double* b, double* x, double rtol, double atol, int maxit,
                             int warm, int* iters, double* relres)
{
    if (!ctx || !m) return IBM_ERR_ARG;
    if (!b || !x) return fail(ctx, IBM_ERR_ARG, "null vector");
    if (b == x) return fail(ctx, IBM_ERR_ARG, "b and x must not alias");
    if (!(atol >= 0)) return fail(ctx, IBM_ERR_ARG, "negative absolute tolerance");
    if (m->precision == 64) return solve_loop<double>(ctx, m, b, x, rtol, atol, maxit, warm != 0, iters, relres);
    return solve_loop<float>(ctx, m, b, x, rtol, atol, maxit, warm != 0, iters, relres);
}

ibm_status ibm_axpby(ibm_ctx* ctx, long long n, double a, const double* x, double b, double* y)
{
    if (!ctx) return IBM_ERR_ARG;
    if (n < 0 || (n && (!x || !y))) return fail(ctx, IBM_ERR_ARG, "bad vector arguments");
    if (!n) return IBM_OK;
    return launch_axpby<double, double>(ctx, (size_t) n, a, x, b, y);
}

ibm_status ibm_mg_solve_fixed(ibm_ctx* ctx, ibm_mg* m, const double* b, double* x, int iters, int warm,
                              double* relres)
{
    if (!ctx || !m) return IBM_ERR_ARG;
    if (!b || !x) return fail(ctx, IBM_ERR_ARG, "null vector");
    if (b == x) return fail(ctx, IBM_ERR_ARG, "b and x must not alias");
    if (iters < 0) return fail(ctx, IBM_ERR_ARG, "negative iteration count");
    if (m->precision == 64) return solve_fixed<double>(ctx, m, b, x, iters, warm != 0, relres);
    return solve_fixed<float>(ctx, m, b, x, iters, warm != 0, relres);
}

// solvers::chebpcg (solvers/chebpcg.h:9-23 -> solvers/pcg.h:11-44): CG
// preconditioned by the Chebyshev polynomial over the Gershgorin interval of
// A, stopping when ||r|| <= rtol ||b||.
ibm_status ibm_chebpcg_solve(ibm_ctx* ctx, const ibm_grid* g, double sigma, double tau, const double* b, double* x,
                             double rtol, int maxit, double* work /* 6 * size */, int* iters, double* relres)
{
    if (!ctx) return IBM_ERR_ARG;
    IBM_TRY(check_grid(ctx, g));
    if (!b || !x || !work) return fail(ctx, IBM_ERR_ARG, "null vector");
    const size_t n = (size_t) ibm_grid_size(g);
    op3<double> A = make_op<double>(*g, sigma, tau);
    double lo, hi, w[4];
    gershgorin_bounds(*g, sigma, tau, &lo, &hi);
    chebyshev_weights(hi, lo, w);     // preconditioners/chebyshev.h:11-12 passes (b, a)
    double* r = work;
    double* d = work + n;
    double* q = work + 2 * n;
    double* s1 = work + 3 * n;
    double* s2 = work + 4 * n;
    double* red = work + 5 * n;       // REDUCE_PARTS + 4 doubles (n >= 1029 not assumed)
    double* partial;
    double* scal;
    void* p;
    IBM_TRY(scratch(ctx, SLOT_REDUCE, sizeof(double) * (REDUCE_PARTS + 8), &p));
    partial = (double*) p;
    scal = partial + REDUCE_PARTS;
    (void) red;
    auto dot = [&](const double* u, const double* v, double* out) -> ibm_status {
        IBM_TRY(launch_dot<double, double>(ctx, n, u, v, partial, scal));
        return read_scalar(ctx, scal, out);
    };
    IBM_TRY(launch_fill<double>(ctx, n, 0.0, x));
    IBM_TRY(launch_convert<double, double>(ctx, n, b, r));
    double nb2;
    IBM_TRY(dot(b, b, &nb2));
    double nb = std::sqrt(nb2), nr = nb;
    IBM_TRY(cheb_apply<double>(ctx, A, w, r, s1, s2, d, false));      // d = P r
    double delta;
    IBM_TRY(dot(r, d, &delta));
    int it = 0;
    if (!std::isfinite(nb)) {
        if (iters) *iters = 0;
        if (relres) *relres = NAN;
        return fail(ctx, IBM_ERR_NOT_CONVERGED, "chebpcg: non-finite right-hand side");
    }
    while (nb > 0 && it < maxit && !(nr <= rtol * nb)) {
        IBM_TRY(launch_stencil<double>(ctx, A, 1.0, d, 0.0, nullptr, 0.0, nullptr, q));   // q = A d
        double nu;
        IBM_TRY(dot(d, q, &nu));
        if (std::fabs(nu) == 0) break;
        double alpha = delta / nu;
        IBM_TRY(launch_axpby<double, double>(ctx, n, alpha, d, 1.0, x));
        IBM_TRY(launch_axpby<double, double>(ctx, n, -alpha, q, 1.0, r));
        IBM_TRY(cheb_apply<double>(ctx, A, w, r, s1, s2, q, false));  // q = P r
        double dold = delta;
        IBM_TRY(dot(r, q, &delta));
        double beta = delta / dold;
        IBM_TRY(launch_axpby<double, double>(ctx, n, 1.0, q, beta, d)); // d = q + beta d
        ++it;
        double rr;
        IBM_TRY(dot(r, r, &rr));
        nr = std::sqrt(rr);
        if (!std::isfinite(nr)) break;
    }
    if (iters) *iters = it;
    if (relres) *relres = nb > 0 ? nr / nb : 0.0;
    if (nb > 0 && !(nr <= rtol * nb)) return fail(ctx, IBM_ERR_NOT_CONVERGED, "chebpcg: relative residual %.3e after %d iterations", nr / nb, it);
    return IBM_OK;
}

ibm_status ibm_dot(ibm_ctx* ctx, int64_t n, const double* x, const double* y, double* out)
{
    if (!ctx || !out) return IBM_ERR_ARG;
    if (n < 0 || (n > 0 && (!x || !y))) return fail(ctx, IBM_ERR_ARG, "bad vectors");
    void* p;
    IBM_TRY(scratch(ctx, SLOT_REDUCE, sizeof(double) * (REDUCE_PARTS + 8), &p));
    double* partial = (double*) p;
    IBM_TRY(launch_dot<double, double>(ctx, (size_t) n, x, y, partial, partial + REDUCE_PARTS));
    return read_scalar(ctx, partial + REDUCE_PARTS, out);
}

} // extern "C"

// ---------------------------------------------------------------------------
// Helmholtz solves on many independent copies of one periodic grid (a batch
// of small boxes): the V-cycle of solvers::multigrid (FP32, degree-4
// Chebyshev smoothing over the upper Gershgorin interval, 4-tap transfers,
// smoother-only coarsest level) inside FP64 defect correction, every
// operation one launch over all copies; a copy converges when its own
// relative FP64 residual reaches rtol, and the solve stops when all have.
// ---------------------------------------------------------------------------
struct ibm_boxes_mg {
    int nbatch = 0;
    double sigma = 0, tau = 0;
    std::vector<ibm::mg_level> lv;
    std::vector<ibm::level_bufs<float>> b32;
    double* r = nullptr;            // FP64 residual, nbatch x n
    double* norms = nullptr;        // device: |b|^2 then |r|^2 per copy
    std::vector<double> host;       // host copy of the norms
    std::vector<void*> owned;
};

namespace ibm {
namespace {

template <typename T>
ibm_status boxes_alloc(ibm_ctx* ctx, ibm_boxes_mg* m, size_t n, T** p)
{
    void* v = nullptr;
    IBM_HIP(ctx, hipMalloc(&v, sizeof(T) * std::max<size_t>(n, 1)));
    m->owned.push_back(v);
    *p = (T*) v;
    return IBM_OK;
}

// out = p(A) v, or out += p(A) v (algo/chebyshev.h:149-163 with chebw<4>).
ibm_status boxes_cheb(ibm_ctx* ctx, const ibm_boxes_mg* m, int l, const float* v, float* out, bool add)
{
    const mg_level& L = m->lv[l];
    const level_bufs<float>& B = m->b32[l];
    const int nb = m->nbatch;
    IBM_TRY(launch_batch_stencil<float>(ctx, L.A32, nb, (float) -L.w[0], v, 0.0f, nullptr, (float) L.w[1], v, B.s1));
    IBM_TRY(launch_batch_stencil<float>(ctx, L.A32, nb, -1.0f, B.s1, 0.0f, nullptr, (float) L.w[2], v, B.s2));
    return launch_batch_stencil<float>(ctx, L.A32, nb, -1.0f, B.s2, add ? 1.0f : 0.0f, add ? out : nullptr,
                                       (float) L.w[3], v, out);
}

// x_l = V(b_l) (solvers/multigrid.h:120-128)
ibm_status boxes_vcycle(ibm_ctx* ctx, const ibm_boxes_mg* m, int l)
{
    const mg_level& L = m->lv[l];
    const level_bufs<float>& B = m->b32[l];
    const int nb = m->nbatch;
    IBM_TRY(boxes_cheb(ctx, m, l, B.b, B.x, false));
    if (l + 1 == (int) m->lv.size()) return IBM_OK;
    IBM_TRY(launch_batch_stencil<float>(ctx, L.A32, nb, -1.0f, B.x, 0.0f, nullptr, 1.0f, B.b, B.r));
    IBM_TRY(launch_batch_restrict<float>(ctx, L.T32, nb, B.r, m->b32[l + 1].b));
    IBM_TRY(boxes_vcycle(ctx, m, l + 1));
    IBM_TRY(launch_batch_prolong_add<float>(ctx, L.T32, nb, m->b32[l + 1].x, B.x));
    IBM_TRY(launch_batch_stencil<float>(ctx, L.A32, nb, -1.0f, B.x, 0.0f, nullptr, 1.0f, B.b, B.r));
    return boxes_cheb(ctx, m, l, B.r, B.x, true);
}

} // namespace
} // namespace ibm

extern "C" {

ibm_status ibm_boxes_mg_create(ibm_ctx* ctx, const ibm_grid* grid, int nbatch, double sigma, double tau,
                               ibm_boxes_mg** out)
{
    using namespace ibm;
    if (!ctx || !out) return IBM_ERR_ARG;
    *out = nullptr;
    if (!grid || grid->dims != 3) return fail(ctx, IBM_ERR_ARG, "batched solves take one 3-dimensional grid");
    if (nbatch < 1) return fail(ctx, IBM_ERR_ARG, "at least one copy");
    for (int d = 0; d < 3; ++d)
        if (grid->axis[d].solid) return fail(ctx, IBM_ERR_BAD_GRID, "batched solves take periodic boxes");
    if (!(sigma != 0.0)) return fail(ctx, IBM_ERR_ARG, "batched solves take Helmholtz operators (sigma != 0)");
    auto* m = new ibm_boxes_mg;
    m->nbatch = nbatch;
    m->sigma = sigma;
    m->tau = tau;
    ibm_grid cur = *grid;
    ibm_status st = IBM_OK;
    for (;;) {
        mg_level L{};
        L.g = cur;
        L.n = (size_t) ibm_grid_size(&cur);
        double a, b;
        smoother_interval(cur, sigma, tau, &a, &b);
        chebyshev_weights(a, b, L.w);
        L.A64 = make_op<double>(cur, sigma, tau);
        L.A32 = make_op<float>(cur, sigma, tau);
        ibm_grid next;
        bool more = refinable(cur) && ibm_grid_coarsen(&cur, &next) == IBM_OK;
        if (more) L.T32 = make_transfer<float>(cur, next);
        m->lv.push_back(L);
        if (!more) break;
        cur = next;
    }
    for (auto& L : m->lv) {
        level_bufs<float> B;
        size_t n = L.n * (size_t) nbatch;
        if ((st = boxes_alloc(ctx, m, n, &B.x)) || (st = boxes_alloc(ctx, m, n, &B.b)) ||
            (st = boxes_alloc(ctx, m, n, &B.r)) || (st = boxes_alloc(ctx, m, n, &B.s1)) ||
            (st = boxes_alloc(ctx, m, n, &B.s2)))
            break;
        m->b32.push_back(B);
    }
    if (st == IBM_OK) st = boxes_alloc(ctx, m, m->lv[0].n * (size_t) nbatch, &m->r);
    if (st == IBM_OK) st = boxes_alloc(ctx, m, 2 * (size_t) nbatch, &m->norms);
    m->host.resize(2 * (size_t) nbatch);
    if (st != IBM_OK) {
        ibm_boxes_mg_destroy(m);
        return st;
    }
    *out = m;
    return IBM_OK;
}

ibm_status ibm_boxes_mg_destroy(ibm_boxes_mg* m)
{
    if (!m) return IBM_OK;
    hipDeviceSynchronize();
    for (void* p : m->owned) hipFree(p);
    delete m;
    return IBM_OK;
}

ibm_status ibm_boxes_mg_solve(ibm_ctx* ctx, ibm_boxes_mg* m, const double* b, double* x, double rtol, int maxit,
                              int* iters, double* max_relres)
{
    using namespace ibm;
    if (!ctx || !m) return IBM_ERR_ARG;
    if (!b || !x) return fail(ctx, IBM_ERR_ARG, "null vector");
    const int nb = m->nbatch;
    const mg_level& L0 = m->lv[0];
    const size_t n = L0.n, total = n * (size_t) nb;
    IBM_TRY(launch_batch_dot<double>(ctx, n, nb, b, b, m->norms));
    int it = 0;
    double worst = INFINITY;
    for (;;) {
        // r = b - A x (FP64), then each copy's |r|^2
        IBM_TRY(launch_batch_stencil<double>(ctx, L0.A64, nb, -1.0, x, 0.0, nullptr, 1.0, b, m->r));
        IBM_TRY(launch_batch_dot<double>(ctx, n, nb, m->r, m->r, m->norms + nb));
        IBM_HIP(ctx, hipMemcpyAsync(m->host.data(), m->norms, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost,
                                    ctx->stream));
        IBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        worst = 0;
        for (int k = 0; k < nb; ++k) {
            double bn = m->host[k], rn = m->host[nb + k];
            double rel = bn > 0 ? std::sqrt(rn / bn) : (rn > 0 ? INFINITY : 0.0);
            if (!(rel <= worst)) worst = rel;      // carries NaN
        }
        if (worst <= rtol || it >= maxit) break;
        IBM_TRY(launch_convert<float, double>(ctx, total, m->r, m->b32[0].b));
        IBM_TRY(boxes_vcycle(ctx, m, 0));
        IBM_TRY(launch_axpby<double, float>(ctx, total, 1.0, m->b32[0].x, 1.0, x));
        ++it;
    }
    if (iters) *iters = it;
    if (max_relres) *max_relres = worst;
    if (!(worst <= rtol))
        return fail(ctx, IBM_ERR_NOT_CONVERGED, "batched solve: worst relative residual %g after %d cycles", worst, it);
    return IBM_OK;
}

} // extern "C"
