// Immersed-boundary timestep: the composer the reference calls as
// ib::solver (test/ib-whole.cc:226,250-275; its header is not part of the
// reference), built from the library's operators.  The scheme:
//
//   1. surface geometry of every cell type from its data sites X:
//      [y, y_theta, y_phi] = E^T X (one RBF launch for all types), unit
//      normals, patch areas and per-cell area / volume in FP64
//      (bases/geometry.h:74-123);
//   2. membrane forces at the sample sites: per cell type the tension law
//      (Skalak or neo-Hookean, forces/tension.h) and bending
//      (forces/bending.h) against the type's undeformed shape, whose
//      mean curvature needs the data-site geometry (five data-to-data
//      operators) and its derivatives at the sample sites (five sample
//      operators on h), and / or a tether f = -k (y - y_ref) to reference
//      sample positions (the spring of test/ib-whole.cc:237-245);
//   3. F = spread(y, f) onto the MAC grid (ib/novel.h);
//   4. H = ins::advection(u) (ins/advection.h:58-125);
//   5. per component, ins::diffusion (ins/diffusion.h:46-56) with gamma = 1:
//        rhs = dt nu (b + L u) + dt I~ (F / rho - H),
//        (I~ - lambda L) du = rhs,  lambda = dt nu / 2,  u* = u + du,
//      b the boundary data of moving walls (fd::upper_boundary x values,
//      test/ib-whole.cc:129-152), solved by multigrid; with
//      IBM_SOLVER_PSEUDO_PRESSURE the new time level's wall values are
//      g + dt grad p (the w = u + k grad phi of ins/boundary.h:16-18, p the
//      previous step's pressure, tangential components only), i.e.
//        rhs += lambda dt B_i (G_i p),  B_i = ins::boundary(grid_i, i);
//   6. projection: L_p phi = (D u*) / dt with L_p = D G the Neumann
//      Laplacian of the cell centres (singular, mean-free solution),
//      u = u* - dt G phi, p = phi;
//   7. FP64 divergence check max |D u|;
//   8. U = interpolate(u) at the data sites (ib/interpolate.h), X += dt U.
//
// With a force scale (params.fscale > 0) the step size follows ins::tracker
// (ins/tracker.h:13-100): e = floor(log2(k_max / k_min)) with
// k_max = sqrt(fscale / (|f| + 1e-3)) / 4 and |f| = sqrt(sum over cell types
// of max|f|^2), dt = k_min 2^min(4, e over the last 16 steps); e < 0 is
// IBM_ERR_DT_TOO_SMALL.  (The reference's retry loop compares the newest
// exponent with a minimum that already includes it, so it never runs.)
// One set of Helmholtz hierarchies is kept per step size in use.
//
// All work is enqueued on the context stream.  ibm_solver_step chooses the
// number of V-cycles of each solve adaptively (host reads of residual norms);
// ibm_solver_step_fixed / ibm_solver_capture use fixed counts (the largest
// the adaptive steps needed so far, plus one) and no host synchronisation,
// so a whole step can be recorded once as a HIP graph and replayed.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#include "context.hpp"
#include "fd_kernels.hpp"

namespace ibm {
namespace {

constexpr int MAX_TYPES = IBM_MAX_CELL_TYPES;

__global__ void tether_kernel(size_t n, double k, const double* __restrict__ y, const double* __restrict__ ref,
                              double* __restrict__ f)
{
    for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t) gridDim.x * blockDim.x)
        f[i] = -k * (y[i] - ref[i]);
}

// Initial guess of a warm-started solve from its last two solutions at the
// same step size: x <- 2 x - prev (linear extrapolation in time) when the
// device step counter *hist says both exist, x unchanged otherwise; prev <-
// the incoming x.  The counter lives on the device so that a replayed graph
// and an eager step decide alike.
__global__ void extrapolate_kernel(size_t n, const int* hist, double* x, double* prev)
{
    const int two = *hist >= 2;
    for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t) gridDim.x * blockDim.x) {
        const double c = x[i];
        const double g = two ? 2.0 * c - prev[i] : c;
        prev[i] = c;
        x[i] = g;
    }
}

// Step counter of the warm starts: set (v >= 0) or advanced towards 2 (v < 0).
__global__ void history_kernel(int* hist, int v)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *hist = v >= 0 ? v : min(*hist + 1, 2);
}

int blocks(size_t n) { return (int) std::min<size_t>((n + 255) / 256, 65536); }

} // namespace
} // namespace ibm

using namespace ibm;

// Adaptive steps whose V-cycle counts set the fixed-count steps'.
constexpr int ITER_WINDOW = 4;

struct ibm_solver {
    ibm_solver_params par{};
    ibm_grid g[3]{}, gp{};
    size_t nu[3]{}, np = 0;
    ibm_mg* helm[3]{};                 // Helmholtz hierarchies of the current step size
    ibm_mg* helm_e[5][3]{};            // per tracker exponent (adaptive step size)
    double dt = 0;                     // current step size
    int hist[16]{};                    // tracker exponents of the last steps
    int hist_len = 0, hist_next = 0;
    ibm_mg* pois = nullptr;
    // Direct FFT x tridiagonal solves (channel grids, unless multigrid is
    // forced): one solver per velocity grid and one for the pressure.
    bool direct = false;
    ibm_fft* fft_u[3]{};
    ibm_fft* fft_p = nullptr;
    double* dpartial[4]{};             // reduction partials of the residual checks (u, v, w, p)
    size_t dparts[4]{};
    ins_boundary_op Bw[3]{};           // ins::boundary(g[i], i), order 2 (pseudo-pressure walls)
    struct type_t {
        ibm_rbf* ops;
        int nd, ns, nc;
        double stiffness;
        double *w, *ref, *Y, *f, *normal, *sigma, *area, *volume, *U;
        // membrane models: Y holds six blocks (position, x_u, x_v, x_uu,
        // x_uv, x_vv); bending adds the data geometry Yd (five blocks of
        // nd x 3 cells), h = H - H0 at the data sites and its sample-site
        // derivatives hd (six ns x nc blocks, the first unused)
        ibm_membrane mem;
        bool membrane, bend;
        ibm_rbf* dops;
        double *table, *h0, *Yd, *h, *hd;
    };
    std::vector<type_t> types;
    double *F[3]{}, *H[3]{}, *rhs[3]{}, *du[3]{}, *div = nullptr;
    bool tile_spread = false;          // spread by tiles (the grids allow it) or by the brick kernel
    double *du_prev[3]{}, *p_prev = nullptr;   // the solutions one step further back
    int* warm_hist = nullptr;          // device: solutions known at step size history_dt (0, 1, 2)
    // Fixed-count steps run the three Helmholtz solves concurrently: the
    // second and third on contexts with streams of their own, forked from and
    // joined back into the caller's stream with events.
    ibm_ctx* aux[2]{};
    hipStream_t aux_stream[2]{};
    hipEvent_t fork_ev = nullptr, join_ev[2]{};
    // ... and evaluate the two tangent operators and the surface geometry
    // (outputs of the step that nothing later in it reads) on the first
    // auxiliary stream, beside the tether forces, spreading and advection.
    hipEvent_t geo_ev = nullptr;
    double history_dt = 0;
    double* info = nullptr;            // device: residual[4], divergence, scratch scalars
    double* partial = nullptr;         // reduction partials (REDUCE_PARTS)
    int max_iters[4]{};                // largest adaptive iteration counts of the last ITER_WINDOW steps
    int last_iters[4]{};
    int recent_iters[ITER_WINDOW][4]{};   // iteration counts of the last ITER_WINDOW adaptive steps
    int recent_next = 0, recent_len = 0;
    int adaptive_steps = 0;
    int last_exponent = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ibm_state captured{};
    double captured_dt = 0;
    std::vector<void*> owned;
};

namespace {

// A status from work on an auxiliary context, with its message copied to
// the caller's context.
ibm_status step_status(ibm_ctx* ctx, ibm_ctx* c, ibm_status st)
{
    if (st != IBM_OK && c != ctx) ctx->error = c->error;
    return st;
}

// Restart the warm-start history when the step size differs from the one
// the stored solutions belong to (host side, outside any captured graph).
ibm_status sync_history(ibm_ctx* ctx, ibm_solver* s, double dt)
{
    if (dt == s->history_dt) return IBM_OK;
    history_kernel<<<1, 64, 0, ctx->stream>>>(s->warm_hist, 0);
    IBM_LAUNCH_CHECK(ctx);
    s->history_dt = dt;
    return IBM_OK;
}

ibm_status alloc(ibm_ctx* ctx, ibm_solver* s, size_t n, double** p)
{
    void* v = nullptr;
    IBM_HIP(ctx, hipMalloc(&v, sizeof(double) * std::max<size_t>(n, 1)));
    s->owned.push_back(v);
    *p = (double*) v;
    return IBM_OK;
}

void release(ibm_solver* s)
{
    if (s->exec) hipGraphExecDestroy(s->exec);
    if (s->graph) hipGraphDestroy(s->graph);
    for (int k = 0; k < 2; ++k) {
        if (s->aux[k]) ibm_destroy(s->aux[k]);
        if (s->aux_stream[k]) (void) hipStreamDestroy(s->aux_stream[k]);
        if (s->join_ev[k]) (void) hipEventDestroy(s->join_ev[k]);
    }
    if (s->fork_ev) (void) hipEventDestroy(s->fork_ev);
    if (s->geo_ev) (void) hipEventDestroy(s->geo_ev);
    if (s->par.fscale > 0) {
        for (int e = 0; e < 5; ++e)
            for (int i = 0; i < 3; ++i)
                if (s->helm_e[e][i]) ibm_mg_destroy(s->helm_e[e][i]);
    } else {
        for (int i = 0; i < 3; ++i)
            if (s->helm[i]) ibm_mg_destroy(s->helm[i]);
    }
    if (s->pois) ibm_mg_destroy(s->pois);
    for (int i = 0; i < 3; ++i)
        if (s->fft_u[i]) ibm_fft_destroy(s->fft_u[i]);
    if (s->fft_p) ibm_fft_destroy(s->fft_p);
    hipDeviceSynchronize();
    for (void* p : s->owned) hipFree(p);
    delete s;
}

ibm_status check_state(ibm_ctx* ctx, const ibm_solver* s, const ibm_state* st)
{
    if (!st) return fail(ctx, IBM_ERR_ARG, "null state");
    for (int i = 0; i < 3; ++i)
        if (!st->u[i]) return fail(ctx, IBM_ERR_ARG, "null velocity component %d", i);
    if (!st->p) return fail(ctx, IBM_ERR_ARG, "null pressure");
    for (size_t t = 0; t < s->types.size(); ++t)
        if (!st->x[t] && s->types[t].nc) return fail(ctx, IBM_ERR_ARG, "null positions of cell type %zu", t);
    return IBM_OK;
}

// ins::tracker: pushes this step's exponent, selects dt and its Helmholtz
// hierarchies (created on first use).
ibm_status track(ibm_ctx* ctx, ibm_solver* s)
{
    const ibm_solver_params& P = s->par;
    double f2 = 0;
    for (auto& T : s->types) {
        if (!T.nc) continue;
        IBM_TRY(launch_max_abs(ctx, (size_t) T.ns * 3 * T.nc, T.f, s->partial, s->info + 6));
        double m = 0;
        IBM_HIP(ctx, hipMemcpyAsync(&m, s->info + 6, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        IBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        f2 += m * m;
    }
    const double fmax = std::sqrt(f2);
    const double kmax = 0.25 * std::sqrt(P.fscale / (fmax + 1e-3));
    const int e_new = (int) std::floor(std::log2(kmax / P.timestep));
    s->hist[s->hist_next] = e_new;
    s->hist_next = (s->hist_next + 1) % 16;
    s->hist_len = std::min(s->hist_len + 1, 16);
    int e = 4;
    for (int i = 0; i < s->hist_len; ++i) e = std::min(e, s->hist[i]);
    s->last_exponent = e;
    if (e < 0) return fail(ctx, IBM_ERR_DT_TOO_SMALL, "time step bound %.3e below kmin %.3e", kmax, P.timestep);
    s->dt = P.timestep * (double) (1 << e);
    if (s->direct) return IBM_OK;
    if (!s->helm_e[e][0]) {
        const double lambda = s->dt * P.diffusivity / 2;
        for (int i = 0; i < 3; ++i) IBM_TRY(ibm_mg_create(ctx, &s->g[i], 1.0, -lambda, P.precision, &s->helm_e[e][i]));
    }
    for (int i = 0; i < 3; ++i) s->helm[i] = s->helm_e[e][i];
    return IBM_OK;
}

// One timestep; `fixed` selects fixed V-cycle counts and no host reads.
ibm_status step_impl(ibm_ctx* ctx, ibm_solver* s, ibm_state* st, const double* const* ub, bool fixed)
{
    const ibm_solver_params& P = s->par;
    const double nu = P.diffusivity;
    const int nt = (int) s->types.size();
    // Fixed-count steps overlap independent work on the auxiliary streams
    // (no host synchronisation inside; profiled steps stay on one stream so
    // that every launch is timed).  Everything forked is joined back into
    // the caller's stream before the projection.
    const bool concurrent = fixed && !ctx->profiling && s->aux[0] && s->aux[1];

    // Bending's chain: the data-site geometry of every bending type (one
    // launch), the mean curvature at the data sites, its derivatives at the
    // sample sites (one launch).  It runs after the sample-site geometry on
    // the same stream: beside it on a second stream both matrix-core
    // launches slowed down more than they overlapped (5.2 against 4.5 ms).
    bool data_done = false;     // bending's data-site geometry computed in the geometry launch
    auto bending_chain = [&](ibm_ctx* bc) -> ibm_status {
        ibm_rbf* dops[MAX_TYPES];
        ibm_rbf* hops[MAX_TYPES];
        int dN[MAX_TYPES], hN[MAX_TYPES];
        const double* dX[MAX_TYPES];
        const double* hX[MAX_TYPES];
        double* dY[MAX_TYPES];
        double* hY[MAX_TYPES];
        int nb = 0;
        for (int t = 0; t < nt; ++t) {
            auto& T = s->types[t];
            if (!T.nc || !T.bend) continue;
            dops[nb] = T.dops;
            dN[nb] = 3 * T.nc;
            dX[nb] = st->x[t];
            dY[nb] = T.Yd;
            hops[nb] = T.ops;
            hN[nb] = T.nc;
            hX[nb] = T.h;
            hY[nb] = T.hd;
            ++nb;
        }
        if (!nb) return IBM_OK;
        unsigned char* dpacked = nullptr;
        if (!data_done)
            IBM_TRY(step_status(ctx, bc, rbf_eval_range(bc, bc->stream, nb, dops, dN, dX, dY, 0, 5, true, &dpacked)));
        for (int t = 0; t < nt; ++t) {
            auto& T = s->types[t];
            if (!T.nc || !T.bend) continue;
            IBM_TRY(step_status(ctx, bc, mean_curvature(bc, bc->stream, T.nd, T.nc, T.Yd,
                                                         T.mem.preferred ? T.h0 : nullptr, T.h)));
        }
        unsigned char* hpacked = nullptr;
        return step_status(ctx, bc, rbf_eval_range(bc, bc->stream, nb, hops, hN, hX, hY, 1, 6, true, &hpacked));
    };

    // 1. geometry: positions, tangents, normals and surface measures of
    // every cell type (the positions in one launch for all types, the
    // tangents in a second one)
    {
        ibm_rbf* ops[2 * MAX_TYPES];
        int N[2 * MAX_TYPES];
        const double* X[2 * MAX_TYPES];
        double* Y[2 * MAX_TYPES];
        int count = 0, groups = 0;
        for (int t = 0; t < nt; ++t) {
            auto& T = s->types[t];
            if (!T.nc) continue;
            ops[count] = T.ops;
            N[count] = 3 * T.nc;
            X[count] = st->x[t];
            Y[count] = T.Y;
            groups += rbf_nops(T.ops);
            ++count;
        }
        bool any_membrane = false;
        for (int t = 0; t < nt; ++t) any_membrane = any_membrane || (s->types[t].membrane && s->types[t].nc);
        // Bending's data-site geometry depends on the data sites alone: it
        // joins the same launch (one grouped launch balances better than
        // two) when the groups fit one launch.
        int dgroups = 0;
        for (int t = 0; t < nt; ++t)
            if (s->types[t].bend && s->types[t].nc) dgroups += rbf_nops(s->types[t].dops);
        if (dgroups && groups + dgroups <= RBF_MAX_GROUPS) {
            for (int t = 0; t < nt; ++t) {
                auto& T = s->types[t];
                if (!T.nc || !T.bend) continue;
                ops[count] = T.dops;
                N[count] = 3 * T.nc;
                X[count] = st->x[t];
                Y[count] = T.Yd;
                ++count;
            }
            data_done = true;
        }
        // Positions (operator 0) first: tether forces and the spreading need
        // only them.  Membrane forces need every operator, so with a
        // membrane model the whole geometry stays on the caller's stream.
        const bool split_geometry = concurrent && !any_membrane;
        ibm_ctx* gc = split_geometry ? s->aux[0] : ctx;
        unsigned char* packed = nullptr;
        if (count) {
            if (!split_geometry) {
                IBM_TRY(rbf_eval_range(ctx, ctx->stream, count, ops, N, X, Y, 0, 8, true, &packed));
            } else {
                IBM_TRY(rbf_eval_range(ctx, ctx->stream, count, ops, N, X, Y, 0, 1, true, &packed));
                IBM_HIP(ctx, hipEventRecord(s->geo_ev, ctx->stream));
                IBM_HIP(ctx, hipStreamWaitEvent(gc->stream, s->geo_ev, 0));
                // One tangent operator per launch: half as many long-lived
                // matrix-core blocks resident beside the force / binning /
                // spreading chain.
                IBM_TRY(step_status(ctx, gc, rbf_eval_range(gc, gc->stream, count, ops, N, X, Y, 1, 2, false, &packed)));
                IBM_TRY(step_status(ctx, gc, rbf_eval_range(gc, gc->stream, count, ops, N, X, Y, 2, 8, false, &packed)));
            }
        }
        for (int t = 0; t < nt; ++t) {
            auto& T = s->types[t];
            if (!T.nc) continue;
            size_t blk = (size_t) T.ns * 3 * T.nc;
            IBM_TRY(step_status(ctx, gc, ibm_surface_geometry(gc, T.ns, T.nc, T.Y, T.Y + blk, T.Y + 2 * blk, T.w,
                                                               T.normal, T.sigma, T.area, T.volume)));
            // 2. tether forces (cell types without a membrane model)
            if (!T.membrane) {
                if (T.stiffness != 0.0) {
                    tether_kernel<<<blocks(blk), 256, 0, ctx->stream>>>(blk, T.stiffness, T.Y, T.ref, T.f);
                    IBM_LAUNCH_CHECK(ctx);
                } else {
                    IBM_TRY(launch_fill<double>(ctx, blk, 0.0, T.f));
                }
            }
        }
        // 2. membrane forces: bending's mean curvature at the data sites
        // (data geometry of every bending type in one launch), its
        // derivatives at the sample sites (one launch), then per type the
        // tension + bending (+ tether) pass
        if (any_membrane) {
            IBM_TRY(bending_chain(ctx));
            for (int t = 0; t < nt; ++t) {
                auto& T = s->types[t];
                if (!T.nc || !T.membrane) continue;
                size_t blk = (size_t) T.ns * 3 * T.nc;
                IBM_TRY(membrane_forces(ctx, ctx->stream, &T.mem, T.stiffness, T.ns, T.nc, T.Y + blk, T.normal,
                                        T.sigma, T.bend ? T.hd + (size_t) T.ns * T.nc : nullptr, T.table, T.Y,
                                        T.ref, 0.0, T.f));
            }
        }
    }
    // 2b. step size (ins::tracker) from the force bound
    if (P.fscale > 0 && !fixed) IBM_TRY(track(ctx, s));
    const double dt = s->dt;
    // Warm starts: extrapolated from the last two solutions once both exist
    // at this step size (sync_history restarts the count before a step at a
    // new size; fixed-count steps and graphs do that in their entry points).
    if (!fixed) IBM_TRY(sync_history(ctx, s, dt));
    // 3. spread all cell types onto the MAC grid
    {
        int counts[MAX_TYPES];
        const double* xs[MAX_TYPES];
        const double* fs[MAX_TYPES];
        for (int t = 0; t < nt; ++t) {
            counts[t] = s->types[t].ns * s->types[t].nc;
            xs[t] = s->types[t].Y;
            fs[t] = s->types[t].f;
        }
        IBM_TRY(spread_sets_step(ctx, s->g, P.delta, s->tile_spread, nt, counts, xs, fs, s->F));
    }
    // 4. advection (on an auxiliary stream beside the geometry it was starved
    // of compute units by the matrix-core blocks: 0.47 ms per component
    // instead of 0.075, no gain)
    {
        const double* u[3] = {st->u[0], st->u[1], st->u[2]};
        IBM_TRY(ibm_advection(ctx, s->g, u, s->H));
    }
    // 5. diffusion, component by component; fixed-count steps run the three
    // components concurrently
    ibm_ctx* cc[3] = {ctx, concurrent ? s->aux[0] : ctx, concurrent ? s->aux[1] : ctx};
    if (concurrent) {
        IBM_HIP(ctx, hipEventRecord(s->fork_ev, ctx->stream));
        for (int k = 0; k < 2; ++k) IBM_HIP(ctx, hipStreamWaitEvent(s->aux[k]->stream, s->fork_ev, 0));
    }
    for (int i = 0; i < 3; ++i) {
        ibm_ctx* c = cc[i];
        const size_t n = s->nu[i];
        // rhs = dt I~ (F/rho - H) + dt nu (b + L u), one pass
        const op3<double> L = make_op<double>(s->g[i], 0.0, 1.0);
        IBM_TRY(step_status(ctx, c, launch_momentum_rhs(c, L, dt, dt * nu, 1.0 / P.density, s->F[i], s->H[i],
                                                         st->u[i], ub ? ub[i] : nullptr, s->rhs[i])));
        if (P.flags & IBM_SOLVER_PSEUDO_PRESSURE) {
            // F[i] (consumed above) <- G_i p; rhs += lambda dt B_i F[i]
            twopoint_op<double> G;
            ibm_grid dst;
            IBM_TRY(step_status(ctx, c, make_twopoint(s->gp, i, 0, &G, &dst)));
            IBM_TRY(step_status(ctx, c, launch_twopoint<double>(c, G, 1.0, st->p, 0.0, s->F[i])));
            IBM_TRY(step_status(ctx, c, launch_ins_boundary(c, s->Bw[i], 0.5 * dt * nu * dt, s->F[i], 1.0,
                                                             s->rhs[i])));
        }
        if (s->direct) {
            // du = A^-1 rhs exactly; then u += du with the FP64 residual
            // check ||rhs - A du|| / ||rhs|| in the same pass.
            const double lambda = dt * nu / 2;
            IBM_TRY(step_status(ctx, c, fft_solve(c, s->fft_u[i], 1.0, -lambda, s->rhs[i], s->du[i])));
            const op3<double> A = make_op<double>(s->g[i], 1.0, -lambda);
            IBM_TRY(step_status(ctx, c, launch_residual_update(c, A, s->du[i], s->rhs[i], st->u[i], s->dpartial[i],
                                                                s->dparts[i], s->info + i)));
            s->last_iters[i] = 1;
            continue;
        }
        extrapolate_kernel<<<blocks(n), 256, 0, c->stream>>>(n, s->warm_hist, s->du[i], s->du_prev[i]);
        IBM_LAUNCH_CHECK(c);
        if (!fixed) {
            int it = 0;
            double rr = 0;
            IBM_TRY(ibm_mg_solve_atol(ctx, s->helm[i], s->rhs[i], s->du[i], P.tolerance, 0.0, P.max_iterations, 1,
                                      &it, &rr));
            s->last_iters[i] = it;
            IBM_HIP(ctx, hipMemcpyAsync(s->info + i, &rr, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
            IBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
            IBM_TRY(launch_axpby<double, double>(ctx, n, 1.0, s->du[i], 1.0, st->u[i]));
        }
    }
    if (fixed && !s->direct) {
        // The three solves cycle by cycle in turn, each on its component's
        // stream (ibm_mg_solve_fixed each).
        ibm_mg* ms[3] = {s->helm[0], s->helm[1], s->helm[2]};
        const double* bs[3] = {s->rhs[0], s->rhs[1], s->rhs[2]};
        double* xs[3] = {s->du[0], s->du[1], s->du[2]};
        double* rr[3] = {s->info + 0, s->info + 1, s->info + 2};
        const int its[3] = {s->max_iters[0] + 1, s->max_iters[1] + 1, s->max_iters[2] + 1};
        for (int i = 1; i < 3; ++i) cc[i]->error.clear();
        const ibm_status mst = mg_solve_fixed_multi(3, cc, ms, bs, xs, its, true, rr, concurrent ? s->fork_ev : nullptr,
                                                    concurrent ? s->join_ev : nullptr);
        if (mst != IBM_OK) {
            // the message is on the context of the solve that failed
            for (int i = 1; i < 3; ++i)
                if (cc[i] != ctx && !cc[i]->error.empty()) ctx->error = cc[i]->error;
            return mst;
        }
        for (int i = 0; i < 3; ++i)
            IBM_TRY(step_status(ctx, cc[i], launch_axpby<double, double>(cc[i], s->nu[i], 1.0, s->du[i], 1.0,
                                                                          st->u[i])));
    }
    if (concurrent) {
        for (int k = 0; k < 2; ++k) {
            IBM_HIP(ctx, hipEventRecord(s->join_ev[k], s->aux[k]->stream));
            IBM_HIP(ctx, hipStreamWaitEvent(ctx->stream, s->join_ev[k], 0));
        }
    }
    // 6. projection
    {
        const double* u[3] = {st->u[0], st->u[1], st->u[2]};
        IBM_TRY(divergence_of(ctx, s->g, u, 1.0 / dt, s->div, nullptr, nullptr));
        if (s->direct) {
            // mean-free p with L p = div - mean(div), and the FP64 check
            // ||div - mean - L p|| / ||div - mean||
            IBM_TRY(fft_solve(ctx, s->fft_p, 0.0, 1.0, s->div, st->p));
            const op3<double> Lp = make_op<double>(s->gp, 0.0, 1.0);
            IBM_TRY(launch_remove_mean<double>(ctx, s->np, s->div, s->partial, s->info + 7));
            IBM_TRY(launch_residual_update(ctx, Lp, st->p, s->div, nullptr, s->dpartial[3], s->dparts[3],
                                           s->info + 3));
            s->last_iters[3] = 1;
        } else {
        extrapolate_kernel<<<blocks(s->np), 256, 0, ctx->stream>>>(s->np, s->warm_hist, st->p, s->p_prev);
        IBM_LAUNCH_CHECK(ctx);
        if (fixed) {
            IBM_TRY(ibm_mg_solve_fixed(ctx, s->pois, s->div, st->p, s->max_iters[3] + 1, 1, s->info + 3));
        } else {
            // Absolute floor: tolerance x the divergence scale of u*,
            // ||D|| ||u*|| / dt with ||D|| <= 2 / h (the right-hand side of a
            // nearly divergence-free u* is rounding noise).
            double uu = 0;
            for (int i = 0; i < 3; ++i) {
                IBM_TRY(launch_dot<double, double>(ctx, s->nu[i], st->u[i], st->u[i], s->partial, s->info + 5));
                double h = 0;
                IBM_HIP(ctx, hipMemcpyAsync(&h, s->info + 5, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
                IBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
                uu += h;
            }
            double rmax = 0;
            for (int d = 0; d < 3; ++d) rmax = std::max(rmax, s->gp.axis[d].resolution);
            const double atol = P.tolerance * 2.0 * rmax * std::sqrt(uu) / dt;
            int it = 0;
            double rr = 0;
            IBM_TRY(ibm_mg_solve_atol(ctx, s->pois, s->div, st->p, P.tolerance, atol, P.max_iterations, 1, &it, &rr));
            s->last_iters[3] = it;
            IBM_HIP(ctx, hipMemcpyAsync(s->info + 3, &rr, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
            IBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        }
        }
        double* out[3] = {st->u[0], st->u[1], st->u[2]};
        IBM_TRY(ibm_gradient(ctx, &s->gp, -dt, st->p, 1.0, out));
        // 7. FP64 divergence check (max |D u|, the field itself not stored)
        IBM_TRY(divergence_of(ctx, s->g, u, 1.0, nullptr, s->partial, s->info + 4));
    }
    // 8. interpolate at the data sites and move them
    {
        int counts[MAX_TYPES];
        const double* xs[MAX_TYPES];
        double* outs[MAX_TYPES];
        for (int t = 0; t < nt; ++t) {
            counts[t] = s->types[t].nd * s->types[t].nc;
            xs[t] = st->x[t];
            outs[t] = s->types[t].U;
        }
        const double* ue[3] = {st->u[0], st->u[1], st->u[2]};
        IBM_TRY(ibm_interpolate_sets(ctx, s->g, (ibm_delta) P.delta, nt, counts, xs, ue, outs));
        for (int t = 0; t < nt; ++t)
            if (counts[t])
                IBM_TRY(launch_axpby<double, double>(ctx, (size_t) 3 * counts[t], dt, s->types[t].U, 1.0, st->x[t]));
    }
    history_kernel<<<1, 64, 0, ctx->stream>>>(s->warm_hist, -1);
    IBM_LAUNCH_CHECK(ctx);
    if (!fixed) {
        // fixed-count steps use the largest counts of the recent adaptive steps
        for (int i = 0; i < 4; ++i) s->recent_iters[s->recent_next][i] = s->last_iters[i];
        s->recent_next = (s->recent_next + 1) % ITER_WINDOW;
        s->recent_len = std::min(s->recent_len + 1, ITER_WINDOW);
        for (int i = 0; i < 4; ++i) {
            s->max_iters[i] = 0;
            for (int k = 0; k < s->recent_len; ++k) s->max_iters[i] = std::max(s->max_iters[i], s->recent_iters[k][i]);
        }
        ++s->adaptive_steps;
    } else {
        for (int i = 0; i < 4; ++i) s->last_iters[i] = s->direct ? 1 : s->max_iters[i] + 1;
    }
    return IBM_OK;
}

} // namespace

extern "C" {

ibm_status ibm_solver_create(ibm_ctx* ctx, const ibm_dimension* domain, int ndims, int refinement,
                             const ibm_solver_params* params, int ntypes, const ibm_cell_type* types,
                             ibm_solver** out)
{
    if (!ctx || !out || !params) return IBM_ERR_ARG;
    *out = nullptr;
    if (ndims != 3) return fail(ctx, IBM_ERR_ARG, "the timestep needs a 3-dimensional domain");
    if (ntypes < 0 || ntypes > IBM_MAX_CELL_TYPES || (ntypes && !types))
        return fail(ctx, IBM_ERR_ARG, "between 0 and %d cell types", IBM_MAX_CELL_TYPES);
    const ibm_solver_params& P = *params;
    if (!(P.timestep > 0) || !(P.diffusivity >= 0) || !(P.density > 0) || !(P.tolerance > 0) || P.max_iterations < 1 ||
        !(P.fscale >= 0))
        return fail(ctx, IBM_ERR_ARG, "bad solver parameters");
    if (P.precision != 64 && P.precision != 32) return fail(ctx, IBM_ERR_ARG, "precision must be 64 or 32");
    if (P.flags & ~(IBM_SOLVER_PSEUDO_PRESSURE | IBM_SOLVER_MULTIGRID))
        return fail(ctx, IBM_ERR_ARG, "unknown solver flags 0x%x", P.flags);
    auto* s = new ibm_solver;
    s->par = P;
    ibm_status st = ibm_mac_grids(domain, ndims, refinement, s->g);
    // Pressure lives on the cell centres with homogeneous Neumann walls:
    // D (on the Dirichlet velocity grids) composed with G is that Laplacian.
    ibm_dimension pd[3];
    for (int d = 0; d < 3; ++d) {
        pd[d] = domain[d];
        if (pd[d].solid) {
            pd[d].lower[0] = pd[d].upper[0] = 0.0;
            pd[d].lower[1] = pd[d].upper[1] = 1.0;
        }
    }
    if (st == IBM_OK) st = ibm_centered_grid(pd, 3, refinement, &s->gp);
    if (st != IBM_OK) {
        release(s);
        return fail(ctx, st, "could not build the grids");
    }
    for (int i = 0; i < 3; ++i) {
        s->nu[i] = (size_t) ibm_grid_size(&s->g[i]);
        s->Bw[i] = make_ins_boundary(s->g[i], i, 2);
    }
    // The tile spread visits every marker once (the brick kernel about 2.6
    // times at C4 density); it needs periodic axes of whole bricks.
    s->tile_spread = ndims == 3 && tiles_usable(ctx, s->g);
    s->np = (size_t) ibm_grid_size(&s->gp);
    s->dt = P.timestep;
    // Channel grids (periodic x and z of power-of-two sizes, walls in y) are
    // solved directly; others, or IBM_SOLVER_MULTIGRID, by multigrid.
    s->direct = ndims == 3 && !(P.flags & IBM_SOLVER_MULTIGRID) && fft_usable(s->gp);
    for (int i = 0; i < 3; ++i) s->direct = s->direct && fft_usable(s->g[i]);
    if (s->direct) {
        for (int i = 0; i < 3 && st == IBM_OK; ++i) st = ibm_fft_create(ctx, &s->g[i], &s->fft_u[i]);
        if (st == IBM_OK) st = ibm_fft_create(ctx, &s->gp, &s->fft_p);
        for (int i = 0; i < 4 && st == IBM_OK; ++i) {
            const ibm_grid& g = i < 3 ? s->g[i] : s->gp;
            s->dparts[i] = residual_parts(g.axis[0].points, g.axis[1].points, g.axis[2].points);
            st = alloc(ctx, s, s->dparts[i], &s->dpartial[i]);
        }
    } else {
        if (!(P.fscale > 0)) {
            const double lambda = P.timestep * P.diffusivity / 2;
            for (int i = 0; i < 3 && st == IBM_OK; ++i)
                st = ibm_mg_create(ctx, &s->g[i], 1.0, -lambda, P.precision, &s->helm[i]);
        }
        if (st == IBM_OK) st = ibm_mg_create(ctx, &s->gp, 0.0, 1.0, P.precision, &s->pois);
    }
    for (int i = 0; i < 3 && st == IBM_OK; ++i) {
        if ((st = alloc(ctx, s, s->nu[i], &s->F[i]))) break;
        if ((st = alloc(ctx, s, s->nu[i], &s->H[i]))) break;
        if ((st = alloc(ctx, s, s->nu[i], &s->rhs[i]))) break;
        if ((st = alloc(ctx, s, s->nu[i], &s->du[i]))) break;
        st = alloc(ctx, s, s->nu[i], &s->du_prev[i]);
    }
    // The velocity increments are the warm start of the first Helmholtz solves.
    for (int i = 0; i < 3 && st == IBM_OK; ++i) {
        st = launch_fill<double>(ctx, s->nu[i], 0.0, s->du[i]);
        if (st == IBM_OK) st = launch_fill<double>(ctx, s->nu[i], 0.0, s->du_prev[i]);
    }
    if (st == IBM_OK) st = alloc(ctx, s, s->np, &s->p_prev);
    if (st == IBM_OK) st = launch_fill<double>(ctx, s->np, 0.0, s->p_prev);
    for (int k = 0; k < 2 && st == IBM_OK; ++k) {
        if (ibm_create(&s->aux[k], ctx->device) != IBM_OK) {
            st = fail(ctx, IBM_ERR_HIP, "solver: auxiliary context");
            break;
        }
        if (hipStreamCreateWithFlags(&s->aux_stream[k], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&s->join_ev[k], hipEventDisableTiming) != hipSuccess) {
            st = fail(ctx, IBM_ERR_HIP, "solver: auxiliary stream");
            break;
        }
        s->aux[k]->stream = s->aux_stream[k];
    }
    if (st == IBM_OK && (hipEventCreateWithFlags(&s->fork_ev, hipEventDisableTiming) != hipSuccess ||
                         hipEventCreateWithFlags(&s->geo_ev, hipEventDisableTiming) != hipSuccess))
        st = fail(ctx, IBM_ERR_HIP, "solver: fork event");
    if (st == IBM_OK) {
        double* h = nullptr;
        st = alloc(ctx, s, 1, &h);
        s->warm_hist = (int*) h;
        if (st == IBM_OK) {
            history_kernel<<<1, 64, 0, ctx->stream>>>(s->warm_hist, 0);
            st = hipGetLastError() == hipSuccess ? IBM_OK : fail(ctx, IBM_ERR_HIP, "history reset");
        }
    }
    if (st == IBM_OK) st = alloc(ctx, s, s->np, &s->div);
    if (st == IBM_OK) st = alloc(ctx, s, 16, &s->info);
    if (st == IBM_OK) st = launch_fill<double>(ctx, 16, 0.0, s->info);
    if (st == IBM_OK) st = alloc(ctx, s, REDUCE_PARTS, &s->partial);
    for (int t = 0; t < ntypes && st == IBM_OK; ++t) {
        const ibm_cell_type& c = types[t];
        if (!c.ops || c.data_sites <= 0 || c.sample_sites <= 0 || c.cells < 0 || !c.weights ||
            (c.stiffness != 0.0 && !c.tether)) {
            st = fail(ctx, IBM_ERR_ARG, "bad cell type %d", t);
            break;
        }
        ibm_solver::type_t T{};
        T.ops = c.ops;
        T.nd = c.data_sites;
        T.ns = c.sample_sites;
        T.nc = c.cells;
        T.stiffness = c.stiffness;
        T.mem = c.membrane;
        T.membrane = c.membrane.tension != IBM_TENSION_NONE || c.membrane.bending != 0.0;
        T.bend = c.membrane.bending != 0.0;
        T.dops = c.data_ops;
        if (c.membrane.tension != IBM_TENSION_NONE && c.membrane.tension != IBM_TENSION_SKALAK &&
            c.membrane.tension != IBM_TENSION_NEOHOOKEAN) {
            st = fail(ctx, IBM_ERR_ARG, "cell type %d: unknown tension law %d", t, c.membrane.tension);
            break;
        }
        if (T.membrane && (rbf_nops(c.ops) < 6 || !c.reference)) {
            st = fail(ctx, IBM_ERR_ARG, "cell type %d: membrane models need the six sample operators (position, two "
                      "first and three second derivatives) and the reference shape", t);
            break;
        }
        if (T.bend && rbf_nops(c.data_ops) < 5) {
            st = fail(ctx, IBM_ERR_ARG, "cell type %d: bending needs the five data-site derivative operators", t);
            break;
        }
        const int nblocks = T.membrane ? 6 : 3;
        size_t blk = (size_t) T.ns * 3 * T.nc;
        if ((st = alloc(ctx, s, T.ns, &T.w)) || (st = alloc(ctx, s, (size_t) nblocks * blk, &T.Y)) ||
            (st = alloc(ctx, s, blk, &T.f)) || (st = alloc(ctx, s, blk, &T.normal)) ||
            (st = alloc(ctx, s, (size_t) T.ns * T.nc, &T.sigma)) || (st = alloc(ctx, s, T.nc, &T.area)) ||
            (st = alloc(ctx, s, T.nc, &T.volume)) || (st = alloc(ctx, s, (size_t) 3 * T.nd * T.nc, &T.U)))
            break;
        if (hipMemcpyAsync(T.w, c.weights, sizeof(double) * T.ns, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess) {
            st = fail(ctx, IBM_ERR_HIP, "copying cell type %d", t);
            break;
        }
        if (c.stiffness != 0.0) {
            if ((st = alloc(ctx, s, blk, &T.ref))) break;
            if (hipMemcpyAsync(T.ref, c.tether, sizeof(double) * blk, hipMemcpyDeviceToDevice, ctx->stream) !=
                hipSuccess) {
                st = fail(ctx, IBM_ERR_HIP, "copying cell type %d", t);
                break;
            }
        }
        if (T.membrane) {
            // The undeformed shape's sample geometry (one cell) and its table
            // of metric terms, areas and mean curvatures.
            double* refY = nullptr;
            if ((st = alloc(ctx, s, 6 * (size_t) T.ns * 3, &refY)) || (st = alloc(ctx, s, 14 * (size_t) T.ns, &T.table)))
                break;
            int three = 3;
            const double* rx = c.reference;
            unsigned char* packed = nullptr;
            if ((st = rbf_eval_range(ctx, ctx->stream, 1, &T.ops, &three, &rx, &refY, 0, 6, true, &packed))) break;
            if ((st = membrane_table(ctx, ctx->stream, T.ns, refY + (size_t) T.ns * 3, T.w, T.table))) break;
        }
        if (T.bend) {
            if ((st = alloc(ctx, s, 5 * (size_t) T.nd * 3 * T.nc, &T.Yd)) ||
                (st = alloc(ctx, s, (size_t) T.nd * T.nc, &T.h)) ||
                (st = alloc(ctx, s, 6 * (size_t) T.ns * T.nc, &T.hd)))
                break;
            if (T.mem.preferred) {
                // H of the undeformed shape at the data sites
                double* refYd = nullptr;
                if ((st = alloc(ctx, s, 5 * (size_t) T.nd * 3, &refYd)) || (st = alloc(ctx, s, T.nd, &T.h0))) break;
                int three = 3;
                const double* rx = c.reference;
                unsigned char* packed = nullptr;
                if ((st = rbf_eval_range(ctx, ctx->stream, 1, &T.dops, &three, &rx, &refYd, 0, 5, true, &packed)))
                    break;
                if ((st = mean_curvature(ctx, ctx->stream, T.nd, 1, refYd, nullptr, T.h0))) break;
            }
        }
        s->types.push_back(T);
    }
    if (st == IBM_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) st = fail(ctx, IBM_ERR_HIP, "solver setup");
    if (st != IBM_OK) {
        release(s);
        return st;
    }
    *out = s;
    return IBM_OK;
}

ibm_status ibm_solver_destroy(ibm_solver* s)
{
    if (s) release(s);
    return IBM_OK;
}

ibm_status ibm_solver_step(ibm_ctx* ctx, ibm_solver* s, ibm_state* st, const double* const* boundary)
{
    if (!ctx || !s) return IBM_ERR_ARG;
    IBM_TRY(check_state(ctx, s, st));
    IBM_TRY(step_impl(ctx, s, st, boundary, false));
    st->t += s->dt;
    return IBM_OK;
}

ibm_status ibm_solver_step_fixed(ibm_ctx* ctx, ibm_solver* s, ibm_state* st, const double* const* boundary)
{
    if (!ctx || !s) return IBM_ERR_ARG;
    IBM_TRY(check_state(ctx, s, st));
    if (!s->adaptive_steps) return fail(ctx, IBM_ERR_ARG, "run an adaptive step first to calibrate the solves");
    IBM_TRY(sync_history(ctx, s, s->dt));
    IBM_TRY(step_impl(ctx, s, st, boundary, true));
    st->t += s->dt;
    return IBM_OK;
}

ibm_status ibm_solver_capture(ibm_ctx* ctx, ibm_solver* s, const ibm_state* st, const double* const* boundary)
{
    if (!ctx || !s) return IBM_ERR_ARG;
    IBM_TRY(check_state(ctx, s, st));
    if (!s->adaptive_steps) return fail(ctx, IBM_ERR_ARG, "run an adaptive step first to calibrate the solves");
    if (!ctx->stream) return fail(ctx, IBM_ERR_ARG, "graph capture needs a created stream, not the null stream");
    if (s->exec) {
        hipGraphExecDestroy(s->exec);
        s->exec = nullptr;
    }
    if (s->graph) {
        hipGraphDestroy(s->graph);
        s->graph = nullptr;
    }
    const bool profiling = ctx->profiling;
    ctx->profiling = false;
    ibm_state copy = *st;
    IBM_TRY(sync_history(ctx, s, s->dt));
    IBM_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ibm_status rc = step_impl(ctx, s, &copy, boundary, true);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    ctx->profiling = profiling;
    if (rc != IBM_OK) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamEndCapture");
    s->graph = g;
    IBM_HIP(ctx, hipGraphInstantiate(&s->exec, g, nullptr, nullptr, 0));
    s->captured = *st;
    s->captured_dt = s->dt;
    return IBM_OK;
}

ibm_status ibm_solver_replay(ibm_ctx* ctx, ibm_solver* s, ibm_state* st)
{
    if (!ctx || !s) return IBM_ERR_ARG;
    if (!s->exec) return fail(ctx, IBM_ERR_ARG, "no captured step");
    if (!st || std::memcmp(st->u, s->captured.u, sizeof(st->u)) || st->p != s->captured.p ||
        std::memcmp(st->x, s->captured.x, sizeof(st->x)))
        return fail(ctx, IBM_ERR_ARG, "the state's buffers differ from the captured ones");
    IBM_TRY(sync_history(ctx, s, s->captured_dt));
    IBM_HIP(ctx, hipGraphLaunch(s->exec, ctx->stream));
    st->t += s->captured_dt;
    return IBM_OK;
}

ibm_status ibm_solver_info(ibm_ctx* ctx, ibm_solver* s, ibm_step_info* out)
{
    if (!ctx || !s || !out) return IBM_ERR_ARG;
    double h[5];
    IBM_HIP(ctx, hipMemcpyAsync(h, s->info, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    IBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < 4; ++i) {
        out->iterations[i] = s->last_iters[i];
        out->residual[i] = h[i];
    }
    out->divergence = h[4];
    out->timestep = s->dt;
    out->exponent = s->last_exponent;
    return IBM_OK;
}

ibm_status ibm_solver_cells(ibm_solver* s, int type, ibm_cell_view* out)
{
    if (!s || !out || type < 0 || type >= (int) s->types.size()) return IBM_ERR_ARG;
    const auto& T = s->types[type];
    size_t blk = (size_t) T.ns * 3 * T.nc;
    out->positions = T.Y;
    out->tangents[0] = T.Y + blk;
    out->tangents[1] = T.Y + 2 * blk;
    out->forces = T.f;
    out->normals = T.normal;
    out->sigma = T.sigma;
    out->area = T.area;
    out->volume = T.volume;
    out->velocity = T.U;
    return IBM_OK;
}

ibm_status ibm_solver_layout(ibm_solver* s, ibm_state_layout* out)
{
    if (!s || !out) return IBM_ERR_ARG;
    *out = ibm_state_layout{};
    out->dims = 3;
    for (int i = 0; i < 3; ++i) out->u_size[i] = (long long) s->nu[i];
    out->p_size = (long long) s->np;
    out->ntypes = (int) s->types.size();
    for (int k = 0; k < out->ntypes; ++k) {
        out->rows[k] = s->types[k].nd;
        out->cols[k] = 3 * s->types[k].nc;
    }
    return IBM_OK;
}

ibm_status ibm_solver_grids(ibm_solver* s, ibm_grid* velocity /* [3] */, ibm_grid* pressure)
{
    if (!s) return IBM_ERR_ARG;
    if (velocity)
        for (int i = 0; i < 3; ++i) velocity[i] = s->g[i];
    if (pressure) *pressure = s->gp;
    return IBM_OK;
}

} // extern "C"
