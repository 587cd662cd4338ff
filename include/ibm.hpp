// C++ API over libibm's C ABI, in the shapes of the reference's headers:
//
//   fd::dimension / fd::domain / fd::mac / fd::grid      fd/dimension.h, fd/domain.h, fd/cell.h, fd/grid.h
//   fd::laplacian / identity / differential / average      fd/laplacian.h, fd/identity.h, fd/differential.h, fd/average.h
//   fd::lower_boundary / upper_boundary                    fd/boundary_ops.h
//   ib::delta::{cosine, peskin, bspline<3>}                ib/cosine.h, ib/peskin.h, ib/bspline.h
//   ib::spread / ib::interpolate                           ib/novel.h:193-250, ib/interpolate.h:88-138
//   algo::gershgorin / algo::chebyshev                     algo/gershgorin.h, algo/chebyshev.h
//   solvers::multigrid / solvers::chebpcg                  solvers/multigrid.h, solvers/chebpcg.h
//   ins::divergence / gradient / advection / diffusion     ins/differential.h, ins/advection.h, ins/diffusion.h
//   bases::operators (slice * matrix), bases::geometry     bases/operators.h:95-109, bases/geometry.h:74-123
//   ib::solver / ib::state                                 test/ib-whole.cc:159,226,275
//
// Differences from the reference's value-semantic device containers: vectors
// are caller-owned device pointers (double*), and every call takes the
// device and stream from an ibm::context (the thread's default one unless
// given).  Status codes become exceptions: ibm::error (a std::runtime_error
// carrying the code), with ins::too_small for IBM_ERR_DT_TOO_SMALL as in
// ins/exceptions.h.
#pragma once
#include <array>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "ibm.h"

namespace ibm {

struct error : std::runtime_error {
    ibm_status code;
    error(ibm_status c, const std::string& what) : std::runtime_error(what), code(c) {}
};

// Owns an ibm_ctx (device + stream + scratch).  Move-only.
class context {
public:
    explicit context(int device = 0)
    {
        ibm_ctx* c = nullptr;
        ibm_status s = ibm_create(&c, device);
        if (s != IBM_OK) throw error(s, "ibm_create failed");
        ctx_.reset(c);
    }
    ibm_ctx* get() const { return ctx_.get(); }
    void set_stream(void* stream) { check(ibm_set_stream(get(), stream)); }
    void synchronize() { check(ibm_synchronize(get())); }
    void check(ibm_status s) const
    {
        if (s == IBM_OK) return;
        throw error(s, ibm_last_error(get()));
    }

    // The calling thread's default context (device 0, null stream).
    static context& current()
    {
        thread_local context c{0};
        return c;
    }

private:
    struct del {
        void operator()(ibm_ctx* c) const { ibm_destroy(c); }
    };
    std::unique_ptr<ibm_ctx, del> ctx_;
};

inline void check(ibm_status s)
{
    if (s != IBM_OK) throw error(s, "ibm call failed");
}

} // namespace ibm

// Units (units.h:16-276): every quantity is a double in library units,
// 1 length unit = 10 um (1 cm = 1000), 1 mass unit = 1 ug (1 g = 10^6),
// 1 time unit = 1 ms (1 s = 1000).  The reference's unit<d, m, t> wrapper
// converts implicitly to and from double; here the types are doubles.
namespace units {

template <long dist, long mass, long time>
using unit = double;
using scalar = unit<0, 0, 0>;
using length = unit<1, 0, 0>;
using mass = unit<0, 1, 0>;
using time = unit<0, 0, 1>;
using force = unit<1, 1, -2>;
using velocity = unit<1, 0, -1>;
using acceleration = unit<1, 0, -2>;
using hertz = unit<0, 0, -1>;
using viscosity = unit<-1, 1, -1>;
using volume = unit<3, 0, 0>;
using area = unit<2, 0, 0>;
using diffusivity = unit<2, 0, -1>;
using density = unit<-3, 1, 0>;
using work = unit<2, 1, -2>;
using energy = work;

// SI units in library units (m = 100 cm of 1000 units; N = kg m / s^2;
// P = 0.1 kg / (m s); J = N m).
inline constexpr double m = 1e5;
inline constexpr double g = 1e6;
inline constexpr double s = 1e3;
inline constexpr double kg = 1e9;
inline constexpr double N = 1e8;
inline constexpr double L = 1e12;
inline constexpr double mL = 1e9;
inline constexpr double P = 1.0;
inline constexpr double J = 1e13;

namespace literals {
#define IBM_UNIT_LITERAL(suffix, factor)                                                   \
    constexpr double operator""_##suffix(long double v) { return (double) (v * (long double) (factor)); } \
    constexpr double operator""_##suffix(unsigned long long v) { return (double) v * (factor); }
IBM_UNIT_LITERAL(m, 1e5)
IBM_UNIT_LITERAL(kg, 1e9)
IBM_UNIT_LITERAL(s, 1e3)
IBM_UNIT_LITERAL(N, 1e8)
IBM_UNIT_LITERAL(cm, 1e3)
IBM_UNIT_LITERAL(mm, 1e2)
IBM_UNIT_LITERAL(um, 0.1)
IBM_UNIT_LITERAL(nm, 1e-4)
IBM_UNIT_LITERAL(L, 1e12)
IBM_UNIT_LITERAL(mL, 1e9)
IBM_UNIT_LITERAL(g, 1e6)
IBM_UNIT_LITERAL(ms, 1.0)
IBM_UNIT_LITERAL(us, 1e-3)
IBM_UNIT_LITERAL(ns, 1e-6)
IBM_UNIT_LITERAL(dyn, 1e3)
IBM_UNIT_LITERAL(P, 1.0)
IBM_UNIT_LITERAL(cP, 1e-2)
IBM_UNIT_LITERAL(J, 1e13)
IBM_UNIT_LITERAL(erg, 1e6)
#undef IBM_UNIT_LITERAL
} // namespace literals
} // namespace units

// As units.h does, the literals are visible everywhere the header is.
using namespace units::literals;

namespace fd {

// Boundary kinds (fd/boundary.h:31-42): `periodic`, `dirichlet` and
// `neumann` are objects, usable as `fd::boundary::periodic` or, like a
// constructor call, `fd::boundary::periodic()`; robin(a, b) describes
// a u + b du/dn = gamma.
namespace boundary {
struct robin {
    double a, b;
    constexpr robin(double alpha, double beta) : a(alpha), b(beta) {}
};
inline constexpr struct periodic_t {
    constexpr periodic_t operator()() const { return *this; }
} periodic{};
inline constexpr struct dirichlet_t : robin {
    constexpr dirichlet_t() : robin(1.0, 0.0) {}
    constexpr dirichlet_t operator()() const { return *this; }
} dirichlet{};
inline constexpr struct neumann_t : robin {
    constexpr neumann_t() : robin(0.0, 1.0) {}
    constexpr neumann_t operator()() const { return *this; }
} neumann{};
} // namespace boundary

struct dimension {
    ibm_dimension d;
    constexpr dimension(double length, boundary::periodic_t) : d{length, 0, 0, {0, 0}, {0, 0}} {}
    constexpr dimension(double length, boundary::robin both) : d{length, 1, 0, {both.a, both.b}, {both.a, both.b}} {}
    constexpr dimension(double length, boundary::robin lower, boundary::robin upper) :
        d{length, 1, 0, {lower.a, lower.b}, {upper.a, upper.b}} {}
    constexpr double length() const { return d.length; }
};

class domain {
public:
    constexpr domain(dimension x, dimension y, dimension z) : dims_{x.d, y.d, z.d} {}
    double unit() const
    {
        double u = 0;
        ibm::check(ibm_domain_unit(dims_.data(), 3, &u));
        return u;
    }
    const ibm_dimension* data() const { return dims_.data(); }

private:
    std::array<ibm_dimension, 3> dims_;
};

struct mac {
    int r;
    constexpr explicit mac(int refinement) : r(refinement) {}
    constexpr int refinement() const { return r; }
};

struct centered {
    int r;
    constexpr explicit centered(int refinement) : r(refinement) {}
    constexpr int refinement() const { return r; }
};

// One discretised component: fd::grid{tag, domain[, component]}.
struct grid {
    ibm_grid g;
    grid() : g{} {}
    explicit grid(const ibm_grid& x) : g(x) {}
    grid(const mac& m, const domain& dom, int component)
    {
        ibm_grid all[3];
        ibm::check(ibm_mac_grids(dom.data(), 3, m.refinement(), all));
        g = all[component];
    }
    grid(const centered& c, const domain& dom) { ibm::check(ibm_centered_grid(dom.data(), 3, c.refinement(), &g)); }
    long long size() const { return ibm_grid_size(&g); }
    const ibm_grid* get() const { return &g; }
};

inline std::array<grid, 3> components(const mac& m, const domain& dom)
{
    ibm_grid all[3];
    ibm::check(ibm_mac_grids(dom.data(), 3, m.refinement(), all));
    return {grid{all[0]}, grid{all[1]}, grid{all[2]}};
}

// sigma * I~ + tau * L, applied matrix-free: y = alpha A x + beta y.
struct operator_ {
    grid g;
    double sigma, tau;
    void operator()(double alpha, const double* x, double beta, double* y,
                    ibm::context& ctx = ibm::context::current()) const
    {
        ctx.check(ibm_op_apply(ctx.get(), g.get(), sigma, tau, alpha, x, beta, y));
    }
};

inline operator_ laplacian(const grid& g) { return {g, 0.0, 1.0}; }
inline operator_ identity(const grid& g) { return {g, 1.0, 0.0}; }
inline operator_ helmholtz(const grid& g, double lambda) { return {g, 1.0, -lambda}; }

struct differential {
    grid g;
    int view;
    grid range() const
    {
        grid out;
        ibm::check(ibm_grid_shift(g.get(), view, &out.g));
        return out;
    }
    void operator()(double alpha, const double* x, double beta, double* y,
                    ibm::context& ctx = ibm::context::current()) const
    {
        ctx.check(ibm_differential(ctx.get(), g.get(), view, alpha, x, beta, y));
    }
};

struct average {
    grid g;
    int view;
    void operator()(double alpha, const double* x, double beta, double* y,
                    ibm::context& ctx = ibm::context::current()) const
    {
        ctx.check(ibm_average(ctx.get(), g.get(), view, alpha, x, beta, y));
    }
};

// fd::lower_boundary / fd::upper_boundary (correction order 0 or 2) times
// face values: y += alpha B g.
inline void apply_boundary(const grid& g, int view, bool upper, int order, double alpha, const double* values,
                           double value, double* y, ibm::context& ctx = ibm::context::current())
{
    ctx.check(ibm_boundary_apply(ctx.get(), g.get(), view, upper ? 1 : 0, order, alpha, values, value, y));
}

} // namespace fd

namespace algo {

inline std::pair<double, double> gershgorin(const fd::operator_& A)
{
    double lo, hi;
    ibm::check(ibm_gershgorin(A.g.get(), A.sigma, A.tau, &lo, &hi));
    return {lo, hi};
}

// algo::chebyshev(a, b, A)(r): the degree-4 Chebyshev polynomial in A
// applied to r.  `work` holds 2 * size doubles.
struct chebyshev {
    double a, b;
    fd::operator_ A;
    void operator()(const double* r, double* out, double* work, ibm::context& ctx = ibm::context::current()) const
    {
        ctx.check(ibm_chebyshev_apply(ctx.get(), A.g.get(), A.sigma, A.tau, a, b, r, out, work));
    }
};

} // namespace algo

namespace ib {

namespace delta {
struct cosine {
    static constexpr ibm_delta kind = IBM_DELTA_COSINE;
};
struct peskin {
    static constexpr ibm_delta kind = IBM_DELTA_PESKIN;
};
template <int n>
struct bspline;
template <>
struct bspline<3> {
    static constexpr ibm_delta kind = IBM_DELTA_BSPLINE3;
};
} // namespace delta

// ib::spread{fd::mac{r}, domain[, delta]}: forces at n markers (SoA x, f)
// onto the three component grids (ib/novel.h:197-250).  Constructible in
// constant expressions, as the reference's is; the component grids are
// derived (on the host) at each call.  The delta defaults to the cosine
// kernel.
template <typename Delta = delta::cosine>
class spread {
public:
    constexpr spread(const fd::mac& m, const fd::domain& dom, Delta = {}) : mac_(m), dom_(dom) {}
    void operator()(int n, const double* x, const double* f, std::array<double*, 3> fe,
                    ibm::context& ctx = ibm::context::current()) const
    {
        const auto c = grids();
        ibm_grid g[3] = {c[0].g, c[1].g, c[2].g};
        ctx.check(ibm_spread(ctx.get(), g, Delta::kind, n, x, f, fe.data()));
    }
    std::array<fd::grid, 3> grids() const { return fd::components(mac_, dom_); }

private:
    fd::mac mac_;
    fd::domain dom_;
};

// ib::interpolate{fd::mac{r}, domain[, delta]} (ib/interpolate.h:92-138).
template <typename Delta = delta::cosine>
class interpolate {
public:
    constexpr interpolate(const fd::mac& m, const fd::domain& dom, Delta = {}) : mac_(m), dom_(dom) {}
    void operator()(int n, const double* x, std::array<const double*, 3> ue, double* ul,
                    ibm::context& ctx = ibm::context::current()) const
    {
        const auto c = fd::components(mac_, dom_);
        ibm_grid g[3] = {c[0].g, c[1].g, c[2].g};
        ctx.check(ibm_interpolate(ctx.get(), g, Delta::kind, n, x, ue.data(), ul));
    }

private:
    fd::mac mac_;
    fd::domain dom_;
};

} // namespace ib

namespace solvers {

// solvers::multigrid over sigma I~ + tau L with Chebyshev smoothing,
// precision 64 or 32 (FP32 V-cycles inside FP64 defect correction).
class multigrid {
public:
    multigrid(const fd::operator_& A, double tol, int precision = 64, ibm::context& ctx = ibm::context::current()) :
        tol_(tol)
    {
        ibm_mg* m = nullptr;
        ctx.check(ibm_mg_create(ctx.get(), A.g.get(), A.sigma, A.tau, precision, &m));
        mg_.reset(m);
    }
    // x = A^-1 b to relative tolerance; returns the V-cycles used.
    int operator()(const double* b, double* x, int maxit = 100, ibm::context& ctx = ibm::context::current()) const
    {
        int it = 0;
        double rr = 0;
        ctx.check(ibm_mg_solve(ctx.get(), mg_.get(), b, x, tol_, maxit, &it, &rr));
        return it;
    }
    ibm_mg* get() const { return mg_.get(); }

private:
    struct del {
        void operator()(ibm_mg* m) const { ibm_mg_destroy(m); }
    };
    double tol_;
    std::unique_ptr<ibm_mg, del> mg_;
};

// solvers::chebpcg(A, tol): `work` holds 6 * size doubles.
struct chebpcg {
    fd::operator_ A;
    double tol;
    int operator()(const double* b, double* x, double* work, int maxit = 10000,
                   ibm::context& ctx = ibm::context::current()) const
    {
        int it = 0;
        double rr = 0;
        ctx.check(ibm_chebpcg_solve(ctx.get(), A.g.get(), A.sigma, A.tau, b, x, tol, maxit, work, &it, &rr));
        return it;
    }
};

} // namespace solvers

namespace ins {

struct too_small : std::runtime_error {
    too_small() : std::runtime_error("timestep too small") {}
};

// ins::divergence{mac, domain}(u) -> out on the cell centres.
struct divergence {
    std::array<fd::grid, 3> g;
    divergence(const fd::mac& m, const fd::domain& dom) : g(fd::components(m, dom)) {}
    void operator()(std::array<const double*, 3> u, double* out, ibm::context& ctx = ibm::context::current()) const
    {
        ibm_grid gg[3] = {g[0].g, g[1].g, g[2].g};
        ctx.check(ibm_divergence(ctx.get(), gg, u.data(), out));
    }
};

// ins::gradient of a cell-centred field: out_i = alpha D_i p + beta out_i.
struct gradient {
    fd::grid p;
    gradient(const fd::grid& centred) : p(centred) {}
    void operator()(double alpha, const double* pv, double beta, std::array<double*, 3> out,
                    ibm::context& ctx = ibm::context::current()) const
    {
        ctx.check(ibm_gradient(ctx.get(), p.get(), alpha, pv, beta, out.data()));
    }
};

struct advection {
    std::array<fd::grid, 3> g;
    advection(const fd::mac& m, const fd::domain& dom) : g(fd::components(m, dom)) {}
    void operator()(std::array<const double*, 3> u, std::array<double*, 3> H,
                    ibm::context& ctx = ibm::context::current()) const
    {
        ibm_grid gg[3] = {g[0].g, g[1].g, g[2].g};
        ctx.check(ibm_advection(ctx.get(), gg, u.data(), H.data()));
    }
};

struct parameters {
    double timestep, diffusivity, density, tolerance;
    double fscale = 0;   // > 0: adaptive step size (ins::tracker), timestep = k_min
};

// ins::diffusion(grid, params)(gamma, u, rhs, f) (ins/diffusion.h:46-56):
// rhs <- k (nu (rhs + L u) + I~ f), k = gamma dt; solves
// (I~ - dt nu / 2 L) du = rhs and writes u_out = u + du.  `rhs` is
// overwritten; `du` is a caller work vector of the grid's size.
class diffusion {
public:
    diffusion(const fd::grid& g, const parameters& p, int precision = 64, ibm::context& ctx = ibm::context::current()) :
        g_(g), p_(p), mg_(fd::helmholtz(g, p.timestep * p.diffusivity / 2), p.tolerance, precision, ctx) {}
    void operator()(double gamma, const double* u, double* rhs, const double* f, double* du, double* u_out,
                    ibm::context& ctx = ibm::context::current()) const
    {
        const double k = gamma * p_.timestep;
        const long long n = g_.size();
        ctx.check(ibm_op_apply(ctx.get(), g_.get(), 0.0, 1.0, k * p_.diffusivity, u, k * p_.diffusivity, rhs));
        ctx.check(ibm_op_apply(ctx.get(), g_.get(), 1.0, 0.0, k, f, 1.0, rhs));
        mg_(rhs, du, 100, ctx);
        if (u_out != u) ctx.check(ibm_axpby(ctx.get(), n, 1.0, u, 0.0, u_out));
        ctx.check(ibm_axpby(ctx.get(), n, 1.0, du, 1.0, u_out));
    }

private:
    fd::grid g_;
    parameters p_;
    solvers::multigrid mg_;
};

} // namespace ins

namespace bases {

// The RBF operators of one cell type packed for evaluation
// (bases::operators: evaluator and first derivatives; slice * matrix).
class operators {
public:
    // E[j]: device (lde x M) column-major, first K rows used.
    operators(int precision, int M, int K, const std::vector<const double*>& E, int lde,
              ibm::context& ctx = ibm::context::current()) :
        M_(M), K_(K), nops_((int) E.size())
    {
        ibm_rbf* op = nullptr;
        ctx.check(ibm_rbf_create(ctx.get(), precision, M, K, nops_, E.data(), lde, &op));
        op_.reset(op);
    }
    // Y = [E_j^T X]_j, X: K x N column-major, Y: nops blocks of M x N.
    void operator()(int N, const double* X, double* Y, ibm::context& ctx = ibm::context::current()) const
    {
        ibm_rbf* o = op_.get();
        ctx.check(ibm_rbf_eval(ctx.get(), 1, &o, &N, &X, &Y));
    }
    ibm_rbf* get() const { return op_.get(); }
    int sample_sites() const { return M_; }
    int data_sites() const { return K_; }

private:
    struct del {
        void operator()(ibm_rbf* o) const { ibm_rbf_destroy(o); }
    };
    int M_, K_, nops_;
    std::unique_ptr<ibm_rbf, del> op_;
};

} // namespace bases

namespace ib {

using state = ibm_state;

// ib::solver{mac, domain, delta, params} with cell types; step(state, ub).
class solver {
public:
    template <typename Delta>
    solver(const fd::mac& m, const fd::domain& dom, Delta, const ins::parameters& p, int precision,
           const std::vector<ibm_cell_type>& types, ibm::context& ctx = ibm::context::current())
    {
        ibm_solver_params par{p.timestep, p.diffusivity, p.density, p.tolerance, (int) Delta::kind, precision, 100, 0,
                              p.fscale};
        ibm_solver* s = nullptr;
        ctx.check(ibm_solver_create(ctx.get(), dom.data(), 3, m.refinement(), &par, (int) types.size(), types.data(),
                                    &s));
        s_.reset(s);
    }
    void step(state& st, const double* const* ub = nullptr, ibm::context& ctx = ibm::context::current()) const
    {
        ibm_status r = ibm_solver_step(ctx.get(), s_.get(), &st, ub);
        if (r == IBM_ERR_DT_TOO_SMALL) throw ins::too_small();
        ctx.check(r);
    }
    void capture(const state& st, const double* const* ub = nullptr, ibm::context& ctx = ibm::context::current()) const
    {
        ctx.check(ibm_solver_capture(ctx.get(), s_.get(), &st, ub));
    }
    void replay(state& st, ibm::context& ctx = ibm::context::current()) const
    {
        ctx.check(ibm_solver_replay(ctx.get(), s_.get(), &st));
    }
    ibm_step_info info(ibm::context& ctx = ibm::context::current()) const
    {
        ibm_step_info i;
        ctx.check(ibm_solver_info(ctx.get(), s_.get(), &i));
        return i;
    }
    // Shapes of this solver's states (checkpoint / resume).
    ibm_state_layout layout() const
    {
        ibm_state_layout l{};
        ibm::check(ibm_solver_layout(s_.get(), &l));
        return l;
    }
    ibm_solver* get() const { return s_.get(); }

private:
    struct del {
        void operator()(ibm_solver* s) const { ibm_solver_destroy(s); }
    };
    std::unique_ptr<ibm_solver, del> s_;
};

} // namespace ib

// Checkpoint / resume in the linalg binary format (linalg/io.h).  The
// reference streams a state with `f << linalg::io::binary << st` and reads
// it back with `f >> linalg::io::binary >> st` (test/ib-whole.cc:184-187);
// here the file is named by its path and the state's shapes come from
// ib::solver::layout().
namespace linalg {
namespace io {

inline constexpr struct binary_t {} binary{};

// Appends (or, with append = false, writes) one state record.
inline void write(const std::string& path, const ib::state& st, const ibm_state_layout& layout, bool append = true,
                  ibm::context& ctx = ibm::context::current())
{
    ctx.check(ibm_state_write(ctx.get(), path.c_str(), append ? 1 : 0, &st, &layout));
}

// Reads record `record` (-1: the last complete one, util/kaboose.cc) into
// the state's buffers; returns the number of complete records in the file.
inline int read(const std::string& path, ib::state& st, const ibm_state_layout& layout, int record = -1,
                ibm::context& ctx = ibm::context::current())
{
    int n = 0;
    ctx.check(ibm_state_read(ctx.get(), path.c_str(), record, &st, &layout, &n));
    return n;
}

} // namespace io
} // namespace linalg
