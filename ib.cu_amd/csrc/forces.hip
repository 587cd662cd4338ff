// Membrane forces of the cell surfaces at their sample sites.
//
// Reference models (include/forces/): tension with a constitutive law
// (forces/tension.h:50-99; Skalak forces/skalak.h:11-36, neo-Hookean
// forces/neohookean.h:10-34) and bending (forces/bending.h:142-223,
// F = -4 kappa s (Lap(H - H0) + 2 (H - H0)(H^2 - K + H H0)) n).  Inputs are
// the surface geometry the RBF operators produce: per point the parameter
// derivatives x_u, x_v, x_uu, x_uv, x_vv (u = theta, v = phi), the unit
// normal and the patch area s, and for bending the parameter derivatives of
// h = H - H0 (H at the data sites, differentiated at the sample sites by the
// same operators).  The reference shape is one undeformed cell shared by all
// cells of a type: point k reads reference point k % ns (forces/load.h:96-110).
//
// Layouts (as bases::geometry holds them): a geometry block of rows sites x
// nc cells stores component i of point k (k = cell * rows + site) at
// [i * P + k], P = rows * nc; the five derivative blocks of a geometry
// follow each other, 3P apart.  Scalar fields (h, s, its derivatives) are
// P-long blocks.
//
// All three kernels are elementwise FP64 passes: one lane per point, the
// reference's formulas term for term (the reference table precomputes the
// per-point metric terms of the undeformed shape once per cell type).
#include <hip/hip_runtime.h>
#include "context.hpp"

namespace ibm {
namespace {

// Reference table rows (one entry per reference sample site).
enum { R_E, R_F, R_G, R_EU, R_FU, R_GU, R_EV, R_FV, R_GV, R_I, R_IU, R_IV, R_S, R_H, R_ROWS };

struct vec3 { double x, y, z; };

__device__ inline vec3 load3(const double* p, size_t P, size_t k) { return {p[k], p[P + k], p[2 * P + k]}; }
__device__ inline double dot3(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline vec3 cross3(vec3 a, vec3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// The metric terms of forces/tension.h:34-48 (`helper`): e, f, g, their u and
// v derivatives through the Christoffel symbols Gamma_{ij,k} = x_ij . x_k
// (forces/load.h:293-333), i = eg - f^2 and its derivatives.
struct metric {
    double e, f, g, eu, fu, gu, ev, fv, gv, i, iu, iv;
};

__device__ inline metric metric_of(vec3 u, vec3 v, vec3 uu, vec3 uv, vec3 vv)
{
    metric m;
    m.e = dot3(u, u);
    m.f = dot3(u, v);
    m.g = dot3(v, v);
    m.eu = 2 * dot3(uu, u);
    m.fu = dot3(uu, v) + dot3(uv, u);
    m.gu = 2 * dot3(uv, v);
    m.ev = 2 * dot3(uv, u);
    m.fv = dot3(uv, v) + dot3(vv, u);
    m.gv = 2 * dot3(vv, v);
    m.i = m.e * m.g - m.f * m.f;
    m.iu = m.eu * m.g + m.e * m.gu - 2 * m.f * m.fu;
    m.iv = m.ev * m.g + m.e * m.gv - 2 * m.f * m.fv;
    return m;
}

// Mean curvature (forces/bending.h:48-59) with n the unit normal.
__device__ inline double mean_of(const metric& m, vec3 uu, vec3 uv, vec3 vv, vec3 n)
{
    double l = dot3(uu, n), k = dot3(uv, n), mm = dot3(vv, n);
    return (l * m.g + mm * m.e - 2 * k * m.f) / (2 * m.i);
}

__device__ inline double gaussian_of(const metric& m, vec3 uu, vec3 uv, vec3 vv, vec3 n)
{
    double l = dot3(uu, n), k = dot3(uv, n), mm = dot3(vv, n);
    return (l * mm - k * k) / m.i;
}

struct geo_point {
    vec3 u, v, uu, uv, vv;
};

__device__ inline geo_point load_geo(const double* geo, size_t P, size_t k)
{
    geo_point g;
    g.u = load3(geo, P, k);
    g.v = load3(geo + 3 * P, P, k);
    g.uu = load3(geo + 6 * P, P, k);
    g.uv = load3(geo + 9 * P, P, k);
    g.vv = load3(geo + 12 * P, P, k);
    return g;
}

__device__ inline vec3 unit_normal(vec3 u, vec3 v)
{
    vec3 c = cross3(u, v);
    double d = sqrt(dot3(c, c));
    return {c.x / d, c.y / d, c.z / d};
}

// Per reference sample site: the metric terms, the patch area s = w |x_u x
// x_v| and the mean curvature (the preferred curvature of bending).
__global__ void reference_table_kernel(int ns, const double* __restrict__ geo, const double* __restrict__ w,
                                       double* __restrict__ table)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    geo_point g = load_geo(geo, ns, k);
    metric m = metric_of(g.u, g.v, g.uu, g.uv, g.vv);
    vec3 c = cross3(g.u, g.v);
    double det = sqrt(dot3(c, c));
    vec3 n = {c.x / det, c.y / det, c.z / det};
    double vals[R_ROWS] = {m.e, m.f, m.g, m.eu, m.fu, m.gu, m.ev, m.fv, m.gv, m.i, m.iu, m.iv,
                           w[k] * det, mean_of(m, g.uu, g.uv, g.vv, n)};
    for (int r = 0; r < R_ROWS; ++r) table[(size_t) r * ns + k] = vals[r];
}

// h = H (minus the reference's H at the same site when h0 is given) at the
// data sites (forces/bending.h:142-158).
__global__ void mean_curvature_kernel(size_t P, int nd, const double* __restrict__ geo, const double* __restrict__ h0,
                                      double* __restrict__ h)
{
    size_t k = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P) return;
    geo_point g = load_geo(geo, P, k);
    metric m = metric_of(g.u, g.v, g.uu, g.uv, g.vv);
    double H = mean_of(m, g.uu, g.uv, g.vv, unit_normal(g.u, g.v));
    if (h0) H -= h0[k % nd];
    h[k] = H;
}

struct membrane_args {
    int tension;          // IBM_TENSION_*
    int preferred;
    double shear, bulk, bending;
    double tether;        // stiffness of f = -k (y - y_ref), 0 = none
};

// Constitutive laws: (w1, w2, w11, w12, w22) of (i1, i2).
__device__ inline void law(int kind, double shear, double bulk, double i1, double i2, double w[5])
{
    if (kind == IBM_TENSION_SKALAK) {                 // forces/skalak.h:20-30
        w[0] = shear * (i1 - 1);
        w[1] = bulk * (i2 - 1) - shear;
        w[2] = shear;
        w[3] = 0.0;
        w[4] = bulk;
    } else {                                          // forces/neohookean.h:20-31
        double j = sqrt(i2);
        double w1 = shear / j;
        w[0] = w1;
        w[1] = bulk * (1 - 1 / j) - 0.5 * w1 * i1 / i2;
        w[2] = 0.0;
        w[3] = -0.5 * w1 / i2;
        w[4] = 0.75 * w1 * i1 / (i2 * i2) + 0.5 * bulk / (i2 * j);
    }
}

// f = beta f + tension + bending + tether, one lane per sample point.
__global__ void __launch_bounds__(256)
membrane_force_kernel(size_t P, int ns, membrane_args a, const double* __restrict__ geo,
                      const double* __restrict__ normal, const double* __restrict__ sigma,
                      const double* __restrict__ hd, const double* __restrict__ table,
                      const double* __restrict__ y, const double* __restrict__ yref, double beta,
                      double* __restrict__ f)
{
    size_t k = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P) return;
    const size_t r = k % (size_t) ns;
    vec3 out = {0.0, 0.0, 0.0};
    geo_point g{};
    metric c{};
    if (a.tension || a.bending != 0.0) {
        g = load_geo(geo, P, k);
        c = metric_of(g.u, g.v, g.uu, g.uv, g.vv);
    }
    if (a.tension) {
        // forces/tension.h:66-96
        const double oe = table[R_E * ns + r], of = table[R_F * ns + r], og = table[R_G * ns + r];
        const double oeu = table[R_EU * ns + r], ofu = table[R_FU * ns + r], ogu = table[R_GU * ns + r];
        const double oev = table[R_EV * ns + r], ofv = table[R_FV * ns + r], ogv = table[R_GV * ns + r];
        const double oi = table[R_I * ns + r], oiu = table[R_IU * ns + r], oiv = table[R_IV * ns + r];
        const double os = table[R_S * ns + r];
        const double i1 = (c.e * og + oe * c.g - 2 * of * c.f) / oi;
        const double i2 = c.i / oi;
        const double i1u = (c.eu * og + c.e * ogu + oeu * c.g + oe * c.gu - 2 * (ofu * c.f + of * c.fu) - i1 * oiu) / oi;
        const double i1v = (c.ev * og + c.e * ogv + oev * c.g + oe * c.gv - 2 * (ofv * c.f + of * c.fv) - i1 * oiv) / oi;
        const double i2u = (c.iu - i2 * oiu) / oi;
        const double i2v = (c.iv - i2 * oiv) / oi;
        double w[5];
        law(a.tension, a.shear, a.bulk, i1, i2, w);
        const double w1 = w[0], w2 = w[1], w11 = w[2], w12 = w[3], w22 = w[4];
        const double w1u = w11 * i1u + w12 * i2u, w1v = w11 * i1v + w12 * i2v;
        const double w2u = w12 * i1u + w22 * i2u, w2v = w12 * i1v + w22 * i2v;
        const double A = w1 * og + w2 * c.g, B = w1 * of + w2 * c.f, C = w1 * oe + w2 * c.e;
        const double cu = -A * oiu / (2 * oi) + B * oiv / (2 * oi) + (w1u * og + w1 * ogu + w2u * c.g + w2 * c.gu) -
                          (w1v * of + w1 * ofv + w2v * c.f + w2 * c.fv);
        const double cv = B * oiu / (2 * oi) - C * oiv / (2 * oi) - (w1u * of + w1 * ofu + w2u * c.f + w2 * c.fu) +
                          (w1v * oe + w1 * oev + w2v * c.e + w2 * c.ev);
        const double sc = os / oi;
        out.x += sc * (A * g.uu.x - 2 * B * g.uv.x + C * g.vv.x + cu * g.u.x + cv * g.v.x);
        out.y += sc * (A * g.uu.y - 2 * B * g.uv.y + C * g.vv.y + cu * g.u.y + cv * g.v.y);
        out.z += sc * (A * g.uu.z - 2 * B * g.uv.z + C * g.vv.z + cu * g.u.z + cv * g.v.z);
    }
    if (a.bending != 0.0) {
        // forces/bending.h:160-223
        const vec3 n = load3(normal, P, k);
        const double s = sigma[k];
        const double hu = hd[k], hv = hd[P + k], huu = hd[2 * P + k], huv = hd[3 * P + k], hvv = hd[4 * P + k];
        const double lh = (((-c.iu * c.g + c.iv * c.f) / (2 * c.i) + c.gu - c.fv) / c.i * hu +
                           ((c.iu * c.f - c.iv * c.e) / (2 * c.i) - c.fu + c.ev) / c.i * hv) +
                          c.g / c.i * huu - 2 * c.f / c.i * huv + c.e / c.i * hvv;
        const double H = mean_of(c, g.uu, g.uv, g.vv, n);
        const double K = gaussian_of(c, g.uu, g.uv, g.vv, n);
        double dh = H, sh = H * H - K;
        if (a.preferred) {
            const double H0 = table[R_H * ns + r];
            dh -= H0;
            sh += H * H0;
        }
        const double mag = -4 * s * a.bending * (lh + 2 * dh * sh);
        out.x += mag * n.x;
        out.y += mag * n.y;
        out.z += mag * n.z;
    }
    if (a.tether != 0.0) {
        out.x += -a.tether * (y[k] - yref[k]);
        out.y += -a.tether * (y[P + k] - yref[P + k]);
        out.z += -a.tether * (y[2 * P + k] - yref[2 * P + k]);
    }
    if (beta != 0.0) {
        out.x += beta * f[k];
        out.y += beta * f[P + k];
        out.z += beta * f[2 * P + k];
    }
    f[k] = out.x;
    f[P + k] = out.y;
    f[2 * P + k] = out.z;
}

int blocks_of(size_t n) { return (int) ((n + 255) / 256); }

} // namespace

ibm_status membrane_table(ibm_ctx* ctx, hipStream_t stream, int ns, const double* geo, const double* w, double* table)
{
    if (ns <= 0 || !geo || !w || !table) return fail(ctx, IBM_ERR_ARG, "bad reference geometry");
    reference_table_kernel<<<blocks_of(ns), 256, 0, stream>>>(ns, geo, w, table);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

ibm_status mean_curvature(ibm_ctx* ctx, hipStream_t stream, int nd, int nc, const double* geo, const double* h0,
                          double* h)
{
    if (nd <= 0 || nc < 0 || !geo || !h) return fail(ctx, IBM_ERR_ARG, "bad data geometry");
    size_t P = (size_t) nd * nc;
    if (!P) return IBM_OK;
    prof_scope prof(ctx, "membrane", 8.0 * (16.0 * P + (h0 ? nd : 0)));
    mean_curvature_kernel<<<blocks_of(P), 256, 0, stream>>>(P, nd, geo, h0, h);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

ibm_status membrane_forces(ibm_ctx* ctx, hipStream_t stream, const ibm_membrane* m, double tether, int ns, int nc,
                           const double* geo, const double* normal, const double* sigma, const double* hd,
                           const double* table, const double* y, const double* yref, double beta, double* f)
{
    if (ns <= 0 || nc < 0 || !f) return fail(ctx, IBM_ERR_ARG, "bad force arguments");
    membrane_args a{};
    if (m) {
        if (m->tension != IBM_TENSION_NONE && m->tension != IBM_TENSION_SKALAK && m->tension != IBM_TENSION_NEOHOOKEAN)
            return fail(ctx, IBM_ERR_ARG, "unknown tension law %d", m->tension);
        a.tension = m->tension;
        a.preferred = m->preferred != 0;
        a.shear = m->shear;
        a.bulk = m->bulk;
        a.bending = m->bending;
    }
    a.tether = tether;
    if ((a.tension || a.bending != 0.0) && (!geo || !table))
        return fail(ctx, IBM_ERR_ARG, "membrane forces need the geometry and the reference table");
    if (a.bending != 0.0 && (!normal || !sigma || !hd)) return fail(ctx, IBM_ERR_ARG, "bending needs n, s and h derivatives");
    if (a.tether != 0.0 && (!y || !yref)) return fail(ctx, IBM_ERR_ARG, "tether needs the positions and their reference");
    size_t P = (size_t) ns * nc;
    if (!P) return IBM_OK;
    // reads: 15 geometry + 3 normal + s + 5 h derivatives + 6 tether per point; writes 3
    double work = 8.0 * P * ((a.tension || a.bending != 0.0 ? 15 : 0) + (a.bending != 0.0 ? 9 : 0) +
                             (a.tether != 0.0 ? 6 : 0) + (beta != 0.0 ? 3 : 0) + 3);
    prof_scope prof(ctx, "membrane", work);
    membrane_force_kernel<<<blocks_of(P), 256, 0, stream>>>(P, ns, a, geo, normal, sigma, hd, table, y, yref, beta, f);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

} // namespace ibm

extern "C" {

ibm_status ibm_membrane_reference(ibm_ctx* ctx, int ns, const double* geo, const double* w, double* table)
{
    if (!ctx) return IBM_ERR_ARG;
    return ibm::membrane_table(ctx, ctx->stream, ns, geo, w, table);
}

ibm_status ibm_mean_curvature(ibm_ctx* ctx, int nd, int nc, const double* geo, const double* h0, double* h)
{
    if (!ctx) return IBM_ERR_ARG;
    return ibm::mean_curvature(ctx, ctx->stream, nd, nc, geo, h0, h);
}

ibm_status ibm_membrane_forces(ibm_ctx* ctx, const ibm_membrane* m, int ns, int nc, const double* geo,
                               const double* normal, const double* sigma, const double* hd, const double* table,
                               double beta, double* f)
{
    if (!ctx) return IBM_ERR_ARG;
    if (!m) return ibm::fail(ctx, IBM_ERR_ARG, "null membrane model");
    return ibm::membrane_forces(ctx, ctx->stream, m, 0.0, ns, nc, geo, normal, sigma, hd, table, nullptr, nullptr,
                                beta, f);
}

} // extern "C"
