// Direct solution of A x = b, A = sigma I~ + tau L, on grids that are
// periodic and uniform in x and z (power-of-two point counts) and walled in
// y -- the channel of the timestep's benchmark.
//
// Row (i, j, k) of the operator (fd/laplacian.h:53-68, fd/identity.h:25-51)
// is sigma Iy_j x + tau (Iy_j Lx x + Ly x + Iy_j Lz x): the identity factors
// of periodic axes are 1 and their 3-point Laplacians are circulant, so the
// discrete Fourier transform in x and z diagonalises them,
//   Lx -> -4 rx^2 sin^2(pi kx / nx),   Lz -> -4 rz^2 sin^2(pi kz / nz),
// and every Fourier mode (kx, kz) leaves one tridiagonal system along y,
//   Iy_j (sigma + tau mu) X_j + tau ry^2 (lo_j X_{j-1} + d_j X_j + hi_j X_{j+1})
//   = B_j,   mu = lambda_x(kx) + lambda_z(kz),
// with the wall rows' ghost coefficients in d_0, d_{n-1} (fd/discretization.h:
// 94-108).  The solve is: real-to-half-complex FFT along x, complex FFT
// along z, one tridiagonal solve per mode (parallel cyclic reduction in
// LDS), and the inverse
// transforms -- the exact solution of the same discrete system the multigrid
// iterates on, in five memory passes.  The singular case (sigma = 0 with
// Neumann walls: the pressure Poisson) solves the mean-free problem, as the
// multigrid does: the (0, 0) mode's right-hand side loses its mean, its last
// unknown is pinned and the solution's mean is removed.
//
// The reference solves these systems with multigrid or PCG
// (solvers/multigrid.h, solvers/pcg.h); this is a different algorithm for the
// same discrete equations, used where the grid allows it.
#include <hip/hip_runtime.h>
#include <cmath>
#include <mutex>
#include <vector>
#include "context.hpp"
#include "fd_kernels.hpp"
#include "fd_ops.hpp"

namespace ibm {
namespace {

constexpr int FFT_THREADS = 256;
constexpr int ZT = 8;              // kx columns per z-transform block
constexpr int MAX_FFT_POINTS = 512;
constexpr int MAX_WALL_POINTS = 800;
constexpr size_t MAX_DYNAMIC_LDS = 160 * 1024;

// Row stride of the half-spectrum in coefficients: nh rounded up to 8 (rows
// start on 128-byte boundaries, so the column tiles of the z transform and
// the cyclic reduction read whole cache lines).
__host__ __device__ constexpr int spec_ld(int nh) { return (nh + 7) & ~7; }

__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

__device__ inline int bitrev(int k, int logn) { return (int) (__brev((unsigned) k) >> (32 - logn)); }

// In-place radix-2 transforms of `count` arrays of n points held in LDS in
// bit-reversed order: point k of array t at a[idx(t, k)] holds input
// bitrev(k).  Rows layout (idx = t n + k, the x transforms) maps consecutive
// lanes to consecutive butterflies of one array; columns layout (idx =
// k count + t, the z transforms, whose rows of count arrays are contiguous)
// maps consecutive lanes to consecutive arrays, so that the lanes of a wave
// touch consecutive LDS words either way.  tw[m] = exp(-2 pi i m / n),
// m < n / 2, also in LDS (a stage's twiddle reads would otherwise be
// dependent global loads); `inverse` conjugates the twiddles (no scaling).
// Every thread of the block takes part in every stage.  Stages are done in
// pairs: a thread loads the quad k0, k0 + h, k0 + 2h, k0 + 3h (h = 2^(s-1)),
// does stage s's two butterflies and stage s + 1's two, and stores it -- the
// radix-2 butterflies and twiddles exactly, with one LDS round trip and one
// barrier per two stages.
__device__ inline void butterfly(double2& u, double2& v, double2 w, bool inverse)
{
    if (inverse) w.y = -w.y;
    const double2 t = cmul(v, w);
    v = make_double2(u.x - t.x, u.y - t.y);
    u = make_double2(u.x + t.x, u.y + t.y);
}

template <bool COLS>
__device__ void fft_stages(double2* a, int count, int n, int logn, const double2* tw, bool inverse)
{
    int s = 1;
    const int quads = count * (n >> 2);
    for (; s + 1 <= logn; s += 2) {
        const int h = 1 << (s - 1);
        const int st0 = n >> s, st1 = n >> (s + 1);       // twiddle index steps of stages s and s + 1
        for (int b = threadIdx.x; b < quads; b += blockDim.x) {
            const int t = COLS ? b % count : b / (n >> 2), bb = COLS ? b / count : b - t * (n >> 2);
            const int grp = bb >> (s - 1), pos = bb & (h - 1);
            const int k0 = grp * 4 * h + pos;
            const int i0 = COLS ? k0 * count + t : t * n + k0;
            const int dh = COLS ? h * count : h;
            double2 q0 = a[i0], q1 = a[i0 + dh], q2 = a[i0 + 2 * dh], q3 = a[i0 + 3 * dh];
            const double2 w0 = tw[pos * st0];
            butterfly(q0, q1, w0, inverse);
            butterfly(q2, q3, w0, inverse);
            butterfly(q0, q2, tw[pos * st1], inverse);
            butterfly(q1, q3, tw[(pos + h) * st1], inverse);
            a[i0] = q0;
            a[i0 + dh] = q1;
            a[i0 + 2 * dh] = q2;
            a[i0 + 3 * dh] = q3;
        }
        __syncthreads();
    }
    if (s == logn) {
        const int half_n = n >> 1, half = 1 << (s - 1), stride = n >> s;
        const int total = count * half_n;
        for (int b = threadIdx.x; b < total; b += blockDim.x) {
            const int t = COLS ? b % count : b / half_n, bb = COLS ? b / count : b - t * half_n;
            const int grp = bb >> (s - 1), pos = bb & (half - 1);
            const int k0 = grp * 2 * half + pos, k1 = k0 + half;
            const int i = COLS ? k0 * count + t : t * n + k0, j = COLS ? k1 * count + t : t * n + k1;
            double2 u = a[i], v = a[j];
            butterfly(u, v, tw[pos * stride], inverse);
            a[i] = u;
            a[j] = v;
        }
        __syncthreads();
    }
}

// Forward transform along x of `rows` rows of nx reals into nh = nx / 2 + 1
// coefficients each (row r of x -> row r of s, x fastest in both).  The
// bit reversal is applied to the (cached) global reads, so the LDS writes
// are consecutive.
__global__ void __launch_bounds__(FFT_THREADS) xfft_forward_kernel(int nx, int logn, int rows,
                                                                   const double* __restrict__ x,
                                                                   double2* __restrict__ s,
                                                                   const double2* __restrict__ tw)
{
    extern __shared__ double2 sh[];
    const int rpb = FFT_THREADS / (nx / 2), nh = nx / 2 + 1;
    double2* stw = sh + rpb * nx;
    const int r0 = blockIdx.x * rpb;
    for (int e = threadIdx.x; e < nx / 2; e += blockDim.x) stw[e] = tw[e];
    for (int e = threadIdx.x; e < rpb * nx; e += blockDim.x) {
        const int t = e / nx, k = e - t * nx, row = r0 + t;
        sh[e] = make_double2(row < rows ? x[(size_t) row * nx + bitrev(k, logn)] : 0.0, 0.0);
    }
    __syncthreads();
    fft_stages<false>(sh, rpb, nx, logn, stw, false);
    for (int e = threadIdx.x; e < rpb * nh; e += blockDim.x) {
        const int t = e / nh, k = e - t * nh, row = r0 + t;
        if (row < rows) s[(size_t) row * spec_ld(nh) + k] = sh[t * nx + k];
    }
}

// Inverse of xfft_forward_kernel (the missing coefficients are the
// conjugates of the stored ones), scaled by `scale`.
__global__ void __launch_bounds__(FFT_THREADS) xfft_inverse_kernel(int nx, int logn, int rows,
                                                                   const double2* __restrict__ s,
                                                                   double* __restrict__ x, double scale,
                                                                   const double2* __restrict__ tw)
{
    extern __shared__ double2 sh[];
    const int rpb = FFT_THREADS / (nx / 2), nh = nx / 2 + 1;
    double2* stw = sh + rpb * nx;
    const int r0 = blockIdx.x * rpb;
    for (int e = threadIdx.x; e < nx / 2; e += blockDim.x) stw[e] = tw[e];
    for (int e = threadIdx.x; e < rpb * nx; e += blockDim.x) {
        const int t = e / nx, k = bitrev(e - t * nx, logn), row = r0 + t;
        double2 v = make_double2(0.0, 0.0);
        if (row < rows) {
            if (k < nh) {
                v = s[(size_t) row * spec_ld(nh) + k];
            } else {
                v = s[(size_t) row * spec_ld(nh) + (nx - k)];
                v.y = -v.y;
            }
        }
        sh[e] = v;
    }
    __syncthreads();
    fft_stages<false>(sh, rpb, nx, logn, stw, true);
    for (int e = threadIdx.x; e < rpb * nx; e += blockDim.x) {
        const int t = e / nx, k = e - t * nx, row = r0 + t;
        if (row < rows) x[(size_t) row * nx + k] = sh[t * nx + k].x * scale;
    }
}

// Transform along z, in place, of s laid out [k][j][kx] (kx fastest, nh of
// them in rows of spec_ld(nh)): block (kx tile, j) transforms ZT columns,
// held in LDS as rows of ZT consecutive kx.
__global__ void __launch_bounds__(FFT_THREADS) zfft_kernel(int nh, int ny, int nz, int logn, double2* __restrict__ s,
                                                           const double2* __restrict__ tw, int inverse)
{
    extern __shared__ double2 sh[];
    double2* stw = sh + ZT * nz;
    const int kx0 = blockIdx.x * ZT, j = blockIdx.y;
    const int ld = spec_ld(nh);
    const size_t plane = (size_t) ld * ny;
    for (int e = threadIdx.x; e < nz / 2; e += blockDim.x) stw[e] = tw[e];
    for (int e = threadIdx.x; e < ZT * nz; e += blockDim.x) {
        const int t = e % ZT, k = e / ZT, kx = kx0 + t, kr = bitrev(k, logn);
        sh[e] = kx < nh ? s[(size_t) kr * plane + (size_t) j * ld + kx] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_stages<true>(sh, ZT, nz, logn, stw, inverse != 0);
    for (int e = threadIdx.x; e < ZT * nz; e += blockDim.x) {
        const int t = e % ZT, k = e / ZT, kx = kx0 + t;
        if (kx < nh) s[(size_t) k * plane + (size_t) j * ld + kx] = sh[e];
    }
}

struct wall_axis {
    int n;
    double r2, lap_lo, lap_hi, id_lo, id_hi;
};

// Row j's coefficients of mode (kx, kz): sub- and super-diagonal lo, hi and
// the diagonal.
__device__ inline void tri_row(const wall_axis& y, double sigma, double tau, double mu, int j, double& lo, double& hi,
                               double& diag)
{
    const double ty = tau * y.r2;
    lo = j > 0 ? ty : 0.0;
    hi = j < y.n - 1 ? ty : 0.0;
    double d = -2.0, id = 1.0;
    if (j == 0) {
        d += y.lap_lo;
        id += y.id_lo;
    }
    if (j == y.n - 1) {
        d += y.lap_hi;
        id += y.id_hi;
    }
    diag = id * (sigma + tau * mu) + ty * d;
}

// Every Fourier mode's tridiagonal system along y by parallel cyclic
// reduction in LDS: block (kx tile of zt modes, kz) holds zt x np2 rows
// (np2 = ny rounded up to a power of two; rows past ny are identity rows),
// 256 threads update PCR_ITEMS rows each per step (row j combined with rows
// j -+ 2^s) until P interleaved subsystems of np2 / P <= PCR_SPLIT rows are
// left (P >= PCR_SPLIT where np2 allows), which one thread each solves by
// the Thomas algorithm.  The singular mode's last
// equation is replaced by x = 0 after its right-hand side loses its mean,
// and its solution loses its mean.
constexpr int PCR_THREADS = 256, PCR_ITEMS = 4, PCR_ROWS = PCR_THREADS * PCR_ITEMS, PCR_SPLIT = 16;

__global__ void __launch_bounds__(PCR_THREADS) pcr_kernel(int nh, int zt, int np2, wall_axis y, double sigma,
                                                          double tau, int singular, const double* __restrict__ lam_x,
                                                          const double* __restrict__ lam_z, double2* __restrict__ s)
{
    extern __shared__ double2 sh[];
    double2* D = sh;
    double* A = (double*) (sh + PCR_ROWS);
    double* B = A + PCR_ROWS;
    double* C = B + PCR_ROWS;
    const int ny = y.n, kx0 = blockIdx.x * zt, kz = blockIdx.y;
    const int ld = spec_ld(nh);
    const size_t plane = (size_t) ld * ny;
    const bool pin_block = singular && kx0 == 0 && kz == 0;     // holds the singular mode (t = 0)
    for (int e = threadIdx.x; e < PCR_ROWS; e += PCR_THREADS) {
        const int t = e % zt, j = e / zt, kx = kx0 + t;
        double a = 0.0, b = 1.0, c = 0.0;
        double2 d = make_double2(0.0, 0.0);
        if (kx < nh && j < ny) {
            tri_row(y, sigma, tau, lam_x[kx] + lam_z[kz], j, a, c, b);
            d = s[(size_t) kz * plane + (size_t) j * ld + kx];
        }
        A[e] = a;
        B[e] = b;
        C[e] = c;
        D[e] = d;
    }
    __syncthreads();
    if (pin_block && threadIdx.x == 0) {
        double2 m = make_double2(0.0, 0.0);
        for (int j = 0; j < ny; ++j) {
            m.x += D[j * zt].x;
            m.y += D[j * zt].y;
        }
        m.x /= ny;
        m.y /= ny;
        for (int j = 0; j < ny - 1; ++j) {
            D[j * zt].x -= m.x;
            D[j * zt].y -= m.y;
        }
        const int last = (ny - 1) * zt;
        A[last] = 0.0;
        B[last] = 1.0;
        C[last] = 0.0;
        D[last] = make_double2(0.0, 0.0);
    }
    __syncthreads();
    // Reduction until the rows left coupled (j -+ P) form P interleaved
    // subsystems of at most PCR_SPLIT rows, then one Thomas sweep per
    // subsystem.
    const int P = max(min(np2, PCR_SPLIT), np2 / PCR_SPLIT);
    for (int sft = 1; sft < P; sft <<= 1) {
        double na[PCR_ITEMS], nb[PCR_ITEMS], nc[PCR_ITEMS];
        double2 nd[PCR_ITEMS];
#pragma unroll
        for (int i = 0; i < PCR_ITEMS; ++i) {
            const int e = threadIdx.x + PCR_THREADS * i;
            const int t = e % zt, j = e / zt;
            double a = A[e], b = B[e], c = C[e];
            double2 d = D[e];
            if (j - sft >= 0) {
                const int m = (j - sft) * zt + t;
                const double al = -a / B[m];
                b += al * C[m];
                d.x += al * D[m].x;
                d.y += al * D[m].y;
                a = al * A[m];
            } else {
                a = 0.0;
            }
            if (j + sft < np2) {
                const int p = (j + sft) * zt + t;
                const double ga = -c / B[p];
                b += ga * A[p];
                d.x += ga * D[p].x;
                d.y += ga * D[p].y;
                c = ga * C[p];
            } else {
                c = 0.0;
            }
            na[i] = a;
            nb[i] = b;
            nc[i] = c;
            nd[i] = d;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < PCR_ITEMS; ++i) {
            const int e = threadIdx.x + PCR_THREADS * i;
            A[e] = na[i];
            B[e] = nb[i];
            C[e] = nc[i];
            D[e] = nd[i];
        }
        __syncthreads();
    }
    const int L = np2 / P, step = P * zt;
    for (int id = threadIdx.x; id < zt * P; id += PCR_THREADS) {
        int e = id;                                  // row r = id / zt of mode t = id % zt
        double inv = 1.0 / B[e];
        double cp = C[e] * inv;
        double2 dp = make_double2(D[e].x * inv, D[e].y * inv);
        C[e] = cp;
        D[e] = dp;
        for (int m = 1; m < L; ++m) {
            e += step;
            const double a = A[e];
            inv = 1.0 / (B[e] - a * cp);
            cp = C[e] * inv;
            dp = make_double2((D[e].x - a * dp.x) * inv, (D[e].y - a * dp.y) * inv);
            C[e] = cp;
            D[e] = dp;
        }
        double2 xv = dp;
        for (int m = L - 2; m >= 0; --m) {
            e -= step;
            xv = make_double2(D[e].x - C[e] * xv.x, D[e].y - C[e] * xv.y);
            D[e] = xv;
        }
    }
    __syncthreads();
    if (pin_block && threadIdx.x == 0) {
        double2 m = make_double2(0.0, 0.0);
        for (int j = 0; j < ny; ++j) {
            m.x += D[j * zt].x;
            m.y += D[j * zt].y;
        }
        m.x /= ny;
        m.y /= ny;
        for (int j = 0; j < ny; ++j) {
            D[j * zt].x -= m.x;
            D[j * zt].y -= m.y;
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < PCR_ROWS; e += PCR_THREADS) {
        const int t = e % zt, j = e / zt, kx = kx0 + t;
        if (kx < nh && j < ny) s[(size_t) kz * plane + (size_t) j * ld + kx] = D[e];
    }
}

int log2_exact(int n)
{
    int l = 0;
    while ((1 << l) < n) ++l;
    return (1 << l) == n ? l : -1;
}

std::vector<double2> twiddles(int n)
{
    std::vector<double2> w(n / 2);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int m = 0; m < n / 2; ++m) {
        const long double a = -2.0L * pi * m / n;
        w[m] = make_double2((double) cosl(a), (double) sinl(a));
    }
    return w;
}

std::vector<double> eigenvalues(int n, double r2, int count)
{
    std::vector<double> l(count);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < count; ++k) {
        const long double sn = sinl(pi * k / n);
        l[k] = (double) (-4.0L * (long double) r2 * sn * sn);
    }
    return l;
}

} // namespace
} // namespace ibm

struct ibm_fft {
    int device;
    int nx, ny, nz, nh, lx, lz;
    ibm::wall_axis y;
    bool neumann;                 // both walls Neumann: sigma = 0 is the singular (mean-free) case
    double2* spec = nullptr;      // nz x ny rows of nh coefficients, row stride spec_ld(nh)
    double2* twx = nullptr;
    double2* twz = nullptr;
    double* lam_x = nullptr;
    double* lam_z = nullptr;
    std::vector<void*> owned;
};

namespace ibm {

// Whether the FFT solver takes grid g (and why not).
static const char* fft_unsupported(const ibm_grid& g)
{
    if (g.axis[0].solid || g.axis[2].solid) return "x and z must be periodic";
    if (!g.axis[1].solid) return "y must be walled";
    const int nx = g.axis[0].points, nz = g.axis[2].points, ny = g.axis[1].points;
    if (log2_exact(nx) < 2 || log2_exact(nz) < 2 || nx > MAX_FFT_POINTS || nz > MAX_FFT_POINTS)
        return "x and z point counts must be powers of two from 4 to 512";
    if (ny < 2 || ny > MAX_WALL_POINTS) return "y must hold 2 to 800 points";
    return nullptr;
}

bool fft_usable(const ibm_grid& g) { return fft_unsupported(g) == nullptr; }

static ibm_status fft_alloc(ibm_ctx* ctx, ibm_fft* p, size_t bytes, void** out)
{
    IBM_HIP(ctx, hipMalloc(out, bytes));
    p->owned.push_back(*out);
    return IBM_OK;
}

static ibm_status fft_lds_attributes(ibm_ctx* ctx)
{
    static std::mutex lock;
    static bool done = false;
    std::lock_guard<std::mutex> guard(lock);
    if (done) return IBM_OK;
    IBM_HIP(ctx, hipFuncSetAttribute((const void*) zfft_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int) MAX_DYNAMIC_LDS));
    IBM_HIP(ctx, hipFuncSetAttribute((const void*) pcr_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int) MAX_DYNAMIC_LDS));
    done = true;
    return IBM_OK;
}

ibm_status fft_solve(ibm_ctx* ctx, ibm_fft* p, double sigma, double tau, const double* b, double* x)
{
    if (!ctx || !p) return IBM_ERR_ARG;
    if (!b || !x) return fail(ctx, IBM_ERR_ARG, "null vector");
    if (p->device != ctx->device) return fail(ctx, IBM_ERR_ARG, "FFT solver lives on another device");
    if (!std::isfinite(sigma) || !std::isfinite(tau)) return fail(ctx, IBM_ERR_ARG, "non-finite operator coefficients");
    const int rows = p->ny * p->nz;
    const double points = (double) p->nx * p->ny * p->nz, coeffs = (double) p->nh * p->ny * p->nz;
    // b read and x written once; the coefficients written and read four times
    prof_scope prof(ctx, "fft_solve", 8.0 * 2.0 * points + 16.0 * 8.0 * coeffs);
    const int rpb = FFT_THREADS / (p->nx / 2);
    const size_t xlds = sizeof(double2) * (rpb * p->nx + p->nx / 2);
    xfft_forward_kernel<<<(rows + rpb - 1) / rpb, FFT_THREADS, xlds, ctx->stream>>>(p->nx, p->lx, rows, b, p->spec,
                                                                                      p->twx);
    IBM_LAUNCH_CHECK(ctx);
    const dim3 zgrid((p->nh + ZT - 1) / ZT, p->ny);
    const size_t zlds = sizeof(double2) * (ZT * p->nz + p->nz / 2);
    zfft_kernel<<<zgrid, FFT_THREADS, zlds, ctx->stream>>>(p->nh, p->ny, p->nz, p->lz, p->spec, p->twz, 0);
    IBM_LAUNCH_CHECK(ctx);
    const int singular = sigma == 0.0 && p->neumann;
    int np2 = 1;
    while (np2 < p->ny) np2 <<= 1;
    const int zt = PCR_ROWS / np2;
    const dim3 tgrid((p->nh + zt - 1) / zt, p->nz);
    const size_t tlds = (sizeof(double2) + 3 * sizeof(double)) * PCR_ROWS;
    pcr_kernel<<<tgrid, PCR_THREADS, tlds, ctx->stream>>>(p->nh, zt, np2, p->y, sigma, tau, singular, p->lam_x, p->lam_z,
                                                          p->spec);
    IBM_LAUNCH_CHECK(ctx);
    zfft_kernel<<<zgrid, FFT_THREADS, zlds, ctx->stream>>>(p->nh, p->ny, p->nz, p->lz, p->spec, p->twz, 1);
    IBM_LAUNCH_CHECK(ctx);
    xfft_inverse_kernel<<<(rows + rpb - 1) / rpb, FFT_THREADS, xlds, ctx->stream>>>(
        p->nx, p->lx, rows, p->spec, x, 1.0 / ((double) p->nx * p->nz), p->twx);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

} // namespace ibm

using namespace ibm;

extern "C" {

ibm_status ibm_fft_create(ibm_ctx* ctx, const ibm_grid* g, ibm_fft** out)
{
    if (!ctx || !g || !out) return IBM_ERR_ARG;
    *out = nullptr;
    if (const char* why = fft_unsupported(*g)) return fail(ctx, IBM_ERR_BAD_GRID, "FFT solver: %s", why);
    IBM_TRY(fft_lds_attributes(ctx));
    auto* p = new ibm_fft;
    p->device = ctx->device;
    p->nx = g->axis[0].points;
    p->ny = g->axis[1].points;
    p->nz = g->axis[2].points;
    p->nh = p->nx / 2 + 1;
    p->lx = log2_exact(p->nx);
    p->lz = log2_exact(p->nz);
    const op_axis<double> ya = make_op_axis<double>(g->axis[1]);
    p->y = wall_axis{ya.n, ya.r2, ya.lap_lo, ya.lap_hi, ya.id_lo, ya.id_hi};
    p->neumann = g->axis[1].interior[0] == 1.0 && g->axis[1].interior[1] == 1.0;
    const std::vector<double2> twx = twiddles(p->nx), twz = twiddles(p->nz);
    const std::vector<double> lx = eigenvalues(p->nx, g->axis[0].resolution * g->axis[0].resolution, p->nh);
    const std::vector<double> lz = eigenvalues(p->nz, g->axis[2].resolution * g->axis[2].resolution, p->nz);
    void* m = nullptr;
    ibm_status st = IBM_OK;
    if ((st = fft_alloc(ctx, p, sizeof(double2) * spec_ld(p->nh) * p->ny * (size_t) p->nz, &m)) == IBM_OK) {
        p->spec = (double2*) m;
        st = fft_alloc(ctx, p, sizeof(double2) * twx.size(), &m);
    }
    if (st == IBM_OK) {
        p->twx = (double2*) m;
        st = fft_alloc(ctx, p, sizeof(double2) * twz.size(), &m);
    }
    if (st == IBM_OK) {
        p->twz = (double2*) m;
        st = fft_alloc(ctx, p, sizeof(double) * lx.size(), &m);
    }
    if (st == IBM_OK) {
        p->lam_x = (double*) m;
        st = fft_alloc(ctx, p, sizeof(double) * lz.size(), &m);
    }
    if (st == IBM_OK) {
        p->lam_z = (double*) m;
        if (hipMemcpy(p->twx, twx.data(), sizeof(double2) * twx.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(p->twz, twz.data(), sizeof(double2) * twz.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(p->lam_x, lx.data(), sizeof(double) * lx.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(p->lam_z, lz.data(), sizeof(double) * lz.size(), hipMemcpyHostToDevice) != hipSuccess)
            st = fail(ctx, IBM_ERR_HIP, "FFT solver: table upload failed");
    }
    if (st != IBM_OK) {
        for (void* q : p->owned) hipFree(q);
        delete p;
        return st;
    }
    *out = p;
    return IBM_OK;
}

ibm_status ibm_fft_destroy(ibm_fft* p)
{
    if (!p) return IBM_OK;
    hipDeviceSynchronize();
    for (void* q : p->owned) hipFree(q);
    delete p;
    return IBM_OK;
}

ibm_status ibm_fft_solve(ibm_ctx* ctx, ibm_fft* p, double sigma, double tau, const double* b, double* x)
{
    return ibm::fft_solve(ctx, p, sigma, tau, b, x);
}

} // extern "C"
