// Matrix-free staggered-grid stencils, multigrid transfers, vector kernels
// and deterministic reductions (FP64 and FP32).
//
// Replaces the CSR sparse matrix-vector products the reference issues through
// cusparseDcsrmv for every operator application (linalg/blas.h:186-199): the
// operators are evaluated from their Kronecker structure (fd_ops.hpp), so a
// stencil pass moves only the vectors (16 B per point in FP64, 8 in FP32,
// +8/+4 per extra input vector) instead of ~100 B of CSR data per row.
#include <hip/hip_runtime.h>
#include <map>
#include <mutex>
#include <algorithm>
#include <type_traits>
#include <utility>
#include <cmath>
#include <vector>
#include "context.hpp"
#include "fd_ops.hpp"
#include "fd_kernels.hpp"

namespace ibm {

// ---------------------------------------------------------------------------
// Stencil: out = alpha * A x + beta * y + gamma * z   (y, z optional)
// Thread (i, j) marches a chunk of KZ planes in k with the k-neighbours in
// registers; i is the fastest index so rows are read coalesced.
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline T lap1(const op_axis<T>& a, int i, T xm, T x0, T xp)
{
    if (a.periodic) return a.r2 * (xm + xp - 2 * x0);
    T d = -2;
    if (i == 0) d += a.lap_lo;
    if (i == a.n - 1) d += a.lap_hi;
    T l = i > 0 ? xm : T(0);
    T r = i < a.n - 1 ? xp : T(0);
    return a.r2 * (l + r + d * x0);
}

template <typename T>
__device__ inline T id1(const op_axis<T>& a, int i)
{
    T d = 1;
    if (i == 0) d += a.id_lo;
    if (i == a.n - 1) d += a.id_hi;
    return d;
}

constexpr int STX = 64, STY = 4, KZ = 8;
// Planes per thread of the fused momentum and residual kernels: they hold
// three or four more vectors per plane than stencil_kernel, so fewer planes
// keep their registers (and so the waves per SIMD) in line with it.
constexpr int KZF = 4;

// One point of out = alpha * A x + beta * y + gamma * z (the expression of
// stencil_kernel, reading the neighbours directly).
template <typename T>
__device__ inline T stencil_at(const op3<T>& A, T alpha, const T* __restrict__ x, T beta, const T* __restrict__ y,
                               T gamma, const T* __restrict__ z, int i, int j, int k)
{
    const int nx = A.a[0].n, ny = A.a[1].n, nz = A.a[2].n;
    const size_t sxy = (size_t) nx * ny;
    const int im = A.a[0].periodic ? (i == 0 ? nx - 1 : i - 1) : (i > 0 ? i - 1 : i);
    const int ip = A.a[0].periodic ? (i == nx - 1 ? 0 : i + 1) : (i < nx - 1 ? i + 1 : i);
    const int jm = A.a[1].periodic ? (j == 0 ? ny - 1 : j - 1) : (j > 0 ? j - 1 : j);
    const int jp = A.a[1].periodic ? (j == ny - 1 ? 0 : j + 1) : (j < ny - 1 ? j + 1 : j);
    const int km = A.a[2].periodic ? (k == 0 ? nz - 1 : k - 1) : (k > 0 ? k - 1 : k);
    const int kp = A.a[2].periodic ? (k == nz - 1 ? 0 : k + 1) : (k < nz - 1 ? k + 1 : k);
    const size_t row = (size_t) nx * j + sxy * (size_t) k;
    const size_t p = (size_t) i + row;
    const T x0 = x[p];
    const T lx = lap1(A.a[0], i, x[row + im], x0, x[row + ip]);
    const T ly = lap1(A.a[1], j, x[(size_t) i + (size_t) nx * jm + sxy * k], x0, x[(size_t) i + (size_t) nx * jp + sxy * k]);
    const T lz = lap1(A.a[2], k, x[(size_t) i + (size_t) nx * j + sxy * km], x0, x[(size_t) i + (size_t) nx * j + sxy * kp]);
    const T Ix = id1(A.a[0], i), Iy = id1(A.a[1], j), Iz = id1(A.a[2], k);
    const T ax = A.sigma * (Ix * Iy * Iz) * x0 + A.tau * (Iy * Iz * lx + Ix * Iz * ly + Ix * Iy * lz);
    T v = alpha * ax;
    if (y) v += beta * y[p];
    if (z) v += gamma * z[p];
    return v;
}

// Thread (i, j) computes KZ consecutive planes.  Every load of the KZ planes
// (the centre column with one plane of halo on each side, the x and y
// neighbours, y and z) is issued before any arithmetic, so a wavefront has
// ~7 KZ independent loads in flight; out-of-range planes of the last chunk
// load clamped (in-bounds) addresses and are not stored.
// OUT selects what is done with each value v: 0 stores out = v; 1 also
// stores a float copy (the FP32 right-hand side of the next V-cycle); 2 adds
// it to the FP64 vector acc instead of storing it (the last smoothing pass of
// a V-cycle folded into the defect correction x += e); 3 stores the float
// copy alone (a residual that nothing reads in FP64).
template <typename T, bool HY, bool HZ, int OUT>
__global__ void __launch_bounds__(STX * STY)
stencil_kernel(op3<T> A, T alpha, const T* __restrict__ x, T beta, const T* __restrict__ y,
               T gamma, const T* __restrict__ z, T* __restrict__ out, float* __restrict__ out32,
               double* __restrict__ acc)
{
    const int nx = A.a[0].n, ny = A.a[1].n, nz = A.a[2].n;
    const int i = blockIdx.x * STX + threadIdx.x;
    const int j = blockIdx.y * STY + threadIdx.y;
    const int k0 = blockIdx.z * KZ;
    if (i >= nx || j >= ny || k0 >= nz) return;
    const size_t sxy = (size_t) nx * ny;
    const int im = A.a[0].periodic ? (i == 0 ? nx - 1 : i - 1) : (i > 0 ? i - 1 : i);
    const int ip = A.a[0].periodic ? (i == nx - 1 ? 0 : i + 1) : (i < nx - 1 ? i + 1 : i);
    const int jm = A.a[1].periodic ? (j == 0 ? ny - 1 : j - 1) : (j > 0 ? j - 1 : j);
    const int jp = A.a[1].periodic ? (j == ny - 1 ? 0 : j + 1) : (j < ny - 1 ? j + 1 : j);
    const size_t col = (size_t) i + (size_t) nx * j;
    const ptrdiff_t dxm = (ptrdiff_t) im - i, dxp = (ptrdiff_t) ip - i;
    const ptrdiff_t dym = ((ptrdiff_t) jm - j) * nx, dyp = ((ptrdiff_t) jp - j) * nx;

    T c[KZ + 2], xl[KZ], xr[KZ], yl[KZ], yr[KZ], yv[KZ], zv[KZ];
    {
        int km = k0 - 1;
        if (km < 0) km = A.a[2].periodic ? nz - 1 : 0;
        c[0] = x[col + sxy * (size_t) km];
    }
#pragma unroll
    for (int q = 0; q <= KZ; ++q) {
        int k = k0 + q;
        if (k >= nz) k = A.a[2].periodic ? k % nz : nz - 1;      // wrap (or clamp; lap1 ignores it)
        c[q + 1] = x[col + sxy * (size_t) k];
    }
#pragma unroll
    for (int q = 0; q < KZ; ++q) {
        const int k = min(k0 + q, nz - 1);
        const size_t p = col + sxy * (size_t) k;
        xl[q] = x[p + dxm];
        xr[q] = x[p + dxp];
        yl[q] = x[p + dym];
        yr[q] = x[p + dyp];
        if (HY) yv[q] = y[p];
        if (HZ) zv[q] = z[p];
    }
    const T Ix = id1(A.a[0], i), Iy = id1(A.a[1], j);
#pragma unroll
    for (int q = 0; q < KZ; ++q) {
        const int k = k0 + q;
        if (k >= nz) break;
        const T x0 = c[q + 1];
        const T xm = c[q], xp = c[q + 2];
        const T lx = lap1(A.a[0], i, xl[q], x0, xr[q]);
        const T ly = lap1(A.a[1], j, yl[q], x0, yr[q]);
        const T lz = lap1(A.a[2], k, xm, x0, xp);
        const T Iz = id1(A.a[2], k);
        const T ax = A.sigma * (Ix * Iy * Iz) * x0 + A.tau * (Iy * Iz * lx + Ix * Iz * ly + Ix * Iy * lz);
        T v = alpha * ax;
        if (HY) v += beta * yv[q];
        if (HZ) v += gamma * zv[q];
        const size_t p = col + sxy * (size_t) k;
        if (OUT == 2) {
            acc[p] += (double) v;
        } else if (OUT == 3) {
            out32[p] = (float) v;
        } else {
            out[p] = v;
            if (OUT == 1) out32[p] = (float) v;
        }
    }
}

// The momentum right-hand side of one velocity component in one pass (the
// loads of stencil_kernel, with u as x):
//   rhs = dt I~ (F inv_rho - H) + dt nu (b + L u),
// I~ the identity weighted on the wall rows (sigma = 1, tau = 0 of A) and L
// the Laplacian (sigma = 0, tau = 1 of A); b is optional.
template <bool HB>
__global__ void __launch_bounds__(STX * STY)
momentum_rhs_kernel(op3<double> A, double dt, double dtnu, double inv_rho, const double* __restrict__ F,
                    const double* __restrict__ H, const double* __restrict__ u, const double* __restrict__ b,
                    double* __restrict__ rhs)
{
    const int nx = A.a[0].n, ny = A.a[1].n, nz = A.a[2].n;
    const int i = blockIdx.x * STX + threadIdx.x;
    const int j = blockIdx.y * STY + threadIdx.y;
    const int k0 = blockIdx.z * KZF;
    if (i >= nx || j >= ny || k0 >= nz) return;
    const size_t sxy = (size_t) nx * ny;
    const int im = A.a[0].periodic ? (i == 0 ? nx - 1 : i - 1) : (i > 0 ? i - 1 : i);
    const int ip = A.a[0].periodic ? (i == nx - 1 ? 0 : i + 1) : (i < nx - 1 ? i + 1 : i);
    const int jm = A.a[1].periodic ? (j == 0 ? ny - 1 : j - 1) : (j > 0 ? j - 1 : j);
    const int jp = A.a[1].periodic ? (j == ny - 1 ? 0 : j + 1) : (j < ny - 1 ? j + 1 : j);
    const size_t col = (size_t) i + (size_t) nx * j;
    const ptrdiff_t dxm = (ptrdiff_t) im - i, dxp = (ptrdiff_t) ip - i;
    const ptrdiff_t dym = ((ptrdiff_t) jm - j) * nx, dyp = ((ptrdiff_t) jp - j) * nx;

    double c[KZF + 2], xl[KZF], xr[KZF], yl[KZF], yr[KZF], fv[KZF], hv[KZF], bv[KZF];
    {
        int km = k0 - 1;
        if (km < 0) km = A.a[2].periodic ? nz - 1 : 0;
        c[0] = u[col + sxy * (size_t) km];
    }
#pragma unroll
    for (int q = 0; q <= KZF; ++q) {
        int k = k0 + q;
        if (k >= nz) k = A.a[2].periodic ? k % nz : nz - 1;
        c[q + 1] = u[col + sxy * (size_t) k];
    }
#pragma unroll
    for (int q = 0; q < KZF; ++q) {
        const int k = min(k0 + q, nz - 1);
        const size_t p = col + sxy * (size_t) k;
        xl[q] = u[p + dxm];
        xr[q] = u[p + dxp];
        yl[q] = u[p + dym];
        yr[q] = u[p + dyp];
        fv[q] = F[p];
        hv[q] = H[p];
        if (HB) bv[q] = b[p];
    }
    const double Ix = id1(A.a[0], i), Iy = id1(A.a[1], j);
#pragma unroll
    for (int q = 0; q < KZF; ++q) {
        const int k = k0 + q;
        if (k >= nz) break;
        const double x0 = c[q + 1];
        const double lx = lap1(A.a[0], i, xl[q], x0, xr[q]);
        const double ly = lap1(A.a[1], j, yl[q], x0, yr[q]);
        const double lz = lap1(A.a[2], k, c[q], x0, c[q + 2]);
        const double Iz = id1(A.a[2], k);
        const double h = fv[q] * inv_rho - hv[q];
        double v = dt * ((Ix * Iy * Iz) * h);
        if (HB) v += dtnu * bv[q];
        rhs[col + sxy * (size_t) k] = dtnu * (Iy * Iz * lx + Ix * Iz * ly + Ix * Iy * lz) + v;
    }
}

// Sums of two values over a (STX x STY)-thread block, each in a fixed order
// (wavefront shuffles, then the wavefronts in index order); thread 0 gets them.
__device__ inline void block_sum2(double& a, double& b)
{
    __shared__ double part[2][STY];
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
    }
    const int t = threadIdx.x + STX * threadIdx.y;
    if ((t & 63) == 0) {
        part[0][t >> 6] = a;
        part[1][t >> 6] = b;
    }
    __syncthreads();
    if (t == 0) {
        a = 0;
        b = 0;
        for (int w = 0; w < STX * STY / 64; ++w) {
            a += part[0][w];
            b += part[1][w];
        }
    }
}

// r = b - A x at every point (stencil_kernel's expression of A x); the sums
// of r^2 and b^2 over each block's points go to partial[block] and
// partial[nblocks + block], and with UPD u += x (a direct solve's update and
// its FP64 residual check in one pass; r itself is not stored).
template <bool UPD>
__global__ void __launch_bounds__(STX * STY)
residual_update_kernel(op3<double> A, const double* __restrict__ x, const double* __restrict__ b,
                       double* __restrict__ u, double* __restrict__ partial)
{
    const int nx = A.a[0].n, ny = A.a[1].n, nz = A.a[2].n;
    const int i = blockIdx.x * STX + threadIdx.x;
    const int j = blockIdx.y * STY + threadIdx.y;
    const int k0 = blockIdx.z * KZF;
    double rr = 0, bb = 0;
    if (i < nx && j < ny && k0 < nz) {
        const size_t sxy = (size_t) nx * ny;
        const int im = A.a[0].periodic ? (i == 0 ? nx - 1 : i - 1) : (i > 0 ? i - 1 : i);
        const int ip = A.a[0].periodic ? (i == nx - 1 ? 0 : i + 1) : (i < nx - 1 ? i + 1 : i);
        const int jm = A.a[1].periodic ? (j == 0 ? ny - 1 : j - 1) : (j > 0 ? j - 1 : j);
        const int jp = A.a[1].periodic ? (j == ny - 1 ? 0 : j + 1) : (j < ny - 1 ? j + 1 : j);
        const size_t col = (size_t) i + (size_t) nx * j;
        const ptrdiff_t dxm = (ptrdiff_t) im - i, dxp = (ptrdiff_t) ip - i;
        const ptrdiff_t dym = ((ptrdiff_t) jm - j) * nx, dyp = ((ptrdiff_t) jp - j) * nx;
        double c[KZF + 2], xl[KZF], xr[KZF], yl[KZF], yr[KZF], bv[KZF], uv[KZF];
        {
            int km = k0 - 1;
            if (km < 0) km = A.a[2].periodic ? nz - 1 : 0;
            c[0] = x[col + sxy * (size_t) km];
        }
#pragma unroll
        for (int q = 0; q <= KZF; ++q) {
            int k = k0 + q;
            if (k >= nz) k = A.a[2].periodic ? k % nz : nz - 1;
            c[q + 1] = x[col + sxy * (size_t) k];
        }
#pragma unroll
        for (int q = 0; q < KZF; ++q) {
            const int k = min(k0 + q, nz - 1);
            const size_t p = col + sxy * (size_t) k;
            xl[q] = x[p + dxm];
            xr[q] = x[p + dxp];
            yl[q] = x[p + dym];
            yr[q] = x[p + dyp];
            bv[q] = b[p];
            if (UPD) uv[q] = u[p];
        }
        const double Ix = id1(A.a[0], i), Iy = id1(A.a[1], j);
#pragma unroll
        for (int q = 0; q < KZF; ++q) {
            const int k = k0 + q;
            if (k >= nz) break;
            const double x0 = c[q + 1];
            const double lx = lap1(A.a[0], i, xl[q], x0, xr[q]);
            const double ly = lap1(A.a[1], j, yl[q], x0, yr[q]);
            const double lz = lap1(A.a[2], k, c[q], x0, c[q + 2]);
            const double Iz = id1(A.a[2], k);
            const double ax = A.sigma * (Ix * Iy * Iz) * x0 + A.tau * (Iy * Iz * lx + Ix * Iz * ly + Ix * Iy * lz);
            const double r = bv[q] - ax;
            rr += r * r;
            bb += bv[q] * bv[q];
            if (UPD) u[col + sxy * (size_t) k] = x0 + uv[q];
        }
    }
    block_sum2(rr, bb);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const int nb = gridDim.x * gridDim.y * gridDim.z;
        const int blk = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
        partial[blk] = rr;
        partial[nb + blk] = bb;
    }
}

// ---------------------------------------------------------------------------
// Degree-4 Chebyshev smoothing in one pass:
//   s1 = w1 v - w0 A v,  s2 = w2 v - A s1,  out = w3 v - A s2 (+ out)
// A block owns a TX x TY column tile of the output and a chunk of z planes
// [z0, z1), and marches through z with rings of x-y planes in LDS: v on the
// tile grown by 3 points per side, s1 (grown by 2) and s2 (grown by 1), all
// stored in the grown-by-3 plane layout.  Phase q computes s1 on plane
// z0-5+q, s2 on plane z0-7+q and the output on plane z0-9+q (whose out / acc
// operands were loaded two phases earlier), then stores v plane z0-3+q
// (loaded two phases earlier) and starts loading plane z0-1+q; every
// stage reads only planes written in earlier phases, so one barrier per
// phase suffices.  A thread's x-y positions (and their wall coefficients)
// are the same on every plane.  Same per-point arithmetic as
// stencil_kernel; points outside a walled axis are not computed (lap1 never
// reads them).  With ACC the result is added to the FP64 vector acc instead
// of being stored.
// The axes' periodicity is a template parameter (bit d of PER: axis d
// periodic), so no branch on it is left in the kernel.  With it as a run-time
// flag the compiler built the branch masks once, inside the first of several
// loops over positions, and reused them in the later loops: lanes that had
// left the first loop early held 0 there, and waves whose active lanes were
// all such lanes took the periodic branch on a walled axis.
// ---------------------------------------------------------------------------
// Tile configurations: WIDE (64 x 8 FP32 / 32 x 8 FP64 columns, 512 threads)
// for large grids; narrow (16 x 8 columns, 256 threads) for the coarser
// grids, whose launches are bound by the phases' latency, not by throughput:
// fewer points per thread and phase, and no wrapped copies of narrow x axes.
template <typename T, bool WIDE>
struct march_tile {
    static constexpr int X = WIDE ? (sizeof(T) == 4 ? 64 : 32) : 16;
    static constexpr int Y = 8;
    static constexpr int THREADS = WIDE ? 512 : 256;
    static constexpr int MIN_Z = WIDE ? 4 : 2;   // shortest z chunk
};
constexpr size_t MARCH_WIDE_POINTS = (size_t) 1 << 19;   // grids from 512K points use the wide tile
// Ring sizes are powers of two and the phase loop is unrolled by MARCH_NV, so
// every ring slot a phase touches is a compile-time constant: the per-phase
// uniform work (issued by the scalar unit, one instruction per SIMD every
// four cycles, like a wave64 VALU op) is then a few compares and adds.
constexpr int MARCH_NV = 8;      // v planes in the ring (7 alive in a phase)
constexpr int MARCH_NS = 4;      // s1 (and s2) planes alive in a phase
constexpr int MARCH_LEAD = 9;    // phases between loading a v plane and storing its output
constexpr int MARCH_TARGET_BLOCKS = MARCH_DEFAULT_TARGET;
constexpr int MARCH_PF = 2;      // phases between loading a plane's operands and using them

namespace {
struct shape_key {
    int kind, elem, nx, ny, nz, per;
    bool operator<(const shape_key& o) const
    {
        const int a[6] = {kind, elem, nx, ny, nz, per}, b[6] = {o.kind, o.elem, o.nx, o.ny, o.nz, o.per};
        for (int i = 0; i < 6; ++i)
            if (a[i] != b[i]) return a[i] < b[i];
        return false;
    }
};
std::mutex tune_lock;
std::map<shape_key, int> tune_table;
} // namespace

int tuned_target(int kind, int elem, int nx, int ny, int nz, int per)
{
    std::lock_guard<std::mutex> guard(tune_lock);
    auto it = tune_table.find(shape_key{kind, elem, nx, ny, nz, per});
    return it == tune_table.end() ? 0 : it->second;
}

void set_tuned_target(int kind, int elem, int nx, int ny, int nz, int per, int target)
{
    std::lock_guard<std::mutex> guard(tune_lock);
    tune_table[shape_key{kind, elem, nx, ny, nz, per}] = target;
}

void clear_tuned_targets()
{
    std::lock_guard<std::mutex> guard(tune_lock);
    tune_table.clear();
}

bool target_tuned(int kind, int elem, int nx, int ny, int nz, int per)
{
    std::lock_guard<std::mutex> guard(tune_lock);
    return tune_table.count(shape_key{kind, elem, nx, ny, nz, per}) != 0;
}

bool march_wide(size_t points) { return points >= MARCH_WIDE_POINTS; }

static int per_mask(const op3<float>& A) { return (A.a[0].periodic ? 1 : 0) | (A.a[1].periodic ? 2 : 0) | (A.a[2].periodic ? 4 : 0); }
static int per_mask(const op3<double>& A) { return (A.a[0].periodic ? 1 : 0) | (A.a[1].periodic ? 2 : 0) | (A.a[2].periodic ? 4 : 0); }
static_assert(MARCH_NV % MARCH_NS == 0 && MARCH_NV % MARCH_PF == 0, "ring periods must divide the unroll");

// Global index of coordinate o - h + c along an axis: wrapped on a periodic
// axis, -1 outside a walled one.
template <bool P, typename T>
__device__ inline int axis_global(const op_axis<T>& a, int o, int h, int c)
{
    int g = o - h + c;
    if constexpr (P) {
        // One period either side without a division (the usual case).
        if (g < 0) g += a.n;
        else if (g >= a.n) g -= a.n;
        if ((unsigned) g >= (unsigned) a.n) {
            g %= a.n;
            if (g < 0) g += a.n;
        }
        return g;
    } else {
        return (g < 0 || g >= a.n) ? -1 : g;
    }
}

// Coefficients of lap1 along one axis at index i: L1 x = r2 (lo xm + hi xp +
// d x0), with lo / hi 0 where the neighbour lies outside a wall.
template <bool P, typename T>
__device__ inline void axis_coef(const op_axis<T>& a, int i, T& d, T& lo, T& hi)
{
    d = -2;
    lo = 1;
    hi = 1;
    if constexpr (!P) {
        if (i == 0) d += a.lap_lo;
        if (i == a.n - 1) d += a.lap_hi;
        lo = i > 0 ? T(1) : T(0);
        hi = i < a.n - 1 ? T(1) : T(0);
    }
}

// A position of the x-y plane (the same on every plane of the march): its
// plane offset (-1: none, or outside a wall) and x / y coefficients.
template <typename T>
struct xy_coef {
    int c;
    T Ix, Iy, IxIy;
    T dx, xlo, xhi, dy, ylo, yhi;
    // The operator on a plane away from z walls (Iz = 1, lz = r2z (zm + zp -
    // 2 x0)), collapsed to seven coefficients:
    // A x = a0 x0 + axm xm + axp xp + aym ym + ayp yp + az (zm + zp).
    T a0, axm, axp, aym, ayp, az;
};

template <int PER, typename T>
__device__ inline xy_coef<T> make_xy_coef(const op3<T>& A, bool valid, int c, int gi, int gj)
{
    xy_coef<T> k;
    k.c = valid ? c : -1;
    gi = valid ? gi : 0;
    gj = valid ? gj : 0;
    k.Ix = id1(A.a[0], gi);
    k.Iy = id1(A.a[1], gj);
    k.IxIy = k.Ix * k.Iy;
    axis_coef<(PER & 1) != 0>(A.a[0], gi, k.dx, k.xlo, k.xhi);
    axis_coef<(PER & 2) != 0>(A.a[1], gj, k.dy, k.ylo, k.yhi);
    const T tx = A.tau * k.Iy * A.a[0].r2, ty = A.tau * k.Ix * A.a[1].r2, tz = A.tau * k.IxIy * A.a[2].r2;
    k.a0 = A.sigma * k.IxIy + tx * k.dx + ty * k.dy - 2 * tz;
    k.axm = tx * k.xlo;
    k.axp = tx * k.xhi;
    k.aym = ty * k.ylo;
    k.ayp = ty * k.yhi;
    k.az = tz;
    return k;
}

// z coefficients of one plane.
template <typename T>
struct z_coef {
    T Iz, dz, zlo, zhi;
};

template <bool P, typename T>
__device__ inline z_coef<T> make_z_coef(const op_axis<T>& a, int k)
{
    z_coef<T> z;
    z.Iz = id1(a, k);
    axis_coef<P>(a, k, z.dz, z.zlo, z.zhi);
    return z;
}

// lap1 from coefficients; periodic axes keep lap1's own expression.
template <bool P, typename T>
__device__ inline T lap_coef(T r2, T lo, T hi, T d, T xm, T x0, T xp)
{
    if constexpr (P) return r2 * (xm + xp - 2 * x0);
    else return r2 * (lo * xm + hi * xp + d * x0);
}

// sigma I~ + tau L at position k of the centre plane Sc (row stride SY),
// Sm / Sp the planes below / above (the expression of stencil_at).
template <int PER, int SY, typename T>
__device__ inline T plane_op(const op3<T>& A, const T* Sm, const T* Sc, const T* Sp, const xy_coef<T>& k,
                             const z_coef<T>& kz)
{
    const int c = k.c;
    const T x0 = Sc[c];
    const T lx = lap_coef<(PER & 1) != 0>(A.a[0].r2, k.xlo, k.xhi, k.dx, Sc[c - 1], x0, Sc[c + 1]);
    const T ly = lap_coef<(PER & 2) != 0>(A.a[1].r2, k.ylo, k.yhi, k.dy, Sc[c - SY], x0, Sc[c + SY]);
    const T lz = lap_coef<(PER & 4) != 0>(A.a[2].r2, kz.zlo, kz.zhi, kz.dz, Sm[c], x0, Sp[c]);
    return A.sigma * (k.IxIy * kz.Iz) * x0 + A.tau * (k.Iy * kz.Iz * lx + k.Ix * kz.Iz * ly + k.IxIy * lz);
}

// The same on a plane away from z walls, from the collapsed coefficients.
template <int SY, typename T>
__device__ inline T plane_op_int(const T* Sm, const T* Sc, const T* Sp, const xy_coef<T>& k)
{
    const int c = k.c;
    return k.a0 * Sc[c] + k.axm * Sc[c - 1] + k.axp * Sc[c + 1] + k.aym * Sc[c - SY] + k.ayp * Sc[c + SY] +
           k.az * (Sm[c] + Sp[c]);
}

// The seven operands of one plane op, loaded before any arithmetic.
template <typename T>
struct nbr7 {
    T x0, xm, xp, ym, yp, zm, zp;
};

template <int SY, typename T>
__device__ inline nbr7<T> load7(const T* Sm, const T* Sc, const T* Sp, int c)
{
    return {Sc[c], Sc[c - 1], Sc[c + 1], Sc[c - SY], Sc[c + SY], Sm[c], Sp[c]};
}

// plane_apply on loaded operands (the same expressions).
template <int PER, typename T>
__device__ inline T apply7(const op3<T>& A, bool interior, const nbr7<T>& n, const xy_coef<T>& k,
                           const z_coef<T>& kz)
{
    if ((PER & 4) != 0 || interior)
        return k.a0 * n.x0 + k.axm * n.xm + k.axp * n.xp + k.aym * n.ym + k.ayp * n.yp + k.az * (n.zm + n.zp);
    const T lx = lap_coef<(PER & 1) != 0>(A.a[0].r2, k.xlo, k.xhi, k.dx, n.xm, n.x0, n.xp);
    const T ly = lap_coef<(PER & 2) != 0>(A.a[1].r2, k.ylo, k.yhi, k.dy, n.ym, n.x0, n.yp);
    const T lz = lap_coef<(PER & 4) != 0>(A.a[2].r2, kz.zlo, kz.zhi, kz.dz, n.zm, n.x0, n.zp);
    return A.sigma * (k.IxIy * kz.Iz) * n.x0 + A.tau * (k.Iy * kz.Iz * lx + k.Ix * kz.Iz * ly + k.IxIy * lz);
}

// Plane op choosing the collapsed form where the plane is away from z walls
// (uniform per plane).
template <int PER, int SY, typename T>
__device__ inline T plane_apply(const op3<T>& A, bool interior, const T* Sm, const T* Sc, const T* Sp,
                                const xy_coef<T>& k, const z_coef<T>& kz)
{
    if ((PER & 4) != 0 || interior) return plane_op_int<SY>(Sm, Sc, Sp, k);
    return plane_op<PER, SY>(A, Sm, Sc, Sp, k, kz);
}

// Two blocks per CU (four waves per SIMD for the wide tile): at most 128 VGPRs.
template <typename T, bool ADD, bool ACC, int PER, bool WIDE>
__global__ void __launch_bounds__((march_tile<T, WIDE>::THREADS))
__attribute__((amdgpu_waves_per_eu(march_tile<T, WIDE>::THREADS / 128)))
cheb_march_kernel(op3<T> A, T w0, T w1, T w2, T w3, const T* __restrict__ v, T* out, const T* addsrc,
                  double* __restrict__ acc, int tiles_x, int tiles_xy, int zchunk, int nblocks, int xcd_chunk)
{
    constexpr bool PX = (PER & 1) != 0, PY = (PER & 2) != 0, PZ = (PER & 4) != 0;
    constexpr int TX = march_tile<T, WIDE>::X, TY = march_tile<T, WIDE>::Y;
    constexpr int MARCH_THREADS = march_tile<T, WIDE>::THREADS;
    constexpr int VX = TX + 6, VY = TY + 6, VP = VX * VY;
    constexpr int NL = (VP + MARCH_THREADS - 1) / MARCH_THREADS;
    constexpr int W1 = TX + 4, P1 = W1 * (TY + 4), N1 = (P1 + MARCH_THREADS - 1) / MARCH_THREADS;
    constexpr int W2 = TX + 2, P2 = W2 * (TY + 2), N2 = (P2 + MARCH_THREADS - 1) / MARCH_THREADS;
    constexpr int PO = TX * TY, NO = (PO + MARCH_THREADS - 1) / MARCH_THREADS;
    // s1 / s2 planes carry 64 slots past the grown plane (one per lane) that
    // positions without a value write to and nothing reads, so the stages run
    // without per-position branches (and so issue all their LDS reads before
    // one wait).  Such positions read at SAFE, whose neighbours lie in the plane.
    constexpr int VPS = VP + 64, SAFE = VX + 1;
    constexpr int NZ0 = (2 * MARCH_NS * VPS + MARCH_THREADS - 1) / MARCH_THREADS;
    __shared__ T Rv[MARCH_NV][VP];
    __shared__ T Rs[2][MARCH_NS][VPS];   // s1, s2
    T(*R1)[VPS] = Rs[0];
    T(*R2)[VPS] = Rs[1];

    int b = blockIdx.x;
    if (xcd_chunk) b = (b & 7) * xcd_chunk + (b >> 3);
    if (b >= nblocks) return;
    const int tile = b % tiles_xy, chunk = b / tiles_xy;
    const int ox = (tile % tiles_x) * TX, oy = (tile / tiles_x) * TY;
    const int nx = A.a[0].n, ny = A.a[1].n, nz = A.a[2].n;
    const int z0 = chunk * zchunk, z1 = min(z0 + zchunk, nz);
    const int tid = threadIdx.x;

    // s1 and s2 start at 0 where they are never computed (outside a wall):
    // the zero neighbour weights there must not meet uninitialised values.
#pragma unroll
    for (int r = 0; r < NZ0; ++r) {
        const int t = tid + r * MARCH_THREADS;
        if (t < 2 * MARCH_NS * VPS) (&Rs[0][0][0])[t] = T(0);
    }

    // Positions of this thread, the same on every plane.
    int lcol[NL];   // v loads: global column gi + nx gj (-1 outside a wall)
#pragma unroll
    for (int r = 0; r < NL; ++r) {
        const int t = tid + r * MARCH_THREADS;
        int col = -1;
        if (t < VP) {
            const int gi = axis_global<PX>(A.a[0], ox, 3, t % VX);
            const int gj = axis_global<PY>(A.a[1], oy, 3, t / VX);
            if (gi >= 0 && gj >= 0) col = gi + nx * gj;
        }
        lcol[r] = col;
    }
    xy_coef<T> k1[N1], k2[N2], ko[NO];
    int ws1[N1], ws2[N2];   // s1 / s2 write slots (past the plane where no value)
    bool ovalid[NO];
    int ocol[NO];   // output: global column gi + nx gj
#pragma unroll
    for (int r = 0; r < N1; ++r) {
        const int t = tid + r * MARCH_THREADS;
        const int cx = 1 + t % W1, cy = 1 + t / W1;
        const int gi = axis_global<PX>(A.a[0], ox, 3, cx), gj = axis_global<PY>(A.a[1], oy, 3, cy);
        k1[r] = make_xy_coef<PER>(A, t < P1 && gi >= 0 && gj >= 0, cx + VX * cy, gi, gj);
        ws1[r] = k1[r].c >= 0 ? k1[r].c : VP + (tid & 63);
        if (k1[r].c < 0) k1[r].c = SAFE;
    }
#pragma unroll
    for (int r = 0; r < N2; ++r) {
        const int t = tid + r * MARCH_THREADS;
        const int cx = 2 + t % W2, cy = 2 + t / W2;
        const int gi = axis_global<PX>(A.a[0], ox, 3, cx), gj = axis_global<PY>(A.a[1], oy, 3, cy);
        k2[r] = make_xy_coef<PER>(A, t < P2 && gi >= 0 && gj >= 0, cx + VX * cy, gi, gj);
        ws2[r] = k2[r].c >= 0 ? k2[r].c : VP + (tid & 63);
        if (k2[r].c < 0) k2[r].c = SAFE;
    }
#pragma unroll
    for (int r = 0; r < NO; ++r) {
        const int t = tid + r * MARCH_THREADS;
        const int tx = t % TX, ty = t / TX;
        const int gi = ox + tx, gj = oy + ty;
        const bool valid = t < PO && gi < nx && gj < ny;
        ko[r] = make_xy_coef<PER>(A, valid, (tx + 3) + VX * (ty + 3), gi, gj);
        ovalid[r] = valid;
        if (!valid) ko[r].c = SAFE;
        ocol[r] = valid ? gi + nx * gj : 0;
    }

    // Plane index along z: wrapped on a periodic axis, -1 outside a walled
    // one (uniform; no division for planes within a period of the grid).
    auto zglobal = [&](int z) -> int {
        if constexpr (PZ) {
            z = z < 0 ? z + nz : z;
            z = z >= nz ? z - nz : z;
            if ((unsigned) z >= (unsigned) nz) {   // z axes shorter than the march lead
                while (z < 0) z += nz;
                while (z >= nz) z -= nz;
            }
            return z;
        } else {
            return (z < 0 || z >= nz) ? -1 : z;
        }
    };
    const unsigned nxy = (unsigned) nx * (unsigned) ny;   // grids hold < 2^32 points

    // Memory operands are loaded MARCH_PF phases ahead of their use, into one
    // of MARCH_PF register sets.  Every load is issued unconditionally, from a
    // valid address (a point outside a wall or the chunk reads point 0, and is
    // masked where it is used), so that the waits before the uses count only
    // the loads issued before them.
    auto fetch = [&](int z, T (&pre)[NL]) {
        const unsigned k = (unsigned) max(zglobal(z), 0) * nxy;
#pragma unroll
        for (int r = 0; r < NL; ++r) pre[r] = v[(unsigned) max(lcol[r], 0) + k];
    };
    auto fetch_out = [&](int z, T (&po)[NO], double (&pa)[NO]) {
        const unsigned k = (unsigned) min(max(z, z0), z1 - 1) * nxy;
#pragma unroll
        for (int r = 0; r < NO; ++r) {
            const unsigned p = (unsigned) ocol[r] + k;
            if (ADD) po[r] = addsrc[p];
            if (ACC) pa[r] = acc[p];
        }
    };

    const int Z = z1 - z0;
    T pre[MARCH_PF][NL], po[MARCH_PF][NO];
    double pa[MARCH_PF][NO];

    // Phase q = (multiple of MARCH_NV) + P.  Ring slot of plane z: (z - z0 + 3)
    // mod the ring size; the v plane stored in phase q is z0 - 3 + q, so the
    // slots below are constants given P.
    auto phase = [&]<int P>(int q) {
        constexpr int F = P % MARCH_PF;
        if (q >= 3 && q <= Z + 6) {          // s1 = w1 v - w0 A v on plane z0-5+q
            const int z = z0 - 5 + q;
            const int gk = zglobal(z);
            if (gk >= 0) {
                const z_coef<T> kz = make_z_coef<PZ>(A.a[2], gk);
                const bool in = gk > 0 && gk < nz - 1;
                const T *Vm = Rv[(P - 3) & (MARCH_NV - 1)], *Vc = Rv[(P - 2) & (MARCH_NV - 1)],
                        *Vp = Rv[(P - 1) & (MARCH_NV - 1)];
                T* dst = R1[(P - 2) & (MARCH_NS - 1)];
                nbr7<T> nb[N1];
#pragma unroll
                for (int r = 0; r < N1; ++r) nb[r] = load7<VX>(Vm, Vc, Vp, k1[r].c);
                __builtin_amdgcn_sched_barrier(0);   // every row's reads in flight before the arithmetic
#pragma unroll
                for (int r = 0; r < N1; ++r)
                    dst[ws1[r]] = -w0 * apply7<PER>(A, in, nb[r], k1[r], kz) + w1 * nb[r].x0;
            }
        }
        if (q >= 6 && q <= Z + 7) {          // s2 = w2 v - A s1 on plane z0-7+q
            const int z = z0 - 7 + q;
            const int gk = zglobal(z);
            if (gk >= 0) {
                const z_coef<T> kz = make_z_coef<PZ>(A.a[2], gk);
                const bool in = gk > 0 && gk < nz - 1;
                const T *Sm = R1[(P - 5) & (MARCH_NS - 1)], *Sc = R1[(P - 4) & (MARCH_NS - 1)],
                        *Sp = R1[(P - 3) & (MARCH_NS - 1)];
                const T* Vc = Rv[(P - 4) & (MARCH_NV - 1)];
                T* dst = R2[(P - 4) & (MARCH_NS - 1)];
                nbr7<T> nb[N2];
                T vc[N2];
#pragma unroll
                for (int r = 0; r < N2; ++r) {
                    nb[r] = load7<VX>(Sm, Sc, Sp, k2[r].c);
                    vc[r] = Vc[k2[r].c];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int r = 0; r < N2; ++r)
                    dst[ws2[r]] = T(-1) * apply7<PER>(A, in, nb[r], k2[r], kz) + w2 * vc[r];
            }
        }
        if (q >= MARCH_LEAD) {               // out = w3 v - A s2 (+ out) on plane z0-9+q
            const int z = z0 - MARCH_LEAD + q;
            const z_coef<T> kz = make_z_coef<PZ>(A.a[2], z);
            const bool in = z > 0 && z < nz - 1;
            const T *Sm = R2[(P - 7) & (MARCH_NS - 1)], *Sc = R2[(P - 6) & (MARCH_NS - 1)],
                    *Sp = R2[(P - 5) & (MARCH_NS - 1)];
            const T* Vc = Rv[(P - 6) & (MARCH_NV - 1)];
            nbr7<T> nb[NO];
            T vc[NO];
#pragma unroll
            for (int r = 0; r < NO; ++r) {
                nb[r] = load7<VX>(Sm, Sc, Sp, ko[r].c);
                vc[r] = Vc[ko[r].c];
            }
            const unsigned kz0 = (unsigned) z * nxy;
#pragma unroll
            for (int r = 0; r < NO; ++r) {
                const unsigned p = (unsigned) ocol[r] + kz0;
                T val = T(-1) * apply7<PER>(A, in, nb[r], ko[r], kz);
                if (ADD) val += T(1) * po[F][r];
                val += w3 * vc[r];
                if (ovalid[r]) {
                    if (ACC) acc[p] = pa[F][r] + (double) val;
                    else out[p] = val;
                }
            }
        }
        // The loads are waited for at the end of the phase, after the stages.
        if (q <= Z + 5) {                    // v plane z0-3+q (0 outside a wall)
            const bool in = zglobal(z0 - 3 + q) >= 0;
            T* dst = Rv[P & (MARCH_NV - 1)];
#pragma unroll
            for (int r = 0; r < NL; ++r) {
                const int t = tid + r * MARCH_THREADS;
                if (t < VP) dst[t] = (in && lcol[r] >= 0) ? pre[F][r] : T(0);
            }
        }
        fetch(z0 - 3 + MARCH_PF + q, pre[F]);   // v plane stored MARCH_PF phases on
        if (ADD || ACC) fetch_out(z0 - MARCH_LEAD + MARCH_PF + q, po[F], pa[F]);   // operands of phase q + MARCH_PF
        __syncthreads();
    };
    auto group = [&]<int... Ps>(int q0, std::integer_sequence<int, Ps...>) {
        ((q0 + Ps < Z + MARCH_LEAD ? (phase.template operator()<Ps>(q0 + Ps), 0) : 0), ...);
    };

#pragma unroll
    for (int f = 0; f < MARCH_PF; ++f) fetch(z0 - 3 + f, pre[f]);
    const int nq = Z + MARCH_LEAD;
    for (int q0 = 0; q0 < nq; q0 += MARCH_NV) group(q0, std::make_integer_sequence<int, MARCH_NV>{});
}

// ---------------------------------------------------------------------------
// Transfers.  Along one axis the fine->coarse weights are w[t]/2 at fine
// index 2c + t - 1 and the coarse->fine weights w[j - 2c + 1]; an index one
// past either end wraps (periodic) or folds onto the end point times the fine
// grid's ghost coefficient (solid), as the reference's boundary entries do
// (solvers/mg/interpolation.h:79-80,116-118).
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline int fold(const transfer_axis<T>& a, int idx, int n, T& w)
{
    if (idx < 0) {
        if (a.periodic) return idx + n;
        w *= a.int_lo;
        return 0;
    }
    if (idx >= n) {
        if (a.periodic) return idx - n;
        w *= a.int_hi;
        return n - 1;
    }
    return idx;
}

template <typename T>
__device__ inline T restrict_at(const transfer3<T>& R, const T* __restrict__ fine, unsigned t)
{
    int c[3];
    const unsigned q = t / (unsigned) R.a[0].nc;
    c[0] = (int) (t - q * (unsigned) R.a[0].nc);
    c[1] = (int) (q % (unsigned) R.a[1].nc);
    c[2] = (int) (q / (unsigned) R.a[1].nc);
    int idx[3][4];
    T wt[3][4];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            T w = R.a[d].w[s] / 2;
            idx[d][s] = fold(R.a[d], 2 * c[d] + s - 1, R.a[d].nf, w);
            wt[d][s] = w;
        }
    const size_t nfx = R.a[0].nf, nfxy = nfx * R.a[1].nf;
    T acc = 0;
#pragma unroll
    for (int sz = 0; sz < 4; ++sz) {
        T az = 0;
#pragma unroll
        for (int sy = 0; sy < 4; ++sy) {
            const T* rowp = fine + nfxy * idx[2][sz] + nfx * idx[1][sy];
            T ax = wt[0][0] * rowp[idx[0][0]] + wt[0][1] * rowp[idx[0][1]] +
                   wt[0][2] * rowp[idx[0][2]] + wt[0][3] * rowp[idx[0][3]];
            az += wt[1][sy] * ax;
        }
        acc += wt[2][sz] * az;
    }
    return acc;
}

template <typename T>
__global__ void restrict_kernel(transfer3<T> R, const T* __restrict__ fine, T* __restrict__ coarse)
{
    const size_t total = (size_t) R.a[0].nc * R.a[1].nc * R.a[2].nc;
    size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    coarse[t] = restrict_at(R, fine, (unsigned) t);
}

// ---------------------------------------------------------------------------
// Residual and restriction in one pass: coarse = R (b - A x).  A block owns a
// CX x CY tile of coarse columns (fine columns 2 cx0 - 1 .. 2 (cx0 + CX)) and
// a chunk of coarse planes, and marches through the fine planes f it needs.
// Phase q computes the fine residual plane f = fs + q into an LDS ring (from
// a 4-plane ring of x with its x-y halo), then each thread with a coarse
// column restricts the previous residual plane over its 4 x 4 x-y taps and
// keeps the last four such partial sums in registers; a coarse plane k is
// emitted once its last fine plane is in (every second fine plane).  Taps
// fold at walls as restrict_at does (index -1 -> 0 times int_lo, n -> n - 1
// times int_hi) and wrap on periodic axes; sums run in restrict_at's order.
// ---------------------------------------------------------------------------
template <typename T, bool WIDE>
struct rr_tile {
    static constexpr int CX = WIDE ? 32 : 8, CY = 4;
    static constexpr int THREADS = WIDE ? 256 : 128;
};
constexpr int RR_TARGET_BLOCKS = TRANSFER_DEFAULT_TARGET;

// Fine tap of coarse index c along an axis: region offset from fine
// coordinate o (unwrapped on periodic axes, folded on walled ones) and weight.
template <typename T>
__device__ inline void rr_tap(const transfer_axis<T>& a, int c, int s, int o, int& off, T& w)
{
    w = a.w[s] / 2;
    int f = 2 * c + s - 1;
    if (!a.periodic) f = fold(a, f, a.nf, w);
    off = f - o;
}

template <typename T, int PER, bool WIDE>
__global__ void __launch_bounds__((rr_tile<T, WIDE>::THREADS))
residual_restrict_kernel(op3<T> A, transfer3<T> R, const T* __restrict__ x, const T* __restrict__ b,
                         T* __restrict__ coarse, int tiles_x, int tiles_xy, int kchunk, int nblocks, int xcd_chunk)
{
    constexpr bool PX = (PER & 1) != 0, PY = (PER & 2) != 0, PZ = (PER & 4) != 0;
    constexpr int CX = rr_tile<T, WIDE>::CX, CY = rr_tile<T, WIDE>::CY, NT = rr_tile<T, WIDE>::THREADS;
    constexpr int XX = 2 * CX + 4, XY = 2 * CY + 4, XP = XX * XY;      // x plane: fine tile grown by 2
    constexpr int RX = 2 * CX + 2, RP = RX * (2 * CY + 2);              // residual plane: grown by 1
    constexpr int NL = (XP + NT - 1) / NT, NR = (RP + NT - 1) / NT;
    static_assert(CX * CY <= NT, "one coarse column per thread");
    // Residual planes carry 64 slots past the plane (one per lane) that
    // positions without a value write to and nothing reads; those positions
    // read at the in-plane position XX + 1, so the rows run unbranched.
    constexpr int RPS = XP + 64;
    __shared__ T Xr[4][XP];
    __shared__ T Rr[2][RPS];

    int blk = blockIdx.x;
    if (xcd_chunk) blk = (blk & 7) * xcd_chunk + (blk >> 3);
    if (blk >= nblocks) return;
    const int tile = blk % tiles_xy, chunk = blk / tiles_xy;
    const int cx0 = (tile % tiles_x) * CX, cy0 = (tile / tiles_x) * CY;
    const int nx = A.a[0].n, ny = A.a[1].n, nz = A.a[2].n;
    const unsigned nxy = (unsigned) nx * (unsigned) ny;   // grids hold < 2^32 points
    const int ncx = R.a[0].nc, ncy = R.a[1].nc, ncz = R.a[2].nc;
    const int k0 = chunk * kchunk, k1 = min(k0 + kchunk, ncz);
    const int tid = threadIdx.x;
    const int fx0 = 2 * cx0 - 2, fy0 = 2 * cy0 - 2;    // fine origin of the x plane region

    // x loads
    int lcol[NL];
#pragma unroll
    for (int r = 0; r < NL; ++r) {
        const int t = tid + r * NT;
        int col = -1;
        if (t < XP) {
            const int gi = axis_global<PX>(A.a[0], fx0, 0, t % XX), gj = axis_global<PY>(A.a[1], fy0, 0, t / XX);
            if (gi >= 0 && gj >= 0) col = gi + nx * gj;
        }
        lcol[r] = col;
    }
    // residual positions (x-plane layout, offset by the halo of 1)
    xy_coef<T> kr[NR];
    int rcol[NR], rws[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int t = tid + r * NT;
        const int i = 1 + t % RX, j = 1 + t / RX;
        const int gi = axis_global<PX>(A.a[0], fx0, 0, i), gj = axis_global<PY>(A.a[1], fy0, 0, j);
        const bool valid = t < RP && gi >= 0 && gj >= 0;
        kr[r] = make_xy_coef<PER>(A, valid, i + XX * j, gi, gj);
        rws[r] = valid ? kr[r].c : XP + (tid & 63);
        if (!valid) kr[r].c = XX + 1;
        rcol[r] = valid ? gi + nx * gj : 0;
    }
    // this thread's coarse column and its x-y taps
    const int ci = cx0 + tid % CX, cj = cy0 + tid / CX;
    const bool has_c = tid < CX * CY && ci < ncx && cj < ncy;
    int tx[4], ty[4];
    T wx[4], wy[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        rr_tap(R.a[0], has_c ? ci : cx0, s, fx0, tx[s], wx[s]);
        rr_tap(R.a[1], has_c ? cj : cy0, s, fy0, ty[s], wy[s]);
    }

    // fine planes of the chunk: fs .. fe (unwrapped on a periodic z axis)
    const int nfz = R.a[2].nf;
    const int fs = PZ ? 2 * k0 - 1 : max(2 * k0 - 1, 0);
    const int fe = PZ ? 2 * k1 : min(2 * k1, nfz - 1);
    auto emit_plane = [&](int k) { return PZ ? 2 * k + 2 : min(2 * k + 2, nfz - 1); };
    auto sx = [&](int f) { return (f - fs + 1) & 3; };     // x ring slot of plane f >= fs - 1

    auto fetch_x = [&](int f, T (&pre)[NL]) {
        const unsigned k = (unsigned) max(axis_global<PZ>(A.a[2], f, 0, 0), 0) * nxy;
#pragma unroll
        for (int r = 0; r < NL; ++r) pre[r] = x[(unsigned) max(lcol[r], 0) + k];
    };
    auto store_x_slot = [&](int f, T* dst, const T (&pre)[NL]) {
        const bool in = axis_global<PZ>(A.a[2], f, 0, 0) >= 0;
#pragma unroll
        for (int r = 0; r < NL; ++r) {
            const int t = tid + r * NT;
            if (t < XP) dst[t] = (in && lcol[r] >= 0) ? pre[r] : T(0);
        }
    };
    auto store_x = [&](int f, const T (&pre)[NL]) { store_x_slot(f, Xr[sx(f)], pre); };
    auto fetch_b = [&](int f, T (&pb)[NR]) {
        const unsigned k = (unsigned) max(axis_global<PZ>(A.a[2], f, 0, 0), 0) * nxy;
#pragma unroll
        for (int r = 0; r < NR; ++r) pb[r] = b[(unsigned) rcol[r] + k];
    };

    // x planes fs - 1 .. fs + 1 before the march
    {
        T p0[NL], p1[NL], p2[NL];
        fetch_x(fs - 1, p0);
        fetch_x(fs, p1);
        fetch_x(fs + 1, p2);
        store_x(fs - 1, p0);
        store_x(fs, p1);
        store_x(fs + 1, p2);
    }
    T P0 = 0, P1 = 0, P2 = 0, P3 = 0;      // partial sums of fine planes f-3 .. f
    int next_k = k0;
    const unsigned ncxy = (unsigned) ncx * (unsigned) ncy;
    const unsigned ccol = has_c ? (unsigned) ci + (unsigned) ncx * cj : 0;

    // Phase q = (even) + S: plane f = fs + q sits in x slot (q + 1) & 3 and in
    // residual slot S (a constant; the loop is unrolled by two).
    T px[2][NL], pb[2][NR];
    auto phase = [&]<int S>(int q) {
        constexpr int F = S & 1;
        const int fp = fs + q - 1;           // residual plane of the previous phase
        if (q >= 1) {
            const T* Rc = Rr[(S - 1) & 1];
            T acc = 0;
#pragma unroll
            for (int sy = 0; sy < 4; ++sy) {
                const T* row = Rc + XX * ty[sy];
                const T ax = wx[0] * row[tx[0]] + wx[1] * row[tx[1]] + wx[2] * row[tx[2]] + wx[3] * row[tx[3]];
                acc += wy[sy] * ax;
            }
            P0 = P1;
            P1 = P2;
            P2 = P3;
            P3 = acc;
            // coarse planes whose last fine plane is fp
            while (next_k < k1 && emit_plane(next_k) == fp) {
                const int k = next_k++;
                T out = 0;
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    T wz = R.a[2].w[s] / 2;
                    int f = 2 * k + s - 1;
                    if (!PZ) f = fold(R.a[2], f, nfz, wz);
                    const int d = fp - f;                 // 0 .. 3
                    const T pv = d == 0 ? P3 : d == 1 ? P2 : d == 2 ? P1 : P0;
                    out += wz * pv;
                }
                if (has_c) coarse[ccol + ncxy * (unsigned) k] = out;
            }
        }
        const int f = fs + q;                // residual plane of this phase
        if (f <= fe) {
            const int gk = axis_global<PZ>(A.a[2], f, 0, 0);
            const z_coef<T> kz = make_z_coef<PZ>(A.a[2], gk);
            const bool in = gk > 0 && gk < nz - 1;
            const T *Xm = Xr[q & 3], *Xc = Xr[(q + 1) & 3], *Xp = Xr[(q + 2) & 3];
            T* dst = Rr[F];
#pragma unroll
            for (int r = 0; r < NR; ++r)
                dst[rws[r]] = T(-1) * plane_apply<PER, XX>(A, in, Xm, Xc, Xp, kr[r], kz) + T(1) * pb[F][r];
        }
        store_x_slot(f + 2, Xr[(q + 3) & 3], px[F]);   // fetched two phases ago
        fetch_x(f + 4, px[F]);
        fetch_b(f + 2, pb[F]);
        __syncthreads();
    };

    fetch_x(fs + 2, px[0]);
    fetch_x(fs + 3, px[1]);
    fetch_b(fs, pb[0]);
    fetch_b(fs + 1, pb[1]);
    __syncthreads();
    const int nq = fe - fs + 2;
    auto group = [&]<int... Ss>(int q0, std::integer_sequence<int, Ss...>) {
        ((q0 + Ss < nq ? (phase.template operator()<Ss>(q0 + Ss), 0) : 0), ...);
    };
    for (int q0 = 0; q0 < nq; q0 += 2) group(q0, std::make_integer_sequence<int, 2>{});
}

// The coarse-to-fine interpolation of one fine point.
template <typename T>
__device__ inline T prolong_at(const transfer3<T>& P, const T* __restrict__ coarse, unsigned t)
{
    int f[3];
    const unsigned q = t / (unsigned) P.a[0].nf;
    f[0] = (int) (t - q * (unsigned) P.a[0].nf);
    f[1] = (int) (q % (unsigned) P.a[1].nf);
    f[2] = (int) (q / (unsigned) P.a[1].nf);
    int idx[3][2];
    T wt[3][2];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        // coarse cells c with t = f - 2c + 1 in [0, 3]
        int j = f[d];
        int ca = (j + 1) >> 1;          // t = j - 2ca + 1 in {0, 1}... its partner is ca - 1
        int cb = ca - 1;
        T wa = P.a[d].w[j - 2 * ca + 1];
        T wb = P.a[d].w[j - 2 * cb + 1];
        idx[d][0] = fold(P.a[d], ca, P.a[d].nc, wa);
        idx[d][1] = fold(P.a[d], cb, P.a[d].nc, wb);
        wt[d][0] = wa;
        wt[d][1] = wb;
    }
    const size_t ncx = P.a[0].nc, ncxy = ncx * P.a[1].nc;
    T acc = 0;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        T az = 0;
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const T* rowp = coarse + ncxy * idx[2][b] + ncx * idx[1][a];
            az += wt[1][a] * (wt[0][0] * rowp[idx[0][0]] + wt[0][1] * rowp[idx[0][1]]);
        }
        acc += wt[2][b] * az;
    }
    return acc;
}

// ---------------------------------------------------------------------------
// Prolongation and residual in one pass: xn = x + P e, r = b - A xn, with xn
// written to its own vector (neighbouring blocks still read x in their
// halos).  A block owns a TX x TY fine column tile and a chunk of fine planes
// and marches through z: phase q forms xn on plane z0-1+q over the tile grown
// by 1 (from x and an LDS ring of e over the matching coarse columns, three
// planes alive in four slots; prolong_at's taps, folds and summation order)
// into a 4-plane ring, and
// computes the residual on plane z0-3+q from the planes of earlier phases.  Points outside walls
// hold 0 in the rings (the collapsed operator weights them by 0).
// ---------------------------------------------------------------------------
template <typename T, bool WIDE>
struct pr_tile {
    static constexpr int X = WIDE ? 64 : 16, Y = 8;
    static constexpr int THREADS = WIDE ? 256 : 128;
};
constexpr int PR_TARGET_BLOCKS = TRANSFER_DEFAULT_TARGET;

// Coarse taps of fine index j along an axis (prolong_at's ca, cb), as region
// offsets from coarse coordinate o (unwrapped on periodic axes, folded on
// walled ones) and weights.
template <typename T>
__device__ inline void pr_taps(const transfer_axis<T>& a, int j, int o, int off[2], T w[2])
{
    const int ca = (j + 1) >> 1, cb = ca - 1;
    T wa = a.w[j - 2 * ca + 1], wb = a.w[j - 2 * cb + 1];
    int ia = ca, ib = cb;
    if (!a.periodic) {
        ia = fold(a, ca, a.nc, wa);
        ib = fold(a, cb, a.nc, wb);
    }
    off[0] = ia - o;
    off[1] = ib - o;
    w[0] = wa;
    w[1] = wb;
}

template <typename T, int PER, bool WIDE>
__global__ void __launch_bounds__((pr_tile<T, WIDE>::THREADS))
prolong_residual_kernel(op3<T> A, transfer3<T> P, const T* __restrict__ e, const T* __restrict__ x,
                        const T* __restrict__ b, T* __restrict__ xn, T* __restrict__ r, int tiles_x, int tiles_xy,
                        int zchunk, int nblocks, int xcd_chunk)
{
    constexpr bool PX = (PER & 1) != 0, PY = (PER & 2) != 0, PZ = (PER & 4) != 0;
    constexpr int TX = pr_tile<T, WIDE>::X, TY = pr_tile<T, WIDE>::Y, NT = pr_tile<T, WIDE>::THREADS;
    constexpr int NXX = TX + 2, NXP = NXX * (TY + 2);            // xn plane: tile grown by 1
    constexpr int CXX = TX / 2 + 3, CP = CXX * (TY / 2 + 3);      // e plane: matching coarse columns
    constexpr int NN = (NXP + NT - 1) / NT, NR = (TX * TY + NT - 1) / NT, NC = (CP + NT - 1) / NT;
    __shared__ T Xn[4][NXP];          // xn planes f - 3 .. f of phase q (f = z0 - 1 + q)
    __shared__ T Ec[4][CP];           // three coarse planes alive; a power-of-two ring

    int blk = blockIdx.x;
    if (xcd_chunk) blk = (blk & 7) * xcd_chunk + (blk >> 3);
    if (blk >= nblocks) return;
    const int tile = blk % tiles_xy, chunk = blk / tiles_xy;
    const int ox = (tile % tiles_x) * TX, oy = (tile / tiles_x) * TY;
    const int nx = A.a[0].n, ny = A.a[1].n, nz = A.a[2].n;
    const unsigned nxy = (unsigned) nx * (unsigned) ny;   // grids hold < 2^32 points
    const int ncx = P.a[0].nc, ncy = P.a[1].nc, ncz = P.a[2].nc;
    const unsigned ncxy = (unsigned) ncx * (unsigned) ncy;
    const int z0 = chunk * zchunk, z1 = min(z0 + zchunk, nz);
    const int tid = threadIdx.x;
    const int cxo = (ox >> 1) - 1, cyo = (oy >> 1) - 1;    // coarse origin of the e region

    // xn positions: fine (ox - 1 + i, oy - 1 + j); their x / y coarse taps
    int ncol[NN], nc_[NN], ntx[NN][2], nty[NN][2];
    T nwx[NN][2], nwy[NN][2];
    bool nout[NN];     // inside the tile: written to xn
#pragma unroll
    for (int q = 0; q < NN; ++q) {
        const int t = tid + q * NT;
        const int i = t % NXX, j = t / NXX;
        const int gi = axis_global<PX>(A.a[0], ox, 1, i), gj = axis_global<PY>(A.a[1], oy, 1, j);
        const bool valid = t < NXP && gi >= 0 && gj >= 0;
        ncol[q] = valid ? gi + nx * gj : -1;
        nc_[q] = t < NXP ? t : -1;
        nout[q] = valid && i >= 1 && i <= TX && j >= 1 && j <= TY;
        // unwrapped fine index on periodic axes, the in-range index on walled ones
        pr_taps(P.a[0], PX ? ox - 1 + i : (gi >= 0 ? gi : 0), cxo, ntx[q], nwx[q]);
        pr_taps(P.a[1], PY ? oy - 1 + j : (gj >= 0 ? gj : 0), cyo, nty[q], nwy[q]);
    }
    // residual positions: the tile (positions without a value read at the
    // in-plane position NXX + 1 and store nothing, so the rows run unbranched)
    xy_coef<T> kr[NR];
    int rcol[NR];
    bool rvalid[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        const int t = tid + q * NT;
        const int i = t % TX, j = t / TX;
        const int gi = ox + i, gj = oy + j;
        const bool valid = t < TX * TY && gi < nx && gj < ny;
        kr[q] = make_xy_coef<PER>(A, valid, (i + 1) + NXX * (j + 1), gi, gj);
        rvalid[q] = valid;
        if (!valid) kr[q].c = NXX + 1;
        rcol[q] = valid ? gi + nx * gj : 0;
    }
    // e loads: coarse region (cxo + u, cyo + v)
    int ecol[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        const int t = tid + q * NT;
        int col = -1;
        if (t < CP) {
            int u = cxo + t % CXX, v = cyo + t / CXX;
            if (PX) { u %= ncx; if (u < 0) u += ncx; } else if (u < 0 || u >= ncx) u = -1;
            if (PY) { v %= ncy; if (v < 0) v += ncy; } else if (v < 0 || v >= ncy) v = -1;
            if (u >= 0 && v >= 0) col = u + ncx * v;
        }
        ecol[q] = col;
    }

    // coarse planes: physical index on a walled z axis, unwrapped on a periodic one
    auto need = [&](int f) { const int ca = (f + 1) >> 1; return PZ ? ca : min(ca, ncz - 1); };
    auto cslot = [&](int c) { return c & 3; };
    auto fetch_e = [&](int c, T (&pe)[NC]) {
        int k = c;
        if (PZ) {
            k = k < 0 ? k + ncz : k;
            k = k >= ncz ? k - ncz : k;
            if ((unsigned) k >= (unsigned) ncz) {   // coarse z axes of one or two planes
                while (k < 0) k += ncz;
                while (k >= ncz) k -= ncz;
            }
        } else {
            k = min(max(k, 0), ncz - 1);
        }
        const unsigned kk = (unsigned) k * ncxy;
#pragma unroll
        for (int q = 0; q < NC; ++q) pe[q] = e[(unsigned) max(ecol[q], 0) + kk];
    };
    auto store_e = [&](int c, const T (&pe)[NC]) {
        T* dst = Ec[cslot(c)];
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            const int t = tid + q * NT;
            if (t < CP) dst[t] = ecol[q] >= 0 ? pe[q] : T(0);
        }
    };
    auto fetch_x = [&](int f, T (&px)[NN]) {
        const unsigned k = (unsigned) max(axis_global<PZ>(A.a[2], f, 0, 0), 0) * nxy;
#pragma unroll
        for (int q = 0; q < NN; ++q) px[q] = x[(unsigned) max(ncol[q], 0) + k];
    };
    auto fetch_b = [&](int f, T (&pb)[NR]) {
        const unsigned k = (unsigned) min(max(f, z0), z1 - 1) * nxy;
#pragma unroll
        for (int q = 0; q < NR; ++q) pb[q] = b[(unsigned) rcol[q] + k];
    };

    // e planes for the first xn plane (z0 - 1) and the next one prefetched
    int cmax = need(z0 - 1);
    T pe[NC];
    fetch_e(cmax - 1, pe);
    store_e(cmax - 1, pe);
    fetch_e(cmax, pe);
    store_e(cmax, pe);
    fetch_e(cmax + 1, pe);

    // xn ring slot of plane f >= z0 - 1: (f - z0 + 1) & 3, so plane z0 - 1 + q
    // of phase q = (multiple of 4) + S sits in slot S, a constant.
    T px[2][NN], pb[2][NR];
    auto phase = [&]<int S>(int q) {
        constexpr int F = S & 1;
        {   // xn on plane f
            const int f = z0 - 1 + q;
            const int gk = axis_global<PZ>(A.a[2], f, 0, 0);
            T* dst = Xn[S];
            if (f <= z1) {
                if (gk < 0) {
#pragma unroll
                    for (int i = 0; i < NN; ++i)
                        if (nc_[i] >= 0) dst[nc_[i]] = T(0);
                } else {
                    // z taps (prolong_at's ca, cb on the fine plane index)
                    int oz[2];
                    T wz[2];
                    pr_taps(P.a[2], PZ ? f : gk, 0, oz, wz);
                    const T* Ea = Ec[cslot(oz[0])];
                    const T* Eb = Ec[cslot(oz[1])];
#pragma unroll
                    for (int i = 0; i < NN; ++i) {
                        if (nc_[i] < 0) continue;
                        T val = T(0);
                        if (ncol[i] >= 0) {
                            T acc = 0;
#pragma unroll
                            for (int bz = 0; bz < 2; ++bz) {
                                const T* E = bz == 0 ? Ea : Eb;
                                T az = 0;
#pragma unroll
                                for (int a = 0; a < 2; ++a) {
                                    const T* row = E + CXX * nty[i][a];
                                    az += nwy[i][a] * (nwx[i][0] * row[ntx[i][0]] + nwx[i][1] * row[ntx[i][1]]);
                                }
                                acc += wz[bz] * az;
                            }
                            val = px[F][i] + acc;
                            if (nout[i] && f >= z0 && f < z1) xn[(unsigned) ncol[i] + (unsigned) gk * nxy] = val;
                        }
                        dst[nc_[i]] = val;
                    }
                }
            }
        }
        if (q >= 2) {   // residual on plane f - 2 (xn planes f - 3 .. f - 1 from earlier phases)
            const int z = z0 - 3 + q;
            if (z >= z0 && z < z1) {
                const z_coef<T> kz = make_z_coef<PZ>(A.a[2], z);
                const bool in = z > 0 && z < nz - 1;
                const T *Sm = Xn[(S - 3) & 3], *Sc = Xn[(S - 2) & 3], *Sp = Xn[(S - 1) & 3];
                nbr7<T> nb[NR];
#pragma unroll
                for (int i = 0; i < NR; ++i) nb[i] = load7<NXX>(Sm, Sc, Sp, kr[i].c);
                const unsigned kz0 = (unsigned) z * nxy;
#pragma unroll
                for (int i = 0; i < NR; ++i) {
                    const T val = T(-1) * apply7<PER>(A, in, nb[i], kr[i], kz) + T(1) * pb[F][i];
                    if (rvalid[i]) r[(unsigned) rcol[i] + kz0] = val;
                }
            }
        }
        // operands two phases on; the e plane the next phase needs
        fetch_x(z0 + 1 + q, px[F]);
        fetch_b(z0 - 1 + q, pb[F]);
        if (need(z0 + q) > cmax) {
            ++cmax;
            store_e(cmax, pe);
            fetch_e(cmax + 1, pe);
        }
        __syncthreads();
    };

    fetch_x(z0 - 1, px[0]);
    fetch_x(z0, px[1]);
    fetch_b(z0 - 3, pb[0]);
    fetch_b(z0 - 2, pb[1]);
    __syncthreads();
    const int nq = (z1 - z0) + 3;
    auto group = [&]<int... Ss>(int q0, std::integer_sequence<int, Ss...>) {
        ((q0 + Ss < nq ? (phase.template operator()<Ss>(q0 + Ss), 0) : 0), ...);
    };
    for (int q0 = 0; q0 < nq; q0 += 4) group(q0, std::make_integer_sequence<int, 4>{});
}

template <typename T>
__global__ void prolong_add_kernel(transfer3<T> P, const T* __restrict__ coarse, T* __restrict__ fine)
{
    const size_t total = (size_t) P.a[0].nf * P.a[1].nf * P.a[2].nf;
    size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    fine[t] += prolong_at(P, coarse, (unsigned) t);
}

// ---------------------------------------------------------------------------
// The coarse end of a V-cycle in one workgroup: every level of `L` (finest
// first) is smoothed, restricted, solved, prolonged and smoothed again with
// the same per-point arithmetic as the separate kernels, one
// __syncthreads() between passes.  Replaces ~10 launches per coarse level.
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline void coarse_stencil_pass(const op3<T>& A, T alpha, const T* x, T beta, const T* y, T gamma,
                                           const T* z, T* out)
{
    const int nx = A.a[0].n, ny = A.a[1].n, nz = A.a[2].n;
    const int n = nx * ny * nz;
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        int i = t % nx, j = (t / nx) % ny, k = t / (nx * ny);
        out[t] = stencil_at(A, alpha, x, beta, y, gamma, z, i, j, k);
    }
    __syncthreads();
}

template <typename T>
struct coarse_bufs {
    T *x, *b, *t1, *t2;
};

// out (+)= p(A) v with temporaries s1, s2 (all distinct from v and out).
template <typename T>
__device__ inline void coarse_cheb(const coarse_levels<T>& L, int l, T* s1, T* s2, const T* v, T* out, bool add)
{
    const op3<T>& A = L.A[l];
    const T* w = L.w[l];
    coarse_stencil_pass(A, -w[0], v, T(0), (const T*) nullptr, w[1], v, s1);
    coarse_stencil_pass(A, T(-1), (const T*) s1, T(0), (const T*) nullptr, w[2], v, s2);
    if (add) coarse_stencil_pass(A, T(-1), (const T*) s2, T(1), (const T*) out, w[3], v, out);
    else coarse_stencil_pass(A, T(-1), (const T*) s2, T(0), (const T*) nullptr, w[3], v, out);
}

// Four LDS vectors per level suffice: going down, x = S b with temporaries
// t1, t2, then the residual in t1 is restricted into the next level's b;
// going up, the residual b - A x goes to t1 and the post-smoothing
// x += S t1 uses t2 and the (no longer needed) b as temporaries.
template <typename T>
__global__ void __launch_bounds__(COARSE_THREADS) coarse_vcycle_kernel(coarse_levels<T> L)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char coarse_smem[];
    T* base = reinterpret_cast<T*>(coarse_smem);
    auto bufs = [&](int l) {
        coarse_bufs<T> c;
        T* p = base + L.off[l];
        const int n = L.n[l];
        c.x = p;
        c.b = p + n;
        c.t1 = p + 2 * n;
        c.t2 = p + 3 * n;
        return c;
    };
    {
        const coarse_bufs<T> B0 = bufs(0);
        for (int t = threadIdx.x; t < L.n[0]; t += blockDim.x) B0.b[t] = L.b0[t];
    }
    __syncthreads();
    const int last = L.count - 1;
    for (int l = 0; l <= last; ++l) {
        const coarse_bufs<T> B = bufs(l);
        coarse_cheb(L, l, B.t1, B.t2, (const T*) B.b, B.x, false);
        if (l == last) break;
        coarse_stencil_pass(L.A[l], T(-1), (const T*) B.x, T(0), (const T*) nullptr, T(1), (const T*) B.b, B.t1);
        const transfer3<T>& R = L.R[l];
        T* nb = bufs(l + 1).b;
        for (int t = threadIdx.x; t < L.n[l + 1]; t += blockDim.x) nb[t] = restrict_at(R, (const T*) B.t1, (unsigned) t);
        __syncthreads();
    }
    for (int l = last - 1; l >= 0; --l) {
        const coarse_bufs<T> B = bufs(l);
        const T* cx = bufs(l + 1).x;
        const transfer3<T>& R = L.R[l];
        for (int t = threadIdx.x; t < L.n[l]; t += blockDim.x) B.x[t] += prolong_at(R, cx, (unsigned) t);
        __syncthreads();
        coarse_stencil_pass(L.A[l], T(-1), (const T*) B.x, T(0), (const T*) nullptr, T(1), (const T*) B.b, B.t1);
        coarse_cheb(L, l, B.t2, B.b, (const T*) B.t1, B.x, true);
    }
    {
        const coarse_bufs<T> B0 = bufs(0);
        for (int t = threadIdx.x; t < L.n[0]; t += blockDim.x) L.x0[t] = B0.x[t];
    }
}

// ---------------------------------------------------------------------------
// Two-point operators (differential / average along one axis)
// out = alpha * op x + beta * out
// ---------------------------------------------------------------------------
// Value of the operator at destination point c.
template <typename T>
__device__ inline T twopoint_value(const twopoint_op<T>& P, const int c[3], const T* __restrict__ x)
{
    // 32-bit indices: the launchers keep every grid below 2^32 points
    const int v = P.view;
    const int cols = P.sn[v];
    const unsigned sn0 = (unsigned) P.sn[0], sn1 = (unsigned) P.sn[1];
    const unsigned stride = v == 0 ? 1u : (v == 1 ? sn0 : sn0 * sn1);
    int base[3] = {c[0], c[1], c[2]};
    base[v] = 0;
    const unsigned b0 = (unsigned) base[0] + sn0 * ((unsigned) base[1] + sn1 * (unsigned) base[2]);
    auto fetch = [&](int idx) -> T {
        if (idx >= 0 && idx < cols) return x[b0 + stride * (unsigned) idx];
        if (P.periodic) return x[b0 + stride * (unsigned) (idx < 0 ? idx + cols : idx - cols)];
        return idx < 0 ? P.int_lo * x[b0] : P.int_hi * x[b0 + stride * (unsigned) (cols - 1)];
    };
    const int r = c[v];
    return P.c0 * fetch(r - P.s1) + P.c1 * fetch(r + 1 - P.s1);
}

template <typename T>
__global__ void twopoint_kernel(twopoint_op<T> P, T alpha, const T* __restrict__ x, T beta, T* __restrict__ out)
{
    const size_t total = (size_t) P.dn[0] * P.dn[1] * P.dn[2];
    size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int c[3];
    {
        const unsigned u = (unsigned) t, q = u / (unsigned) P.dn[0];
        c[0] = (int) (u - q * (unsigned) P.dn[0]);
        c[1] = (int) (q % (unsigned) P.dn[1]);
        c[2] = (int) (q / (unsigned) P.dn[1]);
    }
    T val = twopoint_value(P, c, x);
    T o = alpha * val;
    if (beta != T(0)) o += beta * out[t];
    out[t] = o;
}

// One point of a two-point operator whose source values come from a functor
// src(x, y, z) of source-grid coordinates (same fetch rules as
// twopoint_kernel: periodic wrap, or the ghost coefficient times the end
// value).
template <typename T, typename F>
__device__ inline T twopoint_at(const twopoint_op<T>& P, int x, int y, int z, const F& src)
{
    int c[3] = {x, y, z};
    const int v = P.view, cols = P.sn[v], r = c[v];
    auto fetch = [&](int idx) -> T {
        if (idx < 0 || idx >= cols) {
            if (P.periodic) {
                idx = idx < 0 ? idx + cols : idx - cols;
            } else {
                T g = idx < 0 ? P.int_lo : P.int_hi;
                c[v] = idx < 0 ? 0 : cols - 1;
                return g * src(c[0], c[1], c[2]);
            }
        }
        c[v] = idx;
        return src(c[0], c[1], c[2]);
    };
    T f0 = fetch(r - P.s1);
    T f1 = fetch(r + 1 - P.s1);
    return P.c0 * f0 + P.c1 * f1;
}

// twopoint_at with the view a compile-time constant (no run-time indexing
// of the coordinate triple); the same fetches and the same arithmetic.
template <int V, typename T, typename F>
__device__ inline T twopoint_at_v(const twopoint_op<T>& P, int x, int y, int z, const F& src)
{
    const int cols = P.sn[V], r = V == 0 ? x : (V == 1 ? y : z);
    auto at = [&](int idx) -> T {
        return V == 0 ? src(idx, y, z) : (V == 1 ? src(x, idx, z) : src(x, y, idx));
    };
    auto fetch = [&](int idx) -> T {
        if (idx < 0 || idx >= cols) {
            if (P.periodic) {
                idx = idx < 0 ? idx + cols : idx - cols;
            } else {
                T g = idx < 0 ? P.int_lo : P.int_hi;
                return g * at(idx < 0 ? 0 : cols - 1);
            }
        }
        return at(idx);
    };
    T f0 = fetch(r - P.s1);
    T f1 = fetch(r + 1 - P.s1);
    return P.c0 * f0 + P.c1 * f1;
}

// ins::advection for one component in one pass:
//   H^i = sum_j D_j [ (A_i u^j) o (A_j u^i) ]
// evaluated point by point on u^i's grid (j summed in order 0, 1, 2).  The
// component I is a template parameter, so every operator's axis is known
// at compile time (D_j and A_j along j, A_i along I).
template <typename T, int I>
__global__ void advection_kernel(advection_op<T> Op, const T* __restrict__ u0, const T* __restrict__ u1,
                                 const T* __restrict__ u2, T* __restrict__ H)
{
    const int* dn = Op.D[0].dn;
    const size_t total = (size_t) dn[0] * dn[1] * dn[2];
    size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const unsigned tt = (unsigned) t, q = tt / (unsigned) dn[0];
    const int cx = (int) (tt - q * (unsigned) dn[0]);
    const int cy = (int) (q % (unsigned) dn[1]);
    const int cz = (int) (q / (unsigned) dn[1]);
    const T* ui = I == 0 ? u0 : (I == 1 ? u1 : u2);
    T acc = 0;
    auto term = [&](auto J) {
        constexpr int j = decltype(J)::value;
        const twopoint_op<T>& A = Op.A[j];
        const twopoint_op<T>& B = Op.B[j];
        const T* uj = j == 0 ? u0 : (j == 1 ? u1 : u2);
        // 32-bit indices: the launcher keeps every grid below 2^32 points
        const unsigned as0 = (unsigned) A.sn[0], as1 = (unsigned) A.sn[1];
        const unsigned bs0 = (unsigned) B.sn[0], bs1 = (unsigned) B.sn[1];
        auto prod = [&](int x, int y, int z) -> T {
            T a = twopoint_at_v<I>(A, x, y, z, [&](int p, int q2, int r) -> T {
                return uj[(unsigned) p + as0 * ((unsigned) q2 + as1 * (unsigned) r)];
            });
            T b = twopoint_at_v<j>(B, x, y, z, [&](int p, int q2, int r) -> T {
                return ui[(unsigned) p + bs0 * ((unsigned) q2 + bs1 * (unsigned) r)];
            });
            return a * b;
        };
        T d = twopoint_at_v<j>(Op.D[j], cx, cy, cz, prod);
        acc = j == 0 ? d : acc + d;
    };
    term(std::integral_constant<int, 0>{});
    term(std::integral_constant<int, 1>{});
    term(std::integral_constant<int, 2>{});
    H[t] = acc;
}

// ---------------------------------------------------------------------------
// Vector kernels
// ---------------------------------------------------------------------------
template <typename T, typename U>
__global__ void axpby_kernel(size_t n, T a, const U* __restrict__ x, T b, T* __restrict__ y)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t) gridDim.x * blockDim.x;
    for (; i < n; i += stride) y[i] = a * (T) x[i] + (b == T(0) ? T(0) : b * y[i]);
}

template <typename T, typename U>
__global__ void convert_kernel(size_t n, const U* __restrict__ x, T* __restrict__ y)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t) gridDim.x * blockDim.x;
    for (; i < n; i += stride) y[i] = (T) x[i];
}

template <typename T>
__global__ void fill_kernel(size_t n, T v, T* __restrict__ y)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t) gridDim.x * blockDim.x;
    for (; i < n; i += stride) y[i] = v;
}

// Deterministic reductions: a fixed grid of REDUCE_BLOCKS blocks writes
// partial sums in block order, one block adds them in index order.
constexpr int REDUCE_BLOCKS = 1024;
constexpr int REDUCE_THREADS = 256;

__device__ inline double block_sum(double v)
{
    __shared__ double part[REDUCE_THREADS / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < REDUCE_THREADS / 64; ++w) s += part[w];
    __syncthreads();
    return s;
}

template <typename T, typename U>
__global__ void __launch_bounds__(REDUCE_THREADS)
dot_partial_kernel(size_t n, const T* __restrict__ x, const U* __restrict__ y, double* __restrict__ partial)
{
    double s = 0;
    for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t) gridDim.x * blockDim.x)
        s += (double) x[i] * (double) y[i];
    s = block_sum(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

template <typename T>
__global__ void __launch_bounds__(REDUCE_THREADS)
sum_partial_kernel(size_t n, const T* __restrict__ x, double* __restrict__ partial)
{
    double s = 0;
    for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t) gridDim.x * blockDim.x)
        s += (double) x[i];
    s = block_sum(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void __launch_bounds__(REDUCE_THREADS)
finish_kernel(int nparts, const double* __restrict__ partial, double* __restrict__ out)
{
    double s = 0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += partial[i];
    s = block_sum(s);
    if (threadIdx.x == 0) *out = s;
}

template <typename T>
__global__ void shift_kernel(size_t n, const double* __restrict__ sum, double scale, T* __restrict__ y)
{
    // y -= (sum * scale): removes the mean of a vector on a singular operator
    T m = (T) (*sum * scale);
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t) gridDim.x * blockDim.x;
    for (; i < n; i += stride) y[i] -= m;
}

// Maximum that carries a NaN through (fmax would drop it): a blown-up
// field must not report a finite maximum.
__device__ inline double nan_max(double m, double v)
{
    return (isnan(v) || v > m) ? v : m;
}

__global__ void __launch_bounds__(REDUCE_THREADS)
max_abs_partial_kernel(size_t n, const double* __restrict__ x, double* __restrict__ partial)
{
    __shared__ double part[REDUCE_THREADS / 64];
    double m = 0;
    for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t) gridDim.x * blockDim.x)
        m = nan_max(m, fabs(x[i]));
    for (int off = 32; off > 0; off >>= 1) m = nan_max(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0;
        for (int w = 0; w < REDUCE_THREADS / 64; ++w) r = nan_max(r, part[w]);
        partial[blockIdx.x] = r;
    }
}

__global__ void __launch_bounds__(REDUCE_THREADS)
finish_max_kernel(int nparts, const double* __restrict__ partial, double* __restrict__ out)
{
    __shared__ double part[REDUCE_THREADS / 64];
    double m = 0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) m = nan_max(m, partial[i]);
    for (int off = 32; off > 0; off >>= 1) m = nan_max(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0;
        for (int w = 0; w < REDUCE_THREADS / 64; ++w) r = nan_max(r, part[w]);
        *out = r;
    }
}

// sqrt(num / den); 0 for a zero denominator (an empty right-hand side), NaN
// when either operand is not finite.
__global__ void sqrt_ratio_kernel(const double* __restrict__ num, const double* __restrict__ den,
                                  double* __restrict__ out)
{
    const double d = *den, n = *num;
    if (!isfinite(d) || !isfinite(n)) *out = __longlong_as_double(0x7ff8000000000000ll);
    else *out = d > 0 ? sqrt(n / d) : 0.0;
}

// ||r|| / ||b|| from residual_update_kernel's block sums (r^2 in
// partial[0, nparts), b^2 in partial[nparts, 2 nparts)), with
// sqrt_ratio_kernel's rules.
__global__ void __launch_bounds__(REDUCE_THREADS)
relres_finish_kernel(int nparts, const double* __restrict__ partial, double* __restrict__ out)
{
    double r = 0, b = 0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        r += partial[i];
        b += partial[nparts + i];
    }
    r = block_sum(r);
    b = block_sum(b);
    if (threadIdx.x == 0) {
        if (!isfinite(b) || !isfinite(r)) *out = __longlong_as_double(0x7ff8000000000000ll);
        else *out = b > 0 ? sqrt(r / b) : 0.0;
    }
}

// The divergence D_0 u_0 + D_1 u_1 + D_2 u_2 in one pass (twopoint_kernel's
// value of each term, added in that order), times scale, stored in out; or,
// with MAX, nothing stored and the largest |value| of each block's points
// (NaN carried through) in partial[block].
template <bool MAX>
__global__ void __launch_bounds__(REDUCE_THREADS)
divergence_kernel(twopoint_op<double> P0, twopoint_op<double> P1, twopoint_op<double> P2, double scale,
                  const double* __restrict__ u0, const double* __restrict__ u1, const double* __restrict__ u2,
                  double* __restrict__ out, double* __restrict__ partial)
{
    const unsigned n0 = (unsigned) P0.dn[0], n1 = (unsigned) P0.dn[1];
    const size_t total = (size_t) n0 * n1 * P0.dn[2];
    double m = 0;
    for (size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t) gridDim.x * blockDim.x) {
        int c[3];
        {
            const unsigned u = (unsigned) t, q = u / n0;
            c[0] = (int) (u - q * n0);
            c[1] = (int) (q % n1);
            c[2] = (int) (q / n1);
        }
        double v = twopoint_value(P0, c, u0);
        v = twopoint_value(P1, c, u1) + v;
        v = twopoint_value(P2, c, u2) + v;
        if (MAX) m = nan_max(m, fabs(v));
        else out[t] = v * scale;
    }
    if (MAX) {
        __shared__ double part[REDUCE_THREADS / 64];
        for (int off = 32; off > 0; off >>= 1) m = nan_max(m, __shfl_down(m, off, 64));
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = 0;
            for (int w = 0; w < REDUCE_THREADS / 64; ++w) r = nan_max(r, part[w]);
            partial[blockIdx.x] = r;
        }
    }
}

// y = alpha * B x + beta * y for the diagonal ins::boundary operator: one
// thread per point; beta = 1 leaves the interior untouched (only the points
// on some axis's first or last row are read and written).
__global__ void ins_boundary_kernel(ins_boundary_op B, double alpha, const double* __restrict__ x, double beta,
                                    double* __restrict__ y)
{
    const size_t n = (size_t) B.n[0] * B.n[1] * B.n[2];
    for (size_t k = (size_t) blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t) gridDim.x * blockDim.x) {
        const int i0 = (int) (k % B.n[0]);
        const size_t q = k / B.n[0];
        const int i1 = (int) (q % B.n[1]);
        const int i2 = (int) (q / B.n[1]);
        const int idx[3] = {i0, i1, i2};
        double w = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (idx[d] == 0) w += B.lo[d];
            if (idx[d] == B.n[d] - 1) w += B.hi[d];
        }
        if (beta == 1.0) {
            if (w != 0.0) y[k] += alpha * w * x[k];
        } else {
            const double old = beta == 0.0 ? 0.0 : beta * y[k];
            y[k] = alpha * w * x[k] + old;
        }
    }
}

// y[row] += alpha * scale * g[face] * prod_{d != view} id_d(i_d) on the face
// rows (index `row` along `view`) of a grid; g indexed over the other axes,
// x fastest (the columns of the kron-assembled boundary operator).
__global__ void boundary_kernel(boundary_op B, double alpha, const double* __restrict__ g, double value,
                                double* __restrict__ y)
{
    int a0 = B.view == 0 ? 1 : 0;
    int a1 = B.view == 2 ? 1 : 2;
    long nface = (long) B.n[a0] * B.n[a1];
    for (long f = (long) blockIdx.x * blockDim.x + threadIdx.x; f < nface; f += (long) gridDim.x * blockDim.x) {
        int i0 = (int) (f % B.n[a0]), i1 = (int) (f / B.n[a0]);
        int idx[3];
        idx[B.view] = B.row;
        idx[a0] = i0;
        idx[a1] = i1;
        double w = 1.0;
        if (i0 == 0) w *= B.id_lo[a0];
        if (i0 == B.n[a0] - 1) w *= B.id_hi[a0];
        if (i1 == 0) w *= B.id_lo[a1];
        if (i1 == B.n[a1] - 1) w *= B.id_hi[a1];
        double gv = g ? g[f] : value;
        long at = idx[0] + (long) B.n[0] * (idx[1] + (long) B.n[1] * idx[2]);
        y[at] += alpha * B.scale * gv * w;
    }
}

inline int grid_for(size_t n, int threads = 256)
{
    size_t b = (n + threads - 1) / threads;
    return (int) (b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

} // namespace ibm

// ---------------------------------------------------------------------------
// Host launchers (fd_kernels.hpp)
// ---------------------------------------------------------------------------
namespace ibm {

template <typename T, int OUT>
static ibm_status stencil_dispatch(ibm_ctx* ctx, const op3<T>& A, T alpha, const T* x, T beta, const T* y, T gamma,
                                   const T* z, T* out, float* out32, double* acc)
{
    const int nx = A.a[0].n, ny = A.a[1].n, nz = A.a[2].n;
    if (nx < 1 || ny < 1 || nz < 1) return IBM_OK;
    dim3 block(STX, STY);
    dim3 grid((nx + STX - 1) / STX, (ny + STY - 1) / STY, (nz + KZ - 1) / KZ);
    // x, y, z read and out written; the float copy adds 4 bytes, the FP64
    // read-modify-write of acc replaces the store of out by 16 bytes
    const double streams = (OUT >= 2 ? 1.0 : 2.0) + (y ? 1.0 : 0.0) + (z ? 1.0 : 0.0);
    const double extra = OUT == 1 || OUT == 3 ? 4.0 : OUT == 2 ? 16.0 : 0.0;
    prof_scope prof(ctx, sizeof(T) == 8 ? "stencil_f64" : "stencil_f32",
                    (streams * sizeof(T) + extra) * nx * ny * (double) nz);
    if (y && z) stencil_kernel<T, true, true, OUT><<<grid, block, 0, ctx->stream>>>(A, alpha, x, beta, y, gamma, z, out, out32, acc);
    else if (y) stencil_kernel<T, true, false, OUT><<<grid, block, 0, ctx->stream>>>(A, alpha, x, beta, y, gamma, z, out, out32, acc);
    else if (z) stencil_kernel<T, false, true, OUT><<<grid, block, 0, ctx->stream>>>(A, alpha, x, beta, y, gamma, z, out, out32, acc);
    else stencil_kernel<T, false, false, OUT><<<grid, block, 0, ctx->stream>>>(A, alpha, x, beta, y, gamma, z, out, out32, acc);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

template <typename T>
ibm_status launch_stencil(ibm_ctx* ctx, const op3<T>& A, T alpha, const T* x, T beta, const T* y, T gamma,
                          const T* z, T* out)
{
    return stencil_dispatch<T, 0>(ctx, A, alpha, x, beta, y, gamma, z, out, nullptr, nullptr);
}

// One instantiation per periodicity pattern of the three axes.
template <typename T, bool ADD, bool ACC, bool WIDE>
static void cheb_launch(int per, int blocks, hipStream_t stream, const op3<T>& A, T w0, T w1, T w2, T w3, const T* v,
                        T* out, const T* addsrc, double* acc, int tiles_x, int tiles_xy, int zchunk, int nblocks,
                        int xcd_chunk)
{
    constexpr int threads = march_tile<T, WIDE>::THREADS;
#define IBM_CHEB_CASE(P)                                                                                     \
    case P:                                                                                                  \
        cheb_march_kernel<T, ADD, ACC, P, WIDE><<<blocks, threads, 0, stream>>>(A, w0, w1, w2, w3, v, out, addsrc, acc, \
                                                                               tiles_x, tiles_xy, zchunk,    \
                                                                               nblocks, xcd_chunk);          \
        break;
    switch (per) {
        IBM_CHEB_CASE(0)
        IBM_CHEB_CASE(1)
        IBM_CHEB_CASE(2)
        IBM_CHEB_CASE(3)
        IBM_CHEB_CASE(4)
        IBM_CHEB_CASE(5)
        IBM_CHEB_CASE(6)
        IBM_CHEB_CASE(7)
    }
#undef IBM_CHEB_CASE
}

template <typename T, bool WIDE>
static void cheb_configure(const op3<T>& A, const double w[4], const T* v, T* out, const T* addsrc, bool add,
                           double* acc, hipStream_t stream, int target)
{
    constexpr int TX = march_tile<T, WIDE>::X, TY = march_tile<T, WIDE>::Y;
    const int nx = A.a[0].n, ny = A.a[1].n, nz = A.a[2].n;
    const int tiles_x = (nx + TX - 1) / TX, tiles_xy = tiles_x * ((ny + TY - 1) / TY);
    // Enough z chunks for about `target` blocks (MARCH_TARGET_BLOCKS unless
    // tuned); each chunk costs MARCH_LEAD phases of pipeline fill.
    if (target <= 0) target = MARCH_TARGET_BLOCKS;
    int zchunk = (int) (((long long) nz * tiles_xy + target - 1) / target);
    zchunk = std::min(nz, std::max(zchunk, march_tile<T, WIDE>::MIN_Z));
    const int nblocks = tiles_xy * ((nz + zchunk - 1) / zchunk);
    const int blocks = (nblocks + 7) / 8 * 8;
    const int xcd_chunk = blocks / 8;
    const T w0 = (T) w[0], w1 = (T) w[1], w2 = (T) w[2], w3 = (T) w[3];
    const int per = (A.a[0].periodic ? 1 : 0) | (A.a[1].periodic ? 2 : 0) | (A.a[2].periodic ? 4 : 0);
    if (acc)
        cheb_launch<T, true, true, WIDE>(per, blocks, stream, A, w0, w1, w2, w3, v, out, addsrc, acc, tiles_x, tiles_xy, zchunk,
                                         nblocks, xcd_chunk);
    else if (add)
        cheb_launch<T, true, false, WIDE>(per, blocks, stream, A, w0, w1, w2, w3, v, out, addsrc, acc, tiles_x, tiles_xy, zchunk,
                                          nblocks, xcd_chunk);
    else
        cheb_launch<T, false, false, WIDE>(per, blocks, stream, A, w0, w1, w2, w3, v, out, addsrc, acc, tiles_x, tiles_xy,
                                           zchunk, nblocks, xcd_chunk);
}

bool cheb_fused_fits(int nx, int ny)
{
    return nx >= 1 && ny >= 1 && nx <= 32767 && ny <= 32767;   // packed x-y coordinates
}

template <typename T>
ibm_status launch_cheb_fused(ibm_ctx* ctx, const op3<T>& A, const double w[4], const T* v, T* out, bool add,
                             double* acc, const T* addsrc)
{
    if (!addsrc) addsrc = out;
    const int nx = A.a[0].n, ny = A.a[1].n, nz = A.a[2].n;
    if (nx < 1 || ny < 1 || nz < 1) return IBM_OK;
    if (!cheb_fused_fits(nx, ny)) return fail(ctx, IBM_ERR_ARG, "fused smoothing: %d x %d planes too wide", nx, ny);
    const double n = (double) nx * ny * nz;
    // v read, out written (read too when adding; FP64 read-modify-write with acc)
    const double bytes = n * (sizeof(T) + (acc ? 16.0 + sizeof(T) : (add ? 2.0 : 1.0) * sizeof(T)));
    prof_scope prof(ctx, sizeof(T) == 8 ? "chebyshev_f64" : "chebyshev_f32", bytes);
    const int target = tuned_target(TUNE_CHEB, (int) sizeof(T), nx, ny, nz, per_mask(A));
    if (march_wide((size_t) nx * ny * nz)) cheb_configure<T, true>(A, w, v, out, addsrc, add, acc, ctx->stream, target);
    else cheb_configure<T, false>(A, w, v, out, addsrc, add, acc, ctx->stream, target);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

ibm_status launch_stencil_split(ibm_ctx* ctx, const op3<double>& A, double alpha, const double* x, double beta,
                                const double* y, double gamma, const double* z, double* out, float* out32)
{
    if (!out) return stencil_dispatch<double, 3>(ctx, A, alpha, x, beta, y, gamma, z, nullptr, out32, nullptr);
    return stencil_dispatch<double, 1>(ctx, A, alpha, x, beta, y, gamma, z, out, out32, nullptr);
}

template <typename T>
ibm_status launch_stencil_acc(ibm_ctx* ctx, const op3<T>& A, T alpha, const T* x, T beta, const T* y, T gamma,
                              const T* z, double* acc)
{
    return stencil_dispatch<T, 2>(ctx, A, alpha, x, beta, y, gamma, z, nullptr, nullptr, acc);
}

// Per-point kernels index grids with 32-bit unsigned arithmetic.
constexpr size_t MAX_GRID_POINTS = 0xffffffffull;

template <typename T>
ibm_status launch_restrict(ibm_ctx* ctx, const transfer3<T>& R, const T* fine, T* coarse)
{
    size_t n = (size_t) R.a[0].nc * R.a[1].nc * R.a[2].nc;
    if ((size_t) R.a[0].nf * R.a[1].nf * R.a[2].nf > MAX_GRID_POINTS) return fail(ctx, IBM_ERR_SIZE, "grid too large");
    size_t nf = (size_t) R.a[0].nf * R.a[1].nf * R.a[2].nf;
    prof_scope prof(ctx, "restrict", (double) sizeof(T) * (nf + n));
    restrict_kernel<T><<<(unsigned) ((n + 255) / 256), 256, 0, ctx->stream>>>(R, fine, coarse);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

template <typename T, bool WIDE>
static void rr_configure(const op3<T>& A, const transfer3<T>& R, const T* x, const T* b, T* coarse,
                         hipStream_t stream, int target)
{
    constexpr int CX = rr_tile<T, WIDE>::CX, CY = rr_tile<T, WIDE>::CY;
    const int ncx = R.a[0].nc, ncy = R.a[1].nc, ncz = R.a[2].nc;
    const int tiles_x = (ncx + CX - 1) / CX, tiles_xy = tiles_x * ((ncy + CY - 1) / CY);
    if (target <= 0) target = RR_TARGET_BLOCKS;
    int kchunk = (int) (((long long) ncz * tiles_xy + target - 1) / target);
    kchunk = std::min(ncz, std::max(kchunk, 2));
    const int nblocks = tiles_xy * ((ncz + kchunk - 1) / kchunk);
    const int blocks = (nblocks + 7) / 8 * 8;
    const int per = (A.a[0].periodic ? 1 : 0) | (A.a[1].periodic ? 2 : 0) | (A.a[2].periodic ? 4 : 0);
    constexpr int threads = rr_tile<T, WIDE>::THREADS;
#define IBM_RR_CASE(P)                                                                                       \
    case P:                                                                                                  \
        residual_restrict_kernel<T, P, WIDE><<<blocks, threads, 0, stream>>>(A, R, x, b, coarse, tiles_x,    \
                                                                            tiles_xy, kchunk, nblocks,       \
                                                                            blocks / 8);                     \
        break;
    switch (per) {
        IBM_RR_CASE(0)
        IBM_RR_CASE(1)
        IBM_RR_CASE(2)
        IBM_RR_CASE(3)
        IBM_RR_CASE(4)
        IBM_RR_CASE(5)
        IBM_RR_CASE(6)
        IBM_RR_CASE(7)
    }
#undef IBM_RR_CASE
}

template <typename T>
ibm_status launch_residual_restrict(ibm_ctx* ctx, const op3<T>& A, const transfer3<T>& R, const T* x, const T* b,
                                    T* coarse)
{
    const size_t nf = (size_t) R.a[0].nf * R.a[1].nf * R.a[2].nf;
    const size_t nc = (size_t) R.a[0].nc * R.a[1].nc * R.a[2].nc;
    if (nf > MAX_GRID_POINTS || (size_t) R.a[0].nf * R.a[1].nf > 0x7fffffff)
        return fail(ctx, IBM_ERR_SIZE, "grid too large");
    for (int d = 0; d < 3; ++d)
        if (R.a[d].nf != A.a[d].n || (R.a[d].periodic != 0) != (A.a[d].periodic != 0))
            return fail(ctx, IBM_ERR_ARG, "residual/restriction: operator and transfer grids differ");
    // x and b read, coarse written
    prof_scope prof(ctx, "residual_restrict", (double) sizeof(T) * (2 * nf + nc));
    const int target = tuned_target(TUNE_RR, (int) sizeof(T), A.a[0].n, A.a[1].n, A.a[2].n, per_mask(A));
    if (march_wide(nf)) rr_configure<T, true>(A, R, x, b, coarse, ctx->stream, target);
    else rr_configure<T, false>(A, R, x, b, coarse, ctx->stream, target);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

template <typename T, bool WIDE>
static void pr_configure(const op3<T>& A, const transfer3<T>& P, const T* e, const T* x, const T* b, T* xn, T* r,
                         hipStream_t stream, int target)
{
    constexpr int TX = pr_tile<T, WIDE>::X, TY = pr_tile<T, WIDE>::Y;
    const int nx = A.a[0].n, ny = A.a[1].n, nz = A.a[2].n;
    const int tiles_x = (nx + TX - 1) / TX, tiles_xy = tiles_x * ((ny + TY - 1) / TY);
    if (target <= 0) target = PR_TARGET_BLOCKS;
    int zchunk = (int) (((long long) nz * tiles_xy + target - 1) / target);
    zchunk = std::min(nz, std::max(zchunk, 2));
    const int nblocks = tiles_xy * ((nz + zchunk - 1) / zchunk);
    const int blocks = (nblocks + 7) / 8 * 8;
    const int per = (A.a[0].periodic ? 1 : 0) | (A.a[1].periodic ? 2 : 0) | (A.a[2].periodic ? 4 : 0);
    constexpr int threads = pr_tile<T, WIDE>::THREADS;
#define IBM_PR_CASE(PP)                                                                                      \
    case PP:                                                                                                 \
        prolong_residual_kernel<T, PP, WIDE><<<blocks, threads, 0, stream>>>(A, P, e, x, b, xn, r, tiles_x,  \
                                                                            tiles_xy, zchunk, nblocks,       \
                                                                            blocks / 8);                     \
        break;
    switch (per) {
        IBM_PR_CASE(0)
        IBM_PR_CASE(1)
        IBM_PR_CASE(2)
        IBM_PR_CASE(3)
        IBM_PR_CASE(4)
        IBM_PR_CASE(5)
        IBM_PR_CASE(6)
        IBM_PR_CASE(7)
    }
#undef IBM_PR_CASE
}

template <typename T>
ibm_status launch_prolong_residual(ibm_ctx* ctx, const op3<T>& A, const transfer3<T>& P, const T* e, const T* x,
                                   const T* b, T* xn, T* r)
{
    const size_t nf = (size_t) P.a[0].nf * P.a[1].nf * P.a[2].nf;
    const size_t nc = (size_t) P.a[0].nc * P.a[1].nc * P.a[2].nc;
    if (nf > MAX_GRID_POINTS || (size_t) P.a[0].nf * P.a[1].nf > 0x7fffffff)
        return fail(ctx, IBM_ERR_SIZE, "grid too large");
    for (int d = 0; d < 3; ++d)
        if (P.a[d].nf != A.a[d].n || (P.a[d].periodic != 0) != (A.a[d].periodic != 0))
            return fail(ctx, IBM_ERR_ARG, "prolongation/residual: operator and transfer grids differ");
    // e, x, b read, xn and r written
    prof_scope prof(ctx, "prolong_residual", (double) sizeof(T) * (nc + 4 * nf));
    const int target = tuned_target(TUNE_PR, (int) sizeof(T), A.a[0].n, A.a[1].n, A.a[2].n, per_mask(A));
    if (march_wide(nf)) pr_configure<T, true>(A, P, e, x, b, xn, r, ctx->stream, target);
    else pr_configure<T, false>(A, P, e, x, b, xn, r, ctx->stream, target);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

template <typename T>
ibm_status launch_prolong_add(ibm_ctx* ctx, const transfer3<T>& P, const T* coarse, T* fine)
{
    size_t n = (size_t) P.a[0].nf * P.a[1].nf * P.a[2].nf;
    if (n > MAX_GRID_POINTS) return fail(ctx, IBM_ERR_SIZE, "grid too large");
    size_t nc = (size_t) P.a[0].nc * P.a[1].nc * P.a[2].nc;
    prof_scope prof(ctx, "prolong", (double) sizeof(T) * (nc + 2 * n));
    prolong_add_kernel<T><<<(unsigned) ((n + 255) / 256), 256, 0, ctx->stream>>>(P, coarse, fine);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

template <typename T>
ibm_status launch_twopoint(ibm_ctx* ctx, const twopoint_op<T>& P, T alpha, const T* x, T beta, T* out)
{
    size_t n = (size_t) P.dn[0] * P.dn[1] * P.dn[2];
    if (!n) return IBM_OK;
    if (n > MAX_GRID_POINTS || (size_t) P.sn[0] * P.sn[1] * P.sn[2] > MAX_GRID_POINTS)
        return fail(ctx, IBM_ERR_SIZE, "grid too large");
    size_t ns = (size_t) P.sn[0] * P.sn[1] * P.sn[2];
    prof_scope prof(ctx, "twopoint", (double) sizeof(T) * (ns + n + (beta != T(0) ? n : 0)));
    twopoint_kernel<T><<<(unsigned) ((n + 255) / 256), 256, 0, ctx->stream>>>(P, alpha, x, beta, out);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

template <typename T>
ibm_status launch_advection(ibm_ctx* ctx, const advection_op<T>& Op, const T* const* u, int comp, T* H)
{
    const int* dn = Op.D[0].dn;
    size_t n = (size_t) dn[0] * dn[1] * dn[2];
    if (!n) return IBM_OK;
    if (n > MAX_GRID_POINTS) return fail(ctx, IBM_ERR_SIZE, "grid too large");
    double nsrc = 0;
    for (int d = 0; d < 3; ++d) {
        const size_t na = (size_t) Op.A[d].sn[0] * Op.A[d].sn[1] * Op.A[d].sn[2];
        const size_t nb = (size_t) Op.B[d].sn[0] * Op.B[d].sn[1] * Op.B[d].sn[2];
        if (na > MAX_GRID_POINTS || nb > MAX_GRID_POINTS) return fail(ctx, IBM_ERR_SIZE, "grid too large");
        nsrc += (double) na;
    }
    prof_scope prof(ctx, "advection", (double) sizeof(T) * (nsrc + n));   // u^0..2 read once, H^i written
    if (comp < 0 || comp > 2) return fail(ctx, IBM_ERR_ARG, "advection component %d", comp);
    for (int j = 0; j < 3; ++j)
        if (Op.D[j].view != j || Op.B[j].view != j || Op.A[j].view != comp)
            return fail(ctx, IBM_ERR_ARG, "advection operators along unexpected axes");
    const unsigned blocks = (unsigned) ((n + 255) / 256);
    if (comp == 0) advection_kernel<T, 0><<<blocks, 256, 0, ctx->stream>>>(Op, u[0], u[1], u[2], H);
    else if (comp == 1) advection_kernel<T, 1><<<blocks, 256, 0, ctx->stream>>>(Op, u[0], u[1], u[2], H);
    else advection_kernel<T, 2><<<blocks, 256, 0, ctx->stream>>>(Op, u[0], u[1], u[2], H);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

template <typename T, typename U>
ibm_status launch_axpby(ibm_ctx* ctx, size_t n, T a, const U* x, T b, T* y)
{
    axpby_kernel<T, U><<<grid_for(n), 256, 0, ctx->stream>>>(n, a, x, b, y);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

template <typename T, typename U>
ibm_status launch_convert(ibm_ctx* ctx, size_t n, const U* x, T* y)
{
    convert_kernel<T, U><<<grid_for(n), 256, 0, ctx->stream>>>(n, x, y);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

template <typename T>
ibm_status launch_fill(ibm_ctx* ctx, size_t n, T v, T* y)
{
    fill_kernel<T><<<grid_for(n), 256, 0, ctx->stream>>>(n, v, y);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

template <typename T, typename U>
ibm_status launch_dot(ibm_ctx* ctx, size_t n, const T* x, const U* y, double* partial, double* out)
{
    dot_partial_kernel<T, U><<<REDUCE_BLOCKS, REDUCE_THREADS, 0, ctx->stream>>>(n, x, y, partial);
    IBM_LAUNCH_CHECK(ctx);
    finish_kernel<<<1, REDUCE_THREADS, 0, ctx->stream>>>(REDUCE_BLOCKS, partial, out);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

template <typename T>
ibm_status launch_sum(ibm_ctx* ctx, size_t n, const T* x, double* partial, double* out)
{
    sum_partial_kernel<T><<<REDUCE_BLOCKS, REDUCE_THREADS, 0, ctx->stream>>>(n, x, partial);
    IBM_LAUNCH_CHECK(ctx);
    finish_kernel<<<1, REDUCE_THREADS, 0, ctx->stream>>>(REDUCE_BLOCKS, partial, out);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

template <typename T>
ibm_status launch_remove_mean(ibm_ctx* ctx, size_t n, T* y, double* partial, double* sum)
{
    IBM_TRY(launch_sum<T>(ctx, n, y, partial, sum));
    shift_kernel<T><<<grid_for(n), 256, 0, ctx->stream>>>(n, sum, 1.0 / (double) n, y);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

ibm_status launch_max_abs(ibm_ctx* ctx, size_t n, const double* x, double* partial, double* out)
{
    max_abs_partial_kernel<<<REDUCE_BLOCKS, REDUCE_THREADS, 0, ctx->stream>>>(n, x, partial);
    IBM_LAUNCH_CHECK(ctx);
    finish_max_kernel<<<1, REDUCE_THREADS, 0, ctx->stream>>>(REDUCE_BLOCKS, partial, out);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

ibm_status launch_sqrt_ratio(ibm_ctx* ctx, const double* num, const double* den, double* out)
{
    sqrt_ratio_kernel<<<1, 1, 0, ctx->stream>>>(num, den, out);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

static dim3 fused_grid(int nx, int ny, int nz)
{
    return dim3((nx + STX - 1) / STX, (ny + STY - 1) / STY, (nz + KZF - 1) / KZF);
}

size_t residual_parts(int nx, int ny, int nz)
{
    const dim3 g = fused_grid(nx, ny, nz);
    return 2 * (size_t) g.x * g.y * g.z;
}

ibm_status launch_momentum_rhs(ibm_ctx* ctx, const op3<double>& A, double dt, double dtnu, double inv_rho,
                               const double* F, const double* H, const double* u, const double* b, double* rhs)
{
    const int nx = A.a[0].n, ny = A.a[1].n, nz = A.a[2].n;
    if (nx < 1 || ny < 1 || nz < 1) return IBM_OK;
    prof_scope prof(ctx, "momentum_rhs", (b ? 40.0 : 32.0) * nx * ny * (double) nz);
    const dim3 grid = fused_grid(nx, ny, nz), block(STX, STY);
    if (b) momentum_rhs_kernel<true><<<grid, block, 0, ctx->stream>>>(A, dt, dtnu, inv_rho, F, H, u, b, rhs);
    else momentum_rhs_kernel<false><<<grid, block, 0, ctx->stream>>>(A, dt, dtnu, inv_rho, F, H, u, b, rhs);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

ibm_status launch_residual_update(ibm_ctx* ctx, const op3<double>& A, const double* x, const double* b, double* u,
                                  double* partial, size_t parts, double* out)
{
    const int nx = A.a[0].n, ny = A.a[1].n, nz = A.a[2].n;
    if (nx < 1 || ny < 1 || nz < 1) return IBM_OK;
    const size_t need = residual_parts(nx, ny, nz);
    if (need > parts || need / 2 > (size_t) INT32_MAX)
        return fail(ctx, IBM_ERR_SIZE, "residual check: %zu partial sums, room for %zu", need, parts);
    prof_scope prof(ctx, "residual_update", (u ? 32.0 : 16.0) * nx * ny * (double) nz);
    const dim3 grid = fused_grid(nx, ny, nz), block(STX, STY);
    if (u) residual_update_kernel<true><<<grid, block, 0, ctx->stream>>>(A, x, b, u, partial);
    else residual_update_kernel<false><<<grid, block, 0, ctx->stream>>>(A, x, b, u, partial);
    IBM_LAUNCH_CHECK(ctx);
    relres_finish_kernel<<<1, REDUCE_THREADS, 0, ctx->stream>>>((int) (need / 2), partial, out);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

ibm_status launch_divergence(ibm_ctx* ctx, const twopoint_op<double> P[3], double scale, const double* const* u,
                             double* out, double* partial, double* max_out)
{
    const size_t n = (size_t) P[0].dn[0] * P[0].dn[1] * P[0].dn[2];
    if (!n) return IBM_OK;
    size_t src = 0;
    for (int i = 0; i < 3; ++i) {
        const size_t ns = (size_t) P[i].sn[0] * P[i].sn[1] * P[i].sn[2];
        if (ns > MAX_GRID_POINTS) return fail(ctx, IBM_ERR_SIZE, "grid too large");
        src += ns;
    }
    if (n > MAX_GRID_POINTS) return fail(ctx, IBM_ERR_SIZE, "grid too large");
    prof_scope prof(ctx, "divergence", 8.0 * (double) (src + (out ? n : 0)));
    if (out) {
        divergence_kernel<false><<<grid_for(n), REDUCE_THREADS, 0, ctx->stream>>>(P[0], P[1], P[2], scale, u[0], u[1],
                                                                                  u[2], out, nullptr);
    } else {
        divergence_kernel<true><<<REDUCE_BLOCKS, REDUCE_THREADS, 0, ctx->stream>>>(P[0], P[1], P[2], scale, u[0],
                                                                                   u[1], u[2], nullptr, partial);
        IBM_LAUNCH_CHECK(ctx);
        finish_max_kernel<<<1, REDUCE_THREADS, 0, ctx->stream>>>(REDUCE_BLOCKS, partial, max_out);
    }
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

ibm_status launch_ins_boundary(ibm_ctx* ctx, const ins_boundary_op& B, double alpha, const double* x, double beta,
                               double* y)
{
    const size_t n = (size_t) B.n[0] * B.n[1] * B.n[2];
    if (!n) return IBM_OK;
    prof_scope prof(ctx, "ins_boundary", (beta == 0.0 ? 16.0 : 24.0) * (double) n);
    ins_boundary_kernel<<<grid_for(n), 256, 0, ctx->stream>>>(B, alpha, x, beta, y);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

ibm_status launch_boundary(ibm_ctx* ctx, const boundary_op& B, double alpha, const double* g, double value, double* y)
{
    int a0 = B.view == 0 ? 1 : 0, a1 = B.view == 2 ? 1 : 2;
    size_t nface = (size_t) B.n[a0] * B.n[a1];
    boundary_kernel<<<grid_for(nface), 256, 0, ctx->stream>>>(B, alpha, g, value, y);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

template <typename T>
ibm_status launch_coarse_vcycle(ibm_ctx* ctx, const coarse_levels<T>& L)
{
    prof_scope prof(ctx, "coarse_vcycle");
    coarse_vcycle_kernel<T><<<1, COARSE_THREADS, L.lds_bytes, ctx->stream>>>(L);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

size_t coarse_lds_limit(int device)
{
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess || v <= 0)
        return 64 * 1024;
    return (size_t) v;
}

// The attribute is process-wide per kernel instantiation, while hierarchies
// of different sizes may be created in any order: it is only ever raised,
// so a hierarchy created later never lowers what an earlier one launches with.
template <typename T>
ibm_status prepare_coarse_vcycle(ibm_ctx* ctx, size_t lds_bytes)
{
    static std::mutex lock;
    static size_t set_bytes = 0;
    std::lock_guard<std::mutex> guard(lock);
    if (lds_bytes <= set_bytes) return IBM_OK;
    hipError_t e = hipFuncSetAttribute((const void*) coarse_vcycle_kernel<T>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds_bytes);
    if (e != hipSuccess) return hip_fail(ctx, e, "coarse multigrid kernel LDS size");
    set_bytes = lds_bytes;
    return IBM_OK;
}

#define IBM_INSTANTIATE(T)                                                                              \
    template ibm_status launch_stencil<T>(ibm_ctx*, const op3<T>&, T, const T*, T, const T*, T, const T*, T*); \
    template ibm_status launch_stencil_acc<T>(ibm_ctx*, const op3<T>&, T, const T*, T, const T*, T, const T*, double*); \
    template ibm_status launch_cheb_fused<T>(ibm_ctx*, const op3<T>&, const double*, const T*, T*, bool, double*, const T*); \
    template ibm_status launch_restrict<T>(ibm_ctx*, const transfer3<T>&, const T*, T*);               \
    template ibm_status launch_residual_restrict<T>(ibm_ctx*, const op3<T>&, const transfer3<T>&, const T*, const T*, T*); \
    template ibm_status launch_prolong_add<T>(ibm_ctx*, const transfer3<T>&, const T*, T*);            \
    template ibm_status launch_prolong_residual<T>(ibm_ctx*, const op3<T>&, const transfer3<T>&, const T*, const T*, const T*, T*, T*); \
    template ibm_status launch_fill<T>(ibm_ctx*, size_t, T, T*);                                       \
    template ibm_status launch_sum<T>(ibm_ctx*, size_t, const T*, double*, double*);                   \
    template ibm_status launch_remove_mean<T>(ibm_ctx*, size_t, T*, double*, double*);                 \
    template ibm_status launch_twopoint<T>(ibm_ctx*, const twopoint_op<T>&, T, const T*, T, T*);      \
    template ibm_status launch_coarse_vcycle<T>(ibm_ctx*, const coarse_levels<T>&);                   \
    template ibm_status prepare_coarse_vcycle<T>(ibm_ctx*, size_t);                   \
    template ibm_status launch_advection<T>(ibm_ctx*, const advection_op<T>&, const T* const*, int, T*);
IBM_INSTANTIATE(double)
IBM_INSTANTIATE(float)
template ibm_status launch_axpby<double, double>(ibm_ctx*, size_t, double, const double*, double, double*);
template ibm_status launch_axpby<double, float>(ibm_ctx*, size_t, double, const float*, double, double*);
template ibm_status launch_axpby<float, float>(ibm_ctx*, size_t, float, const float*, float, float*);
template ibm_status launch_convert<float, double>(ibm_ctx*, size_t, const double*, float*);
template ibm_status launch_convert<double, float>(ibm_ctx*, size_t, const float*, double*);
template ibm_status launch_convert<double, double>(ibm_ctx*, size_t, const double*, double*);
template ibm_status launch_dot<double, double>(ibm_ctx*, size_t, const double*, const double*, double*, double*);
template ibm_status launch_dot<float, float>(ibm_ctx*, size_t, const float*, const float*, double*, double*);

ibm_status shifted_grid(const ibm_grid& src, int view, ibm_grid* dst)
{
    ibm_dimension dims[3];
    double al[3];
    for (int i = 0; i < src.dims; ++i) {
        const ibm_axis& a = src.axis[i];
        dims[i].length = a.length;
        dims[i].solid = a.solid;
        dims[i].reserved = 0;
        dims[i].lower[0] = a.robin[0][0]; dims[i].lower[1] = a.robin[0][1];
        dims[i].upper[0] = a.robin[1][0]; dims[i].upper[1] = a.robin[1][1];
        al[i] = a.alignment;
    }
    // fd/differential.h:49-50: the new alignment is shift() + 1/2 (mod 1)
    double sh = src.axis[view].shift + 0.5;
    al[view] = sh - std::floor(sh);
    return ibm_grid_make(dims, src.dims, src.refinement, al, dst);
}


// ---------------------------------------------------------------------------
// Batched operators over nbatch independent copies of one grid, stored one
// after another (vector of copy b at offset b * n): the stencil, restriction
// and prolongation of the kernels above, each point by the same device
// functions (stencil_at, restrict_at, prolong_at), one launch for all copies.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
batch_stencil_kernel(op3<T> A, int nbatch, T alpha, const T* __restrict__ x, T beta, const T* __restrict__ y, T gamma,
                     const T* __restrict__ z, T* __restrict__ out)
{
    const int nx = A.a[0].n, ny = A.a[1].n;
    const size_t n = (size_t) nx * ny * A.a[2].n;
    const size_t total = n * nbatch;
    for (size_t g = (size_t) blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t) gridDim.x * blockDim.x) {
        const size_t b = g / n, t = g - b * n;
        const size_t o = b * n;
        const int i = (int) (t % nx), j = (int) ((t / nx) % ny), k = (int) (t / ((size_t) nx * ny));
        out[g] = stencil_at(A, alpha, x + o, beta, y ? y + o : y, gamma, z ? z + o : z, i, j, k);
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
batch_restrict_kernel(transfer3<T> R, int nbatch, const T* __restrict__ fine, T* __restrict__ coarse)
{
    const size_t nc = (size_t) R.a[0].nc * R.a[1].nc * R.a[2].nc;
    const size_t nf = (size_t) R.a[0].nf * R.a[1].nf * R.a[2].nf;
    const size_t total = nc * nbatch;
    for (size_t g = (size_t) blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t) gridDim.x * blockDim.x) {
        const size_t b = g / nc, t = g - b * nc;
        coarse[g] = restrict_at(R, fine + b * nf, (unsigned) t);
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
batch_prolong_add_kernel(transfer3<T> P, int nbatch, const T* __restrict__ coarse, T* __restrict__ fine)
{
    const size_t nc = (size_t) P.a[0].nc * P.a[1].nc * P.a[2].nc;
    const size_t nf = (size_t) P.a[0].nf * P.a[1].nf * P.a[2].nf;
    const size_t total = nf * nbatch;
    for (size_t g = (size_t) blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t) gridDim.x * blockDim.x) {
        const size_t b = g / nf, t = g - b * nf;
        fine[g] += prolong_at(P, coarse + b * nc, (unsigned) t);
    }
}

// Per copy: out[b] = sum over its n entries of x (x * y when y is given),
// in FP64 with a fixed summation order (one workgroup per copy).
template <typename T>
__global__ void __launch_bounds__(256)
batch_dot_kernel(size_t n, const T* __restrict__ x, const T* __restrict__ y, double* __restrict__ out)
{
    __shared__ double red[256];
    const size_t o = (size_t) blockIdx.x * n;
    double acc = 0;
    for (size_t t = threadIdx.x; t < n; t += blockDim.x) {
        double v = (double) x[o + t];
        acc += y ? v * (double) y[o + t] : v;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int) threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

static int batch_blocks(size_t total)
{
    size_t b = (total + 255) / 256;
    return (int) (b < 65536 ? (b ? b : 1) : 65536);
}

template <typename T>
ibm_status launch_batch_stencil(ibm_ctx* ctx, const op3<T>& A, int nbatch, T alpha, const T* x, T beta, const T* y,
                                T gamma, const T* z, T* out)
{
    size_t total = (size_t) A.a[0].n * A.a[1].n * A.a[2].n * nbatch;
    prof_scope prof(ctx, sizeof(T) == 8 ? "batch_stencil_f64" : "batch_stencil_f32",
                    (double) sizeof(T) * total * (2 + (y ? 1 : 0) + (z ? 1 : 0)));
    batch_stencil_kernel<T><<<batch_blocks(total), 256, 0, ctx->stream>>>(A, nbatch, alpha, x, beta, y, gamma, z, out);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

template <typename T>
ibm_status launch_batch_restrict(ibm_ctx* ctx, const transfer3<T>& R, int nbatch, const T* fine, T* coarse)
{
    size_t total = (size_t) R.a[0].nc * R.a[1].nc * R.a[2].nc * nbatch;
    batch_restrict_kernel<T><<<batch_blocks(total), 256, 0, ctx->stream>>>(R, nbatch, fine, coarse);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

template <typename T>
ibm_status launch_batch_prolong_add(ibm_ctx* ctx, const transfer3<T>& P, int nbatch, const T* coarse, T* fine)
{
    size_t total = (size_t) P.a[0].nf * P.a[1].nf * P.a[2].nf * nbatch;
    batch_prolong_add_kernel<T><<<batch_blocks(total), 256, 0, ctx->stream>>>(P, nbatch, coarse, fine);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

template <typename T>
ibm_status launch_batch_dot(ibm_ctx* ctx, size_t n, int nbatch, const T* x, const T* y, double* out)
{
    batch_dot_kernel<T><<<nbatch, 256, 0, ctx->stream>>>(n, x, y, out);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

#define IBM_INSTANTIATE_BATCH(T)                                                                              \
    template ibm_status launch_batch_stencil<T>(ibm_ctx*, const op3<T>&, int, T, const T*, T, const T*, T, const T*, T*); \
    template ibm_status launch_batch_restrict<T>(ibm_ctx*, const transfer3<T>&, int, const T*, T*);          \
    template ibm_status launch_batch_prolong_add<T>(ibm_ctx*, const transfer3<T>&, int, const T*, T*);       \
    template ibm_status launch_batch_dot<T>(ibm_ctx*, size_t, int, const T*, const T*, double*);
IBM_INSTANTIATE_BATCH(double)
IBM_INSTANTIATE_BATCH(float)

// Gershgorin bounds over row classes (see fd_ops.hpp).
void gershgorin_bounds(const ibm_grid& g, double sigma, double tau, double* lo, double* hi)
{
    struct cls { double ld, lo, id; };
    std::vector<cls> c[3];
    for (int d = 0; d < 3; ++d) {
        const ibm_axis& a = g.axis[d];
        double r2 = a.resolution * a.resolution;
        int n = a.points;
        if (d >= g.dims) { c[d].push_back({0, 0, 1}); continue; }
        if (!a.solid) {
            if (n == 1) c[d].push_back({0, 0, 1});
            else c[d].push_back({-2 * r2, 2 * r2, 1});
            continue;
        }
        if (n == 1) {
            c[d].push_back({(-2 + a.interior[0] + a.interior[1]) * r2, 0, 1 + a.laplacian[0] + a.laplacian[1]});
            continue;
        }
        c[d].push_back({(-2 + a.interior[0]) * r2, r2, 1 + a.laplacian[0]});
        c[d].push_back({(-2 + a.interior[1]) * r2, r2, 1 + a.laplacian[1]});
        if (n >= 3) c[d].push_back({-2 * r2, 2 * r2, 1});
    }
    double mn = INFINITY, mx = -INFINITY;
    for (auto& x : c[0])
        for (auto& y : c[1])
            for (auto& z : c[2]) {
                double center = sigma * x.id * y.id * z.id +
                                tau * (x.ld * y.id * z.id + x.id * y.ld * z.id + x.id * y.id * z.ld);
                double radius = std::fabs(tau) * (x.lo * std::fabs(y.id * z.id) + y.lo * std::fabs(x.id * z.id) +
                                                  z.lo * std::fabs(x.id * y.id));
                mn = std::fmin(mn, center - radius);
                mx = std::fmax(mx, center + radius);
            }
    *lo = mn;
    *hi = mx;
}

} // namespace ibm
