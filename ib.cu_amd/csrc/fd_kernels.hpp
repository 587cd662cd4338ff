// Host launchers of the stencil, transfer, vector and reduction kernels
// (implemented in fd_ops.hip; all enqueue on ctx->stream).
#pragma once
#include "context.hpp"
#include "fd_ops.hpp"

namespace ibm {

// out = alpha * A x + beta * y + gamma * z   (y, z may be null)
template <typename T>
ibm_status launch_stencil(ibm_ctx* ctx, const op3<T>& A, T alpha, const T* x, T beta, const T* y, T gamma,
                          const T* z, T* out);
// Planes the one-pass smoother handles (x and y extents below 32768).
bool cheb_fused_fits(int nx, int ny);
// out = p(A) v, or out = src + p(A) v with add (src = addsrc, or out when
// null), the degree-4 Chebyshev smoothing of weights w in one pass; with acc,
// acc += src + p(A) v instead (add must be true).
template <typename T>
ibm_status launch_cheb_fused(ibm_ctx* ctx, const op3<T>& A, const double w[4], const T* v, T* out, bool add,
                             double* acc, const T* addsrc = nullptr);
// out = alpha A x + beta y + gamma z, also stored as floats in out32 (only
// there when out is null)
ibm_status launch_stencil_split(ibm_ctx* ctx, const op3<double>& A, double alpha, const double* x, double beta,
                                const double* y, double gamma, const double* z, double* out, float* out32);
// acc += alpha A x + beta y + gamma z (acc in FP64)
template <typename T>
ibm_status launch_stencil_acc(ibm_ctx* ctx, const op3<T>& A, T alpha, const T* x, T beta, const T* y, T gamma,
                              const T* z, double* acc);
template <typename T>
ibm_status launch_restrict(ibm_ctx* ctx, const transfer3<T>& R, const T* fine, T* coarse);
// coarse = R (b - A x) in one pass (A on the fine grid of R).
template <typename T>
ibm_status launch_residual_restrict(ibm_ctx* ctx, const op3<T>& A, const transfer3<T>& R, const T* x, const T* b,
                                    T* coarse);
template <typename T>
ibm_status launch_prolong_add(ibm_ctx* ctx, const transfer3<T>& P, const T* coarse, T* fine);
// xn = x + P e and r = b - A xn in one pass (A on the fine grid of P).
template <typename T>
ibm_status launch_prolong_residual(ibm_ctx* ctx, const op3<T>& A, const transfer3<T>& P, const T* e, const T* x,
                                   const T* b, T* xn, T* r);
// y = a x + b y
template <typename T, typename U>
ibm_status launch_axpby(ibm_ctx* ctx, size_t n, T a, const U* x, T b, T* y);
template <typename T, typename U>
ibm_status launch_convert(ibm_ctx* ctx, size_t n, const U* x, T* y);
template <typename T>
ibm_status launch_fill(ibm_ctx* ctx, size_t n, T v, T* y);
// *out = sum x_i y_i in FP64, deterministic order; `partial` holds REDUCE_PARTS doubles
template <typename T, typename U>
ibm_status launch_dot(ibm_ctx* ctx, size_t n, const T* x, const U* y, double* partial, double* out);
template <typename T>
ibm_status launch_sum(ibm_ctx* ctx, size_t n, const T* x, double* partial, double* out);
// y -= mean(y)
template <typename T>
ibm_status launch_remove_mean(ibm_ctx* ctx, size_t n, T* y, double* partial, double* sum);

// The coarse levels of a V-cycle, finest first, run by one workgroup with
// all their vectors (x, b and two temporaries per level) in LDS; only the
// first level's right-hand side is read from and its solution written to
// device memory.
constexpr int MAX_COARSE_LEVELS = 6;
constexpr int COARSE_VECTORS = 4;
// Larger levels run faster as separate grid-wide kernels than in one
// workgroup (measured at 8K points: ~130 us in one workgroup vs ~85 us).
constexpr size_t COARSE_MAX_POINTS = 2048;
constexpr int COARSE_THREADS = 1024;
template <typename T>
struct coarse_levels {
    int count;
    int n[MAX_COARSE_LEVELS];              // points per level
    int off[MAX_COARSE_LEVELS];            // first LDS element of each level's five vectors
    op3<T> A[MAX_COARSE_LEVELS];
    transfer3<T> R[MAX_COARSE_LEVELS];     // level l -> l + 1
    T w[MAX_COARSE_LEVELS][4];             // Chebyshev weights
    const T* b0;                           // first level's right-hand side (device)
    T* x0;                                 // first level's solution (device)
    size_t lds_bytes;
};
template <typename T>
ibm_status launch_coarse_vcycle(ibm_ctx* ctx, const coarse_levels<T>& L);
// Largest dynamic LDS of one workgroup on the device, and the kernel
// attribute that allows a launch to use `lds_bytes` of it.
size_t coarse_lds_limit(int device);
template <typename T>
ibm_status prepare_coarse_vcycle(ibm_ctx* ctx, size_t lds_bytes);

template <typename T>
ibm_status launch_twopoint(ibm_ctx* ctx, const twopoint_op<T>& P, T alpha, const T* x, T beta, T* out);
// ins::advection of one component: A[j] = A_i on u^j's grid, B[j] = A_j on
// u^i's grid, D[j] = D_j from their common grid back to u^i's grid.
template <typename T>
struct advection_op {
    twopoint_op<T> A[3], B[3], D[3];
};
template <typename T>
ibm_status launch_advection(ibm_ctx* ctx, const advection_op<T>& Op, const T* const* u, int comp, T* H);


// *out = max |x_i|
ibm_status launch_max_abs(ibm_ctx* ctx, size_t n, const double* x, double* partial, double* out);
// *out = sqrt(*num / *den), 0 when *den is 0
ibm_status launch_sqrt_ratio(ibm_ctx* ctx, const double* num, const double* den, double* out);

// rhs = dt I~ (F inv_rho - H) + dt nu (b + L u) in one pass: I~ and L are the
// sigma and tau parts of A (the identity with its wall-row weights and the
// Laplacian); b may be null.
ibm_status launch_momentum_rhs(ibm_ctx* ctx, const op3<double>& A, double dt, double dtnu, double inv_rho,
                               const double* F, const double* H, const double* u, const double* b, double* rhs);
// *out = ||b - A x|| / ||b|| (sqrt_ratio's rules), and u += x when u is not
// null, in one pass over the grid; `partial` holds `parts` doubles, at least
// residual_parts(n) of A's extents.
ibm_status launch_residual_update(ibm_ctx* ctx, const op3<double>& A, const double* x, const double* b, double* u,
                                  double* partial, size_t parts, double* out);
size_t residual_parts(int nx, int ny, int nz);
// out = scale (D_0 u_0 + D_1 u_1 + D_2 u_2) for the differentials P (one
// destination grid); with out null, *max_out = max |D_0 u_0 + D_1 u_1 + D_2 u_2|
// instead (`partial`: REDUCE_PARTS doubles).
ibm_status launch_divergence(ibm_ctx* ctx, const twopoint_op<double> P[3], double scale, const double* const* u,
                             double* out, double* partial, double* max_out);

// fd::lower_boundary / fd::upper_boundary applied to face data
// (fd/boundary_ops.h:21-71): the ghost-cell boundary coefficient of the face
// row along `view`, scaled by resolution^order, times the (corrected)
// identity of the other axes.
struct boundary_op {
    int n[3];
    int view, row;
    double scale;
    double id_lo[3], id_hi[3];
};
ibm_status launch_boundary(ibm_ctx* ctx, const boundary_op& B, double alpha, const double* g, double value, double* y);

// ins::boundary(grid, dim) (ins/boundary.h:16-56): a diagonal operator whose
// entry at a point is the sum, over the walled axes other than `dim`, of the
// ghost-cell boundary coefficient (times resolution^order) when the point is
// on that axis's first (lo) or last (hi) row.  y = alpha B x + beta y.
struct ins_boundary_op {
    int n[3];
    double lo[3], hi[3];   // 0 on periodic axes and on `dim`
};
ibm_status launch_ins_boundary(ibm_ctx* ctx, const ins_boundary_op& B, double alpha, const double* x, double beta,
                               double* y);
ins_boundary_op make_ins_boundary(const ibm_grid& g, int dim, int order);

// The differential (kind 0) or average (kind 1) of grid `src` along `view`.
ibm_status make_twopoint(const ibm_grid& src, int view, int kind, twopoint_op<double>* op, ibm_grid* dst);
// ibm_divergence's sum times scale into out, or (out null) its maximum
// modulus into *max_out; one launch.
ibm_status divergence_of(ibm_ctx* ctx, const ibm_grid* u_grids, const double* const* u, double scale, double* out,
                         double* partial, double* max_out);

// Block targets of the z-marching kernels (smoothing, residual + restriction,
// prolongation + residual) per grid: the number of blocks the z chunking aims
// at.  0 means the built-in default; the multigrid's autotuner times the
// candidates on each level and records the fastest here (process-wide, keyed
// by the fine grid's shape, periodicity and element size).  The tile (wide
// for grids of MARCH_WIDE_POINTS points or more, narrow below) stays fixed
// by the grid size: a target only moves chunk boundaries, so every point is
// computed by the same code either way and results do not depend on tuning.
enum { TUNE_CHEB = 0, TUNE_RR = 1, TUNE_PR = 2 };
int tuned_target(int kind, int elem, int nx, int ny, int nz, int per);
void set_tuned_target(int kind, int elem, int nx, int ny, int nz, int per, int target);
bool target_tuned(int kind, int elem, int nx, int ny, int nz, int per);
void clear_tuned_targets();
// Whether launches on an nx x ny x nz grid use the wide tile.
bool march_wide(size_t points);
// The built-in block targets.
constexpr int MARCH_DEFAULT_TARGET = 512;
constexpr int TRANSFER_DEFAULT_TARGET = 1024;

// Batched forms over nbatch copies of one grid stored back to back.
template <typename T>
ibm_status launch_batch_stencil(ibm_ctx* ctx, const op3<T>& A, int nbatch, T alpha, const T* x, T beta, const T* y,
                                T gamma, const T* z, T* out);
template <typename T>
ibm_status launch_batch_restrict(ibm_ctx* ctx, const transfer3<T>& R, int nbatch, const T* fine, T* coarse);
template <typename T>
ibm_status launch_batch_prolong_add(ibm_ctx* ctx, const transfer3<T>& P, int nbatch, const T* coarse, T* fine);
// out[b] = sum of copy b of x (times y when given), FP64, fixed order
template <typename T>
ibm_status launch_batch_dot(ibm_ctx* ctx, size_t n, int nbatch, const T* x, const T* y, double* out);

constexpr int REDUCE_PARTS = 1024;

} // namespace ibm
