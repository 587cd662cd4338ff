/*
 * ibm.h — C ABI of the MI355X immersed-boundary timestep library (libibm.so).
 *
 * Every entry point is plain C: device pointers, sizes and POD descriptors,
 * no C++ or HIP types in the signatures.  Work is enqueued on the context's
 * HIP stream (ibm_set_stream) and is asynchronous unless a function says it
 * returns a host scalar.  Errors are reported as ibm_status codes; the text
 * of the last error on a context is returned by ibm_last_error().
 *
 * Conventions shared with the reference (sonwell/ib.cu):
 *   - marker coordinates and forces are SoA, column-major n x 3:
 *       x[n*i + m] is coordinate i of marker m            (ib/novel.h:83)
 *   - grid vectors are linearised x-fastest:
 *       k = i0 + p0*(i1 + p1*i2), p = points per dimension (ib/indexing.h:27-33)
 *   - lengths are in library units: 1 unit = 10 um, 1 time unit = 1 ms
 *                                                      (units.h:16-18)
 */
#ifndef IBM_H
#define IBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Revision of the entry points' signatures and structures; bindings check it
 * against ibm_abi_version() before calling anything else. */
#define IBM_ABI_VERSION 4
int ibm_abi_version(void);

/* ------------------------------------------------------------------ */
/* Status codes.  include/ibm.hpp turns them into exceptions           */
/* (ibm::error, ins::too_small).                                        */
/* ------------------------------------------------------------------ */
typedef enum ibm_status {
    IBM_OK = 0,
    IBM_ERR_HIP = 1,            /* a HIP runtime call failed                      */
    IBM_ERR_ARG = 2,            /* invalid argument (null pointer, bad enum, ...) */
    IBM_ERR_SIZE = 3,           /* size mismatch           (linalg/size.h:17-23)  */
    IBM_ERR_NOT_CONVERGED = 4,  /* iterative solve hit its iteration cap          */
    IBM_ERR_DT_TOO_SMALL = 5,   /* time step bound below kmin  (ins/tracker.h:93) */
    IBM_ERR_BAD_GRID = 6        /* bad dimension / grid        (fd/dimension.h:29)*/
} ibm_status;

/* Regularised delta kernels on a 4-point stencil.
 *   COSINE   : phi(r) = (1 + cos(pi r / 2)) / 4          (ib/cosine.h:9-16)
 *   PESKIN   : Peskin's 4-point kernel                   (ib/peskin.h:6-14)
 *   BSPLINE3 : cubic cardinal B-spline                   (ib/bspline.h:13-27, n = 3) */
typedef enum ibm_delta {
    IBM_DELTA_COSINE = 0,
    IBM_DELTA_PESKIN = 1,
    IBM_DELTA_BSPLINE3 = 2
} ibm_delta;

/* One dimension of the physical box, fd::dimension (fd/dimension.h:36-88).
 * A periodic dimension has solid = 0.  A solid dimension carries a Robin
 * condition  alpha*u + beta*du/dn = gamma  at each wall (fd/boundary.h:31-42):
 * Dirichlet = (1, 0), Neumann = (0, 1). */
typedef struct ibm_dimension {
    double length;
    int solid;
    int reserved;
    double lower[2];
    double upper[2];
} ibm_dimension;

/* Discretisation of one dimension for one grid, fd::discretization
 * (fd/discretization.h:87-179).  Filled by ibm_grid_make(). */
typedef struct ibm_axis {
    double length;       /* dimension length                                   */
    double resolution;   /* 1/h                                                */
    double alignment;    /* staggering in [0,1): 0 = on cell faces, 0.5 = centres */
    double shift;        /* alignment + (solid && on_boundary)      (:113)     */
    int cells;           /* floor(length*resolution + 0.5)          (:161)     */
    int points;          /* cells - (on_boundary && solid)          (:162)     */
    int solid;
    int reserved;
    double boundary[2];  /* ghost-cell coefficient of gamma   [lower, upper] (:56) */
    double interior[2];  /* ghost-cell coefficient of u_0     [lower, upper] (:60) */
    double laplacian[2]; /* second-order identity correction  [lower, upper] (:66) */
    double robin[2][2];  /* (alpha, beta) at [lower, upper], copied from the dimension */
} ibm_axis;

typedef struct ibm_grid {
    int dims;            /* number of dimensions (1..3)                        */
    int refinement;      /* cells per base unit               (fd/grid.h:43-45) */
    double unit;         /* base unit = floating GCD of the lengths (fd/domain.h:31) */
    ibm_axis axis[3];
} ibm_grid;

/* ------------------------------------------------------------------ */
/* Grids (host only, no device work)                                    */
/* ------------------------------------------------------------------ */

/* Base unit of a domain: GCD of the dimension lengths (fd/domain.h:30-32,
 * algo/gcd.h:27-34). */
ibm_status ibm_domain_unit(const ibm_dimension* dims, int ndims, double* unit);

/* Grid with the given per-dimension alignments and refinement; the
 * counterpart of fd::grid{domain, refinement, cell} (fd/grid.h:77-82). */
ibm_status ibm_grid_make(const ibm_dimension* dims, int ndims, int refinement,
                         const double* alignment, ibm_grid* out);

/* The ndims component grids of a MAC (staggered) tag: component i is aligned
 * on cell faces in dimension i and on cell centres elsewhere
 * (fd/cell.h:106-114,194-196); fd::grid{fd::mac{refinement}, domain, dim_i}. */
ibm_status ibm_mac_grids(const ibm_dimension* dims, int ndims, int refinement,
                         ibm_grid* out /* [ndims] */);

/* The cell-centred grid (pressure), fd::grid{fd::shift::diagonally(mac), domain}
 * (test/ib-whole.cc:179). */
ibm_status ibm_centered_grid(const ibm_dimension* dims, int ndims, int refinement,
                             ibm_grid* out);

/* Coarsened copy of a grid: refinement >> 1 (solvers/multigrid.h:189-190). */
ibm_status ibm_grid_coarsen(const ibm_grid* fine, ibm_grid* coarse);

/* Total number of grid points, fd::size(grid) (fd/size.h:33-41). */
int64_t ibm_grid_size(const ibm_grid* g);

/* ------------------------------------------------------------------ */
/* Context                                                              */
/* ------------------------------------------------------------------ */
typedef struct ibm_ctx ibm_ctx;

/* Creates a context bound to a HIP device.  The context owns its scratch
 * workspaces (grown on demand) and uses the null stream until
 * ibm_set_stream() is called.  One context per host thread. */
ibm_status ibm_create(ibm_ctx** ctx, int device);
ibm_status ibm_destroy(ibm_ctx* ctx);
/* stream is a hipStream_t passed as void* (0 = null stream). */
ibm_status ibm_set_stream(ibm_ctx* ctx, void* stream);
ibm_status ibm_synchronize(ibm_ctx* ctx);
/* Text of the last error reported on ctx (never NULL). */
const char* ibm_last_error(const ibm_ctx* ctx);

/* Kernel timing (no counterpart in the reference, whose cuda::timer of
 * include/cuda/timer.h:9-19 is never instantiated).  While enabled, the
 * library records a hipEvent pair on the context stream around each named
 * kernel launch ("spread_brick", "interp", "bin", ...).  Enabling clears the
 * records.  ibm_profile_read synchronises the stream and returns the summed
 * event time and launch count of one kernel name. */
ibm_status ibm_profile_enable(ibm_ctx* ctx, int on);
ibm_status ibm_profile_read(ibm_ctx* ctx, const char* name, double* total_ms, int* launches);
/* Summed algorithmic work of the recorded launches of one kernel name:
 * compulsory HBM bytes (operands read once, results written once) for the
 * bandwidth-bound kernels, flops (2 M N K) for "rbf_gemm". */
ibm_status ibm_profile_work(ibm_ctx* ctx, const char* name, double* total);

/* ------------------------------------------------------------------ */
/* Immersed-boundary coupling (FP64)                                    */
/* ------------------------------------------------------------------ */

/* Spreads Lagrangian forces onto the component grids:
 *     fe_c(g) = h^-3 * sum_m f_c(m) * prod_j phi((g_j - X_j(m)) / h)
 * Replaces ib::novel::spread<domain, delta>::operator()(n, x, fl, fe)
 * (ib/novel.h:197-227; algorithm :58-168).  x and f are SoA n x 3.
 * Grid points whose stencil index falls outside a solid dimension are
 * dropped (ib/indexing.h:35-62).  Every point of fe[c] is written (no
 * accumulation into existing contents).  Deterministic: the result does not
 * depend on scheduling (no floating-point atomics in global memory). */
ibm_status ibm_spread(ibm_ctx* ctx, const ibm_grid* grids /* [3] */, ibm_delta delta,
                      int n, const double* x, const double* f,
                      double* const* fe /* [3] device pointers */);

/* Batched form: nsets independent marker sets (e.g. every RBC in one set and
 * every platelet in another, each with its own marker count) spread in one
 * pass into the same fe.  x[k], f[k] are SoA n[k] x 3. */
ibm_status ibm_spread_sets(ibm_ctx* ctx, const ibm_grid* grids, ibm_delta delta,
                           int nsets, const int* n, const double* const* x,
                           const double* const* f, double* const* fe);

/* ibm_spread_sets without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * binning buffers live in a fixed per-device workspace of the library
 * (36.6 MB, held until the process ends) instead of the context's scratch,
 * so at most 262,144 markers in total and grids of at most 2,113,536 bins
 * (one bin per face-aligned cell; a 128 x 129 x 128 cell grid) are taken;
 * larger inputs return IBM_ERR_SIZE before any work is queued.  Every call
 * on one GPU shares that workspace: calls that overlap on one device (other
 * streams, other contexts, other host threads) race and must be ordered by
 * the caller.  The result is bitwise that of ibm_spread_sets. */
ibm_status ibm_spread_sets_noalloc(ibm_ctx* ctx, const ibm_grid* grids, ibm_delta delta,
                                   int nsets, const int* n, const double* const* x,
                                   const double* const* f, double* const* fe);

/* Many independent boxes in one launch per operator (no counterpart as
 * one call in the reference, which spreads per instance with
 * ib::novel::spread, ib/novel.h:197-227): nboxes boxes share one set of
 * component grids (`grids`, the geometry of one box); the markers of box b
 * are x[offsets[b] .. offsets[b+1]) of one SoA array of offsets[nboxes]
 * markers (offsets on the host); fe[c] holds nboxes consecutive box grids
 * of component c.  Every periodic axis must have a multiple of 8 cells.
 * The spread bins every marker into 8 x 8 x 8-cell bricks of its box,
 * accumulates each brick's stencils once in a 64-bit fixed-point LDS tile
 * and merges the tiles: deterministic, no floating-point atomics.  The call
 * synchronises the context stream to upload the offsets. */
ibm_status ibm_spread_boxes(ibm_ctx* ctx, const ibm_grid* grids, ibm_delta delta, int nboxes, const int* offsets,
                            const double* x, const double* f, double* const* fe);
/* The matching interpolation: ul (SoA, offsets[nboxes] markers) from
 * ue[c], nboxes consecutive box grids (ib/interpolate.h:92-115 per box). */
ibm_status ibm_interpolate_boxes(ibm_ctx* ctx, const ibm_grid* grids, ibm_delta delta, int nboxes,
                                 const int* offsets, const double* x, const double* const* ue, double* ul);

/* Interpolates grid velocities to markers:
 *     U_c(m) = sum_g u_c(g) * prod_j phi((g_j - X_j(m)) / h)
 * Replaces ib::interpolate<domain, delta>::operator()(n, x, ue, ul)
 * (ib/interpolate.h:92-115; kernel :49-76).  Stencil indices outside a
 * solid dimension are clamped to the nearest grid point (ib/interpolate.h:18).
 * ul is SoA n x 3. */
ibm_status ibm_interpolate(ibm_ctx* ctx, const ibm_grid* grids, ibm_delta delta,
                           int n, const double* x, const double* const* ue,
                           double* ul);

ibm_status ibm_interpolate_sets(ibm_ctx* ctx, const ibm_grid* grids, ibm_delta delta,
                                int nsets, const int* n, const double* const* x,
                                const double* const* ue, double* const* ul);


/* ------------------------------------------------------------------ */
/* Staggered-grid operators (matrix-free)                               */
/*                                                                      */
/* A = sigma * I2 + tau * L on one grid, where L is fd::laplacian(grid) */
/* (fd/laplacian.h:53-68) and I2 is fd::identity(grid, second_order)    */
/* (fd/identity.h:55-66), both as the reference assembles them with     */
/* Kronecker sums/products, x fastest (fd/combine.h:35-49):             */
/*   Laplacian: (0, 1); corrected identity: (1, 0);                     */
/*   Helmholtz I2 - lambda L of ins::diffusion (ins/diffusion.h:21-35): */
/*   (1, -lambda).                                                      */
/* ------------------------------------------------------------------ */

/* y = alpha * A x + beta * y.  Replaces gemv(alpha, A, x, beta, y) on the
 * assembled CSR operator (linalg/blas.h:186-199, cusparseDcsrmv). */
ibm_status ibm_op_apply(ibm_ctx* ctx, const ibm_grid* grid, double sigma, double tau, double alpha,
                        const double* x, double beta, double* y);
ibm_status ibm_op_apply_f32(ibm_ctx* ctx, const ibm_grid* grid, double sigma, double tau, float alpha,
                            const float* x, float beta, float* y);

/* Gershgorin interval of A (algo/gershgorin.h:14-50); host only. */
ibm_status ibm_gershgorin(const ibm_grid* grid, double sigma, double tau, double* lo, double* hi);

/* The four Horner weights of the degree-4 Chebyshev polynomial over [a, b]
 * (algo/chebyshev.h:120-139, chebw<4>); host only. */
ibm_status ibm_chebyshev_weights(double a, double b, double* w /* [4] */);

/* out = p(A) r, algo::chebyshev(a, b, A)(r) (algo/chebyshev.h:145-169).
 * work holds 2 * size(grid) doubles. */
ibm_status ibm_chebyshev_apply(ibm_ctx* ctx, const ibm_grid* grid, double sigma, double tau, double a, double b,
                               const double* r, double* out, double* work);

/* <x, y> in FP64 with a fixed summation order; returns the host scalar
 * (linalg::dot, linalg/blas.h:140-155). */
ibm_status ibm_dot(ibm_ctx* ctx, int64_t n, const double* x, const double* y, double* out);

/* ------------------------------------------------------------------ */
/* Solvers                                                              */
/* ------------------------------------------------------------------ */

/* Geometric multigrid for A on `grid`, solvers::multigrid (solvers/multigrid.h:
 * 176-212): levels by halving the refinement while every axis has an even
 * number of cells (> 2 periodic, > 4 walled), operators re-discretised on
 * each level, degree-4 Chebyshev smoothing over the upper part of the
 * Gershgorin interval (solvers/mg/chebyshev.h:12-51), 4-point transfers
 * (solvers/mg/interpolation.h:19-152), smoother-only coarsest level.
 * precision = 64: FP64 throughout; precision = 32: FP32 V-cycles inside an
 * FP64 defect correction.  A singular operator (sigma = 0 with only periodic
 * and Neumann axes) is solved on mean-free vectors. */
typedef struct ibm_mg ibm_mg;
ibm_status ibm_mg_create(ibm_ctx* ctx, const ibm_grid* grid, double sigma, double tau, int precision,
                         ibm_mg** mg);
ibm_status ibm_mg_destroy(ibm_mg* mg);
/* Number of levels and (optionally) their grids, finest first. */
ibm_status ibm_mg_levels(const ibm_mg* mg, int* levels, ibm_grid* grids, int max_grids);
/* Block-target autotuning of the z-marching smoothing, restriction and
 * prolongation kernels (no counterpart in the reference, whose kernels run
 * with fixed launch shapes).  While enabled (the default), ibm_mg_create
 * times z chunkings aimed at 256, 512, 1024 and 2048 blocks on every level
 * whose grid the process has not seen and keeps the fastest (if it beats the
 * built-in target by 3%) for the rest of the process; targets change
 * timings only, never values.  clear != 0 forgets the recorded targets.
 * ibm_mg_shape reports a level's launch shape for kind 0 (smoothing),
 * 1 (residual + restriction) or 2 (prolongation + residual): *target = the
 * recorded block target (0: none recorded, the built-in one is used),
 * *wide = 1 for the wide tile (grids of 2^19 points or more), 0 narrow. */
ibm_status ibm_mg_autotune(int enable, int clear);
ibm_status ibm_mg_shape(const ibm_mg* mg, int level, int kind, int* target, int* wide);
/* x = V(b): one V-cycle from a zero guess (solvers/multigrid.h:120-128). */
ibm_status ibm_mg_vcycle(ibm_ctx* ctx, ibm_mg* mg, const double* b, double* x);
/* Solves A x = b: x += V(b - A x) until ||b - A x|| <= rtol ||b|| (FP64
 * norms; the reference's loop at solvers/multigrid.h:88-104 with a relative
 * stopping rule).  Returns IBM_ERR_NOT_CONVERGED after maxit cycles. */
ibm_status ibm_mg_solve(ibm_ctx* ctx, ibm_mg* mg, const double* b, double* x, double rtol, int maxit,
                        int* iters, double* relres);

/* Direct solves of A x = b, A = sigma I~ + tau L, on grids periodic in x
 * and z with power-of-two point counts (4 to 512) and walled in y (2 to
 * 800 points): FFTs in x and z diagonalise the periodic parts, leaving one
 * tridiagonal system along y per Fourier mode (Thomas elimination).  The
 * exact solution of the system solvers::multigrid iterates on
 * (solvers/multigrid.h:176-212); no counterpart as an algorithm in the
 * reference.  sigma = 0 with Neumann walls is the singular Poisson case:
 * the mean-free solution of the mean-free problem, as ibm_mg_solve.  b and
 * x may alias.  ibm_fft_create returns IBM_ERR_BAD_GRID for other grids. */
typedef struct ibm_fft ibm_fft;
ibm_status ibm_fft_create(ibm_ctx* ctx, const ibm_grid* grid, ibm_fft** fft);
ibm_status ibm_fft_destroy(ibm_fft* fft);
ibm_status ibm_fft_solve(ibm_ctx* ctx, ibm_fft* fft, double sigma, double tau, const double* b, double* x);

/* Helmholtz solves A x = b, A = sigma I + tau L (sigma != 0), on nbatch
 * independent copies of one periodic grid (a batch of small boxes, or the
 * boxes of one MAC component): solvers::multigrid's V-cycle (as
 * ibm_mg_create) in FP32 inside FP64 defect correction, every operation one
 * launch over all copies.  b and x hold nbatch consecutive grid vectors; x is
 * the starting guess.  Stops when every copy's ||b - A x|| <= rtol ||b||
 * (host reads of the per-copy norms), IBM_ERR_NOT_CONVERGED after maxit
 * cycles.  No counterpart as one call in the reference (one solver per
 * instance there, solvers/multigrid.h:204-212). */
typedef struct ibm_boxes_mg ibm_boxes_mg;
ibm_status ibm_boxes_mg_create(ibm_ctx* ctx, const ibm_grid* grid, int nbatch, double sigma, double tau,
                               ibm_boxes_mg** out);
ibm_status ibm_boxes_mg_destroy(ibm_boxes_mg* mg);
ibm_status ibm_boxes_mg_solve(ibm_ctx* ctx, ibm_boxes_mg* mg, const double* b, double* x, double rtol, int maxit,
                              int* iters, double* max_relres);

/* Chebyshev-preconditioned CG, solvers::chebpcg (solvers/chebpcg.h:9-23,
 * solvers/pcg.h:11-44): the Helmholtz solver of ins::diffusion.  Stops when
 * ||r|| <= rtol ||b||.  work holds 6 * size(grid) doubles. */
ibm_status ibm_chebpcg_solve(ibm_ctx* ctx, const ibm_grid* grid, double sigma, double tau, const double* b,
                             double* x, double rtol, int maxit, double* work, int* iters, double* relres);

/* ------------------------------------------------------------------ */
/* Navier-Stokes building blocks                                        */
/* ------------------------------------------------------------------ */

/* The grid of `src` with axis `view` re-aligned by 1/2: the destination of a
 * difference or average along `view` (fd::shift::directionally,
 * fd/cell.h:165-184; fd/differential.h:49-52).  Host only. */
ibm_status ibm_grid_shift(const ibm_grid* src, int view, ibm_grid* dst);

/* y = alpha * D x + beta * y with D = fd::differential(src, view)
 * (fd/differential.h:86-103), y on ibm_grid_shift(src, view). */
ibm_status ibm_differential(ibm_ctx* ctx, const ibm_grid* src, int view, double alpha, const double* x,
                            double beta, double* y);
/* y = alpha * A x + beta * y with A = fd::average(src, view) (fd/average.h:85-102). */
ibm_status ibm_average(ibm_ctx* ctx, const ibm_grid* src, int view, double alpha, const double* x, double beta,
                       double* y);

/* out = sum_i D_i u_i, ins::divergence (ins/differential.h:46-83). */
ibm_status ibm_divergence(ibm_ctx* ctx, const ibm_grid* u_grids /* [3] */, const double* const* u,
                          double* out);
/* out_i = alpha * D_i p + beta * out_i, ins::gradient (ins/differential.h:88-112). */
ibm_status ibm_gradient(ibm_ctx* ctx, const ibm_grid* p_grid, double alpha, const double* p, double beta,
                        double* const* out /* [3] */);
/* H^i = sum_j D_j[(A_i u^j)(A_j u^i)], ins::advection (ins/advection.h:58-125). */
ibm_status ibm_advection(ibm_ctx* ctx, const ibm_grid* u_grids /* [3] */, const double* const* u,
                         double* const* H /* [3] */);

/* ---------------------------------------------------------------------
 * RBF surface operators (bases/operators.h, bases/geometry.h)
 * ------------------------------------------------------------------- */
typedef struct ibm_rbf ibm_rbf;

/* Arithmetic of an RBF evaluation. */
enum {
    IBM_RBF_F64 = 64,      /* FP64 matrix cores */
    IBM_RBF_F32 = 32,      /* FP32 operands and accumulation */
    IBM_RBF_BF16X3 = 16    /* split BF16 (hi*hi + hi*lo + lo*hi), FP32 accumulation */
};

/* Prepares nops operators of one cell type for evaluation in `precision`:
 * E[j] is a device matrix of lde x M doubles, column-major, whose first K
 * rows are used: the slice E[0:K, 0:M] that bases::operators holds as its
 * evaluator / derivative matrices, K = data sites, M = sample sites
 * (bases/operators.h:145-177; the LU solve that produces E is setup).
 * The operator keeps its own packed copy; E may be freed afterwards. */
ibm_status ibm_rbf_create(ibm_ctx* ctx, int precision, int M, int K, int nops, const double* const* E, int lde,
                          ibm_rbf** out);
ibm_status ibm_rbf_destroy(ibm_rbf* op);

/* Y_b[j] = E_j^T X_b for every batch entry b and operator j of ops[b], in
 * one launch: `slice * matrix` (bases/operators.h:95-109).  X[b] is
 * K x N[b] column-major (ld = K; column comp * ncells + cell, as
 * bases/container.h:54 lays out data-site coordinates); Y[b] receives
 * nops consecutive M x N[b] column-major blocks.  All entries share one
 * precision. */
ibm_status ibm_rbf_eval(ibm_ctx* ctx, int count, ibm_rbf* const* ops, const int* N, const double* const* X,
                        double* const* Y);

/* Kernel variant of the FP64 evaluation (no counterpart in the reference,
 * whose products go to cuBLAS): waves_n = 2 runs four 64 x 64-output waves
 * per 128 x 128 tile, 4 runs eight 64 x 32-output waves (half the
 * registers, twice the waves per SIMD) over two LDS buffers (one barrier
 * per contraction chunk), 0 (the default) times both on the
 * first eager launch of each shape and keeps the faster for the process.
 * Both variants give bitwise the same results.  Process-wide. */
ibm_status ibm_rbf_gemm_variant(int waves_n);

/* Operator construction on the device (bases/operators.h:145-190,
 * bases/fill.h:15-150, bases/phs.h:13-46, solvers/lu.h:14-48): the r^7
 * polyharmonic spline of the chord distance between the parameter-space
 * sites mapped to the unit sphere, augmented with the (degree + 1)^2
 * monomials of the embedding that span the spherical harmonics of degree
 * <= degree; LU of the (nd + np)^2 interpolation matrix (rocSOLVER getrf)
 * and, for each requested kind, E[k] = solve(lu, [Phi_s; P_s]) as a
 * (nd + np) x ns column-major device matrix (lde = nd + np), ready for
 * ibm_rbf_create.  data and sample are device (theta, phi) arrays, n x 2
 * column-major.  Kinds: the evaluator and the first / second derivatives in
 * the sample parameters (bases/rbf.h:29-48). */
enum {
    IBM_RBF_VALUE = 0, IBM_RBF_D_T = 1, IBM_RBF_D_P = 2,
    IBM_RBF_D_TT = 3, IBM_RBF_D_TP = 4, IBM_RBF_D_PP = 5
};
ibm_status ibm_rbf_build(ibm_ctx* ctx, int nd, const double* data, int ns, const double* sample, int degree,
                         int nkinds, const int* kinds, double* const* E);
/* Parameter-space quadrature weights of n sites on a sphere-like closed
 * surface (bases/closed_surface.h:38-68, bases/shapes/sphere.h:101-115):
 * the r^1 interpolation system with a constant augmentation solved for the
 * right-hand side (0, ..., 0, 1), divided by cos(phi) / (4 pi). */
ibm_status ibm_sphere_weights(ibm_ctx* ctx, int n, const double* params, double* w);

/* Surface geometry of nc cells with ns sample sites, in FP64, from the
 * evaluated positions y and tangents t1, t2 (each ns x 3nc as above) and
 * the parameter-space quadrature weights w (ns): unit normals (ns x 3nc),
 * patch areas sigma = w |t1 x t2| (ns x nc) (bases/geometry.h:95-119), and
 * per cell the area sum(sigma) and volume sum(x . n sigma) / 3. */
ibm_status ibm_surface_geometry(ibm_ctx* ctx, int ns, int nc, const double* y, const double* t1, const double* t2,
                                const double* w, double* normal, double* sigma, double* area, double* volume);

/* y = a x + b y over n doubles (linalg::axpy / scal, linalg/blas.h). */
ibm_status ibm_axpby(ibm_ctx* ctx, long long n, double a, const double* x, double b, double* y);

/* ---------------------------------------------------------------------
 * Membrane forces (include/forces/)
 *
 * Geometry blocks follow bases::geometry: for rows sites x nc cells,
 * component i of point k = cell * rows + site is at [i * P + k],
 * P = rows * nc; a "geometry" argument holds the five derivative blocks
 * x_u, x_v, x_uu, x_uv, x_vv (u = theta, v = phi; the order of
 * bases::operators' first and second derivatives, bases/operators.h:
 * 168-190), 3P apart.  Scalar fields are P-long blocks.
 * ------------------------------------------------------------------- */
enum {
    IBM_TENSION_NONE = 0,
    IBM_TENSION_SKALAK = 1,       /* forces::skalak      (forces/skalak.h:11-36)      */
    IBM_TENSION_NEOHOOKEAN = 2    /* forces::neohookean  (forces/neohookean.h:10-34)  */
};

typedef struct {
    int tension;      /* IBM_TENSION_*                                                 */
    int preferred;    /* forces::bending's `preferred`: bend relative to the reference */
    double shear;     /* tension moduli, energy per area                               */
    double bulk;
    double bending;   /* forces::bending modulus (energy); 0 = no bending              */
} ibm_membrane;

/* Per-site table of the undeformed shape (one cell, ns sites): its metric
 * terms, patch areas w |x_u x x_v| and mean curvature, 14 x ns doubles; the
 * reference geometry every cell's forces compare against
 * (object.geometry(reference) of forces/tension.h:56-60). */
ibm_status ibm_membrane_reference(ibm_ctx* ctx, int ns, const double* geometry, const double* weights,
                                  double* table);
/* Mean curvature at the data sites, h = H (minus h0[site] when h0 is given:
 * the preferred curvature), from the data-site geometry (nd x nc cells),
 * forces/bending.h:142-158. */
ibm_status ibm_mean_curvature(ibm_ctx* ctx, int nd, int nc, const double* geometry, const double* h0, double* h);
/* f = beta f + tension + bending at the sample sites (ns x nc cells):
 * forces/tension.h:50-99 with the model's law, plus
 * -4 kappa s (Lap(H - H0) + 2 (H - H0)(H^2 - K + H H0)) n of
 * forces/bending.h:160-223, whose Laplacian takes hd = the five parameter
 * derivatives (hu, hv, huu, huv, hvv) of h at the sample sites (P-long
 * blocks; unused without bending).  normal and sigma are the current
 * geometry's (ibm_surface_geometry). */
ibm_status ibm_membrane_forces(ibm_ctx* ctx, const ibm_membrane* model, int ns, int nc, const double* geometry,
                               const double* normal, const double* sigma, const double* hd, const double* table,
                               double beta, double* f);

/* ---------------------------------------------------------------------
 * Boundary data and fixed-count solves
 * ------------------------------------------------------------------- */
/* y += alpha * B g with B = fd::lower_boundary (side 0) / fd::upper_boundary
 * (side 1) of grid g along `view` with correction order 0 or 2
 * (fd/boundary_ops.h:21-71): the face row's ghost-cell boundary coefficient
 * times resolution^order, kron the order-corrected identity of the other
 * axes.  `values` holds one value per face point (other axes, x fastest),
 * or is null for the constant `value`.  A periodic `view` adds nothing. */
ibm_status ibm_boundary_apply(ibm_ctx* ctx, const ibm_grid* g, int view, int side, int order, double alpha,
                              const double* values, double value, double* y);

/* y = alpha * B x + beta * y with B = ins::boundary(grid, dim)
 * (ins/boundary.h:16-56): the operator that carries the pseudo-pressure wall
 * condition w = u + k grad(phi) of the intermediate velocity.  B is diagonal:
 * on each walled axis j != dim it holds the ghost-cell boundary coefficient
 * (fd/discretization.h:56) on the axis's first and last rows, kron the
 * zeroth-order identity of the other axes, summed over those axes.  order 0
 * is the reference's matrix; order 2 scales axis j's entries by its
 * resolution^2, the Laplacian's factor.  dim = -1 excludes no axis. */
ibm_status ibm_ins_boundary_apply(ibm_ctx* ctx, const ibm_grid* g, int dim, int order, double alpha,
                                  const double* x, double beta, double* y);

/* ibm_mg_solve stopping at ||b - A x|| <= max(rtol ||b||, atol): for right-hand
 * sides whose norm is itself rounding noise (the divergence of a field that
 * is already divergence-free), atol carries the scale of the problem.  With
 * warm != 0 the iteration starts from the x passed in instead of zero. */
ibm_status ibm_mg_solve_atol(ibm_ctx* ctx, ibm_mg* mg, const double* b, double* x, double rtol, double atol,
                             int maxit, int warm, int* iters, double* relres);

/* ibm_mg_solve with exactly `iters` defect-correction cycles and no host
 * synchronisation (recordable in a HIP graph); warm as for ibm_mg_solve_atol.  When `relres` (a device
 * double) is given, ||b - A x|| / ||b|| is written there. */
ibm_status ibm_mg_solve_fixed(ibm_ctx* ctx, ibm_mg* mg, const double* b, double* x, int iters, int warm,
                              double* relres);

/* ---------------------------------------------------------------------
 * Immersed-boundary timestep (the ib::solver of test/ib-whole.cc:226-275;
 * the reference does not ship its implementation, so the scheme is this
 * library's: geometry by RBF evaluation, tether forces, spread, advection,
 * ins::diffusion-form Helmholtz solves, projection onto divergence-free
 * fields with a Neumann pressure Poisson solve, interpolation at the data
 * sites and X += dt U)
 * ------------------------------------------------------------------- */
#define IBM_MAX_CELL_TYPES 8

typedef struct ibm_solver ibm_solver;

typedef struct {
    double timestep;     /* dt; with fscale > 0 the smallest step k_min (library time units) */
    double diffusivity;  /* nu = mu / rho (ins/simulation.h) */
    double density;      /* rho: spread forces are divided by it */
    double tolerance;    /* relative residual of every Helmholtz / Poisson solve */
    int delta;           /* ibm_delta of spread and interpolate */
    int precision;       /* multigrid V-cycles in 64 or 32 bit; defect correction is FP64 */
    int max_iterations;  /* per solve */
    int flags;           /* IBM_SOLVER_* bits */
    double fscale;       /* > 0: step size from ins::tracker's force bound (ins/tracker.h:62-75) */
} ibm_solver_params;

/* Wall condition of the intermediate velocity u* in the Helmholtz solves:
 * u* = g + dt grad(p_prev) on the walls tangential to each component (the
 * w = u + k grad(phi) of ins/boundary.h:16-18, with the previous step's
 * pressure), added to the right-hand side through ins::boundary.  After the
 * projection the tangential wall slip is then dt |grad(p_prev - p)| instead
 * of dt |grad p|.  Without the flag u* takes the wall data g. */
#define IBM_SOLVER_PSEUDO_PRESSURE 1
/* Solve the Helmholtz and pressure systems by multigrid even where the grids
 * allow the direct FFT x tridiagonal solves (ibm_fft_*: periodic x and z of
 * power-of-two sizes, walls in y), which are used there otherwise. */
#define IBM_SOLVER_MULTIGRID 2

typedef struct {
    ibm_rbf* ops;            /* operators data -> sample: position, d/dtheta, d/dphi, and with a
                                membrane model also d2/dtheta2, d2/dtheta dphi, d2/dphi2 (6) */
    int data_sites, sample_sites, cells;
    int reserved;
    const double* weights;   /* sample-site quadrature weights (device; copied) */
    const double* tether;    /* null, or reference sample positions, ns x 3 cells (device; copied) */
    double stiffness;        /* tether force f = -stiffness (y - tether) */
    ibm_membrane membrane;   /* tension / bending models (all zero: none) */
    ibm_rbf* data_ops;       /* bending: 5 operators data -> data: d/dtheta, d/dphi, d2/dtheta2,
                                d2/dtheta dphi, d2/dphi2 (the data geometry of forces/bending.h:146) */
    const double* reference; /* membrane models: the undeformed shape's data sites, nd x 3 (one cell;
                                device; used at creation) */
} ibm_cell_type;

/* ib::state (test/ib-whole.cc:159,180-181): caller-owned device buffers. */
typedef struct {
    double t;
    double* u[3];                    /* MAC velocity components */
    double* p;                       /* pressure on the cell centres */
    double* x[IBM_MAX_CELL_TYPES];   /* data sites per cell type: nd x 3 cells, column comp * cells + cell */
} ibm_state;

typedef struct {
    int iterations[4];    /* V-cycles of the u, v, w Helmholtz solves and the pressure solve */
    double residual[4];   /* their final relative residuals (FP64) */
    double divergence;    /* max |D u| after projection (FP64) */
    double timestep;      /* step size of the last step */
    int exponent;         /* its tracker exponent e (dt = k_min 2^e), 0 without tracking */
    int reserved;
} ibm_step_info;

/* Device views of a cell type's per-step data (owned by the solver). */
typedef struct {
    const double* positions;     /* ns x 3 cells */
    const double* tangents[2];
    const double* forces;
    const double* normals;
    const double* sigma;         /* ns x cells */
    const double* area;          /* cells */
    const double* volume;        /* cells */
    const double* velocity;      /* interpolated at the data sites: nd x 3 cells */
} ibm_cell_view;

/* Shapes of the objects of an ib::state record for checkpoint / resume:
 * u[0..dims) and p are dense vectors of u_size[i] / p_size values, x[k] is
 * a dense rows[k] x cols[k] matrix (data sites x 3 cells, column
 * comp * cells + cell). */
typedef struct {
    int dims;
    int ntypes;
    long long u_size[3];
    long long p_size;
    int rows[IBM_MAX_CELL_TYPES];
    int cols[IBM_MAX_CELL_TYPES];
} ibm_state_layout;

/* ---------------------------------------------------------------------
 * Checkpoint / resume in the reference's linalg binary format
 * (linalg/io.h:67-99 magic numbers, :416-531 writers and readers): a dense
 * object is a 4-byte magic (vector "ve" 0x65760000 or matrix "ma"
 * 0x616d0000, | dense 0x6400, | double 0x64), the int32 header
 * {rows, cols, rows*cols} and the values (column-major, host byte order).
 * Device data is staged through host memory; the calls synchronise the
 * context stream.  Files written here are read by the reference's
 * `f >> linalg::io::binary >> st` (test/ib-whole.cc:184-187) and by
 * util/kaboose.cc, and files they write are read here.
 * ------------------------------------------------------------------- */

/* Appends (append != 0) or writes one dense double vector (is_matrix = 0,
 * cols ignored) or matrix of device data: linalg::io::write(out, binary, v)
 * (linalg/io.h:421-454). */
ibm_status ibm_io_write_dense(ibm_ctx* ctx, const char* path, int append, int is_matrix, long long rows,
                              long long cols, const double* values);
/* Reads the dense object at byte `offset` into device memory, which must
 * have the expected shape (IBM_ERR_SIZE otherwise; IBM_ERR_ARG for a wrong
 * magic): linalg::io::read(in, binary, v) (linalg/io.h:485-518).  *next
 * receives the offset after the object. */
ibm_status ibm_io_read_dense(ibm_ctx* ctx, const char* path, long long offset, int is_matrix, long long rows,
                             long long cols, double* values, long long* next);
/* Header of the object at `offset`: its magic word and shape. */
ibm_status ibm_io_peek_dense(ibm_ctx* ctx, const char* path, long long offset, int* magic, int* rows, int* cols);

/* Writes (append = 0) or appends one ib::state record: t (8 raw bytes),
 * u[0..dims), p, x[0..ntypes) -- the order of ins::state's writer
 * (ins/state.h:48-68) extended by the cell matrices, as test/ib-whole.cc's
 * binary_writer stores every step and util/kaboose.cc:21-31 reads it. */
ibm_status ibm_state_write(ibm_ctx* ctx, const char* path, int append, const ibm_state* st,
                           const ibm_state_layout* layout);
/* Reads record `record` (0-based; negative counts from the end, so -1 is
 * the last complete record, util/kaboose.cc:53-66) into the state's device
 * buffers and sets st->t.  *records (if given) receives the number of
 * complete records; with st == NULL only the records are counted.  A
 * truncated final record (a run killed while writing) is not counted. */
ibm_status ibm_state_read(ibm_ctx* ctx, const char* path, int record, ibm_state* st,
                          const ibm_state_layout* layout, int* records);

ibm_status ibm_solver_create(ibm_ctx* ctx, const ibm_dimension* domain, int ndims, int refinement,
                             const ibm_solver_params* params, int ntypes, const ibm_cell_type* types,
                             ibm_solver** out);
ibm_status ibm_solver_destroy(ibm_solver* s);
/* One step; V-cycle counts chosen by the residual (host reads). `boundary`
 * is null or 3 device vectors of wall data on the velocity grids (e.g. from
 * ibm_boundary_apply), the `ub` of test/ib-whole.cc:250,275. */
ibm_status ibm_solver_step(ibm_ctx* ctx, ibm_solver* s, ibm_state* st, const double* const* boundary);
/* One step with the V-cycle counts fixed (the most an adaptive step has
 * needed, plus one) and no host synchronisation. */
ibm_status ibm_solver_step_fixed(ibm_ctx* ctx, ibm_solver* s, ibm_state* st, const double* const* boundary);
/* Records ibm_solver_step_fixed on the state's buffers as a HIP graph;
 * ibm_solver_replay launches it (the state must hold the same buffers). */
ibm_status ibm_solver_capture(ibm_ctx* ctx, ibm_solver* s, const ibm_state* st, const double* const* boundary);
ibm_status ibm_solver_replay(ibm_ctx* ctx, ibm_solver* s, ibm_state* st);
ibm_status ibm_solver_info(ibm_ctx* ctx, ibm_solver* s, ibm_step_info* out);
ibm_status ibm_solver_cells(ibm_solver* s, int type, ibm_cell_view* out);
ibm_status ibm_solver_grids(ibm_solver* s, ibm_grid* velocity /* [3] */, ibm_grid* pressure);
/* Shapes of this solver's states, for ibm_state_write / ibm_state_read. */
ibm_status ibm_solver_layout(ibm_solver* s, ibm_state_layout* out);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* IBM_H */
