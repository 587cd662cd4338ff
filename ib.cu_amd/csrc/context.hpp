// Library context: device, stream, scratch workspaces and error reporting.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include <cstdio>
#include <cstdarg>
#include "ibm.h"

struct ibm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    std::string error = "";

    // Named scratch buffers, grown on demand and kept until ibm_destroy.
    struct buffer { void* ptr = nullptr; size_t bytes = 0; };
    std::vector<buffer> scratch;

    // Pinned host word used to read device status flags back.
    int* host_flag = nullptr;

    // Optional per-kernel timing: hipEvent pairs recorded around named
    // launches on ctx->stream while profiling is enabled.
    struct prof_record { const char* name; hipEvent_t start, stop; double work; };
    bool profiling = false;
    std::vector<prof_record> prof;
    size_t prof_used = 0;
};

namespace ibm {

enum scratch_slot {
    SLOT_BIN_COUNT = 0,
    SLOT_BIN_START,
    SLOT_KEY,
    SLOT_RANK,
    SLOT_TMP,
    SLOT_PERM,
    SLOT_CODE,
    SLOT_SORTED,
    SLOT_SCAN_TMP,
    SLOT_FLAGS,
    SLOT_REDUCE,        // reduction partials of the solvers (REDUCE_PARTS + 8 doubles)
    SLOT_RBF_X,         // packed data-site coordinates of an RBF evaluation
    SLOT_TILES,         // fixed-point tiles of the tile spread (3 x 12^3 per brick)
    SLOT_TILE_FLAGS,    // per-brick "has markers" flags and the tile scale
    SLOT_BOX_OFF,       // marker offsets of the boxes of a batched spread / interpolation
    SLOT_STEP_BASE,     // timestep workspaces; slots from here on belong to the timestep solver
    SLOT_COUNT = 64
};

ibm_status fail(ibm_ctx* ctx, ibm_status code, const char* fmt, ...);
ibm_status hip_fail(ibm_ctx* ctx, hipError_t e, const char* what);
// Returns a device buffer of at least `bytes` for `slot`, reallocating when
// the request grows.  Contents are not preserved across growth.
ibm_status scratch(ibm_ctx* ctx, int slot, size_t bytes, void** out);
// Looks up, once per device, the device address of the fixed workspace of
// ibm_spread_sets_noalloc; `device` must be current.
ibm_status static_scratch_init(ibm_ctx* ctx, int device);

// Brackets a kernel launch with timing events when profiling is on.
struct prof_scope {
    ibm_ctx* ctx;
    int index = -1;
    // `work`: the launch's algorithmic bytes (or flops for matrix-core
    // kernels), summed by ibm_profile_work.
    prof_scope(ibm_ctx* c, const char* name, double work = 0);
    ~prof_scope();
};

// ibm_rbf_eval restricted to operators [j0, j1) of every batch entry and
// launched on `stream`.  With do_pack the data-site coordinates are packed
// into ctx's SLOT_RBF_X scratch first and *packed receives its address;
// without, *packed must hold an earlier packing of the same batch that no
// later evaluation has overwritten.
ibm_status rbf_eval_range(ibm_ctx* ctx, hipStream_t stream, int count, ibm_rbf* const* ops, const int* N,
                          const double* const* X, double* const* Y, int j0, int j1, bool do_pack,
                          unsigned char** packed);

// Membrane force passes (forces.hip) on `stream`.
ibm_status membrane_table(ibm_ctx* ctx, hipStream_t stream, int ns, const double* geo, const double* w, double* table);
ibm_status mean_curvature(ibm_ctx* ctx, hipStream_t stream, int nd, int nc, const double* geo, const double* h0,
                          double* h);
ibm_status membrane_forces(ibm_ctx* ctx, hipStream_t stream, const ibm_membrane* m, double tether, int ns, int nc,
                           const double* geo, const double* normal, const double* sigma, const double* hd,
                           const double* table, const double* y, const double* yref, double beta, double* f);
// Number of operators an RBF object holds, and the most (entry, operator)
// blocks one rbf_eval_range launch takes.
int rbf_nops(const ibm_rbf* op);
constexpr int RBF_MAX_GROUPS = 24;

// Fixed-count multigrid solves of several independent systems, solve k on
// ctxs[k] (ibm_mg_solve_fixed each), issued cycle by cycle in turn; with
// fork_ev (and count - 1 join events) the streams join into ctxs[0]'s and
// fork again after every cycle.
ibm_status mg_solve_fixed_multi(int count, ibm_ctx* const* ctxs, ibm_mg* const* ms, const double* const* b,
                                double* const* x, const int* iters, bool warm, double* const* relres,
                                hipEvent_t fork_ev = nullptr, const hipEvent_t* join_ev = nullptr);

// The direct channel solver (fft_solve.hip): whether a grid qualifies, and
// the solve on ctx's stream.
bool fft_usable(const ibm_grid& g);
ibm_status fft_solve(ibm_ctx* ctx, ibm_fft* p, double sigma, double tau, const double* b, double* x);

// Whether the tile spread takes these MAC grids, and the timestep's spread
// of nsets marker sets by the tile (tiles = true) or the brick kernel.
bool tiles_usable(ibm_ctx* ctx, const ibm_grid* grids);
ibm_status spread_sets_step(ibm_ctx* ctx, const ibm_grid* grids, int delta, bool tiles, int nsets, const int* n,
                            const double* const* x, const double* const* f, double* const* fe);

} // namespace ibm

#define IBM_HIP(ctx, ...)                                                   \
    do {                                                                    \
        hipError_t e_ = (__VA_ARGS__);                                      \
        if (e_ != hipSuccess) return ibm::hip_fail((ctx), e_, #__VA_ARGS__); \
    } while (0)

#define IBM_TRY(...)                                                        \
    do {                                                                    \
        ibm_status s_ = (__VA_ARGS__);                                      \
        if (s_ != IBM_OK) return s_;                                        \
    } while (0)

#define IBM_LAUNCH_CHECK(ctx)                                               \
    do {                                                                    \
        hipError_t e_ = hipGetLastError();                                  \
        if (e_ != hipSuccess) return ibm::hip_fail((ctx), e_, "kernel launch"); \
    } while (0)
