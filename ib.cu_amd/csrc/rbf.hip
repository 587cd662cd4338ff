// RBF surface evaluation on the matrix cores: Y = E^T X for every cell of a
// cell type at once, and the per-sample surface geometry derived from it.
//
// Restates:
//   bases::impl::slice * matrix (bases/operators.h:95-109): a DGEMM
//     Y[ns x 3nc] = E[0:nd, 0:ns]^T X[nd x 3nc], E being the (nd+np) x ns
//     LU-solved evaluator (or derivative) matrix, X the data-site
//     coordinates with column comp*nc + cell (bases/container.h:54).
//   bases::geometry::compute (bases/geometry.h:74-123): unit normal
//     n = t1 x t2 / |t1 x t2| and patch area sigma = w |t1 x t2|, plus the
//     cell area sum(sigma) and enclosed volume sum(x . n sigma) / 3.
//
// Both operands of E^T X are contiguous along the contraction index, so
// each is packed once into "rows x Kp" tiles (k fastest, zero padded):
// the operator at creation, X per call.  One launch covers every
// (cell type, operator) pair.  Precisions:
//   64  v_mfma_f64_16x16x4 on FP64 operands (the plain path);
//   32  v_mfma_f32_16x16x4 on FP32 operands, FP32 accumulation;
//   16  split BF16: a = a_hi + a_lo with a_hi = bf16(a), a_lo = bf16(a -
//       a_hi); a b ~ a_hi b_hi + a_hi b_lo + a_lo b_hi with
//       v_mfma_f32_16x16x32_bf16 and FP32 accumulation.
// The block tile is 128 (samples) x 128 (columns) with the contraction
// staged through LDS in 128-byte rows (144-byte stride, conflict-free for
// the 16-byte fragment reads), four waves each owning a 64 x 64 quadrant.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>
#include "context.hpp"

namespace ibm {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

constexpr int TILE = 128;        // rows of a block tile (samples and columns)
constexpr int ROWB = 128;        // bytes of one LDS row per contraction chunk
constexpr int NTHREADS = 256;
constexpr int KPAD = 64;         // contraction padded to a multiple of 64 elements
constexpr int MAX_GROUPS = RBF_MAX_GROUPS;

enum { MODE_F64 = 0, MODE_F32 = 1, MODE_BF16X3 = 2 };

__host__ __device__ constexpr int elem_bytes(int mode) { return mode == MODE_F64 ? 8 : mode == MODE_F32 ? 4 : 2; }
__host__ __device__ constexpr int planes(int mode) { return mode == MODE_BF16X3 ? 2 : 1; }

struct gemm_group {
    const unsigned char* E[2];   // operator planes: rows_pad x Kp elements
    const unsigned char* X[2];   // data-site planes: cols_pad x Kp elements
    double* Y;                   // M x N, column-major, ld = M
    int M, N, Kp, tiles_n, first_tile;
};

struct gemm_groups {
    gemm_group g[MAX_GROUPS];
    int count, total_tiles;
};

__device__ inline uint16_t bf16_bits(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);     // round to nearest even (finite inputs)
    return (uint16_t) (u >> 16);
}

__device__ inline float bf16_value(uint16_t b) { return __uint_as_float(((uint32_t) b) << 16); }

// Packs src (column-major, element (k, r) at src[k + ld r]) into dst planes
// laid out as rows_pad x Kp with k fastest, zero outside rows x K.
template <int MODE>
__global__ __launch_bounds__(256) void pack_kernel(const double* __restrict__ src, long ld, int rows, int K,
                                                   int rows_pad, int Kp, unsigned char* __restrict__ hi,
                                                   unsigned char* __restrict__ lo)
{
    long total = (long) rows_pad * Kp;
    for (long i = blockIdx.x * (long) blockDim.x + threadIdx.x; i < total; i += (long) gridDim.x * blockDim.x) {
        // 32-bit quotient where the index fits (a 64-bit division is a long
        // instruction sequence)
        int r, k;
        if (total <= 0xffffffffl) {
            const unsigned u = (unsigned) i;
            r = (int) (u / (unsigned) Kp);
            k = (int) (u - (unsigned) r * (unsigned) Kp);
        } else {
            r = (int) (i / Kp);
            k = (int) (i - (long) r * Kp);
        }
        double v = (r < rows && k < K) ? src[k + ld * r] : 0.0;
        if constexpr (MODE == MODE_F64) {
            ((double*) hi)[i] = v;
        } else if constexpr (MODE == MODE_F32) {
            ((float*) hi)[i] = (float) v;
        } else {
            uint16_t h = bf16_bits((float) v);
            double rest = v - (double) bf16_value(h);
            ((uint16_t*) hi)[i] = h;
            ((uint16_t*) lo)[i] = bf16_bits((float) rest);
        }
    }
}

template <int MODE>
__device__ inline void mfma_step(const u4& xa, const u4& eb, const u4& xa_lo, const u4& eb_lo,
                                 typename std::conditional<MODE == MODE_F64, d4, f4>::type& acc)
{
    if constexpr (MODE == MODE_F64) {
        double a[2], b[2];
        __builtin_memcpy(a, &xa, 16);
        __builtin_memcpy(b, &eb, 16);
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0], b[0], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[1], b[1], acc, 0, 0, 0);
    } else if constexpr (MODE == MODE_F32) {
        float a[4], b[4];
        __builtin_memcpy(a, &xa, 16);
        __builtin_memcpy(b, &eb, 16);
#pragma unroll
        for (int e = 0; e < 4; ++e) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[e], b[e], acc, 0, 0, 0);
    } else {
        bf8 ah, bh, al, bl;
        __builtin_memcpy(&ah, &xa, 16);
        __builtin_memcpy(&bh, &eb, 16);
        __builtin_memcpy(&al, &xa_lo, 16);
        __builtin_memcpy(&bl, &eb_lo, 16);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
    }
}

// One 128 x 128 tile of Y per block.  The MFMA "A" operand is the X tile
// (output rows = columns of Y) and "B" the operator tile, so a lane's four
// results are four columns of Y at one sample row and the 16 lanes of a
// group store 16 consecutive samples (128 contiguous bytes).
//
// Fragment convention (16 x 16 x 4 and 16 x 16 x 32 alike): lane l supplies
// row l % 16 of its operand and the contraction slice owned by lane group
// l / 16; both operands use the same slice assignment, so the sum over the
// chunk is complete whatever order the hardware pairs the slices in.
// Results: lane l holds four rows of column l % 16: rows 4 (l / 16) + r,
// r = 0..3, for the FP32-accumulating instructions, but rows (l / 16) + 4 r
// for v_mfma_f64_16x16x4 (both measured on gfx950).
template <int MODE, int WN, int RB = ROWB, int NB = 1>
__global__ __launch_bounds__(128 * WN, 2) void rbf_gemm_kernel(gemm_groups G)
{
    constexpr int PL = planes(MODE);
    constexpr int THREADS = 128 * WN;                // waves: 2 along the samples x WN along the columns
    constexpr int JN = 8 / WN;                       // 16-column blocks per wave
    constexpr int RS = RB + 16;                      // LDS row stride: 16-row fragment reads hit disjoint banks
    constexpr int LOADS = TILE * RB / 16 / THREADS;  // 16-byte staging loads per thread, tile and plane
    using acc_t = typename std::conditional<MODE == MODE_F64, d4, f4>::type;
    __shared__ __attribute__((aligned(16))) unsigned char ldsb[NB][PL][2][TILE * RS];

    // Blocks reach the XCDs round-robin (block i on XCD i % 8).  XCD x takes
    // the x-th eighth of every group's tiles, in order, so that the XCDs get
    // equal work however the groups' contraction lengths differ, and the
    // tiles sharing an operator row block share an L2.
    const int xcd = blockIdx.x & 7;
    int k = blockIdx.x >> 3, gi = -1, t = 0;
    for (int q = 0; q < G.count; ++q) {
        const int tq = (q + 1 < G.count ? G.g[q + 1].first_tile : G.total_tiles) - G.g[q].first_tile;
        const int lo = xcd * tq / 8, hi = (xcd + 1) * tq / 8;
        if (k < hi - lo) {
            gi = q;
            t = lo + k;
            break;
        }
        k -= hi - lo;
    }
    if (gi < 0) return;                              // whole block exits together
    const gemm_group g = G.g[gi];
    int tm = t / g.tiles_n, tn = t - tm * g.tiles_n;
    const long row_bytes = (long) g.Kp * elem_bytes(MODE);
    const int nchunks = (int) (row_bytes / RB);

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    // Global -> register staging: LOADS x 16 bytes per thread per tile and plane.
    const unsigned char* srcE[PL];
    const unsigned char* srcX[PL];
#pragma unroll
    for (int p = 0; p < PL; ++p) {
        srcE[p] = g.E[p] + (long) tm * TILE * row_bytes;
        srcX[p] = g.X[p] + (long) tn * TILE * row_bytes;
    }
    u4 pe[PL][LOADS], px[PL][LOADS];
    long goff[LOADS];
    int loff[LOADS];
#pragma unroll
    for (int i = 0; i < LOADS; ++i) {
        int idx = tid + THREADS * i;
        int r = idx / (RB / 16), q = idx % (RB / 16);
        goff[i] = (long) r * row_bytes + q * 16;
        loff[i] = r * RS + q * 16;
    }
#define RBF_LOAD(c)                                                                   \
    _Pragma("unroll") for (int p = 0; p < PL; ++p)                                    \
        _Pragma("unroll") for (int i = 0; i < LOADS; ++i) {                           \
            pe[p][i] = *(const u4*) (srcE[p] + goff[i] + (long) (c) * RB);         \
            px[p][i] = *(const u4*) (srcX[p] + goff[i] + (long) (c) * RB);         \
        }

    acc_t acc[JN][4];
#pragma unroll
    for (int j = 0; j < JN; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[j][i] = acc_t{};

    const int frow = lane & 15, fgrp = lane >> 4;
#define RBF_STORE(buf)                                                                \
    _Pragma("unroll") for (int p = 0; p < PL; ++p)                                    \
        _Pragma("unroll") for (int i = 0; i < LOADS; ++i) {                           \
            *(u4*) (&ldsb[buf][p][0][loff[i]]) = pe[p][i];                            \
            *(u4*) (&ldsb[buf][p][1][loff[i]]) = px[p][i];                            \
        }
    RBF_LOAD(0);
    if constexpr (NB == 2) {
        // Two LDS buffers, one barrier per chunk: chunk c + 1 is stored into
        // the buffer chunk c - 1 used (every wave left it at the last
        // barrier) while chunk c is multiplied from the other.
        RBF_STORE(0);
        if (nchunks > 1) {
            RBF_LOAD(1);
        }
        __syncthreads();
    }
    for (int c = 0; c < nchunks; ++c) {
        const int buf = NB == 2 ? (c & 1) : 0;
        if constexpr (NB == 2) {
            if (c + 1 < nchunks) {
                RBF_STORE(buf ^ 1);
            }
            if (c + 2 < nchunks) {
                RBF_LOAD(c + 2);
            }
        } else {
            __syncthreads();
            RBF_STORE(0);
            __syncthreads();
            if (c + 1 < nchunks) {
                RBF_LOAD(c + 1);
            }
        }
        const unsigned char (*lds)[2][TILE * RS] = ldsb[buf];
#pragma unroll
        for (int s = 0; s < RB / 64; ++s) {
            u4 ef[PL][4], xf[PL][JN];
            const int off = s * 64 + fgrp * 16;
#pragma unroll
            for (int p = 0; p < PL; ++p) {
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    ef[p][i] = *(const u4*) (&lds[p][0][(wm * 64 + i * 16 + frow) * RS + off]);
#pragma unroll
                for (int j = 0; j < JN; ++j)
                    xf[p][j] = *(const u4*) (&lds[p][1][(wn * 16 * JN + j * 16 + frow) * RS + off]);
            }
#pragma unroll
            for (int j = 0; j < JN; ++j)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    mfma_step<MODE>(xf[0][j], ef[0][i], xf[PL - 1][j], ef[PL - 1][i], acc[j][i]);
        }
        if constexpr (NB == 2) __syncthreads();
    }
#undef RBF_LOAD
#undef RBF_STORE

    const int m_base = tm * TILE + wm * 64 + frow;
    const int n_base = tn * TILE + wn * 16 * JN + (MODE == MODE_F64 ? fgrp : fgrp * 4);
    constexpr int RSTEP = MODE == MODE_F64 ? 4 : 1;
#pragma unroll
    for (int j = 0; j < JN; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int m = m_base + i * 16;
            if (m >= g.M) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int n = n_base + j * 16 + r * RSTEP;
                if (n < g.N) g.Y[m + (long) g.M * n] = (double) acc[j][i][r];
            }
        }
}

// Surface geometry of one cell per block from the evaluated position and
// the two tangent blocks (each ns x 3nc, column comp*nc + cell).
__global__ __launch_bounds__(256) void geometry_kernel(int ns, int nc, const double* __restrict__ y,
                                                       const double* __restrict__ t1, const double* __restrict__ t2,
                                                       const double* __restrict__ w, double* __restrict__ normal,
                                                       double* __restrict__ sigma, double* __restrict__ area,
                                                       double* __restrict__ volume)
{
    __shared__ double red[2][256];
    const int cell = blockIdx.x;
    const long count = (long) ns * nc;
    double a_sum = 0, v_sum = 0;
    for (int s = threadIdx.x; s < ns; s += blockDim.x) {
        long at = (long) cell * ns + s;
        double a[3], b[3], x[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            a[i] = t1[count * i + at];
            b[i] = t2[count * i + at];
            x[i] = y[count * i + at];
        }
        double n0 = a[1] * b[2] - a[2] * b[1];
        double n1 = a[2] * b[0] - a[0] * b[2];
        double n2 = a[0] * b[1] - a[1] * b[0];
        double det = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
        normal[count * 0 + at] = n0 / det;
        normal[count * 1 + at] = n1 / det;
        normal[count * 2 + at] = n2 / det;
        double sg = w[s] * det;
        sigma[at] = sg;
        a_sum += sg;
        v_sum += w[s] * (x[0] * n0 + x[1] * n1 + x[2] * n2);
    }
    red[0][threadIdx.x] = a_sum;
    red[1][threadIdx.x] = v_sum;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            red[0][threadIdx.x] += red[0][threadIdx.x + h];
            red[1][threadIdx.x] += red[1][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        area[cell] = red[0][0];
        volume[cell] = red[1][0] / 3.0;
    }
}

int mode_of(int precision)
{
    return precision == 64 ? MODE_F64 : precision == 32 ? MODE_F32 : precision == 16 ? MODE_BF16X3 : -1;
}

int round_up(int a, int b) { return (a + b - 1) / b * b; }

template <int MODE>
void launch_pack(hipStream_t s, const double* src, long ld, int rows, int K, int rows_pad, int Kp,
                 unsigned char* hi, unsigned char* lo)
{
    long total = (long) rows_pad * Kp;
    int blocks = (int) std::min<long>((total + 255) / 256, 8192);
    pack_kernel<MODE><<<blocks, 256, 0, s>>>(src, ld, rows, K, rows_pad, Kp, hi, lo);
}

void pack(int mode, hipStream_t s, const double* src, long ld, int rows, int K, int rows_pad, int Kp,
          unsigned char* hi, unsigned char* lo)
{
    if (mode == MODE_F64) launch_pack<MODE_F64>(s, src, ld, rows, K, rows_pad, Kp, hi, lo);
    else if (mode == MODE_F32) launch_pack<MODE_F32>(s, src, ld, rows, K, rows_pad, Kp, hi, lo);
    else launch_pack<MODE_BF16X3>(s, src, ld, rows, K, rows_pad, Kp, hi, lo);
}

} // namespace
} // namespace ibm

using namespace ibm;

struct ibm_rbf {
    int device;
    int mode, M, K, Kp, rows_pad, nops;
    size_t plane_bytes;                  // bytes of one operator plane
    unsigned char* mem;                  // nops x planes(mode) planes
    const unsigned char* plane(int op, int p) const { return mem + (size_t) (op * planes(mode) + p) * plane_bytes; }
};

namespace ibm {

int rbf_nops(const ibm_rbf* op) { return op ? op->nops : 0; }

// FP64 GEMM variants, by waves along the columns: 2 (64 x 64 wave tiles,
// about 250 registers, two waves per SIMD) or 4 (64 x 32, at most 128
// registers, four waves per SIMD, and two LDS buffers so that a chunk takes
// one barrier: 57 against 51 TFLOP/s at the C4 RBC shape; single-buffered it
// reached 55, with 256-byte chunks 52).
// Both give every output the same MFMA sequence over the same chunks, so
// their results are bitwise equal.  Unless forced, the first eager launch of
// a shape (tiles, summed contraction length) times both and keeps the faster
// for the process; launches under graph capture use the recorded choice or
// the 4-wave default.
std::atomic<int> forced_waves_n{0};
std::mutex variant_lock;
std::map<std::pair<int, long>, int> variant_table;
constexpr int DEFAULT_WAVES_N = 4;

void launch_f64(int wn, const gemm_groups& G, int blocks, hipStream_t stream)
{
    if (wn == 4) rbf_gemm_kernel<MODE_F64, 4, ROWB, 2><<<blocks, 512, 0, stream>>>(G);
    else rbf_gemm_kernel<MODE_F64, 2><<<blocks, NTHREADS, 0, stream>>>(G);
}

int f64_variant(ibm_ctx* ctx, hipStream_t stream, const gemm_groups& G, int blocks)
{
    (void) ctx;
    const int forced = forced_waves_n.load();
    if (forced) return forced;
    long ksum = 0;
    for (int i = 0; i < G.count; ++i) ksum += G.g[i].Kp;
    const std::pair<int, long> key{G.total_tiles, ksum};
    {
        std::lock_guard<std::mutex> guard(variant_lock);
        auto it = variant_table.find(key);
        if (it != variant_table.end()) return it->second;
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return DEFAULT_WAVES_N;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) hipEventDestroy(e0);
        return DEFAULT_WAVES_N;
    }
    float ms[2] = {0, 0};
    bool ok = true;
    const int cand[2] = {2, 4};
    for (int c = 0; c < 2 && ok; ++c) {
        launch_f64(cand[c], G, blocks, stream);   // warm-up; every launch writes the same outputs
        ok = hipEventRecord(e0, stream) == hipSuccess;
        for (int r = 0; r < 2 && ok; ++r) launch_f64(cand[c], G, blocks, stream);
        ok = ok && hipGetLastError() == hipSuccess && hipEventRecord(e1, stream) == hipSuccess &&
             hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms[c], e0, e1) == hipSuccess;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    const int pick = !ok ? DEFAULT_WAVES_N : (ms[1] <= ms[0] ? 4 : 2);
    if (ok) {
        std::lock_guard<std::mutex> guard(variant_lock);
        variant_table[key] = pick;
    }
    return pick;
}

ibm_status rbf_eval_range(ibm_ctx* ctx, hipStream_t stream, int count, ibm_rbf* const* ops, const int* N,
                          const double* const* X, double* const* Y, int j0, int j1, bool do_pack,
                          unsigned char** packed)
{
    if (!ctx) return IBM_ERR_ARG;
    if (count <= 0 || !ops || !N || !X || !Y) return fail(ctx, IBM_ERR_ARG, "empty or null batch");
    int mode = ops[0] ? ops[0]->mode : -1;
    int groups = 0;
    size_t xbytes = 0;
    for (int b = 0; b < count; ++b) {
        if (!ops[b] || !X[b] || !Y[b] || N[b] < 0) return fail(ctx, IBM_ERR_ARG, "bad batch entry %d", b);
        if (ops[b]->mode != mode) return fail(ctx, IBM_ERR_ARG, "batch mixes precisions");
        if (ops[b]->device != ctx->device) return fail(ctx, IBM_ERR_ARG, "operator %d lives on another device", b);
        groups += std::max(0, std::min(j1, ops[b]->nops) - j0);
        xbytes += (size_t) round_up(std::max(N[b], 1), TILE) * ops[b]->Kp * elem_bytes(mode) * planes(mode);
    }
    if (groups > MAX_GROUPS) return fail(ctx, IBM_ERR_SIZE, "at most %d operator blocks per launch", MAX_GROUPS);
    if (do_pack) {
        void* scratch_mem;
        IBM_TRY(scratch(ctx, SLOT_RBF_X, xbytes, &scratch_mem));
        *packed = (unsigned char*) scratch_mem;
    } else if (!*packed) {
        return fail(ctx, IBM_ERR_ARG, "no packed operand");
    }
    unsigned char* xp = *packed;

    gemm_groups G{};
    int tiles = 0;
    for (int b = 0; b < count; ++b) {
        const ibm_rbf& op = *ops[b];
        int cols_pad = round_up(std::max(N[b], 1), TILE);
        size_t plane = (size_t) cols_pad * op.Kp * elem_bytes(mode);
        unsigned char* hi = xp;
        unsigned char* lo = planes(mode) > 1 ? xp + plane : nullptr;
        xp += plane * planes(mode);
        if (N[b] == 0) continue;
        if (do_pack) pack(mode, stream, X[b], op.K, N[b], op.K, cols_pad, op.Kp, hi, lo);
        for (int j = j0; j < std::min(j1, op.nops); ++j) {
            gemm_group& g = G.g[G.count++];
            g.E[0] = op.plane(j, 0);
            g.E[1] = planes(mode) > 1 ? op.plane(j, 1) : g.E[0];
            g.X[0] = hi;
            g.X[1] = lo ? lo : hi;
            g.Y = Y[b] + (size_t) j * op.M * N[b];
            g.M = op.M;
            g.N = N[b];
            g.Kp = op.Kp;
            g.tiles_n = cols_pad / TILE;
            g.first_tile = tiles;
            tiles += (op.rows_pad / TILE) * g.tiles_n;
        }
    }
    IBM_LAUNCH_CHECK(ctx);
    if (!tiles) return IBM_OK;
    // Longest contraction first: each XCD takes its share of the groups in
    // group order, so the short tiles come last and fill the tail.
    std::stable_sort(G.g, G.g + G.count, [](const gemm_group& a, const gemm_group& b) { return a.Kp > b.Kp; });
    tiles = 0;
    for (int q = 0; q < G.count; ++q) {
        G.g[q].first_tile = tiles;
        tiles += (round_up(G.g[q].M, TILE) / TILE) * G.g[q].tiles_n;
    }
    G.total_tiles = tiles;
    // blocks: 8 x the most tiles any XCD takes (its share of every group)
    int per_xcd = 0;
    for (int x = 0; x < 8; ++x) {
        int n = 0;
        for (int q = 0; q < G.count; ++q) {
            const int tq = (q + 1 < G.count ? G.g[q + 1].first_tile : tiles) - G.g[q].first_tile;
            n += (x + 1) * tq / 8 - x * tq / 8;
        }
        per_xcd = std::max(per_xcd, n);
    }
    const int blocks = 8 * per_xcd;
    double flops = 0;
    for (int b = 0; b < count; ++b)
        flops += 2.0 * ops[b]->M * (double) N[b] * ops[b]->K * std::max(0, std::min(j1, ops[b]->nops) - j0);
    const int wn = mode == MODE_F64 ? f64_variant(ctx, stream, G, blocks) : 2;
    prof_scope prof(ctx, "rbf_gemm", flops);
    if (mode == MODE_F64) {
        launch_f64(wn, G, blocks, stream);
    } else if (mode == MODE_F32) {
        rbf_gemm_kernel<MODE_F32, 2><<<blocks, NTHREADS, 0, stream>>>(G);
    } else {
        rbf_gemm_kernel<MODE_BF16X3, 2><<<blocks, NTHREADS, 0, stream>>>(G);
    }
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

} // namespace ibm

extern "C" {

ibm_status ibm_rbf_create(ibm_ctx* ctx, int precision, int M, int K, int nops, const double* const* E, int lde,
                          ibm_rbf** out)
{
    if (!ctx || !out) return IBM_ERR_ARG;
    *out = nullptr;
    int mode = mode_of(precision);
    if (mode < 0) return fail(ctx, IBM_ERR_ARG, "precision must be 64, 32 or 16 (split bf16), got %d", precision);
    if (M <= 0 || K <= 0 || nops <= 0 || nops > 8 || lde < K || !E)
        return fail(ctx, IBM_ERR_ARG, "bad operator shape M=%d K=%d nops=%d lde=%d", M, K, nops, lde);
    for (int j = 0; j < nops; ++j)
        if (!E[j]) return fail(ctx, IBM_ERR_ARG, "null operator %d", j);
    ibm_rbf* op = new ibm_rbf{};
    op->device = ctx->device;
    op->mode = mode;
    op->M = M;
    op->K = K;
    op->Kp = round_up(K, KPAD);
    op->rows_pad = round_up(M, TILE);
    op->nops = nops;
    op->plane_bytes = (size_t) op->rows_pad * op->Kp * elem_bytes(mode);
    size_t bytes = op->plane_bytes * nops * planes(mode);
    if (hipMalloc(&op->mem, bytes) != hipSuccess) {
        delete op;
        return fail(ctx, IBM_ERR_HIP, "out of device memory for a %zu-byte operator", bytes);
    }
    for (int j = 0; j < nops; ++j)
        pack(mode, ctx->stream, E[j], lde, M, K, op->rows_pad, op->Kp, (unsigned char*) op->plane(j, 0),
             planes(mode) > 1 ? (unsigned char*) op->plane(j, 1) : nullptr);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess) {
        hipFree(op->mem);
        delete op;
        return fail(ctx, IBM_ERR_HIP, "operator packing failed");
    }
    *out = op;
    return IBM_OK;
}

ibm_status ibm_rbf_gemm_variant(int waves_n)
{
    if (waves_n != 0 && waves_n != 2 && waves_n != 4) return IBM_ERR_ARG;
    ibm::forced_waves_n.store(waves_n);
    return IBM_OK;
}

ibm_status ibm_rbf_destroy(ibm_rbf* op)
{
    if (!op) return IBM_OK;
    hipFree(op->mem);
    delete op;
    return IBM_OK;
}

ibm_status ibm_rbf_eval(ibm_ctx* ctx, int count, ibm_rbf* const* ops, const int* N, const double* const* X,
                        double* const* Y)
{
    if (!ctx) return IBM_ERR_ARG;
    unsigned char* packed = nullptr;
    return ibm::rbf_eval_range(ctx, ctx->stream, count, ops, N, X, Y, 0, 8, true, &packed);
}

ibm_status ibm_surface_geometry(ibm_ctx* ctx, int ns, int nc, const double* y, const double* t1, const double* t2,
                                const double* w, double* normal, double* sigma, double* area, double* volume)
{
    if (!ctx) return IBM_ERR_ARG;
    if (ns <= 0 || nc < 0) return fail(ctx, IBM_ERR_ARG, "bad sizes ns=%d nc=%d", ns, nc);
    if (!nc) return IBM_OK;
    if (!y || !t1 || !t2 || !w || !normal || !sigma || !area || !volume) return fail(ctx, IBM_ERR_ARG, "null array");
    // y, t1, t2 and w read; normals, sigma, area, volume written
    prof_scope prof(ctx, "geometry", 8.0 * ((double) ns * nc * 13 + ns + 2.0 * nc));
    geometry_kernel<<<nc, 256, 0, ctx->stream>>>(ns, nc, y, t1, t2, w, normal, sigma, area, volume);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

} // extern "C"
