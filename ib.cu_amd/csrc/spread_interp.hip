// Spreading and interpolation between Lagrangian markers and the MAC grid.
//
// Reference algorithm (ib/novel.h:58-168, ib/interpolate.h:49-76): per
// component, sort markers by cell, segmented-sum 8 stencil values per sweep
// into 8 plane buffers, combine the planes; interpolation is a per-marker
// 64-point gather.  The values computed are the same:
//     spread:  fe_c(g) = h^-3 sum_m f_c(m) prod_j phi(g_j - X_j(m)/h)
//     interp:  U_c(m)  = sum_g u_c(g) prod_j phi(g_j - X_j(m)/h)
// with the reference's index conventions (geometry.hpp).
//
// MI355X design:
//   * markers are counting-sorted once into grid bins (shared by the three
//     components), canonically ordered inside a bin, so every later pass is
//     independent of scheduling;
//   * spreading is owner-computes: one workgroup per 8x8x8 brick of grid
//     points gathers the markers of the bins whose stencils reach the brick,
//     evaluates their separable weights once per chunk into LDS, and one
//     wavefront per component adds each marker's 4x4x4 = 64 stencil values
//     with one lane per stencil point.  The brick is written once with plain
//     stores: no zero fill, no global atomics, bitwise reproducible.
//   * interpolation is one lane per marker with the separable sum
//     sum_z wz sum_y wy sum_x wx u.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <climits>
#include <mutex>
#include "context.hpp"
#include "deltas.hpp"


// ----------------------------------------------------------------------
// Fixed workspace of ibm_spread_sets_noalloc: the buffers of one binning
// pass for at most STATIC_MAX_MARKERS markers and STATIC_MAX_BINS bins
// (a 128 x 129 x 128 cell grid), 36.6 MB per device, held from module load
// until the process ends.  Host code never passes `scratch` itself to a
// kernel: its device address (one per GPU) comes from hipGetSymbolAddress.
// ----------------------------------------------------------------------
namespace ibm_fixed {
constexpr size_t STATIC_MAX_MARKERS = size_t(1) << 18;
constexpr size_t STATIC_MAX_BINS = size_t(128) * 129 * 128;
constexpr size_t STATIC_SCAN_BYTES = size_t(1) << 20;
constexpr size_t STATIC_STATS_BYTES = 4096;

constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t STATIC_BIN_WORDS = align256(sizeof(int) * (STATIC_MAX_BINS + 1));
constexpr size_t STATIC_INT_ARRAY = align256(sizeof(int) * STATIC_MAX_MARKERS);
constexpr size_t STATIC_SCRATCH_BYTES =
    2 * STATIC_BIN_WORDS +                                   // count, start
    4 * STATIC_INT_ARRAY +                                   // key, rank, perm, code
    align256(sizeof(long long) * STATIC_MAX_MARKERS) +       // tmp
    align256(sizeof(double) * 6 * STATIC_MAX_MARKERS) +      // sorted positions and forces
    STATIC_STATS_BYTES + STATIC_SCAN_BYTES;
} // namespace ibm_fixed

static __device__ __align__(256) unsigned char scratch[ibm_fixed::STATIC_SCRATCH_BYTES];

namespace ibm {
namespace {

constexpr int MAX_SETS = 16;
constexpr int CHUNK = 64;
constexpr int SPREAD_THREADS = 256;

struct kaxis {
    double length, resolution, shift;
    double s_over_r;   // shift / resolution, rounded once as fd::discretization::units does per call
    int cells, points, solid, pad;
};

struct kgrids {
    kaxis a[3][3];     // [component][dimension]
    double res3[3];    // product of the resolutions of each component grid
    int points[3][3];
    int cand[2][3];    // per dimension, the components whose axes are the (at most two) distinct ones
    int sel[3][3];     // sel[c][j]: which candidate component c uses along dimension j
};

struct marker_sets {
    int nsets;
    int total;
    int offset[MAX_SETS + 1];
    const double* x[MAX_SETS];
    const double* f[MAX_SETS];
    double* out[MAX_SETS];
};

struct bin_desc {
    int nb[3];
    int qmin[3];
    int fc[3];
    int fs[3];
    int lo[3], hi[3];
    int nbins;
};

// Statistics gathered while binning; they bound the spread accumulators.
// Maxima are spread over STAT_SLOTS words (by block index) so the atomics
// do not all hit one address; readers reduce the slots.
constexpr int STAT_SLOTS = 64;
struct bin_stats {
    unsigned long long max_abs_f[STAT_SLOTS];   // bit patterns of max |f| (non-negative doubles order as integers)
    unsigned long long max_count[STAT_SLOTS];   // largest number of markers in one bin
};

__device__ inline int set_of(const marker_sets& s, int m)
{
    int k = 0;
    while (k + 1 < s.nsets && m >= s.offset[k + 1]) ++k;
    return k;
}

// axis_stencil() with the per-axis quotient shift/resolution precomputed and
// the periodic wrap (i + cells) % cells written for i in [-cells, 2 cells).
__device__ inline stencil_axis kstencil(const kaxis& a, double x)
{
#pragma clang fp contract(off)
    double y = x - a.s_over_r;
    if (!a.solid) {
        // For y in [-L, 2L), floor(y / L) is decided by comparisons alone
        // (correctly rounded division is monotone and y / L < 1 whenever
        // y < L), so modulo(y, L) = y - L * floor(y / L) is reproduced
        // exactly without the division.
        double L = a.length;
        if (y >= 0 && y < L) {
        } else if (y < 0 && y >= -L) {
            y = y + L;
        } else if (y >= L && y < 2 * L) {
            y = y - L;
        } else {
            y = ref_modulo(y, L);
        }
    }
    double diff = (y * a.resolution + a.shift) - a.shift;
    double fl = ref_floor(diff);
    int i = (int) fl;
    int j = i;
    if (!a.solid) j = i < 0 ? i + a.cells : (i >= a.cells ? i - a.cells : i);
    int w = a.points + a.solid;
    int k = j + a.solid;
    k = k < 0 ? 0 : (k >= w ? w - 1 : k);
    stencil_axis s;
    s.base = k - a.solid;
    s.d = fl - diff;
    return s;
}

// Maximum over the 64 lanes of a wavefront (all lanes must be active).
__device__ inline unsigned long long wave_max(unsigned long long v)
{
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

// ----------------------------------------------------------------------
// Binning: key = bin of the marker's face-aligned cell.  For dimension j the
// face component fc[j] is aligned on cell faces there; its (clamped) base
// cell plus the integer part of its shift is q_j.  Every component's base
// cell then lies in q_j + [lo_j, hi_j], which bounds the bins a brick reads.
// ----------------------------------------------------------------------
// Clears n 32-bit words (used in place of hipMemsetAsync so that a binning
// pass can be recorded in a HIP graph).
__global__ void zero_words_kernel(unsigned int* __restrict__ p, size_t n)
{
    for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t) gridDim.x * blockDim.x)
        p[i] = 0u;
}

__global__ void bin_count_kernel(marker_sets sets, kgrids g, bin_desc bd, int with_forces,
                                 int* __restrict__ key, int* __restrict__ rank,
                                 int* __restrict__ code, int* __restrict__ count,
                                 bin_stats* __restrict__ stats)
{
    // No early return: every lane takes part in the wavefront reduction and
    // in the grouping of the rank atomics.
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long fbits = 0;
    const bool valid = m < sets.total;
    int bin = -1;
    if (valid) {
        int k = set_of(sets, m);
        int l = m - sets.offset[k];
        int n = sets.offset[k + 1] - sets.offset[k];
        const double* x = sets.x[k];
        double p[3] = {x[l], x[n + l], x[2 * n + l]};
        int q[3];
        bin = 0;
        int stride = 1;
        for (int j = 0; j < 3; ++j) {
            stencil_axis st = kstencil(g.a[bd.fc[j]][j], p[j]);
            q[j] = st.base + bd.fs[j];
            int qb = q[j] - bd.qmin[j];
            qb = qb < 0 ? 0 : (qb >= bd.nb[j] ? bd.nb[j] - 1 : qb);
            bin += qb * stride;
            stride *= bd.nb[j];
        }
        int c9 = 0;
        for (int c = 0; c < 3; ++c)
            for (int j = 0; j < 3; ++j) {
                stencil_axis st = kstencil(g.a[c][j], p[j]);
                c9 |= ((st.base - q[j]) & 1) << (3 * c + j);
            }
        key[m] = bin;
        code[m] = c9;
        if (with_forces) {
            const double* f = sets.f[k];
            const double f0 = fabs(f[l]), f1 = fabs(f[n + l]), f2 = fabs(f[2 * n + l]);
            double a = f0 > f1 ? f0 : f1;
            a = a > f2 ? a : f2;
            // A non-finite force is recorded as the canonical NaN, whose bit
            // pattern exceeds every finite and infinite one, so it survives
            // the integer maxima below and poisons the spread result.
            if (!(isfinite(f0) && isfinite(f1) && isfinite(f2))) a = __longlong_as_double(0x7ff8000000000000ll);
            fbits = (unsigned long long) __double_as_longlong(a);
        }
    }
    // Ranks inside a bin: consecutive markers mostly share a bin, so the
    // lanes of a wavefront are grouped by bin (one ballot per distinct bin)
    // and each group's first lane takes the group's ranks with one atomic;
    // the groups' atomics go out together, to distinct addresses.
    {
        const int lane = threadIdx.x & 63;
        const unsigned long long lt = (1ull << lane) - 1;
        unsigned long long todo = __ballot(valid);
        int leader = lane, before = 0, size = 0;
        while (todo) {
            const int l = __ffsll((long long) todo) - 1;
            const int lb = __shfl(bin, l, 64);
            const unsigned long long grp = __ballot(valid && bin == lb) & todo;
            if ((grp >> lane) & 1) {
                leader = l;
                before = __popcll(grp & lt);
            }
            if (lane == l) size = __popcll(grp);
            todo &= ~grp;
        }
        int base = 0;
        if (valid && leader == lane) base = atomicAdd(&count[bin], size);
        base = __shfl(base, leader, 64);
        if (valid) rank[m] = base + before;
    }
    if (with_forces) {
        fbits = wave_max(fbits);
        if ((threadIdx.x & 63) == 0 && fbits) atomicMax(&stats->max_abs_f[blockIdx.x % STAT_SLOTS], fbits);
    }
}

// Places each marker's ordering key (code, marker id) in its bin's range.
__global__ void bin_scatter_kernel(int total, const int* __restrict__ key, const int* __restrict__ rank,
                                   const int* __restrict__ code, const int* __restrict__ start,
                                   long long* __restrict__ tmp)
{
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= total) return;
    tmp[start[key[m]] + rank[m]] = ((long long) code[m] << 32) | (unsigned) m;
}

// Canonical order inside a bin: by (code, marker id).  Writes the permutation
// and sorted copies of positions and forces (SoA, 6 * total doubles).
__global__ void bin_canon_kernel(marker_sets sets, int with_forces,
                                 const int* __restrict__ key, const int* __restrict__ start,
                                 const long long* __restrict__ tmp,
                                 int* __restrict__ perm, double* __restrict__ sorted,
                                 bin_stats* __restrict__ stats)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    int total = sets.total;
    unsigned long long cnt = 0;
    if (p < total) {
        const long long mine = tmp[p];
        int m = (int) (mine & 0xffffffffll);
        int b = key[m];
        int s = start[b], e = start[b + 1];
        int r = 0;
        for (int i = s; i < e; ++i) r += tmp[i] < mine;
        int dst = s + r;
        perm[dst] = m;
        cnt = (unsigned long long) (e - s);
        int k = set_of(sets, m);
        int l = m - sets.offset[k];
        int n = sets.offset[k + 1] - sets.offset[k];
        for (int j = 0; j < 3; ++j) sorted[(size_t) j * total + dst] = sets.x[k][(size_t) j * n + l];
        if (with_forces)
            for (int j = 0; j < 3; ++j) sorted[(size_t) (3 + j) * total + dst] = sets.f[k][(size_t) j * n + l];
    }
    cnt = wave_max(cnt);
    if ((threadIdx.x & 63) == 0 && cnt) atomicMax(&stats->max_count[blockIdx.x % STAT_SLOTS], cnt);
}

// ----------------------------------------------------------------------
// Spreading
//
// One workgroup per brick of BX x BY x BZ grid points (all three
// components).  The markers of the bins whose stencils reach the brick are
// walked in canonical order in chunks of 64; every wavefront owns a stream
// of chunks.  For a chunk, lanes are markers: each computes its separable
// weights once and, if its stencil meets the brick, stores them in the
// wavefront's LDS slice (compacted with a ballot).  Then each quarter of the
// wavefront walks a quarter of the compacted markers with lanes = the 16
// (sy, sz) rows of the stencil, four x points per lane: it sums the
// contributions of consecutive markers of one cell in registers and adds
// them to the brick tile when the cell changes.
//
// The tile holds 64-bit fixed-point integers (two's complement): integer
// addition is associative, so the four wavefronts may add in any order and
// the result is still bitwise reproducible.  The quantum q is a power of
// two chosen so that |sum| < 2^62 q for every point:
//     sum <= 64 cells * max_count * max|f| * h^-3 * max phi^3  (phi^3 <= 0.3),
// and each added cell sum is rounded to the nearest multiple of q, so a
// point's error is at most 64 * q / 2 per flush group, about 2^-55 of that
// bound.
// ----------------------------------------------------------------------
struct region_dim {
    int first;   // first bin coordinate (may be negative before wrapping)
    int count;   // number of bins
    int wrap;    // periodic wrap modulus (0 = none)
};

template <int B>
__device__ inline region_dim brick_region(const bin_desc& bd, const kaxis& fa, int j, int P0)
{
    region_dim r;
    int nb = bd.nb[j];
    if (!fa.solid) {
        int lo = P0 - 2 - bd.hi[j];
        int hi = P0 + B - bd.lo[j];
        if (hi - lo + 1 >= nb) { r.first = 0; r.count = nb; r.wrap = 0; }
        else { r.first = lo - bd.qmin[j]; r.count = hi - lo + 1; r.wrap = nb; }
    } else {
        int lo = P0 - 2 - bd.hi[j] - bd.qmin[j];
        int hi = P0 + B - bd.lo[j] - bd.qmin[j];
        lo = lo < 0 ? 0 : lo;
        hi = hi >= nb ? nb - 1 : hi;
        r.first = lo; r.count = hi >= lo ? hi - lo + 1 : 0; r.wrap = 0;
    }
    return r;
}

// Offset of stencil point (base + o - 1) inside the brick along one axis, or -1.
template <int B>
__device__ inline int brick_offset(const kaxis& a, int base, int o, int P0)
{
    int v = base + o - 1 - P0;
    if (!a.solid) {
        if (v < 0) v += a.cells;
        else if (v >= a.cells) v -= a.cells;
    }
    int lim = a.points - P0;
    lim = lim < B ? lim : B;
    return (v >= 0 && v < lim) ? v : -1;
}

__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int BX = 8, BY = 8, BZ = 8;
constexpr int BRICK_PTS = BX * BY * BZ;
constexpr int WAVES = SPREAD_THREADS / 64;
constexpr int MAX_ROWS = (BY + 4) * (BZ + 4);
constexpr int MAX_SEG2 = 2 * MAX_ROWS;
// Padded weight row: 2*WROW dwords is 2 mod 4, so any 16 consecutive rows
// start in distinct LDS banks, and the 8 extra entries absorb the reads of a
// final partial batch.
constexpr int WROW = CHUNK + 9;
constexpr int VISIT_BATCH = 4;
// Wx row of 6 doubles (48 bytes): the 16-byte stores of 16 consecutive lanes
// start 12 banks apart and cover distinct banks.
constexpr int WX_STRIDE = 6;
constexpr int POS_LIST = 2048;      // region markers listed in LDS; larger regions search the segments
// FAST: every periodic axis has at least BX + 8 cells, so a stencil that
// reaches the brick has a unique brick-relative origin in [-3, B).
constexpr int MIN_FAST_CELLS = BX + 8;
// Bias that keeps the flat brick index of a stencil origin in [-3, B)^3 positive (12 bits).
constexpr int ORG_BIAS = 3 + BX * (3 + BY * 3);
static_assert(ORG_BIAS + BRICK_PTS < 4096, "flat origin must fit in 12 bits");

// Whether the stencil of base cell `base` meets the brick along one axis;
// with FAST, also its brick-relative origin (offset of stencil point 0).
template <bool FAST, int B>
__device__ inline bool axis_hit(const kaxis& a, int base, int P0, int& origin)
{
    int lim = a.points - P0;
    lim = lim < B ? lim : B;
    if (FAST) {
        int v = base - 1 - P0;
        if (!a.solid) {
            if (v < -4) v += a.cells;
            else if (v >= a.cells - 4) v -= a.cells;
        }
        origin = v;
        return v + 3 >= 0 && v < lim;
    } else {
        origin = 0;
        bool m = false;
#pragma unroll
        for (int o = 0; o < 4; ++o) m = m || brick_offset<B>(a, base, o, P0) >= 0;
        return m;
    }
}

template <bool MAC, bool FAST>
__global__ void __launch_bounds__(SPREAD_THREADS)
spread_brick_kernel(kgrids g, bin_desc bd, int delta, int total,
                    const int* __restrict__ start, const double* __restrict__ sorted,
                    const bin_stats* __restrict__ stats,
                    double* __restrict__ fe0, double* __restrict__ fe1, double* __restrict__ fe2,
                    int bricks_x, int bricks_y, int nbricks, int xcd_chunk)
{
    __shared__ unsigned long long acc[3][BRICK_PTS];
    __shared__ double W[WAVES][16][WROW];     // wz[a] * wy[b] * f at row 4a + b
    __shared__ __attribute__((aligned(16))) double Wx[WAVES][CHUNK + VISIT_BATCH][WX_STRIDE];   // wx[0..3]
    __shared__ int org[WAVES][CHUNK + VISIT_BATCH];
    __shared__ int seg_begin[MAX_SEG2];
    __shared__ int seg_pref[MAX_SEG2 + 1];
    __shared__ int pos_list[POS_LIST];
    __shared__ double bound_s;

    const int tid = threadIdx.x;
    int bid = blockIdx.x;
    if (xcd_chunk > 0) bid = (bid & 7) * xcd_chunk + (bid >> 3);
    if (bid >= nbricks) return;
    const int P0x = (bid % bricks_x) * BX;
    const int P0y = ((bid / bricks_x) % bricks_y) * BY;
    const int P0z = (bid / (bricks_x * bricks_y)) * BZ;

    for (int i = tid; i < 3 * BRICK_PTS; i += SPREAD_THREADS) (&acc[0][0])[i] = 0ull;

    region_dim rx = brick_region<BX>(bd, g.a[bd.fc[0]][0], 0, P0x);
    region_dim ry = brick_region<BY>(bd, g.a[bd.fc[1]][1], 1, P0y);
    region_dim rz = brick_region<BZ>(bd, g.a[bd.fc[2]][2], 2, P0z);
    const int rows = ry.count * rz.count;          // <= MAX_ROWS (checked on the host)
    const int nseg = 2 * rows;
    for (int r = tid; r < rows; r += SPREAD_THREADS) {
        int yy = ry.first + r % ry.count;
        int zz = rz.first + r / ry.count;
        if (ry.wrap) yy = (yy % ry.wrap + ry.wrap) % ry.wrap;
        if (rz.wrap) zz = (zz % rz.wrap + rz.wrap) % rz.wrap;
        int row = (zz * bd.nb[1] + yy) * bd.nb[0];
        int x0 = rx.first;
        if (rx.wrap) x0 = (x0 % rx.wrap + rx.wrap) % rx.wrap;
        int c0 = rx.count, c1 = 0;
        if (rx.wrap && x0 + rx.count > bd.nb[0]) { c0 = bd.nb[0] - x0; c1 = rx.count - c0; }
        int s0 = start[row + x0], e0 = start[row + x0 + c0];
        int s1 = start[row], e1 = c1 ? start[row + c1] : s1;
        seg_begin[2 * r] = s0;     seg_pref[2 * r + 1] = e0 - s0;
        seg_begin[2 * r + 1] = s1; seg_pref[2 * r + 2] = e1 - s1;
    }
    __syncthreads();
    if (tid < 64) {
        constexpr int PER = (MAX_SEG2 + 63) / 64;
        int vals[PER];
        int sum = 0;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            int idx = tid * PER + i;
            vals[i] = idx < nseg ? seg_pref[idx + 1] : 0;
            sum += vals[i];
        }
        int incl = sum;
        for (int off = 1; off < 64; off <<= 1) {
            int y = __shfl_up(incl, off, 64);
            if (tid >= off) incl += y;
        }
        int run = incl - sum;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            int idx = tid * PER + i;
            run += vals[i];
            if (idx < nseg) seg_pref[idx + 1] = run;
        }
        if (tid == 0) seg_pref[0] = 0;
        unsigned long long fb = wave_max(stats->max_abs_f[tid]);
        unsigned long long cm = wave_max(stats->max_count[tid]);
        if (tid == 0) {
            double res3 = fmax(g.res3[0], fmax(g.res3[1], g.res3[2]));
            bound_s = 64.0 * (double) cm * __longlong_as_double((long long) fb) * res3 * 0.3;
        }
    }
    __syncthreads();
    const int T = seg_pref[nseg];
    // Flattened sorted positions of the region's markers (when they fit).
    const bool listed = T <= POS_LIST;
    if (listed)
        for (int i = tid; i < nseg; i += SPREAD_THREADS) {
            int b = seg_begin[i], o = seg_pref[i], n = seg_pref[i + 1] - o;
            for (int k = 0; k < n; ++k) pos_list[o + k] = b + k;
        }
    __syncthreads();
    // A NaN or infinite force makes the bound non-finite: every point is then
    // written as NaN (the accumulation below still runs, on a dummy quantum).
    const bool poisoned = !(bound_s <= 1.7976931348623157e308);
    int e2 = 0;
    if (!poisoned && bound_s > 0) frexp(bound_s, &e2);
    const double inv_q = ldexp(1.0, 62 - e2);

    const int wave = tid >> 6;
    const int lane = tid & 63;
    const unsigned long long lanemask_lt = (1ull << lane) - 1;

    // Sorted position of region marker idx (binary search over the segments).
    auto locate = [&](int idx) {
        if (listed) return pos_list[idx];
        int lo = 0, hi = nseg;
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (seg_pref[mid] <= idx) lo = mid; else hi = mid;
        }
        return seg_begin[lo] + (idx - seg_pref[lo]);
    };
    // Positions and forces of the next chunk are loaded while the current
    // one is processed.
    double np_[3] = {0, 0, 0}, nf_[3] = {0, 0, 0};
    {
        int idx = wave * CHUNK + lane;
        if (idx < T) {
            int pp = locate(idx);
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                np_[j] = sorted[(size_t) j * total + pp];
                nf_[j] = sorted[(size_t) (3 + j) * total + pp];
            }
        }
    }
    for (int cb = wave * CHUNK; cb < T; cb += WAVES * CHUNK) {
        // ---- lanes are markers ----
        int idx = cb + lane;
        bool valid = idx < T;
        double p[3] = {np_[0], np_[1], np_[2]}, fv[3] = {nf_[0], nf_[1], nf_[2]};
        {
            int nidx = idx + WAVES * CHUNK;
            if (nidx < T) {
                int pp = locate(nidx);
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    np_[j] = sorted[(size_t) j * total + pp];
                    nf_[j] = sorted[(size_t) (3 + j) * total + pp];
                }
            }
        }
        double w0[3][4], w1[3][4];
        int b0[3], b1[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            stencil_axis st = kstencil(g.a[g.cand[0][j]][j], p[j]);
            b0[j] = st.base;
            delta4(delta, st.d, w0[j]);
            if (MAC || g.cand[1][j] != g.cand[0][j]) {
                st = kstencil(g.a[g.cand[1][j]][j], p[j]);
                b1[j] = st.base;
                delta4(delta, st.d, w1[j]);
            } else {
                b1[j] = b0[j];
#pragma unroll
                for (int o = 0; o < 4; ++o) w1[j][o] = w0[j][o];
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const bool s0 = MAC ? (c == 0) : (g.sel[c][0] != 0);
            const bool s1 = MAC ? (c == 1) : (g.sel[c][1] != 0);
            const bool s2 = MAC ? (c == 2) : (g.sel[c][2] != 0);
            const int bb0 = s0 ? b1[0] : b0[0];
            const int bb1 = s1 ? b1[1] : b0[1];
            const int bb2 = s2 ? b1[2] : b0[2];
            double w[3][4];
#pragma unroll
            for (int o = 0; o < 4; ++o) {
                w[0][o] = s0 ? w1[0][o] : w0[0][o];
                w[1][o] = s1 ? w1[1][o] : w0[1][o];
                w[2][o] = s2 ? w1[2][o] : w0[2][o];
            }
            // Does the stencil meet the brick?  Brick-relative origin of stencil point 0.
            int ow0, ow1, ow2;
            bool hit = valid && axis_hit<FAST, BX>(g.a[c][0], bb0, P0x, ow0)
                             && axis_hit<FAST, BY>(g.a[c][1], bb1, P0y, ow1)
                             && axis_hit<FAST, BZ>(g.a[c][2], bb2, P0z, ow2);
            const unsigned long long hits = __ballot(hit);
            const int nh = __popcll(hits);
            const int slot = __popcll(hits & lanemask_lt);
            if (hit) {
                double fc = fv[c] * g.res3[c];
                double2* wxp = reinterpret_cast<double2*>(&Wx[wave][slot][0]);
                wxp[0] = make_double2(w[0][0], w[0][1]);
                wxp[1] = make_double2(w[0][2], w[0][3]);
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    double wyf = w[1][b] * fc;
#pragma unroll
                    for (int a = 0; a < 4; ++a) W[wave][4 * a + b][slot] = w[2][a] * wyf;
                }
                if (FAST) {
                    // flat brick index of stencil point 0 (biased to stay positive) and,
                    // per axis, which of the four offsets fall inside the brick
                    int lx = min(g.a[c][0].points - P0x, BX), ly = min(g.a[c][1].points - P0y, BY);
                    int lz = min(g.a[c][2].points - P0z, BZ);
                    int mx = 0, my = 0, mz = 0;
#pragma unroll
                    for (int o = 0; o < 4; ++o) {
                        mx |= ((unsigned) (ow0 + o) < (unsigned) lx) << o;
                        my |= ((unsigned) (ow1 + o) < (unsigned) ly) << o;
                        mz |= ((unsigned) (ow2 + o) < (unsigned) lz) << o;
                    }
                    int flat = ow0 + BX * (ow1 + BY * ow2) + ORG_BIAS;
                    org[wave][slot] = flat | (mx << 12) | (my << 16) | (mz << 20);
                } else {
                    org[wave][slot] = (bb0 + 1) | ((bb1 + 1) << 10) | ((bb2 + 1) << 20);
                }
            }
            wave_sync();
            // ---- lanes are (quarter q, sy, sz), each for the four x points ----
            // Quarter q walks compacted markers [q per, (q+1) per) and sums
            // the contributions of consecutive markers of one cell in four
            // registers; a change of cell (stencil origin) adds them to the tile.
            const kaxis& ax = g.a[c][0];
            const kaxis& ay = g.a[c][1];
            const kaxis& az = g.a[c][2];
            const int qy = lane & 3, qz = (lane >> 2) & 3, quarter = lane >> 4;
            const double* wrow = &W[wave][4 * qz + qy][0];
            const int per = (nh + 3) >> 2;
            const int kbeg = quarter * per;
            const int kend = min(kbeg + per, nh);
            const int lane_off = BX * (qy + BY * qz) - ORG_BIAS;
            double P[4] = {0.0, 0.0, 0.0, 0.0};
            int cur = -1;
            auto flush = [&]() {
                if (cur < 0) return;
                if (FAST) {
                    const int o = cur;
                    if ((o >> (16 + qy)) & (o >> (20 + qz)) & 1) {
                        const int t0 = (o & 4095) + lane_off;
#pragma unroll
                        for (int sx = 0; sx < 4; ++sx)
                            if ((o >> (12 + sx)) & 1)
                                atomicAdd(&acc[c][t0 + sx], (unsigned long long) __double2ll_rn(P[sx] * inv_q));
                    }
                } else {
                    const int o = cur;
                    int oy = brick_offset<BY>(ay, ((o >> 10) & 1023) - 1, qy, P0y);
                    int oz = brick_offset<BZ>(az, ((o >> 20) & 1023) - 1, qz, P0z);
                    if (oy >= 0 && oz >= 0) {
#pragma unroll
                        for (int sx = 0; sx < 4; ++sx) {
                            int ox = brick_offset<BX>(ax, (o & 1023) - 1, sx, P0x);
                            if (ox >= 0)
                                atomicAdd(&acc[c][ox + BX * (oy + BY * oz)],
                                          (unsigned long long) __double2ll_rn(P[sx] * inv_q));
                        }
                    }
                }
            };
            for (int k0 = kbeg; k0 < kend; k0 += VISIT_BATCH) {
                // all loads of the batch first (entries past kend are never used)
                int ov[VISIT_BATCH];
                double pv[VISIT_BATCH];
                double2 xa[VISIT_BATCH], xb[VISIT_BATCH];
#pragma unroll
                for (int i = 0; i < VISIT_BATCH; ++i) {
                    ov[i] = org[wave][k0 + i];
                    pv[i] = wrow[k0 + i];
                    const double2* wxp = reinterpret_cast<const double2*>(&Wx[wave][k0 + i][0]);
                    xa[i] = wxp[0];
                    xb[i] = wxp[1];
                }
#pragma unroll
                for (int i = 0; i < VISIT_BATCH; ++i) {
                    if (k0 + i < kend) {
                        if (ov[i] != cur) {
                            flush();
                            P[0] = P[1] = P[2] = P[3] = 0.0;
                            cur = ov[i];
                        }
                        P[0] = fma(xa[i].x, pv[i], P[0]);
                        P[1] = fma(xa[i].y, pv[i], P[1]);
                        P[2] = fma(xb[i].x, pv[i], P[2]);
                        P[3] = fma(xb[i].y, pv[i], P[3]);
                    }
                }
            }
            flush();
            wave_sync();
        }
    }
    __syncthreads();

    // Each grid point belongs to exactly one brick: plain stores.
    const double q = poisoned ? __longlong_as_double(0x7ff8000000000000ll) : ldexp(1.0, e2 - 62);
    double* fe[3] = {fe0, fe1, fe2};
    for (int i = tid; i < BRICK_PTS; i += SPREAD_THREADS) {
        int tx = i % BX, ty = (i / BX) % BY, tz = i / (BX * BY);
        int X = P0x + tx, Y = P0y + ty, Z = P0z + tz;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int* pt = g.points[c];
            if (X < pt[0] && Y < pt[1] && Z < pt[2])
                fe[c][(size_t) X + (size_t) pt[0] * ((size_t) Y + (size_t) pt[1] * Z)] =
                    (double) (long long) acc[c][i] * q;
        }
    }
}

// ----------------------------------------------------------------------
// Interpolation: one lane per marker.
// ----------------------------------------------------------------------
// Four doubles at 8-byte alignment (global loads need only dword alignment).
typedef double d4u8 __attribute__((ext_vector_type(4), aligned(8)));

__global__ void __launch_bounds__(256)
interp_kernel(marker_sets sets, kgrids g, int delta,
              const double* __restrict__ u0, const double* __restrict__ u1,
              const double* __restrict__ u2)
{
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= sets.total) return;
    int k = set_of(sets, m);
    int l = m - sets.offset[k];
    int n = sets.offset[k + 1] - sets.offset[k];
    const double* x = sets.x[k];
    double p[3] = {x[l], x[n + l], x[2 * n + l]};
    for (int c = 0; c < 3; ++c) {
        double w[3][4];
        size_t ix[4], oy[4], oz[4];
        const int* pt = g.points[c];
        for (int j = 0; j < 3; ++j) {
            stencil_axis st = axis_stencil(g.a[c][j], p[j]);
            delta4(delta, st.d, w[j]);
            for (int o = 0; o < 4; ++o) {
                size_t idx = (size_t) axis_interp_index(g.a[c][j], st.base + o - 1);
                if (j == 0) ix[o] = idx;
                else if (j == 1) oy[o] = idx * (size_t) pt[0];
                else oz[o] = idx * (size_t) pt[0] * (size_t) pt[1];
            }
        }
        const double* u = c == 0 ? u0 : (c == 1 ? u1 : u2);    // global loads (an indexed array would be flat)
        // the four x points are adjacent except across a periodic seam or a
        // clamped wall: then one 32-byte load per row instead of four
        const bool adjacent = ix[1] == ix[0] + 1 && ix[2] == ix[0] + 2 && ix[3] == ix[0] + 3;
        double acc = 0.0;
        for (int a = 0; a < 4; ++a) {
            double sy = 0.0;
            for (int b = 0; b < 4; ++b) {
                const double* row = u + oz[a] + oy[b];
                double v[4];
                if (adjacent) {
                    const d4u8 q = *(const d4u8*) (row + ix[0]);
                    v[0] = q.x;
                    v[1] = q.y;
                    v[2] = q.z;
                    v[3] = q.w;
                } else {
#pragma unroll
                    for (int o = 0; o < 4; ++o) v[o] = row[ix[o]];
                }
                double sx = w[0][0] * v[0];
                sx = fma(w[0][1], v[1], sx);
                sx = fma(w[0][2], v[2], sx);
                sx = fma(w[0][3], v[3], sx);
                sy = fma(w[1][b], sx, sy);
            }
            acc = fma(w[2][a], sy, acc);
        }
        sets.out[k][(size_t) c * n + l] = acc;
    }
}

// ----------------------------------------------------------------------
// Host side
// ----------------------------------------------------------------------
kaxis to_kaxis(const ibm_axis& a)
{
    kaxis k;
    k.length = a.length; k.resolution = a.resolution; k.shift = a.shift;
    k.s_over_r = a.shift / a.resolution;
    k.cells = a.cells; k.points = a.points; k.solid = a.solid; k.pad = 0;
    return k;
}

ibm_status make_kgrids(ibm_ctx* ctx, const ibm_grid* grids, kgrids* out)
{
    if (!grids) return fail(ctx, IBM_ERR_ARG, "grids is null");
    for (int c = 0; c < 3; ++c) {
        if (grids[c].dims != 3)
            return fail(ctx, IBM_ERR_ARG, "spread/interpolate support 3-dimensional grids (got %d)", grids[c].dims);
        double r = 1.0;
        for (int j = 0; j < 3; ++j) {
            out->a[c][j] = to_kaxis(grids[c].axis[j]);
            out->points[c][j] = grids[c].axis[j].points;
            r *= grids[c].axis[j].resolution;
            if (grids[c].axis[j].cells < 1 || grids[c].axis[j].points < 1)
                return fail(ctx, IBM_ERR_BAD_GRID, "empty grid axis");
        }
        out->res3[c] = r;
        if ((long long) grids[c].axis[0].points * grids[c].axis[1].points * grids[c].axis[2].points > INT_MAX)
            return fail(ctx, IBM_ERR_SIZE, "spread/interpolate grids hold at most %d points", INT_MAX);
    }
    // Along each dimension the component grids may use at most two distinct
    // discretisations (MAC: the face-aligned one and the centred one).
    for (int j = 0; j < 3; ++j) {
        out->cand[0][j] = 0;
        out->cand[1][j] = 0;
        int ncand = 1;
        for (int c = 0; c < 3; ++c) {
            const ibm_axis& a = grids[c].axis[j];
            int found = -1;
            for (int k = 0; k < ncand; ++k) {
                const ibm_axis& b = grids[out->cand[k][j]].axis[j];
                if (a.shift == b.shift && a.resolution == b.resolution && a.cells == b.cells &&
                    a.points == b.points && a.solid == b.solid && a.length == b.length) {
                    found = k;
                    break;
                }
            }
            if (found < 0) {
                if (ncand == 2)
                    return fail(ctx, IBM_ERR_BAD_GRID, "more than two staggerings along dimension %d", j);
                out->cand[1][j] = c;
                found = ncand++;
            }
            out->sel[c][j] = found;
        }
        // Canonical order for a MAC-shaped axis (component j alone differs):
        // candidate 1 is component j, candidate 0 the shared centred axis.
        if (ncand == 2) {
            int other = (j + 1) % 3, third = (j + 2) % 3;
            if (out->sel[j][j] != out->sel[other][j] && out->sel[other][j] == out->sel[third][j]) {
                out->cand[0][j] = other;
                out->cand[1][j] = j;
                for (int c = 0; c < 3; ++c) out->sel[c][j] = (c == j) ? 1 : 0;
            }
        }
    }
    return IBM_OK;
}

ibm_status make_sets(ibm_ctx* ctx, int nsets, const int* n, const double* const* x,
                     const double* const* f, double* const* out, marker_sets* s)
{
    if (nsets < 0 || nsets > MAX_SETS) return fail(ctx, IBM_ERR_ARG, "nsets must be in [0, %d]", MAX_SETS);
    if (nsets > 0 && (!n || !x)) return fail(ctx, IBM_ERR_ARG, "null marker arrays");
    s->nsets = nsets;
    long long total = 0;
    s->offset[0] = 0;
    for (int k = 0; k < nsets; ++k) {
        if (n[k] < 0) return fail(ctx, IBM_ERR_SIZE, "negative marker count");
        if (n[k] > 0 && !x[k]) return fail(ctx, IBM_ERR_ARG, "null positions for set %d", k);
        if (n[k] > 0 && f && !f[k]) return fail(ctx, IBM_ERR_ARG, "null forces for set %d", k);
        if (n[k] > 0 && out && !out[k]) return fail(ctx, IBM_ERR_ARG, "null output for set %d", k);
        s->x[k] = x[k];
        s->f[k] = f ? f[k] : nullptr;
        s->out[k] = out ? out[k] : nullptr;
        total += n[k];
        s->offset[k + 1] = (int) total;
    }
    if (total > (1LL << 30)) return fail(ctx, IBM_ERR_SIZE, "too many markers");
    s->total = (int) total;
    return IBM_OK;
}

bin_desc make_bins(const ibm_grid* grids)
{
    bin_desc bd;
    long long nbins = 1;
    for (int j = 0; j < 3; ++j) {
        int fc = -1;
        for (int c = 0; c < 3; ++c)
            if (grids[c].axis[j].alignment == 0.0) { fc = c; break; }
        if (fc < 0) fc = 0;
        const ibm_axis& fa = grids[fc].axis[j];
        int fs = (int) fa.shift;                  // 0 or 1 for face grids, 0 otherwise
        bd.fc[j] = fc;
        bd.fs[j] = fs;
        if (fa.solid) { bd.qmin[j] = -1 + fs; bd.nb[j] = fa.points + 1; }
        else { bd.qmin[j] = 0; bd.nb[j] = fa.cells; }
        // base_c - q = floor(y - D) - floor(y) - fs with y = x/h - fs and
        // D = shift_c - fs, i.e. in {-ceil(D) - fs, -floor(D) - fs}.
        int lo = INT_MAX, hi = INT_MIN;
        for (int c = 0; c < 3; ++c) {
            double delta = grids[c].axis[j].shift - fs;
            int a = -(int) ceil(delta) - fs, b = -(int) floor(delta) - fs;
            lo = a < lo ? a : lo;
            hi = b > hi ? b : hi;
        }
        bd.lo[j] = lo;
        bd.hi[j] = hi;
        nbins *= bd.nb[j];
    }
    // Grids are limited to INT_MAX points (make_kgrids), so this only fails
    // with walls adding a bin row; callers check nbins > 0.
    bd.nbins = nbins <= INT_MAX - 1 ? (int) nbins : -1;
    return bd;
}

int blocks_for(long long n, int threads) { return (int) ((n + threads - 1) / threads); }

// Device buffers of one binning pass.
struct bin_work {
    int *count, *start, *key, *rank, *perm, *code;
    long long* tmp;
    double* sorted;          // 6 * total doubles: positions, then forces (SoA; per-marker records in the tile path)
    bin_stats* stats;
    void* scan_tmp;
    size_t scan_bytes;
};

// The binning buffers in the context's scratch slots (grown on demand).
ibm_status ctx_bin_work(ibm_ctx* ctx, int total, int nbins, bin_work* w)
{
    void* p;
    IBM_TRY(scratch(ctx, SLOT_BIN_COUNT, sizeof(int) * (size_t) (nbins + 1), &p)); w->count = (int*) p;
    IBM_TRY(scratch(ctx, SLOT_BIN_START, sizeof(int) * (size_t) (nbins + 1), &p)); w->start = (int*) p;
    IBM_TRY(scratch(ctx, SLOT_KEY, sizeof(int) * (size_t) total, &p)); w->key = (int*) p;
    IBM_TRY(scratch(ctx, SLOT_RANK, sizeof(int) * (size_t) total, &p)); w->rank = (int*) p;
    IBM_TRY(scratch(ctx, SLOT_TMP, sizeof(long long) * (size_t) total, &p)); w->tmp = (long long*) p;
    IBM_TRY(scratch(ctx, SLOT_PERM, sizeof(int) * (size_t) total, &p)); w->perm = (int*) p;
    IBM_TRY(scratch(ctx, SLOT_CODE, sizeof(int) * (size_t) total, &p)); w->code = (int*) p;
    IBM_TRY(scratch(ctx, SLOT_SORTED, sizeof(double) * 6 * (size_t) total, &p)); w->sorted = (double*) p;
    IBM_TRY(scratch(ctx, SLOT_FLAGS, sizeof(bin_stats), &p)); w->stats = (bin_stats*) p;
    size_t tbytes = 0;
    IBM_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tbytes, w->count, w->start, nbins + 1, ctx->stream));
    IBM_TRY(scratch(ctx, SLOT_SCAN_TMP, tbytes, &p));
    w->scan_tmp = p;
    w->scan_bytes = tbytes;
    return IBM_OK;
}

// Counting sort of all markers into bins: leaves the bin starts (nbins+1),
// the permutation and sorted positions (and forces) in `w`.  Every buffer
// that must start at zero is cleared here, by kernels (no memset), so the
// pass can be recorded in a HIP graph and run on reused memory.
ibm_status bin_markers(ibm_ctx* ctx, const marker_sets& sets, const kgrids& kg, const bin_desc& bd,
                       int with_forces, const bin_work& w, bool timed)
{
    const int total = sets.total;
    const int nbins = bd.nbins;
    prof_scope prof(ctx, timed ? "bin" : nullptr);
    {
        size_t words = (size_t) nbins + 1;
        zero_words_kernel<<<blocks_for((long) std::min<size_t>(words, 1u << 22), 256), 256, 0, ctx->stream>>>(
            (unsigned int*) w.count, words);
        IBM_LAUNCH_CHECK(ctx);
        zero_words_kernel<<<1, 256, 0, ctx->stream>>>((unsigned int*) w.stats, sizeof(bin_stats) / 4);
        IBM_LAUNCH_CHECK(ctx);
    }
    if (total > 0) {
        bin_count_kernel<<<blocks_for(total, 256), 256, 0, ctx->stream>>>(sets, kg, bd, with_forces, w.key, w.rank,
                                                                           w.code, w.count, w.stats);
        IBM_LAUNCH_CHECK(ctx);
    }
    size_t tbytes = w.scan_bytes;
    IBM_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(w.scan_tmp, tbytes, w.count, w.start, nbins + 1, ctx->stream));
    if (total > 0) {
        bin_scatter_kernel<<<blocks_for(total, 256), 256, 0, ctx->stream>>>(total, w.key, w.rank, w.code, w.start,
                                                                             w.tmp);
        IBM_LAUNCH_CHECK(ctx);
        bin_canon_kernel<<<blocks_for(total, 256), 256, 0, ctx->stream>>>(sets, with_forces, w.key, w.start, w.tmp,
                                                                           w.perm, w.sorted, w.stats);
        IBM_LAUNCH_CHECK(ctx);
    }
    return IBM_OK;
}

ibm_status check_delta(ibm_ctx* ctx, int delta)
{
    if (delta != IBM_DELTA_COSINE && delta != IBM_DELTA_PESKIN && delta != IBM_DELTA_BSPLINE3)
        return fail(ctx, IBM_ERR_ARG, "unknown delta kind %d", delta);
    return IBM_OK;
}

using namespace ibm_fixed;
static_assert(sizeof(bin_stats) <= STATIC_STATS_BYTES, "binning statistics exceed their fixed slot");
constexpr int STATIC_MAX_DEVICES = 64;
std::mutex static_scratch_mutex;
void* static_scratch_ptr[STATIC_MAX_DEVICES] = {};

// Carves the binning buffers out of the fixed workspace at `base` (a device
// address of `scratch`).
ibm_status static_bin_work(ibm_ctx* ctx, unsigned char* base, int total, int nbins, bin_work* w)
{
    if ((size_t) total > STATIC_MAX_MARKERS)
        return fail(ctx, IBM_ERR_SIZE, "ibm_spread_sets_noalloc takes at most %zu markers (got %d)",
                    STATIC_MAX_MARKERS, total);
    if ((size_t) nbins > STATIC_MAX_BINS)
        return fail(ctx, IBM_ERR_SIZE, "ibm_spread_sets_noalloc takes grids of at most %zu bins (got %d)",
                    STATIC_MAX_BINS, nbins);
    size_t off = 0;
    auto take = [&](size_t bytes) { unsigned char* p = base + off; off += align256(bytes); return p; };
    w->count = (int*) take(STATIC_BIN_WORDS);
    w->start = (int*) take(STATIC_BIN_WORDS);
    w->key = (int*) take(STATIC_INT_ARRAY);
    w->rank = (int*) take(STATIC_INT_ARRAY);
    w->perm = (int*) take(STATIC_INT_ARRAY);
    w->code = (int*) take(STATIC_INT_ARRAY);
    w->tmp = (long long*) take(sizeof(long long) * STATIC_MAX_MARKERS);
    w->sorted = (double*) take(sizeof(double) * 6 * STATIC_MAX_MARKERS);
    w->stats = (bin_stats*) take(STATIC_STATS_BYTES);
    w->scan_tmp = base + off;
    size_t tbytes = 0;
    // A size query: rocPRIM neither allocates nor launches anything here.
    IBM_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tbytes, w->count, w->start, nbins + 1, ctx->stream));
    if (tbytes > STATIC_SCAN_BYTES)
        return fail(ctx, IBM_ERR_SIZE, "scan workspace of %zu bytes exceeds the fixed %zu", tbytes, STATIC_SCAN_BYTES);
    w->scan_bytes = tbytes;
    return IBM_OK;
}

// ----------------------------------------------------------------------
// Tile spreading: every marker is visited once.
//
// The markers of nboxes independent boxes (one set of component grids,
// repeated; the C4 channel is one box, the small periodic boxes of a batch
// are many) are counting-sorted into bins of 8 x 8 x 8 face-aligned cells
// ("bricks"), brick-major.  One workgroup per brick adds each of its
// markers' 64 stencil values per component into a 12 x 12 x 12 tile of
// 64-bit fixed-point integers in LDS (the brick's points plus the reach of
// its stencils), and stores the tile once.  A merge pass then forms every
// grid point from the (at most eight) tiles that cover it.  Integer addition
// is associative, so the result does not depend on scheduling: no
// floating-point atomics, no global atomics, bitwise reproducible.
//
// Tile geometry along axis j: brick b covers bins q in [8 b, 8 b + 8) (bin
// coordinates relative to qmin); a marker of bin q has base cell
// base_c = q + qmin + d_c for component c with d_c in [lo_j, hi_j]
// (hi_j - lo_j <= 1), and touches points base_c - 1 .. base_c + 2, so the
// tile starts at point T0 = 8 b + qmin + lo - 1 and a marker's stencil
// origin lies at tile index (q - 8 b) + (d_c - lo) in [0, 8].
// ----------------------------------------------------------------------
constexpr int TB = 8;                       // brick edge in bins
constexpr int TT = 12;                      // tile edge in points
constexpr int TILE_PTS = TT * TT * TT;
constexpr int TILE_THREADS = 256;

struct tile_desc {
    int nbr[3];         // bricks per axis
    int bricks;         // bricks per box
    int nboxes;
    int keys;           // nboxes * bricks * 512
    int box_points[3];  // points of one box's component grid
};

// Box of marker m: offsets holds nboxes + 1 increasing marker indices.
__device__ inline int box_of(const int* __restrict__ off, int nboxes, int m)
{
    int lo = 0, hi = nboxes;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (off[mid] <= m) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ void tile_count_kernel(marker_sets sets, kgrids g, bin_desc bd, tile_desc td,
                                  const int* __restrict__ box_off, int* __restrict__ key, int* __restrict__ rank,
                                  int* __restrict__ count, bin_stats* __restrict__ stats)
{
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long fbits = 0;
    const bool valid = m < sets.total;
    int bin = -1;
    if (valid) {
        int k = set_of(sets, m);
        int l = m - sets.offset[k];
        int n = sets.offset[k + 1] - sets.offset[k];
        const double* x = sets.x[k];
        double p[3] = {x[l], x[n + l], x[2 * n + l]};
        int q[3], br[3], lb[3];
        for (int j = 0; j < 3; ++j) {
            stencil_axis st = kstencil(g.a[bd.fc[j]][j], p[j]);
            q[j] = st.base + bd.fs[j];
            int qb = q[j] - bd.qmin[j];
            qb = qb < 0 ? 0 : (qb >= bd.nb[j] ? bd.nb[j] - 1 : qb);
            br[j] = qb / TB;
            lb[j] = qb - br[j] * TB;
        }
        const int box = box_of(box_off, td.nboxes, m);
        const int brick = ((box * td.nbr[2] + br[2]) * td.nbr[1] + br[1]) * td.nbr[0] + br[0];
        bin = brick * (TB * TB * TB) + lb[0] + TB * (lb[1] + TB * lb[2]);
        key[m] = bin;
        const double* f = sets.f[k];
        const double f0 = fabs(f[l]), f1 = fabs(f[n + l]), f2 = fabs(f[2 * n + l]);
        double a = f0 > f1 ? f0 : f1;
        a = a > f2 ? a : f2;
        if (!(isfinite(f0) && isfinite(f1) && isfinite(f2))) a = __longlong_as_double(0x7ff8000000000000ll);
        fbits = (unsigned long long) __double_as_longlong(a);
    }
    {
        // ranks inside a bin, one atomic per group of lanes sharing a bin
        const int lane = threadIdx.x & 63;
        const unsigned long long lt = (1ull << lane) - 1;
        unsigned long long todo = __ballot(valid);
        int leader = lane, before = 0, size = 0;
        while (todo) {
            const int l = __ffsll((long long) todo) - 1;
            const int lbin = __shfl(bin, l, 64);
            const unsigned long long grp = __ballot(valid && bin == lbin) & todo;
            if ((grp >> lane) & 1) {
                leader = l;
                before = __popcll(grp & lt);
            }
            if (lane == l) size = __popcll(grp);
            todo &= ~grp;
        }
        int base = 0;
        if (valid && leader == lane) base = atomicAdd(&count[bin], size);
        base = __shfl(base, leader, 64);
        if (valid) rank[m] = base + before;
    }
    fbits = wave_max(fbits);
    if ((threadIdx.x & 63) == 0 && fbits) atomicMax(&stats->max_abs_f[blockIdx.x % STAT_SLOTS], fbits);
}

// Marker m to slot start[key[m]] + rank[m] of its bin, positions and forces
// gathered into `sorted` (six consecutive doubles per slot: x, y, z, fx, fy,
// fz, one 48-byte store); the largest bin count into the stats.  The order inside a bin is that of tile_count_kernel's atomics and
// varies from run to run: the tiles sum integers, so it does not matter.
__global__ void tile_scatter_kernel(marker_sets sets, const int* __restrict__ key, const int* __restrict__ rank,
                                    const int* __restrict__ start, double* __restrict__ sorted,
                                    bin_stats* __restrict__ stats)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const int total = sets.total;
    unsigned long long cnt = 0;
    if (m < total) {
        const int b = key[m];
        const int s0 = start[b];
        const int dst = s0 + rank[m];
        cnt = (unsigned long long) (start[b + 1] - s0);
        const int k = set_of(sets, m);
        const int l = m - sets.offset[k];
        const int n = sets.offset[k + 1] - sets.offset[k];
        const double* x = sets.x[k];
        const double* f = sets.f[k];
        double2* o = (double2*) (sorted + (size_t) 6 * dst);
        o[0] = make_double2(x[l], x[(size_t) n + l]);
        o[1] = make_double2(x[(size_t) 2 * n + l], f[l]);
        o[2] = make_double2(f[(size_t) n + l], f[(size_t) 2 * n + l]);
    }
    cnt = wave_max(cnt);
    if ((threadIdx.x & 63) == 0 && cnt) atomicMax(&stats->max_count[blockIdx.x % STAT_SLOTS], cnt);
}

// The quantum of the fixed-point tiles: q = 2^(e - 62) with e the exponent
// of the bound 64 * max_count * max|f| * max_c res3_c * 0.3 on any point's
// sum; a non-finite bound marks the result as poisoned (all NaN).
struct tile_scale {
    double inv_q, q;
    int poisoned, pad;
};

__global__ void tile_scale_kernel(kgrids g, const bin_stats* __restrict__ stats, tile_scale* __restrict__ out)
{
    const int tid = threadIdx.x;      // one wavefront
    unsigned long long fb = wave_max(stats->max_abs_f[tid]);
    unsigned long long cm = wave_max(stats->max_count[tid]);
    if (tid == 0) {
        double res3 = fmax(g.res3[0], fmax(g.res3[1], g.res3[2]));
        double bound = 64.0 * (double) cm * __longlong_as_double((long long) fb) * res3 * 0.3;
        tile_scale s;
        s.poisoned = !(bound <= 1.7976931348623157e308);
        int e2 = 0;
        if (!s.poisoned && bound > 0) frexp(bound, &e2);
        s.inv_q = ldexp(1.0, 62 - e2);
        s.q = s.poisoned ? __longlong_as_double(0x7ff8000000000000ll) : ldexp(1.0, e2 - 62);
        s.pad = 0;
        *out = s;
    }
}

// One workgroup per brick: lanes are markers, each adds its 3 x 64 stencil
// values to the LDS tile; the tile is then stored (flags[brick] = 0 for a
// brick without markers, whose tile is not written).
__global__ void __launch_bounds__(TILE_THREADS)
tile_spread_kernel(kgrids g, bin_desc bd, tile_desc td, int delta, int total, const int* __restrict__ start,
                   const double* __restrict__ sorted, const tile_scale* __restrict__ scale,
                   unsigned long long* __restrict__ tiles, int* __restrict__ flags)
{
    __shared__ unsigned long long acc[3][TILE_PTS];
    const int brick = blockIdx.x;
    const int s0 = start[(size_t) brick * (TB * TB * TB)];
    const int s1 = start[(size_t) (brick + 1) * (TB * TB * TB)];
    const int tid = threadIdx.x;
    if (s1 == s0) {
        if (tid == 0) flags[brick] = 0;
        return;                                      // whole block
    }
    for (int i = tid; i < 3 * TILE_PTS; i += TILE_THREADS) (&acc[0][0])[i] = 0ull;
    int r = brick % td.bricks;
    int b0 = r % td.nbr[0];
    int b1 = (r / td.nbr[0]) % td.nbr[1];
    int b2 = r / (td.nbr[0] * td.nbr[1]);
    const int bq[3] = {b0 * TB, b1 * TB, b2 * TB};
    const double inv_q = scale->inv_q;
    __syncthreads();
    // Markers are sorted by bin, so neighbours in the sorted order often
    // share a stencil origin, and lanes of one LDS atomic that hit the same
    // address serialise.  Lane l of every wave therefore walks its own
    // segment [l seg, (l + 1) seg) of the brick's markers (the waves
    // interleaved in it): the lanes of one instruction are seg markers apart.
    const int nm = s1 - s0, seg = (nm + 63) / 64, lane = tid & 63;
    for (int r = tid >> 6; r < seg; r += TILE_THREADS / 64) {
        const int li = lane * seg + r;
        if (li >= nm) continue;
        const int i = s0 + li;
        const double2* rec = (const double2*) (sorted + (size_t) 6 * i);
        const double2 r0 = rec[0], r1 = rec[1], r2 = rec[2];
        const double p[3] = {r0.x, r0.y, r1.x}, fv[3] = {r1.y, r2.x, r2.y};
        // bin coordinates of the marker (relative to its brick), as binned
        int lq[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            stencil_axis st = kstencil(g.a[bd.fc[j]][j], p[j]);
            int qb = st.base + bd.fs[j] - bd.qmin[j];
            qb = qb < 0 ? 0 : (qb >= bd.nb[j] ? bd.nb[j] - 1 : qb);
            lq[j] = qb - bq[j];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double w[3][4];
            int o[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const kaxis& a = g.a[c][j];
                stencil_axis st = kstencil(a, p[j]);
                delta4(delta, st.d, w[j]);
                // base_c - q, unwrapped across a periodic seam
                stencil_axis fq = kstencil(g.a[bd.fc[j]][j], p[j]);
                int d = st.base - (fq.base + bd.fs[j]);
                if (!a.solid) {
                    if (d > 2) d -= a.cells;
                    else if (d < -3) d += a.cells;
                }
                o[j] = lq[j] + d - bd.lo[j];
            }
            // A marker outside a walled axis's range (clamped by the index
            // policy) can have an origin off the tile: its points lie outside
            // the grid, which drops them, so it is skipped.
            if (o[0] < 0 || o[0] > TT - 4 || o[1] < 0 || o[1] > TT - 4 || o[2] < 0 || o[2] > TT - 4) continue;
            // The quantum's inverse is a power of two: applied to the force
            // first, it scales every product exactly as it would scale the
            // finished one (a product that ends at 1/2 or more has no
            // subnormal factor on the way), one multiplication fewer per value.
            const double fcq = fv[c] * g.res3[c] * inv_q;
            unsigned long long* t = acc[c];
#pragma unroll
            for (int sz = 0; sz < 4; ++sz) {
#pragma unroll
                for (int sy = 0; sy < 4; ++sy) {
                    const double wyz = w[2][sz] * (w[1][sy] * fcq);
                    const int row = o[0] + TT * ((o[1] + sy) + TT * (o[2] + sz));
#pragma unroll
                    for (int sx = 0; sx < 4; ++sx)
                        atomicAdd(&t[row + sx], (unsigned long long) __double2ll_rn(w[0][sx] * wyz));
                }
            }
        }
    }
    __syncthreads();
    unsigned long long* dst = tiles + (size_t) brick * 3 * TILE_PTS;
    for (int i = tid; i < 3 * TILE_PTS; i += TILE_THREADS) dst[i] = (&acc[0][0])[i];
    if (tid == 0) flags[brick] = 1;
}

// Every grid point of every box and component: the sum of the tile entries
// that cover it, times the quantum.  Block (row, x chunk, box * 3 + c) takes
// MERGE_THREADS points of one x-row (row = y + points_y z) of component c's
// grid, so the y and z coverage is the same for the whole block.
constexpr int MERGE_THREADS = 128;

__global__ void __launch_bounds__(MERGE_THREADS)
tile_merge_kernel(kgrids g, bin_desc bd, tile_desc td, int box0, const unsigned long long* __restrict__ tiles,
                  const int* __restrict__ flags, const tile_scale* __restrict__ scale, double* __restrict__ fe0,
                  double* __restrict__ fe1, double* __restrict__ fe2)
{
    const int bz = blockIdx.z / 3, c = blockIdx.z - 3 * bz, box = box0 + bz;
    const int* pt = g.points[c];
    const int row = blockIdx.x;
    const int x = blockIdx.y * MERGE_THREADS + threadIdx.x;
    if (row >= pt[1] * pt[2] || x >= pt[0]) return;
    const int pp[3] = {x, row % pt[1], row / pt[1]};
    // per axis: up to two (brick, tile index) pairs covering the point
    int nb_c[3], bb[3][2], tt[3][2];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const kaxis& a = g.a[c][j];
        const int t0 = bd.qmin[j] + bd.lo[j] - 1;        // first tile point of brick 0
        int rel = pp[j] - t0;
        int cnt = 0;
        if (!a.solid) {
            // periodic: cells is a multiple of the brick edge here
            rel %= a.cells;
            if (rel < 0) rel += a.cells;
            const int b = rel / TB, t = rel - b * TB;
            bb[j][cnt] = b; tt[j][cnt] = t; ++cnt;
            if (t + TB < TT) {
                bb[j][cnt] = b - 1 < 0 ? td.nbr[j] - 1 : b - 1; tt[j][cnt] = t + TB; ++cnt;
            }
        } else {
            const int b = rel >= 0 ? rel / TB : -1;
            const int t = rel - b * TB;
            if (b >= 0 && b < td.nbr[j]) { bb[j][cnt] = b; tt[j][cnt] = t; ++cnt; }
            if (b - 1 >= 0 && b - 1 < td.nbr[j] && t + TB < TT) { bb[j][cnt] = b - 1; tt[j][cnt] = t + TB; ++cnt; }
        }
        nb_c[j] = cnt;
    }
    long long sum = 0;
    for (int iz = 0; iz < nb_c[2]; ++iz)
        for (int iy = 0; iy < nb_c[1]; ++iy)
            for (int ix = 0; ix < nb_c[0]; ++ix) {
                const int brick = ((box * td.nbr[2] + bb[2][iz]) * td.nbr[1] + bb[1][iy]) * td.nbr[0] + bb[0][ix];
                if (!flags[brick]) continue;
                const size_t at = ((size_t) brick * 3 + c) * TILE_PTS + tt[0][ix] + TT * (tt[1][iy] + TT * tt[2][iz]);
                sum += (long long) tiles[at];
            }
    double* fe = c == 0 ? fe0 : (c == 1 ? fe1 : fe2);
    fe[(size_t) box * td.box_points[c] + (size_t) pt[0] * row + x] = (double) sum * scale->q;
}

// Interpolation over boxes: one lane per marker, grid vectors of box b at
// offset b * box_points[c] (ib/interpolate.h:49-76 per box).
__global__ void __launch_bounds__(256)
interp_boxes_kernel(marker_sets sets, kgrids g, int delta, const int* __restrict__ box_off, int nboxes,
                    const double* __restrict__ u0, const double* __restrict__ u1, const double* __restrict__ u2)
{
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= sets.total) return;
    int k = set_of(sets, m);
    int l = m - sets.offset[k];
    int n = sets.offset[k + 1] - sets.offset[k];
    const int box = box_of(box_off, nboxes, m);
    const double* x = sets.x[k];
    double p[3] = {x[l], x[n + l], x[2 * n + l]};
    const double* ue[3] = {u0, u1, u2};
    for (int c = 0; c < 3; ++c) {
        double w[3][4];
        size_t ix[4], oy[4], oz[4];
        const int* pt = g.points[c];
        for (int j = 0; j < 3; ++j) {
            stencil_axis st = axis_stencil(g.a[c][j], p[j]);
            delta4(delta, st.d, w[j]);
            for (int o = 0; o < 4; ++o) {
                size_t idx = (size_t) axis_interp_index(g.a[c][j], st.base + o - 1);
                if (j == 0) ix[o] = idx;
                else if (j == 1) oy[o] = idx * (size_t) pt[0];
                else oz[o] = idx * (size_t) pt[0] * (size_t) pt[1];
            }
        }
        const double* u = ue[c] + (size_t) box * pt[0] * pt[1] * pt[2];
        double acc = 0.0;
        for (int a = 0; a < 4; ++a) {
            double sy = 0.0;
            for (int b = 0; b < 4; ++b) {
                const double* row = u + oz[a] + oy[b];
                double sx = w[0][0] * row[ix[0]];
                sx = fma(w[0][1], row[ix[1]], sx);
                sx = fma(w[0][2], row[ix[2]], sx);
                sx = fma(w[0][3], row[ix[3]], sx);
                sy = fma(w[1][b], sx, sy);
            }
            acc = fma(w[2][a], sy, acc);
        }
        sets.out[k][(size_t) c * n + l] = acc;
    }
}

} // namespace

// Resolves the device address of `scratch` on `device` (made current by the
// caller) once per device; ibm_create calls it.
ibm_status static_scratch_init(ibm_ctx* ctx, int device)
{
    if (device < 0 || device >= STATIC_MAX_DEVICES)
        return fail(ctx, IBM_ERR_ARG, "device %d beyond the fixed-workspace table", device);
    std::lock_guard<std::mutex> lock(static_scratch_mutex);
    if (static_scratch_ptr[device]) return IBM_OK;
    void* p = nullptr;
    IBM_HIP(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)));
    if (!p) return fail(ctx, IBM_ERR_HIP, "hipGetSymbolAddress returned a null workspace address");
    static_scratch_ptr[device] = p;
    return IBM_OK;
}

ibm_status spread_sets(ibm_ctx* ctx, const ibm_grid* grids, int delta, const marker_sets& sets,
                       double* const* fe, unsigned char* fixed_workspace = nullptr)
{
    kgrids kg;
    IBM_TRY(make_kgrids(ctx, grids, &kg));
    IBM_TRY(check_delta(ctx, delta));
    if (!fe || !fe[0] || !fe[1] || !fe[2]) return fail(ctx, IBM_ERR_ARG, "null output grid");
    for (int c = 0; c < 3; ++c)
        for (int j = 0; j < 3; ++j)
            if (!kg.a[c][j].solid && kg.a[c][j].cells < 4)
                return fail(ctx, IBM_ERR_BAD_GRID, "periodic dimension needs at least 4 cells for a 4-point stencil");
    bin_desc bd = make_bins(grids);
    if (bd.nbins <= 0) return fail(ctx, IBM_ERR_SIZE, "too many grid bins");
    for (int j = 0; j < 3; ++j)
        if (bd.hi[j] - bd.lo[j] > 1 || bd.hi[j] < bd.lo[j])
            return fail(ctx, IBM_ERR_BAD_GRID, "component grids are not a staggering of one cell grid");
    bin_work w;
    if (fixed_workspace) IBM_TRY(static_bin_work(ctx, fixed_workspace, sets.total, bd.nbins, &w));
    else IBM_TRY(ctx_bin_work(ctx, sets.total, bd.nbins, &w));
    IBM_TRY(bin_markers(ctx, sets, kg, bd, 1, w, fixed_workspace == nullptr));
    int maxp[3] = {0, 0, 0};
    for (int c = 0; c < 3; ++c)
        for (int j = 0; j < 3; ++j) maxp[j] = kg.points[c][j] > maxp[j] ? kg.points[c][j] : maxp[j];
    int bx = (maxp[0] + BX - 1) / BX, by = (maxp[1] + BY - 1) / BY, bz = (maxp[2] + BZ - 1) / BZ;
    long long nbricks = (long long) bx * by * bz;
    if (nbricks > 0x7fffffff) return fail(ctx, IBM_ERR_SIZE, "grid too large");
    int xcd_chunk = (nbricks % 8 == 0) ? (int) (nbricks / 8) : 0;
    bool mac = true, fast = true;
    for (int c = 0; c < 3; ++c)
        for (int j = 0; j < 3; ++j) {
            if (kg.sel[c][j] != (c == j ? 1 : 0) || kg.cand[1][j] != j || kg.cand[0][j] == j) mac = false;
            if (!kg.a[c][j].solid && kg.a[c][j].cells < MIN_FAST_CELLS) fast = false;
        }
    if (!fast)
        for (int c = 0; c < 3; ++c)
            for (int j = 0; j < 3; ++j)
                if (kg.a[c][j].points > 1022)
                    return fail(ctx, IBM_ERR_SIZE, "small periodic axes with more than 1022 points on another axis");
    double grid_points = 0;
    for (int c = 0; c < 3; ++c) grid_points += (double) kg.points[c][0] * kg.points[c][1] * kg.points[c][2];
    // positions + forces read once, three grids written once
    prof_scope prof(ctx, fixed_workspace ? nullptr : "spread_brick", 48.0 * sets.total + 8.0 * grid_points);
    auto launch = [&](auto kernel) {
        kernel<<<(int) nbricks, SPREAD_THREADS, 0, ctx->stream>>>(
            kg, bd, delta, sets.total, w.start, w.sorted, w.stats, fe[0], fe[1], fe[2], bx, by, (int) nbricks,
            xcd_chunk);
    };
    if (mac && fast) launch(spread_brick_kernel<true, true>);
    else if (mac) launch(spread_brick_kernel<true, false>);
    else if (fast) launch(spread_brick_kernel<false, true>);
    else launch(spread_brick_kernel<false, false>);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}

ibm_status interp_sets(ibm_ctx* ctx, const ibm_grid* grids, int delta, const marker_sets& sets,
                       const double* const* ue)
{
    kgrids kg;
    IBM_TRY(make_kgrids(ctx, grids, &kg));
    IBM_TRY(check_delta(ctx, delta));
    if (!ue || !ue[0] || !ue[1] || !ue[2]) return fail(ctx, IBM_ERR_ARG, "null input grid");
    if (sets.total == 0) return IBM_OK;
    prof_scope prof(ctx, "interp");
    interp_kernel<<<blocks_for(sets.total, 256), 256, 0, ctx->stream>>>(sets, kg, delta, ue[0], ue[1], ue[2]);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}


// Whether the tile spread takes these grids: 3-dimensional, every periodic
// axis a multiple of the brick edge (so bricks tile it exactly).
bool tiles_fit(const kgrids& kg)
{
    for (int c = 0; c < 3; ++c)
        for (int j = 0; j < 3; ++j)
            if (!kg.a[c][j].solid && kg.a[c][j].cells % TB != 0) return false;
    return true;
}

// Spread of the markers of nboxes boxes (box b: markers [off[b], off[b+1])
// of the sets' concatenation, off on the device) into nboxes consecutive
// copies of each component grid, by bricks and tiles.
ibm_status spread_tiles(ibm_ctx* ctx, const ibm_grid* grids, int delta, const marker_sets& sets, int nboxes,
                        const int* box_off, double* const* fe, bool timed)
{
    kgrids kg;
    IBM_TRY(make_kgrids(ctx, grids, &kg));
    IBM_TRY(check_delta(ctx, delta));
    if (!fe || !fe[0] || !fe[1] || !fe[2]) return fail(ctx, IBM_ERR_ARG, "null output grid");
    if (nboxes < 1) return fail(ctx, IBM_ERR_ARG, "at least one box");
    if (!tiles_fit(kg))
        return fail(ctx, IBM_ERR_BAD_GRID, "batched spreading needs periodic axes of a multiple of %d cells", TB);
    bin_desc bd = make_bins(grids);
    if (bd.nbins <= 0) return fail(ctx, IBM_ERR_SIZE, "too many grid bins");
    for (int j = 0; j < 3; ++j)
        if (bd.hi[j] - bd.lo[j] > 1 || bd.hi[j] < bd.lo[j])
            return fail(ctx, IBM_ERR_BAD_GRID, "component grids are not a staggering of one cell grid");
    tile_desc td{};
    long long bricks = 1;
    for (int j = 0; j < 3; ++j) {
        td.nbr[j] = (bd.nb[j] + TB - 1) / TB;
        bricks *= td.nbr[j];
    }
    long long keys = bricks * nboxes * (TB * TB * TB);
    if (keys > INT_MAX - 1) return fail(ctx, IBM_ERR_SIZE, "too many boxes or bins for one batched spread");
    td.bricks = (int) bricks;
    td.nboxes = nboxes;
    td.keys = (int) keys;
    double grid_points = 0;
    for (int c = 0; c < 3; ++c) {
        td.box_points[c] = kg.points[c][0] * kg.points[c][1] * kg.points[c][2];
        grid_points += (double) td.box_points[c] * nboxes;
    }
    bin_work w;
    IBM_TRY(ctx_bin_work(ctx, sets.total, td.keys, &w));
    void* p;
    const size_t nbricks = (size_t) bricks * nboxes;
    IBM_TRY(scratch(ctx, SLOT_TILES, sizeof(unsigned long long) * 3 * TILE_PTS * nbricks, &p));
    unsigned long long* tiles = (unsigned long long*) p;
    IBM_TRY(scratch(ctx, SLOT_TILE_FLAGS, sizeof(int) * nbricks + 256 + sizeof(tile_scale), &p));
    int* flags = (int*) p;
    tile_scale* sc = (tile_scale*) ((unsigned char*) p + ((sizeof(int) * nbricks + 255) / 256) * 256);
    {
        prof_scope prof(ctx, timed ? "bin" : nullptr);
        size_t words = (size_t) td.keys + 1;
        zero_words_kernel<<<blocks_for((long) std::min<size_t>(words, 1u << 22), 256), 256, 0, ctx->stream>>>(
            (unsigned int*) w.count, words);
        IBM_LAUNCH_CHECK(ctx);
        zero_words_kernel<<<1, 256, 0, ctx->stream>>>((unsigned int*) w.stats, sizeof(bin_stats) / 4);
        IBM_LAUNCH_CHECK(ctx);
        if (sets.total > 0) {
            tile_count_kernel<<<blocks_for(sets.total, 256), 256, 0, ctx->stream>>>(sets, kg, bd, td, box_off, w.key,
                                                                                     w.rank, w.count, w.stats);
            IBM_LAUNCH_CHECK(ctx);
        }
        size_t tbytes = w.scan_bytes;
        IBM_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(w.scan_tmp, tbytes, w.count, w.start, td.keys + 1, ctx->stream));
        if (sets.total > 0) {
            tile_scatter_kernel<<<blocks_for(sets.total, 256), 256, 0, ctx->stream>>>(sets, w.key, w.rank, w.start,
                                                                                       w.sorted, w.stats);
            IBM_LAUNCH_CHECK(ctx);
        }
        tile_scale_kernel<<<1, 64, 0, ctx->stream>>>(kg, w.stats, sc);
        IBM_LAUNCH_CHECK(ctx);
    }
    {
        prof_scope prof(ctx, timed ? "spread_tile" : nullptr, 48.0 * sets.total);
        tile_spread_kernel<<<(int) nbricks, TILE_THREADS, 0, ctx->stream>>>(kg, bd, td, delta, sets.total, w.start,
                                                                            w.sorted, sc, tiles, flags);
        IBM_LAUNCH_CHECK(ctx);
    }
    {
        prof_scope prof(ctx, timed ? "spread_merge" : nullptr, 8.0 * grid_points);
        int rows = 0, xs = 0;
        for (int c = 0; c < 3; ++c) {
            rows = std::max(rows, kg.points[c][1] * kg.points[c][2]);
            xs = std::max(xs, kg.points[c][0]);
        }
        const int xchunks = (xs + MERGE_THREADS - 1) / MERGE_THREADS;
        if (xchunks > 65535) return fail(ctx, IBM_ERR_SIZE, "grid rows of %d points too long for the tile spread", xs);
        constexpr int BOXES_PER_LAUNCH = 65535 / 3;
        for (int b0 = 0; rows > 0 && b0 < td.nboxes; b0 += BOXES_PER_LAUNCH) {
            const int nb = std::min(BOXES_PER_LAUNCH, td.nboxes - b0);
            const dim3 mgrid((unsigned) rows, (unsigned) xchunks, (unsigned) (3 * nb));
            tile_merge_kernel<<<mgrid, MERGE_THREADS, 0, ctx->stream>>>(kg, bd, td, b0, tiles, flags, sc, fe[0], fe[1],
                                                                         fe[2]);
            IBM_LAUNCH_CHECK(ctx);
        }
    }
    return IBM_OK;
}

// The timestep's spread of several marker sets: the tile spread where the
// grids allow it (every periodic axis a multiple of the brick edge), else
// the brick spread.  Both are deterministic; they differ in the last bits
// (their fixed-point scales differ), so one solver always uses one of them.
bool tiles_usable(ibm_ctx* ctx, const ibm_grid* grids)
{
    kgrids kg;
    if (make_kgrids(ctx, grids, &kg) != IBM_OK || !tiles_fit(kg)) return false;
    bin_desc bd = make_bins(grids);
    if (bd.nbins <= 0) return false;
    long long keys = TB * TB * TB;
    for (int j = 0; j < 3; ++j) keys *= (bd.nb[j] + TB - 1) / TB;
    return keys <= INT_MAX - 1;
}

ibm_status spread_sets_step(ibm_ctx* ctx, const ibm_grid* grids, int delta, bool tiles, int nsets, const int* n,
                            const double* const* x, const double* const* f, double* const* fe)
{
    marker_sets s;
    IBM_TRY(make_sets(ctx, nsets, n, x, f, nullptr, &s));
    if (tiles) return spread_tiles(ctx, grids, delta, s, 1, nullptr, fe, true);   // one box: no offsets read
    return spread_sets(ctx, grids, delta, s, fe);
}

// The box offsets (host, nboxes + 1) in the context's scratch.
ibm_status upload_box_offsets(ibm_ctx* ctx, int nboxes, const int* offsets, int total, const int** dev)
{
    if (!offsets) return fail(ctx, IBM_ERR_ARG, "null box offsets");
    if (offsets[0] != 0 || offsets[nboxes] != total)
        return fail(ctx, IBM_ERR_SIZE, "box offsets must run from 0 to the marker count %d", total);
    for (int b = 0; b < nboxes; ++b)
        if (offsets[b + 1] < offsets[b]) return fail(ctx, IBM_ERR_SIZE, "box offsets must not decrease");
    void* p;
    IBM_TRY(scratch(ctx, SLOT_BOX_OFF, sizeof(int) * (size_t) (nboxes + 1), &p));
    IBM_HIP(ctx, hipMemcpyAsync(p, offsets, sizeof(int) * (size_t) (nboxes + 1), hipMemcpyHostToDevice, ctx->stream));
    // the host array may be reused once the call returns
    IBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *dev = (const int*) p;
    return IBM_OK;
}

ibm_status interp_boxes(ibm_ctx* ctx, const ibm_grid* grids, int delta, const marker_sets& sets, int nboxes,
                        const int* box_off, const double* const* ue)
{
    kgrids kg;
    IBM_TRY(make_kgrids(ctx, grids, &kg));
    IBM_TRY(check_delta(ctx, delta));
    if (!ue || !ue[0] || !ue[1] || !ue[2]) return fail(ctx, IBM_ERR_ARG, "null input grid");
    if (sets.total == 0) return IBM_OK;
    prof_scope prof(ctx, "interp");
    interp_boxes_kernel<<<blocks_for(sets.total, 256), 256, 0, ctx->stream>>>(sets, kg, delta, box_off, nboxes, ue[0],
                                                                              ue[1], ue[2]);
    IBM_LAUNCH_CHECK(ctx);
    return IBM_OK;
}
} // namespace ibm

extern "C" {

ibm_status ibm_spread_sets(ibm_ctx* ctx, const ibm_grid* grids, ibm_delta delta, int nsets,
                           const int* n, const double* const* x, const double* const* f,
                           double* const* fe)
{
    if (!ctx) return IBM_ERR_ARG;
    if (nsets > 0 && !f) return ibm::fail(ctx, IBM_ERR_ARG, "null force arrays");
    ibm::marker_sets s;
    IBM_TRY(ibm::make_sets(ctx, nsets, n, x, f, nullptr, &s));
    return ibm::spread_sets(ctx, grids, (int) delta, s, fe);
}

ibm_status ibm_spread_sets_noalloc(ibm_ctx* ctx, const ibm_grid* grids, ibm_delta delta, int nsets,
                                   const int* n, const double* const* x, const double* const* f,
                                   double* const* fe)
{
    if (!ctx) return IBM_ERR_ARG;
    if (nsets > 0 && !f) return ibm::fail(ctx, IBM_ERR_ARG, "null force arrays");
    if (ctx->device < 0 || ctx->device >= ibm::STATIC_MAX_DEVICES || !ibm::static_scratch_ptr[ctx->device])
        return ibm::fail(ctx, IBM_ERR_HIP, "the fixed spreading workspace of device %d is not set up", ctx->device);
    ibm::marker_sets s;
    IBM_TRY(ibm::make_sets(ctx, nsets, n, x, f, nullptr, &s));
    return ibm::spread_sets(ctx, grids, (int) delta, s, fe, (unsigned char*) ibm::static_scratch_ptr[ctx->device]);
}

ibm_status ibm_spread(ibm_ctx* ctx, const ibm_grid* grids, ibm_delta delta, int n,
                      const double* x, const double* f, double* const* fe)
{
    return ibm_spread_sets(ctx, grids, delta, 1, &n, &x, &f, fe);
}

ibm_status ibm_interpolate_sets(ibm_ctx* ctx, const ibm_grid* grids, ibm_delta delta, int nsets,
                                const int* n, const double* const* x, const double* const* ue,
                                double* const* ul)
{
    if (!ctx) return IBM_ERR_ARG;
    if (nsets > 0 && !ul) return ibm::fail(ctx, IBM_ERR_ARG, "null output arrays");
    ibm::marker_sets s;
    IBM_TRY(ibm::make_sets(ctx, nsets, n, x, nullptr, ul, &s));
    return ibm::interp_sets(ctx, grids, (int) delta, s, ue);
}

ibm_status ibm_interpolate(ibm_ctx* ctx, const ibm_grid* grids, ibm_delta delta, int n,
                           const double* x, const double* const* ue, double* ul)
{
    return ibm_interpolate_sets(ctx, grids, delta, 1, &n, &x, ue, &ul);
}

ibm_status ibm_spread_boxes(ibm_ctx* ctx, const ibm_grid* grids, ibm_delta delta, int nboxes, const int* offsets,
                            const double* x, const double* f, double* const* fe)
{
    if (!ctx) return IBM_ERR_ARG;
    if (nboxes < 1) return ibm::fail(ctx, IBM_ERR_ARG, "at least one box");
    if (!offsets) return ibm::fail(ctx, IBM_ERR_ARG, "null box offsets");
    int n = offsets[nboxes];
    ibm::marker_sets s;
    IBM_TRY(ibm::make_sets(ctx, 1, &n, &x, &f, nullptr, &s));
    const int* dev_off = nullptr;
    IBM_TRY(ibm::upload_box_offsets(ctx, nboxes, offsets, s.total, &dev_off));
    return ibm::spread_tiles(ctx, grids, (int) delta, s, nboxes, dev_off, fe, true);
}

ibm_status ibm_interpolate_boxes(ibm_ctx* ctx, const ibm_grid* grids, ibm_delta delta, int nboxes, const int* offsets,
                                 const double* x, const double* const* ue, double* ul)
{
    if (!ctx) return IBM_ERR_ARG;
    if (nboxes < 1) return ibm::fail(ctx, IBM_ERR_ARG, "at least one box");
    if (!offsets) return ibm::fail(ctx, IBM_ERR_ARG, "null box offsets");
    int n = offsets[nboxes];
    ibm::marker_sets s;
    IBM_TRY(ibm::make_sets(ctx, 1, &n, &x, nullptr, &ul, &s));
    const int* dev_off = nullptr;
    IBM_TRY(ibm::upload_box_offsets(ctx, nboxes, offsets, s.total, &dev_off));
    return ibm::interp_boxes(ctx, grids, (int) delta, s, nboxes, dev_off, ue);
}

} // extern "C"
