// Checkpoint / resume in the reference's linear-algebra binary format.
//
// Format (linalg/io.h:67-99 magic numbers, :416-531 writers and readers):
// every dense object is a 4-byte magic word, three int32 header words and
// the raw values, all in host byte order:
//     magic = container | layout | value type
//             container: vector 0x65760000 ("ve"), matrix 0x616d0000 ("ma")
//             layout:    dense  0x00006400 ("d"),  sparse 0x00007300 ("s")
//             value:     double 0x00000064 ("d"),  float  0x00000073 ("s")
//     dense vector: {rows, 1, rows}, then rows values       (:421-430)
//     dense matrix: {rows, cols, rows*cols}, column-major   (:445-454)
// A scalar such as a state's time is its raw 8 bytes with no header
// (:414-419).  An ib::state record is t, u[0..d), p, x[0..n) in that order,
// the order of ins::state's writer (ins/state.h:48-68) extended by the
// cell-type matrices as util/kaboose.cc:21-31 reads them.
//
// Device data is staged through one host buffer per call; the files are
// written and read with stdio.
#include <cstdio>
#include <cstring>
#include <vector>
#include "context.hpp"

namespace ibm {
namespace {

constexpr int MAGIC_VECTOR = 0x65760000;
constexpr int MAGIC_MATRIX = 0x616d0000;
constexpr int MAGIC_DENSE = 0x00006400;
constexpr int MAGIC_DOUBLE = 0x00000064;

struct file_closer {
    FILE* f;
    ~file_closer() { if (f) fclose(f); }
};

int magic_of(int is_matrix) { return (is_matrix ? MAGIC_MATRIX : MAGIC_VECTOR) | MAGIC_DENSE | MAGIC_DOUBLE; }

ibm_status write_dense(ibm_ctx* ctx, FILE* f, int is_matrix, long long rows, long long cols, const double* dev,
                       std::vector<double>& host)
{
    if (rows < 0 || cols < 0 || (!is_matrix && cols != 1))
        return fail(ctx, IBM_ERR_SIZE, "bad dense object shape %lld x %lld", rows, cols);
    long long n = rows * cols;
    if (rows > 0x7fffffffLL || cols > 0x7fffffffLL || n > 0x7fffffffLL)
        return fail(ctx, IBM_ERR_SIZE, "the binary format holds at most 2^31-1 values per object");
    if (n > 0 && !dev) return fail(ctx, IBM_ERR_ARG, "null device data");
    host.resize((size_t) n);
    if (n > 0) {
        IBM_HIP(ctx, hipMemcpyAsync(host.data(), dev, sizeof(double) * (size_t) n, hipMemcpyDeviceToHost, ctx->stream));
        IBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    int head[4] = {magic_of(is_matrix), (int) rows, (int) cols, (int) n};
    if (fwrite(head, sizeof(int), 4, f) != 4 || (n > 0 && fwrite(host.data(), sizeof(double), (size_t) n, f) != (size_t) n))
        return fail(ctx, IBM_ERR_ARG, "write failed");
    return IBM_OK;
}

// Reads one dense object; with dev == nullptr only skips it.  The shape
// must be the expected one.
ibm_status read_dense(ibm_ctx* ctx, FILE* f, int is_matrix, long long rows, long long cols, double* dev,
                      std::vector<double>& host, bool* eof)
{
    int head[4];
    size_t got = fread(head, sizeof(int), 4, f);
    if (got == 0 && feof(f)) { *eof = true; return IBM_OK; }
    if (got != 4) { *eof = true; return IBM_OK; }   // truncated record: treated as the end of the file
    if (head[0] != magic_of(is_matrix))
        return fail(ctx, IBM_ERR_ARG, "bad magic 0x%08x (expected a dense double %s, 0x%08x)", head[0],
                    is_matrix ? "matrix" : "vector", magic_of(is_matrix));
    if (head[1] != rows || head[2] != cols || (long long) head[3] != rows * cols)
        return fail(ctx, IBM_ERR_SIZE, "object of %d x %d (%d values) where %lld x %lld was expected", head[1],
                    head[2], head[3], rows, cols);
    size_t n = (size_t) head[3];
    if (!dev) {
        if (fseek(f, (long) (sizeof(double) * n), SEEK_CUR) != 0) { *eof = true; return IBM_OK; }
        // fseek past the end succeeds; the next read reports the truncation
        return IBM_OK;
    }
    host.resize(n);
    if (n > 0 && fread(host.data(), sizeof(double), n, f) != n) { *eof = true; return IBM_OK; }
    if (n > 0) {
        IBM_HIP(ctx, hipMemcpyAsync(dev, host.data(), sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
        IBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return IBM_OK;
}

ibm_status check_layout(ibm_ctx* ctx, const ibm_state_layout* L)
{
    if (!L) return fail(ctx, IBM_ERR_ARG, "null state layout");
    if (L->dims < 1 || L->dims > 3) return fail(ctx, IBM_ERR_ARG, "state layout: dims must be 1..3");
    if (L->ntypes < 0 || L->ntypes > IBM_MAX_CELL_TYPES)
        return fail(ctx, IBM_ERR_ARG, "state layout: ntypes must be in [0, %d]", IBM_MAX_CELL_TYPES);
    for (int i = 0; i < L->dims; ++i)
        if (L->u_size[i] < 0) return fail(ctx, IBM_ERR_SIZE, "negative velocity size");
    if (L->p_size < 0) return fail(ctx, IBM_ERR_SIZE, "negative pressure size");
    for (int k = 0; k < L->ntypes; ++k)
        if (L->rows[k] < 0 || L->cols[k] < 0) return fail(ctx, IBM_ERR_SIZE, "negative cell matrix shape");
    return IBM_OK;
}

// Reads one record into st (or skips it with st == nullptr).  *complete is
// false at the end of the file or on a truncated record.
ibm_status read_record(ibm_ctx* ctx, FILE* f, const ibm_state_layout* L, ibm_state* st, std::vector<double>& host,
                       bool* complete)
{
    *complete = false;
    double t;
    if (fread(&t, sizeof(double), 1, f) != 1) return IBM_OK;
    bool eof = false;
    for (int i = 0; i < L->dims && !eof; ++i)
        IBM_TRY(read_dense(ctx, f, 0, L->u_size[i], 1, st ? st->u[i] : nullptr, host, &eof));
    if (!eof) IBM_TRY(read_dense(ctx, f, 0, L->p_size, 1, st ? st->p : nullptr, host, &eof));
    for (int k = 0; k < L->ntypes && !eof; ++k)
        IBM_TRY(read_dense(ctx, f, 1, L->rows[k], L->cols[k], st ? st->x[k] : nullptr, host, &eof));
    if (eof) return IBM_OK;
    if (!st) {
        // a skipped record is complete only if its last value exists
        long pos = ftell(f);
        if (fseek(f, 0, SEEK_END) != 0) return fail(ctx, IBM_ERR_ARG, "seek failed");
        long end = ftell(f);
        if (pos > end) return IBM_OK;
        if (fseek(f, pos, SEEK_SET) != 0) return fail(ctx, IBM_ERR_ARG, "seek failed");
    } else {
        st->t = t;
    }
    *complete = true;
    return IBM_OK;
}

} // namespace
} // namespace ibm

extern "C" {

ibm_status ibm_io_write_dense(ibm_ctx* ctx, const char* path, int append, int is_matrix, long long rows,
                              long long cols, const double* values)
{
    if (!ctx) return IBM_ERR_ARG;
    if (!path) return ibm::fail(ctx, IBM_ERR_ARG, "null path");
    ibm::file_closer fc{fopen(path, append ? "ab" : "wb")};
    if (!fc.f) return ibm::fail(ctx, IBM_ERR_ARG, "cannot open %s for writing", path);
    std::vector<double> host;
    IBM_TRY(ibm::write_dense(ctx, fc.f, is_matrix, rows, is_matrix ? cols : 1, values, host));
    if (fflush(fc.f) != 0) return ibm::fail(ctx, IBM_ERR_ARG, "write failed on %s", path);
    return IBM_OK;
}

ibm_status ibm_io_read_dense(ibm_ctx* ctx, const char* path, long long offset, int is_matrix, long long rows,
                             long long cols, double* values, long long* next)
{
    if (!ctx) return IBM_ERR_ARG;
    if (!path) return ibm::fail(ctx, IBM_ERR_ARG, "null path");
    ibm::file_closer fc{fopen(path, "rb")};
    if (!fc.f) return ibm::fail(ctx, IBM_ERR_ARG, "cannot open %s", path);
    if (offset < 0 || fseek(fc.f, (long) offset, SEEK_SET) != 0) return ibm::fail(ctx, IBM_ERR_ARG, "bad offset");
    std::vector<double> host;
    bool eof = false;
    IBM_TRY(ibm::read_dense(ctx, fc.f, is_matrix, rows, is_matrix ? cols : 1, values, host, &eof));
    if (eof) return ibm::fail(ctx, IBM_ERR_SIZE, "%s ends before the object at offset %lld", path, offset);
    if (next) *next = (long long) ftell(fc.f);
    return IBM_OK;
}

ibm_status ibm_io_peek_dense(ibm_ctx* ctx, const char* path, long long offset, int* magic, int* rows, int* cols)
{
    if (!ctx) return IBM_ERR_ARG;
    if (!path || !magic || !rows || !cols) return ibm::fail(ctx, IBM_ERR_ARG, "null argument");
    ibm::file_closer fc{fopen(path, "rb")};
    if (!fc.f) return ibm::fail(ctx, IBM_ERR_ARG, "cannot open %s", path);
    if (offset < 0 || fseek(fc.f, (long) offset, SEEK_SET) != 0) return ibm::fail(ctx, IBM_ERR_ARG, "bad offset");
    int head[4];
    if (fread(head, sizeof(int), 4, fc.f) != 4) return ibm::fail(ctx, IBM_ERR_SIZE, "no object header at %lld", offset);
    *magic = head[0];
    *rows = head[1];
    *cols = head[2];
    return IBM_OK;
}

ibm_status ibm_state_write(ibm_ctx* ctx, const char* path, int append, const ibm_state* st,
                           const ibm_state_layout* layout)
{
    if (!ctx) return IBM_ERR_ARG;
    if (!path || !st) return ibm::fail(ctx, IBM_ERR_ARG, "null argument");
    IBM_TRY(ibm::check_layout(ctx, layout));
    ibm::file_closer fc{fopen(path, append ? "ab" : "wb")};
    if (!fc.f) return ibm::fail(ctx, IBM_ERR_ARG, "cannot open %s for writing", path);
    std::vector<double> host;
    double t = st->t;
    if (fwrite(&t, sizeof(double), 1, fc.f) != 1) return ibm::fail(ctx, IBM_ERR_ARG, "write failed on %s", path);
    for (int i = 0; i < layout->dims; ++i)
        IBM_TRY(ibm::write_dense(ctx, fc.f, 0, layout->u_size[i], 1, st->u[i], host));
    IBM_TRY(ibm::write_dense(ctx, fc.f, 0, layout->p_size, 1, st->p, host));
    for (int k = 0; k < layout->ntypes; ++k)
        IBM_TRY(ibm::write_dense(ctx, fc.f, 1, layout->rows[k], layout->cols[k], st->x[k], host));
    if (fflush(fc.f) != 0) return ibm::fail(ctx, IBM_ERR_ARG, "write failed on %s", path);
    return IBM_OK;
}

ibm_status ibm_state_read(ibm_ctx* ctx, const char* path, int record, ibm_state* st,
                          const ibm_state_layout* layout, int* records)
{
    if (!ctx) return IBM_ERR_ARG;
    if (!path) return ibm::fail(ctx, IBM_ERR_ARG, "null path");
    IBM_TRY(ibm::check_layout(ctx, layout));
    ibm::file_closer fc{fopen(path, "rb")};
    if (!fc.f) return ibm::fail(ctx, IBM_ERR_ARG, "cannot open %s", path);
    std::vector<double> host;
    // count the complete records, remembering where each starts
    std::vector<long> starts;
    for (;;) {
        long pos = ftell(fc.f);
        bool complete = false;
        IBM_TRY(ibm::read_record(ctx, fc.f, layout, nullptr, host, &complete));
        if (!complete) break;
        starts.push_back(pos);
    }
    if (records) *records = (int) starts.size();
    if (!st) return IBM_OK;
    int n = (int) starts.size();
    int r = record < 0 ? n + record : record;
    if (r < 0 || r >= n) return ibm::fail(ctx, IBM_ERR_SIZE, "%s holds %d complete states; record %d asked", path, n, record);
    if (fseek(fc.f, starts[r], SEEK_SET) != 0) return ibm::fail(ctx, IBM_ERR_ARG, "seek failed");
    bool complete = false;
    IBM_TRY(ibm::read_record(ctx, fc.f, layout, st, host, &complete));
    if (!complete) return ibm::fail(ctx, IBM_ERR_ARG, "record %d of %s could not be read", r, path);
    return IBM_OK;
}

} // extern "C"
