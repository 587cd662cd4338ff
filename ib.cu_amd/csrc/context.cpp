// Context lifetime, error reporting and scratch management.
#include "context.hpp"

namespace ibm {

ibm_status fail(ibm_ctx* ctx, ibm_status code, const char* fmt, ...)
{
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->error = buf;
    }
    return code;
}

ibm_status hip_fail(ibm_ctx* ctx, hipError_t e, const char* what)
{
    return fail(ctx, IBM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

ibm_status scratch(ibm_ctx* ctx, int slot, size_t bytes, void** out)
{
    if (slot < 0 || slot >= SLOT_COUNT) return fail(ctx, IBM_ERR_ARG, "bad scratch slot %d", slot);
    if (ctx->scratch.size() < SLOT_COUNT) ctx->scratch.resize(SLOT_COUNT);
    auto& b = ctx->scratch[slot];
    if (bytes == 0) bytes = 256;
    if (b.bytes < bytes) {
        if (b.ptr) {
            // The old buffer may still be read by queued work.
            IBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
            IBM_HIP(ctx, hipFree(b.ptr));
            b.ptr = nullptr;
            b.bytes = 0;
        }
        size_t grown = bytes + bytes / 4;
        IBM_HIP(ctx, hipMalloc(&b.ptr, grown));
        b.bytes = grown;
    }
    *out = b.ptr;
    return IBM_OK;
}

prof_scope::prof_scope(ibm_ctx* c, const char* name, double work) : ctx(c)
{
    if (!ctx->profiling || !name) return;
    if (ctx->prof_used == ctx->prof.size()) {
        ibm_ctx::prof_record r;
        r.name = name;
        if (hipEventCreate(&r.start) != hipSuccess) return;
        if (hipEventCreate(&r.stop) != hipSuccess) { hipEventDestroy(r.start); return; }
        ctx->prof.push_back(r);
    }
    index = (int) ctx->prof_used++;
    ctx->prof[index].name = name;
    ctx->prof[index].work = work;
    hipEventRecord(ctx->prof[index].start, ctx->stream);
}

prof_scope::~prof_scope()
{
    if (index >= 0) hipEventRecord(ctx->prof[index].stop, ctx->stream);
}

} // namespace ibm

extern "C" {

int ibm_abi_version(void) { return IBM_ABI_VERSION; }

ibm_status ibm_profile_enable(ibm_ctx* ctx, int on)
{
    if (!ctx) return IBM_ERR_ARG;
    ctx->profiling = on != 0;
    ctx->prof_used = 0;
    return IBM_OK;
}

ibm_status ibm_profile_read(ibm_ctx* ctx, const char* name, double* total_ms, int* launches)
{
    if (!ctx || !name || !total_ms || !launches) return IBM_ERR_ARG;
    IBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double t = 0;
    int k = 0;
    for (size_t i = 0; i < ctx->prof_used; ++i) {
        if (std::string(ctx->prof[i].name) != name) continue;
        float ms = 0;
        IBM_HIP(ctx, hipEventElapsedTime(&ms, ctx->prof[i].start, ctx->prof[i].stop));
        t += ms;
        ++k;
    }
    *total_ms = t;
    *launches = k;
    return IBM_OK;
}

ibm_status ibm_profile_work(ibm_ctx* ctx, const char* name, double* total)
{
    if (!ctx || !name || !total) return IBM_ERR_ARG;
    double w = 0;
    for (size_t i = 0; i < ctx->prof_used; ++i)
        if (std::string(ctx->prof[i].name) == name) w += ctx->prof[i].work;
    *total = w;
    return IBM_OK;
}

ibm_status ibm_create(ibm_ctx** out, int device)
{
    if (!out) return IBM_ERR_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
        return IBM_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return IBM_ERR_HIP;
    auto* ctx = new ibm_ctx;
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->cu_count = prop.multiProcessorCount;
    if (hipHostMalloc((void**) &ctx->host_flag, sizeof(int) * 16, hipHostMallocDefault) != hipSuccess) {
        delete ctx;
        return IBM_ERR_HIP;
    }
    ctx->scratch.resize(ibm::SLOT_COUNT);
    if (ibm::static_scratch_init(ctx, device) != IBM_OK) {
        hipHostFree(ctx->host_flag);
        delete ctx;
        return IBM_ERR_HIP;
    }
    *out = ctx;
    return IBM_OK;
}

ibm_status ibm_destroy(ibm_ctx* ctx)
{
    if (!ctx) return IBM_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for (auto& b : ctx->scratch)
        if (b.ptr) hipFree(b.ptr);
    for (auto& r : ctx->prof) {
        hipEventDestroy(r.start);
        hipEventDestroy(r.stop);
    }
    if (ctx->host_flag) hipHostFree(ctx->host_flag);
    delete ctx;
    return IBM_OK;
}

ibm_status ibm_set_stream(ibm_ctx* ctx, void* stream)
{
    if (!ctx) return IBM_ERR_ARG;
    ctx->stream = (hipStream_t) stream;
    return IBM_OK;
}

ibm_status ibm_synchronize(ibm_ctx* ctx)
{
    if (!ctx) return IBM_ERR_ARG;
    IBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return IBM_OK;
}

const char* ibm_last_error(const ibm_ctx* ctx)
{
    return ctx ? ctx->error.c_str() : "null context";
}

} // extern "C"
