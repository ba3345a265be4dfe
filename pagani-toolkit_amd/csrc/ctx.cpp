// ctx.cpp -- pg_ctx lifetime, errors, device arena, pointer classification.
#include "pg_internal.hpp"

#include <cstdlib>
#include <cstring>
#include <string>

pg_status pg_fail(pg_ctx *ctx, pg_status st, const char *fmt, ...)
{
    if (ctx) {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return st;
}

pg_status pg_buf(pg_ctx *ctx, const char *name, size_t bytes, void **out)
{
    PgBuf &b = ctx->bufs[name];
    if (b.cap < bytes) {
        if (b.p) {
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipFree(b.p);
            b.p = nullptr;
            b.cap = 0;
        }
        size_t want = bytes + bytes / 8;  // headroom for slightly larger repeats
        hipError_t e = hipMalloc(&b.p, want);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            e = hipMalloc(&b.p, bytes);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                b.p = nullptr;
                return pg_fail(ctx, PG_ENOMEM, "hipMalloc(%zu) for '%s' failed", bytes, name);
            }
            want = bytes;
        }
        b.cap = want;
    }
    *out = b.p;
    return PG_OK;
}

bool pg_is_device_ptr(const void *p)
{
    if (!p) return false;
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

pg_status pg_check_csr_args(pg_ctx *ctx, const int32_t *row, const int32_t *col, int32_t n,
                            int64_t nnz)
{
    if (!ctx) return PG_EINVAL;
    if (n < 0 || nnz < 0 || !row || (nnz > 0 && !col))
        return pg_fail(ctx, PG_EINVAL, "bad CSR arguments (n=%d nnz=%lld)", n, (long long)nnz);
    if (nnz > INT32_MAX) return pg_fail(ctx, PG_ENOTSUP, "nnz %lld exceeds int32", (long long)nnz);
    return PG_OK;
}

extern "C" {

int pg_abi_version(void) { return PAGANI_ABI_VERSION; }

const char *pg_status_string(pg_status s)
{
    switch (s) {
    case PG_OK: return "ok";
    case PG_EINVAL: return "invalid argument";
    case PG_ENOMEM: return "out of device memory";
    case PG_EDEVICE: return "device error";
    case PG_EFORMAT: return "malformed input";
    case PG_ETHRESH: return "threshold error";
    case PG_ENOTSUP: return "not supported";
    case PG_ESTATE: return "bad call order";
    }
    return "unknown status";
}

const char *pg_last_error(const pg_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

pg_status pg_create(pg_ctx **out, int device)
{
    if (!out) return PG_EINVAL;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return PG_EDEVICE;
    }
    if (device < 0 || device >= count) return PG_EINVAL;
    pg_ctx *ctx = new pg_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete ctx;
        return PG_EDEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        ctx->num_cu = prop.multiProcessorCount;
        ctx->lds_per_block = prop.sharedMemPerBlock;
        const std::string a = prop.gcnArchName;
        ctx->arch = a.substr(0, a.find(':'));
    }
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return PG_EDEVICE;
    }
    for (auto &e : ctx->ev) (void)hipEventCreate(&e);
    if (hipHostMalloc(&ctx->h_report, PG_REPORT_BYTES, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        ctx->h_report = nullptr;
        pg_destroy(ctx);
        return PG_ENOMEM;
    }
    // the allocation-free entry points' scratch address on this device
    if (pg_static_scratch(ctx, &ctx->static_scratch) != PG_OK) {
        pg_destroy(ctx);
        return PG_EDEVICE;
    }
    *out = ctx;
    return PG_OK;
}

void pg_destroy(pg_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    pg_emit_free(ctx->emit);
    for (auto &kv : ctx->bufs)
        if (kv.second.p) (void)hipFree(kv.second.p);
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->h_report) (void)hipHostFree(ctx->h_report);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

pg_status pg_synchronize(pg_ctx *ctx)
{
    if (!ctx) return PG_EINVAL;
    PG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PG_OK;
}

pg_status pg_set_timing(pg_ctx *ctx, int32_t level)
{
    if (!ctx) return PG_EINVAL;
    if (level < 0 || level > 2) return pg_fail(ctx, PG_EINVAL, "timing level %d outside 0..2", (int)level);
    ctx->timing = level;
    return PG_OK;
}

}  // extern "C"
