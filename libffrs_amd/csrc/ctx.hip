// ctx.hip -- context, device memory and timing entry points of the C ABI.
#include "common.hpp"

#include <cstring>

namespace ffrs {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

int fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int ctx_set_device(ffrs_ctx *ctx)
{
    FFRS_HIP(hipSetDevice(ctx->device));
    return FFRS_OK;
}

int stage_reserve(ffrs_ctx *ctx, int i, size_t bytes, void **out)
{
    DevBuf &b = ctx->stage[i];
    size_t want = bytes + 64;
    if (b.size < want) {
        if (b.ptr)
            FFRS_HIP(hipFree(b.ptr));
        b.ptr = nullptr;
        b.size = 0;
        size_t sz = want < (1u << 20) ? (1u << 20) : want + want / 4;
        hipError_t e = hipMalloc(&b.ptr, sz);
        if (e != hipSuccess) {
            b.ptr = nullptr;
            return fail(FFRS_ERR_NOMEM, "device staging allocation of %zu bytes failed: %s", sz,
                        hipGetErrorString(e));
        }
        b.size = sz;
    }
    *out = b.ptr;
    return FFRS_OK;
}

}  // namespace ffrs

using namespace ffrs;

extern "C" {

int ffrs_abi_version(void) { return FFRS_ABI_VERSION; }

const char *ffrs_last_error(void) { return g_last_error.c_str(); }

int ffrs_ctx_create(int device, ffrs_ctx **out)
{
    if (!out)
        return fail(FFRS_ERR_INVALID, "out is NULL");
    int n = 0;
    FFRS_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return fail(FFRS_ERR_INVALID, "device %d out of range (%d devices)", device, n);
    auto *ctx = new ffrs_ctx();
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&ctx->cu_count, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipEventCreate(&ctx->ev_begin);
    if (e == hipSuccess)
        e = hipEventCreate(&ctx->ev_end);
    if (e != hipSuccess) {
        delete ctx;
        return fail(FFRS_ERR_DEVICE, "context creation failed: %s", hipGetErrorString(e));
    }
    *out = ctx;
    return FFRS_OK;
}

void ffrs_ctx_destroy(ffrs_ctx *ctx)
{
    if (!ctx)
        return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &b : ctx->stage)
        if (b.ptr)
            (void)hipFree(b.ptr);
    if (ctx->ev_begin)
        (void)hipEventDestroy(ctx->ev_begin);
    if (ctx->ev_end)
        (void)hipEventDestroy(ctx->ev_end);
    for (auto &e : ctx->marks)
        if (e)
            (void)hipEventDestroy(e);
    if (ctx->stream)
        (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int ffrs_ctx_synchronize(ffrs_ctx *ctx)
{
    if (!ctx)
        return fail(FFRS_ERR_INVALID, "NULL context");
    FFRS_HIP(hipSetDevice(ctx->device));
    FFRS_HIP(hipStreamSynchronize(ctx->stream));
    return FFRS_OK;
}

int ffrs_ctx_device_info(ffrs_ctx *ctx, char *name, size_t name_cap, int *cu_count, size_t *free_bytes,
                         size_t *total_bytes)
{
    if (!ctx)
        return fail(FFRS_ERR_INVALID, "NULL context");
    FFRS_HIP(hipSetDevice(ctx->device));
    hipDeviceProp_t prop;
    FFRS_HIP(hipGetDeviceProperties(&prop, ctx->device));
    if (name && name_cap) {
        if (prop.name[0])
            snprintf(name, name_cap, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
        else
            snprintf(name, name_cap, "%s (%d CUs)", prop.gcnArchName, prop.multiProcessorCount);
    }
    if (cu_count)
        *cu_count = ctx->cu_count;
    size_t f = 0, t = 0;
    FFRS_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes)
        *free_bytes = f;
    if (total_bytes)
        *total_bytes = t;
    return FFRS_OK;
}

int ffrs_device_alloc(ffrs_ctx *ctx, size_t bytes, void **d_ptr)
{
    if (!ctx || !d_ptr)
        return fail(FFRS_ERR_INVALID, "NULL argument");
    FFRS_HIP(hipSetDevice(ctx->device));
    hipError_t e = hipMalloc(d_ptr, bytes ? bytes : 1);
    if (e != hipSuccess)
        return fail(FFRS_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return FFRS_OK;
}

int ffrs_host_alloc(ffrs_ctx *ctx, size_t bytes, void **h_ptr)
{
    if (!ctx || !h_ptr)
        return fail(FFRS_ERR_INVALID, "NULL argument");
    FFRS_HIP(hipSetDevice(ctx->device));
    hipError_t e = hipHostMalloc(h_ptr, bytes ? bytes : 1, hipHostMallocDefault);
    if (e != hipSuccess)
        return fail(FFRS_ERR_NOMEM, "hipHostMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return FFRS_OK;
}

int ffrs_host_free(ffrs_ctx *ctx, void *h_ptr)
{
    if (!ctx)
        return fail(FFRS_ERR_INVALID, "NULL context");
    FFRS_HIP(hipSetDevice(ctx->device));
    FFRS_HIP(hipHostFree(h_ptr));
    return FFRS_OK;
}

int ffrs_memcpy_h2d_async(ffrs_ctx *ctx, void *d_dst, const void *h_src, size_t bytes)
{
    if (!ctx)
        return fail(FFRS_ERR_INVALID, "NULL context");
    FFRS_HIP(hipSetDevice(ctx->device));
    FFRS_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return FFRS_OK;
}

int ffrs_memcpy_d2h_async(ffrs_ctx *ctx, void *h_dst, const void *d_src, size_t bytes)
{
    if (!ctx)
        return fail(FFRS_ERR_INVALID, "NULL context");
    FFRS_HIP(hipSetDevice(ctx->device));
    FFRS_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    return FFRS_OK;
}

int ffrs_device_free(ffrs_ctx *ctx, void *d_ptr)
{
    if (!ctx)
        return fail(FFRS_ERR_INVALID, "NULL context");
    FFRS_HIP(hipSetDevice(ctx->device));
    FFRS_HIP(hipFree(d_ptr));
    return FFRS_OK;
}

int ffrs_memcpy_h2d(ffrs_ctx *ctx, void *d_dst, const void *h_src, size_t bytes)
{
    if (!ctx)
        return fail(FFRS_ERR_INVALID, "NULL context");
    FFRS_HIP(hipSetDevice(ctx->device));
    FFRS_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    FFRS_HIP(hipStreamSynchronize(ctx->stream));
    return FFRS_OK;
}

int ffrs_memcpy_d2h(ffrs_ctx *ctx, void *h_dst, const void *d_src, size_t bytes)
{
    if (!ctx)
        return fail(FFRS_ERR_INVALID, "NULL context");
    FFRS_HIP(hipSetDevice(ctx->device));
    FFRS_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipStreamSynchronize(ctx->stream));
    return FFRS_OK;
}

int ffrs_memcpy_d2d(ffrs_ctx *ctx, void *d_dst, const void *d_src, size_t bytes)
{
    if (!ctx)
        return fail(FFRS_ERR_INVALID, "NULL context");
    FFRS_HIP(hipSetDevice(ctx->device));
    FFRS_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return FFRS_OK;
}

int ffrs_timer_begin(ffrs_ctx *ctx)
{
    if (!ctx)
        return fail(FFRS_ERR_INVALID, "NULL context");
    FFRS_HIP(hipSetDevice(ctx->device));
    FFRS_HIP(hipEventRecord(ctx->ev_begin, ctx->stream));
    return FFRS_OK;
}

int ffrs_timer_end(ffrs_ctx *ctx, float *ms)
{
    if (!ctx || !ms)
        return fail(FFRS_ERR_INVALID, "NULL argument");
    FFRS_HIP(hipSetDevice(ctx->device));
    FFRS_HIP(hipEventRecord(ctx->ev_end, ctx->stream));
    FFRS_HIP(hipEventSynchronize(ctx->ev_end));
    FFRS_HIP(hipEventElapsedTime(ms, ctx->ev_begin, ctx->ev_end));
    return FFRS_OK;
}

struct ffrs_graph {
    ffrs_ctx *ctx;
    hipGraphExec_t exec;
};

int ffrs_ctx_capture_begin(ffrs_ctx *ctx)
{
    if (!ctx)
        return fail(FFRS_ERR_INVALID, "NULL context");
    FFRS_HIP(hipSetDevice(ctx->device));
    FFRS_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    return FFRS_OK;
}

int ffrs_ctx_capture_end(ffrs_ctx *ctx, ffrs_graph **out)
{
    if (!ctx || !out)
        return fail(FFRS_ERR_INVALID, "NULL argument");
    FFRS_HIP(hipSetDevice(ctx->device));
    hipGraph_t graph = nullptr;
    FFRS_HIP(hipStreamEndCapture(ctx->stream, &graph));
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess)
        return fail(FFRS_ERR_DEVICE, "graph instantiation failed: %s", hipGetErrorString(e));
    *out = new ffrs_graph{ctx, exec};
    return FFRS_OK;
}

int ffrs_graph_launch(ffrs_graph *g)
{
    if (!g)
        return fail(FFRS_ERR_INVALID, "NULL graph");
    FFRS_HIP(hipSetDevice(g->ctx->device));
    FFRS_HIP(hipGraphLaunch(g->exec, g->ctx->stream));
    return FFRS_OK;
}

void ffrs_graph_destroy(ffrs_graph *g)
{
    if (!g)
        return;
    (void)hipSetDevice(g->ctx->device);
    (void)hipGraphExecDestroy(g->exec);
    delete g;
}

int ffrs_ctx_kernel_timing(ffrs_ctx *ctx, int enable)
{
    if (!ctx)
        return fail(FFRS_ERR_INVALID, "NULL context");
    FFRS_HIP(hipSetDevice(ctx->device));
    if (enable)
        for (auto &e : ctx->marks)
            if (!e)
                FFRS_HIP(hipEventCreate(&e));
    ctx->marks_on = enable != 0;
    ctx->marks_n = 0;
    return FFRS_OK;
}

int ffrs_ctx_kernel_times(ffrs_ctx *ctx, float *ms, int cap, int *count)
{
    if (!ctx || !count || (cap > 0 && !ms))
        return fail(FFRS_ERR_INVALID, "NULL argument");
    FFRS_HIP(hipSetDevice(ctx->device));
    int n = ctx->marks_n > 1 ? ctx->marks_n - 1 : 0;
    *count = n;
    if (n == 0)
        return FFRS_OK;
    FFRS_HIP(hipEventSynchronize(ctx->marks[ctx->marks_n - 1]));
    for (int i = 0; i < n && i < cap; ++i)
        FFRS_HIP(hipEventElapsedTime(&ms[i], ctx->marks[i], ctx->marks[i + 1]));
    return FFRS_OK;
}

}  // extern "C"
