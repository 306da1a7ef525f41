// common.hpp -- shared host/device plumbing for the gfx950 Reed-Solomon library:
// error reporting, the context (device, stream, staging memory), buffer-resource
// loads with hardware bounds checking.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ffrs_gpu.h"

namespace ffrs {

// ---------------------------------------------------------------- errors (host)

void set_error(const char *fmt, ...);
int fail(int code, const char *fmt, ...);

#define FFRS_HIP(call)                                                                            \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return ::ffrs::fail(FFRS_ERR_DEVICE, "%s failed: %s (%s:%d)", #call,                  \
                                hipGetErrorString(e_), __FILE__, __LINE__);                       \
    } while (0)

// ---------------------------------------------------------------- context

// A growable device buffer owned by the context (host entry points stage here).
struct DevBuf {
    void *ptr = nullptr;
    size_t size = 0;
};

}  // namespace ffrs

struct ffrs_ctx {
    int device = 0;
    int cu_count = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;
    ffrs::DevBuf stage[4];
    std::mutex mu;
    // per-kernel timing of the last instrumented call (off unless enabled)
    static constexpr int kMarks = 8;
    bool marks_on = false;
    int marks_n = 0;
    hipEvent_t marks[kMarks] = {};
};

namespace ffrs {

// Ensures ctx->stage[i] holds at least `bytes` (+ 64 bytes of slack for edge loads).
int stage_reserve(ffrs_ctx *ctx, int i, size_t bytes, void **out);
int ctx_set_device(ffrs_ctx *ctx);

// With kernel timing enabled, marks_begin records the first event of a call and
// mark_kernel one event after each kernel; ffrs_ctx_kernel_times reads the gaps.
inline void marks_begin(ffrs_ctx *ctx)
{
    ctx->marks_n = 0;
    if (ctx->marks_on)
        (void)hipEventRecord(ctx->marks[ctx->marks_n++], ctx->stream);
}
inline void mark_kernel(ffrs_ctx *ctx)
{
    if (ctx->marks_on && ctx->marks_n < ffrs_ctx::kMarks)
        (void)hipEventRecord(ctx->marks[ctx->marks_n++], ctx->stream);
}

// Number of workgroups to keep resident for a persistent kernel.
inline int persistent_grid(const ffrs_ctx *ctx, int wgs_per_cu, size_t work_tiles)
{
    size_t g = (size_t)ctx->cu_count * (size_t)wgs_per_cu;
    if (g > work_tiles)
        g = work_tiles;
    return (int)(g ? g : 1);
}

// Raises a kernel's dynamic-LDS limit when a launch needs more than the default 64 KiB
// (gfx950 has 160 KiB of LDS per CU).
template <typename K>
inline int allow_lds(K kernel, size_t bytes)
{
    if (bytes > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess)
            return fail(FFRS_ERR_DEVICE, "hipFuncSetAttribute(%zu bytes of LDS) failed: %s", bytes,
                        hipGetErrorString(e));
    }
    return FFRS_OK;
}

// ---------------------------------------------------------------- device helpers

#if defined(__HIP_DEVICE_COMPILE__)
#define FFRS_RSRC_DWORD3 0x00020000
#else
#define FFRS_RSRC_DWORD3 0
#endif

// Buffer resource over [base, base + bytes): raw (stride 0) buffer loads whose offset
// is at or past `bytes` return zero, which handles ragged edges without branches.
__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void *base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, FFRS_RSRC_DWORD3);
}

__device__ inline uint32_t buf_load_u32(__amdgpu_buffer_rsrc_t r, uint32_t byte_off)
{
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(r, (int)byte_off, 0, 0);
}

// 16 bytes at a 4-byte aligned offset; the caller keeps all of them in range.
__device__ inline uint4 buf_load_u32x4(__amdgpu_buffer_rsrc_t r, uint32_t byte_off)
{
    auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)byte_off, 0, 0);
    return make_uint4(v[0], v[1], v[2], v[3]);
}

__device__ inline uint32_t alignbyte(uint32_t hi, uint32_t lo, uint32_t shift)
{
    return __builtin_amdgcn_alignbyte(hi, lo, shift);
}

}  // namespace ffrs
