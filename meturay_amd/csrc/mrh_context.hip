// mrh_context.hip — context lifetime, error strings, scratch slots and the
// hipEvent kernel profiler.
#include "mrh_internal.hpp"

#include <cstring>

namespace mrh {

int32_t fail(mrh_context* ctx, int32_t code, const char* fmt, ...)
{
    if (ctx) {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

int32_t dmalloc(mrh_context* ctx, void** p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    if (ctx->fail_alloc_after > 0 && --ctx->fail_alloc_after == 0)
        return fail(ctx, MRH_ERR_OOM, "hipMalloc(%zu) failed: injected by fail_alloc_after", bytes);
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(ctx, e == hipErrorOutOfMemory ? MRH_ERR_OOM : MRH_ERR_HIP,
                    "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return MRH_OK;
}

int32_t scratch(mrh_context* ctx, MrhSlot slot, size_t bytes, void** out)
{
    if (bytes > ctx->slot_bytes[slot]) {
        if (ctx->slot_ptr[slot]) {
            MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
            MRH_HIP_CHECK(ctx, hipFree(ctx->slot_ptr[slot]));
            ctx->slot_ptr[slot] = nullptr;
            ctx->slot_bytes[slot] = 0;
        }
        size_t want = bytes + bytes / 8 + 4096;
        MRH_TRY(dmalloc(ctx, &ctx->slot_ptr[slot], want));
        ctx->slot_bytes[slot] = want;
    }
    *out = ctx->slot_ptr[slot];
    return MRH_OK;
}

hipEvent_t prof_begin(mrh_context* ctx)
{
    hipEvent_t e = nullptr;
    if (!ctx->free_events.empty()) {
        e = ctx->free_events.back();
        ctx->free_events.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        return nullptr;
    }
    (void)hipEventRecord(e, ctx->stream);
    return e;
}

void prof_end(mrh_context* ctx, const char* name, hipEvent_t a)
{
    hipEvent_t b = prof_begin(ctx);
    if (!b) return;
    ctx->pending.push_back({name, a, b});
}

int32_t prof_collect(mrh_context* ctx)
{
    if (ctx->pending.empty()) return MRH_OK;
    MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    for (auto& p : ctx->pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            auto& t = ctx->totals[p.name];
            t.first += ms;
            t.second += 1;
        }
        ctx->free_events.push_back(p.a);
        ctx->free_events.push_back(p.b);
    }
    ctx->pending.clear();
    return MRH_OK;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t count)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

int32_t fill_u32(mrh_context* ctx, uint32_t* p, uint32_t value, size_t count, bool plain)
{
    if (count == 0) return MRH_OK;
    if (plain)
        return launch_plain(ctx, "fill_u32", dim3(stride_grid(ctx, count, 256)), dim3(256), 0,
                            k_fill_u32, p, value, count);
    return launch(ctx, "fill_u32", dim3(stride_grid(ctx, count, 256)), dim3(256), 0, k_fill_u32, p,
                  value, count);
}

}  // namespace mrh

using namespace mrh;

extern "C" {

int32_t mrh_context_create(int32_t device, mrh_context** out)
{
    if (!out) return MRH_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return MRH_ERR_HIP;
    }
    if (device < 0 || device >= count) return MRH_ERR_INVALID;
    mrh_context* ctx = new mrh_context();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess) { delete ctx; return MRH_ERR_HIP; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        ctx->num_cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return MRH_ERR_HIP;
    }
    // a failed lookup only disables mrh_partition_by_key_noalloc (its error)
    (void)noalloc_setup(ctx);
    *out = ctx;
    return MRH_OK;
}

int32_t mrh_context_destroy(mrh_context* ctx)
{
    if (!ctx) return MRH_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->guide_svo) mrh_svo_destroy(ctx->guide_svo);
    for (int i = 0; i < SLOT_COUNT; i++)
        if (ctx->slot_ptr[i]) (void)hipFree(ctx->slot_ptr[i]);
    for (auto& p : ctx->pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    for (auto e : ctx->free_events) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return MRH_OK;
}

const char* mrh_last_error(const mrh_context* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int32_t mrh_sync(mrh_context* ctx)
{
    if (!ctx) return MRH_ERR_INVALID;
    MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return MRH_OK;
}

int32_t mrh_stream(mrh_context* ctx, void** stream)
{
    if (!ctx || !stream) return MRH_ERR_INVALID;
    *stream = (void*)ctx->stream;
    return MRH_OK;
}

int32_t mrh_memcpy(mrh_context* ctx, void* dst, const void* src, size_t bytes)
{
    if (!ctx || (bytes && (!dst || !src))) return MRH_ERR_INVALID;
    if (bytes == 0) return MRH_OK;
    MRH_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream));
    MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return MRH_OK;
}

int32_t mrh_device_alloc(mrh_context* ctx, void** out, size_t bytes)
{
    if (!ctx || !out) return MRH_ERR_INVALID;
    MRH_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    return dmalloc(ctx, out, bytes);
}

int32_t mrh_device_free(mrh_context* ctx, void* ptr)
{
    if (!ctx) return MRH_ERR_INVALID;
    if (!ptr) return MRH_OK;
    MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    MRH_HIP_CHECK(ctx, hipFree(ptr));
    return MRH_OK;
}

int32_t mrh_tune(mrh_context* ctx, const char* knob, int64_t value)
{
    if (!ctx || !knob) return MRH_ERR_INVALID;
    const std::string k = knob;
    if (k == "trace_lds_entries") {
        if (value != 0 && value != 8 && value != 12 && value != 16 && value != 24)
            return fail(ctx, MRH_ERR_INVALID, "trace_lds_entries must be 0, 8, 12, 16 or 24");
        ctx->trace_lds_entries = (uint32_t)value;
        return MRH_OK;
    }
    if (k == "trace_blocks_per_cu") {
        if (value < 1 || value > 64) return fail(ctx, MRH_ERR_INVALID, "trace_blocks_per_cu must be in [1, 64]");
        ctx->trace_blocks_per_cu = (uint32_t)value;
        return MRH_OK;
    }
    if (k == "trace_width") {
        if (value != 2 && value != 4) return fail(ctx, MRH_ERR_INVALID, "trace_width must be 2 or 4");
        ctx->trace_width = (uint32_t)value;
        return MRH_OK;
    }
    if (k == "shade_partition") {
        if (value != 0 && value != 1) return fail(ctx, MRH_ERR_INVALID, "shade_partition must be 0 or 1");
        ctx->shade_partition = (uint32_t)value;
        return MRH_OK;
    }
    if (k == "trace_block") {
        if (value != 128 && value != 256 && value != 512)
            return fail(ctx, MRH_ERR_INVALID, "trace_block must be 128, 256 or 512");
        ctx->trace_block = (uint32_t)value;
        return MRH_OK;
    }
    if (k == "shade_tile") {
        if (value != 512 && value != 1024 && value != 2048)
            return fail(ctx, MRH_ERR_INVALID, "shade_tile must be 512, 1024 or 2048");
        ctx->shade_tile = (uint32_t)value;
        return MRH_OK;
    }
    if (k == "fail_alloc_after") {
        if (value < 0) return fail(ctx, MRH_ERR_INVALID, "fail_alloc_after must be >= 0");
        ctx->fail_alloc_after = (uint64_t)value;
        return MRH_OK;
    }
    return fail(ctx, MRH_ERR_INVALID, "unknown knob %s", knob);
}

int32_t mrh_tune_get(mrh_context* ctx, const char* knob, int64_t* value)
{
    if (!ctx || !knob || !value) return MRH_ERR_INVALID;
    const std::string k = knob;
    if (k == "trace_lds_entries") *value = ctx->trace_lds_entries;
    else if (k == "trace_blocks_per_cu") *value = ctx->trace_blocks_per_cu;
    else if (k == "trace_width") *value = ctx->trace_width;
    else if (k == "shade_partition") *value = ctx->shade_partition;
    else if (k == "trace_block") *value = ctx->trace_block;
    else if (k == "shade_tile") *value = ctx->shade_tile;
    else if (k == "fail_alloc_after") *value = (int64_t)ctx->fail_alloc_after;
    else return fail(ctx, MRH_ERR_INVALID, "unknown knob %s", knob);
    return MRH_OK;
}

int32_t mrh_profile_enable(mrh_context* ctx, int32_t on)
{
    if (!ctx) return MRH_ERR_INVALID;
    ctx->profile = on != 0;
    return MRH_OK;
}

int32_t mrh_profile_reset(mrh_context* ctx)
{
    if (!ctx) return MRH_ERR_INVALID;
    MRH_TRY(prof_collect(ctx));
    ctx->totals.clear();
    return MRH_OK;
}

int32_t mrh_profile_get(mrh_context* ctx, const char* kernel, double* total_ms, uint64_t* launches)
{
    if (!ctx || !kernel) return MRH_ERR_INVALID;
    MRH_TRY(prof_collect(ctx));
    auto it = ctx->totals.find(kernel);
    if (total_ms) *total_ms = (it == ctx->totals.end()) ? 0.0 : it->second.first;
    if (launches) *launches = (it == ctx->totals.end()) ? 0 : it->second.second;
    return MRH_OK;
}

}  // extern "C"
