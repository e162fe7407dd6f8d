// mrh_internal.hpp — context, error reporting, scratch memory and profiled
// kernel launches shared by every translation unit of libmrh.
#pragma once

#include "mrh.h"
#include "mrh_math.hpp"

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <vector>

// Named scratch slots of a context (grow-only device buffers).
enum MrhSlot : int {
    SLOT_SORT_KEYS_A = 0,
    SLOT_SORT_KEYS_B,
    SLOT_SORT_VALS_A,
    SLOT_SORT_VALS_B,
    SLOT_SORT_META,
    SLOT_MORTON_AABB,
    SLOT_MORTON_EXT,
    SLOT_PART_KEYS,
    SLOT_FRAME,
    SLOT_TRACE_OVERFLOW,
    SLOT_GUIDE_REC,
    SLOT_FILTER,
    SLOT_SAMPLES,
    SLOT_BOUNDS_PART,
    SLOT_VOXEL,
    SLOT_COUNT
};

struct MrhEventPair {
    const char* name;
    hipEvent_t a, b;
};

struct mrh_context {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 256;
    std::string err;
    void* slot_ptr[SLOT_COUNT] = {};
    size_t slot_bytes[SLOT_COUNT] = {};
    bool profile = false;
    std::vector<MrhEventPair> pending;
    std::vector<hipEvent_t> free_events;
    std::map<std::string, std::pair<double, uint64_t>> totals;
    // tuning knobs (mrh_tune): traversal stack entries kept in LDS per
    // thread (0 = whole stack in private scratch; 8 is what the autotuner
    // picks on MI355X for 1080p frames of a 2 M-triangle scene) and blocks
    // per CU of the persistent trace grid
    uint32_t trace_lds_entries = 8;
    uint32_t trace_blocks_per_cu = 8;
    // traversal tree: 2 = binary LBVH, 4 = 4-wide tree collapsed from it
    uint32_t trace_width = 4;
    // material partition of a bounce: 0 = global radix pass, 1 = per-block tiles
    uint32_t shade_partition = 1;
    // threads per block of the frame renders' traversal kernel (128, 256, 512)
    uint32_t trace_block = 256;
    // rays per tile of the per-block partition (512, 1024, 2048)
    uint32_t shade_tile = 1024;
    // failure-handling tests: when k > 0, the k-th following dmalloc on this
    // context fails with MRH_ERR_OOM (knob "fail_alloc_after"; 0 = off)
    uint64_t fail_alloc_after = 0;
    // octree reused by guided renders that do not bring their own
    struct mrh_svo* guide_svo = nullptr;
    // this device's fixed partition workspace (mrh_partition_by_key_noalloc);
    // null when its address lookup failed (the reason in noalloc_err)
    void* noalloc_scratch = nullptr;
    std::string noalloc_err;
};

namespace mrh {

int32_t fail(mrh_context* ctx, int32_t code, const char* fmt, ...);

#define MRH_HIP_CHECK(ctx, call)                                                      \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            return mrh::fail((ctx), e_ == hipErrorOutOfMemory ? MRH_ERR_OOM : MRH_ERR_HIP, \
                             "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),   \
                             __FILE__, __LINE__);                                     \
    } while (0)

#define MRH_TRY(expr)                      \
    do {                                   \
        int32_t rc_ = (expr);              \
        if (rc_ != MRH_OK) return rc_;     \
    } while (0)

// Grow-only scratch buffer of a slot; the returned pointer stays valid until
// the next call that grows the same slot.
int32_t scratch(mrh_context* ctx, MrhSlot slot, size_t bytes, void** out);

// Device allocation with error reporting.
int32_t dmalloc(mrh_context* ctx, void** p, size_t bytes);

// Begin/end event pair around a launch when profiling is enabled.
hipEvent_t prof_begin(mrh_context* ctx);
void prof_end(mrh_context* ctx, const char* name, hipEvent_t a);
// Folds finished event pairs into totals (synchronises the stream).
int32_t prof_collect(mrh_context* ctx);

template <class Kernel, class... Args>
int32_t launch(mrh_context* ctx, const char* name, dim3 grid, dim3 block, size_t shmem,
               Kernel kernel, Args... args)
{
    if (grid.x == 0) return MRH_OK;
    hipEvent_t a = ctx->profile ? prof_begin(ctx) : nullptr;
    hipLaunchKernelGGL(kernel, grid, block, shmem, ctx->stream, args...);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(ctx, MRH_ERR_HIP, "launch of %s failed: %s", name, hipGetErrorString(e));
    if (a) prof_end(ctx, name, a);
    return MRH_OK;
}

// launch() without the profiler's events: for entry points that must not
// create runtime objects (mrh_partition_by_key_noalloc).
template <class Kernel, class... Args>
int32_t launch_plain(mrh_context* ctx, const char* name, dim3 grid, dim3 block, size_t shmem,
                     Kernel kernel, Args... args)
{
    if (grid.x == 0) return MRH_OK;
    hipLaunchKernelGGL(kernel, grid, block, shmem, ctx->stream, args...);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(ctx, MRH_ERR_HIP, "launch of %s failed: %s", name, hipGetErrorString(e));
    return MRH_OK;
}

// Fill kernel launcher (graph-safe replacement for hipMemsetAsync); plain =
// no profiler events.
int32_t fill_u32(mrh_context* ctx, uint32_t* p, uint32_t value, size_t count, bool plain = false);

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Grid for grid-stride kernels: enough blocks to fill every CU several times.
inline uint32_t stride_grid(mrh_context* ctx, uint64_t items, uint32_t block)
{
    uint64_t want = (items + block - 1) / block;
    uint64_t cap = (uint64_t)ctx->num_cus * 16u;
    if (want > cap) want = cap;
    return (uint32_t)(want ? want : 1);
}

// ---------------------------------------------------------------- sorting
// Onesweep LSD radix sort over bits [begin, end) of 32- or 64-bit keys.
// d_vals_in may be null (iota).  d_keys_out may be null when only the
// permutation is wanted.  n_dev (optional) is a device-side element count
// <= n_max; n_max bounds the launch.  part (optional) receives the offsets
// of the digits [part_begin, end) (see mrh_partition_by_key).
struct SortArgs {
    const void* keys_in = nullptr;
    const uint32_t* vals_in = nullptr;
    void* keys_out = nullptr;
    uint32_t* vals_out = nullptr;
    uint32_t n_max = 0;
    const uint32_t* n_dev = nullptr;
    uint32_t key_bytes = 4;
    uint32_t begin_bit = 0, end_bit = 32;
    bool descending = false;
    uint32_t* part_offsets = nullptr;
    uint32_t part_begin = 0;
};
// Workspace of a sort: the meta words (digit histograms, partition counts,
// tile counters, look-back status) and the ping-pong keys/values of the
// intermediate passes.  radix_sort takes them from the context's grow-only
// slots when no workspace is given.
struct SortSpace {
    uint32_t* meta = nullptr;
    size_t meta_words = 0;
    void* keys_a = nullptr;
    void* keys_b = nullptr;
    uint32_t* vals_a = nullptr;
    uint32_t* vals_b = nullptr;
    uint32_t max_n = 0;          // elements each ping-pong buffer holds
};
int32_t radix_sort(mrh_context* ctx, const SortArgs& a, const SortSpace* space = nullptr);

// Device address of this device's fixed partition workspace, looked up once
// per device (mrh_context_create calls it with the device current).
int32_t noalloc_setup(mrh_context* ctx);

}  // namespace mrh
