// mrh_trace.hip — closest-hit and any-hit ray casting over the LBVH.
//
// One thread per ray in a persistent, XCD-sliced grid; SoA ray planes read
// coalesced, SoA hit planes written coalesced.  The traversal stack keeps
// its top entries in LDS (see LdsStack) with a global overflow tail.
#include "mrh_bvh.hpp"

namespace mrh {
namespace {

template <int W, bool ANY, bool STATS, int CAP, int L>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(lds_stack_waves(L, 256, 8)))) void k_trace(const float4* __restrict__ tnodes,
                                               const float4* __restrict__ leaves, uint32_t root,
                                               uint32_t n_leaves, const float* __restrict__ rays,
                                               uint32_t n, uint32_t pitch, float tmin,
                                               uint32_t* __restrict__ out_prim, float* __restrict__ out_t,
                                               float* __restrict__ out_bary, uint32_t* __restrict__ stats,
                                               uint2* __restrict__ overflow)
{
    __shared__ uint32_t s_node[L > 0 ? L * 256 : 1];
    __shared__ float s_t[L > 0 ? L * 256 : 1];
    const uint32_t lb = xcd_block();
    const size_t ostride = (size_t)gridDim.x * blockDim.x;
    const size_t oslot = (size_t)lb * blockDim.x + threadIdx.x;
    for (uint32_t i = lb * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const V3 o = v3(rays[i], rays[(size_t)pitch + i], rays[2 * (size_t)pitch + i]);
        const V3 d = v3(rays[3 * (size_t)pitch + i], rays[4 * (size_t)pitch + i], rays[5 * (size_t)pitch + i]);
        const float tmax = rays[6 * (size_t)pitch + i];
        TraceResult r;
        if constexpr (L > 0) {
            LdsStack<L, CAP, 256> st{s_node, s_t, threadIdx.x, overflow, ostride, oslot};
            r = traverse_w<W, ANY, STATS>(st, tnodes, leaves, root, n_leaves, o, d, tmin, tmax);
        } else {
            PrivateStack<CAP> st;
            r = traverse_w<W, ANY, STATS>(st, tnodes, leaves, root, n_leaves, o, d, tmin, tmax);
        }
        if (ANY) {
            out_prim[i] = (r.leaf != NO_HIT) ? 1u : 0u;
        } else {
            uint32_t prim = NO_HIT;
            if (r.leaf != NO_HIT) prim = __float_as_uint(leaves[3 * (size_t)r.leaf].w);
            out_prim[i] = prim;
            out_t[i] = (prim == NO_HIT) ? 0.0f : r.t;
            out_bary[i] = (prim == NO_HIT) ? 0.0f : r.ba;
            out_bary[(size_t)n + i] = (prim == NO_HIT) ? 0.0f : r.bb;
            if (STATS) {
                stats[2 * (size_t)i] = r.visits;
                stats[2 * (size_t)i + 1] = r.tests;
            }
        }
    }
}

template <int W, bool ANY, bool STATS, int CAP>
int32_t trace_cap(mrh_context* ctx, const mrh_lbvh* b, const TreeSel& ts, const float* rays, uint32_t n,
                  uint32_t pitch, float tmin, uint32_t* prim, float* t, float* bary, uint32_t* stats)
{
    const uint32_t grid = trace_grid(ctx, n);
    const char* name = ANY ? "trace_any" : (STATS ? "trace_closest_stats" : "trace_closest");
    const uint32_t lds = ctx->trace_lds_entries;
    uint2* ovf = nullptr;
    MRH_TRY(trace_overflow(ctx, grid, CAP, lds, &ovf));
#define MRH_TRACE_L(LV)                                                                                     \
    return launch(ctx, name, dim3(grid), dim3(256), 0, k_trace<W, ANY, STATS, CAP, LV>, ts.nodes, b->leaves, \
                  ts.root, b->n, rays, n, pitch, tmin, prim, t, bary, stats, ovf)
    switch (lds) {
    case 0: MRH_TRACE_L(0);
    case 8: MRH_TRACE_L(8);
    case 12: MRH_TRACE_L(12);
    case 16: MRH_TRACE_L(16);
    case 24: MRH_TRACE_L(24);
    default: return fail(ctx, MRH_ERR_INVALID, "trace_lds_entries must be 0, 8, 12, 16 or 24");
    }
#undef MRH_TRACE_L
}

template <bool ANY, bool STATS>
int32_t trace_launch(mrh_context* ctx, const mrh_lbvh* b, const float* rays, uint32_t n, uint32_t pitch,
                     float tmin, uint32_t* prim, float* t, float* bary, uint32_t* stats)
{
    TreeSel ts;
    MRH_TRY(select_tree(ctx, b, ts));
    if (ts.width == 4) {
        if (ts.cap == 64) return trace_cap<4, ANY, STATS, 64>(ctx, b, ts, rays, n, pitch, tmin, prim, t, bary, stats);
        return trace_cap<4, ANY, STATS, 128>(ctx, b, ts, rays, n, pitch, tmin, prim, t, bary, stats);
    }
    if (ts.cap == 64) return trace_cap<2, ANY, STATS, 64>(ctx, b, ts, rays, n, pitch, tmin, prim, t, bary, stats);
    return trace_cap<2, ANY, STATS, 128>(ctx, b, ts, rays, n, pitch, tmin, prim, t, bary, stats);
}

}  // namespace
}  // namespace mrh

using namespace mrh;

extern "C" {

int32_t mrh_trace_closest(mrh_context* ctx, const mrh_lbvh* bvh, const float* d_rays, uint32_t n,
                          uint32_t pitch, float tmin, uint32_t* d_prim, float* d_t, float* d_bary,
                          uint32_t* d_stats)
{
    if (!ctx || !bvh) return MRH_ERR_INVALID;
    if (n == 0) return MRH_OK;
    if (!d_rays || !d_prim || !d_t || !d_bary) return fail(ctx, MRH_ERR_INVALID, "mrh_trace_closest: null pointer");
    if (pitch < n) return fail(ctx, MRH_ERR_INVALID, "mrh_trace_closest: pitch < n");
    if (d_stats) return trace_launch<false, true>(ctx, bvh, d_rays, n, pitch, tmin, d_prim, d_t, d_bary, d_stats);
    return trace_launch<false, false>(ctx, bvh, d_rays, n, pitch, tmin, d_prim, d_t, d_bary, nullptr);
}

int32_t mrh_trace_any(mrh_context* ctx, const mrh_lbvh* bvh, const float* d_rays, uint32_t n,
                      uint32_t pitch, float tmin, uint32_t* d_hit)
{
    if (!ctx || !bvh) return MRH_ERR_INVALID;
    if (n == 0) return MRH_OK;
    if (!d_rays || !d_hit) return fail(ctx, MRH_ERR_INVALID, "mrh_trace_any: null pointer");
    if (pitch < n) return fail(ctx, MRH_ERR_INVALID, "mrh_trace_any: pitch < n");
    return trace_launch<true, false>(ctx, bvh, d_rays, n, pitch, tmin, d_hit, nullptr, nullptr, nullptr);
}

}  // extern "C"
