// mrh_trace.hip — closest-hit and any-hit ray casting over the LBVH.
//
// One thread per ray, SoA ray planes read coalesced, SoA hit planes written
// coalesced.  The traversal keeps a private (scratch) stack sized from the
// tree depth measured at build time.
#include "mrh_bvh.hpp"

namespace mrh {
namespace {

template <bool ANY, bool STATS, int STACK>
__global__ __launch_bounds__(256) void k_trace(const float4* __restrict__ tnodes,
                                               const float4* __restrict__ leaves, uint32_t root, uint32_t n_leaves,
                                               const float* __restrict__ rays, uint32_t n,
                                               uint32_t pitch, float tmin, uint32_t* __restrict__ out_prim,
                                               float* __restrict__ out_t, float* __restrict__ out_bary,
                                               uint32_t* __restrict__ stats)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const V3 o = v3(rays[i], rays[(size_t)pitch + i], rays[2 * (size_t)pitch + i]);
        const V3 d = v3(rays[3 * (size_t)pitch + i], rays[4 * (size_t)pitch + i], rays[5 * (size_t)pitch + i]);
        const float tmax = rays[6 * (size_t)pitch + i];
        TraceResult r = traverse<ANY, STATS, STACK>(tnodes, leaves, root, n_leaves, o, d, tmin, tmax);
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

template <bool ANY, bool STATS>
int32_t trace_launch(mrh_context* ctx, const mrh_lbvh* b, const float* rays, uint32_t n, uint32_t pitch,
                     float tmin, uint32_t* prim, float* t, float* bary, uint32_t* stats)
{
    const uint32_t root = root_code(b);
    const dim3 grid(stride_grid(ctx, n, 256)), block(256);
    const char* name = ANY ? "trace_any" : (STATS ? "trace_closest_stats" : "trace_closest");
    if (b->max_depth <= 64)
        return launch(ctx, name, grid, block, 0, k_trace<ANY, STATS, 64>, b->tnodes, b->leaves, root, b->n, rays,
                      n, pitch, tmin, prim, t, bary, stats);
    if (b->max_depth <= 128)
        return launch(ctx, name, grid, block, 0, k_trace<ANY, STATS, 128>, b->tnodes, b->leaves, root, b->n,
                      rays, n, pitch, tmin, prim, t, bary, stats);
    return fail(ctx, MRH_ERR_UNSUPPORTED, "trace: tree depth %u exceeds the traversal stack (128)",
                b->max_depth);
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
