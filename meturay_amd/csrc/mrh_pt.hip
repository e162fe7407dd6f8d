// mrh_pt.hip — the plain wavefront path tracer (PathTracer::Render,
// TracerCUDA/PathTracer.cu:144-291) over resident scenes.
//
// Per bounce, with every count kept on the device (no host round trip, so
// the loop can be captured or replayed without synchronising):
//   trace_pt   : live rays [0, n) -> leaf, t, barycentrics, HitKey
//   partition  : one onesweep pass, stable descending by the key's batch
//                byte -> ray order + batch offsets
//   shade      : one launch over all partitions; rays are visited in
//                partition order so a wave nearly always runs one material;
//                emitters add Le*T to the image, Lambert / GGX conductor /
//                dielectric sample the next direction, Russian roulette runs
//                after rr_start, survivors are appended (wave-aggregated
//                atomics) into the other path buffer, which becomes the next
//                bounce's contiguous live range.
// Paths own a PCG32 stream keyed by (seed, global sample id), so results do
// not depend on the order rays are processed in.
#include "mrh_filter.hpp"
#include "mrh_path.hpp"

#include <cfloat>
#include <cmath>
#include <cstring>

namespace mrh {
namespace {

// GenerationKernels.cuh:29-108 + GPUCameraPinhole.cuh:209-240
// (coords, when set, receives every sample's image coordinate)
__global__ __launch_bounds__(256) void k_raygen(Cam cam, uint32_t w, uint32_t h, uint32_t S,
                                                uint32_t seed, uint32_t sample_base, uint32_t n_total,
                                                PathBuf out, float2* __restrict__ coords)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_total; i += gridDim.x * blockDim.x) {
        V3 o, d;
        uint64_t st;
        // the global sample id keys the stream
        camera_ray(cam, w, h, S, seed, sample_base + i, i, o, d, st, coords ? coords + i : nullptr);
        out.ox[i] = o.x; out.oy[i] = o.y; out.oz[i] = o.z;
        out.dx[i] = d.x; out.dy[i] = d.y; out.dz[i] = d.z;
        out.tr[i] = 1.0f; out.tg[i] = 1.0f; out.tb[i] = 1.0f;
        out.sample[i] = i;
        out.dm[i] = 1u;
        out.rng[i] = st;
    }
}

// counts layout: [0, 64) live rays per bounce, [64, 66) u64 node visits,
// [66, 68) u64 triangle tests
constexpr uint32_t COUNTS_WORDS = 128;

template <int W, int CAP, bool STATS, int L, int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(lds_stack_waves(L, BS, 8)))) void k_trace_pt(const float4* __restrict__ tnodes,
                                                  const float4* __restrict__ leaves, uint32_t root, uint32_t n_leaves,
                                                  PathBuf in, uint32_t* __restrict__ counts,
                                                  uint32_t bounce, float tmin, float tmax, HitBuf hit,
                                                  uint2* __restrict__ overflow)
{
    __shared__ uint32_t s_node[L > 0 ? L * BS : 1];
    __shared__ float s_t[L > 0 ? L * BS : 1];
    const uint32_t n = counts[bounce];
    const uint32_t lb = xcd_block();
    const size_t ostride = (size_t)gridDim.x * blockDim.x;
    const size_t oslot = (size_t)lb * blockDim.x + threadIdx.x;
    uint64_t visits = 0, tests = 0;
    for (uint32_t i = lb * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const V3 o = v3(in.ox[i], in.oy[i], in.oz[i]);
        const V3 d = v3(in.dx[i], in.dy[i], in.dz[i]);
        TraceResult r;
        if constexpr (L > 0) {
            LdsStack<L, CAP, BS> st{s_node, s_t, threadIdx.x, overflow, ostride, oslot};
            r = traverse_w<W, false, STATS>(st, tnodes, leaves, root, n_leaves, o, d, tmin, tmax);
        } else {
            PrivateStack<CAP> st;
            r = traverse_w<W, false, STATS>(st, tnodes, leaves, root, n_leaves, o, d, tmin, tmax);
        }
        hit.leaf[i] = r.leaf;
        hit.t[i] = r.t;
        hit.ba[i] = r.ba;
        hit.bb[i] = r.bb;
        hit.key[i] = (r.leaf == NO_HIT) ? 0u : __float_as_uint(leaves[3 * (size_t)r.leaf + 1].w);
        if (STATS) {
            visits += r.visits;
            tests += r.tests;
        }
    }
    if (STATS) {
        for (int off = 32; off > 0; off >>= 1) {
            visits += __shfl_xor(visits, off, 64);
            tests += __shfl_xor(tests, off, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            atomicAdd((unsigned long long*)(counts + 64), (unsigned long long)visits);
            atomicAdd((unsigned long long*)(counts + 66), (unsigned long long)tests);
        }
    }
}

template <int W, int CAP, bool STATS>
int32_t launch_trace_pt_w(mrh_context* ctx, uint32_t grid, const mrh_lbvh* b, const TreeSel& ts,
                          const PathBuf& in, uint32_t* counts, uint32_t bounce, float tmin, float tmax,
                          const HitBuf& hit)
{
    const uint32_t root = ts.root;
    const uint32_t lds = ctx->trace_lds_entries;
    const uint32_t bs = ctx->trace_block;
    uint2* ovf = nullptr;
    MRH_TRY(trace_overflow(ctx, grid, CAP, lds, &ovf, bs));
    const char* name = STATS ? "trace_pt_stats" : "trace_pt";
    const float4* tn = ts.nodes;
    const float4* lv = (const float4*)b->leaves;
#define MRH_TRACE_PT_L(LV, BS)                                                                            \
    return launch(ctx, name, dim3(grid), dim3(BS), 0, k_trace_pt<W, CAP, STATS, LV, BS>, tn, lv, root, b->n, \
                  in, counts, bounce, tmin, tmax, hit, ovf)
#define MRH_TRACE_PT_B(BS)                                                                   \
    switch (lds) {                                                                           \
    case 0: MRH_TRACE_PT_L(0, BS);                                                           \
    case 8: MRH_TRACE_PT_L(8, BS);                                                           \
    case 12: MRH_TRACE_PT_L(12, BS);                                                         \
    case 16: MRH_TRACE_PT_L(16, BS);                                                         \
    case 24: MRH_TRACE_PT_L(24, BS);                                                         \
    default: return fail(ctx, MRH_ERR_INVALID, "trace_lds_entries must be 0, 8, 12, 16 or 24"); \
    }
    switch (bs) {
    case 128: MRH_TRACE_PT_B(128);
    case 256: MRH_TRACE_PT_B(256);
    case 512: MRH_TRACE_PT_B(512);
    default: return fail(ctx, MRH_ERR_INVALID, "trace_block must be 128, 256 or 512");
    }
#undef MRH_TRACE_PT_B
#undef MRH_TRACE_PT_L
}

template <bool STATS>
int32_t launch_trace_pt(mrh_context* ctx, uint32_t grid, const mrh_lbvh* b, const PathBuf& in, uint32_t* counts,
                        uint32_t bounce, float tmin, float tmax, const HitBuf& hit)
{
    TreeSel ts;
    MRH_TRY(select_tree(ctx, b, ts));
    if (ts.width == 4) {
        if (ts.cap == 64) return launch_trace_pt_w<4, 64, STATS>(ctx, grid, b, ts, in, counts, bounce, tmin, tmax, hit);
        return launch_trace_pt_w<4, 128, STATS>(ctx, grid, b, ts, in, counts, bounce, tmin, tmax, hit);
    }
    if (ts.cap == 64) return launch_trace_pt_w<2, 64, STATS>(ctx, grid, b, ts, in, counts, bounce, tmin, tmax, hit);
    return launch_trace_pt_w<2, 128, STATS>(ctx, grid, b, ts, in, counts, bounce, tmin, tmax, hit);
}

// Two-level scenes: the top tree over the instances, then the hit mesh's
// tree in local space (traverse_two_level); writes the instance too, and the
// instance's HitKey override where it has one.  Five waves per SIMD: the
// walk needs 90 registers; capped at 80 (six waves) it spills and runs 3 %
// slower, at 72 or 64 it spills far more (1.4x and 2.9x the frame time).
template <int W, int CAP, bool STATS, int L>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(lds_stack_waves(L, 256, 5)))) void k_trace_pt2(const float4* __restrict__ top4,
                                                   const float4* __restrict__ top_leaves, uint32_t top_root,
                                                   uint32_t n_inst, const float4* __restrict__ inst,
                                                   const float4* __restrict__ mesh_nodes,
                                                   const float4* __restrict__ leaves, PathBuf in,
                                                   uint32_t* __restrict__ counts, uint32_t bounce, float tmin,
                                                   float tmax, HitBuf hit, uint2* __restrict__ overflow)
{
    __shared__ uint32_t s_node[L > 0 ? L * 256 : 1];
    __shared__ float s_t[L > 0 ? L * 256 : 1];
    const uint32_t n = counts[bounce];
    const uint32_t lb = xcd_block();
    const size_t ostride = (size_t)gridDim.x * blockDim.x;
    const size_t oslot = (size_t)lb * blockDim.x + threadIdx.x;
    uint64_t visits = 0, tests = 0;
    for (uint32_t i = lb * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const V3 o = v3(in.ox[i], in.oy[i], in.oz[i]);
        const V3 d = v3(in.dx[i], in.dy[i], in.dz[i]);
        TraceResult r;
        uint32_t ii;
        auto world = [&](V3& wo, V3& wd) {
            wo = v3(in.ox[i], in.oy[i], in.oz[i]);
            wd = v3(in.dx[i], in.dy[i], in.dz[i]);
        };
        if constexpr (L > 0) {
            LdsStack<L, CAP, 256> st{s_node, s_t, threadIdx.x, overflow, ostride, oslot};
            r = two_level_w<W, false, STATS>(st, top4, top_leaves, top_root, n_inst, inst, mesh_nodes, leaves, o, d,
                                             tmin, tmax, ii, world);
        } else {
            PrivateStack<CAP> st;
            r = two_level_w<W, false, STATS>(st, top4, top_leaves, top_root, n_inst, inst, mesh_nodes, leaves, o, d,
                                             tmin, tmax, ii, world);
        }
        hit.leaf[i] = r.leaf;
        hit.t[i] = r.t;
        hit.ba[i] = r.ba;
        hit.bb[i] = r.bb;
        hit.inst[i] = ii;
        uint32_t key = 0u;
        if (r.leaf != NO_HIT) {
            const uint32_t ko = __float_as_uint(inst[INST_F4 * (size_t)ii + 6].y);
            key = ko ? ko : __float_as_uint(leaves[3 * (size_t)r.leaf + 1].w);
        }
        hit.key[i] = key;
        if (STATS) {
            visits += r.visits;
            tests += r.tests;
        }
    }
    if (STATS) {
        for (int off = 32; off > 0; off >>= 1) {
            visits += __shfl_xor(visits, off, 64);
            tests += __shfl_xor(tests, off, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            atomicAdd((unsigned long long*)(counts + 64), (unsigned long long)visits);
            atomicAdd((unsigned long long*)(counts + 66), (unsigned long long)tests);
        }
    }
}

template <int W, int CAP, bool STATS>
int32_t launch_trace_pt2_w(mrh_context* ctx, uint32_t grid, const mrh_scene* s, const PathBuf& in,
                           uint32_t* counts, uint32_t bounce, float tmin, float tmax, const HitBuf& hit)
{
    const mrh_lbvh* top = s->bvh;
    const uint32_t lds = ctx->trace_lds_entries;
    uint2* ovf = nullptr;
    MRH_TRY(trace_overflow(ctx, grid, CAP, lds, &ovf));
    const char* name = STATS ? "trace_pt_stats" : "trace_pt";
    const float4* t4 = top->nodes4;
    const float4* tl = (const float4*)top->leaves;
    const float4* mn = (W == 4) ? s->mesh_nodes4 : s->mesh_nodes2;
#define MRH_TRACE_PT2_L(LV)                                                                                    \
    return launch(ctx, name, dim3(grid), dim3(256), 0, k_trace_pt2<W, CAP, STATS, LV>, t4, tl, top->root4,      \
                  s->n_inst, (const float4*)s->inst, mn, (const float4*)s->leaves, in, counts, bounce, tmin, tmax, \
                  hit, ovf)
    switch (lds) {
    case 0: MRH_TRACE_PT2_L(0);
    case 8: MRH_TRACE_PT2_L(8);
    case 12: MRH_TRACE_PT2_L(12);
    case 16: MRH_TRACE_PT2_L(16);
    case 24: MRH_TRACE_PT2_L(24);
    default: return fail(ctx, MRH_ERR_INVALID, "trace_lds_entries must be 0, 8, 12, 16 or 24");
    }
#undef MRH_TRACE_PT2_L
}

template <bool STATS>
int32_t launch_trace_pt2(mrh_context* ctx, uint32_t grid, const mrh_scene* s, const PathBuf& in, uint32_t* counts,
                         uint32_t bounce, float tmin, float tmax, const HitBuf& hit)
{
    const uint32_t w = (ctx->trace_width == 4) ? 4u : 2u;
    const uint32_t need = two_level_stack_need(s->bvh->max_depth, s->mesh_depth, w);
    if (need > 128)
        return fail(ctx, MRH_ERR_UNSUPPORTED, "two-level trees need %u stack entries (at most 128)", need);
    if (w == 4) {
        if (need <= 64) return launch_trace_pt2_w<4, 64, STATS>(ctx, grid, s, in, counts, bounce, tmin, tmax, hit);
        return launch_trace_pt2_w<4, 128, STATS>(ctx, grid, s, in, counts, bounce, tmin, tmax, hit);
    }
    if (need <= 64) return launch_trace_pt2_w<2, 64, STATS>(ctx, grid, s, in, counts, bounce, tmin, tmax, hit);
    return launch_trace_pt2_w<2, 128, STATS>(ctx, grid, s, in, counts, bounce, tmin, tmax, hit);
}

// One launch over all partitions of a bounce (see shade_ray); with NEE a
// path appends its next-event ray beside its continuing ray.
template <bool LEARN, bool GUIDE, bool NEE>
__global__ __launch_bounds__(256) void k_shade(PathBuf in, PathBuf out, HitBuf hit, const uint32_t* __restrict__ order,
                                               uint32_t* counts, uint32_t bounce, ShadeReq q, GuideArgs ga)
{
    __shared__ uint32_t s_wave[4];
    const uint32_t n = counts[bounce];
    const bool with_pdf = NEE && q.mis;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const uint32_t j = base + threadIdx.x;
        bool alive = false, has_nee = false;
        RayNext nx, ne;
        uint32_t sample = 0;
        if (j < n) {
            const uint32_t rid = order[j];
            sample = in.sample[rid];
            alive = shade_ray<LEARN, GUIDE, NEE>(in, hit, rid, q, ga, true, true, nx, ne, has_nee);
        }
        if (NEE) append_ray_block(has_nee, ne, sample, out, &counts[bounce + 1], false, s_wave);
        append_ray_block(alive, nx, sample, out, &counts[bounce + 1], with_pdf, s_wave);
    }
}


// The same shade step with the partition done per block: a block takes a
// tile of SHADE_TILE consecutive rays (coalesced key loads; 512, 1024 or
// 2048, the "shade_tile" knob), counts them per HitKey batch in LDS, lays
// the tile's ray indices out batch by batch and shades them in that order,
// so a wave nearly always runs one material while all ray-state lines of the
// tile are fetched by the one block, once.  No global sort, no second pass
// over the keys.
template <bool LEARN, bool GUIDE, bool NEE, uint32_t SHADE_TILE>
__global__ __launch_bounds__(256) void k_shade_tiles(PathBuf in, PathBuf out, HitBuf hit, uint32_t* counts,
                                                     uint32_t bounce, ShadeReq q, GuideArgs ga)
{
    __shared__ uint16_t s_perm[SHADE_TILE];
    __shared__ uint32_t s_cnt[8], s_off[8], s_wave[4];
    const uint32_t n = counts[bounce];
    const bool with_pdf = NEE && q.mis;
    for (uint32_t base = blockIdx.x * SHADE_TILE; base < n; base += gridDim.x * SHADE_TILE) {
        if (threadIdx.x < 8) s_cnt[threadIdx.x] = 0;
        __syncthreads();
        uint32_t b[SHADE_TILE / 256];
#pragma unroll
        for (uint32_t k = 0; k < SHADE_TILE / 256; k++) {
            const uint32_t j = base + threadIdx.x + 256 * k;
            b[k] = (j < n) ? min(hit.key[j] >> 24, 7u) : 7u;
            atomicAdd(&s_cnt[b[k]], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            // batches in descending order, as the global partition orders them
            uint32_t acc = 0;
            for (int c = 7; c >= 0; c--) {
                s_off[c] = acc;
                acc += s_cnt[c];
            }
        }
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < SHADE_TILE / 256; k++) {
            const uint32_t pos = atomicAdd(&s_off[b[k]], 1u);
            s_perm[pos] = (uint16_t)(threadIdx.x + 256 * k);
        }
        __syncthreads();
#pragma unroll 1
        for (uint32_t k = 0; k < SHADE_TILE / 256; k++) {
            const uint32_t rid = base + s_perm[threadIdx.x + 256 * k];
            bool alive = false, has_nee = false;
            RayNext nx, ne;
            uint32_t sample = 0;
            if (rid < n) {
                sample = in.sample[rid];
                alive = shade_ray<LEARN, GUIDE, NEE>(in, hit, rid, q, ga, true, true, nx, ne, has_nee);
            }
            if (NEE) append_ray_block(has_nee, ne, sample, out, &counts[bounce + 1], false, s_wave);
            append_ray_block(alive, nx, sample, out, &counts[bounce + 1], with_pdf, s_wave);
        }
    }
}

__global__ void k_scale(float* __restrict__ p, size_t count, float s)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        p[i] = p[i] * s;
}

constexpr uint32_t MAX_BOUNCES = 62;

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Live rays a frame of `total` paths can hold: every path has one ray in
// flight, and with NEE a second (its next-event ray).
uint64_t frame_cap(uint64_t total, uint32_t nee) { return nee ? 2 * total : total; }

size_t frame_bytes(uint64_t cap)
{
    // 2 path buffers (9 floats + sample + dm + pdf = 12 words, + u64 rng),
    // hits (4 words) + key + instance + order, plus counts/offsets
    size_t per = 0;
    per += 2 * (12 * align_up(cap * 4) + align_up(cap * 8));
    per += 7 * align_up(cap * 4);
    per += 4096;
    return per;
}

// One frame in flight: path/hit buffers carved from an arena, the bounce
// counter and the frame's constants.
struct Frame {
    PathBuf bufs[2];
    HitBuf hit;
    uint32_t *order = nullptr, *counts = nullptr, *offsets = nullptr;
    Cam cam;
    uint64_t total = 0, cap = 0, spp = 0;
    uint32_t w = 0, h = 0, S = 0, max_depth = 0, rr_start = 0, seed = 0, sample_base = 0;
    uint32_t nee = 0, mis = 0;
    uint32_t per_sample = 0;   // radiance goes to a per-sample buffer (filtered frames)
    uint32_t count_traversal = 0, bounce = 0;
    uint32_t learn = 0, guide = 0;   // record path vertices / sample the guide
    GuideArgs ga = {};
    struct WfState* wf = nullptr;    // field-traced guide (render_wfpg), sampled on guide frames
};

void carve_frame(void* base, uint64_t cap, Frame& f)
{
    char* p = (char*)base;
    auto take = [&](size_t bytes) {
        char* r = p;
        p += align_up(bytes);
        return r;
    };
    for (int b = 0; b < 2; b++) {
        PathBuf& q = f.bufs[b];
        q.ox = (float*)take(cap * 4); q.oy = (float*)take(cap * 4); q.oz = (float*)take(cap * 4);
        q.dx = (float*)take(cap * 4); q.dy = (float*)take(cap * 4); q.dz = (float*)take(cap * 4);
        q.tr = (float*)take(cap * 4); q.tg = (float*)take(cap * 4); q.tb = (float*)take(cap * 4);
        q.sample = (uint32_t*)take(cap * 4);
        q.dm = (uint32_t*)take(cap * 4);
        q.rng = (uint64_t*)take(cap * 8);
        q.pdf = (float*)take(cap * 4);
    }
    f.hit.leaf = (uint32_t*)take(cap * 4);
    f.hit.t = (float*)take(cap * 4);
    f.hit.ba = (float*)take(cap * 4);
    f.hit.bb = (float*)take(cap * 4);
    f.hit.key = (uint32_t*)take(cap * 4);
    f.hit.inst = (uint32_t*)take(cap * 4);
    f.order = (uint32_t*)take(cap * 4);
    f.counts = (uint32_t*)take(COUNTS_WORDS * 4);
    f.offsets = (uint32_t*)take(260 * 4);
}

// Light list of the scene (the primitive-backed lights of
// CPULightGroup::InitializeGroup, GPULightPrimitive.hpp:42-110: one light per
// emissive triangle, in primitive order): rank[p] is the light index of
// primitive p (~0 for other primitives); lights[rank] = its sorted leaf.
__global__ void k_light_table(const float4* __restrict__ leaves, uint32_t n, const uint32_t* __restrict__ rank,
                              uint32_t* __restrict__ lights)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t r = rank[__float_as_uint(leaves[3 * (size_t)i].w)];
        if (r != 0xFFFFFFFFu) lights[r] = i;
    }
}

template <class T>
int32_t scene_upload(mrh_scene* s, T** dst, const T* src, size_t count)
{
    *dst = nullptr;
    size_t bytes = count * sizeof(T);
    void* p = nullptr;
    MRH_TRY(dmalloc(s->ctx, &p, bytes ? bytes : 16));
    s->allocs.push_back(p);
    s->bytes += bytes;
    if (bytes) MRH_HIP_CHECK(s->ctx, hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
    *dst = (T*)p;
    return MRH_OK;
}

void scene_free(mrh_scene* s)
{
    if (!s) return;
    if (s->ctx) (void)hipStreamSynchronize(s->ctx->stream);
    if (s->bvh) lbvh_free(s->bvh);
    for (void* p : s->allocs) (void)hipFree(p);
    delete s;
}

int32_t scene_create(mrh_context* ctx, const mrh_scene_desc* d, mrh_scene** out)
{
    if (d->n == 0 || !d->tris || !d->prim_key)
        return fail(ctx, MRH_ERR_INVALID, "mrh_scene_create: empty scene");
    // validate keys against the material tables (SceneError on bad ids)
    for (uint32_t i = 0; i < d->n; i++) {
        uint32_t b = d->prim_key[i] >> 24, id = d->prim_key[i] & 0xFFFFFFu;
        uint32_t lim = (b == MRH_BATCH_LAMBERT) ? d->n_lambert
                     : (b == MRH_BATCH_METAL) ? d->n_metal
                     : (b == MRH_BATCH_DIELECTRIC) ? d->n_dielectric
                     : (b == MRH_BATCH_EMISSIVE) ? d->n_emissive
                     : (b == MRH_BATCH_REFLECT) ? d->n_reflect : 0u;
        if (b == MRH_BATCH_NULL || b > MRH_BATCH_REFLECT || id >= lim)
            return fail(ctx, MRH_ERR_INVALID, "mrh_scene_create: primitive %u has invalid key 0x%08x", i,
                        d->prim_key[i]);
    }
    if (d->n_dielectric > MAX_DIELECTRICS)
        return fail(ctx, MRH_ERR_INVALID, "mrh_scene_create: at most %u dielectric materials", MAX_DIELECTRICS);
    const bool two = d->instancing != nullptr;
    if (d->normals && two)
        return fail(ctx, MRH_ERR_UNSUPPORTED, "mrh_scene_create: per-corner normals in a two-level scene");
    // smooth shading: every corner's world-to-tangent rotation
    std::vector<float> tbn;
    if (d->normals) {
        tbn.resize((size_t)d->n * 12);
        for (size_t k = 0; k < (size_t)d->n * 3; k++) {
            const float* nv = d->normals + 3 * k;
            const float l2 = nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2];
            if (!(l2 > 0.0f) || !std::isfinite(l2))
                return fail(ctx, MRH_ERR_INVALID, "mrh_scene_create: corner normal %zu of triangle %zu is not a direction",
                            k % 3, k / 3);
            const Q4 q = vertex_frame(normalize(v3(nv[0], nv[1], nv[2])));
            tbn[4 * k + 0] = q.w;
            tbn[4 * k + 1] = q.x;
            tbn[4 * k + 2] = q.y;
            tbn[4 * k + 3] = q.z;
        }
    }
    std::vector<uint32_t> rank(two ? 0 : d->n);
    uint32_t n_lights = 0;
    for (uint32_t i = 0; i < d->n && !two; i++)
        rank[i] = ((d->prim_key[i] >> 24) == MRH_BATCH_EMISSIVE) ? n_lights++ : 0xFFFFFFFFu;
    mrh_scene* s = new mrh_scene();
    s->ctx = ctx;
    s->n_lights = n_lights;
    s->n_lambert = d->n_lambert;
    s->n_metal = d->n_metal;
    s->n_dielectric = d->n_dielectric;
    s->n_emissive = d->n_emissive;
    s->n_reflect = d->n_reflect;
    float* dtris = nullptr;
    uint32_t* dkeys = nullptr;
    int32_t rc = MRH_OK;
    do {
        if ((rc = scene_upload(s, &s->lambert, d->lambert, (size_t)d->n_lambert * 3)) != MRH_OK) break;
        if ((rc = scene_upload(s, &s->metal, d->metal, (size_t)d->n_metal * 10)) != MRH_OK) break;
        if ((rc = scene_upload(s, &s->dielectric, d->dielectric, (size_t)d->n_dielectric * 4)) != MRH_OK) break;
        if ((rc = scene_upload(s, &s->emissive, d->emissive, (size_t)d->n_emissive * 3)) != MRH_OK) break;
        if ((rc = scene_upload(s, &s->reflect, d->reflect, (size_t)d->n_reflect * 4)) != MRH_OK) break;
        if (two) {
            rc = two_level_build(ctx, d, s);
            break;
        }
        if (!tbn.empty()) {
            float* dtbn = nullptr;
            if ((rc = scene_upload(s, &dtbn, tbn.data(), tbn.size())) != MRH_OK) break;
            s->tbn = (float4*)dtbn;
        }
        if ((rc = scene_upload(s, &dtris, d->tris, (size_t)d->n * 9)) != MRH_OK) break;
        if ((rc = scene_upload(s, &dkeys, d->prim_key, (size_t)d->n)) != MRH_OK) break;
        uint32_t bits = d->morton_bits ? d->morton_bits : 30u;
        if ((rc = lbvh_create(ctx, dtris, d->n, bits, dkeys, &s->bvh)) != MRH_OK) break;
        s->bytes += (size_t)d->n * (36 + 24 + 8 + 8 + 4 + 12 + 8 + 48 + 64 + 4);
        if (n_lights) {
            uint32_t* drank = nullptr;
            if ((rc = dmalloc(ctx, (void**)&drank, (size_t)d->n * 4)) != MRH_OK) break;
            void* lp = nullptr;
            if ((rc = dmalloc(ctx, &lp, (size_t)n_lights * 4)) != MRH_OK) {
                (void)hipFree(drank);
                break;
            }
            s->allocs.push_back(lp);
            s->bytes += (size_t)n_lights * 4;
            s->lights = (uint32_t*)lp;
            if (hipMemcpy(drank, rank.data(), (size_t)d->n * 4, hipMemcpyHostToDevice) != hipSuccess)
                rc = fail(ctx, MRH_ERR_HIP, "mrh_scene_create: light rank upload failed");
            if (rc == MRH_OK)
                rc = launch(ctx, "light_table", dim3(stride_grid(ctx, d->n, 256)), dim3(256), 0, k_light_table,
                            (const float4*)s->bvh->leaves, d->n, (const uint32_t*)drank, s->lights);
            if (rc == MRH_OK && hipStreamSynchronize(ctx->stream) != hipSuccess)
                rc = fail(ctx, MRH_ERR_HIP, "mrh_scene_create: light table failed");
            (void)hipFree(drank);
        }
    } while (0);
    if (rc != MRH_OK) {
        scene_free(s);
        return rc;
    }
    *out = s;
    return MRH_OK;
}

int32_t validate_params(mrh_context* ctx, const mrh_scene* s, const mrh_render_params* p, uint64_t* total)
{
    if (p->nee > 1 || p->direct_light_mis > 1)
        return fail(ctx, MRH_ERR_INVALID, "render: nee and direct_light_mis are 0 or 1");
    if (p->width == 0 || p->height == 0 || p->samples == 0)
        return fail(ctx, MRH_ERR_INVALID, "render: empty frame");
    if (p->max_depth == 0 || p->max_depth > MAX_BOUNCES)
        return fail(ctx, MRH_ERR_INVALID, "render: max_depth must be in [1, %u]", MAX_BOUNCES);
    const uint64_t t = (uint64_t)p->width * p->height * p->samples * p->samples;
    if (t >= (1ull << 30))
        return fail(ctx, MRH_ERR_INVALID, "render: %llu paths exceed 2^30", (unsigned long long)t);
    // the global material partition sorts a whole queue: a path and its
    // next-event ray each, and the sort takes fewer than 2^30 keys
    if (ctx->shade_partition == 0 && frame_cap(t, p->nee) >= (1ull << 30))
        return fail(ctx, MRH_ERR_INVALID,
                    "render: %llu rays per bounce exceed the global partition's 2^30 - 1 keys "
                    "(use the tile partition, shade_partition = 1, or fewer paths)",
                    (unsigned long long)frame_cap(t, p->nee));
    if (s->bvh->max_depth > 128)
        return fail(ctx, MRH_ERR_UNSUPPORTED, "tree depth %u exceeds the traversal stack", s->bvh->max_depth);
    *total = t;
    return MRH_OK;
}

// GenerateWork: clear counts, write the camera rays of one frame; with
// coords the frame keeps every sample's radiance apart (the image pointer
// of frame_bounce is then a per-sample buffer) and its image coordinate.
int32_t frame_begin(mrh_context* ctx, const mrh_camera_desc* cd, const mrh_render_params* p,
                    void* arena, uint32_t sample_base, Frame& f, float2* coords = nullptr)
{
    f.w = p->width;
    f.h = p->height;
    f.S = p->samples;
    f.spp = (uint64_t)p->samples * p->samples;
    f.total = (uint64_t)p->width * p->height * f.spp;
    f.max_depth = p->max_depth;
    f.rr_start = p->rr_start;
    f.seed = p->seed;
    f.sample_base = sample_base;
    f.count_traversal = p->count_traversal;
    f.nee = p->nee;
    f.mis = p->nee && p->direct_light_mis;
    f.cap = frame_cap(f.total, f.nee);
    f.per_sample = coords != nullptr;
    f.bounce = 0;
    f.learn = f.guide = 0;
    f.cam = make_camera(*cd, p->width, p->height);
    carve_frame(arena, f.cap, f);
    MRH_TRY(fill_u32(ctx, f.counts, 0u, COUNTS_WORDS));
    MRH_TRY(fill_u32(ctx, f.counts, (uint32_t)f.total, 1));
    return launch(ctx, "raygen", dim3(stride_grid(ctx, f.total, 256)), dim3(256), 0, k_raygen, f.cam,
                  f.w, f.h, f.S, f.seed, f.sample_base, (uint32_t)f.total, f.bufs[0], coords);
}

// ------------------------------------------------- field-traced guiding
// State of one render_wfpg call: the scene's voxel octree and its payload,
// the per-ray arrays of a guided bounce and the per-bin fields.
struct WfState {
    mrh_wfpg_params wp = {};
    mrh_svo* svo = nullptr;
    float cube[4] = {0, 0, 0, 1};
    float* tris = nullptr;           // [n_tris][9] the scene's triangles
    uint32_t n_tris = 0;
    uint32_t *vcodes = nullptr, *vprims = nullptr;   // voxel records
    uint32_t n_vox = 0, nodes = 0;
    float *irr = nullptr, *nrm = nullptr;            // node payload
    bool have_payload = false;
    // per ray of a bounce (cap entries)
    uint32_t *codes = nullptr, *flag = nullptr, *bins = nullptr, *order = nullptr, *offs = nullptr,
             *pbins = nullptr, *pcount = nullptr;
    float4* org = nullptr;
    float2 *xi = nullptr, *euv = nullptr, *guv = nullptr;
    float *bdir = nullptr, *gdir = nullptr, *gpdf = nullptr, *epdf = nullptr;
    // per bin
    float4* origins = nullptr;
    float2* jitter = nullptr;
    float* fields = nullptr;
    size_t field_floats = 0, bin_cap = 0;
    unsigned long long* minkey = nullptr;   // per node: the bin origin's (sample << 32 | ray)
    float* vtx = nullptr;            // recorded path vertices, [record slots][7]
    float* bfield = nullptr;         // product mode: per ray the BxDF's 8x8 cell luminances
    mrh_wfpg_stats st = {};
    std::vector<void*> allocs;
};

// Per-ray parameters of the bounce's guide preparation
struct WfPrep {
    float lox, loy, loz, inv_cell;   // leaf grid of the voxel octree
    uint32_t depth;
    float tmin;                      // field cone start distance
    float ratio;                     // probability of the guide's sample
    uint32_t *codes, *flag;
    float4* org;
    float2 *xi, *euv;
    float* bdir;
};

__device__ __forceinline__ uint32_t wf_hash(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ float wf_unit(uint32_t h) { return (float)(h >> 8) * (1.0f / 16777216.0f); }

// The half of KCGenAndSampleDistribution that happens per ray before the
// field is sampled (WFPGTracerKC.cuh:1480-1522): for a non-specular surface
// hit (Lambert, or a conductor of roughness >= 0.01 — the vertices the
// guide of this library serves) the ray's cell code at the hit point, the
// one-sample mixture choice (guide with probability ratio, else the BxDF's
// own sample, whose direction and concentric-octahedral (u,v) are kept for
// the guide pdf there), the guide sample's two uniforms and the bin origin
// candidate (hit point pushed off the surface, cone start tmin).  Draws come
// from the path's stream, which is advanced in place.  Flag 3 = the BxDF
// sample failed (the path ends as it would without the guide).
__global__ __launch_bounds__(256) void k_wf_prep(PathBuf in, HitBuf hit, uint32_t n, ShadeReq q, WfPrep w)
{
    for (uint32_t rid = blockIdx.x * blockDim.x + threadIdx.x; rid < n; rid += gridDim.x * blockDim.x) {
        const uint32_t key = hit.key[rid];
        const uint32_t batch = key >> 24, id = key & 0xFFFFFFu;
        const uint32_t dm = in.dm[rid];
        const uint32_t depth = dm & 0xFFu;
        const bool ok = !(dm & DM_NEE) && depth < q.max_depth &&
                        (batch == MRH_BATCH_LAMBERT ||
                         (batch == MRH_BATCH_METAL && (1.0f - q.mats.metal[(size_t)id * 10]) < 0.99f));
        if (!ok) {
            w.codes[rid] = SVO_NO_NODE;
            w.flag[rid] = 0u;
            continue;
        }
        V3 c0, c1, c2;
        const uint32_t li = hit.leaf[rid];
        leaf_corners(q, li, hit.inst ? hit.inst[rid] : 0u, c0, c1, c2);
        const V3 d = v3(in.dx[rid], in.dy[rid], in.dz[rid]);
        const Surface s = q.tbn ? make_surface_tbn(c0, c1, c2, hit.ba[rid], hit.bb[rid], d,
                                                   q.tbn + 3 * (size_t)__float_as_uint(q.leaves[3 * (size_t)li].w))
                                : make_surface(c0, c1, c2, hit.ba[rid], hit.bb[rid], d);
        const V3 wi = -normalize(d);
        Rng rng{in.rng[rid], pcg32_inc(q.sample_base + in.sample[rid])};
        uint32_t fl = 1u;
        float2 e = make_float2(0.5f, 0.5f);
        if (rng.uniform() >= w.ratio) {
            BsdfOut o;
            o.medium = 0u;
            if (batch == MRH_BATCH_LAMBERT) {
                const float* a = q.mats.lambert + (size_t)id * 3;
                lambert_sample(s, v3(a[0], a[1], a[2]), rng, o);
            } else {
                metal_sample(s, wi, q.mats.metal + (size_t)id * 10, rng, o);
            }
            const float len = sqrtf(dot(o.d, o.d));
            if (o.pdf > 0.0f && len > 0.0f && isfinite(len)) {
                const V3 bd = scale(o.d, 1.0f / len);
                dir_to_octa(v3(bd.z, bd.x, bd.y), e.x, e.y);   // y-up direction -> the map's z-up frame
                w.bdir[3 * (size_t)rid] = o.d.x;
                w.bdir[3 * (size_t)rid + 1] = o.d.y;
                w.bdir[3 * (size_t)rid + 2] = o.d.z;
                fl = 2u;
            } else {
                fl = 3u;
            }
        }
        const float x0 = rng.uniform(), x1 = rng.uniform();
        w.xi[rid] = make_float2(x0, x1);
        w.euv[rid] = e;
        in.rng[rid] = rng.state;
        w.flag[rid] = fl;
        w.codes[rid] = svo_code(s.pos, w.lox, w.loy, w.loz, w.inv_cell, w.depth);
        const V3 p = nudge(s.pos, s.geo_n);
        w.org[rid] = make_float4(p.x, p.y, p.z, w.tmin);
    }
}

// Per node of the octree, the binned ray of the smallest path (sample) id:
// (sample << 32 | ray) by a 64-bit atomic minimum, so the choice does not
// depend on the order of the rays in the bounce's buffer.  The rays are read
// in bin order, so a wave holds one or two bins: each distinct bin of the
// wave is reduced across the wave first and takes one atomic.  The loop
// bound is uniform over the block, so every lane joins the shuffles.
__global__ __launch_bounds__(256) void k_wf_minray(const uint32_t* __restrict__ count,
                                                   const uint32_t* __restrict__ offs,
                                                   const uint32_t* __restrict__ order,
                                                   const uint32_t* __restrict__ bins,
                                                   const uint32_t* __restrict__ sample,
                                                   unsigned long long* __restrict__ minkey)
{
    const uint32_t m = offs[*count];   // binned rays
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t base = blockIdx.x * blockDim.x; base < m; base += gridDim.x * blockDim.x) {
        const uint32_t p = base + threadIdx.x;
        const bool valid = p < m;
        uint32_t node = SVO_NO_NODE, lo = 0xFFFFFFFFu, hi = 0xFFFFFFFFu;
        if (valid) {
            const uint32_t r = order[p];
            node = bins[r];
            hi = sample[r];
            lo = r;
        }
        uint64_t pend = __ballot(valid);
        while (pend) {
            const int l = __ffsll((unsigned long long)pend) - 1;
            const uint32_t ln = __shfl(node, l, 64);
            const bool mine = valid && node == ln;
            const uint64_t same = __ballot(mine);
            uint32_t h = mine ? hi : 0xFFFFFFFFu, w = mine ? lo : 0xFFFFFFFFu;
            for (int off = 32; off > 0; off >>= 1) {
                const uint32_t oh = __shfl_xor(h, off, 64), ow = __shfl_xor(w, off, 64);
                if (oh < h || (oh == h && ow < w)) {
                    h = oh;
                    w = ow;
                }
            }
            if ((int)lane == l) atomicMin(minkey + ln, ((unsigned long long)h << 32) | w);
            pend &= ~same;
        }
    }
}

// Origin and jitter of every bin: the pushed-off hit point of its ray of the
// smallest path id (the reference takes a random ray of the bin,
// WFPGTracerKC.cuh:1132-1192) and two hashed uniforms of (key, bin).
__global__ void k_wf_bins(const uint32_t* __restrict__ count, const uint32_t* __restrict__ part_bins,
                          const unsigned long long* __restrict__ minkey, const float4* __restrict__ org,
                          float4* __restrict__ origins, float2* __restrict__ jitter, uint32_t key)
{
    const uint32_t nb = *count;
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        origins[b] = org[(uint32_t)minkey[part_bins[b]]];
        const uint32_t h = wf_hash(key ^ wf_hash(2u * b + 1u));
        jitter[b] = make_float2(wf_unit(h), wf_unit(wf_hash(h ^ 0x9e3779b9u)));
    }
}

// Product mode: per guided ray the luminance of its BxDF (f cos) toward the
// 64 cell centres of the 8x8 concentric-octahedral map, each direction tilted
// toward the shading normal by the cell's solid angle in radians (4 pi / 64)
// about cross(direction, normal) (BxDFProjFunc, WFPGTracerKC.cuh:1604-1629;
// ProductSampler.cuh:270-308).  One wave per ray, lane = cell (row-major,
// 8 per row), so the 256 bytes of a ray are one coalesced store.
__global__ __launch_bounds__(256) void k_wf_bxdf_field(PathBuf in, HitBuf hit, uint32_t n, ShadeReq q,
                                                       const uint32_t* __restrict__ flag, float* __restrict__ bfield)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t rid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; rid < n; rid += waves) {
        if (flag[rid] == 0u) continue;   // wave-uniform
        const uint32_t key = hit.key[rid];
        const uint32_t batch = key >> 24, id = key & 0xFFFFFFu;
        V3 c0, c1, c2;
        const uint32_t li = hit.leaf[rid];
        leaf_corners(q, li, hit.inst ? hit.inst[rid] : 0u, c0, c1, c2);
        const V3 d = v3(in.dx[rid], in.dy[rid], in.dz[rid]);
        const Surface s = q.tbn ? make_surface_tbn(c0, c1, c2, hit.ba[rid], hit.bb[rid], d,
                                                   q.tbn + 3 * (size_t)__float_as_uint(q.leaves[3 * (size_t)li].w))
                                : make_surface(c0, c1, c2, hit.ba[rid], hit.bb[rid], d);
        const V3 wi = -normalize(d);
        const V3 ns = rotate(conj(s.w2t), v3(0.0f, 0.0f, 1.0f));
        const float u = ((float)(lane & 7u) + 0.5f) * 0.125f, v = ((float)(lane >> 3) + 0.5f) * 0.125f;
        const V3 dz = octa_to_dir(u, v);
        V3 w = v3(dz.y, dz.z, dz.x);
        V3 k = cross(w, ns);
        const float kl = sqrtf(dot(k, k));
        if (kl > 0.0f) {
            // Rodrigues: w cos a + (k x w) sin a + k (k . w)(1 - cos a), k unit
            k = scale(k, 1.0f / kl);
            const float a = 4.0f * PI_F / 64.0f;
            const float ca = cosf(a), sa = sinf(a);
            w = (scale(w, ca) + scale(cross(k, w), sa)) + scale(k, dot(k, w) * (1.0f - ca));
        }
        float pb;
        V3 fc;
        if (batch == MRH_BATCH_LAMBERT) {
            const float* al = q.mats.lambert + (size_t)id * 3;
            fc = lambert_eval(s, v3(al[0], al[1], al[2]), w, pb);
        } else {
            fc = metal_eval(s, wi, q.mats.metal + (size_t)id * 10, w, pb);
        }
        const float g = luminance(fc);
        bfield[64 * (size_t)rid + lane] = isfinite(g) ? g : 0.0f;
    }
}

__global__ void k_wf_tris(const float4* __restrict__ leaves, uint32_t n, float* __restrict__ tris)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        for (int c = 0; c < 3; c++) {
            const float4 l = leaves[3 * (size_t)i + c];
            tris[9 * (size_t)i + 3 * c] = l.x;
            tris[9 * (size_t)i + 3 * c + 1] = l.y;
            tris[9 * (size_t)i + 3 * c + 2] = l.z;
        }
    }
}

int32_t wf_alloc(mrh_context* ctx, WfState& w, void** p, size_t bytes)
{
    MRH_TRY(dmalloc(ctx, p, bytes ? bytes : 4));
    w.allocs.push_back(*p);
    return MRH_OK;
}

void wf_free(WfState& w)
{
    for (void* p : w.allocs) (void)hipFree(p);
    w.allocs.clear();
    if (w.fields) (void)hipFree(w.fields);
    w.fields = nullptr;
    if (w.origins) (void)hipFree(w.origins);
    w.origins = nullptr;
    if (w.jitter) (void)hipFree(w.jitter);
    w.jitter = nullptr;
    if (w.svo) (void)mrh_svo_destroy(w.svo);
    w.svo = nullptr;
}

// GenerateGuidedDirections (WFPGTracer.cu:590-760) for the rays of one
// bounce: prepare every ray, bin the rays by octree node
// (mrh_svo_bin_rays), cone-trace each bin's field from its origin
// (mrh_wfpg_trace_fields; 64^2 texels at the first bounce, then 32^2,
// 16^2 and 8^2 from the fourth on, as the reference's kernel list by depth,
// WFPGTracer.cu:676-692) and sample every binned ray
// (mrh_wfpg_sample_fields).  Host reads: the bounce's ray count, the bin
// count and the number of binned rays.
int32_t wf_bounce(mrh_context* ctx, Frame& f, const PathBuf& in, const HitBuf& hb, const ShadeReq& q)
{
    WfState& w = *f.wf;
    uint32_t n = 0;
    MRH_HIP_CHECK(ctx, hipMemcpyAsync(&n, f.counts + f.bounce, 4, hipMemcpyDeviceToHost, ctx->stream));
    MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (n == 0) return MRH_OK;
    const uint32_t D = w.wp.svo_depth;
    WfPrep pr;
    pr.lox = w.cube[0];
    pr.loy = w.cube[1];
    pr.loz = w.cube[2];
    pr.inv_cell = (float)(1u << D) / w.cube[3];
    pr.depth = D;
    pr.tmin = 0.15f * (w.cube[3] / (float)(1u << D)) + 1e-5f;
    pr.ratio = w.wp.guide_ratio;
    pr.codes = w.codes;
    pr.flag = w.flag;
    pr.org = w.org;
    pr.xi = w.xi;
    pr.euv = w.euv;
    pr.bdir = w.bdir;
    MRH_TRY(launch(ctx, "wfpg_prep", dim3(stride_grid(ctx, n, 256)), dim3(256), 0, k_wf_prep, in, hb, n, q, pr));
    if (w.wp.product)
        MRH_TRY(launch(ctx, "wfpg_bxdf_field", dim3(stride_grid(ctx, (uint64_t)n * 64, 256)), dim3(256), 0,
                       k_wf_bxdf_field, in, hb, n, q, (const uint32_t*)w.flag, w.bfield));
    MRH_TRY(mrh_svo_bin_rays(ctx, w.svo, w.codes, n, w.wp.min_bin_level, w.wp.bin_ray_count, w.bins, w.order,
                             w.offs, w.pbins, w.pcount));
    uint32_t nb = 0;
    MRH_HIP_CHECK(ctx, hipMemcpyAsync(&nb, w.pcount, 4, hipMemcpyDeviceToHost, ctx->stream));
    MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (nb == 0) return MRH_OK;
    uint32_t binned = 0;
    MRH_HIP_CHECK(ctx, hipMemcpyAsync(&binned, w.offs + nb, 4, hipMemcpyDeviceToHost, ctx->stream));
    MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    const uint32_t side = 64u >> (f.bounce < 3u ? f.bounce : 3u);
    const size_t need = (size_t)nb * side * side;
    if (need > w.field_floats) {
        if (w.fields) MRH_HIP_CHECK(ctx, hipFree(w.fields));
        w.fields = nullptr;
        w.field_floats = 0;
        MRH_TRY(dmalloc(ctx, (void**)&w.fields, need * 4));
        w.field_floats = need;
    }
    if (nb > w.bin_cap) {
        if (w.origins) MRH_HIP_CHECK(ctx, hipFree(w.origins));
        if (w.jitter) MRH_HIP_CHECK(ctx, hipFree(w.jitter));
        w.origins = nullptr;
        w.jitter = nullptr;
        w.bin_cap = 0;
        MRH_TRY(dmalloc(ctx, (void**)&w.origins, (size_t)nb * 16));
        MRH_TRY(dmalloc(ctx, (void**)&w.jitter, (size_t)nb * 8));
        w.bin_cap = nb;
    }
    const uint32_t key = (f.seed * 0x9e3779b9u) ^ (f.sample_base * 0x85ebca6bu) ^ (f.bounce * 0xc2b2ae35u);
    MRH_TRY(fill_u32(ctx, (uint32_t*)w.minkey, 0xFFFFFFFFu, 2 * (size_t)w.nodes));
    MRH_TRY(launch(ctx, "wfpg_minray", dim3(stride_grid(ctx, binned, 256)), dim3(256), 0, k_wf_minray,
                   (const uint32_t*)w.pcount, (const uint32_t*)w.offs, (const uint32_t*)w.order,
                   (const uint32_t*)w.bins, (const uint32_t*)in.sample, w.minkey));
    MRH_TRY(launch(ctx, "wfpg_bins", dim3(stride_grid(ctx, nb, 256)), dim3(256), 0, k_wf_bins,
                   (const uint32_t*)w.pcount, (const uint32_t*)w.pbins, (const unsigned long long*)w.minkey,
                   (const float4*)w.org, w.origins, w.jitter, key));
    mrh_wfpg_trace_desc td;
    memset(&td, 0, sizeof(td));
    td.field_x = td.field_y = side;
    td.bin_count = nb;
    td.max_query_level_offset = 0;
    td.boundary_radiance = w.wp.boundary_radiance;
    td.bin_origins = (const float*)w.origins;
    td.bin_jitter = (const float*)w.jitter;
    td.node_irradiance = w.irr;
    td.node_normals = w.nrm;
    td.out_fields = w.fields;
    MRH_TRY(mrh_wfpg_trace_fields(ctx, w.svo, &td));
    mrh_wfpg_fields_desc fd;
    memset(&fd, 0, sizeof(fd));
    fd.field_x = fd.field_y = side;
    fd.bin_count = nb;
    fd.ray_count = n;
    fd.filter_alpha = w.wp.filter_alpha;
    fd.fields = w.fields;
    fd.bin_offsets = w.offs;
    fd.ray_ids = w.order;
    fd.xi = (const float*)w.xi;
    fd.out_uv = (float*)w.guv;
    fd.out_dir = w.gdir;
    fd.out_pdf = w.gpdf;
    fd.eval_uv = (const float*)w.euv;
    fd.out_eval_pdf = w.epdf;
    if (w.wp.product) {   // the product distribution takes the field's place in the mixture
        fd.bxdf_field = w.bfield;
        fd.seed = key;
    }
    MRH_TRY(mrh_wfpg_sample_fields(ctx, &fd));
    w.st.bins += nb;
    w.st.binned_rays += binned;
    w.st.guided_bounces++;
    return MRH_OK;
}

// One iteration of PathTracer::Render: trace, partition, shade
int32_t frame_bounce(mrh_context* ctx, const mrh_scene* s, Frame& f, float* image)
{
    const mrh_lbvh* b = s->bvh;
    const uint32_t bounce = f.bounce;
    const PathBuf& in = f.bufs[bounce & 1];
    const PathBuf& outb = f.bufs[(bounce + 1) & 1];
    const float tmin = (bounce == 0) ? f.cam.near_d : 0.0f;
    const float tmax = (bounce == 0) ? f.cam.far_d : FLT_MAX;
    const uint32_t grid = stride_grid(ctx, f.cap, 256);
    const uint32_t tgrid = trace_grid(ctx, f.cap, s->two_level ? 256u : ctx->trace_block);
    const float4* lv = s->two_level ? (const float4*)s->leaves : (const float4*)b->leaves;
    HitBuf hb = f.hit;
    if (s->two_level) {
        if (f.count_traversal) MRH_TRY(launch_trace_pt2<true>(ctx, tgrid, s, in, f.counts, bounce, tmin, tmax, hb));
        else MRH_TRY(launch_trace_pt2<false>(ctx, tgrid, s, in, f.counts, bounce, tmin, tmax, hb));
    } else {
        hb.inst = nullptr;
        if (f.count_traversal) MRH_TRY(launch_trace_pt<true>(ctx, tgrid, b, in, f.counts, bounce, tmin, tmax, hb));
        else MRH_TRY(launch_trace_pt<false>(ctx, tgrid, b, in, f.counts, bounce, tmin, tmax, hb));
    }
    const bool tiles = ctx->shade_partition == 1;
    if (!tiles) {
        SortArgs sa;
        sa.keys_in = f.hit.key;
        sa.vals_out = f.order;
        sa.n_max = (uint32_t)f.cap;
        sa.n_dev = f.counts + bounce;
        sa.key_bytes = 4;
        sa.begin_bit = 24;
        sa.end_bit = 32;
        sa.descending = true;
        sa.part_offsets = f.offsets;
        sa.part_begin = 24;
        MRH_TRY(radix_sort(ctx, sa));
    }
    ShadeReq q;
    q.mats = Materials{s->lambert, s->metal, s->dielectric, s->emissive, s->reflect};
    q.leaves = lv;
    q.image = image;
    q.spp = f.per_sample ? 1u : (uint32_t)f.spp;
    q.sample_base = f.sample_base;
    q.rr_start = f.rr_start;
    q.max_depth = f.max_depth;
    q.lights = s->lights;
    q.n_lights = s->n_lights;
    q.nee = f.nee;
    q.mis = f.mis;
    q.inst = s->two_level ? (const float4*)s->inst : nullptr;
    q.light_inst = s->two_level ? s->light_inst : nullptr;
    q.tbn = s->tbn;
    if (f.wf && f.guide) MRH_TRY(wf_bounce(ctx, f, in, hb, q));
    auto shade = [&](auto kernel) {
        return launch(ctx, "shade", dim3(grid), dim3(256), 0, kernel, in, outb, hb, (const uint32_t*)f.order,
                      f.counts, bounce, q, f.ga);
    };
    auto shade_t = [&](auto kernel, uint32_t tile) {
        const uint32_t tgrid_s = stride_grid(ctx, (f.cap + tile - 1) / tile * 256, 256);
        return launch(ctx, "shade", dim3(tgrid_s), dim3(256), 0, kernel, in, outb, hb, f.counts, bounce, q, f.ga);
    };
    const uint32_t v = (f.learn ? 1u : 0u) | (f.guide ? 2u : 0u) | (f.nee ? 4u : 0u);
    if (tiles) {
#define MRH_SHADE_TILES(T)                                                                     \
    switch (v) {                                                                               \
    case 0: MRH_TRY(shade_t(k_shade_tiles<false, false, false, T>, T)); break;                 \
    case 1: MRH_TRY(shade_t(k_shade_tiles<true, false, false, T>, T)); break;                  \
    case 2: MRH_TRY(shade_t(k_shade_tiles<false, true, false, T>, T)); break;                  \
    case 3: MRH_TRY(shade_t(k_shade_tiles<true, true, false, T>, T)); break;                   \
    case 4: MRH_TRY(shade_t(k_shade_tiles<false, false, true, T>, T)); break;                  \
    case 5: MRH_TRY(shade_t(k_shade_tiles<true, false, true, T>, T)); break;                   \
    case 6: MRH_TRY(shade_t(k_shade_tiles<false, true, true, T>, T)); break;                   \
    default: MRH_TRY(shade_t(k_shade_tiles<true, true, true, T>, T)); break;                   \
    }
        switch (ctx->shade_tile) {
        case 512: MRH_SHADE_TILES(512u); break;
        case 1024: MRH_SHADE_TILES(1024u); break;
        case 2048: MRH_SHADE_TILES(2048u); break;
        default: return fail(ctx, MRH_ERR_INVALID, "shade_tile must be 512, 1024 or 2048");
        }
#undef MRH_SHADE_TILES
    } else {
        switch (v) {
        case 0: MRH_TRY(shade(k_shade<false, false, false>)); break;
        case 1: MRH_TRY(shade(k_shade<true, false, false>)); break;
        case 2: MRH_TRY(shade(k_shade<false, true, false>)); break;
        case 3: MRH_TRY(shade(k_shade<true, true, false>)); break;
        case 4: MRH_TRY(shade(k_shade<false, false, true>)); break;
        case 5: MRH_TRY(shade(k_shade<true, false, true>)); break;
        case 6: MRH_TRY(shade(k_shade<false, true, true>)); break;
        default: MRH_TRY(shade(k_shade<true, true, true>)); break;
        }
    }
    f.bounce++;
    return MRH_OK;
}

// Waits for the frame and reads its per-bounce counts
int32_t frame_counts(mrh_context* ctx, const Frame& f, mrh_render_stats* stats)
{
    uint32_t hc[68];
    MRH_HIP_CHECK(ctx, hipMemcpyAsync(hc, f.counts, sizeof(hc), hipMemcpyDeviceToHost, ctx->stream));
    MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (stats) {
        stats->bounces = f.bounce;
        for (uint32_t i = 0; i < f.bounce; i++) {
            stats->segments += hc[i];
            if (i < 32) stats->bounce_rays[i] = hc[i];
        }
        stats->node_visits += (uint64_t)hc[64] | ((uint64_t)hc[65] << 32);
        stats->tri_tests += (uint64_t)hc[66] | ((uint64_t)hc[67] << 32);
    }
    return MRH_OK;
}

// Per-sample buffers of a filtered frame: radiance, image coordinates and
// the filter's two sums
struct SampleBufs {
    float* values = nullptr;     // [samples][3]
    float2* coords = nullptr;    // [samples]
    float* weighted = nullptr;   // [pixels][3], sum of w * value
    float* weights = nullptr;    // [pixels], sum of w
};

size_t sample_bytes(uint64_t total, uint64_t pixels)
{
    return align_up(total * 12) + align_up(total * 8) + align_up(pixels * 12) + align_up(pixels * 4);
}

SampleBufs carve_samples(void* base, uint64_t total, uint64_t pixels)
{
    SampleBufs b;
    char* q = (char*)base;
    b.values = (float*)q; q += align_up(total * 12);
    b.coords = (float2*)q; q += align_up(total * 8);
    b.weighted = (float*)q; q += align_up(pixels * 12);
    b.weights = (float*)q;
    return b;
}

// One frame (PathTracer::Render until done, then Finalize).  Without a
// filter every sample adds to its pixel and the image is the per-pixel
// mean; with one (RayTracer::Finalize -> FilterToImg, RayTracer.cu:81-95)
// every sample keeps its radiance and image coordinate and the filter
// gathers them into the image.
int32_t render(mrh_context* ctx, mrh_scene* s, const mrh_camera_desc* cd, const mrh_render_params* p,
               const mrh_filter_desc* fd, float* d_image, mrh_render_stats* stats)
{
    uint64_t total = 0;
    MRH_TRY(validate_params(ctx, s, p, &total));
    FilterDev filt;
    if (fd) MRH_TRY(filter_make(ctx, fd, &filt));
    void* arena = nullptr;
    MRH_TRY(scratch(ctx, SLOT_FRAME, frame_bytes(frame_cap(total, p->nee)), &arena));
    const size_t pixels = (size_t)p->width * p->height;
    const size_t img_count = pixels * 3;
    SampleBufs sb;
    if (fd) {
        void* mem = nullptr;
        MRH_TRY(scratch(ctx, SLOT_SAMPLES, sample_bytes(total, pixels), &mem));
        sb = carve_samples(mem, total, pixels);
        MRH_TRY(fill_u32(ctx, (uint32_t*)sb.values, 0u, total * 3));
        MRH_TRY(fill_u32(ctx, (uint32_t*)sb.weighted, 0u, img_count));
        MRH_TRY(fill_u32(ctx, (uint32_t*)sb.weights, 0u, pixels));
    } else {
        MRH_TRY(fill_u32(ctx, (uint32_t*)d_image, 0u, img_count));
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    MRH_HIP_CHECK(ctx, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(ctx, MRH_ERR_HIP, "mrh_render: event create failed");
    }
    MRH_HIP_CHECK(ctx, hipEventRecord(e0, ctx->stream));
    Frame f;
    int32_t rc = frame_begin(ctx, cd, p, arena, 0u, f, fd ? sb.coords : nullptr);
    while (rc == MRH_OK && f.bounce < f.max_depth) rc = frame_bounce(ctx, s, f, fd ? sb.values : d_image);
    if (rc == MRH_OK && fd) {
        rc = filter_gather(ctx, filt, sb.values, sb.coords, nullptr, nullptr, (uint32_t)f.spp, p->width, p->height,
                           sb.weighted, sb.weights);
        if (rc == MRH_OK) rc = filter_resolve(ctx, sb.weighted, sb.weights, pixels, d_image);
    } else if (rc == MRH_OK) {
        rc = launch(ctx, "finalize", dim3(stride_grid(ctx, img_count, 256)), dim3(256), 0, k_scale, d_image,
                    img_count, 1.0f / (float)f.spp);
    }
    if (rc == MRH_OK && hipEventRecord(e1, ctx->stream) != hipSuccess)
        rc = fail(ctx, MRH_ERR_HIP, "mrh_render: event record failed");
    if (rc == MRH_OK) {
        mrh_render_stats local;
        memset(&local, 0, sizeof(local));
        rc = frame_counts(ctx, f, &local);
        if (rc == MRH_OK && stats) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) ms = 0.0f;
            local.render_ms = ms;
            *stats = local;
        }
    } else {
        (void)hipStreamSynchronize(ctx->stream);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

// Learn/guide state of one guided render: path-vertex records of the frame
// being learned from, the octree built from the last learned frame, and the
// cube of its voxel grid.
struct GuideState {
    mrh_guide_params gp = {};
    float cube[4] = {0, 0, 0, 1};
    uint64_t n_rec = 0;
    void* recmem = nullptr;
    RecBuf rec = {};
    uint32_t *c_codes = nullptr, *c_bins = nullptr;
    float* c_rad = nullptr;
    mrh_svo* svo = nullptr;
    bool own_svo = false, own_records = false, have_guide = false;
};

int32_t guide_validate(mrh_context* ctx, const mrh_guide_params* gp, uint64_t total, uint32_t max_depth)
{
    if (gp->iterations == 0 || gp->iterations > 64)
        return fail(ctx, MRH_ERR_INVALID, "guide: iterations must be in [1, 64]");
    if (gp->svo_depth == 0 || gp->svo_depth > 10)
        return fail(ctx, MRH_ERR_INVALID, "guide: svo_depth must be in [1, 10]");
    if (!(gp->guide_ratio >= 0.0f && gp->guide_ratio <= 1.0f))
        return fail(ctx, MRH_ERR_INVALID, "guide: guide_ratio must be in [0, 1]");
    if ((uint64_t)gp->iterations * total >= (1ull << 32))
        return fail(ctx, MRH_ERR_INVALID, "guide: iterations x paths exceed 2^32 samples");
    if (total * max_depth >= (1ull << 31))
        return fail(ctx, MRH_ERR_INVALID, "guide: %llu path vertices exceed 2^31",
                    (unsigned long long)(total * max_depth));
    return MRH_OK;
}

void guide_free(GuideState& g)
{
    if (g.recmem && g.own_records) (void)hipFree(g.recmem);
    g.recmem = nullptr;
    if (g.own_svo && g.svo) svo_free(g.svo);
    g.svo = nullptr;
}

// Cube = scene bounds padded by one voxel; records for frames of `total`
// paths; `svo` (may be null) receives the guides.  With `persistent` the
// state owns its records and octree (the incremental tracer keeps them
// across calls); otherwise they live in the context's reusable scratch.
int32_t guide_init(mrh_context* ctx, const mrh_scene* s, uint64_t total, uint32_t max_depth,
                   const mrh_guide_params* gp, mrh_svo* svo, bool persistent, GuideState& g)
{
    g.gp = *gp;
    const mrh_lbvh* b = s->bvh;
    float span = 0.0f;
    for (int a = 0; a < 3; a++) span = fmaxf(span, b->hi[a] - b->lo[a]);
    if (!(span > 0.0f)) span = 1.0f;
    const float cells = (float)(1u << gp->svo_depth);
    const float pad = span / (cells - 2.0f > 1.0f ? cells - 2.0f : 1.0f);
    g.cube[0] = b->lo[0] - pad;
    g.cube[1] = b->lo[1] - pad;
    g.cube[2] = b->lo[2] - pad;
    g.cube[3] = span + 2.0f * pad;
    if (svo) {
        g.svo = svo;
    } else if (persistent) {
        MRH_TRY(mrh_svo_create(ctx, &g.svo));
        g.own_svo = true;
    } else {
        if (!ctx->guide_svo) MRH_TRY(mrh_svo_create(ctx, &ctx->guide_svo));
        g.svo = ctx->guide_svo;
    }
    if (gp->iterations > 1) {
        g.n_rec = total * max_depth;
        if (persistent) {
            MRH_TRY(dmalloc(ctx, &g.recmem, (size_t)g.n_rec * 36));
            g.own_records = true;
        } else {
            MRH_TRY(scratch(ctx, SLOT_GUIDE_REC, (size_t)g.n_rec * 36, &g.recmem));
        }
        char* q = (char*)g.recmem;
        g.rec.code = (uint32_t*)q; q += g.n_rec * 4;
        g.rec.bin = (uint32_t*)q; q += g.n_rec * 4;
        g.rec.tr = (float*)q; q += g.n_rec * 4;
        g.rec.tg = (float*)q; q += g.n_rec * 4;
        g.rec.tb = (float*)q; q += g.n_rec * 4;
        g.rec.rad = (float*)q; q += g.n_rec * 4;
        g.c_codes = (uint32_t*)q; q += g.n_rec * 4;
        g.c_bins = (uint32_t*)q; q += g.n_rec * 4;
        g.c_rad = (float*)q;
        g.rec.max_depth = max_depth;
        g.rec.stride = (uint32_t)total;
    }
    return MRH_OK;
}

// Sets the frame's learn / guide switches; frame `index` (0-based) learns
// while index + 1 < iterations.
int32_t guide_frame(mrh_context* ctx, GuideState& g, uint32_t index, Frame& f)
{
    f.learn = g.recmem && index + 1 < g.gp.iterations;
    f.guide = g.have_guide;
    f.ga = GuideArgs{};
    f.ga.rec = g.rec;
    f.ga.lox = g.cube[0];
    f.ga.loy = g.cube[1];
    f.ga.loz = g.cube[2];
    f.ga.inv_cell = (float)(1u << g.gp.svo_depth) / g.cube[3];
    f.ga.depth = g.gp.svo_depth;
    f.ga.ratio = g.gp.guide_ratio;
    f.ga.min_samples = g.gp.min_samples;
    if (g.have_guide) f.ga.svo = svo_view(g.svo);
    if (f.learn) MRH_TRY(fill_u32(ctx, g.rec.code, SVO_NO_NODE, g.n_rec));
    return MRH_OK;
}

// After a learning frame: rebuild the octree over the recorded cells, splat
// and filter (WFPGTracer::Finalize, WFPGTracer.cu:1699-1790)
int32_t guide_learn(mrh_context* ctx, GuideState& g)
{
    uint32_t kept = 0;
    MRH_TRY(svo_compact_records(ctx, g.svo, g.rec.code, g.rec.bin, g.rec.rad, (uint32_t)g.n_rec, g.c_codes, g.c_bins,
                                g.c_rad, &kept));
    if (kept == 0) return MRH_OK;   // nothing learned: keep the previous guide
    MRH_TRY(svo_build(ctx, g.svo, g.c_codes, kept, g.cube, g.gp.svo_depth, true));
    MRH_TRY(svo_splat(ctx, g.svo, g.c_codes, g.c_bins, g.c_rad, kept));
    MRH_TRY(svo_filter(ctx, g.svo));
    g.have_guide = true;
    return MRH_OK;
}

// WFPGTracer::Render / Finalize (WFPGTracer.cu:1590-1790) over the structure
// of this library: `iterations` frames of the same view, each with
// samples^2 spp and its own sample range.  Every frame but the last records
// its path vertices; after it the octree is rebuilt over the recorded
// vertices' cells, their radiance splatted and filtered, and the next frame
// samples directions from the one-sample mixture of that guide and the BxDF.
// The image is the mean over all frames.
int32_t render_guided(mrh_context* ctx, mrh_scene* s, const mrh_camera_desc* cd, const mrh_render_params* p,
                      const mrh_guide_params* gp, mrh_svo* svo, float* d_image, mrh_render_stats* stats)
{
    uint64_t total = 0;
    MRH_TRY(validate_params(ctx, s, p, &total));
    MRH_TRY(guide_validate(ctx, gp, total, p->max_depth));
    void* arena = nullptr;
    MRH_TRY(scratch(ctx, SLOT_FRAME, frame_bytes(frame_cap(total, p->nee)), &arena));
    const size_t img_count = (size_t)p->width * p->height * 3;
    MRH_TRY(fill_u32(ctx, (uint32_t*)d_image, 0u, img_count));
    GuideState g;
    int32_t rc = guide_init(ctx, s, total, p->max_depth, gp, svo, false, g);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == MRH_OK && hipEventCreate(&e0) != hipSuccess) rc = fail(ctx, MRH_ERR_HIP, "event create");
    if (rc == MRH_OK && hipEventCreate(&e1) != hipSuccess) rc = fail(ctx, MRH_ERR_HIP, "event create");
    if (rc == MRH_OK) (void)hipEventRecord(e0, ctx->stream);
    mrh_render_stats local;
    memset(&local, 0, sizeof(local));
    for (uint32_t it = 0; it < gp->iterations && rc == MRH_OK; it++) {
        Frame f;
        rc = frame_begin(ctx, cd, p, arena, (uint32_t)(it * total), f);
        if (rc == MRH_OK) rc = guide_frame(ctx, g, it, f);
        while (rc == MRH_OK && f.bounce < f.max_depth) rc = frame_bounce(ctx, s, f, d_image);
        if (rc == MRH_OK) rc = frame_counts(ctx, f, &local);
        if (rc == MRH_OK && f.learn) rc = guide_learn(ctx, g);
    }
    if (rc == MRH_OK)
        rc = launch(ctx, "finalize", dim3(stride_grid(ctx, img_count, 256)), dim3(256), 0, k_scale, d_image,
                    img_count, 1.0f / ((float)((uint64_t)p->samples * p->samples) * (float)gp->iterations));
    if (rc == MRH_OK) {
        (void)hipEventRecord(e1, ctx->stream);
        const hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = fail(ctx, MRH_ERR_HIP, "render_guided: %s", hipGetErrorString(e));
        float ms = 0.0f;
        if (rc == MRH_OK) (void)hipEventElapsedTime(&ms, e0, e1);
        local.render_ms = ms;
        if (rc == MRH_OK && stats) *stats = local;
    }
    (void)hipStreamSynchronize(ctx->stream);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    guide_free(g);
    return rc;
}

// WFPGTracer with its own guide (WFPGTracer.cu:590-760, 1590-1790): the
// scene's triangles are voxelised into an octree once; `iterations` frames
// of the same view follow, each with samples^2 spp and its own sample range.
// Every frame but the last records its path vertices (position, direction
// toward the previous vertex, luminance of the radiance that came back), and
// the octree's payload is formed from them when it ends; every frame after
// the first guides its non-specular vertices with the fields cone-traced
// through that payload (wf_bounce), one-sample MIS with the BxDF.  The image
// is the mean over all frames.
// Checks of the field-traced guide's parameters (render and tracer)
int32_t wf_validate(mrh_context* ctx, const mrh_scene* s, const mrh_wfpg_params* wp, uint64_t total,
                    uint32_t max_depth, mrh_guide_params& gp)
{
    if (s->two_level) return fail(ctx, MRH_ERR_UNSUPPORTED, "wfpg: two-level scenes are not supported");
    if (wp->min_bin_level > 10) return fail(ctx, MRH_ERR_INVALID, "wfpg: min_bin_level must be in [0, 10]");
    if (wp->bin_ray_count == 0 || wp->bin_ray_count > MRH_BIN_RAY_COUNT_MAX)
        return fail(ctx, MRH_ERR_INVALID, "wfpg: bin_ray_count must be in [1, %u]", MRH_BIN_RAY_COUNT_MAX);
    if (!(wp->filter_alpha > 0.0f) || !std::isfinite(wp->filter_alpha))
        return fail(ctx, MRH_ERR_INVALID, "wfpg: filter_alpha must be finite and > 0");
    if (wp->product > 1) return fail(ctx, MRH_ERR_INVALID, "wfpg: product must be 0 or 1");
    if (!(wp->boundary_radiance >= 0.0f) || !std::isfinite(wp->boundary_radiance))
        return fail(ctx, MRH_ERR_INVALID, "wfpg: boundary_radiance must be finite and >= 0");
    memset(&gp, 0, sizeof(gp));
    gp.iterations = wp->iterations;
    gp.svo_depth = wp->svo_depth;
    gp.guide_ratio = wp->guide_ratio;
    MRH_TRY(guide_validate(ctx, &gp, total, max_depth));
    if (wp->iterations > 1 && total * max_depth >= (1ull << 30))
        return fail(ctx, MRH_ERR_INVALID, "wfpg: %llu path-vertex slots exceed the payload's 2^30",
                    (unsigned long long)(total * max_depth));
    return MRH_OK;
}

// The voxel octree over the scene's triangles and the per-ray, per-node and
// per-vertex buffers of the field-traced guide (rays per bounce `cap`,
// path-vertex slots `n_rec`)
int32_t wf_setup(mrh_context* ctx, const mrh_scene* s, const mrh_wfpg_params* wp, uint64_t cap, uint64_t n_rec,
                 WfState& w)
{
    w.wp = *wp;
    // AnisoSVOctreeCPU::Constrcut (AnisoSVO.cu:670-678): the scene's
    // largest span padded by one leaf on every side
    const mrh_lbvh* b = s->bvh;
    float span = 0.0f;
    for (int a = 0; a < 3; a++) span = fmaxf(span, b->hi[a] - b->lo[a]);
    const float pad = span / (float)(1u << wp->svo_depth);
    for (int a = 0; a < 3; a++) w.cube[a] = b->lo[a] - pad;
    w.cube[3] = span + 2.0f * pad;
    w.n_tris = b->n;
    MRH_TRY(wf_alloc(ctx, w, (void**)&w.tris, (size_t)w.n_tris * 36));
    MRH_TRY(launch(ctx, "wfpg_tris", dim3(stride_grid(ctx, w.n_tris, 256)), dim3(256), 0, k_wf_tris,
                   (const float4*)b->leaves, w.n_tris, w.tris));
    MRH_TRY(mrh_svo_voxelize(ctx, w.tris, w.n_tris, w.cube, wp->svo_depth, nullptr, nullptr, 0, &w.n_vox));
    if (w.n_vox == 0) return fail(ctx, MRH_ERR_INVALID, "mrh_render_wfpg: the scene voxelises to nothing");
    MRH_TRY(wf_alloc(ctx, w, (void**)&w.vcodes, (size_t)w.n_vox * 4));
    MRH_TRY(wf_alloc(ctx, w, (void**)&w.vprims, (size_t)w.n_vox * 4));
    uint32_t again = 0;
    MRH_TRY(mrh_svo_voxelize(ctx, w.tris, w.n_tris, w.cube, wp->svo_depth, w.vcodes, w.vprims, w.n_vox,
                             &again));
    MRH_TRY(mrh_svo_create(ctx, &w.svo));
    MRH_TRY(mrh_svo_build(ctx, w.svo, w.vcodes, w.n_vox, w.cube, wp->svo_depth));
    const SvoView v = svo_view(w.svo);
    w.nodes = v.lvl_off[v.depth + 1];
    MRH_TRY(wf_alloc(ctx, w, (void**)&w.irr, (size_t)w.nodes * 8));
    MRH_TRY(wf_alloc(ctx, w, (void**)&w.nrm, (size_t)w.nodes * 12));
    MRH_TRY(wf_alloc(ctx, w, (void**)&w.minkey, (size_t)w.nodes * 8));
    MRH_TRY(wf_alloc(ctx, w, (void**)&w.codes, cap * 4));
    MRH_TRY(wf_alloc(ctx, w, (void**)&w.flag, cap * 4));
    MRH_TRY(wf_alloc(ctx, w, (void**)&w.bins, cap * 4));
    MRH_TRY(wf_alloc(ctx, w, (void**)&w.order, cap * 4));
    MRH_TRY(wf_alloc(ctx, w, (void**)&w.offs, (cap + 1) * 4));
    MRH_TRY(wf_alloc(ctx, w, (void**)&w.pbins, cap * 4));
    MRH_TRY(wf_alloc(ctx, w, (void**)&w.pcount, 4));
    MRH_TRY(wf_alloc(ctx, w, (void**)&w.org, cap * 16));
    MRH_TRY(wf_alloc(ctx, w, (void**)&w.xi, cap * 8));
    MRH_TRY(wf_alloc(ctx, w, (void**)&w.euv, cap * 8));
    MRH_TRY(wf_alloc(ctx, w, (void**)&w.guv, cap * 8));
    MRH_TRY(wf_alloc(ctx, w, (void**)&w.bdir, cap * 12));
    MRH_TRY(wf_alloc(ctx, w, (void**)&w.gdir, cap * 12));
    MRH_TRY(wf_alloc(ctx, w, (void**)&w.gpdf, cap * 4));
    MRH_TRY(wf_alloc(ctx, w, (void**)&w.epdf, cap * 4));
    if (wp->product) MRH_TRY(wf_alloc(ctx, w, (void**)&w.bfield, cap * 256));
    if (n_rec) MRH_TRY(wf_alloc(ctx, w, (void**)&w.vtx, (size_t)n_rec * 28));
    return MRH_OK;
    return MRH_OK;
}

// Points a frame at the field-traced guide: guided once a payload exists,
// its path vertices recorded beside the records
void wf_frame(WfState& w, Frame& f)
{
    f.wf = &w;
    f.guide = w.have_payload ? 1u : 0u;
    f.ga.rec.vtx = w.vtx;
    f.ga.wf_flag = w.flag;
    f.ga.wf_gdir = w.gdir;
    f.ga.wf_gpdf = w.gpdf;
    f.ga.wf_bdir = w.bdir;
    f.ga.wf_epdf = w.epdf;
}

// After a learning frame: the octree's payload from its path vertices
int32_t wf_learn(mrh_context* ctx, WfState& w, const GuideState& g)
{
    mrh_svo_payload_desc pd;
    memset(&pd, 0, sizeof(pd));
    pd.tris = w.tris;
    pd.codes = w.vcodes;
    pd.prims = w.vprims;
    pd.n_records = w.n_vox;
    pd.vertices = w.vtx;
    pd.n_vertices = (uint32_t)g.n_rec;
    pd.vertex_mask = g.rec.code;   // slots without a vertex keep SVO_NO_NODE
    pd.node_irradiance = w.irr;
    pd.node_normals = w.nrm;
    MRH_TRY(mrh_svo_payload(ctx, w.svo, &pd));
    w.have_payload = true;
    return MRH_OK;
}

int32_t render_wfpg(mrh_context* ctx, mrh_scene* s, const mrh_camera_desc* cd, const mrh_render_params* p,
                    const mrh_wfpg_params* wp, float* d_image, mrh_render_stats* stats, mrh_wfpg_stats* wstats)
{
    uint64_t total = 0;
    MRH_TRY(validate_params(ctx, s, p, &total));
    mrh_guide_params gp;
    MRH_TRY(wf_validate(ctx, s, wp, total, p->max_depth, gp));
    void* arena = nullptr;
    MRH_TRY(scratch(ctx, SLOT_FRAME, frame_bytes(frame_cap(total, p->nee)), &arena));
    const size_t img_count = (size_t)p->width * p->height * 3;
    MRH_TRY(fill_u32(ctx, (uint32_t*)d_image, 0u, img_count));
    GuideState g;
    WfState w;
    int32_t rc = guide_init(ctx, s, total, p->max_depth, &gp, nullptr, false, g);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == MRH_OK && hipEventCreate(&e0) != hipSuccess) rc = fail(ctx, MRH_ERR_HIP, "event create");
    if (rc == MRH_OK && hipEventCreate(&e1) != hipSuccess) rc = fail(ctx, MRH_ERR_HIP, "event create");
    if (rc == MRH_OK) (void)hipEventRecord(e0, ctx->stream);
    if (rc == MRH_OK) rc = wf_setup(ctx, s, wp, frame_cap(total, p->nee), g.n_rec, w);
    mrh_render_stats local;
    memset(&local, 0, sizeof(local));
    for (uint32_t it = 0; it < wp->iterations && rc == MRH_OK; it++) {
        Frame f;
        rc = frame_begin(ctx, cd, p, arena, (uint32_t)(it * total), f);
        if (rc == MRH_OK) rc = guide_frame(ctx, g, it, f);
        if (rc != MRH_OK) break;
        wf_frame(w, f);
        while (rc == MRH_OK && f.bounce < f.max_depth) rc = frame_bounce(ctx, s, f, d_image);
        if (rc == MRH_OK) rc = frame_counts(ctx, f, &local);
        if (rc == MRH_OK && f.learn) rc = wf_learn(ctx, w, g);
    }
    if (rc == MRH_OK)
        rc = launch(ctx, "finalize", dim3(stride_grid(ctx, img_count, 256)), dim3(256), 0, k_scale, d_image,
                    img_count, 1.0f / ((float)((uint64_t)p->samples * p->samples) * (float)wp->iterations));
    if (rc == MRH_OK) {
        (void)hipEventRecord(e1, ctx->stream);
        const hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = fail(ctx, MRH_ERR_HIP, "render_wfpg: %s", hipGetErrorString(e));
        float ms = 0.0f;
        if (rc == MRH_OK) (void)hipEventElapsedTime(&ms, e0, e1);
        local.render_ms = ms;
        if (rc == MRH_OK && stats) *stats = local;
        w.st.octree_nodes = w.nodes;
        w.st.voxel_records = w.n_vox;
        if (rc == MRH_OK && wstats) *wstats = w.st;
    }
    (void)hipStreamSynchronize(ctx->stream);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    wf_free(w);
    guide_free(g);
    return rc;
}

__global__ void k_mean(const float* __restrict__ acc, float* __restrict__ out, size_t count, float s)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        out[i] = acc[i] * s;
}

}  // namespace
}  // namespace mrh

using namespace mrh;

extern "C" {

int32_t mrh_scene_create(mrh_context* ctx, const mrh_scene_desc* desc, mrh_scene** out)
{
    if (!ctx || !desc || !out) return MRH_ERR_INVALID;
    *out = nullptr;
    return scene_create(ctx, desc, out);
}

int32_t mrh_scene_destroy(mrh_scene* scene)
{
    if (!scene) return MRH_ERR_INVALID;
    scene_free(scene);
    return MRH_OK;
}

int32_t mrh_scene_bvh(const mrh_scene* scene, const mrh_lbvh** out)
{
    if (!scene || !out) return MRH_ERR_INVALID;
    if (scene->two_level) return fail(scene->ctx, MRH_ERR_UNSUPPORTED, "mrh_scene_bvh: a two-level scene has no single LBVH");
    *out = scene->bvh;
    return MRH_OK;
}

int32_t mrh_scene_bytes(const mrh_scene* scene, uint64_t* bytes)
{
    if (!scene || !bytes) return MRH_ERR_INVALID;
    *bytes = scene->bytes;
    return MRH_OK;
}

int32_t mrh_render(mrh_context* ctx, mrh_scene* scene, const mrh_camera_desc* cam,
                   const mrh_render_params* params, float* d_image, mrh_render_stats* stats)
{
    if (!ctx || !scene || !cam || !params || !d_image) return MRH_ERR_INVALID;
    if (scene->ctx != ctx) return fail(ctx, MRH_ERR_INVALID, "mrh_render: scene belongs to another context");
    return render(ctx, scene, cam, params, nullptr, d_image, stats);
}

int32_t mrh_render_filtered(mrh_context* ctx, mrh_scene* scene, const mrh_camera_desc* cam,
                            const mrh_render_params* params, const mrh_filter_desc* filter, float* d_image,
                            mrh_render_stats* stats)
{
    if (!ctx || !scene || !cam || !params || !filter || !d_image) return MRH_ERR_INVALID;
    if (scene->ctx != ctx)
        return fail(ctx, MRH_ERR_INVALID, "mrh_render_filtered: scene belongs to another context");
    return render(ctx, scene, cam, params, filter, d_image, stats);
}

int32_t mrh_render_wfpg(mrh_context* ctx, mrh_scene* scene, const mrh_camera_desc* cam,
                        const mrh_render_params* params, const mrh_wfpg_params* wfpg, float* d_image,
                        mrh_render_stats* stats, mrh_wfpg_stats* wstats)
{
    if (!ctx || !scene || !cam || !params || !wfpg || !d_image) return MRH_ERR_INVALID;
    if (scene->ctx != ctx) return fail(ctx, MRH_ERR_INVALID, "mrh_render_wfpg: scene belongs to another context");
    return render_wfpg(ctx, scene, cam, params, wfpg, d_image, stats, wstats);
}

int32_t mrh_render_guided(mrh_context* ctx, mrh_scene* scene, const mrh_camera_desc* cam,
                          const mrh_render_params* params, const mrh_guide_params* guide, mrh_svo* svo,
                          float* d_image, mrh_render_stats* stats)
{
    if (!ctx || !scene || !cam || !params || !guide || !d_image) return MRH_ERR_INVALID;
    if (scene->ctx != ctx) return fail(ctx, MRH_ERR_INVALID, "mrh_render_guided: scene belongs to another context");
    if (svo && svo_ctx(svo) != ctx)
        return fail(ctx, MRH_ERR_INVALID, "mrh_render_guided: svo belongs to another context");
    return render_guided(ctx, scene, cam, params, guide, svo, d_image, stats);
}

int32_t mrh_render_bytes(const mrh_render_params* params, uint64_t* bytes)
{
    if (!params || !bytes) return MRH_ERR_INVALID;
    const uint64_t total = (uint64_t)params->width * params->height * params->samples * params->samples;
    const uint64_t cap = frame_cap(total, params->nee);
    *bytes = frame_bytes(cap) + cap * 16;   // + sort scratch
    return MRH_OK;
}

}  // extern "C"

// --------------------------------------------------------------------------
// Incremental tracer: the GPUTracerI protocol (GenerateWork, Render() until it
// returns false, Finalize) over a resident scene, with progressive
// accumulation across frames.
// --------------------------------------------------------------------------
struct mrh_tracer {
    mrh_context* ctx = nullptr;
    mrh_scene* scene = nullptr;
    mrh_render_params p = {};
    void* arena = nullptr;
    float* accum = nullptr;
    uint64_t total = 0;
    uint32_t frames = 0;
    bool active = false;
    mrh::Frame f;
    mrh_render_stats stats = {};
    bool guided = false;
    uint32_t guide_frames = 0;   // frames rendered since guiding was set
    mrh::GuideState g;
    bool wfpg = false;           // the guide is the field-traced one (mrh_tracer_set_wfpg)
    mrh::WfState w;
    // attached reconstruction filter: per-sample buffers of the frame in
    // flight and the accumulated filter sums
    bool filtered = false;
    mrh::FilterDev filt = {};
    void* smem = nullptr;
    mrh::SampleBufs sb;
};

extern "C" {

int32_t mrh_tracer_create(mrh_context* ctx, mrh_scene* scene, const mrh_render_params* params,
                          mrh_tracer** out)
{
    if (!ctx || !scene || !params || !out) return MRH_ERR_INVALID;
    *out = nullptr;
    if (scene->ctx != ctx) return fail(ctx, MRH_ERR_INVALID, "mrh_tracer_create: scene belongs to another context");
    uint64_t total = 0;
    MRH_TRY(validate_params(ctx, scene, params, &total));
    mrh_tracer* t = new mrh_tracer();
    t->ctx = ctx;
    t->scene = scene;
    t->p = *params;
    t->total = total;
    const size_t img = (size_t)params->width * params->height * 3;
    int32_t rc = dmalloc(ctx, &t->arena, frame_bytes(frame_cap(total, params->nee)));
    if (rc == MRH_OK) rc = dmalloc(ctx, (void**)&t->accum, img * 4);
    if (rc == MRH_OK) rc = fill_u32(ctx, (uint32_t*)t->accum, 0u, img);
    if (rc != MRH_OK) {
        (void)hipFree(t->arena);
        (void)hipFree(t->accum);
        delete t;
        return rc;
    }
    *out = t;
    return MRH_OK;
}

int32_t mrh_tracer_generate_work(mrh_tracer* t, const mrh_camera_desc* cam)
{
    if (!t || !cam) return MRH_ERR_INVALID;
    mrh_context* ctx = t->ctx;
    if (t->active) return fail(ctx, MRH_ERR_INVALID, "mrh_tracer_generate_work: the previous frame is still rendering");
    if (t->filtered) MRH_TRY(fill_u32(ctx, (uint32_t*)t->sb.values, 0u, t->total * 3));
    MRH_TRY(frame_begin(ctx, cam, &t->p, t->arena, (uint32_t)(t->frames * t->total), t->f,
                        t->filtered ? t->sb.coords : nullptr));
    if (t->guided) MRH_TRY(guide_frame(ctx, t->g, t->guide_frames, t->f));
    if (t->wfpg) wf_frame(t->w, t->f);
    t->active = true;
    return MRH_OK;
}

int32_t mrh_tracer_render(mrh_tracer* t, int32_t* more)
{
    if (!t || !more) return MRH_ERR_INVALID;
    *more = 0;
    if (!t->active) return MRH_OK;
    MRH_TRY(frame_bounce(t->ctx, t->scene, t->f, t->filtered ? t->sb.values : t->accum));
    if (t->f.bounce < t->f.max_depth) {
        *more = 1;
        return MRH_OK;
    }
    if (t->filtered)
        MRH_TRY(filter_gather(t->ctx, t->filt, t->sb.values, t->sb.coords, nullptr, nullptr, (uint32_t)t->f.spp,
                              t->p.width, t->p.height, t->sb.weighted, t->sb.weights));
    MRH_TRY(frame_counts(t->ctx, t->f, &t->stats));
    t->active = false;
    t->frames++;
    if (t->guided) {
        t->guide_frames++;
        if (t->f.learn) MRH_TRY(t->wfpg ? wf_learn(t->ctx, t->w, t->g) : guide_learn(t->ctx, t->g));
    }
    return MRH_OK;
}

int32_t mrh_tracer_finalize(mrh_tracer* t, float* d_image)
{
    if (!t || !d_image) return MRH_ERR_INVALID;
    mrh_context* ctx = t->ctx;
    if (t->active) return fail(ctx, MRH_ERR_INVALID, "mrh_tracer_finalize: frame still rendering");
    const size_t img = (size_t)t->p.width * t->p.height * 3;
    if (t->filtered) {
        MRH_TRY(filter_resolve(ctx, t->sb.weighted, t->sb.weights, img / 3, d_image));
    } else {
        const float s = t->frames ? 1.0f / (float)((uint64_t)t->frames * t->f.spp) : 0.0f;
        MRH_TRY(launch(ctx, "finalize", dim3(stride_grid(ctx, img, 256)), dim3(256), 0, k_mean,
                       (const float*)t->accum, d_image, img, s));
    }
    MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return MRH_OK;
}

int32_t mrh_tracer_set_guide(mrh_tracer* t, const mrh_guide_params* guide, mrh_svo* svo)
{
    if (!t || !guide) return MRH_ERR_INVALID;
    mrh_context* ctx = t->ctx;
    if (t->active) return fail(ctx, MRH_ERR_INVALID, "mrh_tracer_set_guide: frame still rendering");
    if (svo && svo_ctx(svo) != ctx) return fail(ctx, MRH_ERR_INVALID, "mrh_tracer_set_guide: svo of another context");
    MRH_TRY(guide_validate(ctx, guide, t->total, t->p.max_depth));
    MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    guide_free(t->g);
    t->g = GuideState{};
    wf_free(t->w);
    t->w = WfState{};
    t->guided = t->wfpg = false;
    MRH_TRY(guide_init(ctx, t->scene, t->total, t->p.max_depth, guide, svo, true, t->g));
    t->guided = true;
    t->guide_frames = 0;
    return MRH_OK;
}

int32_t mrh_tracer_set_wfpg(mrh_tracer* t, const mrh_wfpg_params* wfpg)
{
    if (!t || !wfpg) return MRH_ERR_INVALID;
    mrh_context* ctx = t->ctx;
    if (t->active) return fail(ctx, MRH_ERR_INVALID, "mrh_tracer_set_wfpg: frame still rendering");
    mrh_guide_params gp;
    MRH_TRY(wf_validate(ctx, t->scene, wfpg, t->total, t->p.max_depth, gp));
    MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    guide_free(t->g);
    t->g = GuideState{};
    wf_free(t->w);
    t->w = WfState{};
    t->guided = t->wfpg = false;
    MRH_TRY(guide_init(ctx, t->scene, t->total, t->p.max_depth, &gp, nullptr, true, t->g));
    const int32_t rc = wf_setup(ctx, t->scene, wfpg, frame_cap(t->total, t->p.nee), t->g.n_rec, t->w);
    if (rc != MRH_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        wf_free(t->w);
        t->w = WfState{};
        guide_free(t->g);
        t->g = GuideState{};
        return rc;
    }
    t->guided = t->wfpg = true;
    t->guide_frames = 0;
    return MRH_OK;
}

int32_t mrh_tracer_reset_image(mrh_tracer* t)
{
    if (!t) return MRH_ERR_INVALID;
    if (t->active) return fail(t->ctx, MRH_ERR_INVALID, "mrh_tracer_reset_image: frame still rendering");
    t->frames = 0;
    t->stats = mrh_render_stats{};
    const size_t pixels = (size_t)t->p.width * t->p.height;
    if (t->filtered) {
        MRH_TRY(fill_u32(t->ctx, (uint32_t*)t->sb.weighted, 0u, pixels * 3));
        MRH_TRY(fill_u32(t->ctx, (uint32_t*)t->sb.weights, 0u, pixels));
    }
    return fill_u32(t->ctx, (uint32_t*)t->accum, 0u, pixels * 3);
}

int32_t mrh_tracer_set_filter(mrh_tracer* t, const mrh_filter_desc* filter)
{
    if (!t) return MRH_ERR_INVALID;
    mrh_context* ctx = t->ctx;
    if (t->active) return fail(ctx, MRH_ERR_INVALID, "mrh_tracer_set_filter: frame still rendering");
    FilterDev f = {};
    if (filter) MRH_TRY(filter_make(ctx, filter, &f));
    MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (t->smem) (void)hipFree(t->smem);
    t->smem = nullptr;
    t->sb = SampleBufs{};
    t->filtered = false;
    if (filter) {
        const uint64_t pixels = (uint64_t)t->p.width * t->p.height;
        MRH_TRY(dmalloc(ctx, &t->smem, sample_bytes(t->total, pixels)));
        t->sb = carve_samples(t->smem, t->total, pixels);
        t->filt = f;
        t->filtered = true;
    }
    return mrh_tracer_reset_image(t);
}

int32_t mrh_tracer_stats(const mrh_tracer* t, mrh_render_stats* out, uint32_t* frames)
{
    if (!t) return MRH_ERR_INVALID;
    if (out) *out = t->stats;
    if (frames) *frames = t->frames;
    return MRH_OK;
}

int32_t mrh_tracer_destroy(mrh_tracer* t)
{
    if (!t) return MRH_ERR_INVALID;
    (void)hipStreamSynchronize(t->ctx->stream);
    (void)hipFree(t->arena);
    (void)hipFree(t->accum);
    if (t->smem) (void)hipFree(t->smem);
    guide_free(t->g);
    wf_free(t->w);
    delete t;
    return MRH_OK;
}

}  // extern "C"
