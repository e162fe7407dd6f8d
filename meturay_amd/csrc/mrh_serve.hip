// mrh_serve.hip — the serving engine: render requests of different scenes,
// resolutions, sample counts and guiding settings share one set of
// persistent, compacting wavefront kernels (continuous batching).
//
// The reference renders one scene per TracerSystem and drives one tracer's
// Render() loop per frame (TracerThread / TracerSystemCUDA,
// PathTracer.cu:144-291); it has no request loop.  This engine keeps the
// same per-path semantics (a request's paths are exactly the paths
// mrh_render / mrh_render_guided trace for it: same PCG32 keys, same
// arithmetic) and changes only how work reaches the GPU:
//   * a request occupies one of <= 64 slots; a device slot table holds its
//     scene (BVH, materials), camera, frame constants, image and guide;
//   * every step is one wavefront over the shared queue: camera rays of newly
//     injected samples are appended behind the surviving rays of all
//     requests, then trace (per-ray scene), partition by material batch,
//     shade (per-ray request constants), survivors compacted into the other
//     queue; per-slot live and traced counts come from LDS histograms;
//   * the host admits queued requests into free slots while the HBM budget
//     allows, injects new samples smallest-remaining-work first, and
//     finishes a request (or runs its guide's learn pass) when its last path
//     ends;
//   * the launch sequence of a step is captured once per queue parity in a
//     hipGraph and replayed (all per-step values live in device memory).
// The engine works on a private context (own stream and scratch), so the
// graphs' baked-in buffers cannot be moved by other work.
#include "mrh_path.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>

namespace mrh {
namespace {

constexpr uint32_t SERVE_MAX_SLOTS = 64;
// step results: [0] rays appended to the other queue, [RES_LIVE + s] of
// them slot s's, [RES_SEG + s] rays of slot s traced, [RES_PATHS] path rays
// among the appended (the rest are next-event rays)
constexpr uint32_t RES_LIVE = 1, RES_SEG = 1 + SERVE_MAX_SLOTS, RES_PATHS = 1 + 2 * SERVE_MAX_SLOTS,
                   RES_WORDS = 2 + 2 * SERVE_MAX_SLOTS;

// Device view of one admitted request
struct SlotDev {
    const float4* tnodes;   // binary traversal nodes
    const float4* nodes4;   // 4-wide traversal nodes (two-level: the top tree's)
    uint32_t root, root4, n_leaves, pad2;   // (two-level: n_leaves = instances)
    // two-level scenes: top-tree leaf records, instance records, mesh trees
    const float4* top_leaves;
    const float4* inst;
    const float4* mesh_nodes2;
    const float4* mesh_nodes4;
    uint32_t two, pad3, pad4, pad5;
    ShadeReq q;
    GuideArgs ga;
    Cam cam;
    uint32_t w, h, S, seed;
    uint32_t learn, guide, pad0, pad1;
};

struct Inject {
    uint32_t slot, first, count, offset;   // offset: first injected index of the chunk
};

struct StepHdr {
    uint32_t n_live;    // surviving rays at the head of the queue
    uint32_t n_total;   // n_live + injected
    uint32_t n_chunks, pad;
    Inject chunks[SERVE_MAX_SLOTS];
};

// Camera rays of the injected chunks, behind the live rays
__global__ __launch_bounds__(256) void k_serve_raygen(const StepHdr* __restrict__ hdr,
                                                      const SlotDev* __restrict__ slots, PathBuf out,
                                                      uint32_t* __restrict__ out_slot)
{
    const uint32_t n_live = hdr->n_live, n_inj = hdr->n_total - hdr->n_live, nc = hdr->n_chunks;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n_inj; t += gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = nc;   // last chunk with offset <= t
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (hdr->chunks[mid].offset <= t) lo = mid;
            else hi = mid;
        }
        const Inject c = hdr->chunks[lo];
        const SlotDev& sd = slots[c.slot];
        const uint32_t i = c.first + (t - c.offset);
        V3 o, d;
        uint64_t st;
        camera_ray(sd.cam, sd.w, sd.h, sd.S, sd.seed, sd.q.sample_base + i, i, o, d, st);
        const uint32_t k = n_live + t;
        out.ox[k] = o.x; out.oy[k] = o.y; out.oz[k] = o.z;
        out.dx[k] = d.x; out.dy[k] = d.y; out.dz[k] = d.z;
        out.tr[k] = 1.0f; out.tg[k] = 1.0f; out.tb[k] = 1.0f;
        out.sample[k] = i;
        out.dm[k] = 1u;
        out.rng[k] = st;
        out_slot[k] = c.slot;
    }
}

// Closest hit of every queued ray in its own request's scene; camera rays
// (depth 1) are clipped to the request's near/far planes.
template <int W, int CAP, int L>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(lds_stack_waves(L, 256, 8)))) void k_serve_trace(const StepHdr* __restrict__ hdr,
                                                     const SlotDev* __restrict__ slots, PathBuf in,
                                                     const uint32_t* __restrict__ in_slot, HitBuf hit,
                                                     uint2* __restrict__ overflow)
{
    __shared__ uint32_t s_node[L > 0 ? L * 256 : 1];
    __shared__ float s_t[L > 0 ? L * 256 : 1];
    const uint32_t n = hdr->n_total;
    const uint32_t lb = xcd_block();
    const size_t ostride = (size_t)gridDim.x * blockDim.x;
    const size_t oslot = (size_t)lb * blockDim.x + threadIdx.x;
    for (uint32_t i = lb * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const SlotDev& sd = slots[in_slot[i]];
        if (sd.two) continue;   // traced by k_serve_trace2
        const V3 o = v3(in.ox[i], in.oy[i], in.oz[i]);
        const V3 d = v3(in.dx[i], in.dy[i], in.dz[i]);
        const bool cam_ray = (in.dm[i] & (0xFFu | DM_NEE)) == 1u;
        const float tmin = cam_ray ? sd.cam.near_d : 0.0f;
        const float tmax = cam_ray ? sd.cam.far_d : FLT_MAX;
        const float4* leaves = sd.q.leaves;
        const float4* nodes = (W == 4) ? sd.nodes4 : sd.tnodes;
        const uint32_t root = (W == 4) ? sd.root4 : sd.root;
        TraceResult r;
        if constexpr (L > 0) {
            LdsStack<L, CAP, 256> st{s_node, s_t, threadIdx.x, overflow, ostride, oslot};
            r = traverse_w<W, false, false>(st, nodes, leaves, root, sd.n_leaves, o, d, tmin, tmax);
        } else {
            PrivateStack<CAP> st;
            r = traverse_w<W, false, false>(st, nodes, leaves, root, sd.n_leaves, o, d, tmin, tmax);
        }
        hit.leaf[i] = r.leaf;
        hit.t[i] = r.t;
        hit.ba[i] = r.ba;
        hit.bb[i] = r.bb;
        hit.key[i] = (r.leaf == NO_HIT) ? 0u : __float_as_uint(leaves[3 * (size_t)r.leaf + 1].w);
        hit.inst[i] = 0u;
    }
}

// The rays of two-level requests (launched only when one is admitted): the
// top tree, then the hit mesh's tree in local space (two_level_w).
template <int W, int L>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(lds_stack_waves(L, 256, 5)))) void k_serve_trace2(
    const StepHdr* __restrict__ hdr, const SlotDev* __restrict__ slots, PathBuf in,
    const uint32_t* __restrict__ in_slot, HitBuf hit, uint2* __restrict__ overflow)
{
    constexpr int CAP = 128;
    __shared__ uint32_t s_node[L > 0 ? L * 256 : 1];
    __shared__ float s_t[L > 0 ? L * 256 : 1];
    const uint32_t n = hdr->n_total;
    const uint32_t lb = xcd_block();
    const size_t ostride = (size_t)gridDim.x * blockDim.x;
    const size_t oslot = (size_t)lb * blockDim.x + threadIdx.x;
    for (uint32_t i = lb * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const SlotDev& sd = slots[in_slot[i]];
        if (!sd.two) continue;
        const V3 o = v3(in.ox[i], in.oy[i], in.oz[i]);
        const V3 d = v3(in.dx[i], in.dy[i], in.dz[i]);
        const bool cam_ray = (in.dm[i] & (0xFFu | DM_NEE)) == 1u;
        const float tmin = cam_ray ? sd.cam.near_d : 0.0f;
        const float tmax = cam_ray ? sd.cam.far_d : FLT_MAX;
        const float4* mesh_nodes = (W == 4) ? sd.mesh_nodes4 : sd.mesh_nodes2;
        auto world = [&](V3& wo, V3& wd) {
            wo = v3(in.ox[i], in.oy[i], in.oz[i]);
            wd = v3(in.dx[i], in.dy[i], in.dz[i]);
        };
        TraceResult r;
        uint32_t ii;
        if constexpr (L > 0) {
            LdsStack<L, CAP, 256> st{s_node, s_t, threadIdx.x, overflow, ostride, oslot};
            r = two_level_w<W, false, false>(st, sd.nodes4, sd.top_leaves, sd.root4, sd.n_leaves, sd.inst, mesh_nodes,
                                             sd.q.leaves, o, d, tmin, tmax, ii, world);
        } else {
            PrivateStack<CAP> st;
            r = two_level_w<W, false, false>(st, sd.nodes4, sd.top_leaves, sd.root4, sd.n_leaves, sd.inst, mesh_nodes,
                                             sd.q.leaves, o, d, tmin, tmax, ii, world);
        }
        hit.leaf[i] = r.leaf;
        hit.t[i] = r.t;
        hit.ba[i] = r.ba;
        hit.bb[i] = r.bb;
        hit.inst[i] = ii;
        uint32_t key = 0u;
        if (r.leaf != NO_HIT) {
            const uint32_t ko = __float_as_uint(sd.inst[INST_F4 * (size_t)ii + 6].y);
            key = ko ? ko : __float_as_uint(sd.q.leaves[3 * (size_t)r.leaf + 1].w);
        }
        hit.key[i] = key;
    }
}

// Shade every queued ray with its request's constants (shade_ray), append
// the survivors (and, with NEE, the next-event rays) to the other queue;
// counts go to res (RES_* above).  TILES: each block partitions its own tile
// of SERVE_TILE rays by HitKey batch in LDS (no global sort; see
// k_shade_tiles); otherwise rays are visited through the global sort's order.
constexpr uint32_t SERVE_TILE = 1024;

template <bool G, bool NEE, bool TILES>
__global__ __launch_bounds__(256) void k_serve_shade(const StepHdr* __restrict__ hdr,
                                                     const SlotDev* __restrict__ slots, PathBuf in,
                                                     const uint32_t* __restrict__ in_slot, PathBuf out,
                                                     uint32_t* __restrict__ out_slot, HitBuf hit,
                                                     const uint32_t* __restrict__ order, uint32_t* res)
{
    __shared__ uint32_t s_live[SERVE_MAX_SLOTS], s_seg[SERVE_MAX_SLOTS], s_paths;
    __shared__ uint16_t s_perm[TILES ? SERVE_TILE : 1];
    __shared__ uint32_t s_cnt[8], s_off[8], s_wave[4];
    if (threadIdx.x < SERVE_MAX_SLOTS) {
        s_live[threadIdx.x] = 0;
        s_seg[threadIdx.x] = 0;
    }
    if (threadIdx.x == 0) s_paths = 0;
    __syncthreads();
    const uint32_t n = hdr->n_total;
    constexpr uint32_t STEP = TILES ? SERVE_TILE : 256;
    for (uint32_t base = blockIdx.x * STEP; base < n; base += gridDim.x * STEP) {
        if constexpr (TILES) {
            if (threadIdx.x < 8) s_cnt[threadIdx.x] = 0;
            __syncthreads();
            uint32_t b[SERVE_TILE / 256];
#pragma unroll
            for (uint32_t k = 0; k < SERVE_TILE / 256; k++) {
                const uint32_t j = base + threadIdx.x + 256 * k;
                b[k] = (j < n) ? min(hit.key[j] >> 24, 7u) : 7u;
                atomicAdd(&s_cnt[b[k]], 1u);
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                uint32_t acc = 0;
                for (int c = 7; c >= 0; c--) {
                    s_off[c] = acc;
                    acc += s_cnt[c];
                }
            }
            __syncthreads();
#pragma unroll
            for (uint32_t k = 0; k < SERVE_TILE / 256; k++)
                s_perm[atomicAdd(&s_off[b[k]], 1u)] = (uint16_t)(threadIdx.x + 256 * k);
            __syncthreads();
        }
#pragma unroll 1
        for (uint32_t k = 0; k < STEP / 256; k++) {
            uint32_t rid = 0;
            bool valid;
            if constexpr (TILES) {
                rid = base + s_perm[threadIdx.x + 256 * k];
                valid = rid < n;
            } else {
                const uint32_t j = base + threadIdx.x;
                valid = j < n;
                if (valid) rid = order[j];
            }
            bool alive = false, has_nee = false;
            RayNext nx, ne;
            uint32_t sample = 0, slot = 0;
            if (valid) {
                slot = in_slot[rid];
                const SlotDev& sd = slots[slot];
                sample = in.sample[rid];
                alive = shade_ray<G, G, NEE>(in, hit, rid, sd.q, sd.ga, G && sd.learn != 0u, G && sd.guide != 0u,
                                             nx, ne, has_nee);
                atomicAdd(&s_seg[slot], 1u);
                const uint32_t made = (alive ? 1u : 0u) + (has_nee ? 1u : 0u);
                if (made) atomicAdd(&s_live[slot], made);
            }
            if (NEE) append_ray_block(has_nee, ne, sample, out, &res[0], false, s_wave, out_slot, slot);
            const uint64_t m = __ballot(alive);
            if ((threadIdx.x & 63) == 0 && m) atomicAdd(&s_paths, (uint32_t)__popcll(m));
            append_ray_block(alive, nx, sample, out, &res[0], NEE, s_wave, out_slot, slot);
        }
    }
    __syncthreads();
    if (threadIdx.x < SERVE_MAX_SLOTS) {
        if (s_live[threadIdx.x]) atomicAdd(&res[RES_LIVE + threadIdx.x], s_live[threadIdx.x]);
        if (s_seg[threadIdx.x]) atomicAdd(&res[RES_SEG + threadIdx.x], s_seg[threadIdx.x]);
    }
    if (threadIdx.x == 0 && s_paths) atomicAdd(&res[RES_PATHS], s_paths);
}

__global__ void k_serve_scale(float* __restrict__ p, size_t count, float s)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        p[i] *= s;
}

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace
}  // namespace mrh

using namespace mrh;

// Host state of one request
struct ServeReq {
    uint64_t ticket = 0;
    mrh_request r = {};
    uint32_t iterations = 1, it = 0, S = 1;
    uint64_t total = 0;        // paths per iteration
    uint64_t injected = 0;     // of the current iteration
    uint64_t live = 0;         // rays in flight (paths and their next-event rays)
    uint64_t segments = 0;
    uint64_t charge = 0;       // budget bytes
    int slot = -1;
    bool learning = false, have_guide = false;
    bool in_learn = false;     // its learn pass is queued or running on the learner
    bool learned_ok = true;    // result of the last learn pass
    int32_t learn_rc = MRH_OK;
    int32_t cancel_status = MRH_OK;   // MRH_ERR_CANCELLED / MRH_ERR_TIMEOUT once cancelled
    uint32_t kept = 0;
    void* recmem = nullptr;
    size_t recmem_bytes = 0;
    RecBuf rec = {};
    uint32_t *c_codes = nullptr, *c_bins = nullptr;
    float* c_rad = nullptr;
    mrh_svo* svo = nullptr;
    float cube[4] = {0, 0, 0, 1};
    double t_submit = 0, t_admit = 0;
};

struct mrh_server {
    mrh_context* user = nullptr;   // errors are mirrored here
    mrh_context* ctx = nullptr;    // private context: stream, scratch, graphs
    mrh_server_params p = {};
    uint32_t cap = 0;              // paths in flight (wave_paths)
    uint32_t buf_cap = 0;          // rays a queue holds: a path and its next-event ray
    uint32_t partition = 1;        // material partition of a step (the shade_partition knob at creation)
    PathBuf bufs[2] = {};
    uint32_t* slot_of[2] = {};
    HitBuf hit = {};
    uint32_t* order = nullptr;
    uint32_t* d_res = nullptr;
    StepHdr* d_hdr = nullptr;
    SlotDev* d_slots = nullptr;
    std::vector<void*> allocs;
    // pinned host mirrors
    StepHdr* h_hdr = nullptr;
    SlotDev* h_slots = nullptr;
    uint32_t* h_res = nullptr;
    uint32_t parity = 0;           // queue holding the live rays
    uint32_t n_live = 0;           // rays in that queue
    uint32_t n_paths = 0;          // path rays among them
    uint64_t used = 0;             // budget bytes held: queues, resident scenes, admitted requests
    uint64_t fixed_bytes = 0;      // the queues and step buffers (charged from creation on)
    std::map<const mrh_scene*, uint32_t> resident;   // scenes of admitted requests (charged once each)
    uint64_t next_ticket = 1;
    uint64_t steps = 0, segments = 0;
    std::deque<ServeReq*> queue;
    ServeReq* slots[SERVE_MAX_SLOTS] = {};
    std::vector<mrh_request_done> done;
    // graphs per shade variant (bit 0: learn-or-guide, bit 1: NEE) and queue
    // parity, with what they were captured against
    // (bit 2: some admitted request has a two-level scene)
    hipGraphExec_t exec[8][2] = {};
    void* graph_scratch[8][2][SLOT_COUNT] = {};
    uint32_t graph_cap[8][2] = {}, graph_w[8][2] = {}, graph_lds[8][2] = {}, graph_bpc[8][2] = {};
    uint32_t max_tree_depth = 0;    // deepest LBVH among submitted requests
    uint32_t max_tree_leaves = 0;   // largest LBVH among submitted requests
    // learner: guide learn passes run on their own context and host thread,
    // overlapping the wavefront steps of other requests
    mrh_context* lctx = nullptr;
    std::thread learner;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<ServeReq*> learn_q;   // waiting for the learner
    std::vector<ServeReq*> learned;  // finished, waiting for the engine thread
    bool stop = false;
    std::string learn_err;
    // reusable octrees (on lctx) and record buffers
    std::vector<mrh_svo*> free_svos;
    std::vector<std::pair<size_t, void*>> free_recs;
    size_t pooled_bytes = 0;
};

namespace {

int32_t mirror_error(mrh_server* s, int32_t rc)
{
    if (rc != MRH_OK && s->user) s->user->err = s->ctx->err;
    return rc;
}

// device bytes a request holds while admitted: its image, learn records and
// an estimate of its guide (nodes <= min(records, 16 * 4^depth), 1300 B each)
uint64_t request_charge(const ServeReq* q)
{
    uint64_t b = (uint64_t)q->r.params.width * q->r.params.height * 12;
    if (q->iterations > 1) {
        const uint64_t rec = q->total * q->r.params.max_depth;
        uint64_t nodes = std::min<uint64_t>(rec, 16ull << (2 * q->r.guide.svo_depth));
        b += rec * 36 + nodes * 1300;
    }
    return b;
}

// keeps up to this many bytes of released record buffers for reuse
constexpr size_t REC_POOL_BYTES = 64ull << 30;

void release(mrh_server* s, ServeReq* q)
{
    if (q->svo) s->free_svos.push_back(q->svo);
    q->svo = nullptr;
    if (q->recmem) {
        (void)hipStreamSynchronize(s->ctx->stream);
        if (s->pooled_bytes + q->recmem_bytes <= REC_POOL_BYTES) {
            s->free_recs.emplace_back(q->recmem_bytes, q->recmem);
            s->pooled_bytes += q->recmem_bytes;
        } else {
            (void)hipFree(q->recmem);
        }
    }
    q->recmem = nullptr;
    if (q->slot >= 0) {
        s->slots[q->slot] = nullptr;
        s->used -= q->charge;
        auto it = s->resident.find(q->r.scene);
        if (it != s->resident.end() && --it->second == 0) {
            s->used -= q->r.scene->bytes;
            s->resident.erase(it);
        }
    }
    delete q;
}

// budget bytes admitting q adds: its own charge and its scene's bytes when no
// admitted request holds that scene yet
uint64_t admit_cost(const mrh_server* s, const ServeReq* q)
{
    return q->charge + (s->resident.count(q->r.scene) ? 0u : q->r.scene->bytes);
}

SlotDev& host_slot(mrh_server* s, const ServeReq* q) { return s->h_slots[q->slot]; }

int32_t upload_slot(mrh_server* s, const ServeReq* q)
{
    MRH_HIP_CHECK(s->ctx, hipMemcpyAsync(s->d_slots + q->slot, s->h_slots + q->slot, sizeof(SlotDev),
                                         hipMemcpyHostToDevice, s->ctx->stream));
    // the pinned entry may be rewritten before the next step's wait
    MRH_HIP_CHECK(s->ctx, hipStreamSynchronize(s->ctx->stream));
    return MRH_OK;
}

// slot descriptor for the request's current iteration
void fill_slot(mrh_server* s, ServeReq* q)
{
    SlotDev& d = host_slot(s, q);
    const mrh_scene* sc = q->r.scene;
    const mrh_lbvh* b = sc->bvh;
    memset(&d, 0, sizeof(d));
    d.tnodes = (const float4*)b->tnodes;
    d.nodes4 = (const float4*)b->nodes4;
    d.root = root_code(b);
    d.root4 = b->root4;
    d.n_leaves = b->n;
    d.q.mats = Materials{sc->lambert, sc->metal, sc->dielectric, sc->emissive, sc->reflect};
    d.q.leaves = (const float4*)b->leaves;
    d.q.tbn = sc->tbn;
    if (sc->two_level) {
        d.two = 1u;
        d.n_leaves = sc->n_inst;
        d.top_leaves = (const float4*)b->leaves;
        d.inst = (const float4*)sc->inst;
        d.mesh_nodes2 = (const float4*)sc->mesh_nodes2;
        d.mesh_nodes4 = (const float4*)sc->mesh_nodes4;
        d.q.leaves = (const float4*)sc->leaves;
        d.q.inst = (const float4*)sc->inst;
        d.q.light_inst = sc->light_inst;
    }
    d.q.image = q->r.d_image;
    d.q.spp = q->S * q->S;
    d.q.sample_base = (uint32_t)(q->it * q->total);
    d.q.rr_start = q->r.params.rr_start;
    d.q.max_depth = q->r.params.max_depth;
    d.q.lights = sc->lights;
    d.q.n_lights = sc->n_lights;
    d.q.nee = q->r.params.nee;
    d.q.mis = q->r.params.nee && q->r.params.direct_light_mis;
    d.cam = make_camera(q->r.camera, q->r.params.width, q->r.params.height);
    d.w = q->r.params.width;
    d.h = q->r.params.height;
    d.S = q->S;
    d.seed = q->r.params.seed;
    d.learn = q->learning ? 1u : 0u;
    d.guide = q->have_guide ? 1u : 0u;
    d.ga.rec = q->rec;
    d.ga.lox = q->cube[0];
    d.ga.loy = q->cube[1];
    d.ga.loz = q->cube[2];
    d.ga.inv_cell = (float)(1u << q->r.guide.svo_depth) / q->cube[3];
    d.ga.depth = q->r.guide.svo_depth;
    d.ga.ratio = q->r.guide.guide_ratio;
    d.ga.min_samples = q->r.guide.min_samples;
    if (q->have_guide) d.ga.svo = svo_view(q->svo);
}

int32_t admit(mrh_server* s, ServeReq* q, int slot)
{
    mrh_context* ctx = s->ctx;
    s->used += admit_cost(s, q);
    s->resident[q->r.scene]++;
    q->slot = slot;
    s->slots[slot] = q;
    q->t_admit = now_s();
    const size_t img = (size_t)q->r.params.width * q->r.params.height * 3;
    MRH_TRY(fill_u32(ctx, (uint32_t*)q->r.d_image, 0u, img));
    if (q->iterations > 1) {
        const mrh_lbvh* b = q->r.scene->bvh;
        float span = 0.0f;
        for (int a = 0; a < 3; a++) span = fmaxf(span, b->hi[a] - b->lo[a]);
        if (!(span > 0.0f)) span = 1.0f;
        const float cells = (float)(1u << q->r.guide.svo_depth);
        const float pad = span / (cells - 2.0f > 1.0f ? cells - 2.0f : 1.0f);
        q->cube[0] = b->lo[0] - pad;
        q->cube[1] = b->lo[1] - pad;
        q->cube[2] = b->lo[2] - pad;
        q->cube[3] = span + 2.0f * pad;
        const uint64_t n_rec = q->total * q->r.params.max_depth;
        const size_t need = (size_t)n_rec * 36;
        // smallest pooled buffer that fits, else a new one
        int best = -1;
        for (size_t k = 0; k < s->free_recs.size(); k++)
            if (s->free_recs[k].first >= need && (best < 0 || s->free_recs[k].first < s->free_recs[best].first))
                best = (int)k;
        if (best >= 0) {
            q->recmem_bytes = s->free_recs[best].first;
            q->recmem = s->free_recs[best].second;
            s->pooled_bytes -= q->recmem_bytes;
            s->free_recs.erase(s->free_recs.begin() + best);
        } else {
            MRH_TRY(dmalloc(ctx, &q->recmem, need));
            q->recmem_bytes = need;
        }
        char* m = (char*)q->recmem;
        q->rec.code = (uint32_t*)m; m += n_rec * 4;
        q->rec.bin = (uint32_t*)m; m += n_rec * 4;
        q->rec.tr = (float*)m; m += n_rec * 4;
        q->rec.tg = (float*)m; m += n_rec * 4;
        q->rec.tb = (float*)m; m += n_rec * 4;
        q->rec.rad = (float*)m; m += n_rec * 4;
        q->c_codes = (uint32_t*)m; m += n_rec * 4;
        q->c_bins = (uint32_t*)m; m += n_rec * 4;
        q->c_rad = (float*)m;
        q->rec.max_depth = q->r.params.max_depth;
        q->rec.stride = (uint32_t)q->total;
        MRH_TRY(fill_u32(ctx, q->rec.code, SVO_NO_NODE, n_rec));
        if (!s->free_svos.empty()) {
            q->svo = s->free_svos.back();
            s->free_svos.pop_back();
        } else {
            MRH_TRY(mrh_svo_create(s->lctx, &q->svo));
        }
        q->learning = true;
    }
    fill_slot(s, q);
    return upload_slot(s, q);
}

// The learn pass of one request on the learner's context (learner thread):
// compact the records, rebuild, splat and filter the octree, reset the
// records for the next learning frame; returns when the device is done.
int32_t learn_pass(mrh_server* s, ServeReq* q)
{
    mrh_context* lc = s->lctx;
    const uint32_t n_rec = (uint32_t)(q->total * q->r.params.max_depth);
    uint32_t kept = 0;
    MRH_TRY(svo_compact_records(lc, q->svo, q->rec.code, q->rec.bin, q->rec.rad, n_rec, q->c_codes, q->c_bins,
                                q->c_rad, &kept));
    if (kept) {
        MRH_TRY(svo_build(lc, q->svo, q->c_codes, kept, q->cube, q->r.guide.svo_depth, true));
        MRH_TRY(svo_splat(lc, q->svo, q->c_codes, q->c_bins, q->c_rad, kept));
        MRH_TRY(svo_filter(lc, q->svo));
    }
    MRH_TRY(fill_u32(lc, q->rec.code, SVO_NO_NODE, n_rec));
    MRH_HIP_CHECK(lc, hipStreamSynchronize(lc->stream));
    q->kept = kept;
    return MRH_OK;
}

void learner_main(mrh_server* s)
{
    (void)hipSetDevice(s->lctx->device);
    for (;;) {
        ServeReq* q = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->stop || !s->learn_q.empty(); });
            if (s->stop && s->learn_q.empty()) return;
            q = s->learn_q.front();
            s->learn_q.pop_front();
        }
        const int32_t rc = learn_pass(s, q);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            q->learned_ok = rc == MRH_OK;
            q->learn_rc = rc;
            if (rc != MRH_OK) s->learn_err = s->lctx->err;
            s->learned.push_back(q);
        }
        s->cv.notify_all();
    }
}

int32_t finish_request(mrh_server* s, ServeReq* q);

// The request's next frame: guide from the last learn pass, fresh sample range
int32_t next_iteration(mrh_server* s, ServeReq* q)
{
    q->it++;
    q->injected = 0;
    q->learning = q->it + 1 < q->iterations;
    fill_slot(s, q);
    return upload_slot(s, q);
}

// All paths of the request's current iteration have ended
int32_t iteration_done(mrh_server* s, ServeReq* q)
{
    if (q->learning) {
        // hand the records to the learner; injection waits until it is done
        q->in_learn = true;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->learn_q.push_back(q);
        }
        s->cv.notify_one();
        return MRH_OK;
    }
    if (q->it + 1 < q->iterations) return next_iteration(s, q);
    return finish_request(s, q);
}

// Ends a request that failed: reported with its error, slot and budget
// released.  None of its paths may be in flight.
void end_request(mrh_server* s, ServeReq* q, int32_t status)
{
    mrh_request_done d = {};
    d.ticket = q->ticket;
    d.user_tag = q->r.user_tag;
    d.status = status;
    d.segments = q->segments;
    d.submit_s = q->t_submit;
    d.admit_s = q->t_admit;
    d.done_s = now_s();
    s->done.push_back(d);
    release(s, q);
}

int32_t finish_request(mrh_server* s, ServeReq* q)
{
    mrh_context* ctx = s->ctx;
    q->it++;
    // finished: the image is the mean over all samples of all iterations
    const size_t img = (size_t)q->r.params.width * q->r.params.height * 3;
    MRH_TRY(launch(ctx, "serve_finalize", dim3(stride_grid(ctx, img, 256)), dim3(256), 0, k_serve_scale,
                   q->r.d_image, img, 1.0f / ((float)(q->S * q->S) * (float)q->iterations)));
    MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    mrh_request_done d = {};
    d.ticket = q->ticket;
    d.user_tag = q->r.user_tag;
    d.status = MRH_OK;
    d.segments = q->segments;
    d.submit_s = q->t_submit;
    d.admit_s = q->t_admit;
    d.done_s = now_s();
    s->done.push_back(d);
    release(s, q);
    return MRH_OK;
}

// Tree width and stack capacity that fit every submitted scene (see
// select_tree: 4-wide needs 3 * ceil(depth / 2) + 1 stack entries)
void tree_shape(const mrh_server* s, uint32_t& w, uint32_t& cap)
{
    const uint32_t d = s->max_tree_depth;
    w = (s->ctx->trace_width == 4 && d <= 84 && s->max_tree_leaves <= (1u << 25)) ? 4u : 2u;
    cap = (w == 4) ? (d <= 42 ? 64u : 128u) : (d <= 64 ? 64u : 128u);
}

// The launches of one step over queue parity `par` (captured or eager);
// variant bit 0: some admitted request learns or samples a guide, bit 1:
// some admitted request sends next-event rays
int32_t step_launches(mrh_server* s, uint32_t par, uint32_t v)
{
    mrh_context* ctx = s->ctx;
    const PathBuf& in = s->bufs[par];
    const PathBuf& out = s->bufs[par ^ 1];
    const uint32_t* in_slot = s->slot_of[par];
    uint32_t* out_slot = s->slot_of[par ^ 1];
    MRH_TRY(fill_u32(ctx, s->d_res, 0u, RES_WORDS));
    MRH_TRY(launch(ctx, "serve_raygen", dim3(stride_grid(ctx, s->cap, 256)), dim3(256), 0, k_serve_raygen,
                   (const StepHdr*)s->d_hdr, (const SlotDev*)s->d_slots, in, s->slot_of[par]));
    const uint32_t tgrid = trace_grid(ctx, s->buf_cap);
    const uint32_t lds = ctx->trace_lds_entries;
    uint32_t w, cap;
    tree_shape(s, w, cap);
    uint2* ovf = nullptr;
    MRH_TRY(trace_overflow(ctx, tgrid, (v & 4u) ? 128u : cap, lds, &ovf));
#define MRH_SERVE_TRACE(W, C, LV)                                                                             \
    MRH_TRY(launch(ctx, "serve_trace", dim3(tgrid), dim3(256), 0, k_serve_trace<W, C, LV>,                     \
                   (const StepHdr*)s->d_hdr, (const SlotDev*)s->d_slots, in, in_slot, s->hit, ovf))
#define MRH_SERVE_TRACE_L(W, C)                                                  \
    switch (lds) {                                                               \
    case 0: MRH_SERVE_TRACE(W, C, 0); break;                                     \
    case 8: MRH_SERVE_TRACE(W, C, 8); break;                                     \
    case 12: MRH_SERVE_TRACE(W, C, 12); break;                                   \
    case 16: MRH_SERVE_TRACE(W, C, 16); break;                                   \
    case 24: MRH_SERVE_TRACE(W, C, 24); break;                                   \
    default: return fail(ctx, MRH_ERR_INVALID, "trace_lds_entries must be 0, 8, 12, 16 or 24"); \
    }
    if (w == 4 && cap == 64) {
        MRH_SERVE_TRACE_L(4, 64)
    } else if (w == 4) {
        MRH_SERVE_TRACE_L(4, 128)
    } else if (cap == 64) {
        MRH_SERVE_TRACE_L(2, 64)
    } else {
        MRH_SERVE_TRACE_L(2, 128)
    }
#undef MRH_SERVE_TRACE_L
#undef MRH_SERVE_TRACE
    if (v & 4u) {
#define MRH_SERVE_TRACE2(W, LV)                                                                                 \
    MRH_TRY(launch(ctx, "serve_trace", dim3(tgrid), dim3(256), 0, k_serve_trace2<W, LV>, (const StepHdr*)s->d_hdr, \
                   (const SlotDev*)s->d_slots, in, in_slot, s->hit, ovf))
#define MRH_SERVE_TRACE2_L(W)                                                    \
    switch (lds) {                                                               \
    case 0: MRH_SERVE_TRACE2(W, 0); break;                                       \
    case 8: MRH_SERVE_TRACE2(W, 8); break;                                       \
    case 12: MRH_SERVE_TRACE2(W, 12); break;                                     \
    case 16: MRH_SERVE_TRACE2(W, 16); break;                                     \
    case 24: MRH_SERVE_TRACE2(W, 24); break;                                     \
    default: return fail(ctx, MRH_ERR_INVALID, "trace_lds_entries must be 0, 8, 12, 16 or 24"); \
    }
        if (w == 4) {
            MRH_SERVE_TRACE2_L(4)
        } else {
            MRH_SERVE_TRACE2_L(2)
        }
#undef MRH_SERVE_TRACE2_L
#undef MRH_SERVE_TRACE2
    }
    const bool tiles = s->partition == 1;
    if (!tiles) {
        SortArgs sa;
        sa.keys_in = s->hit.key;
        sa.vals_out = s->order;
        sa.n_max = s->buf_cap;
        sa.n_dev = &s->d_hdr->n_total;
        sa.key_bytes = 4;
        sa.begin_bit = 24;
        sa.end_bit = 32;
        sa.descending = true;
        MRH_TRY(radix_sort(ctx, sa));
    }
    const uint32_t grid = tiles ? stride_grid(ctx, (uint64_t)(s->buf_cap + SERVE_TILE - 1) / SERVE_TILE * 256, 256)
                                : stride_grid(ctx, s->buf_cap, 256);
    auto shade = [&](auto kernel) {
        return launch(ctx, "serve_shade", dim3(grid), dim3(256), 0, kernel, (const StepHdr*)s->d_hdr,
                      (const SlotDev*)s->d_slots, in, in_slot, out, out_slot, s->hit, (const uint32_t*)s->order,
                      s->d_res);
    };
    if (tiles) {
        switch (v) {
        case 0: return shade(k_serve_shade<false, false, true>);
        case 1: return shade(k_serve_shade<true, false, true>);
        case 2: return shade(k_serve_shade<false, true, true>);
        default: return shade(k_serve_shade<true, true, true>);
        }
    }
    switch (v) {
    case 0: return shade(k_serve_shade<false, false, false>);
    case 1: return shade(k_serve_shade<true, false, false>);
    case 2: return shade(k_serve_shade<false, true, false>);
    default: return shade(k_serve_shade<true, true, false>);
    }
}

bool graph_valid(mrh_server* s, uint32_t v, uint32_t par)
{
    if (!s->exec[v][par]) return false;
    for (int k = 0; k < SLOT_COUNT; k++)
        if (s->graph_scratch[v][par][k] != s->ctx->slot_ptr[k]) return false;
    uint32_t w, cap;
    tree_shape(s, w, cap);
    return s->graph_cap[v][par] == cap && s->graph_w[v][par] == w &&
           s->graph_lds[v][par] == s->ctx->trace_lds_entries && s->graph_bpc[v][par] == s->ctx->trace_blocks_per_cu;
}

int32_t run_step(mrh_server* s, uint32_t par, uint32_t v)
{
    mrh_context* ctx = s->ctx;
    if (!s->p.use_graphs) return step_launches(s, par, v);
    if (!graph_valid(s, v, par)) {
        if (s->exec[v][par]) {
            (void)hipGraphExecDestroy(s->exec[v][par]);
            s->exec[v][par] = nullptr;
        }
        // one eager step first so every scratch buffer exists at its final size
        MRH_TRY(step_launches(s, par, v));
        MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        const bool prof = ctx->profile;
        ctx->profile = false;
        hipGraph_t graph = nullptr;
        MRH_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int32_t rc = step_launches(s, par, v);
        hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
        ctx->profile = prof;
        if (rc != MRH_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return fail(ctx, MRH_ERR_HIP, "step capture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&s->exec[v][par], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(ctx, MRH_ERR_HIP, "step graph instantiate: %s", hipGetErrorString(e));
        for (int k = 0; k < SLOT_COUNT; k++) s->graph_scratch[v][par][k] = ctx->slot_ptr[k];
        tree_shape(s, s->graph_w[v][par], s->graph_cap[v][par]);
        s->graph_lds[v][par] = ctx->trace_lds_entries;
        s->graph_bpc[v][par] = ctx->trace_blocks_per_cu;
        // the eager step above already did this step's work
        return MRH_OK;
    }
    MRH_HIP_CHECK(ctx, hipGraphLaunch(s->exec[v][par], ctx->stream));
    return MRH_OK;
}

// Queued: reported at once.  Admitted: takes no new samples, and its paths in
// flight end at their next shading (max depth 0 in its slot); it is reported
// with `status` when the last has ended or its learn pass has returned.
int32_t cancel_request(mrh_server* s, ServeReq* q, int32_t status)
{
    if (q->slot < 0) {
        s->queue.erase(std::find(s->queue.begin(), s->queue.end(), q));
        end_request(s, q, status);
        return MRH_OK;
    }
    if (q->cancel_status != MRH_OK) return MRH_OK;
    q->cancel_status = status;
    q->injected = q->total;
    q->learning = false;
    fill_slot(s, q);
    host_slot(s, q).q.max_depth = 0;
    return upload_slot(s, q);
}

int32_t server_step(mrh_server* s, uint32_t* active, uint32_t* queued)
{
    mrh_context* ctx = s->ctx;
    if (s->p.timeout_ms) {
        const double t = now_s(), lim = s->p.timeout_ms * 1e-3;
        std::vector<ServeReq*> late;
        for (ServeReq* q : s->queue)
            if (t - q->t_submit > lim) late.push_back(q);
        for (uint32_t k = 0; k < SERVE_MAX_SLOTS; k++)
            if (s->slots[k] && s->slots[k]->cancel_status == MRH_OK && t - s->slots[k]->t_submit > lim)
                late.push_back(s->slots[k]);
        for (ServeReq* q : late) MRH_TRY(cancel_request(s, q, MRH_ERR_TIMEOUT));
    }
    // admission, first come first served, while slots and budget allow
    while (!s->queue.empty()) {
        ServeReq* q = s->queue.front();
        int slot = -1;
        uint32_t n_active = 0;
        for (uint32_t k = 0; k < s->p.max_slots; k++) {
            if (s->slots[k]) n_active++;
            else if (slot < 0) slot = (int)k;
        }
        if (slot < 0 || (n_active > 0 && s->used + admit_cost(s, q) > s->p.hbm_budget)) break;
        s->queue.pop_front();
        // a request that cannot be admitted (its record buffer does not
        // fit, say) ends alone with that error; the others go on
        const int32_t rc = admit(s, q, slot);
        if (rc != MRH_OK) end_request(s, q, rc);
    }
    // requests whose learn pass finished: next frame, guided
    std::vector<ServeReq*> ready;
    std::string lerr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        ready.swap(s->learned);
        lerr = s->learn_err;
    }
    for (ServeReq* q : ready) {
        q->in_learn = false;
        if (q->cancel_status != MRH_OK) {
            end_request(s, q, q->cancel_status);
            continue;
        }
        if (!q->learned_ok) {
            // none of its paths is in flight (its frame had ended): it ends
            // alone with the learn pass's error, the others go on
            fail(ctx, MRH_ERR_HIP, "guide learn pass failed: %s", lerr.c_str());
            end_request(s, q, q->learn_rc);
            continue;
        }
        if (q->kept) q->have_guide = true;
        MRH_TRY(next_iteration(s, q));
    }
    // injection, smallest remaining work first
    std::vector<ServeReq*> act;
    for (uint32_t k = 0; k < s->p.max_slots; k++)
        if (s->slots[k] && s->slots[k]->injected < s->slots[k]->total) act.push_back(s->slots[k]);
    if (s->p.policy == MRH_SERVE_ARRIVAL) {
        std::sort(act.begin(), act.end(), [](const ServeReq* a, const ServeReq* b) { return a->ticket < b->ticket; });
    } else {
        std::sort(act.begin(), act.end(), [](const ServeReq* a, const ServeReq* b) {
            const uint64_t ra = (a->total - a->injected) + (uint64_t)(a->iterations - a->it - 1) * a->total;
            const uint64_t rb = (b->total - b->injected) + (uint64_t)(b->iterations - b->it - 1) * b->total;
            return ra != rb ? ra < rb : a->ticket < b->ticket;
        });
    }
    StepHdr& h = *s->h_hdr;
    h.n_live = s->n_live;
    h.n_chunks = 0;
    // at most cap paths in flight: with a next-event ray each the other
    // queue still holds every ray a step appends
    uint64_t room = s->cap - s->n_paths, inj = 0;
    for (ServeReq* q : act) {
        if (!room) break;
        const uint64_t k = std::min<uint64_t>(room, q->total - q->injected);
        Inject& c = h.chunks[h.n_chunks++];
        c.slot = (uint32_t)q->slot;
        c.first = (uint32_t)q->injected;
        c.count = (uint32_t)k;
        c.offset = (uint32_t)inj;
        q->injected += k;
        q->live += k;
        inj += k;
        room -= k;
    }
    h.n_total = s->n_live + (uint32_t)inj;
    if (h.n_total > 0) {
        MRH_HIP_CHECK(ctx, hipMemcpyAsync(s->d_hdr, s->h_hdr, sizeof(StepHdr), hipMemcpyHostToDevice, ctx->stream));
        uint32_t v = 0;
        for (uint32_t k = 0; k < s->p.max_slots; k++) {
            const ServeReq* q = s->slots[k];
            if (!q) continue;
            if (q->learning || q->have_guide) v |= 1u;
            if (q->r.params.nee) v |= 2u;
            if (q->r.scene->two_level) v |= 4u;
        }
        MRH_TRY(run_step(s, s->parity, v));
        MRH_HIP_CHECK(ctx, hipMemcpyAsync(s->h_res, s->d_res, RES_WORDS * 4, hipMemcpyDeviceToHost, ctx->stream));
        MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        s->steps++;
        s->segments += h.n_total;
        s->n_live = s->h_res[0];
        s->n_paths = s->h_res[RES_PATHS];
        s->parity ^= 1u;
        for (uint32_t k = 0; k < s->p.max_slots; k++) {
            ServeReq* q = s->slots[k];
            if (!q) continue;
            q->live = s->h_res[RES_LIVE + k];
            q->segments += s->h_res[RES_SEG + k];
        }
    }
    // iterations / requests whose paths have all ended
    for (uint32_t k = 0; k < s->p.max_slots; k++) {
        ServeReq* q = s->slots[k];
        if (!q || q->in_learn || q->injected != q->total || q->live != 0) continue;
        if (q->cancel_status != MRH_OK) end_request(s, q, q->cancel_status);
        else MRH_TRY(iteration_done(s, q));
    }
    uint32_t a = 0;
    for (uint32_t k = 0; k < s->p.max_slots; k++) a += s->slots[k] ? 1u : 0u;
    // nothing to trace while learn passes run: wait for one instead of spinning
    if (h.n_total == 0 && a > 0) {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv.wait_for(lk, std::chrono::milliseconds(2), [&] { return !s->learned.empty(); });
    }
    if (active) *active = a;
    if (queued) *queued = (uint32_t)s->queue.size();
    return MRH_OK;
}

void server_free(mrh_server* s)
{
    if (s->learner.joinable()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv.notify_all();
        s->learner.join();
    }
    if (s->ctx) (void)hipStreamSynchronize(s->ctx->stream);
    for (uint32_t k = 0; k < SERVE_MAX_SLOTS; k++)
        if (s->slots[k]) release(s, s->slots[k]);
    for (ServeReq* q : s->queue) delete q;
    for (int v = 0; v < 8; v++)
        for (int par = 0; par < 2; par++)
            if (s->exec[v][par]) (void)hipGraphExecDestroy(s->exec[v][par]);
    for (mrh_svo* v : s->free_svos) svo_free(v);
    for (auto& r : s->free_recs) (void)hipFree(r.second);
    for (void* p : s->allocs) (void)hipFree(p);
    if (s->h_hdr) (void)hipHostFree(s->h_hdr);
    if (s->h_slots) (void)hipHostFree(s->h_slots);
    if (s->h_res) (void)hipHostFree(s->h_res);
    if (s->ctx) mrh_context_destroy(s->ctx);
    if (s->lctx) mrh_context_destroy(s->lctx);
    delete s;
}

int32_t server_create(mrh_context* user, const mrh_server_params* p, mrh_server** out)
{
    if (p->max_slots == 0 || p->max_slots > SERVE_MAX_SLOTS)
        return fail(user, MRH_ERR_INVALID, "server: max_slots must be in [1, %u]", SERVE_MAX_SLOTS);
    if (p->policy > MRH_SERVE_ARRIVAL) return fail(user, MRH_ERR_INVALID, "server: unknown policy %u", p->policy);
    if (p->wave_paths < 1024 || p->wave_paths >= (1u << 30))
        return fail(user, MRH_ERR_INVALID, "server: wave_paths must be in [1024, 2^30)");
    // a queue holds 2 * wave_paths rays; the global partition sorts it whole
    if (user->shade_partition == 0 && p->wave_paths >= (1u << 29))
        return fail(user, MRH_ERR_INVALID,
                    "server: with the global partition (shade_partition = 0) wave_paths must be below 2^29");
    mrh_server* s = new mrh_server();
    s->user = user;
    s->p = *p;
    s->cap = p->wave_paths;
    s->buf_cap = 2 * p->wave_paths;
    int32_t rc = mrh_context_create(user->device, &s->ctx);
    if (rc != MRH_OK) {
        delete s;
        return fail(user, rc, "server: context creation failed");
    }
    rc = mrh_context_create(user->device, &s->lctx);
    if (rc != MRH_OK) {
        mrh_context_destroy(s->ctx);
        delete s;
        return fail(user, rc, "server: learner context creation failed");
    }
    mrh_context* ctx = s->ctx;
    ctx->trace_lds_entries = user->trace_lds_entries;
    ctx->trace_blocks_per_cu = user->trace_blocks_per_cu;
    ctx->trace_width = user->trace_width;
    ctx->shade_partition = user->shade_partition;
    s->partition = user->shade_partition;
    auto take = [&](size_t bytes) -> void* {
        void* q = nullptr;
        if (rc == MRH_OK) rc = dmalloc(ctx, &q, bytes);
        if (q) s->allocs.push_back(q);
        s->fixed_bytes += bytes;
        return q;
    };
    const size_t n = s->buf_cap;
    for (int b = 0; b < 2; b++) {
        PathBuf& q = s->bufs[b];
        q.ox = (float*)take(n * 4); q.oy = (float*)take(n * 4); q.oz = (float*)take(n * 4);
        q.dx = (float*)take(n * 4); q.dy = (float*)take(n * 4); q.dz = (float*)take(n * 4);
        q.tr = (float*)take(n * 4); q.tg = (float*)take(n * 4); q.tb = (float*)take(n * 4);
        q.sample = (uint32_t*)take(n * 4);
        q.dm = (uint32_t*)take(n * 4);
        q.rng = (uint64_t*)take(n * 8);
        q.pdf = (float*)take(n * 4);
        s->slot_of[b] = (uint32_t*)take(n * 4);
    }
    s->hit.leaf = (uint32_t*)take(n * 4);
    s->hit.t = (float*)take(n * 4);
    s->hit.ba = (float*)take(n * 4);
    s->hit.bb = (float*)take(n * 4);
    s->hit.key = (uint32_t*)take(n * 4);
    s->hit.inst = (uint32_t*)take(n * 4);
    s->order = (uint32_t*)take(n * 4);
    s->d_res = (uint32_t*)take(RES_WORDS * 4);
    s->d_hdr = (StepHdr*)take(sizeof(StepHdr));
    s->d_slots = (SlotDev*)take(sizeof(SlotDev) * SERVE_MAX_SLOTS);
    if (rc == MRH_OK && hipHostMalloc((void**)&s->h_hdr, sizeof(StepHdr), hipHostMallocDefault) != hipSuccess)
        rc = fail(ctx, MRH_ERR_OOM, "server: pinned header");
    if (rc == MRH_OK &&
        hipHostMalloc((void**)&s->h_slots, sizeof(SlotDev) * SERVE_MAX_SLOTS, hipHostMallocDefault) != hipSuccess)
        rc = fail(ctx, MRH_ERR_OOM, "server: pinned slot table");
    if (rc == MRH_OK && hipHostMalloc((void**)&s->h_res, RES_WORDS * 4, hipHostMallocDefault) != hipSuccess)
        rc = fail(ctx, MRH_ERR_OOM, "server: pinned results");
    if (rc == MRH_OK) {
        memset(s->h_hdr, 0, sizeof(StepHdr));
        memset(s->h_slots, 0, sizeof(SlotDev) * SERVE_MAX_SLOTS);
        rc = fill_u32(ctx, (uint32_t*)s->d_slots, 0u, sizeof(SlotDev) * SERVE_MAX_SLOTS / 4);
    }
    if (rc == MRH_OK) rc = hipStreamSynchronize(ctx->stream) == hipSuccess ? MRH_OK : MRH_ERR_HIP;
    s->used = s->fixed_bytes;
    if (rc == MRH_OK && s->fixed_bytes > p->hbm_budget)
        rc = fail(ctx, MRH_ERR_BUDGET, "server: its queues need %llu bytes, budget is %llu",
                  (unsigned long long)s->fixed_bytes, (unsigned long long)p->hbm_budget);
    if (rc == MRH_OK) s->learner = std::thread(learner_main, s);
    if (rc != MRH_OK) {
        if (s->user != nullptr) mirror_error(s, rc);
        server_free(s);
        return rc;
    }
    *out = s;
    return MRH_OK;
}

int32_t server_submit(mrh_server* s, const mrh_request* r, uint64_t* ticket)
{
    mrh_context* ctx = s->ctx;
    if (!r->scene || !r->d_image) return fail(ctx, MRH_ERR_INVALID, "submit: scene and d_image are required");
    const mrh_render_params& p = r->params;
    if (p.width == 0 || p.height == 0 || p.samples == 0)
        return fail(ctx, MRH_ERR_INVALID, "submit: empty frame");
    if (p.max_depth == 0 || p.max_depth > 62) return fail(ctx, MRH_ERR_INVALID, "submit: max_depth must be in [1, 62]");
    if (p.nee > 1 || p.direct_light_mis > 1)
        return fail(ctx, MRH_ERR_INVALID, "submit: nee and direct_light_mis are 0 or 1");
    if (r->scene->two_level) {
        const uint32_t need = std::max(two_level_stack_need(r->scene->bvh->max_depth, r->scene->mesh_depth, 4),
                                       two_level_stack_need(r->scene->bvh->max_depth, r->scene->mesh_depth, 2));
        if (need > 128)
            return fail(ctx, MRH_ERR_UNSUPPORTED, "submit: two-level trees need %u stack entries (at most 128)", need);
    } else if (r->scene->bvh->max_depth > 128)
        return fail(ctx, MRH_ERR_UNSUPPORTED, "submit: tree depth %u exceeds the traversal stack",
                    r->scene->bvh->max_depth);
    ServeReq* q = new ServeReq();
    q->r = *r;
    q->S = p.samples;
    q->iterations = r->guide.iterations > 1 ? r->guide.iterations : 1u;
    q->total = (uint64_t)p.width * p.height * p.samples * p.samples;
    if (q->iterations > 1) {
        if (q->iterations > 64 || r->guide.svo_depth == 0 || r->guide.svo_depth > 10 ||
            !(r->guide.guide_ratio >= 0.0f && r->guide.guide_ratio <= 1.0f)) {
            delete q;
            return fail(ctx, MRH_ERR_INVALID, "submit: bad guide parameters");
        }
    }
    if (q->total * q->iterations >= (1ull << 32) || q->total * p.max_depth >= (1ull << 31)) {
        delete q;
        return fail(ctx, MRH_ERR_INVALID, "submit: request too large");
    }
    q->charge = request_charge(q);
    const uint64_t need = s->fixed_bytes + r->scene->bytes + q->charge;
    if (need > s->p.hbm_budget) {
        delete q;
        return fail(ctx, MRH_ERR_BUDGET,
                    "submit: request needs %llu bytes with its scene and the queues, budget is %llu",
                    (unsigned long long)need, (unsigned long long)s->p.hbm_budget);
    }
    if (!r->scene->two_level) {
        s->max_tree_depth = std::max(s->max_tree_depth, r->scene->bvh->max_depth);
        s->max_tree_leaves = std::max(s->max_tree_leaves, r->scene->bvh->n);
    }
    q->ticket = s->next_ticket++;
    q->t_submit = now_s();
    s->queue.push_back(q);
    if (ticket) *ticket = q->ticket;
    return MRH_OK;
}

}  // namespace

extern "C" {

int32_t mrh_server_create(mrh_context* ctx, const mrh_server_params* params, mrh_server** out)
{
    if (!ctx || !params || !out) return MRH_ERR_INVALID;
    *out = nullptr;
    return server_create(ctx, params, out);
}

int32_t mrh_server_context(mrh_server* server, mrh_context** out)
{
    if (!server || !out) return MRH_ERR_INVALID;
    *out = server->ctx;
    return MRH_OK;
}

int32_t mrh_server_submit(mrh_server* server, const mrh_request* request, uint64_t* ticket)
{
    if (!server || !request) return MRH_ERR_INVALID;
    return mirror_error(server, server_submit(server, request, ticket));
}

int32_t mrh_server_step(mrh_server* server, uint32_t* active, uint32_t* queued)
{
    if (!server) return MRH_ERR_INVALID;
    return mirror_error(server, server_step(server, active, queued));
}

int32_t mrh_server_poll(mrh_server* server, mrh_request_done* out, uint32_t max, uint32_t* n)
{
    if (!server || !n || (max && !out)) return MRH_ERR_INVALID;
    const uint32_t k = (uint32_t)std::min<size_t>(max, server->done.size());
    for (uint32_t i = 0; i < k; i++) out[i] = server->done[i];
    server->done.erase(server->done.begin(), server->done.begin() + k);
    *n = k;
    return MRH_OK;
}

int32_t mrh_server_cancel(mrh_server* server, uint64_t ticket)
{
    if (!server) return MRH_ERR_INVALID;
    for (ServeReq* q : server->queue)
        if (q->ticket == ticket) return mirror_error(server, cancel_request(server, q, MRH_ERR_CANCELLED));
    for (uint32_t k = 0; k < SERVE_MAX_SLOTS; k++)
        if (server->slots[k] && server->slots[k]->ticket == ticket)
            return mirror_error(server, cancel_request(server, server->slots[k], MRH_ERR_CANCELLED));
    return mirror_error(server, fail(server->ctx, MRH_ERR_INVALID, "cancel: no queued or active request %llu",
                                     (unsigned long long)ticket));
}

int32_t mrh_server_stats(const mrh_server* server, mrh_server_info* out)
{
    if (!server || !out) return MRH_ERR_INVALID;
    memset(out, 0, sizeof(*out));
    out->steps = server->steps;
    out->segments = server->segments;
    out->budget_used = server->used;
    out->fixed_bytes = server->fixed_bytes;
    for (const auto& r : server->resident) out->scene_bytes += r.first->bytes;
    out->queued = (uint32_t)server->queue.size();
    for (uint32_t k = 0; k < SERVE_MAX_SLOTS; k++) out->active += server->slots[k] ? 1u : 0u;
    out->live_paths = server->n_live;
    return MRH_OK;
}

int32_t mrh_server_destroy(mrh_server* server)
{
    if (!server) return MRH_ERR_INVALID;
    server_free(server);
    return MRH_OK;
}

}  // extern "C"
