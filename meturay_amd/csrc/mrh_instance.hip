// mrh_instance.hip — two-level scenes (instancing) and ray casting against a
// whole scene.
//
// Build: every mesh gets its own LBVH over its local-space triangles (the
// same bit-exact builder as single-level scenes); the meshes' leaf records
// and binary / 4-wide traversal nodes are concatenated so one kernel argument
// reaches all of them, with each instance record holding its mesh's offsets
// and root codes.  A top-level LBVH is then built over the instances' world
// boxes: each box is given to the builder as the degenerate triangle
// (min, max, max), whose bounds and centroid are the box's own, and the leaf
// record's primitive index is the instance.  The reference builds one base
// BVH over its surface accelerators and resumes it per leaf from a separate
// launch (GPUAcceleratorBVHKC.cuh:513-652); here the top tree and the mesh
// trees are walked in one launch on one stack (traverse_two_level).
#include "mrh_path.hpp"

#include <cmath>
#include <cstring>

namespace mrh {

// Inverse of the affine local-to-world matrix (row-major 3x4) in double,
// rounded to float: adjugate / determinant, translation -A^-1 t.  Returns
// false for a singular or non-finite matrix.
bool affine_inverse(const float* m, float* out)
{
    const double a00 = m[0], a01 = m[1], a02 = m[2], t0 = m[3];
    const double a10 = m[4], a11 = m[5], a12 = m[6], t1 = m[7];
    const double a20 = m[8], a21 = m[9], a22 = m[10], t2 = m[11];
    const double c00 = a11 * a22 - a12 * a21;
    const double c01 = a12 * a20 - a10 * a22;
    const double c02 = a10 * a21 - a11 * a20;
    const double det = a00 * c00 + a01 * c01 + a02 * c02;
    if (!(det != 0.0) || !std::isfinite(det)) return false;
    const double id = 1.0 / det;
    double inv[3][3];
    inv[0][0] = c00 * id;
    inv[0][1] = (a02 * a21 - a01 * a22) * id;
    inv[0][2] = (a01 * a12 - a02 * a11) * id;
    inv[1][0] = c01 * id;
    inv[1][1] = (a00 * a22 - a02 * a20) * id;
    inv[1][2] = (a02 * a10 - a00 * a12) * id;
    inv[2][0] = c02 * id;
    inv[2][1] = (a01 * a20 - a00 * a21) * id;
    inv[2][2] = (a00 * a11 - a01 * a10) * id;
    for (int r = 0; r < 3; r++) {
        const double tr = -(inv[r][0] * t0 + inv[r][1] * t1 + inv[r][2] * t2);
        for (int c = 0; c < 3; c++) out[4 * r + c] = (float)inv[r][c];
        out[4 * r + 3] = (float)tr;
    }
    for (int i = 0; i < 12; i++)
        if (!std::isfinite(out[i])) return false;
    return true;
}

namespace {

// Host twin of xf_row (the same float operations in the same order)
float xf_row_host(const float* m, float x, float y, float z, float w)
{
    float r = 0.0f;
    r = r + m[0] * x;
    r = r + m[1] * y;
    r = r + m[2] * z;
    r = r + m[3] * w;
    return r;
}

// GPUTransformSingle::LocalToWorld(AABB3f) (GPUTransformSingle.cuh:180-197):
// the eight corners taken to the world, min / max over them.
void world_box(const float* l2w, const float* lo, const float* hi, float* out)
{
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < 8; i++) {
        const float x = (i & 1) ? hi[0] : lo[0];
        const float y = (i & 2) ? hi[1] : lo[1];
        const float z = (i & 4) ? hi[2] : lo[2];
        for (int a = 0; a < 3; a++) {
            const float v = xf_row_host(l2w + 4 * a, x, y, z, 1.0f);
            mn[a] = std::fmin(mn[a], v);
            mx[a] = std::fmax(mx[a], v);
        }
    }
    for (int a = 0; a < 3; a++) {
        out[a] = mn[a];
        out[3 + a] = mx[a];
    }
}

__global__ void k_leaf_prim_offset(float4* __restrict__ leaves, uint32_t n, uint32_t first)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float4 l0 = leaves[3 * (size_t)i];
        l0.w = __uint_as_float(__float_as_uint(l0.w) + first);
        leaves[3 * (size_t)i] = l0;
    }
}

bool key_valid(const mrh_scene_desc* d, uint32_t key)
{
    const uint32_t b = key >> 24, id = key & 0xFFFFFFu;
    const uint32_t lim = (b == MRH_BATCH_LAMBERT) ? d->n_lambert
                       : (b == MRH_BATCH_METAL) ? d->n_metal
                       : (b == MRH_BATCH_DIELECTRIC) ? d->n_dielectric
                       : (b == MRH_BATCH_EMISSIVE) ? d->n_emissive
                       : (b == MRH_BATCH_REFLECT) ? d->n_reflect : 0u;
    return b != MRH_BATCH_NULL && b <= MRH_BATCH_REFLECT && id < lim;
}

template <class T>
int32_t upload(mrh_scene* s, T** dst, const T* src, size_t count)
{
    *dst = nullptr;
    const size_t bytes = count * sizeof(T);
    void* p = nullptr;
    MRH_TRY(dmalloc(s->ctx, &p, bytes ? bytes : 16));
    s->allocs.push_back(p);
    s->bytes += bytes;
    if (bytes) MRH_HIP_CHECK(s->ctx, hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
    *dst = (T*)p;
    return MRH_OK;
}

int32_t device_buf(mrh_scene* s, void** dst, size_t bytes)
{
    MRH_TRY(dmalloc(s->ctx, dst, bytes ? bytes : 16));
    s->allocs.push_back(*dst);
    s->bytes += bytes;
    return MRH_OK;
}

int32_t build_meshes(mrh_context* ctx, const mrh_scene_desc* d, mrh_scene* s, std::vector<mrh_lbvh*>& mb)
{
    const mrh_instancing_desc* I = d->instancing;
    const uint32_t bits = d->morton_bits ? d->morton_bits : 30u;
    // the triangles and keys are staged only for the mesh builds (each
    // LBVH copies its triangles, and the leaf records carry corners and keys)
    float* dtris = nullptr;
    uint32_t* dkeys = nullptr;
    MRH_TRY(dmalloc(ctx, (void**)&dtris, (size_t)d->n * 36));
    int32_t rc = dmalloc(ctx, (void**)&dkeys, (size_t)d->n * 4);
    if (rc == MRH_OK && (hipMemcpy(dtris, d->tris, (size_t)d->n * 36, hipMemcpyHostToDevice) != hipSuccess ||
                         hipMemcpy(dkeys, d->prim_key, (size_t)d->n * 4, hipMemcpyHostToDevice) != hipSuccess))
        rc = fail(ctx, MRH_ERR_HIP, "mrh_scene_create: triangle upload failed");
    size_t nodes = 0;
    for (uint32_t m = 0; m < I->n_meshes && rc == MRH_OK; m++) {
        const uint32_t f = I->mesh_first[m], n = I->mesh_first[m + 1] - f;
        mrh_lbvh* b = nullptr;
        rc = lbvh_create(ctx, dtris + (size_t)f * 9, n, bits, dkeys + f, &b);
        if (rc != MRH_OK) break;
        mb[m] = b;
        s->mesh_depth = std::max(s->mesh_depth, b->max_depth);
        nodes += n - 1;
    }
    (void)hipStreamSynchronize(ctx->stream);
    if (dtris) (void)hipFree(dtris);
    if (dkeys) (void)hipFree(dkeys);
    MRH_TRY(rc);
    void *lv = nullptr, *n2 = nullptr, *n4 = nullptr;
    MRH_TRY(device_buf(s, &lv, (size_t)d->n * 48));
    MRH_TRY(device_buf(s, &n2, nodes * 64));
    MRH_TRY(device_buf(s, &n4, nodes * 128));
    s->leaves = (float4*)lv;
    s->mesh_nodes2 = (float4*)n2;
    s->mesh_nodes4 = (float4*)n4;
    size_t nb = 0;
    for (uint32_t m = 0; m < I->n_meshes; m++) {
        const uint32_t f = I->mesh_first[m], n = I->mesh_first[m + 1] - f;
        const mrh_lbvh* b = mb[m];
        MRH_HIP_CHECK(ctx, hipMemcpyAsync(s->leaves + 3 * (size_t)f, b->leaves, (size_t)n * 48,
                                          hipMemcpyDeviceToDevice, ctx->stream));
        if (n > 1) {
            MRH_HIP_CHECK(ctx, hipMemcpyAsync(s->mesh_nodes2 + 4 * nb, b->tnodes, (size_t)(n - 1) * 64,
                                              hipMemcpyDeviceToDevice, ctx->stream));
            MRH_HIP_CHECK(ctx, hipMemcpyAsync(s->mesh_nodes4 + 8 * nb, b->nodes4, (size_t)(n - 1) * 128,
                                              hipMemcpyDeviceToDevice, ctx->stream));
        }
        if (f)
            MRH_TRY(launch(ctx, "leaf_prim_offset", dim3(stride_grid(ctx, n, 256)), dim3(256), 0, k_leaf_prim_offset,
                           s->leaves + 3 * (size_t)f, n, f));
        nb += n - 1;
    }
    MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return MRH_OK;
}

}  // namespace

int32_t two_level_build(mrh_context* ctx, const mrh_scene_desc* d, mrh_scene* s)
{
    const mrh_instancing_desc* I = d->instancing;
    if (I->n_meshes == 0 || !I->mesh_first || I->n_instances == 0 || !I->xform || !I->inst_mesh)
        return fail(ctx, MRH_ERR_INVALID, "mrh_scene_create: instancing needs meshes and instances");
    if (I->mesh_first[0] != 0 || I->mesh_first[I->n_meshes] != d->n)
        return fail(ctx, MRH_ERR_INVALID, "mrh_scene_create: mesh ranges must cover the %u triangles", d->n);
    for (uint32_t m = 0; m < I->n_meshes; m++) {
        if (I->mesh_first[m + 1] <= I->mesh_first[m])
            return fail(ctx, MRH_ERR_INVALID, "mrh_scene_create: mesh %u is empty or out of order", m);
        // 4-wide node records are addressed by 32-bit byte offsets per tree
        if (I->mesh_first[m + 1] - I->mesh_first[m] > (1u << 25))
            return fail(ctx, MRH_ERR_UNSUPPORTED, "mrh_scene_create: mesh %u has more than 2^25 triangles", m);
    }
    if (I->n_instances > (1u << 25))
        return fail(ctx, MRH_ERR_UNSUPPORTED, "mrh_scene_create: more than 2^25 instances");
    const uint32_t ni = I->n_instances;
    std::vector<float> w2l((size_t)ni * 12);
    for (uint32_t i = 0; i < ni; i++) {
        if (I->inst_mesh[i] >= I->n_meshes)
            return fail(ctx, MRH_ERR_INVALID, "mrh_scene_create: instance %u names mesh %u of %u", i,
                        I->inst_mesh[i], I->n_meshes);
        if (I->inst_key && I->inst_key[i] && !key_valid(d, I->inst_key[i]))
            return fail(ctx, MRH_ERR_INVALID, "mrh_scene_create: instance %u has invalid key 0x%08x", i,
                        I->inst_key[i]);
        for (int k = 0; k < 12; k++)
            if (!std::isfinite(I->xform[(size_t)i * 12 + k]))
                return fail(ctx, MRH_ERR_INVALID, "mrh_scene_create: instance %u transform is not finite", i);
        if (!affine_inverse(I->xform + (size_t)i * 12, w2l.data() + (size_t)i * 12))
            return fail(ctx, MRH_ERR_INVALID, "mrh_scene_create: instance %u transform is singular", i);
    }
    std::vector<mrh_lbvh*> mb(I->n_meshes, nullptr);
    int32_t rc = build_meshes(ctx, d, s, mb);
    std::vector<uint32_t> lights, light_inst;
    std::vector<float> boxes((size_t)ni * 9);
    std::vector<float> rec((size_t)ni * INST_F4 * 4, 0.0f);
    if (rc == MRH_OK) {
        // instance records and world boxes
        size_t nb = 0;
        std::vector<size_t> node_base(I->n_meshes);
        for (uint32_t m = 0; m < I->n_meshes; m++) {
            node_base[m] = nb;
            nb += I->mesh_first[m + 1] - I->mesh_first[m] - 1;
        }
        for (uint32_t i = 0; i < ni; i++) {
            const uint32_t m = I->inst_mesh[i];
            const mrh_lbvh* b = mb[m];
            const float* l2w = I->xform + (size_t)i * 12;
            float* r = rec.data() + (size_t)i * INST_F4 * 4;
            std::memcpy(r, w2l.data() + (size_t)i * 12, 12 * sizeof(float));
            std::memcpy(r + 12, l2w, 12 * sizeof(float));
            const uint32_t u6[4] = {m, I->inst_key ? I->inst_key[i] : 0u, I->mesh_first[m], (uint32_t)node_base[m]};
            const uint32_t u7[4] = {(uint32_t)node_base[m], b->root4, root_code(b), b->n};
            std::memcpy(r + 24, u6, 16);
            std::memcpy(r + 28, u7, 16);
            float wb[6];
            world_box(l2w, b->lo, b->hi, wb);
            float* t = boxes.data() + (size_t)i * 9;
            t[0] = wb[0]; t[1] = wb[1]; t[2] = wb[2];
            t[3] = wb[3]; t[4] = wb[4]; t[5] = wb[5];
            t[6] = wb[3]; t[7] = wb[4]; t[8] = wb[5];
        }
        // lights: every emissive triangle of every instance, instance-major,
        // each mesh's triangles in primitive order
        for (uint32_t i = 0; i < ni && rc == MRH_OK; i++) {
            const uint32_t m = I->inst_mesh[i];
            const uint32_t f = I->mesh_first[m], n = I->mesh_first[m + 1] - f;
            const uint32_t ko = I->inst_key ? I->inst_key[i] : 0u;
            bool any = false;
            for (uint32_t p = 0; p < n && !any; p++) any = ((ko ? ko : d->prim_key[f + p]) >> 24) == MRH_BATCH_EMISSIVE;
            if (!any) continue;
            std::vector<uint32_t> sorted(n), pos(n);
            if (hipMemcpy(sorted.data(), mb[m]->sorted_idx, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess) {
                rc = fail(ctx, MRH_ERR_HIP, "mrh_scene_create: sorted index download failed");
                break;
            }
            for (uint32_t k = 0; k < n; k++) pos[sorted[k]] = k;
            for (uint32_t p = 0; p < n; p++) {
                if (((ko ? ko : d->prim_key[f + p]) >> 24) != MRH_BATCH_EMISSIVE) continue;
                lights.push_back(f + pos[p]);
                light_inst.push_back(i);
            }
        }
    }
    for (mrh_lbvh* b : mb)
        if (b) lbvh_free(b);
    if (rc != MRH_OK) return rc;
    if (lights.size() >= (1ull << 32))
        return fail(ctx, MRH_ERR_UNSUPPORTED, "mrh_scene_create: too many instanced lights");
    MRH_TRY(upload(s, &s->inst, (const float4*)rec.data(), (size_t)ni * INST_F4));
    float* dboxes = nullptr;
    MRH_TRY(dmalloc(ctx, (void**)&dboxes, boxes.size() * 4));
    rc = MRH_OK;
    if (hipMemcpy(dboxes, boxes.data(), boxes.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(ctx, MRH_ERR_HIP, "mrh_scene_create: instance box upload failed");
    if (rc == MRH_OK) rc = lbvh_create(ctx, dboxes, ni, d->morton_bits ? d->morton_bits : 30u, nullptr, &s->bvh);
    (void)hipFree(dboxes);
    if (rc != MRH_OK) return rc;
    s->bytes += (size_t)ni * (36 + 24 + 8 + 8 + 4 + 12 + 8 + 48 + 64 + 4);
    if (!lights.empty()) {
        MRH_TRY(upload(s, &s->lights, lights.data(), lights.size()));
        MRH_TRY(upload(s, &s->light_inst, light_inst.data(), light_inst.size()));
    }
    s->n_lights = (uint32_t)lights.size();
    s->n_inst = ni;
    s->n_tris = d->n;
    s->two_level = true;
    return MRH_OK;
}

namespace {

// Closest hit of a ray batch against a scene (single- or two-level): the
// primitive is the leaf record's scene index.
template <int W, int CAP, bool TWO>
__global__ __launch_bounds__(256) void k_scene_trace(const float4* __restrict__ nodes,
                                                     const float4* __restrict__ top_leaves, uint32_t root,
                                                     uint32_t n_top, const float4* __restrict__ inst,
                                                     const float4* __restrict__ mesh_nodes,
                                                     const float4* __restrict__ leaves, const float* __restrict__ rays,
                                                     uint32_t n, uint32_t pitch, float tmin,
                                                     uint32_t* __restrict__ out_inst, uint32_t* __restrict__ out_prim,
                                                     float* __restrict__ out_t, float* __restrict__ out_bary,
                                                     uint32_t* __restrict__ stats)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const V3 o = v3(rays[i], rays[pitch + i], rays[2 * (size_t)pitch + i]);
        const V3 d = v3(rays[3 * (size_t)pitch + i], rays[4 * (size_t)pitch + i], rays[5 * (size_t)pitch + i]);
        const float tmax = rays[6 * (size_t)pitch + i];
        PrivateStack<CAP> st;
        TraceResult r;
        uint32_t ii = 0;
        if constexpr (TWO) {
            auto world = [&](V3& wo, V3& wd) {
                wo = v3(rays[i], rays[pitch + i], rays[2 * (size_t)pitch + i]);
                wd = v3(rays[3 * (size_t)pitch + i], rays[4 * (size_t)pitch + i], rays[5 * (size_t)pitch + i]);
            };
            r = two_level_w<W, false, true>(st, nodes, top_leaves, root, n_top, inst, mesh_nodes, leaves, o, d, tmin,
                                            tmax, ii, world);
        } else {
            r = traverse_w<W, false, true>(st, nodes, leaves, root, n_top, o, d, tmin, tmax);
        }
        const bool h = r.leaf != NO_HIT;
        if (out_inst) out_inst[i] = h ? ii : NO_HIT;
        out_prim[i] = h ? __float_as_uint(leaves[3 * (size_t)r.leaf].w) : NO_HIT;
        out_t[i] = h ? r.t : 0.0f;
        out_bary[i] = h ? r.ba : 0.0f;
        out_bary[(size_t)n + i] = h ? r.bb : 0.0f;
        if (stats) {
            stats[2 * (size_t)i] = r.visits;
            stats[2 * (size_t)i + 1] = r.tests;
        }
    }
}

template <int W, int CAP>
int32_t scene_trace_w(mrh_context* ctx, const mrh_scene* s, const float* rays, uint32_t n, uint32_t pitch,
                      float tmin, uint32_t* inst, uint32_t* prim, float* t, float* bary, uint32_t* stats)
{
    const uint32_t grid = stride_grid(ctx, n, 256);
    const mrh_lbvh* b = s->bvh;
    if (s->two_level) {
        const float4* mn = (W == 4) ? s->mesh_nodes4 : s->mesh_nodes2;
        return launch(ctx, "scene_trace", dim3(grid), dim3(256), 0, k_scene_trace<W, CAP, true>, (const float4*)b->nodes4,
                      (const float4*)b->leaves, b->root4, s->n_inst, (const float4*)s->inst, mn,
                      (const float4*)s->leaves, rays, n, pitch, tmin, inst, prim, t, bary, stats);
    }
    const float4* nodes = (W == 4) ? (const float4*)b->nodes4 : (const float4*)b->tnodes;
    const uint32_t root = (W == 4) ? b->root4 : root_code(b);
    return launch(ctx, "scene_trace", dim3(grid), dim3(256), 0, k_scene_trace<W, CAP, false>, nodes,
                  (const float4*)nullptr, root, b->n, (const float4*)nullptr, (const float4*)nullptr,
                  (const float4*)b->leaves, rays, n, pitch, tmin, inst, prim, t, bary, stats);
}

}  // namespace
}  // namespace mrh

using namespace mrh;

extern "C" int32_t mrh_scene_trace_closest(mrh_context* ctx, const mrh_scene* s, const float* d_rays, uint32_t n,
                                           uint32_t pitch, float tmin, uint32_t* d_inst, uint32_t* d_prim,
                                           float* d_t, float* d_bary, uint32_t* d_stats)
{
    if (!ctx) return MRH_ERR_INVALID;
    if (!s || (n && (!d_rays || !d_prim || !d_t || !d_bary)) || pitch < n)
        return fail(ctx, MRH_ERR_INVALID, "mrh_scene_trace_closest: bad arguments");
    if (n == 0) return MRH_OK;
    const uint32_t w = (ctx->trace_width == 4) ? 4u : 2u;
    uint32_t need;
    if (s->two_level) {
        need = two_level_stack_need(s->bvh->max_depth, s->mesh_depth, w);
    } else {
        need = (w == 4) ? 3 * ((s->bvh->max_depth + 1) / 2) + 1 : s->bvh->max_depth + 1;
    }
    if (need > 128) return fail(ctx, MRH_ERR_UNSUPPORTED, "trees need %u stack entries (at most 128)", need);
    int32_t rc;
    if (w == 4) {
        rc = (need <= 64) ? scene_trace_w<4, 64>(ctx, s, d_rays, n, pitch, tmin, d_inst, d_prim, d_t, d_bary, d_stats)
                          : scene_trace_w<4, 128>(ctx, s, d_rays, n, pitch, tmin, d_inst, d_prim, d_t, d_bary, d_stats);
    } else {
        rc = (need <= 64) ? scene_trace_w<2, 64>(ctx, s, d_rays, n, pitch, tmin, d_inst, d_prim, d_t, d_bary, d_stats)
                          : scene_trace_w<2, 128>(ctx, s, d_rays, n, pitch, tmin, d_inst, d_prim, d_t, d_bary, d_stats);
    }
    MRH_TRY(rc);
    MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return MRH_OK;
}
