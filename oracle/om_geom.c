/*
 * om_geom.c — oracle: Morton codes, sorts, LBVH hierarchy + refit, ray casts.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Plain serial C; the product's
 * HIP kernels are compared against these functions bit-for-bit.
 */
#include "oracle.h"
#include "om.h"

#include <stdlib.h>
#include <string.h>
#include <float.h>

/* ------------------------------------------------------------------ */
/* Morton codes                                                        */
/* ------------------------------------------------------------------ */

/* RayLib/MortonCode.h:60-70 (32-bit Expand3D): 10 bits spread to every 3rd bit */
uint32_t om_morton_expand10(uint32_t x)
{
    x &= 0x000003ffu;
    x = (x ^ (x << 16)) & 0xff0000ffu;
    x = (x ^ (x << 8)) & 0x0300f00fu;
    x = (x ^ (x << 4)) & 0x030c30c3u;
    x = (x ^ (x << 2)) & 0x09249249u;
    return x;
}

/* RayLib/MortonCode.h:45-57 (64-bit Expand3D): 21 bits spread to every 3rd bit */
uint64_t om_morton_expand21(uint32_t v)
{
    uint64_t x = v;
    x &= 0x1fffffu;
    x = (x | x << 32) & 0x001f00000000ffffull;
    x = (x | x << 16) & 0x001f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

/* RayLib/MortonCode.h:152-158: x -> bit 0, y -> bit 1, z -> bit 2 */
uint32_t om_morton_compose30(uint32_t x, uint32_t y, uint32_t z)
{
    return om_morton_expand10(x) | (om_morton_expand10(y) << 1) | (om_morton_expand10(z) << 2);
}

uint64_t om_morton_compose63(uint32_t x, uint32_t y, uint32_t z)
{
    return om_morton_expand21(x) | (om_morton_expand21(y) << 1) | (om_morton_expand21(z) << 2);
}

/* Triangle AABB: component-wise min/max of the three corners
 * (RayLib/Triangle.h:56-70, Triangle::BoundingBox). */
void om_tri_aabbs(const float* tris, uint32_t n, float* aabbs)
{
    for (uint32_t i = 0; i < n; i++) {
        const float* t = tris + (size_t)i * 9;
        float* b = aabbs + (size_t)i * 6;
        for (int a = 0; a < 3; a++) {
            b[a] = fminf(fminf(t[a], t[3 + a]), t[6 + a]);
            b[3 + a] = fmaxf(fmaxf(t[a], t[3 + a]), t[6 + a]);
        }
    }
}

/* Extent of the leaf cloud = union of all leaf AABBs
 * (LinearBVH.hpp:459-467, ReduceArrayGPU with ReduceAABB3f). */
void om_extent(const float* aabbs, uint32_t n, float ext[6])
{
    ext[0] = ext[1] = ext[2] = INFINITY;
    ext[3] = ext[4] = ext[5] = -INFINITY;
    for (uint32_t i = 0; i < n; i++) {
        const float* b = aabbs + (size_t)i * 6;
        for (int a = 0; a < 3; a++) {
            ext[a] = fminf(ext[a], b[a]);
            ext[3 + a] = fmaxf(ext[3 + a], b[3 + a]);
        }
    }
}

/* LinearBVH.hpp:400-414 (FindMortonDelta): largest span * 2^-bits computed in
 * double, then rounded to float.  The reference fixes bits = 21; 30-bit codes
 * use 10 bits per dimension. */
float om_morton_delta(const float ext[6], uint32_t bits_per_dim)
{
    double res = 1.0 / (double)(1u << bits_per_dim);
    double best = 0.0;
    for (int a = 0; a < 3; a++) {
        float span = ext[3 + a] - ext[a];
        double d = (double)span * res;
        if (d > best) best = d;
    }
    return (float)best;
}

/* LinearBVH.hpp:143-165 (KCGenMortonCodes body) with the two recorded
 * decisions: IEEE division (the reference GPU build uses fast-math division)
 * and a clamp to 2^bits - 1 (the reference wraps 2^bits to 0 through the
 * Expand3D mask).  Centroid = min + span * 0.5 (RayLib/AABB.hpp:79-82).
 * A zero delta (all leaves at one point) quantises everything to 0. */
void om_quantise(const float aabb[6], const float ext[6], float delta,
                 uint32_t bits_per_dim, uint32_t q[3])
{
    uint32_t qmax = (1u << bits_per_dim) - 1u;
    for (int a = 0; a < 3; a++) {
        float span = aabb[3 + a] - aabb[a];
        float c = aabb[a] + span * 0.5f;
        float rel = c - ext[a];
        uint32_t v = 0;
        if (delta > 0.0f) {
            float f = floorf(rel / delta);
            v = (f >= (float)qmax) ? qmax : (uint32_t)f;
        }
        q[a] = v;
    }
}

void om_morton_codes(const float* tris, uint32_t n, uint32_t bits, void* codes)
{
    if (n == 0) return;
    float* aabbs = (float*)malloc((size_t)n * 6 * sizeof(float));
    float ext[6];
    om_tri_aabbs(tris, n, aabbs);
    om_extent(aabbs, n, ext);
    uint32_t bpd = (bits == 30) ? 10u : 21u;
    float delta = om_morton_delta(ext, bpd);
    for (uint32_t i = 0; i < n; i++) {
        uint32_t q[3];
        om_quantise(aabbs + (size_t)i * 6, ext, delta, bpd, q);
        if (bits == 30) ((uint32_t*)codes)[i] = om_morton_compose30(q[0], q[1], q[2]);
        else ((uint64_t*)codes)[i] = om_morton_compose63(q[0], q[1], q[2]);
    }
    free(aabbs);
}

/* ------------------------------------------------------------------ */
/* Sorting                                                             */
/* ------------------------------------------------------------------ */
typedef struct { uint64_t key; uint32_t idx; } om_pair;

static int om_pair_cmp(const void* a, const void* b)
{
    const om_pair* x = (const om_pair*)a;
    const om_pair* y = (const om_pair*)b;
    if (x->key != y->key) return (x->key < y->key) ? -1 : 1;
    return (x->idx < y->idx) ? -1 : (x->idx > y->idx);
}

/* Stable ascending sort by key with the input index as value: ordering by
 * (key, index) is exactly what a stable sort of an iota value array yields
 * (LinearBVH.hpp:489-492, IotaGPU + RadixSortValKeyGPU). */
void om_sort_pairs_u64(const uint64_t* keys, uint32_t n, uint64_t* sorted_keys, uint32_t* sorted_idx)
{
    om_pair* p = (om_pair*)malloc((size_t)(n ? n : 1) * sizeof(om_pair));
    for (uint32_t i = 0; i < n; i++) { p[i].key = keys[i]; p[i].idx = i; }
    qsort(p, n, sizeof(om_pair), om_pair_cmp);
    for (uint32_t i = 0; i < n; i++) {
        if (sorted_keys) sorted_keys[i] = p[i].key;
        sorted_idx[i] = p[i].idx;
    }
    free(p);
}

void om_sort_pairs_u32(const uint32_t* keys, uint32_t n, uint32_t* sorted_keys, uint32_t* sorted_idx)
{
    om_pair* p = (om_pair*)malloc((size_t)(n ? n : 1) * sizeof(om_pair));
    for (uint32_t i = 0; i < n; i++) { p[i].key = keys[i]; p[i].idx = i; }
    qsort(p, n, sizeof(om_pair), om_pair_cmp);
    for (uint32_t i = 0; i < n; i++) {
        if (sorted_keys) sorted_keys[i] = (uint32_t)p[i].key;
        sorted_idx[i] = p[i].idx;
    }
    free(p);
}

/* RayMemory.cu:187-228 sorts (key, rayId) stably in DESCENDING key order
 * over the used bit range; RayMemory.cu:230-312 then cuts the sorted list
 * where the key changes.  A stable counting sort over the digit is the
 * textbook equivalent. */
void om_partition_desc(const uint32_t* keys, uint32_t n, uint32_t begin_bit, uint32_t end_bit,
                       uint32_t* order, uint32_t* offsets)
{
    uint32_t nbits = end_bit - begin_bit;
    uint32_t nd = 1u << nbits;
    uint32_t mask = (nbits >= 32) ? 0xFFFFFFFFu : (nd - 1u);
    uint32_t* count = (uint32_t*)calloc(nd, sizeof(uint32_t));
    for (uint32_t i = 0; i < n; i++) count[(keys[i] >> begin_bit) & mask]++;
    /* descending digit order: digit nd-1 first */
    uint32_t* start = (uint32_t*)malloc(nd * sizeof(uint32_t));
    uint32_t run = 0;
    for (uint32_t j = 0; j < nd; j++) {
        uint32_t d = nd - 1u - j;
        start[d] = run;
        if (offsets) offsets[j] = run;
        run += count[d];
    }
    if (offsets) offsets[nd] = n;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t d = (keys[i] >> begin_bit) & mask;
        order[start[d]++] = i;
    }
    free(count);
    free(start);
}

/* ------------------------------------------------------------------ */
/* LBVH                                                                */
/* ------------------------------------------------------------------ */
static inline uint32_t om_clz32(uint32_t x) { return x ? (uint32_t)__builtin_clz(x) : 32u; }
static inline uint32_t om_clz64(uint64_t x) { return x ? (uint32_t)__builtin_clzll(x) : 64u; }

/* LinearBVH.hpp:30-39 (Delta) extended with the (code, index) tie-break for
 * duplicate codes (Karras 2012 sec. 4): the augmented key is code || index,
 * so equal codes compare by the common prefix of the 32-bit indices. */
static inline uint32_t om_delta(const void* codes, uint32_t bits, uint32_t i)
{
    if (bits == 30) {
        const uint32_t* c = (const uint32_t*)codes;
        uint32_t x = c[i] ^ c[i + 1];
        return x ? om_clz32(x) : 32u + om_clz32(i ^ (i + 1u));
    } else {
        const uint64_t* c = (const uint64_t*)codes;
        uint64_t x = c[i] ^ c[i + 1];
        return x ? om_clz64(x) : 64u + om_clz32(i ^ (i + 1u));
    }
}

int om_lbvh_build(const float* tris, uint32_t n, uint32_t bits,
                  void* sorted_codes_out, uint32_t* sorted_idx,
                  uint32_t* left, uint32_t* right, uint32_t* parent,
                  float* node_aabb, uint32_t* root)
{
    if (n == 0 || (bits != 30 && bits != 63)) return -1;
    size_t csz = (bits == 30) ? 4 : 8;
    void* codes = malloc((size_t)n * csz);
    void* scodes = malloc((size_t)n * csz);
    om_morton_codes(tris, n, bits, codes);
    if (bits == 30) om_sort_pairs_u32((const uint32_t*)codes, n, (uint32_t*)scodes, sorted_idx);
    else om_sort_pairs_u64((const uint64_t*)codes, n, (uint64_t*)scodes, sorted_idx);
    if (sorted_codes_out) memcpy(sorted_codes_out, scodes, (size_t)n * csz);

    float* tri_aabb = (float*)malloc((size_t)n * 6 * sizeof(float));
    om_tri_aabbs(tris, n, tri_aabb);
    /* leaves: node i = sorted leaf i (LinearBVH.hpp:183-200) */
    for (uint32_t i = 0; i < n; i++)
        memcpy(node_aabb + (size_t)i * 6, tri_aabb + (size_t)sorted_idx[i] * 6, 6 * sizeof(float));

    if (n == 1) {
        parent[0] = OM_INVALID;
        *root = 0;
        free(codes); free(scodes); free(tri_aabb);
        return 0;
    }

    uint32_t ni = n - 1;
    uint32_t* rl = (uint32_t*)malloc(ni * sizeof(uint32_t));
    uint32_t* rr = (uint32_t*)malloc(ni * sizeof(uint32_t));
    uint8_t* arrived = (uint8_t*)calloc(ni, 1);
    for (uint32_t i = 0; i < ni; i++) { left[i] = right[i] = OM_INVALID; }
    *root = OM_INVALID;

    /* Bottom-up emission (Apetrei 2014), restating ChooseParentAndUpdate
     * (LinearBVH.hpp:43-100) + KCConstructLBVHInter root handling
     * (LinearBVH.hpp:219-229) + BottomUpAABBInter (LinearBVH.hpp:310-365).
     * Each leaf walks up; the first child to reach a parent stops, the second
     * continues with the parent's full range.  Serial order does not change
     * the result: the parent choice of a range is a pure function of the
     * codes, and min/max unions are exact. */
    for (uint32_t leaf = 0; leaf < n; leaf++) {
        uint32_t l = leaf, r = leaf, node = leaf;
        for (;;) {
            if (l == 0 && r == n - 1) { parent[node] = OM_INVALID; *root = node; break; }
            uint32_t ld = (l == 0) ? 0u : om_delta(scodes, bits, l - 1);
            uint32_t rd = (r == n - 1) ? 0u : om_delta(scodes, bits, r);
            uint32_t p;
            if (rd > ld) { p = r; left[p] = node; rl[p] = l; }
            else { p = l - 1; right[p] = node; rr[p] = r; }
            parent[node] = n + p;
            if (!arrived[p]) { arrived[p] = 1; break; }
            /* second arrival: both children known */
            l = rl[p]; r = rr[p]; node = n + p;
            const float* a = node_aabb + (size_t)left[p] * 6;
            const float* b = node_aabb + (size_t)right[p] * 6;
            float* o = node_aabb + (size_t)node * 6;
            for (int k = 0; k < 3; k++) {
                o[k] = fminf(a[k], b[k]);
                o[3 + k] = fmaxf(a[3 + k], b[3 + k]);
            }
        }
    }
    free(rl); free(rr); free(arrived);
    free(codes); free(scodes); free(tri_aabb);
    return (*root == OM_INVALID) ? -2 : 0;
}

/* ------------------------------------------------------------------ */
/* Ray casting                                                         */
/* ------------------------------------------------------------------ */

/* RayLib/Ray.hpp:143-185 (Moller-Trumbore, cullFace = false).
 * bary = (1-u-v, u, v). */
int om_intersect_triangle(const float o[3], const float d[3], const float tri[9],
                          float* t_out, float bary[3])
{
    om_v3 p0 = v3(tri[0], tri[1], tri[2]);
    om_v3 p1 = v3(tri[3], tri[4], tri[5]);
    om_v3 p2 = v3(tri[6], tri[7], tri[8]);
    om_v3 dir = v3(d[0], d[1], d[2]);
    om_v3 pos = v3(o[0], o[1], o[2]);
    om_v3 e0 = v3_sub(p1, p0);
    om_v3 e1 = v3_sub(p2, p0);
    om_v3 p = v3_cross(dir, e1);
    float det = v3_dot(e0, p);
    if (fabsf(det) < 1.0e-7f) return 0;
    float inv_det = 1.0f / det;
    om_v3 tv = v3_sub(pos, p0);
    float u = v3_dot(tv, p) * inv_det;
    if (u < 0.0f || u > 1.0f) return 0;
    om_v3 q = v3_cross(tv, e0);
    float v = v3_dot(dir, q) * inv_det;
    if (v < 0.0f || (v + u) > 1.0f) return 0;
    float t = v3_dot(e1, q) * inv_det;
    if (t <= 1.0e-7f) return 0;
    float c = 1.0f - u - v;
    *t_out = t;
    bary[0] = c; bary[1] = u; bary[2] = v;
    return 1;
}

/* RayLib/Ray.hpp:248-270 (slab test).  min/max follow IEEE minNum/maxNum
 * (fminf/fmaxf), which also fixes the result for 0*inf = NaN slabs. */
int om_intersect_aabb(const float o[3], const float inv_d[3], const float box[6],
                      float tmin, float tmax, float* t_enter)
{
    float te = tmin, tx = tmax;
    for (int a = 0; a < 3; a++) {
        float t0 = (box[a] - o[a]) * inv_d[a];
        float t1 = (box[3 + a] - o[a]) * inv_d[a];
        te = fmaxf(te, fminf(t0, t1));
        tx = fminf(tx, fmaxf(t0, t1));
    }
    *t_enter = te;
    return tx >= te;
}

typedef struct {
    uint32_t node;
    float    t;
} om_stack_entry;

/* The box test of the 4-wide walks (the traversal kernels' node4_test): the
 * slab test with each axis's planes ordered by the direction's sign.  The
 * near plane (the min plane where the reciprocal direction is non-negative,
 * else the max plane) gives the axis's entry distance and the far plane its
 * exit, each as fma(plane, inv, -o * inv); the box is entered when
 * max(tmin, entries) <= min(tmax, exits).  Direction components below
 * 2^-64 in magnitude become +-2^-64 (sign kept), so every distance is
 * finite.  Entry and exit are taken max / min left to right: x, y, z, then
 * tmin / tmax (exact, in any order, for non-NaN values). */
typedef struct {
    float inv[3], noi[3];
    int   near_max[3];   /* 1: the near plane of the axis is the max plane */
} om_slab;

static float slab_rcp(float d)
{
    const float tiny = 0x1p-64f;
    return 1.0f / (fabsf(d) < tiny ? copysignf(tiny, d) : d);
}

static void om_slab_setup(const float o[3], const float d[3], om_slab* s)
{
    for (int a = 0; a < 3; a++) {
        s->inv[a] = slab_rcp(d[a]);
        s->noi[a] = -(o[a] * s->inv[a]);
        s->near_max[a] = signbit(s->inv[a]) ? 1 : 0;
    }
}

static int om_slab_box(const om_slab* s, const float box[6], float tmin, float tmax, float* t_enter)
{
    float te = 0.0f, tx = 0.0f;
    for (int a = 0; a < 3; a++) {
        const float np = s->near_max[a] ? box[3 + a] : box[a];
        const float fp = s->near_max[a] ? box[a] : box[3 + a];
        const float tn = fmaf(np, s->inv[a], s->noi[a]);
        const float tf = fmaf(fp, s->inv[a], s->noi[a]);
        te = (a == 0) ? tn : fmaxf(te, tn);
        tx = (a == 0) ? tf : fminf(tx, tf);
    }
    te = fmaxf(te, tmin);
    tx = fminf(tx, tmax);
    *t_enter = te;
    return tx >= te;
}

int om_slab_box_test(const float o[3], const float d[3], const float box[6], float tmin, float tmax,
                     float* t_enter)
{
    om_slab s;
    om_slab_setup(o, d, &s);
    return om_slab_box(&s, box, tmin, tmax, t_enter);
}

/* Closest hit over the LBVH.  Semantics follow KCIntersectBVH
 * (GPUAcceleratorBVHKC.cuh:229-346) + AcceptHit
 * (GPUAcceleratorCommonKC.cuh:33-78): accept t in [tmin, tmax), tmax shrinks
 * on every accepted hit, the first hit found wins ties.  Visit order is the
 * specified one: both child boxes are tested at their parent, the child with
 * the smaller entry t goes first (ties: left), the other is pushed with its
 * entry t and culled on pop when that entry exceeds the current tmax
 * (equivalent to re-testing its box). */
uint32_t om_trace_one(const float* tris, uint32_t n, const uint32_t* sorted_idx,
                             const uint32_t* left, const uint32_t* right, const float* node_aabb,
                             uint32_t root, const float o[3], const float d[3], float tmin,
                             float* tmax_io, float bary_out[2], int any_hit,
                             uint32_t* visits, uint32_t* tests)
{
    uint32_t hit = OM_INVALID;
    float tmax = *tmax_io;
    float inv_d[3] = {1.0f / d[0], 1.0f / d[1], 1.0f / d[2]};
    uint32_t nv = 0, nt = 0;
    if (n == 0) goto done;

    om_stack_entry stack[128];
    int sp = 0;
    uint32_t next = root;
    for (;;) {
        /* resolve `next`: test leaves, pop until an internal node is found */
        for (;;) {
            if (next == OM_INVALID) {
                for (;;) {
                    if (sp == 0) goto done;
                    om_stack_entry e = stack[--sp];
                    if (!(e.t > tmax)) { next = e.node; break; }
                }
            }
            if (next < n) {
                uint32_t prim = sorted_idx[next];
                float t, b[3];
                nt++;
                if (om_intersect_triangle(o, d, tris + (size_t)prim * 9, &t, b) &&
                    t >= tmin && t < tmax) {
                    tmax = t;
                    hit = prim;
                    bary_out[0] = b[0];
                    bary_out[1] = b[1];
                    if (any_hit) goto done;
                }
                next = OM_INVALID;
                continue;
            }
            break;
        }
        /* internal node */
        uint32_t p = next - n;
        nv++;
        uint32_t c0 = left[p], c1 = right[p];
        float e0, e1;
        int h0 = om_intersect_aabb(o, inv_d, node_aabb + (size_t)c0 * 6, tmin, tmax, &e0);
        int h1 = om_intersect_aabb(o, inv_d, node_aabb + (size_t)c1 * 6, tmin, tmax, &e1);
        if (h0 && h1) {
            if (e1 < e0) { next = c1; stack[sp].node = c0; stack[sp].t = e0; sp++; }
            else { next = c0; stack[sp].node = c1; stack[sp].t = e1; sp++; }
        } else if (h0) next = c0;
        else if (h1) next = c1;
        else next = OM_INVALID;
    }
done:
    *tmax_io = tmax;
    if (visits) *visits = nv;
    if (tests) *tests = nt;
    return hit;
}

void om_trace_closest(const float* tris, uint32_t n, const uint32_t* sorted_idx,
                      const uint32_t* left, const uint32_t* right, const float* node_aabb,
                      uint32_t root, const float* rays, uint32_t nrays, uint32_t pitch,
                      float tmin, uint32_t* hit_prim, float* hit_t, float* hit_bary,
                      uint32_t* stats)
{
    for (uint32_t i = 0; i < nrays; i++) {
        float o[3] = {rays[i], rays[pitch + i], rays[2 * (size_t)pitch + i]};
        float d[3] = {rays[3 * (size_t)pitch + i], rays[4 * (size_t)pitch + i], rays[5 * (size_t)pitch + i]};
        float tmax = rays[6 * (size_t)pitch + i];
        float b[2] = {0.0f, 0.0f};
        uint32_t nv, nt;
        uint32_t h = om_trace_one(tris, n, sorted_idx, left, right, node_aabb, root, o, d, tmin,
                                  &tmax, b, 0, &nv, &nt);
        hit_prim[i] = h;
        hit_t[i] = (h == OM_INVALID) ? 0.0f : tmax;
        hit_bary[i] = (h == OM_INVALID) ? 0.0f : b[0];
        hit_bary[(size_t)nrays + i] = (h == OM_INVALID) ? 0.0f : b[1];
        if (stats) { stats[2 * (size_t)i] = nv; stats[2 * (size_t)i + 1] = nt; }
    }
}

void om_trace_any(const float* tris, uint32_t n, const uint32_t* sorted_idx,
                  const uint32_t* left, const uint32_t* right, const float* node_aabb,
                  uint32_t root, const float* rays, uint32_t nrays, uint32_t pitch,
                  float tmin, uint32_t* hit_any)
{
    for (uint32_t i = 0; i < nrays; i++) {
        float o[3] = {rays[i], rays[pitch + i], rays[2 * (size_t)pitch + i]};
        float d[3] = {rays[3 * (size_t)pitch + i], rays[4 * (size_t)pitch + i], rays[5 * (size_t)pitch + i]};
        float tmax = rays[6 * (size_t)pitch + i];
        float b[2];
        uint32_t h = om_trace_one(tris, n, sorted_idx, left, right, node_aabb, root, o, d, tmin,
                                  &tmax, b, 1, NULL, NULL);
        hit_any[i] = (h != OM_INVALID) ? 1u : 0u;
    }
}

/* ------------------------------------------------------------------------
 * 4-wide traversal tree collapsed from the LBVH.  Every internal node at even
 * depth (root = 0) becomes a 4-wide node whose children are its binary
 * children, each internal child replaced by that child's two children:
 * slots [c0 or c0.left, c0.right, c1 or c1.left, c1.right], compacted in
 * that order, empty slots last.  Layout per node (32 floats): min x[4],
 * min y[4], min z[4], max x[4], max y[4], max z[4], child code[4] (uint
 * bits: LEAF_BIT | sorted leaf, internal index, or 0xFFFFFFFF = empty),
 * 4 pad.  Nodes are indexed by their binary internal index (odd-depth
 * entries are unused). */
#define OM_LEAF_BIT 0x80000000u

static uint32_t code_of(uint32_t g, uint32_t n) { return g < n ? (OM_LEAF_BIT | g) : (g - n); }

uint32_t om_bvh4_build(uint32_t n, const uint32_t* left, const uint32_t* right, const float* node_aabb,
                       uint32_t root, float* nodes4)
{
    if (n <= 1) return OM_LEAF_BIT;
    /* depth of every internal node, top-down from the root */
    uint8_t* depth = (uint8_t*)calloc(n - 1, 1);
    uint32_t* todo = (uint32_t*)malloc((size_t)n * 4);
    int top = 0;
    todo[top++] = root - n;
    depth[root - n] = 0;
    while (top) {
        uint32_t p = todo[--top];
        uint32_t cs[2] = {left[p], right[p]};
        for (int k = 0; k < 2; k++)
            if (cs[k] >= n) {
                depth[cs[k] - n] = (uint8_t)(depth[p] + 1);
                todo[top++] = cs[k] - n;
            }
    }
    for (uint32_t p = 0; p + 1 < n; p++) {
        if (depth[p] & 1u) continue;
        float* q = nodes4 + (size_t)p * 32;
        uint32_t code[4];
        int k = 0;
        uint32_t cs[2] = {left[p], right[p]};
        for (int c = 0; c < 2; c++) {
            uint32_t kids[2];
            int nk = 0;
            if (cs[c] < n) kids[nk++] = cs[c];
            else { kids[nk++] = left[cs[c] - n]; kids[nk++] = right[cs[c] - n]; }
            for (int j = 0; j < nk; j++, k++) {
                const float* b = node_aabb + (size_t)kids[j] * 6;
                for (int a = 0; a < 6; a++) q[a * 4 + k] = b[a];
                code[k] = code_of(kids[j], n);
            }
        }
        for (; k < 4; k++) {
            for (int a = 0; a < 3; a++) { q[a * 4 + k] = INFINITY; q[(a + 3) * 4 + k] = -INFINITY; }
            code[k] = OM_INVALID;
        }
        for (int j = 0; j < 4; j++) q[24 + j] = om_u2f(code[j]);
        for (int j = 0; j < 4; j++) q[28 + j] = 0.0f;
    }
    free(depth);
    free(todo);
    return root - n;
}

/* Closest / any hit over the 4-wide tree.  At a node the occupied slots
 * whose boxes the ray enters (om_slab_box, within [tmin, current tmax]) are ordered by
 * (entry distance, slot); the nearest is taken, the others pushed farthest
 * first with their entry distance and culled on pop when it exceeds tmax.
 * Acceptance as om_trace_one. */
uint32_t om_trace4_one(const float* tris, uint32_t n, const uint32_t* sorted_idx, const float* nodes4,
                       uint32_t root4, const float o[3], const float d[3], float tmin, float* tmax_io,
                       float bary_out[2], int any_hit, uint32_t* visits, uint32_t* tests)
{
    uint32_t hit = OM_INVALID;
    float tmax = *tmax_io;
    om_slab sl;
    om_slab_setup(o, d, &sl);
    uint32_t nv = 0, nt = 0;
    om_stack_entry stack[256];
    int sp = 0;
    uint32_t next = root4;
    if (n == 0) goto done4;
    for (;;) {
        for (;;) {
            if (next == OM_INVALID) {
                for (;;) {
                    if (sp == 0) goto done4;
                    om_stack_entry e = stack[--sp];
                    if (!(e.t > tmax)) { next = e.node; break; }
                }
            }
            if (next & OM_LEAF_BIT) {
                uint32_t prim = sorted_idx[next & ~OM_LEAF_BIT];
                float t, b[3];
                nt++;
                if (om_intersect_triangle(o, d, tris + (size_t)prim * 9, &t, b) && t >= tmin && t < tmax) {
                    tmax = t;
                    hit = prim;
                    bary_out[0] = b[0];
                    bary_out[1] = b[1];
                    if (any_hit) goto done4;
                }
                next = OM_INVALID;
                continue;
            }
            break;
        }
        nv++;
        const float* q = nodes4 + (size_t)next * 32;
        float key[4];
        uint32_t slot[4], code[4];
        int cnt = 0;
        for (int k = 0; k < 4; k++) {
            uint32_t c = om_f2u(q[24 + k]);
            float box[6] = {q[k], q[4 + k], q[8 + k], q[12 + k], q[16 + k], q[20 + k]};
            float e;
            int h = (c != OM_INVALID) && om_slab_box(&sl, box, tmin, tmax, &e);
            key[k] = h ? e : INFINITY;
            slot[k] = (uint32_t)k;
            code[k] = c;
            cnt += h;
        }
        /* sorting network (0,1)(2,3)(0,2)(1,3)(1,2) on (key, slot) */
        static const int net[5][2] = {{0, 1}, {2, 3}, {0, 2}, {1, 3}, {1, 2}};
        for (int s = 0; s < 5; s++) {
            int a = net[s][0], b = net[s][1];
            if (key[a] > key[b] || (key[a] == key[b] && slot[a] > slot[b])) {
                float tk = key[a]; key[a] = key[b]; key[b] = tk;
                uint32_t ts = slot[a]; slot[a] = slot[b]; slot[b] = ts;
                uint32_t tc = code[a]; code[a] = code[b]; code[b] = tc;
            }
        }
        if (cnt == 0) { next = OM_INVALID; continue; }
        for (int k = cnt - 1; k >= 1; k--) { stack[sp].node = code[k]; stack[sp].t = key[k]; sp++; }
        next = code[0];
    }
done4:
    *tmax_io = tmax;
    if (visits) *visits = nv;
    if (tests) *tests = nt;
    return hit;
}

void om_trace4_closest(const float* tris, uint32_t n, const uint32_t* sorted_idx, const float* nodes4,
                       uint32_t root4, const float* rays, uint32_t nrays, uint32_t pitch, float tmin,
                       uint32_t* hit_prim, float* hit_t, float* hit_bary, uint32_t* stats)
{
    for (uint32_t i = 0; i < nrays; i++) {
        float o[3] = {rays[i], rays[pitch + i], rays[2 * (size_t)pitch + i]};
        float d[3] = {rays[3 * (size_t)pitch + i], rays[4 * (size_t)pitch + i], rays[5 * (size_t)pitch + i]};
        float tmax = rays[6 * (size_t)pitch + i];
        float b[2] = {0.0f, 0.0f};
        uint32_t nv, nt;
        uint32_t h = om_trace4_one(tris, n, sorted_idx, nodes4, root4, o, d, tmin, &tmax, b, 0, &nv, &nt);
        hit_prim[i] = h;
        hit_t[i] = (h == OM_INVALID) ? 0.0f : tmax;
        hit_bary[i] = (h == OM_INVALID) ? 0.0f : b[0];
        hit_bary[(size_t)nrays + i] = (h == OM_INVALID) ? 0.0f : b[1];
        if (stats) { stats[2 * (size_t)i] = nv; stats[2 * (size_t)i + 1] = nt; }
    }
}

void om_trace4_any(const float* tris, uint32_t n, const uint32_t* sorted_idx, const float* nodes4,
                   uint32_t root4, const float* rays, uint32_t nrays, uint32_t pitch, float tmin,
                   uint32_t* hit_any)
{
    for (uint32_t i = 0; i < nrays; i++) {
        float o[3] = {rays[i], rays[pitch + i], rays[2 * (size_t)pitch + i]};
        float d[3] = {rays[3 * (size_t)pitch + i], rays[4 * (size_t)pitch + i], rays[5 * (size_t)pitch + i]};
        float tmax = rays[6 * (size_t)pitch + i];
        float b[2];
        uint32_t h = om_trace4_one(tris, n, sorted_idx, nodes4, root4, o, d, tmin, &tmax, b, 1, NULL, NULL);
        hit_any[i] = (h != OM_INVALID) ? 1u : 0u;
    }
}

/* Brute force over all primitives in index order (the semantics of
 * KCIntersectLinear, GPUAcceleratorLinearKC.cuh:24-108). */
void om_trace_brute(const float* tris, uint32_t n, const float* rays, uint32_t nrays,
                    uint32_t pitch, float tmin, uint32_t* hit_prim, float* hit_t,
                    float* hit_bary)
{
    for (uint32_t i = 0; i < nrays; i++) {
        float o[3] = {rays[i], rays[pitch + i], rays[2 * (size_t)pitch + i]};
        float d[3] = {rays[3 * (size_t)pitch + i], rays[4 * (size_t)pitch + i], rays[5 * (size_t)pitch + i]};
        float tmax = rays[6 * (size_t)pitch + i];
        uint32_t best = OM_INVALID;
        float bb[2] = {0.0f, 0.0f};
        for (uint32_t k = 0; k < n; k++) {
            float t, b[3];
            if (om_intersect_triangle(o, d, tris + (size_t)k * 9, &t, b) && t >= tmin && t < tmax) {
                tmax = t; best = k; bb[0] = b[0]; bb[1] = b[1];
            }
        }
        hit_prim[i] = best;
        hit_t[i] = (best == OM_INVALID) ? 0.0f : tmax;
        hit_bary[i] = bb[0];
        hit_bary[(size_t)nrays + i] = bb[1];
    }
}

/* ------------------------------------------------------------- two-level */

/* Matrix<4>::operator*(Vector<4>) (RayLib/Matrix.hpp:268-289) for one row
 * of an affine matrix: 0 + m0 v0 + m1 v1 + m2 v2 + m3 w, left to right. */
static float xf_row(const float* m, float x, float y, float z, float w)
{
    float r = 0.0f;
    r = r + m[0] * x;
    r = r + m[1] * y;
    r = r + m[2] * z;
    r = r + m[3] * w;
    return r;
}

/* Ray::Transform (RayLib/Ray.hpp:461-465): position with w = 1, direction
 * with w = 0, the direction not renormalised. */
void om_xf_ray(const float m12[12], const float o[3], const float d[3], float out_o[3], float out_d[3])
{
    for (int r = 0; r < 3; r++) {
        out_o[r] = xf_row(m12 + 4 * r, o[0], o[1], o[2], 1.0f);
        out_d[r] = xf_row(m12 + 4 * r, d[0], d[1], d[2], 0.0f);
    }
}

void om_xf_point(const float m12[12], const float p[3], float out[3])
{
    for (int r = 0; r < 3; r++) out[r] = xf_row(m12 + 4 * r, p[0], p[1], p[2], 1.0f);
}

/* World-to-local matrix: the inverse of the affine local-to-world matrix in
 * double (adjugate over determinant, translation -A^-1 t), rounded to
 * float.  Returns 0 when singular or not finite. */
int om_affine_inverse(const float m[12], float out[12])
{
    const double a00 = (double)m[0], a01 = (double)m[1], a02 = (double)m[2], t0 = (double)m[3];
    const double a10 = (double)m[4], a11 = (double)m[5], a12 = (double)m[6], t1 = (double)m[7];
    const double a20 = (double)m[8], a21 = (double)m[9], a22 = (double)m[10], t2 = (double)m[11];
    const double c00 = a11 * a22 - a12 * a21;
    const double c01 = a12 * a20 - a10 * a22;
    const double c02 = a10 * a21 - a11 * a20;
    const double det = a00 * c00 + a01 * c01 + a02 * c02;
    if (!(det != 0.0) || !isfinite(det)) return 0;
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
        if (!isfinite(out[i])) return 0;
    return 1;
}

/* GPUTransformSingle::LocalToWorld(AABB3f) (GPUTransformSingle.cuh:180-197):
 * the box's eight corners (bit 0 -> x max, bit 1 -> y, bit 2 -> z) taken
 * to the world, min / max over them. */
void om_world_box(const float l2w[12], const float box[6], float out[6])
{
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < 8; i++) {
        float x = (i & 1) ? box[3] : box[0];
        float y = (i & 2) ? box[4] : box[1];
        float z = (i & 4) ? box[5] : box[2];
        for (int a = 0; a < 3; a++) {
            float v = xf_row(l2w + 4 * a, x, y, z, 1.0f);
            mn[a] = fminf(mn[a], v);
            mx[a] = fmaxf(mx[a], v);
        }
    }
    for (int a = 0; a < 3; a++) {
        out[a] = mn[a];
        out[3 + a] = mx[a];
    }
}

/* Closest / any hit over a two-level scene: the top tree over the instance
 * boxes walked 4-wide as om_trace4_one; at an instance the ray goes to local
 * space by the instance's world-to-local matrix and the mesh's tree (4-wide
 * or binary by width) is walked with the current closest t as its tmax
 * (the reference resumes its base BVH per leaf, GPUAcceleratorBVHKC.cuh:
 * 513-652, and intersects the leaf's accelerator with the transformed ray,
 * GPUAcceleratorBVHKC.cuh:274-345).  Returns the scene primitive; *inst_out
 * the instance.  Counts: top-tree internal nodes plus the meshes' nodes and
 * triangle tests. */
uint32_t om_trace2l_one(const om_two_level* tl, const float* tris, uint32_t width, const float o[3],
                        const float d[3], float tmin, float* tmax_io, float bary_out[2], int any_hit,
                        uint32_t* inst_out, uint32_t* visits, uint32_t* tests)
{
    uint32_t hit = OM_INVALID, hinst = OM_INVALID;
    float tmax = *tmax_io;
    om_slab sl;
    om_slab_setup(o, d, &sl);
    uint32_t nv = 0, nt = 0;
    om_stack_entry stack[128];
    int sp = 0;
    uint32_t next = tl->top_root4;
    for (;;) {
        for (;;) {
            if (next == OM_INVALID) {
                for (;;) {
                    if (sp == 0) goto done2;
                    om_stack_entry e = stack[--sp];
                    if (!(e.t > tmax)) { next = e.node; break; }
                }
            }
            if (next & OM_LEAF_BIT) {
                uint32_t ii = tl->top_sorted_idx[next & ~OM_LEAF_BIT];
                next = OM_INVALID;
                float lo[3], ld[3];
                om_xf_ray(tl->w2l + (size_t)ii * 12, o, d, lo, ld);
                uint32_t m = tl->inst_mesh[ii];
                uint32_t f = tl->mesh_first[m], n = tl->mesh_first[m + 1] - f;
                float bt = tmax, bb[2] = {0.0f, 0.0f};
                uint32_t bv = 0, bn = 0, h;
                if (width == 4)
                    h = om_trace4_one(tris + (size_t)f * 9, n, tl->mesh_sorted_idx + f,
                                      tl->mesh_nodes4 + (size_t)tl->mesh_ibase[m] * 32, tl->mesh_root4[m], lo, ld,
                                      tmin, &bt, bb, any_hit, &bv, &bn);
                else
                    h = om_trace_one(tris + (size_t)f * 9, n, tl->mesh_sorted_idx + f, tl->mesh_left + tl->mesh_ibase[m],
                                     tl->mesh_right + tl->mesh_ibase[m],
                                     tl->mesh_node_aabb + (size_t)tl->mesh_abase[m] * 6, tl->mesh_root[m], lo, ld,
                                     tmin, &bt, bb, any_hit, &bv, &bn);
                nv += bv;
                nt += bn;
                if (h != OM_INVALID) {
                    tmax = bt;
                    hit = f + h;
                    hinst = ii;
                    bary_out[0] = bb[0];
                    bary_out[1] = bb[1];
                    if (any_hit) goto done2;
                }
                continue;
            }
            break;
        }
        nv++;
        const float* q = tl->top_nodes4 + (size_t)next * 32;
        float key[4];
        uint32_t slot[4], code[4];
        int cnt = 0;
        for (int k = 0; k < 4; k++) {
            uint32_t c = om_f2u(q[24 + k]);
            float box[6] = {q[k], q[4 + k], q[8 + k], q[12 + k], q[16 + k], q[20 + k]};
            float e;
            int h = (c != OM_INVALID) && om_slab_box(&sl, box, tmin, tmax, &e);
            key[k] = h ? e : INFINITY;
            slot[k] = (uint32_t)k;
            code[k] = c;
            cnt += h;
        }
        static const int net[5][2] = {{0, 1}, {2, 3}, {0, 2}, {1, 3}, {1, 2}};
        for (int s = 0; s < 5; s++) {
            int a = net[s][0], b = net[s][1];
            if (key[a] > key[b] || (key[a] == key[b] && slot[a] > slot[b])) {
                float tk = key[a]; key[a] = key[b]; key[b] = tk;
                uint32_t ts = slot[a]; slot[a] = slot[b]; slot[b] = ts;
                uint32_t tc = code[a]; code[a] = code[b]; code[b] = tc;
            }
        }
        if (cnt == 0) { next = OM_INVALID; continue; }
        for (int k = cnt - 1; k >= 1; k--) { stack[sp].node = code[k]; stack[sp].t = key[k]; sp++; }
        next = code[0];
    }
done2:
    *tmax_io = tmax;
    if (inst_out) *inst_out = hinst;
    if (visits) *visits = nv;
    if (tests) *tests = nt;
    return hit;
}

void om_trace2l_closest(const om_two_level* tl, const float* tris, uint32_t width, const float* rays, uint32_t nrays,
                        uint32_t pitch, float tmin, uint32_t* hit_inst, uint32_t* hit_prim, float* hit_t,
                        float* hit_bary, uint32_t* stats)
{
    for (uint32_t i = 0; i < nrays; i++) {
        float o[3] = {rays[i], rays[pitch + i], rays[2 * (size_t)pitch + i]};
        float d[3] = {rays[3 * (size_t)pitch + i], rays[4 * (size_t)pitch + i], rays[5 * (size_t)pitch + i]};
        float tmax = rays[6 * (size_t)pitch + i];
        float b[2] = {0.0f, 0.0f};
        uint32_t nv, nt, ii;
        uint32_t h = om_trace2l_one(tl, tris, width, o, d, tmin, &tmax, b, 0, &ii, &nv, &nt);
        hit_inst[i] = (h == OM_INVALID) ? OM_INVALID : ii;
        hit_prim[i] = h;
        hit_t[i] = (h == OM_INVALID) ? 0.0f : tmax;
        hit_bary[i] = (h == OM_INVALID) ? 0.0f : b[0];
        hit_bary[(size_t)nrays + i] = (h == OM_INVALID) ? 0.0f : b[1];
        if (stats) { stats[2 * (size_t)i] = nv; stats[2 * (size_t)i + 1] = nt; }
    }
}
