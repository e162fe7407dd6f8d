// mrh_bvh.hpp — LBVH storage and the closest/any-hit traversal routine
// shared by the ray-cast entry points and the path tracer.
//
// Two layouts are kept per tree:
//  * the reference's node numbering (leaves [0,n), internal node p at n+p,
//    left/right/parent in global node indices, 6-float AABBs per node) —
//    the bit-exact build output;
//  * the traversal layout: one 64-byte record per internal node holding both
//    child boxes and both child codes (bit 31 = leaf, else internal index p),
//    and one 48-byte record per leaf in sorted order holding the triangle's
//    three corners, its primitive index and its HitKey.  One 64-byte load
//    answers both child slab tests; a leaf needs no indirection.
#pragma once

#include "mrh_internal.hpp"

#include <cfloat>

struct mrh_lbvh {
    mrh_context* ctx = nullptr;
    uint32_t n = 0, bits = 30, root = 0, max_depth = 0;
    float* tris = nullptr;          // [n][9] primitive order
    float* leaf_aabb = nullptr;     // [n][6] primitive order
    void* codes = nullptr;          // unsorted codes
    void* sorted_codes = nullptr;
    uint32_t* sorted_idx = nullptr;
    uint32_t* left = nullptr;       // [n-1]
    uint32_t* right = nullptr;      // [n-1]
    uint32_t* parent = nullptr;     // [2n-1]
    float* node_aabb = nullptr;     // [2n-1][6]
    uint32_t* ranges = nullptr;     // [n-1][2]
    uint32_t* counters = nullptr;   // [n-1]
    float4* tnodes = nullptr;       // [n-1][4]
    float4* leaves = nullptr;       // [n][3]
    uint32_t* prim_key = nullptr;   // [n] (optional HitKeys, primitive order)
    uint32_t* meta = nullptr;       // [0] root, [1] max depth, [2..7] extent encodings
    std::vector<void*> allocs;
};

namespace mrh {

constexpr uint32_t LEAF_BIT = 0x80000000u;
constexpr uint32_t NO_HIT = 0xFFFFFFFFu;

int32_t morton_codes(mrh_context* ctx, const float* d_tris, uint32_t n, uint32_t bits, void* codes,
                     float* aabb_out, uint32_t* meta_out);
// Builds an LBVH over device triangles; d_prim_key (optional) is stored in
// the leaf records.
int32_t lbvh_create(mrh_context* ctx, const float* d_tris, uint32_t n, uint32_t bits,
                    const uint32_t* d_prim_key, mrh_lbvh** out);
int32_t lbvh_build_into(mrh_lbvh* b, const float* d_tris_src);
void lbvh_free(mrh_lbvh* b);

// Traversal entry code of the root
inline uint32_t root_code(const mrh_lbvh* b) { return (b->n == 1) ? LEAF_BIT : (b->root - b->n); }

struct TraceResult {
    uint32_t leaf;   // sorted leaf index or NO_HIT
    float t, ba, bb;
    uint32_t visits, tests;
};

// Closest (ANY = false) or any (ANY = true) hit.  Visit order: both child
// boxes are tested at the parent; the child with the smaller entry distance
// is taken first (ties: left); the other is pushed with its entry distance
// and dropped on pop when that distance exceeds the current tmax (the same
// outcome as re-testing its box).  Accepts t in [tmin, tmax); the first hit
// found at a given t wins.  STACK must be >= the tree depth.
template <bool ANY, bool STATS, int STACK>
__device__ __forceinline__ TraceResult traverse(const float4* __restrict__ tnodes,
                                                const float4* __restrict__ leaves, uint32_t root,
                                                uint32_t n_leaves, V3 o, V3 d, float tmin,
                                                float tmax)
{
    TraceResult r;
    r.leaf = NO_HIT;
    r.t = tmax;
    r.ba = r.bb = 0.0f;
    r.visits = r.tests = 0;
    const V3 inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    uint32_t st_node[STACK];
    float st_t[STACK];
    int sp = 0;
    uint32_t next = root;
    for (;;) {
        // resolve `next`: intersect leaves, pop until an internal node
        for (;;) {
            if (next == NO_HIT) {
                bool found = false;
                while (sp > 0) {
                    --sp;
                    if (!(st_t[sp] > r.t)) {
                        next = st_node[sp];
                        found = true;
                        break;
                    }
                }
                if (!found) return r;
            }
            if (next & LEAF_BIT) {
                const uint32_t li = next & ~LEAF_BIT;
                if (li >= n_leaves) {   // unreachable on a well-formed tree
                    next = NO_HIT;
                    continue;
                }
                const float4 a = leaves[3 * li + 0];
                const float4 b = leaves[3 * li + 1];
                const float4 c = leaves[3 * li + 2];
                float t, ba, bb;
                if (STATS) r.tests++;
                if (intersect_tri(o, d, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), t,
                                  ba, bb) &&
                    t >= tmin && t < r.t) {
                    r.t = t;
                    r.leaf = li;
                    r.ba = ba;
                    r.bb = bb;
                    if (ANY) return r;
                }
                next = NO_HIT;
                continue;
            }
            if (next + 1u >= n_leaves) {   // internal index out of range: unreachable
                next = NO_HIT;
                continue;
            }
            break;
        }
        if (STATS) r.visits++;
        const float4* nd = tnodes + 4 * (size_t)next;
        const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2], n3 = nd[3];
        float e0, e1;
        const bool h0 = intersect_box(o, inv, n0.x, n0.y, n0.z, n0.w, n1.x, n1.y, tmin, r.t, e0);
        const bool h1 = intersect_box(o, inv, n1.z, n1.w, n2.x, n2.y, n2.z, n2.w, tmin, r.t, e1);
        const uint32_t c0 = __float_as_uint(n3.x), c1 = __float_as_uint(n3.y);
        if (h0 && h1) {
            // the host only launches with STACK >= tree depth; the guard keeps
            // a malformed tree from writing past the private array
            const bool room = sp < STACK;
            if (e1 < e0) {
                next = c1;
                if (room) { st_node[sp] = c0; st_t[sp] = e0; }
            } else {
                next = c0;
                if (room) { st_node[sp] = c1; st_t[sp] = e1; }
            }
            sp += room ? 1 : 0;
        } else if (h0) {
            next = c0;
        } else if (h1) {
            next = c1;
        } else {
            next = NO_HIT;
        }
    }
}

}  // namespace mrh
