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
    float* leaf_aabb = nullptr;     // [n][6] primitive order
    void* codes = nullptr;          // unsorted codes
    void* sorted_codes = nullptr;
    uint32_t* sorted_idx = nullptr;
    uint32_t* left = nullptr;       // [n-1]
    uint32_t* right = nullptr;      // [n-1]
    uint32_t* parent = nullptr;     // [2n-1]
    float* node_aabb = nullptr;     // [2n-1][6]
    uint32_t* halves = nullptr;     // [n-1][16] build handshake: each child's box, node and range end
    uint32_t* counters = nullptr;   // [n-1]
    float4* tnodes = nullptr;       // [n-1][4]
    float4* leaves = nullptr;       // [n][3]
    uint32_t* prim_key = nullptr;   // [n] (optional HitKeys, primitive order)
    uint32_t* meta = nullptr;       // [0] root, [1] max depth, [2..7] extent encodings
    float lo[3] = {}, hi[3] = {};   // scene bounds (host copy)
    float4* nodes4 = nullptr;       // [n-1][8] 4-wide traversal nodes (even-depth entries used)
    uint32_t root4 = 0;             // traversal code of the 4-wide root
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
int32_t lbvh_build_into(mrh_lbvh* b, const float* d_tris);
void lbvh_free(mrh_lbvh* b);

// Persistent grid of the traversal kernels: a multiple of 8 blocks (one
// slice per XCD) of `block` threads.
inline uint32_t trace_grid(mrh_context* ctx, uint64_t rays, uint32_t block = 256)
{
    uint64_t want = (rays + block - 1) / block;
    uint64_t cap = (uint64_t)ctx->num_cus * ctx->trace_blocks_per_cu;
    if (want > cap) want = cap;
    want = (want + 7) & ~7ull;
    return (uint32_t)(want ? want : 8);
}

// Overflow buffer for stack entries beyond the LDS part (per grid thread).
inline int32_t trace_overflow(mrh_context* ctx, uint32_t grid, uint32_t cap, uint32_t lds, uint2** out,
                              uint32_t block = 256)
{
    *out = nullptr;
    if (lds == 0 || cap <= lds) return MRH_OK;
    void* p = nullptr;
    MRH_TRY(scratch(ctx, SLOT_TRACE_OVERFLOW, (size_t)(cap - lds) * grid * block * sizeof(uint2), &p));
    *out = (uint2*)p;
    return MRH_OK;
}

// Traversal entry code of the root
inline uint32_t root_code(const mrh_lbvh* b) { return (b->n == 1) ? LEAF_BIT : (b->root - b->n); }

// Traversal tree chosen for a launch: the binary LBVH (width 2) or the
// 4-wide tree collapsed from it (width 4; its stack needs
// 3 * ceil(depth / 2) + 1 entries, so trees deeper than 84 use width 2).
struct TreeSel {
    const float4* nodes;
    uint32_t root, width, cap;
};

inline int32_t select_tree(mrh_context* ctx, const mrh_lbvh* b, TreeSel& t)
{
    uint32_t w = ctx->trace_width;
    // the 4-wide stack bound, and 32-bit byte offsets of its node records
    if (w == 4 && (b->max_depth > 84 || b->n > (1u << 25))) w = 2;
    t.width = w;
    if (w == 4) {
        t.nodes = b->nodes4;
        t.root = b->root4;
        t.cap = (b->max_depth <= 42) ? 64 : 128;
        return MRH_OK;
    }
    if (b->max_depth > 128)
        return fail(ctx, MRH_ERR_UNSUPPORTED, "tree depth %u exceeds the traversal stack (128)", b->max_depth);
    t.nodes = b->tnodes;
    t.root = root_code(b);
    t.cap = (b->max_depth <= 64) ? 64 : 128;
    return MRH_OK;
}

struct TraceResult {
    uint32_t leaf;   // sorted leaf index or NO_HIT
    float t, ba, bb;
    uint32_t visits, tests;
};

// Traversal stacks.  An entry is (child code, entry distance).
//
// PrivateStack keeps the whole stack in a private array (scratch memory,
// which the L2 writes back to HBM).  LdsStack keeps the first L entries of
// every thread in LDS, entry-major and thread-minor so a wave's accesses
// spread over the banks, and spills deeper entries to the thread's own slice
// of a global overflow buffer.
template <int N>
struct PrivateStack {
    static constexpr int CAP = N;
    static constexpr int LDS_ENTRIES = 0;
    uint32_t node[N];
    float t[N];
    __device__ __forceinline__ void put(int i, uint32_t n, float d)
    {
        node[i] = n;
        t[i] = d;
    }
    __device__ __forceinline__ void get(int i, uint32_t& n, float& d) const
    {
        n = node[i];
        d = t[i];
    }
};

// LDS pointers with their address space spelled out, so no access through
// them is compiled as a generic (flat) access.
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(3))) float lds_f32;

// Waves per SIMD a trace kernel asks the register allocator for: `want`,
// or fewer when its LDS stack (L entries of 8 bytes per thread, BLOCK
// threads per block, 160 KiB of LDS per CU, four SIMDs) already holds the
// CU below that; asking for more would only cap registers for nothing.
constexpr int lds_stack_waves(int L, int BLOCK, int want)
{
    if (L <= 0) return want;
    const int blocks = (160 * 1024) / (L * BLOCK * 8);
    const int waves = blocks * (BLOCK / 64) / 4;
    return waves < 1 ? 1 : (waves < want ? waves : want);
}

template <int L, int N, int BLOCK>
struct LdsStack {
    static constexpr int CAP = N;
    static constexpr int LDS_ENTRIES = L;
    lds_u32* s_node;      // [L][BLOCK]
    lds_f32* s_t;         // [L][BLOCK]
    uint32_t slot;        // threadIdx.x
    uint2* overflow;      // [(N-L)][ostride]
    size_t ostride, oslot;
    __device__ __forceinline__ LdsStack(uint32_t* node, float* t, uint32_t sl, uint2* ovf, size_t ost, size_t osl)
        : s_node((lds_u32*)node), s_t((lds_f32*)t), slot(sl), overflow(ovf), ostride(ost), oslot(osl)
    {
    }
    __device__ __forceinline__ void put(int i, uint32_t n, float d)
    {
        if (i < L) {
            s_node[i * BLOCK + slot] = n;
            s_t[i * BLOCK + slot] = d;
        } else {
            overflow[(size_t)(i - L) * ostride + oslot] = make_uint2(n, __float_as_uint(d));
        }
    }
    // entry i < L, no check
    __device__ __forceinline__ void put_lds(int i, uint32_t n, float d)
    {
        s_node[i * BLOCK + slot] = n;
        s_t[i * BLOCK + slot] = d;
    }
    __device__ __forceinline__ void get(int i, uint32_t& n, float& d) const
    {
        if (i < L) {
            n = s_node[i * BLOCK + slot];
            d = s_t[i * BLOCK + slot];
        } else {
            const uint2 e = overflow[(size_t)(i - L) * ostride + oslot];
            n = e.x;
            d = __uint_as_float(e.y);
        }
    }
};

// Logical block index that gives each of the 8 XCDs (workgroups are handed
// to XCDs round-robin) a contiguous slice of the ray range, so neighbouring
// rays share an L2.  Requires gridDim.x % 8 == 0.
__device__ __forceinline__ uint32_t xcd_block()
{
    return (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
}

// Closest (ANY = false) or any (ANY = true) hit.  Visit order: both child
// boxes are tested at the parent; the child with the smaller entry distance
// is taken first (ties: left); the other is pushed with its entry distance
// and dropped on pop when that distance exceeds the current tmax (the same
// outcome as re-testing its box).  Accepts t in [tmin, tmax); the first hit
// found at a given t wins.  Stack::CAP must be >= the tree depth.
template <bool ANY, bool STATS, class Stack>
__device__ __forceinline__ TraceResult traverse(Stack& st, const float4* __restrict__ tnodes,
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
    int sp = 0;
    uint32_t next = root;
    for (;;) {
        // resolve `next`: intersect leaves, pop until an internal node
        for (;;) {
            if (next == NO_HIT) {
                bool found = false;
                while (sp > 0) {
                    --sp;
                    uint32_t nn;
                    float et;
                    st.get(sp, nn, et);
                    if (!(et > r.t)) {
                        next = nn;
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
            // the host only launches with CAP >= tree depth; the guard keeps a
            // malformed tree from writing past the stack
            const bool room = sp < Stack::CAP;
            if (e1 < e0) {
                next = c1;
                if (room) st.put(sp, c0, e0);
            } else {
                next = c0;
                if (room) st.put(sp, c1, e1);
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

// Per-ray constants of the 4-wide node test, the slab test with each
// axis's two planes ordered by the sign of the direction: the near plane
// (the min plane where the reciprocal direction is non-negative, else the
// max plane) gives that axis's entry distance and the far plane its exit,
// each as fma(plane, inv, -o * inv).  A box then costs six fused
// multiply-adds and no per-axis min/max; the near/far choice is made once
// per ray, as byte offsets into the node record.  Direction components
// below 2^-64 in magnitude are replaced by +-2^-64 (sign kept), so every
// plane distance is finite and no 0 * inf arises.
struct RaySlab {
    V3 inv, noi;             // reciprocal direction, -o * inv
    uint32_t nx, ny, nz;     // byte offsets of the near planes in a 4-wide node (far = near ^ FAR_*)
};
// near <-> far plane offsets: x 0 <-> 48, y 16 <-> 64, z 32 <-> 80
constexpr uint32_t FAR_X = 48u, FAR_Y = 80u, FAR_Z = 112u;

__device__ __forceinline__ float slab_rcp(float d)
{
    const float tiny = 0x1p-64f;
    return 1.0f / (fabsf(d) < tiny ? copysignf(tiny, d) : d);
}

__device__ __forceinline__ RaySlab make_slab(V3 o, V3 d)
{
    RaySlab s;
    s.inv = v3(slab_rcp(d.x), slab_rcp(d.y), slab_rcp(d.z));
    s.noi = v3(-(o.x * s.inv.x), -(o.y * s.inv.y), -(o.z * s.inv.z));
    // node layout: min x/y/z at bytes 0/16/32, max x/y/z at 48/64/80
    s.nx = signbit(s.inv.x) ? 48u : 0u;
    s.ny = signbit(s.inv.y) ? 64u : 16u;
    s.nz = signbit(s.inv.z) ? 80u : 32u;
    return s;
}

// The 4 slots of the node at byte offset nb from base: slot j is entered
// when max(tmin, entry distances) <= min(tmax, exit distances); k.j = its
// entry distance.  Empty slots hold (+inf, -inf) boxes, whose entry is +inf
// and exit -inf for every ray, so they are never entered.
struct Node4Test {
    float4 k;
    uint4 c;
    bool h0, h1, h2, h3;
};

__device__ __forceinline__ Node4Test node4_test(const char* __restrict__ base, uint32_t nb, const RaySlab& s,
                                                float tmin, float tmax)
{
    // nb has its low 7 bits clear, so | and ^ act on the offsets alone
    const uint32_t ax = nb | s.nx, ay = nb | s.ny, az = nb | s.nz;
    const float4 ex = *(const float4*)(base + (size_t)ax);
    const float4 ey = *(const float4*)(base + (size_t)ay);
    const float4 ez = *(const float4*)(base + (size_t)az);
    const float4 xx = *(const float4*)(base + (size_t)(ax ^ FAR_X));
    const float4 xy = *(const float4*)(base + (size_t)(ay ^ FAR_Y));
    const float4 xz = *(const float4*)(base + (size_t)(az ^ FAR_Z));
    const float4 cc = *(const float4*)(base + (size_t)(nb | 96u));
    Node4Test q;
    q.c = make_uint4(__float_as_uint(cc.x), __float_as_uint(cc.y), __float_as_uint(cc.z), __float_as_uint(cc.w));
    auto slot = [&](float ax, float ay, float az, float bx, float by, float bz, float& k) {
        const float te = fmaxf(fmaxf(fmaxf(__builtin_fmaf(ax, s.inv.x, s.noi.x), __builtin_fmaf(ay, s.inv.y, s.noi.y)),
                                     __builtin_fmaf(az, s.inv.z, s.noi.z)),
                               tmin);
        const float tx = fminf(fminf(fminf(__builtin_fmaf(bx, s.inv.x, s.noi.x), __builtin_fmaf(by, s.inv.y, s.noi.y)),
                                     __builtin_fmaf(bz, s.inv.z, s.noi.z)),
                               tmax);
        k = te;
        return tx >= te;
    };
    q.h0 = slot(ex.x, ey.x, ez.x, xx.x, xy.x, xz.x, q.k.x);
    q.h1 = slot(ex.y, ey.y, ez.y, xx.y, xy.y, xz.y, q.k.y);
    q.h2 = slot(ex.z, ey.z, ez.z, xx.z, xy.z, xz.z, q.k.z);
    q.h3 = slot(ex.w, ey.w, ez.w, xx.w, xy.w, xz.w, q.k.w);
    return q;
}

// Orders the entered slots of a node by (entry distance, slot), pushes all
// but the nearest farthest first, and returns the nearest (NO_HIT if none).
template <class Stack>
__device__ __forceinline__ uint32_t node4_order_push(Stack& st, int& sp, const Node4Test& q)
{
    const int cnt = (int)q.h0 + (int)q.h1 + (int)q.h2 + (int)q.h3;
    if (cnt == 0) return NO_HIT;
    float k0 = q.h0 ? q.k.x : INFINITY, k1 = q.h1 ? q.k.y : INFINITY, k2 = q.h2 ? q.k.z : INFINITY,
          k3 = q.h3 ? q.k.w : INFINITY;
    uint32_t c0 = q.c.x, c1 = q.c.y, c2 = q.c.z, c3 = q.c.w;
    uint32_t s0 = 0, s1 = 1, s2 = 2, s3 = 3;
    // (key, slot) lexicographic compare-exchange: the order is unique
    auto cx = [](float& ka, uint32_t& sa, uint32_t& ca, float& kb, uint32_t& sb, uint32_t& cb) {
        const bool sw = (ka > kb) || (ka == kb && sa > sb);
        const float tk = sw ? kb : ka;
        kb = sw ? ka : kb;
        ka = tk;
        const uint32_t ts = sw ? sb : sa;
        sb = sw ? sa : sb;
        sa = ts;
        const uint32_t tc = sw ? cb : ca;
        cb = sw ? ca : cb;
        ca = tc;
    };
    cx(k0, s0, c0, k1, s1, c1);
    cx(k2, s2, c2, k3, s3, c3);
    cx(k0, s0, c0, k2, s2, c2);
    cx(k1, s1, c1, k3, s3, c3);
    cx(k1, s1, c1, k2, s2, c2);
    if constexpr (Stack::LDS_ENTRIES >= 3) {
        // Room for three more entries in the LDS part (nearly always): write
        // three entries without branches, the pushed ones farthest first
        // and then whatever; only the cnt - 1 pushed ones are below the new
        // top, so the rest is never read.
        if (sp + 3 <= Stack::LDS_ENTRIES) {
            const bool four = cnt >= 4, three = cnt == 3;
            st.put_lds(sp, four ? c3 : (three ? c2 : c1), four ? k3 : (three ? k2 : k1));
            st.put_lds(sp + 1, four ? c2 : c1, four ? k2 : k1);
            st.put_lds(sp + 2, c1, k1);
            sp += cnt - 1;
            return c0;
        }
    }
    // push farthest first; the guard keeps a malformed tree from writing
    // past the stack (the host sizes CAP for the tree depth)
    if (cnt >= 4 && sp < Stack::CAP) st.put(sp++, c3, k3);
    if (cnt >= 3 && sp < Stack::CAP) st.put(sp++, c2, k2);
    if (cnt >= 2 && sp < Stack::CAP) st.put(sp++, c1, k1);
    return c0;
}

// Closest / any hit over the 4-wide tree.  Node layout (8 float4, 128 B):
// min x[4], min y[4], min z[4], max x[4], max y[4], max z[4], child code[4]
// (LEAF_BIT | leaf, internal index, or NO_HIT for an empty slot), pad.  The
// slots whose boxes the ray enters (node4_test) are ordered by (entry
// distance, slot) with a 5-comparator network; the nearest is taken next
// and the others are pushed farthest first, culled on pop once their entry
// distance exceeds tmax.  Acceptance as traverse().  Node records are
// addressed by 32-bit byte offsets (internal index << 7), so a tree holds
// at most 2^25 internal nodes (select_tree falls back to the binary tree).
template <bool ANY, bool STATS, class Stack>
__device__ __forceinline__ TraceResult traverse4(Stack& st, const float4* __restrict__ nodes4,
                                                 const float4* __restrict__ leaves, uint32_t root,
                                                 uint32_t n_leaves, V3 o, V3 d, float tmin, float tmax)
{
    TraceResult r;
    r.leaf = NO_HIT;
    r.t = tmax;
    r.ba = r.bb = 0.0f;
    r.visits = r.tests = 0;
    const RaySlab sl = make_slab(o, d);
    const char* nbase = (const char*)nodes4;
    int sp = 0;
    uint32_t next = root;
    for (;;) {
        for (;;) {
            if (next == NO_HIT) {
                bool found = false;
                while (sp > 0) {
                    --sp;
                    uint32_t nn;
                    float et;
                    st.get(sp, nn, et);
                    if (!(et > r.t)) {
                        next = nn;
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
                if (intersect_tri(o, d, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), t, ba, bb) &&
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
        next = node4_order_push(st, sp, node4_test(nbase, next << 7, sl, tmin, r.t));
    }
}

// Either traversal, by tree width
template <int W, bool ANY, bool STATS, class Stack>
__device__ __forceinline__ TraceResult traverse_w(Stack& st, const float4* __restrict__ nodes,
                                                  const float4* __restrict__ leaves, uint32_t root,
                                                  uint32_t n_leaves, V3 o, V3 d, float tmin, float tmax)
{
    if constexpr (W == 4) return traverse4<ANY, STATS>(st, nodes, leaves, root, n_leaves, o, d, tmin, tmax);
    else return traverse<ANY, STATS>(st, nodes, leaves, root, n_leaves, o, d, tmin, tmax);
}

// ------------------------------------------------------------ two-level

// Instance record, 8 float4 (128 B):
//   [0..2] world-to-local rows (m[r][0..3]), [3..5] local-to-world rows,
//   [6] (mesh, HitKey override or 0, first leaf of the mesh in the
//        concatenated leaf records, first node of the mesh's binary tree)
//   [7] (first node of the mesh's 4-wide tree, 4-wide root code, binary
//        root code, triangle count)
constexpr uint32_t INST_F4 = 8;

// Matrix<4>::operator*(Vector<4>) (RayLib/Matrix.hpp:268-289) for one row
// of an affine matrix: 0 + m0 v0 + m1 v1 + m2 v2 + m3 w, summed left to right.
__device__ __forceinline__ float xf_row(float4 m, V3 v, float w)
{
    float r = 0.0f;
    r = r + m.x * v.x;
    r = r + m.y * v.y;
    r = r + m.z * v.z;
    r = r + m.w * w;
    return r;
}
__device__ __forceinline__ V3 xf_point(float4 r0, float4 r1, float4 r2, V3 p)
{
    return v3(xf_row(r0, p, 1.0f), xf_row(r1, p, 1.0f), xf_row(r2, p, 1.0f));
}
__device__ __forceinline__ V3 xf_dir(float4 r0, float4 r1, float4 r2, V3 d)
{
    return v3(xf_row(r0, d, 0.0f), xf_row(r1, d, 0.0f), xf_row(r2, d, 0.0f));
}

// A stack that continues another one above its entry `base` (a mesh's
// traversal inside the top-level walk).  Entries past the parent's capacity
// are dropped on put and read back as culled, so a malformed tree cannot
// index past the stack (the host sizes CAP for both trees' depths).
template <class P>
struct OffsetStack {
    static constexpr int CAP = P::CAP;
    static constexpr int LDS_ENTRIES = 0;
    P& p;
    int base;
    __device__ __forceinline__ void put(int i, uint32_t n, float d)
    {
        if (base + i < CAP) p.put(base + i, n, d);
    }
    __device__ __forceinline__ void get(int i, uint32_t& n, float& d) const
    {
        if (base + i < CAP) {
            p.get(base + i, n, d);
        } else {
            n = NO_HIT;
            d = INFINITY;
        }
    }
};

// Closest / any hit over a two-level scene.  The top tree is walked 4-wide
// as traverse4 (its leaf records hold the instance index in the first
// corner's w); at an instance the ray goes to local space
// (GPUTransformSingle::WorldToLocal = Ray::Transform by the inverse,
// RayLib/Ray.hpp:461-465) and the mesh's tree (width W) is walked with the
// current closest t as its tmax on the same stack, above the top-level
// entries.  r.leaf is the index into the concatenated leaf records; inst
// receives the instance.  Counts add the top tree's internal nodes to the
// meshes' nodes and triangle tests.
template <int W, bool ANY, bool STATS, class Stack>
__device__ __forceinline__ TraceResult traverse_two_level(Stack& st, const float4* __restrict__ top4,
                                                          const float4* __restrict__ top_leaves, uint32_t top_root,
                                                          uint32_t n_inst, const float4* __restrict__ inst,
                                                          const float4* __restrict__ mesh_nodes,
                                                          const float4* __restrict__ leaves, V3 o, V3 d,
                                                          float tmin, float tmax, uint32_t& hit_inst)
{
    TraceResult r;
    r.leaf = NO_HIT;
    r.t = tmax;
    r.ba = r.bb = 0.0f;
    r.visits = r.tests = 0;
    hit_inst = NO_HIT;
    const RaySlab sl = make_slab(o, d);
    const char* tbase = (const char*)top4;
    int sp = 0;
    uint32_t next = top_root;
    for (;;) {
        for (;;) {
            if (next == NO_HIT) {
                bool found = false;
                while (sp > 0) {
                    --sp;
                    uint32_t nn;
                    float et;
                    st.get(sp, nn, et);
                    if (!(et > r.t)) {
                        next = nn;
                        found = true;
                        break;
                    }
                }
                if (!found) return r;
            }
            if (next & LEAF_BIT) {
                const uint32_t li = next & ~LEAF_BIT;
                next = NO_HIT;
                if (li >= n_inst) continue;   // unreachable on a well-formed tree
                const uint32_t ii = __float_as_uint(top_leaves[3 * (size_t)li].w);
                if (ii >= n_inst) continue;
                const float4* rec = inst + INST_F4 * (size_t)ii;
                const float4 w0 = rec[0], w1 = rec[1], w2 = rec[2];
                const float4 m6 = rec[6], m7 = rec[7];
                const V3 lo = xf_point(w0, w1, w2, o);
                const V3 ld = xf_dir(w0, w1, w2, d);
                const uint32_t leaf_base = __float_as_uint(m6.z);
                const uint32_t n_tris = __float_as_uint(m7.w);
                const float4* nodes;
                uint32_t root;
                if constexpr (W == 4) {
                    nodes = mesh_nodes + 8 * (size_t)__float_as_uint(m7.x);
                    root = __float_as_uint(m7.y);
                } else {
                    nodes = mesh_nodes + 4 * (size_t)__float_as_uint(m6.w);
                    root = __float_as_uint(m7.z);
                }
                OffsetStack<Stack> sub{st, sp};
                const TraceResult b = traverse_w<W, ANY, STATS>(sub, nodes, leaves + 3 * (size_t)leaf_base, root,
                                                                n_tris, lo, ld, tmin, r.t);
                if (STATS) {
                    r.visits += b.visits;
                    r.tests += b.tests;
                }
                if (b.leaf != NO_HIT) {
                    r.t = b.t;
                    r.leaf = leaf_base + b.leaf;
                    r.ba = b.ba;
                    r.bb = b.bb;
                    hit_inst = ii;
                    if (ANY) return r;
                }
                continue;
            }
            if (next + 1u >= n_inst) {   // internal index out of range: unreachable
                next = NO_HIT;
                continue;
            }
            break;
        }
        if (STATS) r.visits++;
        next = node4_order_push(st, sp, node4_test(tbase, next << 7, sl, tmin, r.t));
    }
}

// The same walk as traverse_two_level with 4-wide mesh trees, in one loop:
// entering an instance swaps the ray to local space and the node / leaf
// bases to the mesh's, and records the stack depth; when the stack drains
// back to that depth the world ray is reloaded (world(o, d)) and the top
// tree continues.  Mesh entries sit above the top-level ones, so the visit
// order, culling and counts are those of the nested walk; one copy of the
// node test keeps the register count of the single-level kernel.
template <bool ANY, bool STATS, class Stack, class World>
__device__ __forceinline__ TraceResult traverse_two_level4(Stack& st, const float4* __restrict__ top4,
                                                           const float4* __restrict__ top_leaves, uint32_t top_root,
                                                           uint32_t n_inst, const float4* __restrict__ inst,
                                                           const float4* __restrict__ mesh_nodes,
                                                           const float4* __restrict__ leaves, V3 o, V3 d,
                                                           float tmin, float tmax, uint32_t& hit_inst,
                                                           World&& world)
{
    TraceResult r;
    r.leaf = NO_HIT;
    r.t = tmax;
    r.ba = r.bb = 0.0f;
    r.visits = r.tests = 0;
    hit_inst = NO_HIT;
    RaySlab sl = make_slab(o, d);
    int sp = 0;
    int base = -1;                  // stack depth at the current instance's entry (-1: top tree)
    uint32_t cur = 0;               // current instance
    uint32_t n_cur = n_inst;        // leaves of the current tree
    const float4* nodes = top4;
    const float4* lv = top_leaves;
    uint32_t next = top_root;
    bool done = false;
    for (;;) {
        for (;;) {
            if (next == NO_HIT) {
                bool found = false;
                while (!found) {
                    if (sp == base) {   // the instance's tree is done: back to the world ray
                        world(o, d);
                        sl = make_slab(o, d);
                        base = -1;
                        nodes = top4;
                        lv = top_leaves;
                        n_cur = n_inst;
                    }
                    if (sp == 0) break;
                    --sp;
                    uint32_t nn;
                    float et;
                    st.get(sp, nn, et);
                    if (!(et > r.t)) {
                        next = nn;
                        found = true;
                    }
                }
                if (!found) {
                    done = true;
                    break;
                }
            }
            if (next & LEAF_BIT) {
                const uint32_t li = next & ~LEAF_BIT;
                next = NO_HIT;
                if (li >= n_cur) continue;   // unreachable on a well-formed tree
                if (base >= 0) {
                    const float4 a = lv[3 * li + 0];
                    const float4 b = lv[3 * li + 1];
                    const float4 c = lv[3 * li + 2];
                    float t, ba, bb;
                    if (STATS) r.tests++;
                    if (intersect_tri(o, d, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), t, ba, bb) &&
                        t >= tmin && t < r.t) {
                        r.t = t;
                        r.leaf = li;     // mesh-local here; made global at the end
                        r.ba = ba;
                        r.bb = bb;
                        hit_inst = cur;
                        if (ANY) break;
                    }
                    continue;
                }
                const uint32_t ii = __float_as_uint(lv[3 * (size_t)li].w);
                if (ii >= n_inst) continue;
                const float4* rec = inst + INST_F4 * (size_t)ii;
                const float4 w0 = rec[0], w1 = rec[1], w2 = rec[2];
                const float4 m6 = rec[6], m7 = rec[7];
                o = xf_point(w0, w1, w2, o);
                d = xf_dir(w0, w1, w2, d);
                sl = make_slab(o, d);
                base = sp;
                cur = ii;
                n_cur = __float_as_uint(m7.w);
                nodes = mesh_nodes + 8 * (size_t)__float_as_uint(m7.x);
                lv = leaves + 3 * (size_t)__float_as_uint(m6.z);
                next = __float_as_uint(m7.y);
                continue;
            }
            if (next + 1u >= n_cur) {   // internal index out of range: unreachable
                next = NO_HIT;
                continue;
            }
            break;
        }
        if (done || (ANY && r.leaf != NO_HIT)) break;
        if (STATS) r.visits++;
        next = node4_order_push(st, sp, node4_test((const char*)nodes, next << 7, sl, tmin, r.t));
    }
    // any hit: leave with the hit (the leaf is made global below)
    if (r.leaf != NO_HIT) r.leaf += __float_as_uint(inst[INST_F4 * (size_t)hit_inst + 6].z);
    return r;
}

// Either two-level walk, by mesh-tree width (world reloads the world ray)
template <int W, bool ANY, bool STATS, class Stack, class World>
__device__ __forceinline__ TraceResult two_level_w(Stack& st, const float4* __restrict__ top4,
                                                   const float4* __restrict__ top_leaves, uint32_t top_root,
                                                   uint32_t n_inst, const float4* __restrict__ inst,
                                                   const float4* __restrict__ mesh_nodes,
                                                   const float4* __restrict__ leaves, V3 o, V3 d, float tmin,
                                                   float tmax, uint32_t& hit_inst, World&& world)
{
    if constexpr (W == 4)
        return traverse_two_level4<ANY, STATS>(st, top4, top_leaves, top_root, n_inst, inst, mesh_nodes, leaves, o, d,
                                               tmin, tmax, hit_inst, world);
    else
        return traverse_two_level<W, ANY, STATS>(st, top4, top_leaves, top_root, n_inst, inst, mesh_nodes, leaves, o,
                                                 d, tmin, tmax, hit_inst);
}

// Stack entries a two-level walk needs: the top tree's 4-wide need plus the
// deepest mesh tree's need at width w.
inline uint32_t two_level_stack_need(uint32_t top_depth, uint32_t mesh_depth, uint32_t w)
{
    const uint32_t top = 3 * ((top_depth + 1) / 2) + 1;
    const uint32_t mesh = (w == 4) ? 3 * ((mesh_depth + 1) / 2) + 1 : mesh_depth + 1;
    return top + mesh;
}

}  // namespace mrh
