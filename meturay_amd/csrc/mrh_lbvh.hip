// mrh_lbvh.hip — Morton codes and the one-pass LBVH builder.
//
// Build = 5 kernels + the onesweep sort:
//   k_tri_bounds   : triangle AABBs + per-block extents, reduced by
//                    k_extent_reduce (order-preserving integer min / max,
//                    exact and order independent)
//   k_morton       : quantised centroids -> 30- or 63-bit codes
//   radix sort     : (code, primitive) ascending, stable
//   k_lbvh_emit    : one thread per sorted leaf walks up (Apetrei 2014): the
//                    parent of a range is chosen by the reference's rule, the
//                    first child to reach a parent stops, the second
//                    continues, so hierarchy, AABB refit and the traversal
//                    records are written in a single pass with no relaunch
//                    loop and no host round trip.
//   k_bvh_depth_collapse : node depths -> the deepest leaf (bounds the
//                    traversal stack) and the 4-wide traversal nodes.
#include "mrh_bvh.hpp"

#include <cstring>

namespace mrh {
namespace {

__device__ __forceinline__ uint32_t enc_ordered(float f)
{
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ float dec_ordered(uint32_t e)
{
    uint32_t b = (e & 0x80000000u) ? (e & 0x7FFFFFFFu) : ~e;
    return __uint_as_float(b);
}

__device__ __forceinline__ void st_agent(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p)
{
    return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// meta[0..1] = root / depth, [2..4] = the encoded extent min (starts at the
// largest code), [5..7] = the encoded extent max (starts at 0), [8..15] = the
// hierarchy flags and spare words; one launch instead of four fills
__global__ void k_meta_init(uint32_t* __restrict__ meta)
{
    const uint32_t i = threadIdx.x;
    if (i < 16) meta[i] = (i >= 2 && i < 5) ? 0xFFFFFFFFu : 0u;
}

// meta[2..4] = encoded extent min, meta[5..7] = encoded extent max.  A
// block stages a tile of 256 triangles through LDS: the corners come in as
// consecutive words (every load instruction covers 1 KiB of the array) and
// the boxes leave the same way, instead of 36- and 24-byte strided accesses.
constexpr uint32_t BOUNDS_TILE = 256;

__global__ __launch_bounds__(256) void k_tri_bounds(const float* __restrict__ tris, uint32_t n,
                                                    float* __restrict__ aabb, uint32_t* __restrict__ part_out)
{
    __shared__ float s_in[BOUNDS_TILE * 9];
    __shared__ float s_out[BOUNDS_TILE * 6];
    float mn[3] = {INFINITY, INFINITY, INFINITY};
    float mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    const uint32_t t = threadIdx.x;
    for (uint32_t base = blockIdx.x * BOUNDS_TILE; base < n; base += gridDim.x * BOUNDS_TILE) {
        const uint32_t cnt = min(BOUNDS_TILE, n - base);
        const float* src = tris + (size_t)base * 9;
#pragma unroll
        for (uint32_t j = 0; j < 9; j++) {
            const uint32_t k = t + j * BOUNDS_TILE;
            if (k < cnt * 9) s_in[k] = src[k];
        }
        __syncthreads();
        if (t < cnt) {
            const float* c = s_in + 9 * t;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const float lo = fminf(fminf(c[a], c[3 + a]), c[6 + a]);
                const float hi = fmaxf(fmaxf(c[a], c[3 + a]), c[6 + a]);
                s_out[6 * t + a] = lo;
                s_out[6 * t + 3 + a] = hi;
                mn[a] = fminf(mn[a], lo);
                mx[a] = fmaxf(mx[a], hi);
            }
        }
        __syncthreads();
        float* dst = aabb + (size_t)base * 6;
#pragma unroll
        for (uint32_t j = 0; j < 6; j++) {
            const uint32_t k = t + j * BOUNDS_TILE;
            if (k < cnt * 6) dst[k] = s_out[k];
        }
        __syncthreads();
    }
    // wave, then block reduction into this block's six encoded words
    // (reduced by k_extent_reduce: no atomics on one line from every block)
    __shared__ uint32_t part[4][6];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        for (int off = 32; off > 0; off >>= 1) {
            mn[a] = fminf(mn[a], __shfl_xor(mn[a], off, 64));
            mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], off, 64));
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            part[wave][a] = enc_ordered(mn[a]);
            part[wave][3 + a] = enc_ordered(mx[a]);
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const uint32_t k = threadIdx.x;
        uint32_t v = part[0][k];
        for (uint32_t w = 1; w < blockDim.x / 64; w++) v = (k < 3) ? min(v, part[w][k]) : max(v, part[w][k]);
        part_out[6 * (size_t)blockIdx.x + k] = v;
    }
}

// meta[2..4] / meta[5..7] = the encoded extent over the blocks' partials
// (one block; order-preserving integer min / max, exact)
__global__ __launch_bounds__(256) void k_extent_reduce(const uint32_t* __restrict__ part, uint32_t blocks,
                                                       uint32_t* __restrict__ meta)
{
    __shared__ uint32_t s[6][4];
    uint32_t v[6];
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = (k < 3) ? 0xFFFFFFFFu : 0u;
    for (uint32_t b = threadIdx.x; b < blocks; b += blockDim.x) {
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const uint32_t x = part[6 * (size_t)b + k];
            v[k] = (k < 3) ? min(v[k], x) : max(v[k], x);
        }
    }
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        for (int off = 32; off > 0; off >>= 1) {
            const uint32_t y = (uint32_t)__shfl_xor((int)v[k], off, 64);
            v[k] = (k < 3) ? min(v[k], y) : max(v[k], y);
        }
        if (lane == 0) s[k][wave] = v[k];
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const uint32_t k = threadIdx.x;
        uint32_t r = s[k][0];
        for (uint32_t w = 1; w < blockDim.x / 64; w++) r = (k < 3) ? min(r, s[k][w]) : max(r, s[k][w]);
        meta[2 + k] = r;
    }
}

// LinearBVH.hpp:143-165 with IEEE division and the clamp to 2^b - 1;
// delta = float(max span * 2^-b) as FindMortonDelta (LinearBVH.hpp:400-414).
template <class CodeT>
__global__ __launch_bounds__(256) void k_morton(const float* __restrict__ aabb, uint32_t n,
                                                const uint32_t* __restrict__ meta, uint32_t bpd,
                                                CodeT* __restrict__ codes)
{
    float emin[3], emax[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        emin[a] = dec_ordered(meta[2 + a]);
        emax[a] = dec_ordered(meta[5 + a]);
    }
    const double res = 1.0 / (double)(1u << bpd);
    double best = 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float span = emax[a] - emin[a];
        double dd = (double)span * res;
        best = (dd > best) ? dd : best;
    }
    const float delta = (float)best;
    const uint32_t qmax = (1u << bpd) - 1u;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float* b = aabb + (size_t)i * 6;
        uint32_t q[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            float span = b[3 + a] - b[a];
            float c = b[a] + span * 0.5f;
            float rel = c - emin[a];
            uint32_t v = 0;
            if (delta > 0.0f) {
                float f = floorf(rel / delta);
                v = (f >= (float)qmax) ? qmax : (uint32_t)f;
            }
            q[a] = v;
        }
        if (sizeof(CodeT) == 4)
            codes[i] = (CodeT)(expand10(q[0]) | (expand10(q[1]) << 1) | (expand10(q[2]) << 2));
        else
            codes[i] = (CodeT)(expand21(q[0]) | (expand21(q[1]) << 1) | (expand21(q[2]) << 2));
    }
}

// LinearBVH.hpp:30-39 + the (code, index) tie-break for equal codes.
template <class CodeT>
__device__ __forceinline__ uint32_t delta_at(const CodeT* __restrict__ c, uint32_t i)
{
    CodeT x = c[i] ^ c[i + 1];
    if (sizeof(CodeT) == 4)
        return x ? (uint32_t)__clz((uint32_t)x) : 32u + (uint32_t)__clz(i ^ (i + 1u));
    return x ? (uint32_t)__clzll((unsigned long long)x) : 64u + (uint32_t)__clz(i ^ (i + 1u));
}

__device__ __forceinline__ uint32_t child_code(uint32_t node, uint32_t n)
{
    return (node < n) ? (node | LEAF_BIT) : (node - n);
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// 32 coherent bytes: two 16-byte stores / loads that go through to memory
// past this XCD's L2 (cache policy sc1, what agent-scope atomic stores and
// loads of a word use), so the thread that reads a half-record on another
// XCD sees it once the writer's counter increment is visible.  The loads
// wait for their data inside the asm block.
__device__ __forceinline__ void st_half(uint32_t* p, u32x4 a, u32x4 b)
{
    asm volatile("global_store_dwordx4 %0, %1, off sc1\n\t"
                 "global_store_dwordx4 %0, %2, off offset:16 sc1"
                 :
                 : "v"(p), "v"(a), "v"(b)
                 : "memory");
}

__device__ __forceinline__ void ld_half(const uint32_t* p, u32x4& a, u32x4& b)
{
    asm volatile("global_load_dwordx4 %0, %2, off sc1\n\t"
                 "global_load_dwordx4 %1, %2, off offset:16 sc1\n\t"
                 "s_waitcnt vmcnt(0)"
                 : "=&v"(a), "=&v"(b)
                 : "v"(p)
                 : "memory");
}

// One thread per sorted leaf walks up (Apetrei 2014) and the second child
// to reach a parent continues.  A child reaching parent p publishes one
// 32-byte half of p's 64-byte record (left child: words 0-7, right child:
// 8-15): its box, its node index and its range end (l for a left child, r
// for a right one), then increments p's counter.  The second arrival reads
// its sibling's half and has everything p needs: both children, both boxes
// and p's range.
//
// A block owns 1024 consecutive leaves [lo, hi].  A node whose range lies
// inside them is reached only by the block's own threads, so its record and
// counter live in LDS; the others use the global record and counter, with
// stores and loads that go through to memory.  Node p (split between p and
// p + 1) has range [l, r] with l - 1 the last j < p and r the first j > p
// where delta(j) < delta(p) (delta 0 past either end), so p is local when
// such j exist in [lo - 1, p - 1] and in [p + 1, hi]: a prefix and a suffix
// minimum of the block's deltas.  The arrays the view reads (links,
// parents, boxes, binary traversal nodes) are plain stores, read only after
// the kernel.
constexpr uint32_t EMIT_BLOCK = 1024;
constexpr uint32_t EMIT_WAVES = EMIT_BLOCK / 64;

template <class CodeT>
__global__ __launch_bounds__(EMIT_BLOCK) void k_lbvh_emit(
    uint32_t n, const CodeT* __restrict__ scodes, const uint32_t* __restrict__ sidx,
    const float* __restrict__ tris,
    const uint32_t* __restrict__ prim_key, uint32_t* __restrict__ left, uint32_t* __restrict__ right,
    uint32_t* __restrict__ parent, float* __restrict__ node_aabb, uint32_t* halves, uint32_t* counters,
    float4* __restrict__ tnodes, float4* __restrict__ leaves, uint32_t* meta)
{
    __shared__ uint32_t s_delta[EMIT_BLOCK + 1];   // delta(lo - 1 + k)
    __shared__ uint32_t s_suf[EMIT_BLOCK];         // min of s_delta[k .. EMIT_BLOCK]
    __shared__ uint32_t s_wave[2 * EMIT_WAVES];
    __shared__ uint8_t s_local[EMIT_BLOCK];
    __shared__ uint32_t s_cnt[EMIT_BLOCK];
    __shared__ u32x4 s_rec[EMIT_BLOCK * 4];        // [node - lo][4 x 16 B]
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    const uint32_t lo = blockIdx.x * EMIT_BLOCK;
    const uint32_t hi = min(lo + EMIT_BLOCK - 1u, n - 1u);
    auto delta_or0 = [&](uint32_t k) -> uint32_t {   // delta(lo - 1 + k), 0 past the ends
        if (k == 0 && lo == 0) return 0u;
        const uint32_t j = lo - 1u + k;
        return (j + 1u >= n) ? 0u : delta_at(scodes, j);
    };
    s_delta[t] = delta_or0(t);
    if (t == 0) s_delta[EMIT_BLOCK] = delta_or0(EMIT_BLOCK);
    s_cnt[t] = 0u;
    __syncthreads();
    // inclusive prefix minimum of s_delta[0 .. t]
    uint32_t pm = s_delta[t];
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)pm, off, 64);
        if (lane >= off) pm = min(pm, y);
    }
    // suffix minimum of s_delta[t .. EMIT_BLOCK - 1]
    uint32_t sm = s_delta[t];
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_down((int)sm, off, 64);
        if (lane + off < 64) sm = min(sm, y);
    }
    if (lane == 63) s_wave[wave] = pm;
    if (lane == 0) s_wave[EMIT_WAVES + wave] = sm;
    __syncthreads();
    for (uint32_t w = 0; w < wave; w++) pm = min(pm, s_wave[w]);
    for (uint32_t w = wave + 1; w < EMIT_WAVES; w++) sm = min(sm, s_wave[EMIT_WAVES + w]);
    s_suf[t] = min(sm, s_delta[EMIT_BLOCK]);
    __syncthreads();
    {
        // node p = lo + t: local when p < hi and both nearest smaller deltas are in the block
        const uint32_t dp = s_delta[t + 1u];
        const uint32_t right_min = (t + 2u < EMIT_BLOCK) ? s_suf[t + 2u]
                                 : (t + 2u == EMIT_BLOCK ? s_delta[EMIT_BLOCK] : 0xFFFFFFFFu);
        s_local[t] = (lo + t < hi && pm < dp && right_min < dp) ? 1 : 0;
    }
    __syncthreads();
    const uint32_t i = lo + t;
    if (i >= n) return;
    const uint32_t prim = sidx[i];
    // leaf record, and the leaf's box from the same corners (the values
    // k_tri_bounds computed: one gather instead of two)
    float box[6];
    {
        const float* tr = tris + (size_t)prim * 9;
        float c[9];
#pragma unroll
        for (int k = 0; k < 9; k++) c[k] = tr[k];
        const uint32_t key = prim_key ? prim_key[prim] : 0u;
        leaves[3 * (size_t)i + 0] = make_float4(c[0], c[1], c[2], __uint_as_float(prim));
        leaves[3 * (size_t)i + 1] = make_float4(c[3], c[4], c[5], __uint_as_float(key));
        leaves[3 * (size_t)i + 2] = make_float4(c[6], c[7], c[8], 0.0f);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            box[a] = fminf(fminf(c[a], c[3 + a]), c[6 + a]);
            box[3 + a] = fmaxf(fmaxf(c[a], c[3 + a]), c[6 + a]);
        }
    }
#pragma unroll
    for (int a = 0; a < 6; a++) node_aabb[(size_t)i * 6 + a] = box[a];

    uint32_t node = i, l = i, r = i;
    // a walk is at most tree-depth steps (< 100 for 30/63-bit codes with
    // the index tie-break); the bound only matters for a corrupt hierarchy
    for (uint32_t step = 0;; step++) {
        if (step == 256u) {
            atomicOr(&meta[8], 4u);
            break;
        }
        if (l == 0 && r == n - 1) {
            parent[node] = 0xFFFFFFFFu;
            meta[0] = node;
            break;
        }
        // ChooseParentAndUpdate (LinearBVH.hpp:43-100)
        const uint32_t ld = (l == 0) ? 0u : (l >= lo ? s_delta[l - lo] : delta_at(scodes, l - 1));
        const uint32_t rd = (r == n - 1) ? 0u : (r <= hi ? s_delta[r - lo + 1u] : delta_at(scodes, r));
        const bool is_left = rd > ld;
        const uint32_t p = is_left ? r : l - 1;
        parent[node] = n + p;
        const u32x4 ha_out = u32x4{__float_as_uint(box[0]), __float_as_uint(box[1]), __float_as_uint(box[2]),
                                   __float_as_uint(box[3])};
        const u32x4 hb_out = u32x4{__float_as_uint(box[4]), __float_as_uint(box[5]), node, is_left ? l : r};
        u32x4 ha, hb;
        const bool local = p >= lo && p < hi && s_local[p - lo];
        if (local) {
            // LDS handshake: the half is written before the counter
            // increment, and LDS serves one wave's requests in order, so a
            // sibling that sees the increment reads the half
            u32x4* rec = s_rec + 4 * (p - lo);
            rec[is_left ? 0 : 2] = ha_out;
            rec[is_left ? 1 : 3] = hb_out;
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0)
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            const uint32_t old = __hip_atomic_fetch_add(&s_cnt[p - lo], 1u, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_WORKGROUP);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            if (old == 0u) break;   // sibling not done yet: it will continue
            ha = rec[is_left ? 2 : 0];
            hb = rec[is_left ? 3 : 1];
        } else {
            uint32_t* rec = halves + 16 * (size_t)p;
            st_half(rec + (is_left ? 0 : 8), ha_out, hb_out);
            // The half-record went through to memory (sc1) and the counter
            // is an agent-scope atomic, so the handshake only needs the
            // stores to have completed before the increment.  A
            // __threadfence() would also write back the whole L2 twice per
            // step; an agent-scope release/acquire pair costs an L2
            // write-back per step on this chip.  The signal fences keep the
            // compiler from moving memory operations across the wait
            // (s_waitcnt alone is no compiler barrier); the wait keeps the
            // hardware from doing so.
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __builtin_amdgcn_s_waitcnt(0);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            const uint32_t old = __hip_atomic_fetch_add(&counters[p], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            if (old == 0u) break;   // sibling not done yet: it will continue
            ld_half(rec + (is_left ? 8 : 0), ha, hb);
        }
        // second arrival: the sibling's half holds its box, node and end
        const uint32_t sib = hb.z;
        if (sib >= 2 * n - 1) {   // never taken on a correct build; reported to the host
            atomicOr(&meta[8], 1u);
            break;
        }
        const float sbox[6] = {__uint_as_float(ha.x), __uint_as_float(ha.y), __uint_as_float(ha.z),
                               __uint_as_float(ha.w), __uint_as_float(hb.x), __uint_as_float(hb.y)};
        if (is_left) r = hb.w;
        else l = hb.w;
        if (l > r || r >= n) {
            atomicOr(&meta[8], 2u);
            break;
        }
        const uint32_t lc = is_left ? node : sib;
        const uint32_t rc = is_left ? sib : node;
        left[p] = lc;
        right[p] = rc;
        const float* lb = is_left ? box : sbox;
        const float* rb = is_left ? sbox : box;
        float4* tn = tnodes + 4 * (size_t)p;
        tn[0] = make_float4(lb[0], lb[1], lb[2], lb[3]);
        tn[1] = make_float4(lb[4], lb[5], rb[0], rb[1]);
        tn[2] = make_float4(rb[2], rb[3], rb[4], rb[5]);
        tn[3] = make_float4(__uint_as_float(child_code(lc, n)), __uint_as_float(child_code(rc, n)), 0.0f, 0.0f);
        float ub[6];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            ub[a] = fminf(lb[a], rb[a]);
            ub[3 + a] = fmaxf(lb[3 + a], rb[3 + a]);
        }
        node = n + p;
#pragma unroll
        for (int a = 0; a < 6; a++) {
            box[a] = ub[a];
            node_aabb[(size_t)node * 6 + a] = ub[a];
        }
    }
}

// Depth of every internal node, then both uses of it: the deepest leaf (a
// child of a node at depth d is at d + 1; one atomic per block when it
// raises meta[1]) bounds the traversal stack, and every node at even depth
// (root = 0) becomes a 4-wide traversal node gathering its children, an
// internal child replaced by that child's two children, from the binary
// nodes' child boxes (layout in traverse4).
//
// A block takes 256 consecutive internal nodes.  Those whose parent is
// among them form a forest in LDS; pointer jumping (8 rounds) gives every
// node its forest root and the distance to it, and only the forest roots
// walk the parent links to the tree's root.
constexpr uint32_t DEPTH_BLOCK = 256;

__global__ __launch_bounds__(DEPTH_BLOCK) void k_bvh_depth_collapse(uint32_t n, const float4* __restrict__ tnodes,
                                                                    const uint32_t* __restrict__ parent,
                                                                    float4* __restrict__ nodes4,
                                                                    uint32_t* __restrict__ meta)
{
    constexpr uint32_t NONE = 0xFFFFFFFFu;
    __shared__ uint32_t s_jump[DEPTH_BLOCK], s_dist[DEPTH_BLOCK], s_done[DEPTH_BLOCK], s_root_depth[DEPTH_BLOCK];
    const uint32_t t = threadIdx.x;
    const uint32_t lo = blockIdx.x * DEPTH_BLOCK;
    const uint32_t p = lo + t;
    const bool valid = p + 1 < n;
    uint32_t up = NONE;   // parent's slot in this block
    if (valid) {
        const uint32_t par = parent[n + p];
        if (par != NONE && par - n - lo < DEPTH_BLOCK) up = par - n - lo;
    }
    uint32_t jump = (up != NONE) ? up : t, dist = (up != NONE) ? 1u : 0u, done = (up == NONE) ? 1u : 0u;
    s_jump[t] = jump;
    s_dist[t] = dist;
    s_done[t] = done;
    __syncthreads();
    for (int round = 0; round < 8; round++) {
        if (!done) {
            dist += s_dist[jump];
            done = s_done[jump];
            jump = s_jump[jump];
        }
        __syncthreads();
        s_jump[t] = jump;
        s_dist[t] = dist;
        s_done[t] = done;
        __syncthreads();
    }
    // forest roots: depth by walking the parent links
    if (valid && up == NONE) {
        uint32_t d = 0, x = n + p;
        while (parent[x] != NONE && d < 4096u) {
            x = parent[x];
            d++;
        }
        s_root_depth[t] = d;
    }
    __syncthreads();
    uint32_t best = 0;
    if (valid) {
        const uint32_t d = dist + s_root_depth[jump];
        const float4* tn = tnodes + 4 * (size_t)p;
        const float4 a3 = tn[3];
        const uint32_t code[2] = {__float_as_uint(a3.x), __float_as_uint(a3.y)};
        if ((code[0] | code[1]) & LEAF_BIT) best = d + 1u;
        if (!(d & 1u)) {
            const float4 a0 = tn[0], a1 = tn[1], a2 = tn[2];
            float* q = (float*)(nodes4 + 8 * (size_t)p);
            uint32_t k = 0;
            const float box[2][6] = {{a0.x, a0.y, a0.z, a0.w, a1.x, a1.y}, {a1.z, a1.w, a2.x, a2.y, a2.z, a2.w}};
            for (int c = 0; c < 2; c++) {
                if (code[c] & LEAF_BIT) {
                    for (int a = 0; a < 6; a++) q[a * 4 + k] = box[c][a];
                    q[24 + k] = __uint_as_float(code[c]);
                    k++;
                } else {
                    const float4* tq = tnodes + 4 * (size_t)code[c];
                    const float4 b0 = tq[0], b1 = tq[1], b2 = tq[2], b3 = tq[3];
                    const float gb[2][6] = {{b0.x, b0.y, b0.z, b0.w, b1.x, b1.y},
                                            {b1.z, b1.w, b2.x, b2.y, b2.z, b2.w}};
                    const uint32_t gc[2] = {__float_as_uint(b3.x), __float_as_uint(b3.y)};
                    for (int g = 0; g < 2; g++) {
                        for (int a = 0; a < 6; a++) q[a * 4 + k] = gb[g][a];
                        q[24 + k] = __uint_as_float(gc[g]);
                        k++;
                    }
                }
            }
            for (; k < 4; k++) {
                for (int a = 0; a < 3; a++) {
                    q[a * 4 + k] = INFINITY;
                    q[(a + 3) * 4 + k] = -INFINITY;
                }
                q[24 + k] = __uint_as_float(0xFFFFFFFFu);
            }
            for (int j = 28; j < 32; j++) q[j] = 0.0f;
        }
    }
    __shared__ uint32_t part[DEPTH_BLOCK / 64];
    for (int off = 32; off > 0; off >>= 1) best = max(best, (uint32_t)__shfl_xor((int)best, off, 64));
    if ((t & 63) == 0) part[t >> 6] = best;
    __syncthreads();
    if (t == 0) {
        for (uint32_t w = 1; w < DEPTH_BLOCK / 64; w++) best = max(best, part[w]);
        if (best > ld_agent(&meta[1])) atomicMax(&meta[1], best);
    }
}

template <class T>
int32_t alloc_into(mrh_lbvh* b, T** p, size_t count)
{
    void* q = nullptr;
    MRH_TRY(dmalloc(b->ctx, &q, count * sizeof(T)));
    b->allocs.push_back(q);
    *p = (T*)q;
    return MRH_OK;
}

}  // namespace

// Morton codes of n triangles into `codes` (extent + AABBs in scratch).
int32_t morton_codes(mrh_context* ctx, const float* d_tris, uint32_t n, uint32_t bits, void* codes,
                     float* aabb_out, uint32_t* meta_out)
{
    if (n == 0) return MRH_OK;
    float* aabb = aabb_out;
    uint32_t* meta = meta_out;
    if (!aabb) {
        void* p = nullptr;
        MRH_TRY(scratch(ctx, SLOT_MORTON_AABB, (size_t)n * 24, &p));
        aabb = (float*)p;
    }
    if (!meta) {
        void* p = nullptr;
        MRH_TRY(scratch(ctx, SLOT_MORTON_EXT, 64, &p));
        meta = (uint32_t*)p;
    }
    const uint32_t bgrid = stride_grid(ctx, n, 256);
    void* part = nullptr;
    MRH_TRY(scratch(ctx, SLOT_BOUNDS_PART, (size_t)bgrid * 24, &part));
    MRH_TRY(launch(ctx, "lbvh_meta_init", dim3(1), dim3(64), 0, k_meta_init, meta));
    MRH_TRY(launch(ctx, "tri_bounds", dim3(bgrid), dim3(256), 0, k_tri_bounds, d_tris, n, aabb, (uint32_t*)part));
    MRH_TRY(launch(ctx, "extent_reduce", dim3(1), dim3(256), 0, k_extent_reduce, (const uint32_t*)part, bgrid, meta));
    if (bits == 30)
        return launch(ctx, "morton", dim3(stride_grid(ctx, n, 256)), dim3(256), 0, k_morton<uint32_t>,
                      (const float*)aabb, n, (const uint32_t*)meta, 10u, (uint32_t*)codes);
    return launch(ctx, "morton", dim3(stride_grid(ctx, n, 256)), dim3(256), 0, k_morton<uint64_t>,
                  (const float*)aabb, n, (const uint32_t*)meta, 21u, (uint64_t*)codes);
}

// The triangles are read during the build only (bounds, codes, leaf
// records), so they are not copied.
int32_t lbvh_build_into(mrh_lbvh* b, const float* d_tris)
{
    mrh_context* ctx = b->ctx;
    const uint32_t n = b->n;
    if (!d_tris) return fail(ctx, MRH_ERR_INVALID, "lbvh build: null triangles");
    MRH_TRY(morton_codes(ctx, d_tris, n, b->bits, b->codes, b->leaf_aabb, b->meta));
    SortArgs sa;
    sa.keys_in = b->codes;
    sa.keys_out = b->sorted_codes;
    sa.vals_out = b->sorted_idx;
    sa.n_max = n;
    sa.key_bytes = (b->bits == 30) ? 4 : 8;
    sa.begin_bit = 0;
    sa.end_bit = b->bits;
    MRH_TRY(radix_sort(ctx, sa));
    if (n > 1) MRH_TRY(fill_u32(ctx, b->counters, 0u, n - 1));
    const uint32_t egrid = div_up(n, EMIT_BLOCK);   // one thread per leaf
    if (b->bits == 30)
        MRH_TRY(launch(ctx, "lbvh_emit", dim3(egrid), dim3(EMIT_BLOCK), 0, k_lbvh_emit<uint32_t>, n,
                       (const uint32_t*)b->sorted_codes, (const uint32_t*)b->sorted_idx,
                       d_tris,
                       (const uint32_t*)b->prim_key, b->left, b->right, b->parent, b->node_aabb,
                       b->halves, b->counters, b->tnodes, b->leaves, b->meta));
    else
        MRH_TRY(launch(ctx, "lbvh_emit", dim3(egrid), dim3(EMIT_BLOCK), 0, k_lbvh_emit<uint64_t>, n,
                       (const uint64_t*)b->sorted_codes, (const uint32_t*)b->sorted_idx,
                       d_tris,
                       (const uint32_t*)b->prim_key, b->left, b->right, b->parent, b->node_aabb,
                       b->halves, b->counters, b->tnodes, b->leaves, b->meta));
    if (n > 1)
        MRH_TRY(launch(ctx, "bvh_depth_collapse", dim3(div_up(n - 1, DEPTH_BLOCK)), dim3(DEPTH_BLOCK), 0,
                       k_bvh_depth_collapse, n,
                       (const float4*)b->tnodes, (const uint32_t*)b->parent, b->nodes4, b->meta));
    uint32_t host_meta[9] = {};
    MRH_HIP_CHECK(ctx, hipMemcpyAsync(host_meta, b->meta, sizeof(host_meta), hipMemcpyDeviceToHost,
                                      ctx->stream));
    MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    b->root = host_meta[0];
    b->max_depth = host_meta[1];
    for (int a = 0; a < 3; a++) {
        const uint32_t e0 = host_meta[2 + a], e1 = host_meta[5 + a];
        uint32_t u0 = (e0 & 0x80000000u) ? (e0 & 0x7FFFFFFFu) : ~e0;
        uint32_t u1 = (e1 & 0x80000000u) ? (e1 & 0x7FFFFFFFu) : ~e1;
        std::memcpy(&b->lo[a], &u0, 4);
        std::memcpy(&b->hi[a], &u1, 4);
    }
    if (host_meta[8])
        return fail(ctx, MRH_ERR_HIP, "lbvh build: inconsistent hierarchy (flags 0x%x)", host_meta[8]);
    if (b->root >= 2 * n - 1)
        return fail(ctx, MRH_ERR_HIP, "lbvh build produced no root (root = %u, n = %u)", b->root, n);
    b->root4 = root_code(b);
    return MRH_OK;
}

int32_t lbvh_create(mrh_context* ctx, const float* d_tris, uint32_t n, uint32_t bits,
                    const uint32_t* d_prim_key, mrh_lbvh** out)
{
    *out = nullptr;
    if (n == 0) return fail(ctx, MRH_ERR_INVALID, "mrh_lbvh_build: n = 0");
    if (bits != 30 && bits != 63) return fail(ctx, MRH_ERR_INVALID, "mrh_lbvh_build: bits must be 30 or 63");
    if (n >= (1u << 30)) return fail(ctx, MRH_ERR_INVALID, "mrh_lbvh_build: n too large");
    mrh_lbvh* b = new mrh_lbvh();
    b->ctx = ctx;
    b->n = n;
    b->bits = bits;
    const size_t ni = (n > 1) ? n - 1 : 1;
    const size_t cb = (bits == 30) ? 4 : 8;
    int32_t rc = MRH_OK;
    do {
        if ((rc = alloc_into(b, &b->leaf_aabb, (size_t)n * 6)) != MRH_OK) break;
        if ((rc = alloc_into(b, (char**)&b->codes, (size_t)n * cb)) != MRH_OK) break;
        if ((rc = alloc_into(b, (char**)&b->sorted_codes, (size_t)n * cb)) != MRH_OK) break;
        if ((rc = alloc_into(b, &b->sorted_idx, n)) != MRH_OK) break;
        if ((rc = alloc_into(b, &b->left, ni)) != MRH_OK) break;
        if ((rc = alloc_into(b, &b->right, ni)) != MRH_OK) break;
        if ((rc = alloc_into(b, &b->parent, 2 * (size_t)n - 1)) != MRH_OK) break;
        if ((rc = alloc_into(b, &b->node_aabb, (2 * (size_t)n - 1) * 6)) != MRH_OK) break;
        if ((rc = alloc_into(b, &b->halves, 16 * ni)) != MRH_OK) break;
        if ((rc = alloc_into(b, &b->counters, ni)) != MRH_OK) break;
        if ((rc = alloc_into(b, &b->tnodes, 4 * ni)) != MRH_OK) break;
        if ((rc = alloc_into(b, &b->nodes4, 8 * ni)) != MRH_OK) break;
        if ((rc = alloc_into(b, &b->leaves, 3 * (size_t)n)) != MRH_OK) break;
        if ((rc = alloc_into(b, &b->meta, 16)) != MRH_OK) break;
        if (d_prim_key) {
            if ((rc = alloc_into(b, &b->prim_key, n)) != MRH_OK) break;
            if (hipMemcpyAsync(b->prim_key, d_prim_key, (size_t)n * 4, hipMemcpyDeviceToDevice,
                               ctx->stream) != hipSuccess) {
                rc = fail(ctx, MRH_ERR_HIP, "prim key copy failed");
                break;
            }
        }
        rc = lbvh_build_into(b, d_tris);
    } while (0);
    if (rc != MRH_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        for (void* p : b->allocs) (void)hipFree(p);
        delete b;
        return rc;
    }
    *out = b;
    return MRH_OK;
}

void lbvh_free(mrh_lbvh* b)
{
    if (!b) return;
    if (b->ctx) (void)hipStreamSynchronize(b->ctx->stream);
    for (void* p : b->allocs) (void)hipFree(p);
    delete b;
}

}  // namespace mrh

using namespace mrh;

extern "C" {

int32_t mrh_morton_codes(mrh_context* ctx, const float* d_tris, uint32_t n, uint32_t bits, void* d_codes)
{
    if (!ctx) return MRH_ERR_INVALID;
    if (bits != 30 && bits != 63) return fail(ctx, MRH_ERR_INVALID, "mrh_morton_codes: bits must be 30 or 63");
    if (n && (!d_tris || !d_codes)) return fail(ctx, MRH_ERR_INVALID, "mrh_morton_codes: null pointer");
    return morton_codes(ctx, d_tris, n, bits, d_codes, nullptr, nullptr);
}

int32_t mrh_lbvh_build(mrh_context* ctx, const float* d_tris, uint32_t n, uint32_t bits, mrh_lbvh** out)
{
    if (!ctx || !out) return MRH_ERR_INVALID;
    if (!d_tris) return fail(ctx, MRH_ERR_INVALID, "mrh_lbvh_build: null triangles");
    return lbvh_create(ctx, d_tris, n, bits, nullptr, out);
}

int32_t mrh_lbvh_rebuild(mrh_context* ctx, mrh_lbvh* bvh, const float* d_tris)
{
    if (!ctx || !bvh || bvh->ctx != ctx) return MRH_ERR_INVALID;
    return lbvh_build_into(bvh, d_tris);
}

int32_t mrh_lbvh_view_get(const mrh_lbvh* b, mrh_lbvh_view* v)
{
    if (!b || !v) return MRH_ERR_INVALID;
    v->n = b->n;
    v->bits = b->bits;
    v->root = b->root;
    v->max_depth = b->max_depth;
    v->d_sorted_codes = b->sorted_codes;
    v->d_sorted_idx = b->sorted_idx;
    v->d_left = b->left;
    v->d_right = b->right;
    v->d_parent = b->parent;
    v->d_node_aabb = b->node_aabb;
    v->d_nodes4 = (const float*)b->nodes4;
    v->root4 = b->root4;
    v->reserved = 0;
    return MRH_OK;
}

int32_t mrh_lbvh_destroy(mrh_lbvh* bvh)
{
    if (!bvh) return MRH_ERR_INVALID;
    lbvh_free(bvh);
    return MRH_OK;
}

}  // extern "C"
