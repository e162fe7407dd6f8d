// mrh_guide.hip — the path-guiding SVO: build from surface-point voxel
// codes, radiance splat, filter, lookup and product sampling.
//
// Build (replaces AnisoSVOctreeCPU::Constrcut's sort / RLE / per-level
// passes, AnisoSVO.cu:657-1188, with the structure config 4 asks for):
//   sort the leaf codes (onesweep), then for each level from the leaves up:
//   unique (flag + scan + scatter) and parent links; then child masks and
//   first-child indices with atomics.  Levels are stored root first.
// Splat (KCAccumulateRadianceToLeaf, AnisoSVO.cu:225-263): atomic add of a
//   radiance sample into one of the leaf's 8x8 direction bins + a count.
// Filter (NormalizeAndFilterRadiance, AnisoSVO.cu:1415-1483): sums and counts
//   are pushed up the tree, bins become means, bins without samples take their
//   parent's value top-down, then a 3x3 box blur over the bin grid and a
//   1e-4 floor.
#include "mrh_guide.hpp"
#include "mrh_bvh.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

struct mrh_svo {
    mrh_context* ctx = nullptr;
    uint32_t depth = 0, nodes = 0, cap = 0;
    uint32_t lvl_count[11] = {}, lvl_off[12] = {};
    float cube[4] = {0, 0, 0, 1};   // lo xyz, size
    uint32_t* codes = nullptr;       // [cap] level code of each node
    uint32_t* first_child = nullptr; // [cap]
    uint32_t* child_mask = nullptr;  // [cap]
    uint32_t* parent = nullptr;      // [cap]
    float* sum = nullptr;            // [cap][64] sums, then means (filter)
    unsigned long long* acc = nullptr; // [cap][64] fixed-point splat sums
    uint32_t* cnt = nullptr;         // [cap][64]
    float* radiance = nullptr;       // [cap][64]
    uint32_t* node_count = nullptr;  // [cap] samples in the subtree
    uint4* pack = nullptr;           // [cap] first child, child mask, subtree samples (lookup records)
    // grow-only build scratch: sort/flag/scan buffers, per-level arrays, counter
    void* work_a = nullptr;
    size_t work_a_bytes = 0;
    void* work_b = nullptr;
    size_t work_b_bytes = 0;
    uint32_t* counter = nullptr;     // [16]
    // grow-only workspace of mrh_svo_bin_rays
    void* bin_work = nullptr;
    size_t bin_work_bytes = 0;
};

namespace mrh {
namespace {

// ---------------------------------------------------------------- scan
constexpr uint32_t SCAN_ITEMS = 4, SCAN_BLOCK = 256, SCAN_TILE = SCAN_ITEMS * SCAN_BLOCK;

__global__ __launch_bounds__(256) void k_scan_tiles(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                    uint32_t* __restrict__ tile_sums, uint32_t n)
{
    __shared__ uint32_t s_tmp[8];
    const uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS], t = 0;
#pragma unroll
    for (uint32_t k = 0; k < SCAN_ITEMS; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        t += v[k];
    }
    // exclusive scan of the per-thread totals over the block
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = t;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(incl, off, 64);
        if (lane >= (uint32_t)off) incl += y;
    }
    if (lane == 63) s_tmp[wave] = incl;
    __syncthreads();
    uint32_t wp = 0;
    for (uint32_t w = 0; w < wave; w++) wp += s_tmp[w];
    uint32_t run = incl - t + wp;
#pragma unroll
    for (uint32_t k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == SCAN_BLOCK - 1) tile_sums[blockIdx.x] = run;
}

__global__ __launch_bounds__(256) void k_scan_add(uint32_t* __restrict__ out, const uint32_t* __restrict__ tile_pre,
                                                  uint32_t n)
{
    const uint32_t add = tile_pre[blockIdx.x];
    const uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
#pragma unroll
    for (uint32_t k = 0; k < SCAN_ITEMS; k++)
        if (base + k < n) out[base + k] += add;
}

// words of scratch scan_exclusive needs for n elements
size_t scan_tmp_words(uint32_t n)
{
    size_t w = 0;
    while (n > 1) {
        const uint32_t tiles = div_up(n, SCAN_TILE);
        w += 2 * ((size_t)tiles + 1);
        n = tiles;
    }
    return w + 2;
}

// exclusive scan in -> out (out may alias in); tmp: scan_tmp_words(n) words
int32_t scan_exclusive(mrh_context* ctx, const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tmp)
{
    if (n == 0) return MRH_OK;
    const uint32_t tiles = div_up(n, SCAN_TILE);
    uint32_t* sums = tmp;
    uint32_t* pre = sums + tiles + 1;
    MRH_TRY(launch(ctx, "scan_tiles", dim3(tiles), dim3(SCAN_BLOCK), 0, k_scan_tiles, in, out, sums, n));
    if (tiles > 1) {
        MRH_TRY(scan_exclusive(ctx, sums, pre, tiles, pre + tiles + 1));
        MRH_TRY(launch(ctx, "scan_add", dim3(tiles), dim3(SCAN_BLOCK), 0, k_scan_add, out, (const uint32_t*)pre, n));
    }
    return MRH_OK;
}

// ---------------------------------------------------------------- unique
// flags[i] = 1 where (in[i] >> shift) starts a new run
__global__ void k_flag_runs(const uint32_t* __restrict__ in, uint32_t n, uint32_t shift, uint32_t* __restrict__ flags)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        flags[i] = (i == 0 || (in[i] >> shift) != (in[i - 1] >> shift)) ? 1u : 0u;
}

// out[pos[i]] = in[i] >> shift where flags[i]; seg[i] = pos[i] + flags[i] - 1
__global__ void k_compact_runs(const uint32_t* __restrict__ in, const uint32_t* __restrict__ flags,
                               const uint32_t* __restrict__ pos, uint32_t n, uint32_t shift,
                               uint32_t* __restrict__ out, uint32_t* __restrict__ seg)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (flags[i]) out[pos[i]] = in[i] >> shift;
        if (seg) seg[i] = pos[i] + flags[i] - 1u;
    }
}

// distinct values of in >> shift (in sorted) into out; seg[i] = index of
// in[i]'s run; no host round trip
int32_t unique_runs(mrh_context* ctx, const uint32_t* in, uint32_t n, uint32_t shift, uint32_t* out,
                    uint32_t* seg, uint32_t* flags, uint32_t* pos, uint32_t* tmp)
{
    if (n == 0) return MRH_OK;
    const uint32_t g = stride_grid(ctx, n, 256);
    MRH_TRY(launch(ctx, "svo_flag", dim3(g), dim3(256), 0, k_flag_runs, in, n, shift, flags));
    MRH_TRY(scan_exclusive(ctx, flags, pos, n, tmp));
    return launch(ctx, "svo_compact", dim3(g), dim3(256), 0, k_compact_runs, in, (const uint32_t*)flags,
                  (const uint32_t*)pos, n, shift, out, seg);
}

// Histogram of the coarsest level at which each sorted code starts a new
// prefix (level l's node count = sum of bins 0..l); codes < 2^(3 depth).
__global__ __launch_bounds__(256) void k_level_hist(const uint32_t* __restrict__ sorted, uint32_t n, uint32_t depth,
                                                    uint32_t* __restrict__ hist)
{
    __shared__ uint32_t h[12];
    if (threadIdx.x < 12) h[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t m = 0;
        if (i > 0) {
            const uint32_t x = sorted[i] ^ sorted[i - 1];
            m = x ? depth - (31u - (uint32_t)__clz(x)) / 3u : 12u;
        }
        if (m <= depth) atomicAdd(&h[m], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 12 && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

// Appends the kept records (code != SVO_NO_NODE) at *ctr.  Each block
// takes chunks of 4096 records, ranks the kept ones with a block scan and
// claims their output range with one atomic per chunk.  The order of the
// output depends on scheduling; everything built from it (sorted cells,
// fixed-point sums) does not.
constexpr uint32_t APPEND_ITEMS = 16, APPEND_CHUNK = APPEND_ITEMS * 256;

__global__ __launch_bounds__(256) void k_append_records(const uint32_t* __restrict__ codes,
                                                        const uint32_t* __restrict__ bins,
                                                        const float* __restrict__ rad, uint32_t n,
                                                        uint32_t* __restrict__ oc, uint32_t* __restrict__ ob,
                                                        float* __restrict__ orad, uint32_t* ctr)
{
    __shared__ uint32_t s_wave[4];
    __shared__ uint32_t s_base;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t c0 = blockIdx.x * APPEND_CHUNK; c0 < n; c0 += gridDim.x * APPEND_CHUNK) {
        uint32_t code[APPEND_ITEMS];
        uint32_t t = 0;
#pragma unroll
        for (uint32_t k = 0; k < APPEND_ITEMS; k++) {
            const uint32_t i = c0 + k * 256 + threadIdx.x;
            code[k] = (i < n) ? codes[i] : SVO_NO_NODE;
            t += (code[k] != SVO_NO_NODE) ? 1u : 0u;
        }
        // exclusive scan of the per-thread counts over the block
        uint32_t incl = t;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off, 64);
            if (lane >= (uint32_t)off) incl += y;
        }
        if (lane == 63) s_wave[wave] = incl;
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (uint32_t w = 0; w < 4; w++) {
            before += (w < wave) ? s_wave[w] : 0u;
            total += s_wave[w];
        }
        if (threadIdx.x == 0) s_base = total ? atomicAdd(ctr, total) : 0u;
        __syncthreads();
        uint32_t k_out = s_base + before + incl - t;
#pragma unroll
        for (uint32_t k = 0; k < APPEND_ITEMS; k++) {
            if (code[k] == SVO_NO_NODE) continue;
            const uint32_t i = c0 + k * 256 + threadIdx.x;
            oc[k_out] = code[k];
            ob[k_out] = bins[i];
            orad[k_out] = rad[i];
            k_out++;
        }
        __syncthreads();   // s_wave / s_base are reused by the next chunk
    }
}

// ---------------------------------------------------------------- build
__global__ void k_svo_codes(const float* __restrict__ pts, uint32_t n, uint32_t pitch, float lox, float loy,
                            float loz, float inv_cell, uint32_t depth, uint32_t* __restrict__ codes)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        codes[i] = svo_code(v3(pts[i], pts[(size_t)pitch + i], pts[2 * (size_t)pitch + i]), lox, loy, loz,
                            inv_cell, depth);
}

// node records of one level: code, parent (global), child links into parent
__global__ void k_svo_link(const uint32_t* __restrict__ lvl_codes, const uint32_t* __restrict__ parent_local,
                           uint32_t count, uint32_t off, uint32_t parent_off, uint32_t* __restrict__ codes,
                           uint32_t* __restrict__ parent, uint32_t* __restrict__ first_child,
                           uint32_t* __restrict__ child_mask)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t g = off + i;
        const uint32_t c = lvl_codes[i];
        codes[g] = c;
        first_child[g] = SVO_NO_NODE;
        child_mask[g] = 0u;
        if (parent_local) {
            const uint32_t p = parent_off + parent_local[i];
            parent[g] = p;
            atomicOr(&child_mask[p], 1u << (c & 7u));
            atomicMin(&first_child[p], g);
        } else {
            parent[g] = SVO_NO_NODE;
        }
    }
}

// lookup records: first child, child mask, samples in the subtree
__global__ void k_svo_pack(uint32_t nodes, const uint32_t* __restrict__ first_child,
                           const uint32_t* __restrict__ child_mask, const uint32_t* __restrict__ node_count,
                           uint4* __restrict__ pack)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nodes; i += gridDim.x * blockDim.x)
        pack[i] = make_uint4(first_child[i], child_mask[i], node_count ? node_count[i] : 0u, 0u);
}

// ---------------------------------------------------------------- splat / filter
__global__ void k_svo_splat(SvoView s, const uint32_t* __restrict__ codes, const uint32_t* __restrict__ bins,
                            const float* __restrict__ values, uint32_t n, unsigned long long* __restrict__ acc,
                            uint32_t* __restrict__ cnt)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t c = codes[i];
        if (c == SVO_NO_NODE) continue;
        const uint32_t node = svo_lookup(s, c);
        const uint32_t b = bins[i] & (GUIDE_BINS - 1u);
        atomicAdd(&acc[(size_t)node * GUIDE_BINS + b], (unsigned long long)guide_fix(values[i]));
        atomicAdd(&cnt[(size_t)node * GUIDE_BINS + b], 1u);
    }
}

__global__ void k_svo_fix(size_t count, const unsigned long long* __restrict__ acc, float* __restrict__ sum)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        sum[i] = (float)acc[i] * (1.0f / GUIDE_FIX_ONE);
}

// interior sums = sums of children (one level, children already final)
__global__ void k_svo_up(uint32_t off, uint32_t count, const uint32_t* __restrict__ first_child,
                         const uint32_t* __restrict__ child_mask, float* __restrict__ sum, uint32_t* __restrict__ cnt)
{
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < count * GUIDE_BINS; t += gridDim.x * blockDim.x) {
        const uint32_t node = off + t / GUIDE_BINS, b = t % GUIDE_BINS;
        const uint32_t m = child_mask[node];
        if (!m) continue;
        const uint32_t fc = first_child[node];
        const uint32_t k = (uint32_t)__popc(m);
        float s = sum[(size_t)node * GUIDE_BINS + b];   // samples splatted here directly, if any
        uint32_t c = cnt[(size_t)node * GUIDE_BINS + b];
        for (uint32_t j = 0; j < k; j++) {
            s += sum[(size_t)(fc + j) * GUIDE_BINS + b];
            c += cnt[(size_t)(fc + j) * GUIDE_BINS + b];
        }
        sum[(size_t)node * GUIDE_BINS + b] = s;
        cnt[(size_t)node * GUIDE_BINS + b] = c;
    }
}

// samples under each node (counts already summed up the tree)
__global__ void k_svo_node_count(uint32_t nodes, const uint32_t* __restrict__ cnt, uint32_t* __restrict__ out)
{
    for (uint32_t node = blockIdx.x * blockDim.x + threadIdx.x; node < nodes; node += gridDim.x * blockDim.x) {
        uint32_t c = 0;
        for (uint32_t b = 0; b < GUIDE_BINS; b++) c += cnt[(size_t)node * GUIDE_BINS + b];
        out[node] = c;
    }
}

// bins become means; a bin without samples takes its parent's value (root: 0)
__global__ void k_svo_down(uint32_t off, uint32_t count, const uint32_t* __restrict__ parent,
                           const uint32_t* __restrict__ cnt, float* __restrict__ mean)
{
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < count * GUIDE_BINS; t += gridDim.x * blockDim.x) {
        const uint32_t node = off + t / GUIDE_BINS, b = t % GUIDE_BINS;
        const size_t k = (size_t)node * GUIDE_BINS + b;
        const uint32_t c = cnt[k];
        if (c) {
            mean[k] = mean[k] / (float)c;
        } else {
            const uint32_t p = parent[node];
            mean[k] = (p == SVO_NO_NODE) ? 0.0f : mean[(size_t)p * GUIDE_BINS + b];
        }
    }
}

// 3x3 box blur over the bin grid (clamped at the borders) + floor
__global__ void k_svo_blur(uint32_t nodes, const float* __restrict__ mean, float* __restrict__ rad)
{
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nodes * GUIDE_BINS; t += gridDim.x * blockDim.x) {
        const uint32_t node = t / GUIDE_BINS, b = t % GUIDE_BINS;
        const int ix = (int)(b % GUIDE_RES), iy = (int)(b / GUIDE_RES);
        float s = 0.0f;
        int k = 0;
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                const int x = ix + dx, y = iy + dy;
                if (x < 0 || y < 0 || x >= (int)GUIDE_RES || y >= (int)GUIDE_RES) continue;
                s += mean[(size_t)node * GUIDE_BINS + y * GUIDE_RES + x];
                k++;
            }
        rad[t] = s / (float)k + GUIDE_EPS;
    }
}


// ---------------------------------------------------------------- voxeliser
// The geometry of the reference's guiding octree (TriFunctions::Voxelize,
// GPUPrimitiveTriangle.h:384-620): each triangle is projected onto the plane
// of its normal's dominant axis (the cube mapped to [-1,1]^3, the dominant
// axis turned to +z by the exact axis permutation that the reference's
// quarter- or half-turn quaternion stands for), rasterised conservatively at
// the leaf resolution (edges pushed out by half a pixel along their L1
// normal), and every covered pixel emits the leaf holding the triangle's
// plane point under the pixel centre.  One wave per triangle: the pixels of
// its rectangle 64 at a time, row by row, coverage by ballot; a count pass,
// a scan, and a write pass that places each record at its triangle's offset
// plus its rank, so records come out triangle by triangle, row by row.
struct VoxSetup {
    V3 p[3];
    float q[3][2];
    float c0[2], ec0[2], ec1[2], denom_c;
    float e0[2], e1[2], denom;
    int x0, x1, y0, y1;
};

__device__ inline void vox_setup(const float* __restrict__ tri, const float cube[4], uint32_t R, VoxSetup& s)
{
    for (int k = 0; k < 3; k++) s.p[k] = v3(tri[3 * k], tri[3 * k + 1], tri[3 * k + 2]);
    const float ea[3] = {s.p[1].x - s.p[0].x, s.p[1].y - s.p[0].y, s.p[1].z - s.p[0].z};
    const float eb[3] = {s.p[2].x - s.p[0].x, s.p[2].y - s.p[0].y, s.p[2].z - s.p[0].z};
    float n[3] = {ea[1] * eb[2] - ea[2] * eb[1], ea[2] * eb[0] - ea[0] * eb[2], ea[0] * eb[1] - ea[1] * eb[0]};
    const float inv = 1.0f / sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    n[0] *= inv;
    n[1] *= inv;
    n[2] *= inv;
    int dom = 0;
    float best = fabsf(n[0]);
    if (fabsf(n[1]) > best) {
        dom = 1;
        best = fabsf(n[1]);
    }
    if (fabsf(n[2]) > best) dom = 2;
    const bool neg = signbit(n[dom]);
    const float P[3][3] = {{s.p[0].x, s.p[0].y, s.p[0].z}, {s.p[1].x, s.p[1].y, s.p[1].z},
                           {s.p[2].x, s.p[2].y, s.p[2].z}};
    for (int k = 0; k < 3; k++) {
        float ndc[3];
        for (int a = 0; a < 3; a++) {
            const float lo = cube[a], hi = cube[a] + cube[3];
            const float sc = 2.0f / (hi - lo), tr = -((hi + lo) / (hi - lo));
            ndc[a] = sc * P[k][a] + tr;
        }
        float u, v;
        if (dom == 0) {
            u = neg ? ndc[2] : -ndc[2];
            v = ndc[1];
        } else if (dom == 1) {
            u = ndc[0];
            v = neg ? ndc[2] : -ndc[2];
        } else {
            u = neg ? -ndc[0] : ndc[0];
            v = ndc[1];
        }
        s.q[k][0] = u;
        s.q[k][1] = v;
    }
    float mn[2] = {s.q[0][0], s.q[0][1]}, mx[2] = {s.q[0][0], s.q[0][1]};
    for (int k = 1; k < 3; k++)
        for (int a = 0; a < 2; a++) {
            mn[a] = fminf(mn[a], s.q[k][a]);
            mx[a] = fmaxf(mx[a], s.q[k][a]);
        }
    const float Rf = (float)R;
    int r[4] = {(int)floorf((0.5f + 0.5f * mn[0]) * Rf), (int)ceilf((0.5f + 0.5f * mx[0]) * Rf),
                (int)floorf((0.5f + 0.5f * mn[1]) * Rf), (int)ceilf((0.5f + 0.5f * mx[1]) * Rf)};
    for (int i = 0; i < 4; i++) r[i] = r[i] < 0 ? 0 : (r[i] > (int)R ? (int)R : r[i]);
    s.x0 = r[0];
    s.x1 = r[1];
    s.y0 = r[2];
    s.y1 = r[3];
    const float hp = 1.0f / Rf;
    float pl[3][3];
    for (int k = 0; k < 3; k++) {
        const int k1 = (k + 1) % 3;
        const float dx = s.q[k1][0] - s.q[k][0], dy = s.q[k1][1] - s.q[k][1];
        pl[k][0] = dy;
        pl[k][1] = -dx;
        pl[k][2] = dx * s.q[k][1] - dy * s.q[k][0];
        pl[k][2] -= hp * fabsf(pl[k][0]) + hp * fabsf(pl[k][1]);
    }
    const int pair[3][2] = {{0, 2}, {0, 1}, {1, 2}};
    float cc[3][2];
    for (int k = 0; k < 3; k++) {
        const float* A = pl[pair[k][0]];
        const float* B = pl[pair[k][1]];
        const float X = A[1] * B[2] - A[2] * B[1], Y = A[2] * B[0] - A[0] * B[2], Z = A[0] * B[1] - A[1] * B[0];
        cc[k][0] = X / Z;
        cc[k][1] = Y / Z;
    }
    s.c0[0] = cc[0][0];
    s.c0[1] = cc[0][1];
    s.ec0[0] = cc[1][0] - cc[0][0];
    s.ec0[1] = cc[1][1] - cc[0][1];
    s.ec1[0] = cc[2][0] - cc[0][0];
    s.ec1[1] = cc[2][1] - cc[0][1];
    s.denom_c = 1.0f / (s.ec0[0] * s.ec1[1] - s.ec1[0] * s.ec0[1]);
    s.e0[0] = s.q[1][0] - s.q[0][0];
    s.e0[1] = s.q[1][1] - s.q[0][1];
    s.e1[0] = s.q[2][0] - s.q[0][0];
    s.e1[1] = s.q[2][1] - s.q[0][1];
    s.denom = 1.0f / (s.e0[0] * s.e1[1] - s.e1[0] * s.e0[1]);
}

// GPUPrimitiveTriangle.h:548-612: conservative barycentrics of the pixel
// centre in range -> the leaf of the original triangle's plane point there
__device__ inline bool vox_pixel(const VoxSetup& s, const float cube[4], uint32_t R, int x, int y, uint32_t& code)
{
    const float dp = 2.0f / (float)R;
    const float px = ((float)x + 0.5f) * dp - 1.0f, py = ((float)y + 0.5f) * dp - 1.0f;
    const float g0 = px - s.c0[0], g1 = py - s.c0[1];
    const float b = (g0 * s.ec1[1] - s.ec1[0] * g1) * s.denom_c;
    const float c = (s.ec0[0] * g1 - g0 * s.ec0[1]) * s.denom_c;
    const float a = 1.0f - b - c;
    if (!(b >= 0.0f && b <= 1.0f && c >= 0.0f && c <= 1.0f && a >= 0.0f && a <= 1.0f)) return false;
    const float h0 = px - s.q[0][0], h1 = py - s.q[0][1];
    const float B = (h0 * s.e1[1] - s.e1[0] * h1) * s.denom;
    const float C = (s.e0[0] * h1 - h0 * s.e0[1]) * s.denom;
    const float A = 1.0f - B - C;
    const float pos[3] = {s.p[0].x * A + s.p[1].x * B + s.p[2].x * C, s.p[0].y * A + s.p[1].y * B + s.p[2].y * C,
                          s.p[0].z * A + s.p[1].z * B + s.p[2].z * C};
    uint32_t q[3];
    for (int k = 0; k < 3; k++) {
        const float f = ((pos[k] - cube[k]) / cube[3]) * (float)R;
        q[k] = !(f > 0.0f) ? 0u : (f >= (float)(R - 1) ? R - 1 : (uint32_t)f);
    }
    code = expand10(q[0]) | (expand10(q[1]) << 1) | (expand10(q[2]) << 2);
    return true;
}

// offsets == nullptr: out[t] = records of triangle t; else out[offsets[t] +
// rank] = the records
__global__ __launch_bounds__(256) void k_voxelize(const float* __restrict__ tris, uint32_t n, float4 cube4,
                                                  uint32_t depth, const uint32_t* __restrict__ offsets,
                                                  uint32_t* __restrict__ out, uint32_t* __restrict__ prims,
                                                  unsigned long long* __restrict__ total)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
    const uint32_t R = 1u << depth;
    const float cube[4] = {cube4.x, cube4.y, cube4.z, cube4.w};
    const uint64_t below = (1ull << lane) - 1ull;
    unsigned long long wave_total = 0ull;   // one atomic per wave, not per triangle
    for (uint32_t t = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; t < n; t += nw) {
        VoxSetup s;
        vox_setup(tris + 9 * (size_t)t, cube, R, s);
        const int w = s.x1 - s.x0, h = s.y1 - s.y0;
        const uint32_t area = (w > 0 && h > 0) ? (uint32_t)(w * h) : 0u;
        uint32_t run = offsets ? offsets[t] : 0u;
        for (uint32_t base = 0; base < area; base += 64u) {
            const uint32_t k = base + lane;
            uint32_t code = 0u;
            bool hit = false;
            if (k < area) hit = vox_pixel(s, cube, R, s.x0 + (int)(k % (uint32_t)w), s.y0 + (int)(k / (uint32_t)w), code);
            const uint64_t m = __ballot(hit);
            if (offsets && hit) {
                const uint32_t at = run + (uint32_t)__popcll(m & below);
                out[at] = code;
                if (prims) prims[at] = t;
            }
            run += (uint32_t)__popcll(m);
        }
        if (!offsets && lane == 0) {
            out[t] = run;
            wave_total += run;
        }
    }
    if (!offsets && lane == 0 && wave_total) atomicAdd(total, wave_total);
}

// ---------------------------------------------------------------- cone trace
// AnisoSVOctreeGPU::ConeTraceRay (AnisoSVO.cuh:767-1032) on this octree: the
// ray in the cube's unit space, mirrored to a positive direction; a Morton
// bit stack of the child indices taken and the parent links instead of a
// node stack; descending into an existing child stops at a leaf, where the
// cone's disc (diameter^2 = 4 t^2 aperture / pi) is wider than the child's
// voxel, or `offset` levels above the leaves; a missing child is crossed
// past its far face.  A walk longer than CONE_MAX_STEPS steps ends as a miss.
constexpr uint32_t CONE_MAX_STEPS = 1u << 16;

__device__ inline float neg_exp2i(uint32_t p) { return p == 0 ? 1.0f : __int_as_float((int)((127u - p) << 23)); }

struct ConeTree {
    const uint32_t* first_child;
    const uint32_t* child_mask;
    const uint32_t* parent;
    uint32_t depth;
    float lo[3], size;
};

// One cone from o along dir (world space) starting at tmin: the node it
// stops at (SVO_NO_NODE: left the cube), whether it is a leaf; returns t.
__device__ inline float cone_trace_one(const ConeTree& tr, const float o[3], const float dir[3], float tmin,
                                       float aperture, uint32_t offset, uint32_t& node, uint32_t& is_leaf)
{
    const float EPS = 1.0e-5f;
    const float scale = 1.0f / tr.size;
    const float leaf = tr.size / (float)(1u << tr.depth);
    float coeff[3], bias[3];
    uint32_t mirror = 0;
    for (int a = 0; a < 3; a++) {
        float d = (fabsf(dir[a]) < EPS) ? copysignf(EPS, dir[a]) : dir[a];
        d = d * scale;
        float p = (o[a] - tr.lo[a]) * scale;
        if (d < 0.0f) {
            p = 1.0f - p;
            mirror |= 1u << a;
        }
        coeff[a] = 1.0f / fabsf(d);
        bias[a] = p * coeff[a];
    }
    const float te0 = -bias[0], te1 = -bias[1], te2 = -bias[2];
    const float tsvo = fmaxf(fmaxf(te0, te1), te2);
    if (!(te0 < tmin && te1 < tmin && te2 < tmin)) tmin = nextafterf(tsvo, INFINITY);
    uint64_t stack = 0;
    uint32_t cnt = 0, nid = 0;
    node = SVO_NO_NODE;
    is_leaf = 0;
    float corner[3] = {0.0f, 0.0f, 0.0f};
    bool empty = false;
    for (uint32_t step = 0; nid != SVO_NO_NODE; step++) {
        if (step >= CONE_MAX_STEPS) break;
        const float cs = neg_exp2i(cnt + 1), vs = neg_exp2i(cnt);
        float tmax = INFINITY;
        for (int a = 0; a < 3; a++) tmax = fminf(tmax, (corner[a] + vs) * coeff[a] - bias[a]);
        if (empty || tmax < tmin) {
            if (empty) tmin = nextafterf(tmax, INFINITY);
            const uint32_t bits = (uint32_t)(stack & 7u);
            for (int a = 0; a < 3; a++)
                if ((bits >> a) & 1u) corner[a] -= vs;
            stack >>= 3;
            if (!empty) {
                if (cnt == 0) break;
                nid = tr.parent[nid];
            }
            cnt--;
            empty = false;
            continue;
        }
        uint32_t child = 0;
        for (int a = 0; a < 3; a++)
            if ((corner[a] + cs) * coeff[a] - bias[a] < tmin) child |= 1u << a;
        const uint32_t mir = child ^ mirror;
        const uint32_t mask = tr.child_mask[nid];
        const bool has = (mask >> mir) & 1u;
        if (has) {
            nid = tr.first_child[nid] + (uint32_t)__popc(mask & ((1u << mir) - 1u));
            const uint32_t level = cnt + 1;
            const float vsz = leaf * (float)(1u << (tr.depth - level));
            const float diam2 = 4.0f * ((tmin * tmin * aperture) * 0.318309886183790671538f);
            const bool child_leaf = (level == tr.depth);
            if (child_leaf || diam2 > vsz * vsz || (tr.depth - level) <= offset) {
                node = nid;
                is_leaf = child_leaf ? 1u : 0u;
                return tmin;
            }
        }
        empty = !has;
        for (int a = 0; a < 3; a++)
            if ((child >> a) & 1u) corner[a] += cs;
        stack = (stack << 3) | child;
        cnt++;
    }
    node = SVO_NO_NODE;
    is_leaf = 0;
    return tmin;
}

__global__ __launch_bounds__(256) void k_svo_cone_trace(ConeTree tr, const float* __restrict__ rays, uint32_t n,
                                                        float aperture, uint32_t offset, uint32_t* __restrict__ out_node,
                                                        float* __restrict__ out_t, uint32_t* __restrict__ out_leaf)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float* r = rays + 7 * (size_t)i;
        const float o[3] = {r[0], r[1], r[2]}, d[3] = {r[3], r[4], r[5]};
        uint32_t node, leaf;
        const float t = cone_trace_one(tr, o, d, r[6], aperture, offset, node, leaf);
        out_node[i] = node;
        out_t[i] = t;
        if (out_leaf) out_leaf[i] = leaf;
    }
}

// GenerateRadianceField (WFPGTracerKC.cuh:1252-1318) for every texel of
// every bin: the texel's direction with the bin's jitter on the concentric
// octahedral map (y up), a cone of the texel's solid angle from the bin's
// origin, and the radiance the stopped node sends back along the cone
// (front or back side by the node's normal; the boundary radiance on a miss).
__global__ __launch_bounds__(256) void k_wfpg_trace_fields(ConeTree tr, uint32_t X, uint32_t Y, uint32_t lx,
                                                           uint32_t bins, const float4* __restrict__ origins,
                                                           const float2* __restrict__ jitter, float aperture,
                                                           uint32_t offset, const float2* __restrict__ irr,
                                                           const float* __restrict__ normals, float boundary,
                                                           float* __restrict__ fields, uint32_t* __restrict__ out_node,
                                                           float* __restrict__ out_t)
{
    const uint32_t per = X * Y;
    const uint64_t total = (uint64_t)bins * per;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t b = (uint32_t)(i / per), p = (uint32_t)(i % per);
        const uint32_t x = p & (X - 1u), y = p >> lx;
        const float4 og = origins[b];
        const float2 jt = jitter[b];
        const float u = ((float)x + jt.x) / (float)X, v = ((float)y + jt.y) / (float)Y;
        const V3 dz = octa_to_dir(u, v);
        const float o[3] = {og.x, og.y, og.z}, d[3] = {dz.y, dz.z, dz.x};
        uint32_t node, leaf;
        const float t = cone_trace_one(tr, o, d, og.w, aperture, offset, node, leaf);
        float rad = boundary;
        if (node != SVO_NO_NODE) {
            const float* nrm = normals + 3 * (size_t)node;
            const bool front = (nrm[0] * -d[0] + nrm[1] * -d[1] + nrm[2] * -d[2]) >= 0.0f;
            const float2 two = irr[node];
            rad = front ? two.x : two.y;
        }
        fields[i] = rad;
        if (out_node) out_node[i] = node;
        if (out_t) out_t[i] = t;
    }
}

// ---------------------------------------------------------------- payload
// The octree's node payload as the reference forms it (VoxelPayload,
// AnisoSVO.cuh:23-115): normals from a 2-means clustering of the voxel
// records' triangle normals (KCReduceVoxelPayload, AnisoSVOKC.cuh:8-230) and
// of the children's average normals bottom-up (KCAverageNormals,
// AnisoSVO.cu:419-520); path-vertex radiance splatted into the nearest
// leaf's front or back side (KCAccumulateRadianceToLeaf, :225-263; 48.16
// fixed point, order independent), averaged, averaged bottom-up over the
// children's valid sides (swapping a child whose normal opposes) and filled
// top-down (NormalizeAndFilterRadiance, :1415-1483).
constexpr float PAYLOAD_INVALID = 65504.0f;   // the reference's half maximum

struct Vec3 {
    float v[3];
};

// 2-means of up to n vectors (zero vectors skipped), 4 rounds seeded with
// the first non-zero vector and its negation; a vector joins the mean it is
// farther from (the reference's rule: the labels swap each round).  Writes
// the larger cluster's average as unit normal + length.
template <class Get>
__device__ inline float4 kmeans2_dev(uint32_t n, Get get)
{
    uint32_t first = n;
    for (uint32_t i = 0; i < n; i++) {
        const Vec3 a = get(i);
        if (!(a.v[0] == 0.0f && a.v[1] == 0.0f && a.v[2] == 0.0f)) {
            first = i;
            break;
        }
    }
    if (first == n) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const Vec3 f = get(first);
    float m[2][3] = {{f.v[0], f.v[1], f.v[2]}, {-f.v[0], -f.v[1], -f.v[2]}};
    uint32_t cnt[2] = {0, 0};
    for (int it = 0; it < 4; it++) {
        for (int k = 0; k < 2; k++) {
            const float inv = 1.0f / sqrtf(m[k][0] * m[k][0] + m[k][1] * m[k][1] + m[k][2] * m[k][2]);
            m[k][0] *= inv;
            m[k][1] *= inv;
            m[k][2] *= inv;
        }
        float acc[2][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
        cnt[0] = cnt[1] = 0;
        for (uint32_t i = 0; i < n; i++) {
            const Vec3 a = get(i);
            if (a.v[0] == 0.0f && a.v[1] == 0.0f && a.v[2] == 0.0f) continue;
            const float d0 = 1.0f - (m[0][0] * a.v[0] + m[0][1] * a.v[1] + m[0][2] * a.v[2]);
            const float d1 = 1.0f - (m[1][0] * a.v[0] + m[1][1] * a.v[1] + m[1][2] * a.v[2]);
            const int k = (d0 >= d1) ? 0 : 1;
            acc[k][0] += a.v[0];
            acc[k][1] += a.v[1];
            acc[k][2] += a.v[2];
            cnt[k]++;
        }
        for (int k = 0; k < 2; k++)
            for (int c = 0; c < 3; c++) m[k][c] = acc[k][c];
    }
    const int k = (cnt[0] >= cnt[1]) ? 0 : 1;
    const float r = 1.0f / (float)cnt[k];
    const float mean[3] = {m[k][0] * r, m[k][1] * r, m[k][2] * r};
    const float len = sqrtf(mean[0] * mean[0] + mean[1] * mean[1] + mean[2] * mean[2]);
    const float il = 1.0f / len;
    return make_float4(mean[0] * il, mean[1] * il, mean[2] * il, len);
}

// the leaf of a cell code, SVO_NO_NODE when the tree lacks it
__device__ inline uint32_t leaf_of_code(const uint32_t* __restrict__ first_child, const uint32_t* __restrict__ mask,
                                        uint32_t depth, uint32_t code)
{
    uint32_t node = 0;
    for (uint32_t l = 0; l < depth; l++) {
        const uint32_t oct = (code >> (3u * (depth - 1u - l))) & 7u;
        const uint32_t m = mask[node];
        if (!((m >> oct) & 1u)) return SVO_NO_NODE;
        node = first_child[node] + (uint32_t)__popc(m & ((1u << oct) - 1u));
    }
    return node;
}

__global__ void k_pl_keys(const uint32_t* __restrict__ codes, uint32_t n, const uint32_t* __restrict__ first_child,
                          const uint32_t* __restrict__ mask, uint32_t depth, uint32_t leaf0, uint32_t* __restrict__ keys)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t lf = leaf_of_code(first_child, mask, depth, codes[i]);
        keys[i] = (lf == SVO_NO_NODE) ? 0xFFFFFFFFu : lf - leaf0;
    }
}

// [begin, end) of each leaf's run in the sorted keys
__global__ void k_pl_runs(const uint32_t* __restrict__ keys, uint32_t n, uint32_t* __restrict__ begin,
                          uint32_t* __restrict__ end)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t k = keys[i];
        if (k == 0xFFFFFFFFu) continue;
        if (i == 0 || keys[i - 1] != k) begin[k] = i;
        if (i == n - 1 || keys[i + 1] != k) end[k] = i + 1;
    }
}

__global__ void k_pl_leaf_normals(uint32_t leaves, uint32_t leaf0, const uint32_t* __restrict__ begin,
                                  const uint32_t* __restrict__ end, const uint32_t* __restrict__ order,
                                  const uint32_t* __restrict__ prims, const float* __restrict__ tris,
                                  float4* __restrict__ nrm)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < leaves; i += gridDim.x * blockDim.x) {
        const uint32_t b = begin[i], e = end[i];
        nrm[leaf0 + i] = kmeans2_dev(e - b, [&](uint32_t k) {
            const float* t = tris + 9 * (size_t)prims[order[b + k]];
            const float ea[3] = {t[3] - t[0], t[4] - t[1], t[5] - t[2]};
            const float eb[3] = {t[6] - t[0], t[7] - t[1], t[8] - t[2]};
            const float nn[3] = {ea[1] * eb[2] - ea[2] * eb[1], ea[2] * eb[0] - ea[0] * eb[2],
                                 ea[0] * eb[1] - ea[1] * eb[0]};
            const float inv = 1.0f / sqrtf(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
            Vec3 r;
            r.v[0] = nn[0] * inv;
            r.v[1] = nn[1] * inv;
            r.v[2] = nn[2] * inv;
            return r;
        });
    }
}

__global__ void k_pl_node_normals(uint32_t off, uint32_t count, const uint32_t* __restrict__ first_child,
                                  const uint32_t* __restrict__ mask, float4* __restrict__ nrm)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t nd = off + i, fc = first_child[nd], k = (uint32_t)__popc(mask[nd]);
        nrm[nd] = kmeans2_dev(k, [&](uint32_t c) {
            const float4 q = nrm[fc + c];
            Vec3 r;
            r.v[0] = q.x * q.w;
            r.v[1] = q.y * q.w;
            r.v[2] = q.z * q.w;
            return r;
        });
    }
}

__global__ void k_pl_splat(const float* __restrict__ verts, uint32_t n, const uint32_t* __restrict__ vmask,
                           const uint32_t* __restrict__ first_child,
                           const uint32_t* __restrict__ mask, uint32_t depth, float4 cube, const float4* __restrict__ nrm,
                           unsigned long long* __restrict__ tot, uint32_t* __restrict__ num)
{
    const uint32_t R = 1u << depth;
    const float lo[3] = {cube.x, cube.y, cube.z};
    const float leaf = cube.w / (float)R;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        if (vmask && vmask[v] == 0xFFFFFFFFu) continue;
        const float* x = verts + 7 * (size_t)v;
        bool inside = true;
        for (int a = 0; a < 3; a++) inside = inside && !(x[a] < lo[a] || x[a] > lo[a] + cube.w);
        if (!inside) continue;
        float fi[3], dist[3];
        int inc[3];
        for (int a = 0; a < 3; a++) {
            const float l = (x[a] - lo[a]) / leaf;
            fi[a] = floorf(l);
            const float fr = l - fi[a];
            inc[a] = (fr < 0.5f) ? -1 : 1;
            dist[a] = 0.5f - fabsf(fr - 0.5f);
        }
        int m0 = 0;
        if (dist[1] < dist[m0]) m0 = 1;
        if (dist[2] < dist[m0]) m0 = 2;
        const int o1 = (m0 + 1) % 3, o2 = (m0 + 2) % 3;
        const int m1 = (dist[o2] < dist[o1]) ? o2 : o1;
        const int m2 = 3 - m0 - m1;
        uint32_t lf = SVO_NO_NODE;
        for (uint32_t i = 0; i < 8 && lf == SVO_NO_NODE; i++) {
            int sh[3];
            sh[m0] = (int)(i & 1u);
            sh[m1] = (int)((i >> 1) & 1u);
            sh[m2] = (int)((i >> 2) & 1u);
            long long q[3];
            bool ok = true;
            for (int a = 0; a < 3; a++) {
                q[a] = (long long)fi[a] + sh[a] * inc[a];
                ok = ok && q[a] >= 0 && q[a] < (long long)R;
            }
            if (!ok) continue;
            const uint32_t code =
                expand10((uint32_t)q[0]) | (expand10((uint32_t)q[1]) << 1) | (expand10((uint32_t)q[2]) << 2);
            lf = leaf_of_code(first_child, mask, depth, code);
        }
        if (lf == SVO_NO_NODE) continue;
        const float4 nn = nrm[lf];
        const int side = (nn.x * x[3] + nn.y * x[4] + nn.z * x[5] >= 0.0f) ? 0 : 1;
        const float lum = x[6];
        const unsigned long long f =
            (lum > 0.0f) ? (unsigned long long)(fminf(lum, 2147483647.0f) * 65536.0f) : 0ull;
        // the lanes still here that splat into the same leaf side add up in
        // the wave first (path vertices of neighbouring samples share
        // leaves), one atomic pair per distinct side; integer sums, so the
        // result is the same
        const uint32_t slot = 2u * lf + (uint32_t)side;
        const uint32_t flo = (uint32_t)f, fhi = (uint32_t)(f >> 32);
        uint64_t pend = __ballot(1);
        while (pend) {
            const int l = __ffsll((unsigned long long)pend) - 1;
            const uint32_t ls = __shfl(slot, l, 64);
            const uint64_t same = __ballot(slot == ls);
            unsigned long long sum = 0;
            for (uint64_t it = same; it; it &= it - 1ull) {
                const int j = __ffsll((unsigned long long)it) - 1;
                sum += ((unsigned long long)(uint32_t)__shfl(fhi, j, 64) << 32) | (uint32_t)__shfl(flo, j, 64);
            }
            if ((int)(threadIdx.x & 63u) == l) {
                atomicAdd(&tot[ls], sum);
                atomicAdd(&num[ls], (uint32_t)__popcll(same));
            }
            pend &= ~same;
        }
    }
}

__global__ void k_pl_leaf_avg(uint32_t leaf0, uint32_t leaves, const unsigned long long* __restrict__ tot,
                              const uint32_t* __restrict__ num, float2* __restrict__ irr)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < leaves; i += gridDim.x * blockDim.x) {
        const size_t lf = leaf0 + i;
        float r[2];
        for (int j = 0; j < 2; j++) {
            const uint32_t c = num[2 * lf + j];
            r[j] = c ? ((float)tot[2 * lf + j] * (1.0f / 65536.0f)) / (float)c : PAYLOAD_INVALID;
        }
        irr[lf] = make_float2(r[0], r[1]);
    }
}

__global__ void k_pl_bottom_up(uint32_t off, uint32_t count, const uint32_t* __restrict__ first_child,
                               const uint32_t* __restrict__ mask, const float4* __restrict__ nrm, float2* __restrict__ irr)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t nd = off + i, fc = first_child[nd], k = (uint32_t)__popc(mask[nd]);
        const float4 pn = nrm[nd];
        float sum[2] = {0.0f, 0.0f};
        uint32_t c[2] = {0, 0};
        for (uint32_t q = 0; q < k; q++) {
            const float4 cn = nrm[fc + q];
            const bool swap = (cn.x * pn.x + cn.y * pn.y + cn.z * pn.z) < 0.0f;
            const float2 ci = irr[fc + q];
            const float vals[2] = {swap ? ci.y : ci.x, swap ? ci.x : ci.y};
            for (int j = 0; j < 2; j++)
                if (vals[j] != PAYLOAD_INVALID) {
                    sum[j] += vals[j];
                    c[j]++;
                }
        }
        irr[nd] = make_float2(c[0] ? sum[0] * (1.0f / (float)c[0]) : PAYLOAD_INVALID,
                              c[1] ? sum[1] * (1.0f / (float)c[1]) : PAYLOAD_INVALID);
    }
}

__global__ void k_pl_top_down(uint32_t off, uint32_t count, const uint32_t* __restrict__ parent, float2* __restrict__ irr)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t nd = off + i;
        float2 v = irr[nd];
        const float2 p = irr[parent[nd]];
        if (v.x == PAYLOAD_INVALID) v.x = p.x;
        if (v.y == PAYLOAD_INVALID) v.y = p.y;
        irr[nd] = v;
    }
}

__global__ void k_pl_out(uint32_t nodes, const float4* __restrict__ nrm, float* __restrict__ normals,
                         float* __restrict__ lengths)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nodes; i += gridDim.x * blockDim.x) {
        const float4 q = nrm[i];
        normals[3 * (size_t)i] = q.x;
        normals[3 * (size_t)i + 1] = q.y;
        normals[3 * (size_t)i + 2] = q.z;
        if (lengths) lengths[i] = q.w;
    }
}
}  // namespace

SvoView svo_view(const mrh_svo* s)
{
    SvoView v;
    v.depth = s->depth;
    for (int i = 0; i < 12; i++) v.lvl_off[i] = s->lvl_off[i];
    v.pack = s->pack;
    v.radiance = s->radiance;
    v.lox = s->cube[0];
    v.loy = s->cube[1];
    v.loz = s->cube[2];
    v.inv_cell = (float)(1u << s->depth) / s->cube[3];
    return v;
}

static int32_t svo_reserve(mrh_svo* s, uint32_t nodes)
{
    mrh_context* ctx = s->ctx;
    if (nodes <= s->cap) return MRH_OK;
    (void)hipStreamSynchronize(ctx->stream);
    void* ptrs[] = {s->codes, s->first_child, s->child_mask, s->parent, s->sum, s->acc, s->cnt, s->radiance,
                    s->node_count, s->pack};
    for (void* p : ptrs) (void)hipFree(p);
    s->codes = s->first_child = s->child_mask = s->parent = s->node_count = nullptr;
    s->pack = nullptr;
    s->sum = s->radiance = nullptr;
    s->acc = nullptr;
    s->cnt = nullptr;
    s->cap = 0;
    const uint32_t cap = nodes + nodes / 4 + 64;
    MRH_TRY(dmalloc(ctx, (void**)&s->codes, (size_t)cap * 4));
    MRH_TRY(dmalloc(ctx, (void**)&s->first_child, (size_t)cap * 4));
    MRH_TRY(dmalloc(ctx, (void**)&s->child_mask, (size_t)cap * 4));
    MRH_TRY(dmalloc(ctx, (void**)&s->parent, (size_t)cap * 4));
    MRH_TRY(dmalloc(ctx, (void**)&s->sum, (size_t)cap * GUIDE_BINS * 4));
    MRH_TRY(dmalloc(ctx, (void**)&s->acc, (size_t)cap * GUIDE_BINS * 8));
    MRH_TRY(dmalloc(ctx, (void**)&s->cnt, (size_t)cap * GUIDE_BINS * 4));
    MRH_TRY(dmalloc(ctx, (void**)&s->radiance, (size_t)cap * GUIDE_BINS * 4));
    MRH_TRY(dmalloc(ctx, (void**)&s->node_count, (size_t)cap * 4));
    MRH_TRY(dmalloc(ctx, (void**)&s->pack, (size_t)cap * 16));
    s->cap = cap;
    return MRH_OK;
}

int32_t svo_codes(mrh_context* ctx, const float* d_points, uint32_t n, uint32_t pitch, const float cube[4],
                  uint32_t depth, uint32_t* d_codes)
{
    const float inv_cell = (float)(1u << depth) / cube[3];
    return launch(ctx, "svo_codes", dim3(stride_grid(ctx, n, 256)), dim3(256), 0, k_svo_codes, d_points, n, pitch,
                  cube[0], cube[1], cube[2], inv_cell, depth, d_codes);
}

// grow-only device buffer
static int32_t grow(mrh_context* ctx, void** p, size_t* have, size_t want)
{
    if (want <= *have) return MRH_OK;
    if (*p) {
        MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        (void)hipFree(*p);
        *p = nullptr;
        *have = 0;
    }
    want += want / 8 + 4096;
    MRH_TRY(dmalloc(ctx, p, want));
    *have = want;
    return MRH_OK;
}

// Sort the cells, count every level's distinct prefixes in one pass (one
// host round trip), then for each level from the leaves up: unique prefixes
// and the parent index of the level below; then links, lookup records and
// fresh radiance.
int32_t svo_build(mrh_context* ctx, mrh_svo* s, const uint32_t* d_codes, uint32_t n, const float cube[4],
                  uint32_t depth, bool filtered)
{
    if (depth < 1 || depth > 10) return fail(ctx, MRH_ERR_INVALID, "svo depth must be in [1, 10]");
    if (n == 0) return fail(ctx, MRH_ERR_INVALID, "svo build: no points");
    if (!(cube[3] > 0.0f)) return fail(ctx, MRH_ERR_INVALID, "svo build: empty cube");
    for (int i = 0; i < 4; i++) s->cube[i] = cube[i];
    s->depth = depth;
    // work_a: sorted codes, sort values / flags, scan positions, scan scratch
    const size_t tmp_words = scan_tmp_words(n);
    MRH_TRY(grow(ctx, &s->work_a, &s->work_a_bytes, ((size_t)n * 3 + tmp_words) * 4));
    if (!s->counter) MRH_TRY(dmalloc(ctx, (void**)&s->counter, 64));
    uint32_t* sorted = (uint32_t*)s->work_a;
    uint32_t* flags = sorted + n;   // also the sort's value output
    uint32_t* pos = flags + n;
    uint32_t* tmp = pos + n;
    SortArgs sa;
    sa.keys_in = d_codes;
    sa.keys_out = sorted;
    sa.vals_out = flags;
    sa.n_max = n;
    sa.key_bytes = 4;
    sa.begin_bit = 0;
    sa.end_bit = 3 * depth;
    MRH_TRY(radix_sort(ctx, sa));
    MRH_TRY(fill_u32(ctx, s->counter, 0u, 16));
    MRH_TRY(launch(ctx, "svo_level_hist", dim3(stride_grid(ctx, n, 256)), dim3(256), 0, k_level_hist,
                   (const uint32_t*)sorted, n, depth, s->counter));
    uint32_t hist[12] = {};
    MRH_HIP_CHECK(ctx, hipMemcpyAsync(hist, s->counter, sizeof(hist), hipMemcpyDeviceToHost, ctx->stream));
    MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    uint32_t run = 0;
    for (uint32_t l = 0; l <= depth; l++) {
        run += hist[l];
        s->lvl_count[l] = run;
    }
    for (uint32_t l = depth + 1; l < 11; l++) s->lvl_count[l] = 0;
    if (s->lvl_count[0] != 1) return fail(ctx, MRH_ERR_HIP, "svo build: %u roots", s->lvl_count[0]);
    s->lvl_off[0] = 0;
    for (uint32_t l = 0; l <= depth; l++) s->lvl_off[l + 1] = s->lvl_off[l] + s->lvl_count[l];
    for (uint32_t l = depth + 1; l < 11; l++) s->lvl_off[l + 1] = s->lvl_off[l];
    s->nodes = s->lvl_off[depth + 1];
    // work_b: codes of every level and parent indices of every level but the root's
    MRH_TRY(grow(ctx, &s->work_b, &s->work_b_bytes, (size_t)s->nodes * 2 * 4));
    uint32_t* lvl_codes[11] = {};
    uint32_t* lvl_parent[11] = {};   // parent_local of level l's nodes (index into level l-1)
    uint32_t* q = (uint32_t*)s->work_b;
    for (uint32_t l = 0; l <= depth; l++) {
        lvl_codes[l] = q;
        q += s->lvl_count[l];
        lvl_parent[l] = l ? q : nullptr;
        q += l ? s->lvl_count[l] : 0;
    }
    MRH_TRY(unique_runs(ctx, sorted, n, 0, lvl_codes[depth], nullptr, flags, pos, tmp));
    for (int l = (int)depth - 1; l >= 0; l--)
        MRH_TRY(unique_runs(ctx, lvl_codes[l + 1], s->lvl_count[l + 1], 3, lvl_codes[l], lvl_parent[l + 1], flags,
                            pos, tmp));
    MRH_TRY(svo_reserve(s, s->nodes));
    for (uint32_t l = 0; l <= depth; l++) {
        const uint32_t c = s->lvl_count[l];
        MRH_TRY(launch(ctx, "svo_link", dim3(stride_grid(ctx, c, 256)), dim3(256), 0, k_svo_link,
                       (const uint32_t*)lvl_codes[l], (const uint32_t*)lvl_parent[l], c, s->lvl_off[l],
                       l ? s->lvl_off[l - 1] : 0u, s->codes, s->parent, s->first_child, s->child_mask));
    }
    MRH_TRY(launch(ctx, "svo_pack", dim3(stride_grid(ctx, s->nodes, 256)), dim3(256), 0, k_svo_pack, s->nodes,
                   (const uint32_t*)s->first_child, (const uint32_t*)s->child_mask, (const uint32_t*)nullptr,
                   s->pack));
    // fresh radiance
    MRH_TRY(fill_u32(ctx, (uint32_t*)s->acc, 0u, (size_t)s->nodes * GUIDE_BINS * 2));
    MRH_TRY(fill_u32(ctx, s->cnt, 0u, (size_t)s->nodes * GUIDE_BINS));
    if (filtered) return MRH_OK;
    MRH_TRY(fill_u32(ctx, (uint32_t*)s->sum, 0u, (size_t)s->nodes * GUIDE_BINS));
    uint32_t eps_bits;
    std::memcpy(&eps_bits, &GUIDE_EPS, 4);
    MRH_TRY(fill_u32(ctx, (uint32_t*)s->radiance, eps_bits, (size_t)s->nodes * GUIDE_BINS));
    return fill_u32(ctx, s->node_count, 0u, s->nodes);
}

int32_t svo_splat(mrh_context* ctx, mrh_svo* s, const uint32_t* codes, const uint32_t* bins, const float* values,
                  uint32_t n)
{
    if (!s->nodes) return fail(ctx, MRH_ERR_INVALID, "svo splat: empty svo");
    if (n == 0) return MRH_OK;
    return launch(ctx, "svo_splat", dim3(stride_grid(ctx, n, 256)), dim3(256), 0, k_svo_splat, svo_view(s), codes,
                  bins, values, n, s->acc, s->cnt);
}

int32_t svo_filter(mrh_context* ctx, mrh_svo* s)
{
    if (!s->nodes) return fail(ctx, MRH_ERR_INVALID, "svo filter: empty svo");
    const size_t bins = (size_t)s->nodes * GUIDE_BINS;
    MRH_TRY(launch(ctx, "svo_fix", dim3(stride_grid(ctx, bins, 256)), dim3(256), 0, k_svo_fix, bins,
                   (const unsigned long long*)s->acc, s->sum));
    for (int l = (int)s->depth - 1; l >= 0; l--) {
        const uint32_t c = s->lvl_count[l];
        MRH_TRY(launch(ctx, "svo_up", dim3(stride_grid(ctx, (uint64_t)c * GUIDE_BINS, 256)), dim3(256), 0, k_svo_up,
                       s->lvl_off[l], c, (const uint32_t*)s->first_child, (const uint32_t*)s->child_mask, s->sum,
                       s->cnt));
    }
    MRH_TRY(launch(ctx, "svo_node_count", dim3(stride_grid(ctx, s->nodes, 256)), dim3(256), 0, k_svo_node_count,
                   s->nodes, (const uint32_t*)s->cnt, s->node_count));
    MRH_TRY(launch(ctx, "svo_pack", dim3(stride_grid(ctx, s->nodes, 256)), dim3(256), 0, k_svo_pack, s->nodes,
                   (const uint32_t*)s->first_child, (const uint32_t*)s->child_mask, (const uint32_t*)s->node_count,
                   s->pack));
    for (uint32_t l = 0; l <= s->depth; l++) {
        const uint32_t c = s->lvl_count[l];
        MRH_TRY(launch(ctx, "svo_down", dim3(stride_grid(ctx, (uint64_t)c * GUIDE_BINS, 256)), dim3(256), 0,
                       k_svo_down, s->lvl_off[l], c, (const uint32_t*)s->parent, (const uint32_t*)s->cnt, s->sum));
    }
    return launch(ctx, "svo_blur", dim3(stride_grid(ctx, (uint64_t)s->nodes * GUIDE_BINS, 256)), dim3(256), 0,
                  k_svo_blur, s->nodes, (const float*)s->sum, s->radiance);
}

int32_t svo_compact_records(mrh_context* ctx, mrh_svo* s, const uint32_t* codes, const uint32_t* bins,
                            const float* rad, uint32_t n, uint32_t* out_codes, uint32_t* out_bins, float* out_rad,
                            uint32_t* count)
{
    *count = 0;
    if (n == 0) return MRH_OK;
    if (!s->counter) MRH_TRY(dmalloc(ctx, (void**)&s->counter, 64));
    MRH_TRY(fill_u32(ctx, s->counter + 12, 0u, 1));
    uint32_t grid = div_up(n, APPEND_CHUNK);
    if (grid > (uint32_t)ctx->num_cus * 8u) grid = (uint32_t)ctx->num_cus * 8u;
    MRH_TRY(launch(ctx, "svo_rec_append", dim3(grid), dim3(256), 0, k_append_records, codes, bins, rad, n, out_codes,
                   out_bins, out_rad, s->counter + 12));
    MRH_HIP_CHECK(ctx, hipMemcpyAsync(count, s->counter + 12, 4, hipMemcpyDeviceToHost, ctx->stream));
    MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return MRH_OK;
}

mrh_context* svo_ctx(const mrh_svo* s) { return s->ctx; }

void svo_free(mrh_svo* s)
{
    if (!s) return;
    if (s->ctx) (void)hipStreamSynchronize(s->ctx->stream);
    void* ptrs[] = {s->codes, s->first_child, s->child_mask, s->parent, s->sum,    s->acc,
                    s->cnt,   s->radiance,    s->pack,       s->node_count, s->work_a, s->work_b,
                    s->counter, s->bin_work};
    for (void* p : ptrs) (void)hipFree(p);
    delete s;
}

namespace {

__global__ void k_svo_lookup(SvoView s, const uint32_t* __restrict__ codes, uint32_t n, uint32_t min_count,
                             uint32_t* __restrict__ out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = min_count ? svo_lookup_guide(s, codes[i], min_count) : svo_lookup(s, codes[i]);
}

__global__ void k_svo_sample(SvoView s, const uint32_t* __restrict__ nodes, const float* __restrict__ nrm,
                             const float* __restrict__ xi, uint32_t n, float* __restrict__ dirs,
                             float* __restrict__ pdf)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const V3 nn = v3(nrm[i], nrm[(size_t)n + i], nrm[2 * (size_t)n + i]);
        float pd;
        const GuideBins g = guide_bins(s, nodes[i]);
        const V3 d = guide_sample(g, nn, guide_total(g, nn), xi[i], xi[(size_t)n + i], xi[2 * (size_t)n + i], pd);
        dirs[i] = d.x;
        dirs[(size_t)n + i] = d.y;
        dirs[2 * (size_t)n + i] = d.z;
        pdf[i] = pd;
    }
}

__global__ void k_svo_pdf(SvoView s, const uint32_t* __restrict__ nodes, const float* __restrict__ nrm,
                          const float* __restrict__ dirs, uint32_t n, float* __restrict__ pdf)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const V3 nn = v3(nrm[i], nrm[(size_t)n + i], nrm[2 * (size_t)n + i]);
        const V3 d = v3(dirs[i], dirs[(size_t)n + i], dirs[2 * (size_t)n + i]);
        pdf[i] = guide_pdf(s, nodes[i], nn, guide_total(guide_bins(s, nodes[i]), nn), d);
    }
}

// ---------------------------------------------------------------- ray bins
// Start node of each ray (the deepest existing node of its cell), and the
// (key, ray) pairs of the first sort; rays without a code get key `skip`.
__global__ void k_bin_start(SvoView s, const uint32_t* __restrict__ codes, uint32_t n, uint32_t skip,
                            uint32_t* __restrict__ start, uint32_t* __restrict__ own)
{
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t c = codes[i];
        const uint32_t v = (c == SVO_NO_NODE) ? skip : svo_lookup(s, c);
        start[i] = v;
        // the node's own ray count: one atomic per distinct node of the wave
        const bool has = v != skip;
        uint64_t pend = __ballot(has);
        while (pend) {
            const int l = __ffsll((unsigned long long)pend) - 1;
            const uint32_t lv = __shfl(v, l, 64);
            const uint64_t same = __ballot(has && v == lv);
            if ((int)lane == l) atomicAdd(own + lv, (uint32_t)__popcll(same));
            pend &= ~same;
        }
    }
}

constexpr uint32_t BIN_MARK = 1u << 31;

// One level of the collapse (KCCollapseRayCounts / KCCollapseRayCountsLeaf,
// AnisoSVO.cu:265-350), as a gather: a node's rays are its own and those its
// unmarked children handed up; it is marked at min_level or with at least
// min_rays rays.  tot[v] = rays | BIN_MARK.  Integer sums: no atomics, no
// order dependence.
__global__ void k_bin_collapse(uint32_t off, uint32_t count, uint32_t at_min, uint32_t min_rays,
                               const uint32_t* __restrict__ own, const uint32_t* __restrict__ first_child,
                               const uint32_t* __restrict__ child_mask, uint32_t* __restrict__ tot)
{
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += gridDim.x * blockDim.x) {
        const uint32_t v = off + t;
        uint32_t total = own[v];
        const uint32_t fc = first_child[v];
        if (fc != SVO_NO_NODE) {
            const uint32_t nc = (uint32_t)__popc(child_mask[v]);
            for (uint32_t k = 0; k < nc; k++) {
                const uint32_t c = tot[fc + k];
                if (!(c & BIN_MARK)) total += c;
            }
        }
        const bool mark = at_min || total >= min_rays;
        tot[v] = (total & ~BIN_MARK) | (mark ? BIN_MARK : 0u);
    }
}

// Each ray climbs from its start node to the first marked node
// (AnisoSVOctreeGPU::FindMarkedBin, AnisoSVO.cuh:1244-1262); keys of the
// partition sort.
__global__ void k_bin_assign(const uint32_t* __restrict__ start, uint32_t n, uint32_t skip,
                             const uint32_t* __restrict__ tot, const uint32_t* __restrict__ parent,
                             uint32_t* __restrict__ bins, uint32_t* __restrict__ key, uint32_t* __restrict__ val)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t v = start[i];
        uint32_t b = SVO_NO_NODE;
        if (v != skip) {
            uint32_t u = v;
            while (!(tot[u] & BIN_MARK) && parent[u] != SVO_NO_NODE) u = parent[u];
            b = (tot[u] & BIN_MARK) ? u : v;
        }
        bins[i] = b;
        if (key) {
            key[i] = (b == SVO_NO_NODE) ? skip : b;
            val[i] = i;
        }
    }
}

__global__ void k_bin_flags(const uint32_t* __restrict__ skey, uint32_t n, uint32_t skip, uint32_t* __restrict__ flags)
{
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x)
        flags[p] = (skey[p] != skip && (p == 0 || skey[p - 1] != skey[p])) ? 1u : 0u;
}

// Partition runs: first position and node of each bin, the end of the
// binned rays and the run count
__global__ void k_bin_runs(const uint32_t* __restrict__ skey, const uint32_t* __restrict__ flags,
                           const uint32_t* __restrict__ pos, uint32_t n, uint32_t skip,
                           uint32_t* __restrict__ offsets, uint32_t* __restrict__ part_bins,
                           uint32_t* __restrict__ part_count)
{
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const uint32_t k = skey[p];
        if (k == skip) continue;
        if (flags[p]) {
            if (offsets) offsets[pos[p]] = p;
            if (part_bins) part_bins[pos[p]] = k;
        }
        if (p == n - 1 || skey[p + 1] == skip) {
            const uint32_t runs = pos[p] + flags[p];
            if (offsets) offsets[runs] = p + 1;
            if (part_count) *part_count = runs;
        }
    }
}

int32_t svo_bin_rays(mrh_context* ctx, mrh_svo* s, const uint32_t* codes, uint32_t n, uint32_t min_level,
                     uint32_t min_rays, uint32_t* bins, uint32_t* order, uint32_t* offsets, uint32_t* part_bins,
                     uint32_t* part_count)
{
    const uint32_t N = s->nodes;
    const uint32_t skip = N;                    // sorts after every node
    const uint32_t bits = 32u - (uint32_t)__builtin_clz(N);   // holds 0..N
    if (min_level > s->depth) min_level = s->depth;   // WFPGTracer.cu:612-613
    // workspace: start, key, val, skey, flags, pos [n]; own, tot [N]; scan tmp
    const size_t tmp_words = scan_tmp_words(n ? n : 1);
    const size_t words = 6 * (size_t)n + 2 * (size_t)N + tmp_words + 64;
    MRH_TRY(grow(ctx, &s->bin_work, &s->bin_work_bytes, words * 4));
    uint32_t* w = (uint32_t*)s->bin_work;
    uint32_t* start = w;
    uint32_t* key = start + n;
    uint32_t* val = key + n;
    uint32_t* skey = val + n;
    uint32_t* flags = skey + n;
    uint32_t* pos = flags + n;
    uint32_t* own = pos + n;
    uint32_t* tot = own + N;
    uint32_t* tmp = tot + N;
    if (order) {
        if (part_count) MRH_TRY(fill_u32(ctx, part_count, 0u, 1));
        if (offsets) MRH_TRY(fill_u32(ctx, offsets, 0u, 1));
    }
    if (n == 0) return MRH_OK;
    const uint32_t g = stride_grid(ctx, n, 256);
    const SvoView view = svo_view(s);
    MRH_TRY(fill_u32(ctx, own, 0u, N));
    MRH_TRY(fill_u32(ctx, tot, 0u, N));
    MRH_TRY(launch(ctx, "bin_start", dim3(g), dim3(256), 0, k_bin_start, view, codes, n, skip, start, own));
    for (int32_t l = (int32_t)s->depth; l >= (int32_t)min_level; l--) {
        const uint32_t cnt = s->lvl_count[l];
        if (!cnt) continue;
        MRH_TRY(launch(ctx, "bin_collapse", dim3(stride_grid(ctx, cnt, 256)), dim3(256), 0, k_bin_collapse,
                       s->lvl_off[l], cnt, (uint32_t)((uint32_t)l == min_level), min_rays, (const uint32_t*)own,
                       (const uint32_t*)s->first_child, (const uint32_t*)s->child_mask, tot));
    }
    MRH_TRY(launch(ctx, "bin_assign", dim3(g), dim3(256), 0, k_bin_assign, (const uint32_t*)start, n, skip,
                   (const uint32_t*)tot, (const uint32_t*)s->parent, bins, order ? key : nullptr, val));
    if (!order) return MRH_OK;
    SortArgs sa;
    sa.keys_in = key;
    sa.vals_in = val;
    sa.keys_out = skey;
    sa.vals_out = order;
    sa.n_max = n;
    sa.key_bytes = 4;
    sa.begin_bit = 0;
    sa.end_bit = bits;
    MRH_TRY(radix_sort(ctx, sa));
    MRH_TRY(launch(ctx, "bin_flags", dim3(g), dim3(256), 0, k_bin_flags, (const uint32_t*)skey, n, skip, flags));
    MRH_TRY(scan_exclusive(ctx, flags, pos, n, tmp));
    return launch(ctx, "bin_runs", dim3(g), dim3(256), 0, k_bin_runs, (const uint32_t*)skey, (const uint32_t*)flags,
                  (const uint32_t*)pos, n, skip, offsets, part_bins, part_count);
}

}  // namespace
}  // namespace mrh

using namespace mrh;

extern "C" {

int32_t mrh_svo_voxelize(mrh_context* ctx, const float* d_tris, uint32_t n, const float cube[4], uint32_t depth,
                         uint32_t* d_codes, uint32_t* d_prims, uint32_t capacity, uint32_t* out_count)
{
    if (!ctx || !cube || !out_count) return MRH_ERR_INVALID;
    *out_count = 0;
    if (depth < 1 || depth > 10) return fail(ctx, MRH_ERR_INVALID, "mrh_svo_voxelize: depth must be in [1, 10]");
    for (int a = 0; a < 4; a++)
        if (!std::isfinite(cube[a])) return fail(ctx, MRH_ERR_INVALID, "mrh_svo_voxelize: cube must be finite");
    if (!(cube[3] > 0.0f)) return fail(ctx, MRH_ERR_INVALID, "mrh_svo_voxelize: cube size must be > 0");
    if (n == 0) return MRH_OK;
    if (!d_tris) return fail(ctx, MRH_ERR_INVALID, "mrh_svo_voxelize: null triangles");
    if (n >= (1u << 30)) return fail(ctx, MRH_ERR_INVALID, "mrh_svo_voxelize: at most 2^30 - 1 triangles");
    // counts[n + 1] (scanned in place into offsets), scan scratch, a 64-bit total
    const size_t words = 2 + (size_t)n + 1 + scan_tmp_words(n + 1);
    void* mem = nullptr;
    MRH_TRY(scratch(ctx, SLOT_VOXEL, words * 4, &mem));
    unsigned long long* total = (unsigned long long*)mem;
    uint32_t* counts = (uint32_t*)mem + 2;
    uint32_t* tmp = counts + n + 1;
    const float4 c4 = make_float4(cube[0], cube[1], cube[2], cube[3]);
    const uint32_t grid = (uint32_t)std::min<uint64_t>(div_up(n, 4), (uint64_t)ctx->num_cus * 8u);
    MRH_TRY(fill_u32(ctx, (uint32_t*)mem, 0u, 2));
    MRH_TRY(fill_u32(ctx, counts + n, 0u, 1));
    MRH_TRY(launch(ctx, "voxelize_count", dim3(grid), dim3(256), 0, k_voxelize, d_tris, n, c4, depth,
                   (const uint32_t*)nullptr, counts, (uint32_t*)nullptr, total));
    unsigned long long h_total = 0;
    MRH_HIP_CHECK(ctx, hipMemcpyAsync(&h_total, total, 8, hipMemcpyDeviceToHost, ctx->stream));
    MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (h_total > 0xFFFFFFFFull)
        return fail(ctx, MRH_ERR_INVALID, "mrh_svo_voxelize: %llu voxel records exceed 2^32 - 1", h_total);
    *out_count = (uint32_t)h_total;
    if (!d_codes || h_total > capacity || h_total == 0) return MRH_OK;   // a size query
    MRH_TRY(scan_exclusive(ctx, counts, counts, n + 1, tmp));
    return launch(ctx, "voxelize", dim3(grid), dim3(256), 0, k_voxelize, d_tris, n, c4, depth,
                  (const uint32_t*)counts, d_codes, d_prims, (unsigned long long*)nullptr);
}

static ConeTree cone_tree(const mrh_svo* s)
{
    ConeTree t;
    t.first_child = s->first_child;
    t.child_mask = s->child_mask;
    t.parent = s->parent;
    t.depth = s->depth;
    t.lo[0] = s->cube[0];
    t.lo[1] = s->cube[1];
    t.lo[2] = s->cube[2];
    t.size = s->cube[3];
    return t;
}

int32_t mrh_svo_cone_trace(mrh_context* ctx, const mrh_svo* svo, const float* d_rays, uint32_t n, float aperture,
                           uint32_t max_query_level_offset, uint32_t* d_nodes, float* d_t, uint32_t* d_is_leaf)
{
    if (!ctx || !svo || svo->ctx != ctx) return MRH_ERR_INVALID;
    if (!svo->nodes) return fail(ctx, MRH_ERR_INVALID, "mrh_svo_cone_trace: the octree is not built");
    if (!(aperture >= 0.0f) || !std::isfinite(aperture))
        return fail(ctx, MRH_ERR_INVALID, "mrh_svo_cone_trace: aperture must be finite and >= 0");
    if (n == 0) return MRH_OK;
    if (!d_rays || !d_nodes || !d_t) return fail(ctx, MRH_ERR_INVALID, "mrh_svo_cone_trace: null rays or outputs");
    return launch(ctx, "svo_cone_trace", dim3(stride_grid(ctx, n, 256)), dim3(256), 0, k_svo_cone_trace,
                  cone_tree(svo), d_rays, n, aperture, max_query_level_offset, d_nodes, d_t, d_is_leaf);
}

int32_t mrh_svo_payload(mrh_context* ctx, const mrh_svo* svo, const mrh_svo_payload_desc* d)
{
    if (!ctx || !svo || !d || svo->ctx != ctx) return MRH_ERR_INVALID;
    if (!svo->nodes) return fail(ctx, MRH_ERR_INVALID, "mrh_svo_payload: the octree is not built");
    if (!d->node_irradiance || !d->node_normals)
        return fail(ctx, MRH_ERR_INVALID, "mrh_svo_payload: null outputs");
    if (d->n_records && (!d->tris || !d->codes || !d->prims))
        return fail(ctx, MRH_ERR_INVALID, "mrh_svo_payload: null records");
    if (d->n_vertices && !d->vertices) return fail(ctx, MRH_ERR_INVALID, "mrh_svo_payload: null vertices");
    if (d->n_records >= (1u << 30) || d->n_vertices >= (1u << 30))
        return fail(ctx, MRH_ERR_INVALID, "mrh_svo_payload: at most 2^30 - 1 records and vertices");
    const uint32_t nodes = svo->nodes, D = svo->depth, leaf0 = svo->lvl_off[D], leaves = nodes - leaf0;
    const uint32_t nr = d->n_records;
    // workspace: keys, sorted keys, order (nr each), begin/end (leaves each),
    // normals (4 per node), sums (2 x u64 per node), counts (2 per node)
    const size_t w_keys = nr ? nr : 1, w_leaves = leaves ? leaves : 1;
    const size_t words = 3 * w_keys + 2 * w_leaves + 4 * (size_t)nodes + 4 * (size_t)nodes + 2 * (size_t)nodes + 8;
    void* mem = nullptr;
    MRH_TRY(scratch(ctx, SLOT_VOXEL, words * 4, &mem));
    unsigned long long* tot = (unsigned long long*)mem;                  // 8-byte aligned first
    float4* nrm = (float4*)(tot + 2 * (size_t)nodes);
    uint32_t* num = (uint32_t*)(nrm + nodes);
    uint32_t* keys = num + 2 * (size_t)nodes;
    uint32_t* skeys = keys + w_keys;
    uint32_t* order = skeys + w_keys;
    uint32_t* begin = order + w_keys;
    uint32_t* end = begin + w_leaves;
    MRH_TRY(fill_u32(ctx, begin, 0u, 2 * w_leaves));
    if (nr) {
        MRH_TRY(launch(ctx, "payload_keys", dim3(stride_grid(ctx, nr, 256)), dim3(256), 0, k_pl_keys, d->codes, nr,
                       (const uint32_t*)svo->first_child, (const uint32_t*)svo->child_mask, D, leaf0, keys));
        SortArgs sa;
        sa.keys_in = keys;
        sa.keys_out = skeys;
        sa.vals_out = order;
        sa.n_max = nr;
        sa.key_bytes = 4;
        sa.begin_bit = 0;
        sa.end_bit = 32;
        MRH_TRY(radix_sort(ctx, sa));
        MRH_TRY(launch(ctx, "payload_runs", dim3(stride_grid(ctx, nr, 256)), dim3(256), 0, k_pl_runs,
                       (const uint32_t*)skeys, nr, begin, end));
    }
    MRH_TRY(launch(ctx, "payload_leaf_normals", dim3(stride_grid(ctx, leaves, 256)), dim3(256), 0, k_pl_leaf_normals,
                   leaves, leaf0, (const uint32_t*)begin, (const uint32_t*)end, (const uint32_t*)order, d->prims,
                   d->tris, nrm));
    for (int l = (int)D - 1; l >= 0; l--) {
        const uint32_t off = svo->lvl_off[l], cnt = svo->lvl_off[l + 1] - off;
        MRH_TRY(launch(ctx, "payload_node_normals", dim3(stride_grid(ctx, cnt, 256)), dim3(256), 0,
                       k_pl_node_normals, off, cnt, (const uint32_t*)svo->first_child,
                       (const uint32_t*)svo->child_mask, nrm));
    }
    MRH_TRY(fill_u32(ctx, (uint32_t*)tot, 0u, 4 * (size_t)nodes));
    MRH_TRY(fill_u32(ctx, num, 0u, 2 * (size_t)nodes));
    if (d->n_vertices)
        MRH_TRY(launch(ctx, "payload_splat", dim3(stride_grid(ctx, d->n_vertices, 256)), dim3(256), 0, k_pl_splat,
                       d->vertices, d->n_vertices, d->vertex_mask, (const uint32_t*)svo->first_child, (const uint32_t*)svo->child_mask,
                       D, make_float4(svo->cube[0], svo->cube[1], svo->cube[2], svo->cube[3]), (const float4*)nrm, tot,
                       num));
    float2* irr = (float2*)d->node_irradiance;
    MRH_TRY(launch(ctx, "payload_leaf_avg", dim3(stride_grid(ctx, leaves, 256)), dim3(256), 0, k_pl_leaf_avg, leaf0,
                   leaves, (const unsigned long long*)tot, (const uint32_t*)num, irr));
    for (int l = (int)D - 1; l >= 0; l--) {
        const uint32_t off = svo->lvl_off[l], cnt = svo->lvl_off[l + 1] - off;
        MRH_TRY(launch(ctx, "payload_bottom_up", dim3(stride_grid(ctx, cnt, 256)), dim3(256), 0, k_pl_bottom_up, off,
                       cnt, (const uint32_t*)svo->first_child, (const uint32_t*)svo->child_mask, (const float4*)nrm,
                       irr));
    }
    for (uint32_t l = 1; l <= D; l++) {
        const uint32_t off = svo->lvl_off[l], cnt = svo->lvl_off[l + 1] - off;
        MRH_TRY(launch(ctx, "payload_top_down", dim3(stride_grid(ctx, cnt, 256)), dim3(256), 0, k_pl_top_down, off,
                       cnt, (const uint32_t*)svo->parent, irr));
    }
    return launch(ctx, "payload_out", dim3(stride_grid(ctx, nodes, 256)), dim3(256), 0, k_pl_out, nodes,
                  (const float4*)nrm, d->node_normals, d->node_normal_length);
}

int32_t mrh_wfpg_trace_fields(mrh_context* ctx, const mrh_svo* svo, const mrh_wfpg_trace_desc* d)
{
    if (!ctx || !svo || !d || svo->ctx != ctx) return MRH_ERR_INVALID;
    if (!svo->nodes) return fail(ctx, MRH_ERR_INVALID, "mrh_wfpg_trace_fields: the octree is not built");
    const auto side_ok = [](uint32_t s) { return s == 8 || s == 16 || s == 32 || s == 64; };
    if (!side_ok(d->field_x) || !side_ok(d->field_y))
        return fail(ctx, MRH_ERR_INVALID, "mrh_wfpg_trace_fields: field sides must be 8, 16, 32 or 64");
    if (!std::isfinite(d->boundary_radiance))
        return fail(ctx, MRH_ERR_INVALID, "mrh_wfpg_trace_fields: boundary radiance must be finite");
    if (d->bin_count == 0) return MRH_OK;
    if (!d->bin_origins || !d->bin_jitter || !d->node_irradiance || !d->node_normals || !d->out_fields)
        return fail(ctx, MRH_ERR_INVALID, "mrh_wfpg_trace_fields: null input or output");
    // OctohedralConeAperture (WFPGTracer.cu:134-141): 4 pi over the texel
    // count, which the reference forms as field_x * field_x
    const float aperture = (4.0f * 3.14159265358979323846f) / ((float)d->field_x * (float)d->field_x);
    const uint64_t total = (uint64_t)d->bin_count * d->field_x * d->field_y;
    return launch(ctx, "wfpg_trace_fields", dim3(stride_grid(ctx, total, 256)), dim3(256), 0, k_wfpg_trace_fields,
                  cone_tree(svo), d->field_x, d->field_y, (uint32_t)__builtin_ctz(d->field_x), d->bin_count,
                  (const float4*)d->bin_origins, (const float2*)d->bin_jitter, aperture, d->max_query_level_offset,
                  (const float2*)d->node_irradiance, d->node_normals, d->boundary_radiance, d->out_fields,
                  d->out_nodes, d->out_t);
}

int32_t mrh_svo_bin_rays(mrh_context* ctx, mrh_svo* svo, const uint32_t* d_codes, uint32_t n, uint32_t min_level,
                         uint32_t bin_ray_count, uint32_t* d_bins, uint32_t* d_order, uint32_t* d_part_offsets,
                         uint32_t* d_part_bins, uint32_t* d_part_count)
{
    if (!ctx || !svo || svo->ctx != ctx) return MRH_ERR_INVALID;
    if (!svo->nodes) return fail(ctx, MRH_ERR_INVALID, "mrh_svo_bin_rays: the octree is not built");
    if (bin_ray_count < 1 || bin_ray_count > MRH_BIN_RAY_COUNT_MAX)
        return fail(ctx, MRH_ERR_INVALID, "mrh_svo_bin_rays: bin_ray_count must be in [1, %u]",
                    (unsigned)MRH_BIN_RAY_COUNT_MAX);
    if (n >= (1u << 30)) return fail(ctx, MRH_ERR_INVALID, "mrh_svo_bin_rays: at most 2^30 - 1 rays");
    if (n && (!d_codes || !d_bins)) return fail(ctx, MRH_ERR_INVALID, "mrh_svo_bin_rays: null codes or bins");
    if (!d_order && (d_part_offsets || d_part_bins || d_part_count))
        return fail(ctx, MRH_ERR_INVALID, "mrh_svo_bin_rays: the partition needs d_order");
    return svo_bin_rays(ctx, svo, d_codes, n, min_level, bin_ray_count, d_bins, d_order, d_part_offsets,
                        d_part_bins, d_part_count);
}

int32_t mrh_svo_create(mrh_context* ctx, mrh_svo** out)
{
    if (!ctx || !out) return MRH_ERR_INVALID;
    mrh_svo* s = new mrh_svo();
    s->ctx = ctx;
    *out = s;
    return MRH_OK;
}

int32_t mrh_svo_destroy(mrh_svo* svo)
{
    if (!svo) return MRH_ERR_INVALID;
    svo_free(svo);
    return MRH_OK;
}

int32_t mrh_svo_codes(mrh_context* ctx, const float* d_points, uint32_t n, uint32_t pitch, const float cube[4],
                      uint32_t depth, uint32_t* d_codes)
{
    if (!ctx || !cube) return MRH_ERR_INVALID;
    if (n == 0) return MRH_OK;
    if (!d_points || !d_codes || pitch < n) return fail(ctx, MRH_ERR_INVALID, "mrh_svo_codes: bad arguments");
    if (depth < 1 || depth > 10) return fail(ctx, MRH_ERR_INVALID, "svo depth must be in [1, 10]");
    return svo_codes(ctx, d_points, n, pitch, cube, depth, d_codes);
}

int32_t mrh_svo_build(mrh_context* ctx, mrh_svo* svo, const uint32_t* d_codes, uint32_t n, const float cube[4],
                      uint32_t depth)
{
    if (!ctx || !svo || !cube || svo->ctx != ctx) return MRH_ERR_INVALID;
    if (n && !d_codes) return fail(ctx, MRH_ERR_INVALID, "mrh_svo_build: null codes");
    return svo_build(ctx, svo, d_codes, n, cube, depth);
}

int32_t mrh_svo_splat(mrh_context* ctx, mrh_svo* svo, const uint32_t* d_codes, const uint32_t* d_bins,
                      const float* d_values, uint32_t n)
{
    if (!ctx || !svo || svo->ctx != ctx) return MRH_ERR_INVALID;
    if (n && (!d_codes || !d_bins || !d_values)) return fail(ctx, MRH_ERR_INVALID, "mrh_svo_splat: null input");
    return svo_splat(ctx, svo, d_codes, d_bins, d_values, n);
}

int32_t mrh_svo_filter(mrh_context* ctx, mrh_svo* svo)
{
    if (!ctx || !svo || svo->ctx != ctx) return MRH_ERR_INVALID;
    return svo_filter(ctx, svo);
}

int32_t mrh_svo_lookup(mrh_context* ctx, const mrh_svo* svo, const uint32_t* d_codes, uint32_t n, uint32_t min_count,
                       uint32_t* d_nodes)
{
    if (!ctx || !svo) return MRH_ERR_INVALID;
    if (n == 0) return MRH_OK;
    if (!svo->nodes || !d_codes || !d_nodes) return fail(ctx, MRH_ERR_INVALID, "mrh_svo_lookup: bad arguments");
    return launch(ctx, "svo_lookup", dim3(stride_grid(ctx, n, 256)), dim3(256), 0, k_svo_lookup, svo_view(svo),
                  d_codes, n, min_count, d_nodes);
}

int32_t mrh_svo_sample(mrh_context* ctx, const mrh_svo* svo, const uint32_t* d_nodes, const float* d_normals,
                       const float* d_xi, uint32_t n, float* d_dirs, float* d_pdf)
{
    if (!ctx || !svo) return MRH_ERR_INVALID;
    if (n == 0) return MRH_OK;
    if (!svo->nodes || !d_nodes || !d_normals || !d_xi || !d_dirs || !d_pdf)
        return fail(ctx, MRH_ERR_INVALID, "mrh_svo_sample: bad arguments");
    return launch(ctx, "svo_sample", dim3(stride_grid(ctx, n, 256)), dim3(256), 0, k_svo_sample, svo_view(svo),
                  d_nodes, d_normals, d_xi, n, d_dirs, d_pdf);
}

int32_t mrh_svo_pdf(mrh_context* ctx, const mrh_svo* svo, const uint32_t* d_nodes, const float* d_normals,
                    const float* d_dirs, uint32_t n, float* d_pdf)
{
    if (!ctx || !svo) return MRH_ERR_INVALID;
    if (n == 0) return MRH_OK;
    if (!svo->nodes || !d_nodes || !d_normals || !d_dirs || !d_pdf)
        return fail(ctx, MRH_ERR_INVALID, "mrh_svo_pdf: bad arguments");
    return launch(ctx, "svo_pdf", dim3(stride_grid(ctx, n, 256)), dim3(256), 0, k_svo_pdf, svo_view(svo), d_nodes,
                  d_normals, d_dirs, n, d_pdf);
}

int32_t mrh_svo_view_get(const mrh_svo* s, mrh_svo_view* v)
{
    if (!s || !v) return MRH_ERR_INVALID;
    v->depth = s->depth;
    v->nodes = s->nodes;
    for (int i = 0; i < 11; i++) v->level_count[i] = s->lvl_count[i];
    for (int i = 0; i < 12; i++) v->level_offset[i] = s->lvl_off[i];
    for (int i = 0; i < 4; i++) v->cube[i] = s->cube[i];
    v->d_codes = s->codes;
    v->d_first_child = s->first_child;
    v->d_child_mask = s->child_mask;
    v->d_parent = s->parent;
    v->d_sum = s->sum;
    v->d_count = s->cnt;
    v->d_radiance = s->radiance;
    v->d_node_count = s->node_count;
    return MRH_OK;
}

}  // extern "C"
