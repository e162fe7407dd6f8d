// mrh_guide.hpp — sparse voxel octree (SVO) radiance cache for path guiding:
// device-side lookup and product sampling shared by the guide kernels and
// the guided shade kernel.
//
// Directions are binned on an 8x8 grid of the equal-area concentric
// octahedral map (Clarberg 2008), the parameterisation the reference's
// guided tracer uses for its radiance fields
// (RayLib/CoordinateConversion.h:112-180): uniform (u,v) maps to uniform
// solid angle, so every bin spans 4*pi/64 steradians.
#pragma once

#include "mrh_internal.hpp"

namespace mrh {

constexpr uint32_t GUIDE_RES = 8;                  // bins per axis
constexpr uint32_t GUIDE_BINS = GUIDE_RES * GUIDE_RES;
constexpr uint32_t SVO_NO_NODE = 0xFFFFFFFFu;
constexpr float GUIDE_EPS = 1.0e-4f;               // radiance floor of a bin

// Splatted radiance is summed in 48.16 fixed point so that the sums, and every
// guide built from them, do not depend on the order of the atomic adds.
// Samples are clamped to [0, 32767].
constexpr float GUIDE_FIX_ONE = 65536.0f;
constexpr float GUIDE_FIX_MAX = 32767.0f;
MRH_HD uint64_t guide_fix(float v)
{
    return (v > 0.0f) ? (uint64_t)(fminf(v, GUIDE_FIX_MAX) * GUIDE_FIX_ONE) : 0ull;
}

// (u,v) in [0,1]^2 -> unit direction, z up
MRH_HD V3 octa_to_dir(float u, float v)
{
    const float a = 2.0f * u - 1.0f, b = 2.0f * v - 1.0f;
    const float ua = fabsf(a), va = fabsf(b);
    const float d = 1.0f - (ua + va);
    const float r = 1.0f - fabsf(d);
    const float phin = (r == 0.0f) ? 0.0f : ((va - ua) / r + 1.0f) * 0.125f;   // phi / (2 pi)
    const float z = copysignf(1.0f - r * r, d);
    float s, c;
    sincos_2pi(phin, s, c);
    const float k = r * sqrtf(fmaxf(0.0f, 2.0f - r * r));
    return v3(copysignf(k * c, a), copysignf(k * s, b), z);
}

// atan2(y, x) / (pi / 2) for y, x >= 0: the angle as a fraction of a quadrant
MRH_HD float quadrant_angle(float y, float x)
{
    if (y == 0.0f && x == 0.0f) return 0.0f;
    if (y <= x) return atan_unit(y / x) * 0.63661977236758134f;
    return 1.0f - atan_unit(x / y) * 0.63661977236758134f;
}

// unit direction -> (u,v)
MRH_HD void dir_to_octa(V3 w, float& u, float& v)
{
    const float r = sqrtf(fmaxf(0.0f, 1.0f - fabsf(w.z)));
    const float t = quadrant_angle(fabsf(w.y), fabsf(w.x));   // 2 phi / pi
    float b = r * t, a = r - b;
    if (w.z < 0.0f) {
        const float a2 = 1.0f - b, b2 = 1.0f - a;
        a = a2;
        b = b2;
    }
    u = (copysignf(a, w.x) + 1.0f) * 0.5f;
    v = (copysignf(b, w.y) + 1.0f) * 0.5f;
}

MRH_HD uint32_t dir_to_bin(V3 w)
{
    float u, v;
    dir_to_octa(w, u, v);
    uint32_t ix = (uint32_t)(u * (float)GUIDE_RES), iy = (uint32_t)(v * (float)GUIDE_RES);
    ix = ix > GUIDE_RES - 1 ? GUIDE_RES - 1 : ix;
    iy = iy > GUIDE_RES - 1 ? GUIDE_RES - 1 : iy;
    return iy * GUIDE_RES + ix;
}

// Directions of the 64 bin centres, octa_to_dir((ix + 1/2) / 8, (iy + 1/2) / 8)
// for bin iy * 8 + ix, as the exact float32 values that function returns (the
// tests recompute them).  Written out so that the unrolled product loops use
// them as immediate operands instead of loading them.
constexpr float GUIDE_CENTERS[GUIDE_BINS][3] = {
    {-0x1.f7efc0p-3f, -0x1.f7efc0p-3f, -0x1.ep-1f},
    {-0x1.03323cp-2f, -0x1.38e0a2p-1f, -0x1.8p-1f},
    {-0x1.dca414p-3f, -0x1.bcb64cp-1f, -0x1.cp-2f},
    {-0x1.8f8b84p-3f, -0x1.f6297cp-1f, 0x0.0p+0f},
    {0x1.8f8b84p-3f, -0x1.f6297cp-1f, 0x0.0p+0f},
    {0x1.dca414p-3f, -0x1.bcb64cp-1f, -0x1.cp-2f},
    {0x1.03323cp-2f, -0x1.38e0a2p-1f, -0x1.8p-1f},
    {0x1.f7efc0p-3f, -0x1.f7efc0p-3f, -0x1.ep-1f},
    {-0x1.38e0a2p-1f, -0x1.03323cp-2f, -0x1.8p-1f},
    {-0x1.458d48p-1f, -0x1.458d48p-1f, -0x1.cp-2f},
    {-0x1.1c73b4p-1f, -0x1.a9b662p-1f, 0x0.0p+0f},
    {-0x1.dca414p-3f, -0x1.bcb64cp-1f, 0x1.cp-2f},
    {0x1.dca414p-3f, -0x1.bcb64cp-1f, 0x1.cp-2f},
    {0x1.1c73b4p-1f, -0x1.a9b662p-1f, 0x0.0p+0f},
    {0x1.458d48p-1f, -0x1.458d48p-1f, -0x1.cp-2f},
    {0x1.38e0a2p-1f, -0x1.03323cp-2f, -0x1.8p-1f},
    {-0x1.bcb64cp-1f, -0x1.dca414p-3f, -0x1.cp-2f},
    {-0x1.a9b662p-1f, -0x1.1c73b4p-1f, 0x0.0p+0f},
    {-0x1.458d48p-1f, -0x1.458d48p-1f, 0x1.cp-2f},
    {-0x1.03323cp-2f, -0x1.38e0a2p-1f, 0x1.8p-1f},
    {0x1.03323cp-2f, -0x1.38e0a2p-1f, 0x1.8p-1f},
    {0x1.458d48p-1f, -0x1.458d48p-1f, 0x1.cp-2f},
    {0x1.a9b662p-1f, -0x1.1c73b4p-1f, 0x0.0p+0f},
    {0x1.bcb64cp-1f, -0x1.dca414p-3f, -0x1.cp-2f},
    {-0x1.f6297cp-1f, -0x1.8f8b84p-3f, 0x0.0p+0f},
    {-0x1.bcb64cp-1f, -0x1.dca414p-3f, 0x1.cp-2f},
    {-0x1.38e0a2p-1f, -0x1.03323cp-2f, 0x1.8p-1f},
    {-0x1.f7efc0p-3f, -0x1.f7efc0p-3f, 0x1.ep-1f},
    {0x1.f7efc0p-3f, -0x1.f7efc0p-3f, 0x1.ep-1f},
    {0x1.38e0a2p-1f, -0x1.03323cp-2f, 0x1.8p-1f},
    {0x1.bcb64cp-1f, -0x1.dca414p-3f, 0x1.cp-2f},
    {0x1.f6297cp-1f, -0x1.8f8b84p-3f, 0x0.0p+0f},
    {-0x1.f6297cp-1f, 0x1.8f8b84p-3f, 0x0.0p+0f},
    {-0x1.bcb64cp-1f, 0x1.dca414p-3f, 0x1.cp-2f},
    {-0x1.38e0a2p-1f, 0x1.03323cp-2f, 0x1.8p-1f},
    {-0x1.f7efc0p-3f, 0x1.f7efc0p-3f, 0x1.ep-1f},
    {0x1.f7efc0p-3f, 0x1.f7efc0p-3f, 0x1.ep-1f},
    {0x1.38e0a2p-1f, 0x1.03323cp-2f, 0x1.8p-1f},
    {0x1.bcb64cp-1f, 0x1.dca414p-3f, 0x1.cp-2f},
    {0x1.f6297cp-1f, 0x1.8f8b84p-3f, 0x0.0p+0f},
    {-0x1.bcb64cp-1f, 0x1.dca414p-3f, -0x1.cp-2f},
    {-0x1.a9b662p-1f, 0x1.1c73b4p-1f, 0x0.0p+0f},
    {-0x1.458d48p-1f, 0x1.458d48p-1f, 0x1.cp-2f},
    {-0x1.03323cp-2f, 0x1.38e0a2p-1f, 0x1.8p-1f},
    {0x1.03323cp-2f, 0x1.38e0a2p-1f, 0x1.8p-1f},
    {0x1.458d48p-1f, 0x1.458d48p-1f, 0x1.cp-2f},
    {0x1.a9b662p-1f, 0x1.1c73b4p-1f, 0x0.0p+0f},
    {0x1.bcb64cp-1f, 0x1.dca414p-3f, -0x1.cp-2f},
    {-0x1.38e0a2p-1f, 0x1.03323cp-2f, -0x1.8p-1f},
    {-0x1.458d48p-1f, 0x1.458d48p-1f, -0x1.cp-2f},
    {-0x1.1c73b4p-1f, 0x1.a9b662p-1f, 0x0.0p+0f},
    {-0x1.dca414p-3f, 0x1.bcb64cp-1f, 0x1.cp-2f},
    {0x1.dca414p-3f, 0x1.bcb64cp-1f, 0x1.cp-2f},
    {0x1.1c73b4p-1f, 0x1.a9b662p-1f, 0x0.0p+0f},
    {0x1.458d48p-1f, 0x1.458d48p-1f, -0x1.cp-2f},
    {0x1.38e0a2p-1f, 0x1.03323cp-2f, -0x1.8p-1f},
    {-0x1.f7efc0p-3f, 0x1.f7efc0p-3f, -0x1.ep-1f},
    {-0x1.03323cp-2f, 0x1.38e0a2p-1f, -0x1.8p-1f},
    {-0x1.dca414p-3f, 0x1.bcb64cp-1f, -0x1.cp-2f},
    {-0x1.8f8b84p-3f, 0x1.f6297cp-1f, 0x0.0p+0f},
    {0x1.8f8b84p-3f, 0x1.f6297cp-1f, 0x0.0p+0f},
    {0x1.dca414p-3f, 0x1.bcb64cp-1f, -0x1.cp-2f},
    {0x1.03323cp-2f, 0x1.38e0a2p-1f, -0x1.8p-1f},
    {0x1.f7efc0p-3f, 0x1.f7efc0p-3f, -0x1.ep-1f},
};

// Voxel code of a point: 30-bit Morton code of its cell at depth `depth`
// (<= 10) of the cube [lo, lo + size)^3; points outside clamp to the border.
MRH_HD uint32_t svo_code(V3 p, float lox, float loy, float loz, float inv_cell, uint32_t depth)
{
    const uint32_t qmax = (1u << depth) - 1u;
    const float f[3] = {(p.x - lox) * inv_cell, (p.y - loy) * inv_cell, (p.z - loz) * inv_cell};
    uint32_t q[3];
    for (int a = 0; a < 3; a++) {
        float g = floorf(f[a]);
        q[a] = (g <= 0.0f) ? 0u : (g >= (float)qmax ? qmax : (uint32_t)g);
    }
    return expand10(q[0]) | (expand10(q[1]) << 1) | (expand10(q[2]) << 2);
}

// Read-only view of a built SVO for device code.  Level l holds the unique
// prefixes code >> 3(depth - l); node records of all levels are
// concatenated (level offsets in lvl_off).
struct SvoView {
    uint32_t depth;
    uint32_t lvl_off[12];            // first global node of each level, [depth+1] = total
    const uint4* pack;               // [nodes] first child, child mask, samples in the subtree, 0
    const float* radiance;           // [nodes][64] filtered radiance
    float lox, loy, loz, inv_cell;   // cube of the leaf grid
};

// Deepest node on the path from the root to `code`'s leaf cell whose
// subtree holds >= min_count samples (min_count = 0: the deepest existing
// node).  With min_count > 0 this is the guide lookup: the role of the
// reference's collapse of sparse bins to their ancestors
// (AnisoSVO.cu:1485-1529, BinRayCount); the root when no child qualifies.
// One 16-byte record per level.
__device__ __forceinline__ uint32_t svo_lookup_guide(const SvoView& s, uint32_t code, uint32_t min_count)
{
    uint32_t node = 0;   // root
    uint4 rec = s.pack[0];
    for (uint32_t l = 0; l < s.depth; l++) {
        const uint32_t m = rec.y;
        const uint32_t oct = (code >> (3u * (s.depth - 1u - l))) & 7u;
        if (!((m >> oct) & 1u)) break;
        const uint32_t child = rec.x + (uint32_t)__popc(m & ((1u << oct) - 1u));
        const uint4 crec = s.pack[child];
        if (crec.z < min_count) break;
        node = child;
        rec = crec;
    }
    return node;
}

__device__ __forceinline__ uint32_t svo_lookup(const SvoView& s, uint32_t code) { return svo_lookup_guide(s, code, 0u); }

// A node's 64 filtered radiances, held in registers
struct GuideBins {
    float4 r[GUIDE_BINS / 4];
};

__device__ __forceinline__ GuideBins guide_bins(const SvoView& s, uint32_t node)
{
    const float4* p = (const float4*)(s.radiance + (size_t)node * GUIDE_BINS);
    GuideBins g;
#pragma unroll
    for (int k = 0; k < (int)GUIDE_BINS / 4; k++) g.r[k] = p[k];
    return g;
}

template <int B>
__device__ __forceinline__ float bin_radiance(const GuideBins& g)
{
    const float4 v = g.r[B >> 2];
    return (B & 3) == 0 ? v.x : ((B & 3) == 1 ? v.y : ((B & 3) == 2 ? v.z : v.w));
}

// Product of bin B's filtered radiance and the clamped cosine of its centre
// around n
template <int B>
__device__ __forceinline__ float guide_weight(const GuideBins& g, V3 n)
{
    return bin_radiance<B>(g) * fmaxf(dot(n, v3(GUIDE_CENTERS[B][0], GUIDE_CENTERS[B][1], GUIDE_CENTERS[B][2])), 0.0f);
}

template <int B>
__device__ __forceinline__ void guide_total_acc(const GuideBins& g, V3 n, float& t)
{
    t += guide_weight<B>(g, n);
    if constexpr (B + 1 < (int)GUIDE_BINS) guide_total_acc<B + 1>(g, n, t);
}

// Sum of the product weights over the 64 bins, in bin order
__device__ __forceinline__ float guide_total(const GuideBins& g, V3 n)
{
    float t = 0.0f;
    guide_total_acc<0>(g, n, t);
    return t;
}

// Solid-angle pdf of one bin's share of the product (equal-area bins)
constexpr float GUIDE_BIN_SR_INV = (float)GUIDE_BINS * 0.07957747154594767f;   // 64 / (4 pi)

// running-sum scan: the first bin (of positive weight) where the sum
// exceeds target, else the last bin of positive weight
template <int B>
__device__ __forceinline__ void guide_pick(const GuideBins& g, V3 n, float target, bool& found, float& acc,
                                           uint32_t& pick, float& wp)
{
    const float w = guide_weight<B>(g, n);
    if (!found && w > 0.0f) {
        pick = B;
        wp = w;
        acc += w;
        found = acc > target;
    }
    if constexpr (B + 1 < (int)GUIDE_BINS) guide_pick<B + 1>(g, n, target, found, acc, pick, wp);
}

// Samples a bin with probability proportional to its product weight, then a
// uniform point of the bin in the map.  pdf = 0 when the product vanishes.
// `total` is guide_total(g, n).
__device__ __forceinline__ V3 guide_sample(const GuideBins& g, V3 n, float total, float xi0, float xi1, float xi2,
                                           float& pdf)
{
    if (!(total > 0.0f)) {
        pdf = 0.0f;
        return v3(0.0f, 0.0f, 1.0f);
    }
    const float target = xi0 * total;
    float acc = 0.0f, wp = 0.0f;
    uint32_t pick = 0;
    bool found = false;
    guide_pick<0>(g, n, target, found, acc, pick, wp);
    const float u = ((float)(pick % GUIDE_RES) + xi1) / (float)GUIDE_RES;
    const float v = ((float)(pick / GUIDE_RES) + xi2) / (float)GUIDE_RES;
    pdf = wp / total * GUIDE_BIN_SR_INV;
    return octa_to_dir(u, v);
}

// pdf of direction d under the node's product distribution
__device__ __forceinline__ float guide_pdf(const SvoView& s, uint32_t node, V3 n, float total, V3 d)
{
    if (!(total > 0.0f)) return 0.0f;
    const uint32_t b = dir_to_bin(d);
    const float w = s.radiance[(size_t)node * GUIDE_BINS + b] *
                    fmaxf(dot(n, v3(GUIDE_CENTERS[b][0], GUIDE_CENTERS[b][1], GUIDE_CENTERS[b][2])), 0.0f);
    return w / total * GUIDE_BIN_SR_INV;
}

// host-side handles shared by the guide and path tracer units
SvoView svo_view(const mrh_svo* s);
int32_t svo_codes(mrh_context* ctx, const float* d_points, uint32_t n, uint32_t pitch, const float cube[4],
                  uint32_t depth, uint32_t* d_codes);
// filtered = true: a splat and a filter follow, which rewrite the means,
// the filtered radiance and the node counts, so only the splat
// accumulators are cleared
int32_t svo_build(mrh_context* ctx, mrh_svo* s, const uint32_t* d_codes, uint32_t n, const float cube[4],
                  uint32_t depth, bool filtered = false);
int32_t svo_splat(mrh_context* ctx, mrh_svo* s, const uint32_t* codes, const uint32_t* bins, const float* values,
                  uint32_t n);
int32_t svo_filter(mrh_context* ctx, mrh_svo* s);
// Compaction of the learning records whose code is not SVO_NO_NODE (in no
// particular order); *count receives the number kept (synchronous).
int32_t svo_compact_records(mrh_context* ctx, mrh_svo* s, const uint32_t* codes, const uint32_t* bins,
                            const float* rad, uint32_t n, uint32_t* out_codes, uint32_t* out_bins, float* out_rad,
                            uint32_t* count);
void svo_free(mrh_svo* s);
mrh_context* svo_ctx(const mrh_svo* s);

}  // namespace mrh
