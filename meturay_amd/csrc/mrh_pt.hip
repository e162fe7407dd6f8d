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
#include "mrh_bvh.hpp"
#include "mrh_guide.hpp"

#include <cfloat>
#include <cstring>

struct mrh_scene {
    mrh_context* ctx = nullptr;
    mrh_lbvh* bvh = nullptr;
    float* lambert = nullptr;
    float* metal = nullptr;
    float* dielectric = nullptr;
    float* emissive = nullptr;
    uint32_t n_lambert = 0, n_metal = 0, n_dielectric = 0, n_emissive = 0;
    uint64_t bytes = 0;
    std::vector<void*> allocs;
};

namespace mrh {
namespace {

constexpr float INV_PI = 0.31830988618379067154f;
constexpr float PI_F = 3.14159265358979323846f;

struct Cam {
    V3 pos, bl, right, up;
    float pw, ph, near_d, far_d;
};

struct PathBuf {
    float *ox, *oy, *oz, *dx, *dy, *dz;
    float *tr, *tg, *tb;
    uint32_t* sample;
    uint32_t* dm;        // depth (bits 0..7) | medium << 8
    uint64_t* rng;
};

struct HitBuf {
    uint32_t* leaf;
    float *t, *ba, *bb;
    uint32_t* key;
};

struct Materials {
    const float* lambert;
    const float* metal;
    const float* dielectric;
    const float* emissive;
};

// GPUCameraPinhole.cuh:140-170, evaluated on the host with the same float
// operations as the oracle.
Cam make_camera(const mrh_camera_desc& c, uint32_t w, uint32_t h)
{
    float fov_y = (float)(2.0 * atan(tan((double)c.fov_x * 0.5) * (double)h / (double)w));
    V3 p = v3(c.position[0], c.position[1], c.position[2]);
    V3 g = v3(c.gaze[0], c.gaze[1], c.gaze[2]);
    V3 u = v3(c.up[0], c.up[1], c.up[2]);
    float wh = tanf(c.fov_x * 0.5f) * c.near_plane;
    float hh = tanf(fov_y * 0.5f) * c.near_plane;
    V3 gd = g - p;
    V3 right = normalize(cross(gd, u));
    u = normalize(cross(right, gd));
    gd = normalize(cross(u, right));
    Cam cam;
    cam.pos = p;
    cam.bl = ((p - scale(right, wh)) - scale(u, hh)) + scale(gd, c.near_plane);
    cam.right = right;
    cam.up = u;
    cam.pw = wh * 2.0f;
    cam.ph = hh * 2.0f;
    cam.near_d = c.near_plane;
    cam.far_d = c.far_plane;
    return cam;
}

// GenerationKernels.cuh:29-108 + GPUCameraPinhole.cuh:209-240
__global__ __launch_bounds__(256) void k_raygen(Cam cam, uint32_t w, uint32_t h, uint32_t S,
                                                uint32_t seed, uint32_t sample_base, uint32_t n_total,
                                                PathBuf out)
{
    const uint32_t spp = S * S;
    const float dxp = cam.pw / (float)(w * S);
    const float dyp = cam.ph / (float)(h * S);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_total; i += gridDim.x * blockDim.x) {
        const uint32_t pix = i / spp, s = i - pix * spp;
        const uint32_t px = pix % w, py = pix / w, sx = s % S, sy = s / S;
        const uint32_t gid = sample_base + i;   // global sample id keys the stream
        Rng rng{path_rng_init(seed, gid), pcg32_inc(gid)};
        const float jx = rng.uniform();
        const float jy = rng.uniform();
        float sdx = (float)(px * S + sx) * dxp;
        float sdy = (float)(py * S + sy) * dyp;
        sdx += jx * dxp;
        sdy += jy * dyp;
        const V3 sp = cam.bl + (scale(cam.right, sdx) + scale(cam.up, sdy));
        const V3 d = normalize(sp - cam.pos);
        out.ox[i] = cam.pos.x; out.oy[i] = cam.pos.y; out.oz[i] = cam.pos.z;
        out.dx[i] = d.x; out.dy[i] = d.y; out.dz[i] = d.z;
        out.tr[i] = 1.0f; out.tg[i] = 1.0f; out.tb[i] = 1.0f;
        out.sample[i] = i;
        out.dm[i] = 1u;
        out.rng[i] = rng.state;
    }
}

// counts layout: [0, 64) live rays per bounce, [64, 66) u64 node visits,
// [66, 68) u64 triangle tests
constexpr uint32_t COUNTS_WORDS = 128;

template <int CAP, bool STATS, int L>
__global__ __launch_bounds__(256) void k_trace_pt(const float4* __restrict__ tnodes,
                                                  const float4* __restrict__ leaves, uint32_t root, uint32_t n_leaves,
                                                  PathBuf in, uint32_t* __restrict__ counts,
                                                  uint32_t bounce, float tmin, float tmax, HitBuf hit,
                                                  uint2* __restrict__ overflow)
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
        if constexpr (L > 0) {
            LdsStack<L, CAP, 256> st{s_node, s_t, threadIdx.x, overflow, ostride, oslot};
            r = traverse<false, STATS>(st, tnodes, leaves, root, n_leaves, o, d, tmin, tmax);
        } else {
            PrivateStack<CAP> st;
            r = traverse<false, STATS>(st, tnodes, leaves, root, n_leaves, o, d, tmin, tmax);
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

template <int CAP, bool STATS>
int32_t launch_trace_pt(mrh_context* ctx, uint32_t grid, const mrh_lbvh* b, const PathBuf& in,
                        uint32_t* counts, uint32_t bounce, float tmin, float tmax, const HitBuf& hit)
{
    const uint32_t root = root_code(b);
    const uint32_t lds = ctx->trace_lds_entries;
    uint2* ovf = nullptr;
    MRH_TRY(trace_overflow(ctx, grid, CAP, lds, &ovf));
    const char* name = STATS ? "trace_pt_stats" : "trace_pt";
    const float4* tn = (const float4*)b->tnodes;
    const float4* lv = (const float4*)b->leaves;
#define MRH_TRACE_PT_L(LV)                                                                         \
    return launch(ctx, name, dim3(grid), dim3(256), 0, k_trace_pt<CAP, STATS, LV>, tn, lv, root, b->n, \
                  in, counts, bounce, tmin, tmax, hit, ovf)
    switch (lds) {
    case 0: MRH_TRACE_PT_L(0);
    case 8: MRH_TRACE_PT_L(8);
    case 12: MRH_TRACE_PT_L(12);
    case 16: MRH_TRACE_PT_L(16);
    case 24: MRH_TRACE_PT_L(24);
    default: return fail(ctx, MRH_ERR_INVALID, "trace_lds_entries must be 0, 8, 12, 16 or 24");
    }
#undef MRH_TRACE_PT_L
}

// ------------------------------------------------------------------ BxDFs
// TracerFunctions.cuh:8-30
__device__ __forceinline__ float fresnel_dielectric(float cos_in, float ior_in, float ior_out)
{
    float sin_in = sqrtf(fmaxf(0.0f, 1.0f - cos_in * cos_in));
    float sin_out = ior_in / ior_out * sin_in;
    if (sin_out >= 1.0f) return 1.0f;
    float cos_out = sqrtf(fmaxf(0.0f, 1.0f - sin_out * sin_out));
    float par = ((ior_out * cos_in - ior_in * cos_out) / (ior_out * cos_in + ior_in * cos_out));
    par = par * par;
    float perp = ((ior_in * cos_in - ior_out * cos_out) / (ior_in * cos_in + ior_out * cos_out));
    perp = perp * perp;
    return (par + perp) * 0.5f;
}

// TracerFunctions.cuh:33-79, one channel
__device__ __forceinline__ float fresnel_conductor(float cos_in, float eta, float k)
{
    float cos2 = cos_in * cos_in;
    float sin2 = fmaxf(0.0f, 1.0f - cos2);
    float sin2_2 = sin2 * sin2;
    float sin2_4 = sin2_2 * sin2_2;
    float eta2 = eta * eta;
    float k2 = k * k;
    float diff = eta2 - k2 - sin2;
    float a2b2 = diff * diff;
    a2b2 += 4.0f * eta2 * k2;
    a2b2 = sqrtf(fmaxf(a2b2, 0.0f));
    float a = 0.5f * (a2b2 + diff);
    a = sqrtf(fmaxf(a, 0.0f));
    float pt1 = a2b2 + cos2;
    float pt2 = a * (2.0f * cos_in);
    float perp = (pt1 - pt2) / (pt1 + pt2);
    float qt1 = a2b2 * cos2 + sin2_4;
    float qt2 = pt2 * sin2_2;
    float par = perp * (qt1 - qt2) / (qt1 + qt2);
    return (par + perp) * 0.5f;
}

// TracerFunctions.cuh:81-90
__device__ __forceinline__ float dggx(float ndh, float alpha)
{
    float a2 = alpha * alpha;
    float denom = ndh * ndh * (a2 - 1.0f) + 1.0f;
    denom = denom * denom;
    denom *= PI_F;
    return a2 / denom;
}

// TracerFunctions.cuh:92-116
__device__ __forceinline__ float lambda_smith(V3 v, float alpha)
{
    V3 s = mul(v, v);
    float a2 = alpha * alpha;
    float inner = a2 * (s.x + s.y) / s.z;
    float l = sqrtf(1.0f + inner) - 1.0f;
    l *= 0.5f;
    return l;
}
__device__ __forceinline__ float gsmith_single(V3 v, float alpha) { return 1.0f / (1.0f + lambda_smith(v, alpha)); }
__device__ __forceinline__ float gsmith_corr(V3 wo, V3 wi, float alpha)
{
    return 1.0f / (lambda_smith(wo, alpha) + lambda_smith(wi, alpha) + 1.0f);
}

// TracerFunctions.cuh:186-196
__device__ __forceinline__ float vndf_pdf(V3 v, V3 h, float alpha)
{
    float vdh = fmaxf(0.0f, dot(h, v));
    float ndh = fmaxf(0.0f, h.z);
    float ndv = fmaxf(0.0f, v.z);
    float dd = dggx(ndh, alpha);
    float g1 = gsmith_single(v, alpha);
    return (ndv == 0.0f) ? 0.0f : (vdh * dd * g1 / ndv);
}

// TracerFunctions.cuh:198-258 (Heitz 2018)
__device__ __forceinline__ V3 vndf_sample(V3 v, float a, float xi0, float xi1, float& pdf)
{
    V3 vh = normalize(v3(a * v.x, a * v.y, v.z));
    float lensq = 0.0f;
    lensq += vh.x * vh.x;
    lensq += vh.y * vh.y;
    V3 t1v = (lensq > 0.0f) ? divs(v3(-vh.y, vh.x, 0.0f), sqrtf(lensq)) : v3(1.0f, 0.0f, 0.0f);
    V3 t2v = cross(vh, t1v);
    float r = sqrtf(xi0);
    float sp, cp;
    sincos_2pi(xi1, sp, cp);
    float t1 = r * cp;
    float t2 = r * sp;
    float s = 0.5f * (1.0f + vh.z);
    t2 = (1.0f - s) * sqrtf(1.0f - t1 * t1) + s * t2;
    float val = 1.0f - t1 * t1 - t2 * t2;
    float sq = sqrtf(fmaxf(0.0f, val));
    V3 nh = (scale(t1v, t1) + scale(t2v, t2)) + scale(vh, sq);
    V3 nm = v3(a * nh.x, a * nh.y, fmaxf(0.0f, nh.z));
    float len2 = dot(nm, nm);
    if (len2 < 1.0e-5f) nm = v3(0.0f, 0.0f, 1.0f);
    else nm = scale(nm, 1.0f / sqrtf(len2));
    pdf = vndf_pdf(v, nm, a);
    return nm;
}

struct Surface {
    V3 pos, geo_n;
    Q4 w2t;
    bool back;
};

// GPUPrimitiveTriangle.h:626-670 for flat-shaded two-sided triangles (the
// face frame conj(RotationBetweenZAxis(n)) at every corner).
__device__ __forceinline__ Surface make_surface(V3 p0, V3 p1, V3 p2, float ba, float bb, V3 dir)
{
    Surface s;
    float c = 1.0f - ba - bb;
    s.pos = (scale(p0, ba) + scale(p1, bb)) + scale(p2, c);
    V3 n = normalize(cross(p1 - p0, p2 - p0));
    s.w2t = conj(rot_between_z(n));
    s.back = dot(n, dir) > 0.0f;
    if (s.back) {
        n = -n;
        s.w2t = qmul(qmake(0.0f, 1.0f, 0.0f, 0.0f), s.w2t);
    }
    s.geo_n = n;
    return s;
}

struct BsdfOut {
    V3 o, d, refl;
    float pdf;
    bool specular;
    uint32_t medium;
};

// SimpleMaterialsKC.cuh:16-60
__device__ __forceinline__ void lambert_sample(const Surface& s, V3 albedo, Rng& rng, BsdfOut& out)
{
    const V3 normal = rotate(conj(s.w2t), v3(0.0f, 0.0f, 1.0f));
    const float xi0 = rng.uniform();
    const float xi1 = rng.uniform();
    const float c0 = sqrtf(fmaxf(0.0f, 1.0f - xi0 * xi0));
    float sn, cs;
    sincos_2pi(xi1, sn, cs);
    V3 dir = normalize(v3(c0 * cs, c0 * sn, xi0));
    dir = rotate(rot_between_z(normal), dir);
    const float ndl = fabsf(dot(normal, dir));
    out.o = nudge(s.pos, s.geo_n);
    out.d = dir;
    out.refl = scale(scale(albedo, ndl), INV_PI);
    out.pdf = INV_PI * 0.5f;
    out.specular = false;
}

// MetalGlassMaterialsKC.cuh:24-150
__device__ __forceinline__ void metal_sample(const Surface& s, V3 wi, const float* __restrict__ m,
                                             Rng& rng, BsdfOut& out)
{
    const float rough = m[0];
    const float alpha = rough * rough;
    const V3 V = rotate(s.w2t, wi);
    const float xi0 = rng.uniform();
    const float xi1 = rng.uniform();
    float pdf;
    const V3 H = vndf_sample(V, alpha, xi0, xi1, pdf);
    const float vdh = fmaxf(0.0f, dot(V, H));
    V3 L = scale(H, 2.0f * vdh) - V;
    pdf = (vdh == 0.0f) ? 0.0f : pdf / (4.0f * vdh);
    const float ldh = fmaxf(0.0f, dot(L, H));
    const float ndh = fmaxf(0.0f, H.z);
    const float ndv = fmaxf(0.0f, V.z);
    V3 F = v3(fresnel_conductor(vdh, m[1], m[4]), fresnel_conductor(vdh, m[2], m[5]),
              fresnel_conductor(vdh, m[3], m[6]));
    F = mul(F, v3(m[7], m[8], m[9]));
    const float D = dggx(ndh, alpha);
    float G = gsmith_corr(V, L, alpha);
    G = (ldh == 0.0f) ? 0.0f : G;
    G = (vdh == 0.0f) ? 0.0f : G;
    V3 spec = v3(((D * F.x) * G) * 0.25f, ((D * F.y) * G) * 0.25f, ((D * F.z) * G) * 0.25f);
    spec = divs(spec, ndv);
    if (ndv == 0.0f) spec = v3(0.0f, 0.0f, 0.0f);
    if (isinf(D) || isnan(D)) {
        if (rough == 0.0f) L = scale(v3(0.0f, 0.0f, 1.0f), 2.0f * V.z) - V;
        spec = scale(F, G);
        const float gs = gsmith_single(V, alpha);
        spec = (gs != 0.0f) ? divs(spec, gs) : v3(0.0f, 0.0f, 0.0f);
        pdf = 1.0f;
    }
    out.d = rotate(conj(s.w2t), L);
    out.o = nudge(s.pos, s.geo_n);
    out.refl = spec;
    out.pdf = pdf;
    out.specular = (1.0f - rough) >= 0.99f;
}

// SimpleMaterialsKC.cuh:172-290; medium 0 = vacuum, k+1 = inside dielectric k
__device__ __forceinline__ void dielectric_sample(const Surface& s, V3 wi, const float* __restrict__ m,
                                                  uint32_t mat_id, uint32_t medium,
                                                  const float* __restrict__ diel, Rng& rng, BsdfOut& out)
{
    const V3 albedo = v3(m[0], m[1], m[2]);
    const float ior = m[3];
    const V3 normal = rotate(conj(s.w2t), v3(0.0f, 0.0f, 1.0f));
    const bool entering = !s.back;
    const float from = (medium == 0u) ? 1.0f : diel[(size_t)(medium - 1u) * 4 + 3];
    const float to = entering ? ior : 1.0f;
    const float f = fresnel_dielectric(fabsf(dot(wi, normal)), from, to);
    const float xi = rng.uniform();
    out.specular = true;
    if (xi < f) {
        const float k = 2.0f * dot(wi, normal);
        out.d = scale(normal, k) - wi;
        out.o = nudge(s.pos, s.geo_n);
        out.pdf = f;
        out.medium = medium;
        out.refl = scale(albedo, f);
    } else {
        out.medium = entering ? (mat_id + 1u) : 0u;
        out.pdf = 1.0f - f;
        const float ratio = from / to;
        const float cos_in = dot(normal, wi);
        const float sin_in2 = fmaxf(0.0f, 1.0f - cos_in * cos_in);
        const float sin_out2 = ratio * ratio * sin_in2;
        const float cos_out = sqrtf(fmaxf(0.0f, 1.0f - sin_out2));
        if (sin_out2 >= 1.0f) {
            out.pdf = 0.0f;
            out.refl = v3(0.0f, 0.0f, 0.0f);
            out.d = wi;
            out.o = s.pos;
            return;
        }
        const float kk = ratio * cos_in - cos_out;
        out.d = scale(-wi, ratio) + scale(normal, kk);
        out.o = nudge(s.pos, -s.geo_n);
        out.refl = scale(scale(albedo, 1.0f), 1.0f - f);
    }
}

// Lambert f*cos and the pdf of its (uniform hemisphere) sampler
// (SimpleMaterialsKC.cuh:80-101 Evaluate; sampling pdf 1/(2 pi) on the
// normal's side).  The reference's LambertC::Pdf returns the cosine pdf, which
// does not match its uniform sampler; the mixture below needs the sampler's
// own pdf, so that is what this returns.
__device__ __forceinline__ V3 lambert_eval(const Surface& s, V3 albedo, V3 dir, float& pdf)
{
    const V3 normal = rotate(conj(s.w2t), v3(0.0f, 0.0f, 1.0f));
    const float ndl = dot(normal, dir);
    pdf = (ndl > 0.0f) ? INV_PI * 0.5f : 0.0f;
    return scale(scale(albedo, fmaxf(ndl, 0.0f)), INV_PI);
}

// MetalMatFuncs::Evaluate / Pdf (MetalGlassMaterialsKC.cuh:152-220)
__device__ __forceinline__ V3 metal_eval(const Surface& s, V3 wi, const float* __restrict__ m, V3 dir, float& pdf)
{
    const float rough = m[0];
    const float alpha = rough * rough;
    const V3 L = rotate(s.w2t, dir);
    const V3 V = rotate(s.w2t, wi);
    const V3 H = normalize(L + V);
    const float ndh = fmaxf(0.0f, H.z);
    const float vdh = fmaxf(0.0f, dot(V, H));
    const float ldh = fmaxf(0.0f, dot(L, H));
    const float ndv = fmaxf(0.0f, V.z);
    float pspec = vndf_pdf(V, H, alpha);
    float D = dggx(ndh, alpha);
    pspec = (isnan(D) || isinf(D)) ? 0.0f : pspec;
    pdf = (vdh == 0.0f) ? 0.0f : pspec / (4.0f * vdh);
    D = (isnan(D) || isinf(D)) ? 0.0f : D;
    float G = gsmith_corr(V, L, alpha);
    G = (ldh == 0.0f) ? 0.0f : G;
    G = (vdh == 0.0f) ? 0.0f : G;
    V3 F = v3(fresnel_conductor(vdh, m[1], m[4]), fresnel_conductor(vdh, m[2], m[5]),
              fresnel_conductor(vdh, m[3], m[6]));
    F = mul(F, v3(m[7], m[8], m[9]));
    V3 spec = v3(((D * F.x) * G) * 0.25f, ((D * F.y) * G) * 0.25f, ((D * F.z) * G) * 0.25f);
    spec = divs(spec, ndv);
    return (ndv == 0.0f) ? v3(0.0f, 0.0f, 0.0f) : spec;
}

// Utility::RGBToLuminance (RayLib/ColorConversion.h:15-22, its coefficients)
__device__ __forceinline__ float luminance(V3 c) { return (0.2120f * c.x + 0.7150f * c.y) + 0.0722f * c.z; }

// Path-vertex records of the learning pass, indexed [sample * max_depth + bounce]
struct RecBuf {
    uint32_t* code;    // voxel code of the vertex, SVO_NO_NODE = no record
    uint32_t* bin;     // direction bin of the continuing ray
    float* tr;         // throughput leaving the vertex (PathNode.cuh:116 radFactor), R
    float* tg;         // G
    float* tb;         // B
    float* rad;        // luminance of the radiance arriving along that ray
    uint32_t max_depth;
};

struct GuideArgs {
    RecBuf rec;
    SvoView svo;        // valid when GUIDE
    float lox, loy, loz, inv_cell;   // leaf grid of the records
    uint32_t depth;
    float ratio;        // probability of sampling the guide
    uint32_t min_samples;   // guide lookup threshold
};

// PathTracerKC.cuh:337-426 (light hit) and 430-648 (surface work); with
// LEARN the WFPG path-node bookkeeping (WFPGTracerKC.cuh:202-520) and with
// GUIDE the one-sample MIS of guide and BxDF (WFPGTracerKC.cuh:1480-1553,
// pdf = r p_guide + (1 - r) p_bxdf, weight = f cos / pdf).
template <bool LEARN, bool GUIDE>
__global__ __launch_bounds__(256) void k_shade(
    PathBuf in, PathBuf out, HitBuf hit, const uint32_t* __restrict__ order, uint32_t* counts,
    uint32_t bounce, uint32_t last, uint32_t rr_start, uint32_t spp, uint32_t sample_base,
    Materials mats, const float4* __restrict__ leaves, float* __restrict__ image, GuideArgs ga)
{
    const uint32_t n = counts[bounce];
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t lanes_below = (1ull << lane) - 1ull;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const uint32_t j = base + threadIdx.x;
        bool alive = false;
        V3 no = v3(0.0f, 0.0f, 0.0f), nd = no, nT = no;
        uint32_t sample = 0, ndm = 0;
        uint64_t nrng = 0;
        if (j < n) {
            const uint32_t rid = order[j];
            const uint32_t key = hit.key[rid];
            const uint32_t batch = key >> 24, id = key & 0xFFFFFFu;
            if (batch != MRH_BATCH_NULL) {
                const V3 T = v3(in.tr[rid], in.tg[rid], in.tb[rid]);
                sample = in.sample[rid];
                if (batch == MRH_BATCH_EMISSIVE) {
                    const float* le = mats.emissive + (size_t)id * 3;
                    float* px = image + (size_t)(sample / spp) * 3;
                    atomicAdd(px + 0, le[0] * T.x);
                    atomicAdd(px + 1, le[1] * T.y);
                    atomicAdd(px + 2, le[2] * T.z);
                    if (LEARN) {
                        // radiance arriving along each earlier vertex's outgoing ray:
                        // path contribution over that vertex's throughput
                        // (PathNode.cuh:152-190, per channel, then luminance)
                        const V3 c = mul(v3(le[0], le[1], le[2]), T);
                        const size_t r0 = (size_t)sample * ga.rec.max_depth;
                        for (uint32_t k = 0; k < bounce; k++) {
                            if (ga.rec.code[r0 + k] == SVO_NO_NODE) continue;
                            const float fr = ga.rec.tr[r0 + k], fg = ga.rec.tg[r0 + k], fb = ga.rec.tb[r0 + k];
                            const V3 L = v3(fr == 0.0f ? 0.0f : c.x / fr, fg == 0.0f ? 0.0f : c.y / fg,
                                            fb == 0.0f ? 0.0f : c.z / fb);
                            ga.rec.rad[r0 + k] = luminance(L);
                        }
                    }
                } else if (!last) {
                    const uint32_t li = hit.leaf[rid];
                    const float4 l0 = leaves[3 * (size_t)li + 0];
                    const float4 l1 = leaves[3 * (size_t)li + 1];
                    const float4 l2 = leaves[3 * (size_t)li + 2];
                    const V3 d = v3(in.dx[rid], in.dy[rid], in.dz[rid]);
                    const Surface s = make_surface(v3(l0.x, l0.y, l0.z), v3(l1.x, l1.y, l1.z),
                                                   v3(l2.x, l2.y, l2.z), hit.ba[rid], hit.bb[rid], d);
                    const V3 wi = -normalize(d);
                    Rng rng{in.rng[rid], pcg32_inc(sample_base + sample)};
                    const uint32_t dm = in.dm[rid];
                    const uint32_t depth = dm & 0xFFu, medium = dm >> 8;
                    BsdfOut o;
                    o.medium = medium;
                    bool guided = false;
                    uint32_t gnode = 0;
                    float gtotal = 0.0f;
                    V3 ns = v3(0.0f, 0.0f, 1.0f);
                    if (GUIDE && (batch == MRH_BATCH_LAMBERT ||
                                  (batch == MRH_BATCH_METAL && (1.0f - mats.metal[(size_t)id * 10]) < 0.99f))) {
                        ns = rotate(conj(s.w2t), v3(0.0f, 0.0f, 1.0f));
                        gnode = svo_lookup_guide(ga.svo,
                                                 svo_code(s.pos, ga.svo.lox, ga.svo.loy, ga.svo.loz,
                                                          ga.svo.inv_cell, ga.svo.depth),
                                                 ga.min_samples);
                        gtotal = guide_total(ga.svo, gnode, ns);
                        guided = gtotal > 0.0f;
                    }
                    if (guided) {
                        const float xi = rng.uniform();
                        V3 dir;
                        float pg = 0.0f, pb = 0.0f;
                        V3 fcos;
                        if (xi < ga.ratio) {
                            const float u0 = rng.uniform(), u1 = rng.uniform(), u2 = rng.uniform();
                            dir = guide_sample(ga.svo, gnode, ns, gtotal, u0, u1, u2, pg);
                            if (batch == MRH_BATCH_LAMBERT) {
                                const float* a = mats.lambert + (size_t)id * 3;
                                fcos = lambert_eval(s, v3(a[0], a[1], a[2]), dir, pb);
                            } else {
                                fcos = metal_eval(s, wi, mats.metal + (size_t)id * 10, dir, pb);
                            }
                        } else {
                            if (batch == MRH_BATCH_LAMBERT) {
                                const float* a = mats.lambert + (size_t)id * 3;
                                lambert_sample(s, v3(a[0], a[1], a[2]), rng, o);
                            } else {
                                metal_sample(s, wi, mats.metal + (size_t)id * 10, rng, o);
                            }
                            dir = o.d;
                            fcos = o.refl;
                            pb = o.pdf;
                            pg = guide_pdf(ga.svo, gnode, ns, gtotal, dir);
                        }
                        o.d = dir;
                        o.o = nudge(s.pos, s.geo_n);
                        o.refl = fcos;
                        o.pdf = ga.ratio * pg + (1.0f - ga.ratio) * pb;
                        o.specular = false;
                        o.medium = medium;
                    } else if (batch == MRH_BATCH_LAMBERT) {
                        const float* a = mats.lambert + (size_t)id * 3;
                        lambert_sample(s, v3(a[0], a[1], a[2]), rng, o);
                    } else if (batch == MRH_BATCH_METAL) {
                        metal_sample(s, wi, mats.metal + (size_t)id * 10, rng, o);
                    } else {
                        dielectric_sample(s, wi, mats.dielectric + (size_t)id * 4, id, medium,
                                          mats.dielectric, rng, o);
                    }
                    V3 t2 = mul(T, o.refl);
                    t2 = (o.pdf == 0.0f) ? v3(0.0f, 0.0f, 0.0f) : divs(t2, o.pdf);
                    const float avg = dot(t2, v3(0.333f, 0.333f, 0.333f));
                    bool terminate = false;
                    if (!o.specular && depth > rr_start) {
                        // TracerFunctions.cuh:321-330
                        const float pr = avg < 0.005f ? 0.005f : (avg > 1.0f ? 1.0f : avg);
                        if (rng.uniform() >= pr) terminate = true;
                        else t2 = scale(t2, 1.0f / pr);
                    }
                    alive = !((terminate && !o.specular) || is_zero(t2));
                    if (LEARN && alive) {
                        // path node of this vertex (PPGTracerKC.cuh:605-620)
                        const size_t r = (size_t)sample * ga.rec.max_depth + bounce;
                        ga.rec.code[r] = svo_code(s.pos, ga.lox, ga.loy, ga.loz, ga.inv_cell, ga.depth);
                        ga.rec.bin[r] = dir_to_bin(normalize(o.d));
                        ga.rec.tr[r] = t2.x;
                        ga.rec.tg[r] = t2.y;
                        ga.rec.tb[r] = t2.z;
                        ga.rec.rad[r] = 0.0f;
                    }
                    no = o.o;
                    nd = o.d;
                    nT = t2;
                    ndm = ((depth + 1u) & 0xFFu) | (o.medium << 8);
                    nrng = rng.state;
                }
            }
        }
        // wave-aggregated append into the next bounce's live range
        const uint64_t m = __ballot(alive);
        if (m) {
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1u;
            uint32_t slot0 = 0;
            if (lane == leader) slot0 = atomicAdd(&counts[bounce + 1], (uint32_t)__popcll(m));
            slot0 = __shfl(slot0, (int)leader, 64);
            if (alive) {
                const uint32_t k = slot0 + (uint32_t)__popcll(m & lanes_below);
                out.ox[k] = no.x; out.oy[k] = no.y; out.oz[k] = no.z;
                out.dx[k] = nd.x; out.dy[k] = nd.y; out.dz[k] = nd.z;
                out.tr[k] = nT.x; out.tg[k] = nT.y; out.tb[k] = nT.z;
                out.sample[k] = sample;
                out.dm[k] = ndm;
                out.rng[k] = nrng;
            }
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

size_t frame_bytes(uint64_t cap)
{
    // 2 path buffers (9 floats + sample + dm = 11 words, + u64 rng),
    // hits (4 words) + key + order, plus counts/offsets
    size_t per = 0;
    per += 2 * (11 * align_up(cap * 4) + align_up(cap * 8));
    per += 6 * align_up(cap * 4);
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
    uint64_t total = 0, spp = 0;
    uint32_t w = 0, h = 0, S = 0, max_depth = 0, rr_start = 0, seed = 0, sample_base = 0;
    uint32_t count_traversal = 0, bounce = 0;
    uint32_t learn = 0, guide = 0;   // record path vertices / sample the guide
    GuideArgs ga = {};
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
    }
    f.hit.leaf = (uint32_t*)take(cap * 4);
    f.hit.t = (float*)take(cap * 4);
    f.hit.ba = (float*)take(cap * 4);
    f.hit.bb = (float*)take(cap * 4);
    f.hit.key = (uint32_t*)take(cap * 4);
    f.order = (uint32_t*)take(cap * 4);
    f.counts = (uint32_t*)take(COUNTS_WORDS * 4);
    f.offsets = (uint32_t*)take(260 * 4);
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
    if (s->ctx) hipStreamSynchronize(s->ctx->stream);
    if (s->bvh) lbvh_free(s->bvh);
    for (void* p : s->allocs) hipFree(p);
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
                     : (b == MRH_BATCH_EMISSIVE) ? d->n_emissive : 0u;
        if (b == MRH_BATCH_NULL || b > MRH_BATCH_EMISSIVE || id >= lim)
            return fail(ctx, MRH_ERR_INVALID, "mrh_scene_create: primitive %u has invalid key 0x%08x", i,
                        d->prim_key[i]);
    }
    mrh_scene* s = new mrh_scene();
    s->ctx = ctx;
    s->n_lambert = d->n_lambert;
    s->n_metal = d->n_metal;
    s->n_dielectric = d->n_dielectric;
    s->n_emissive = d->n_emissive;
    float* dtris = nullptr;
    uint32_t* dkeys = nullptr;
    int32_t rc = MRH_OK;
    do {
        if ((rc = scene_upload(s, &s->lambert, d->lambert, (size_t)d->n_lambert * 3)) != MRH_OK) break;
        if ((rc = scene_upload(s, &s->metal, d->metal, (size_t)d->n_metal * 10)) != MRH_OK) break;
        if ((rc = scene_upload(s, &s->dielectric, d->dielectric, (size_t)d->n_dielectric * 4)) != MRH_OK) break;
        if ((rc = scene_upload(s, &s->emissive, d->emissive, (size_t)d->n_emissive * 3)) != MRH_OK) break;
        if ((rc = scene_upload(s, &dtris, d->tris, (size_t)d->n * 9)) != MRH_OK) break;
        if ((rc = scene_upload(s, &dkeys, d->prim_key, (size_t)d->n)) != MRH_OK) break;
        uint32_t bits = d->morton_bits ? d->morton_bits : 30u;
        if ((rc = lbvh_create(ctx, dtris, d->n, bits, dkeys, &s->bvh)) != MRH_OK) break;
        s->bytes += (size_t)d->n * (36 + 24 + 8 + 8 + 4 + 12 + 8 + 48 + 64 + 4);
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
    if (p->nee) return fail(ctx, MRH_ERR_UNSUPPORTED, "NextEventEstimation is not implemented");
    if (p->width == 0 || p->height == 0 || p->samples == 0)
        return fail(ctx, MRH_ERR_INVALID, "render: empty frame");
    if (p->max_depth == 0 || p->max_depth > MAX_BOUNCES)
        return fail(ctx, MRH_ERR_INVALID, "render: max_depth must be in [1, %u]", MAX_BOUNCES);
    const uint64_t t = (uint64_t)p->width * p->height * p->samples * p->samples;
    if (t >= (1ull << 30))
        return fail(ctx, MRH_ERR_INVALID, "render: %llu paths exceed 2^30", (unsigned long long)t);
    if (s->bvh->max_depth > 128)
        return fail(ctx, MRH_ERR_UNSUPPORTED, "tree depth %u exceeds the traversal stack", s->bvh->max_depth);
    *total = t;
    return MRH_OK;
}

// GenerateWork: clear counts, write the camera rays of one frame
int32_t frame_begin(mrh_context* ctx, const mrh_camera_desc* cd, const mrh_render_params* p,
                    void* arena, uint32_t sample_base, Frame& f)
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
    f.bounce = 0;
    f.learn = f.guide = 0;
    f.cam = make_camera(*cd, p->width, p->height);
    carve_frame(arena, f.total, f);
    MRH_TRY(fill_u32(ctx, f.counts, 0u, COUNTS_WORDS));
    MRH_TRY(fill_u32(ctx, f.counts, (uint32_t)f.total, 1));
    return launch(ctx, "raygen", dim3(stride_grid(ctx, f.total, 256)), dim3(256), 0, k_raygen, f.cam,
                  f.w, f.h, f.S, f.seed, f.sample_base, (uint32_t)f.total, f.bufs[0]);
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
    const uint32_t grid = stride_grid(ctx, f.total, 256);
    const uint32_t tgrid = trace_grid(ctx, f.total);
    const float4* lv = (const float4*)b->leaves;
    if (f.count_traversal) {
        if (b->max_depth <= 64) MRH_TRY((launch_trace_pt<64, true>(ctx, tgrid, b, in, f.counts, bounce, tmin, tmax, f.hit)));
        else MRH_TRY((launch_trace_pt<128, true>(ctx, tgrid, b, in, f.counts, bounce, tmin, tmax, f.hit)));
    } else {
        if (b->max_depth <= 64) MRH_TRY((launch_trace_pt<64, false>(ctx, tgrid, b, in, f.counts, bounce, tmin, tmax, f.hit)));
        else MRH_TRY((launch_trace_pt<128, false>(ctx, tgrid, b, in, f.counts, bounce, tmin, tmax, f.hit)));
    }
    SortArgs sa;
    sa.keys_in = f.hit.key;
    sa.vals_out = f.order;
    sa.n_max = (uint32_t)f.total;
    sa.n_dev = f.counts + bounce;
    sa.key_bytes = 4;
    sa.begin_bit = 24;
    sa.end_bit = 32;
    sa.descending = true;
    sa.part_offsets = f.offsets;
    sa.part_begin = 24;
    MRH_TRY(radix_sort(ctx, sa));
    const Materials mats{s->lambert, s->metal, s->dielectric, s->emissive};
    auto shade = [&](auto kernel) {
        return launch(ctx, "shade", dim3(grid), dim3(256), 0, kernel, in, outb, f.hit, (const uint32_t*)f.order,
                      f.counts, bounce, (uint32_t)(bounce + 1 == f.max_depth), f.rr_start, (uint32_t)f.spp,
                      f.sample_base, mats, lv, image, f.ga);
    };
    if (f.learn && f.guide) MRH_TRY(shade(k_shade<true, true>));
    else if (f.learn) MRH_TRY(shade(k_shade<true, false>));
    else if (f.guide) MRH_TRY(shade(k_shade<false, true>));
    else MRH_TRY(shade(k_shade<false, false>));
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

int32_t render(mrh_context* ctx, mrh_scene* s, const mrh_camera_desc* cd, const mrh_render_params* p,
               float* d_image, mrh_render_stats* stats)
{
    uint64_t total = 0;
    MRH_TRY(validate_params(ctx, s, p, &total));
    void* arena = nullptr;
    MRH_TRY(scratch(ctx, SLOT_FRAME, frame_bytes(total), &arena));
    const size_t img_count = (size_t)p->width * p->height * 3;
    MRH_TRY(fill_u32(ctx, (uint32_t*)d_image, 0u, img_count));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    MRH_HIP_CHECK(ctx, hipEventCreate(&e0));
    MRH_HIP_CHECK(ctx, hipEventCreate(&e1));
    MRH_HIP_CHECK(ctx, hipEventRecord(e0, ctx->stream));
    Frame f;
    int32_t rc = frame_begin(ctx, cd, p, arena, 0u, f);
    while (rc == MRH_OK && f.bounce < f.max_depth) rc = frame_bounce(ctx, s, f, d_image);
    if (rc == MRH_OK)
        rc = launch(ctx, "finalize", dim3(stride_grid(ctx, img_count, 256)), dim3(256), 0, k_scale, d_image,
                    img_count, 1.0f / (float)f.spp);
    if (rc == MRH_OK) {
        hipEventRecord(e1, ctx->stream);
        mrh_render_stats local;
        memset(&local, 0, sizeof(local));
        rc = frame_counts(ctx, f, &local);
        if (rc == MRH_OK && stats) {
            float ms = 0.0f;
            hipEventElapsedTime(&ms, e0, e1);
            local.render_ms = ms;
            *stats = local;
        }
    } else {
        hipStreamSynchronize(ctx->stream);
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc;
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
    if (gp->iterations == 0 || gp->iterations > 64)
        return fail(ctx, MRH_ERR_INVALID, "render_guided: iterations must be in [1, 64]");
    if (gp->svo_depth == 0 || gp->svo_depth > 10)
        return fail(ctx, MRH_ERR_INVALID, "render_guided: svo_depth must be in [1, 10]");
    if (!(gp->guide_ratio >= 0.0f && gp->guide_ratio <= 1.0f))
        return fail(ctx, MRH_ERR_INVALID, "render_guided: guide_ratio must be in [0, 1]");
    if ((uint64_t)gp->iterations * total >= (1ull << 32))
        return fail(ctx, MRH_ERR_INVALID, "render_guided: iterations x paths exceed 2^32 samples");
    const uint64_t n_rec = total * p->max_depth;
    if (n_rec >= (1ull << 32))
        return fail(ctx, MRH_ERR_INVALID, "render_guided: %llu path vertices exceed 2^32",
                    (unsigned long long)n_rec);
    void* arena = nullptr;
    MRH_TRY(scratch(ctx, SLOT_FRAME, frame_bytes(total), &arena));
    const size_t img_count = (size_t)p->width * p->height * 3;
    MRH_TRY(fill_u32(ctx, (uint32_t*)d_image, 0u, img_count));

    // cube of the voxel grid: scene bounds padded by one voxel on each side
    const mrh_lbvh* b = s->bvh;
    float span = 0.0f;
    for (int a = 0; a < 3; a++) span = fmaxf(span, b->hi[a] - b->lo[a]);
    if (!(span > 0.0f)) span = 1.0f;
    const float cells = (float)(1u << gp->svo_depth);
    const float pad = span / (cells - 2.0f > 1.0f ? cells - 2.0f : 1.0f);
    const float cube[4] = {b->lo[0] - pad, b->lo[1] - pad, b->lo[2] - pad, span + 2.0f * pad};

    mrh_svo* own = nullptr;
    if (!svo) {
        MRH_TRY(mrh_svo_create(ctx, &own));
        svo = own;
    }
    // records [n_rec] x (code, bin, tr, tg, tb, rad) + compacted (code, bin, rad)
    void* recmem = nullptr;
    int32_t rc = MRH_OK;
    const bool learns = gp->iterations > 1;
    if (learns) rc = dmalloc(ctx, &recmem, (size_t)n_rec * 36);
    RecBuf rec = {};
    uint32_t *c_codes = nullptr, *c_bins = nullptr;
    float* c_rad = nullptr;
    if (recmem) {
        char* q = (char*)recmem;
        rec.code = (uint32_t*)q; q += n_rec * 4;
        rec.bin = (uint32_t*)q; q += n_rec * 4;
        rec.tr = (float*)q; q += n_rec * 4;
        rec.tg = (float*)q; q += n_rec * 4;
        rec.tb = (float*)q; q += n_rec * 4;
        rec.rad = (float*)q; q += n_rec * 4;
        c_codes = (uint32_t*)q; q += n_rec * 4;
        c_bins = (uint32_t*)q; q += n_rec * 4;
        c_rad = (float*)q;
        rec.max_depth = p->max_depth;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == MRH_OK && hipEventCreate(&e0) != hipSuccess) rc = fail(ctx, MRH_ERR_HIP, "event create");
    if (rc == MRH_OK && hipEventCreate(&e1) != hipSuccess) rc = fail(ctx, MRH_ERR_HIP, "event create");
    if (rc == MRH_OK) (void)hipEventRecord(e0, ctx->stream);
    mrh_render_stats local;
    memset(&local, 0, sizeof(local));
    bool have_guide = false;
    for (uint32_t it = 0; it < gp->iterations && rc == MRH_OK; it++) {
        const bool learn = learns && it + 1 < gp->iterations;
        if (learn) rc = fill_u32(ctx, rec.code, SVO_NO_NODE, n_rec);
        Frame f;
        if (rc == MRH_OK) rc = frame_begin(ctx, cd, p, arena, (uint32_t)(it * total), f);
        if (rc != MRH_OK) break;
        f.learn = learn;
        f.guide = have_guide;
        f.ga.rec = rec;
        f.ga.lox = cube[0];
        f.ga.loy = cube[1];
        f.ga.loz = cube[2];
        f.ga.inv_cell = cells / cube[3];
        f.ga.depth = gp->svo_depth;
        f.ga.ratio = gp->guide_ratio;
        f.ga.min_samples = gp->min_samples;
        if (have_guide) f.ga.svo = svo_view(svo);
        while (rc == MRH_OK && f.bounce < f.max_depth) rc = frame_bounce(ctx, s, f, d_image);
        if (rc == MRH_OK) rc = frame_counts(ctx, f, &local);
        if (rc != MRH_OK || !learn) continue;
        uint32_t kept = 0;
        rc = svo_compact_records(ctx, rec.code, rec.bin, rec.rad, (uint32_t)n_rec, c_codes, c_bins, c_rad, &kept);
        if (rc != MRH_OK || kept == 0) continue;   // nothing learned: keep the previous guide
        rc = svo_build(ctx, svo, c_codes, kept, cube, gp->svo_depth);
        if (rc == MRH_OK) rc = svo_splat(ctx, svo, c_codes, c_bins, c_rad, kept);
        if (rc == MRH_OK) rc = svo_filter(ctx, svo);
        if (rc == MRH_OK) have_guide = true;
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
    (void)hipFree(recmem);
    if (own) svo_free(own);
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
    return render(ctx, scene, cam, params, d_image, stats);
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
    *bytes = frame_bytes(total) + total * 16;   // + sort scratch
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
    int32_t rc = dmalloc(ctx, &t->arena, frame_bytes(total));
    if (rc == MRH_OK) rc = dmalloc(ctx, (void**)&t->accum, img * 4);
    if (rc == MRH_OK) rc = fill_u32(ctx, (uint32_t*)t->accum, 0u, img);
    if (rc != MRH_OK) {
        hipFree(t->arena);
        hipFree(t->accum);
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
    MRH_TRY(frame_begin(ctx, cam, &t->p, t->arena, (uint32_t)(t->frames * t->total), t->f));
    t->active = true;
    return MRH_OK;
}

int32_t mrh_tracer_render(mrh_tracer* t, int32_t* more)
{
    if (!t || !more) return MRH_ERR_INVALID;
    *more = 0;
    if (!t->active) return MRH_OK;
    MRH_TRY(frame_bounce(t->ctx, t->scene, t->f, t->accum));
    if (t->f.bounce < t->f.max_depth) {
        *more = 1;
        return MRH_OK;
    }
    MRH_TRY(frame_counts(t->ctx, t->f, &t->stats));
    t->active = false;
    t->frames++;
    return MRH_OK;
}

int32_t mrh_tracer_finalize(mrh_tracer* t, float* d_image)
{
    if (!t || !d_image) return MRH_ERR_INVALID;
    mrh_context* ctx = t->ctx;
    if (t->active) return fail(ctx, MRH_ERR_INVALID, "mrh_tracer_finalize: frame still rendering");
    const size_t img = (size_t)t->p.width * t->p.height * 3;
    const float s = t->frames ? 1.0f / (float)((uint64_t)t->frames * t->f.spp) : 0.0f;
    MRH_TRY(launch(ctx, "finalize", dim3(stride_grid(ctx, img, 256)), dim3(256), 0, k_mean,
                   (const float*)t->accum, d_image, img, s));
    MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return MRH_OK;
}

int32_t mrh_tracer_reset_image(mrh_tracer* t)
{
    if (!t) return MRH_ERR_INVALID;
    if (t->active) return fail(t->ctx, MRH_ERR_INVALID, "mrh_tracer_reset_image: frame still rendering");
    t->frames = 0;
    t->stats = mrh_render_stats{};
    return fill_u32(t->ctx, (uint32_t*)t->accum, 0u, (size_t)t->p.width * t->p.height * 3);
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
    hipStreamSynchronize(t->ctx->stream);
    hipFree(t->arena);
    hipFree(t->accum);
    delete t;
    return MRH_OK;
}

}  // extern "C"
