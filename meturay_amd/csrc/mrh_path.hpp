// mrh_path.hpp — device building blocks of the wavefront path tracer shared
// by the single-request tracer (mrh_pt.hip) and the multi-request serving
// engine (mrh_serve.hip): camera, path / hit buffers, BxDFs, the guided
// mixture, and the shade step of one ray.
#pragma once

#include "mrh_bvh.hpp"
#include "mrh_guide.hpp"

#include <cfloat>
#include <vector>

// Resident scene: triangles + LBVH, HitKey-indexed material tables
struct mrh_scene {
    mrh_context* ctx = nullptr;
    mrh_lbvh* bvh = nullptr;
    float* lambert = nullptr;
    float* metal = nullptr;
    float* dielectric = nullptr;
    float* emissive = nullptr;
    float* reflect = nullptr;
    uint32_t n_lambert = 0, n_metal = 0, n_dielectric = 0, n_emissive = 0, n_reflect = 0;
    uint32_t* lights = nullptr;     // leaf index of every emissive triangle, in primitive order
    uint32_t n_lights = 0;
    uint64_t bytes = 0;
    std::vector<void*> allocs;
    // Two-level scenes: bvh is the top-level tree over the instances' world
    // boxes; the meshes' trees and leaf records are concatenated.
    bool two_level = false;
    float4* inst = nullptr;         // [n_inst][INST_F4] instance records
    uint32_t n_inst = 0;
    float4* mesh_nodes2 = nullptr;  // binary trees of all meshes
    float4* mesh_nodes4 = nullptr;  // 4-wide trees of all meshes
    float4* leaves = nullptr;       // [n][3] leaf records, mesh after mesh (prim = scene index)
    uint32_t n_tris = 0;
    uint32_t mesh_depth = 0;        // deepest mesh tree
    uint32_t* light_inst = nullptr; // instance of every light (lights hold leaf indices)
    // smooth shading (single-level scenes with per-corner normals): the
    // corners' world-to-tangent rotations, [n][3] float4 (w, x, y, z), primitive order
    float4* tbn = nullptr;
};

namespace mrh {

// Builds the meshes' trees, the top-level tree, the instance records and the
// light list of a two-level scene (mrh_instance.hip); s holds the material
// tables already.
int32_t two_level_build(mrh_context* ctx, const mrh_scene_desc* d, mrh_scene* s);
// World-to-local matrix of a row-major 3x4 local-to-world one (double
// adjugate, rounded to float); false when singular or not finite.
bool affine_inverse(const float* l2w, float* w2l);

namespace {

constexpr float INV_PI = 0.31830988618379067154f;
constexpr float PI_F = 3.14159265358979323846f;

struct Cam {
    V3 pos, bl, right, up;
    float pw, ph, near_d, far_d;
};

// Ray word dm: depth (bits 0..7; 1 = camera ray), the ray type (RayAuxPath
// type: bit 8 = SPECULAR_PATH_RAY, bit 9 = NEE_RAY, neither = camera or
// PATH_RAY) and the medium (bits 10..31; 0 = vacuum, k + 1 = inside
// dielectric k).
constexpr uint32_t DM_SPEC = 1u << 8;
constexpr uint32_t DM_NEE = 1u << 9;
constexpr uint32_t DM_MEDIUM_SHIFT = 10;
constexpr uint32_t MAX_DIELECTRICS = (1u << 22) - 1u;

struct PathBuf {
    float *ox, *oy, *oz, *dx, *dy, *dz;
    float *tr, *tg, *tb;
    uint32_t* sample;
    uint32_t* dm;        // depth | type | medium (above)
    uint64_t* rng;       // PCG32 state; a NEE ray holds its light's leaf (low word) and instance (high word) instead
    float* pdf;          // BxDF pdf of a path ray's direction (RayAuxPath prevPDF), with MIS only
};

struct HitBuf {
    uint32_t* leaf;
    float *t, *ba, *bb;
    uint32_t* key;
    uint32_t* inst = nullptr;   // instance of the hit (two-level scenes only)
};

struct Materials {
    const float* lambert;
    const float* metal;
    const float* dielectric;
    const float* emissive;
    const float* reflect;
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

// GPUPrimitiveTriangle.h:626-670 for smooth-shaded triangles: the
// corners' world-to-tangent rotations q0..q2 (vertex_frame) blended by
// BarySLerp and normalised, flipped by the tangent-axis rotation on the back
// side; position, geometric normal and side as make_surface.
__device__ __forceinline__ Surface make_surface_tbn(V3 p0, V3 p1, V3 p2, float ba, float bb, V3 dir,
                                                    const float4* __restrict__ tbn)
{
    Surface s = make_surface(p0, p1, p2, ba, bb, dir);
    const float4 a = tbn[0], b = tbn[1], c = tbn[2];
    Q4 w = qnormalize(bary_slerp(qmake(a.x, a.y, a.z, a.w), qmake(b.x, b.y, b.z, b.w), qmake(c.x, c.y, c.z, c.w),
                                 ba, bb));
    if (s.back) w = qmul(qmake(0.0f, 1.0f, 0.0f, 0.0f), w);
    s.w2t = w;
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

// ReflectMatFuncs::Sample (SimpleMaterialsKC.cuh:105-151): a perfect mirror
// about the shading normal weighted by the albedo at roughness 0; any other
// roughness reflects nothing (the path ends)
__device__ __forceinline__ void reflect_sample(const Surface& s, V3 wi, const float* __restrict__ m, BsdfOut& out)
{
    const V3 normal = rotate(conj(s.w2t), v3(0.0f, 0.0f, 1.0f));
    out.specular = true;
    out.pdf = 1.0f;
    out.o = s.pos;
    out.d = wi;
    if (m[3] == 0.0f) {
        const float k = 2.0f * dot(wi, normal);
        out.d = scale(normal, k) - wi;
        out.o = nudge(s.pos, s.geo_n);
        out.refl = v3(m[0], m[1], m[2]);
    } else {
        out.refl = v3(0.0f, 0.0f, 0.0f);
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

// Path-vertex records of the learning pass, bounce-major: [bounce * stride +
// sample], so the records a wave writes at one bounce (nearby samples) share
// cache lines
struct RecBuf {
    uint32_t* code;    // voxel code of the vertex, SVO_NO_NODE = no record
    uint32_t* bin;     // direction bin of the continuing ray
    float* tr;         // throughput leaving the vertex (PathNode.cuh:116 radFactor), R
    float* tg;         // G
    float* tb;         // B
    float* rad;        // luminance of the radiance arriving along that ray
    float* vtx;        // optional [stride * max_depth][7]: position, direction toward the
                       // previous vertex, rad (the octree payload's path vertices; code marks the valid ones)
    uint32_t max_depth;
    uint32_t stride;   // samples per frame
};

struct GuideArgs {
    RecBuf rec;
    SvoView svo;        // valid when GUIDE
    float lox, loy, loz, inv_cell;   // leaf grid of the records
    uint32_t depth;
    float ratio;        // probability of sampling the guide
    uint32_t min_samples;   // guide lookup threshold
    // Per-ray guided directions of the field-traced guide, indexed like the
    // rays of the bounce (set when the bounce's fields were sampled):
    // wf_flag 0 = not guided, 1 = the guide's sample, 2 = the BxDF's sample;
    // wf_gdir / wf_gpdf the guide's sample and pdf, wf_bdir / wf_epdf the
    // BxDF's sample and the guide pdf there.  Null: the octree guide.
    const uint32_t* wf_flag;
    const float* wf_gdir;
    const float* wf_gpdf;
    const float* wf_bdir;
    const float* wf_epdf;
};

// Camera ray of sample i of a frame (GenerationKernels.cuh:29-108 +
// GPUCameraPinhole.cuh:209-240): stratum (sx, sy) of pixel (px, py), jittered
// by the first two draws of the path's stream, keyed by the global sample id
// gid.  coord (may be null) receives the sample's image coordinate
// ((stratum + jitter) / strata * resolution, GPUCameraPinhole.cuh:236-237
// and GenerationKernels.cuh:77-78), kept inside its own pixel.
__device__ __forceinline__ void camera_ray(const Cam& cam, uint32_t w, uint32_t h, uint32_t S, uint32_t seed,
                                           uint32_t gid, uint32_t i, V3& o, V3& d, uint64_t& rng_state,
                                           float2* coord = nullptr)
{
    const uint32_t spp = S * S;
    const float dxp = cam.pw / (float)(w * S);
    const float dyp = cam.ph / (float)(h * S);
    const uint32_t pix = i / spp, s = i - pix * spp;
    const uint32_t px = pix % w, py = pix / w, sx = s % S, sy = s / S;
    Rng rng{path_rng_init(seed, gid), pcg32_inc(gid)};
    const float jx = rng.uniform();
    const float jy = rng.uniform();
    float sdx = (float)(px * S + sx) * dxp;
    float sdy = (float)(py * S + sy) * dyp;
    sdx += jx * dxp;
    sdy += jy * dyp;
    const V3 sp = cam.bl + (scale(cam.right, sdx) + scale(cam.up, sdy));
    d = normalize(sp - cam.pos);
    o = cam.pos;
    rng_state = rng.state;
    if (coord) {
        float cx = ((float)(px * S + sx) + jx) / (float)(w * S) * (float)w;
        float cy = ((float)(py * S + sy) + jy) / (float)(h * S) * (float)h;
        // rounding may land on a neighbour's edge: keep [px, px + 1)
        const float ex = __int_as_float(__float_as_int((float)(px + 1)) - 1);
        const float ey = __int_as_float(__float_as_int((float)(py + 1)) - 1);
        cx = fminf(fmaxf(cx, (float)px), ex);
        cy = fminf(fmaxf(cy, (float)py), ey);
        *coord = make_float2(cx, cy);
    }
}

// Per-request constants of the shade step
struct ShadeReq {
    Materials mats;
    const float4* leaves;
    float* image;            // accumulation, [pixels][3]
    uint32_t spp, sample_base, rr_start, max_depth;
    const uint32_t* lights;  // leaf index of every emissive triangle, in primitive order
    uint32_t n_lights;
    uint32_t nee;            // next-event estimation ("NextEventEstimation")
    uint32_t mis;            // power-heuristic MIS of NEE and BxDF rays ("DirectLightMIS"), with nee only
    const float4* inst = nullptr;          // two-level scenes: instance records
    const uint32_t* light_inst = nullptr;  // two-level scenes: instance of every light
    const float4* tbn = nullptr;           // smooth shading: [primitive][3] corner rotations (w, x, y, z)
};

// World-space corners of leaf li (of instance ii in a two-level scene: the
// local corners taken to the world by the instance's local-to-world matrix,
// GPUTransformSingle::LocalToWorld, GPUTransformSingle.cuh:160-166).
template <class Q>
__device__ __forceinline__ void leaf_corners(const Q& q, uint32_t li, uint32_t ii, V3& p0, V3& p1, V3& p2)
{
    const float4 l0 = q.leaves[3 * (size_t)li + 0];
    const float4 l1 = q.leaves[3 * (size_t)li + 1];
    const float4 l2 = q.leaves[3 * (size_t)li + 2];
    p0 = v3(l0.x, l0.y, l0.z);
    p1 = v3(l1.x, l1.y, l1.z);
    p2 = v3(l2.x, l2.y, l2.z);
    if (q.inst) {
        const float4* rec = q.inst + INST_F4 * (size_t)ii;
        const float4 m0 = rec[3], m1 = rec[4], m2 = rec[5];
        p0 = xf_point(m0, m1, m2, p0);
        p1 = xf_point(m0, m1, m2, p1);
        p2 = xf_point(m0, m1, m2, p2);
    }
}

// Next-bounce state of a surviving path
struct RayNext {
    V3 o, d, T;
    uint32_t dm;
    uint64_t rng;
    float pdf;
};

// TracerFunctions.cuh:302-309 (power 2, one sample per strategy)
__device__ __forceinline__ float power_heuristic(float pdf0, float pdf1)
{
    const float w0 = 1.0f * pdf0;
    const float w1 = 1.0f * pdf1;
    return (w0 * w0) / (w0 * w0 + w1 * w1);
}

// TriFunctions::Area (GPUPrimitiveTriangle.h:321-340)
__device__ __forceinline__ float tri_area(V3 p0, V3 p1, V3 p2)
{
    const V3 c = cross(p1 - p0, p2 - p0);
    return sqrtf(dot(c, c)) * 0.5f;
}

// Solid-angle pdf of reaching a point of light triangle (p0, p1, p2) at
// distance t along unit direction d: GPULight<PGroup>::Pdf, hit form
// (GPULightPrimitive.cuh:245-256, PositionPdfFromHit = 1 / area) — the
// conditional of GPULightSamplerUniform::Pdf (GPULightSamplerUniform.cuh:166-192).
__device__ __forceinline__ float light_pdf_hit(V3 p0, V3 p1, V3 p2, float t, V3 d)
{
    const float pdf = 1.0f / tri_area(p0, p1, p2);
    const V3 n = normalize(cross(p1 - p0, p2 - p0));
    const float ndl = fabsf(dot(n, -d));
    return pdf * t * t / ndl;
}

// GPULightSamplerUniform::SampleLight (GPULightSamplerUniform.cuh:82-130)
// over the scene's emissive triangles, with GPULight<PGroup>::Sample
// (GPULightPrimitive.cuh:139-166) and the triangle's uniform point
// (GPUPrimitiveTriangle.h:65-110, Osada 2002).  Draws three numbers: light
// index, then the two barycentric numbers.  Returns the light's leaf, the
// unit direction from p, the squared distance and the pdf (solid angle x
// 1 / light count).  A primitive-backed light is traced to its first hit
// (tMax = FLT_MAX), so no distance is returned.
__device__ __forceinline__ uint32_t sample_light(const ShadeReq& q, V3 p, Rng& rng, V3& dir, float& dist2,
                                                 float& pdf, uint32_t& inst)
{
    float r1 = rng.uniform();
    r1 *= (float)q.n_lights;
    uint32_t idx = (uint32_t)r1;
    if (idx == q.n_lights) idx--;
    const uint32_t leaf = q.lights[idx];
    const uint32_t li_inst = q.light_inst ? q.light_inst[idx] : 0u;
    V3 p0, p1, p2;
    leaf_corners(q, leaf, li_inst, p0, p1, p2);
    const float xi0 = rng.uniform();
    const float xi1 = rng.uniform();
    const float s1 = sqrtf(xi0);
    const float s2 = xi1;
    const float a = 1.0f - s1;
    const float b = (1.0f - s2) * s1;
    const float c = s1 * s2;
    pdf = 1.0f / tri_area(p0, p1, p2);
    const V3 n = normalize(cross(p1 - p0, p2 - p0));
    const V3 lp = (scale(p0, a) + scale(p1, b)) + scale(p2, c);
    dir = lp - p;
    dist2 = dot(dir, dir);
    const float dist = sqrtf(dist2);
    dir = scale(dir, 1.0f / dist);
    const float ndl = fabsf(dot(n, -dir));
    pdf *= dist2 / ndl;
    pdf *= 1.0f / (float)q.n_lights;
    inst = li_inst;
    return leaf;
}

__device__ __forceinline__ void add_radiance(const ShadeReq& q, uint32_t sample, V3 c)
{
    float* px = q.image + (size_t)(sample / q.spp) * 3;
    atomicAdd(px + 0, c.x);
    atomicAdd(px + 1, c.y);
    atomicAdd(px + 2, c.z);
}

// PathTracerKC.cuh:337-426 (light hit) and 430-648 (surface work) for ray
// rid; with learn the WFPG path-node bookkeeping (WFPGTracerKC.cuh:202-520)
// and with guide the one-sample MIS of guide and BxDF
// (WFPGTracerKC.cuh:1480-1553, pdf = r p_guide + (1 - r) p_bxdf,
// weight = f cos / pdf).  LEARN / GUIDE / NEE compile the three parts in;
// learn / guide / q.nee switch them per request.  The ray's depth is dm & 0xFF (1 = camera ray); at depth
// max_depth only emission is collected.  Returns true when the path
// continues, with its next state in nx; with NEE, has_nee says whether a
// next-event ray was spawned, in nee (traced with the next bounce).
template <bool LEARN, bool GUIDE, bool NEE, class Q, class G>
__device__ __forceinline__ bool shade_ray(const PathBuf& in, const HitBuf& hit, uint32_t rid, const Q& q,
                                          const G& ga, bool learn, bool guide, RayNext& nx, RayNext& nee,
                                          bool& has_nee)
{
    has_nee = false;
    const uint32_t key = hit.key[rid];
    const uint32_t batch = key >> 24, id = key & 0xFFFFFFu;
    if (batch == MRH_BATCH_NULL) return false;
    const V3 T = v3(in.tr[rid], in.tg[rid], in.tb[rid]);
    const uint32_t sample = in.sample[rid];
    const uint32_t dm = in.dm[rid];
    const uint32_t depth = dm & 0xFFu, medium = dm >> DM_MEDIUM_SHIFT;
    const Materials& mats = q.mats;
    if (NEE && (dm & DM_NEE)) {
        // a NEE ray counts only where it hits the light it was sent to
        // (isCorrectNEERay); any other hit, or a miss, means occluded
        const uint64_t want = in.rng[rid];
        if (batch == MRH_BATCH_EMISSIVE && hit.leaf[rid] == (uint32_t)want &&
            (!hit.inst || hit.inst[rid] == (uint32_t)(want >> 32))) {
            const float* le = mats.emissive + (size_t)id * 3;
            add_radiance(q, sample, v3(le[0] * T.x, le[1] * T.y, le[2] * T.z));
        }
        return false;
    }
    if (batch == MRH_BATCH_EMISSIVE) {
        const float* le = mats.emissive + (size_t)id * 3;
        V3 c = v3(le[0] * T.x, le[1] * T.y, le[2] * T.z);
        bool counts = true;
        if (NEE && q.nee && depth > 1u && !(dm & DM_SPEC)) {
            // a path ray that left a non-specular vertex, which already sent
            // a NEE ray: without MIS the light was counted by that ray, with
            // MIS this hit takes the BxDF strategy's power-heuristic weight
            if (!q.mis) {
                counts = false;
            } else {
                V3 p0, p1, p2;
                leaf_corners(q, hit.leaf[rid], hit.inst ? hit.inst[rid] : 0u, p0, p1, p2);
                const V3 d = normalize(v3(in.dx[rid], in.dy[rid], in.dz[rid]));
                const float pc = light_pdf_hit(p0, p1, p2, hit.t[rid], d);
                const float pm = 1.0f / (float)q.n_lights;
                c = scale(c, power_heuristic(in.pdf[rid], pc * pm));
            }
        }
        if (counts) add_radiance(q, sample, c);
        if (LEARN && learn) {
            // radiance arriving along each earlier vertex's outgoing ray:
            // path contribution over that vertex's throughput
            // (PathNode.cuh:152-190, per channel, then luminance).  The
            // guide learns the radiance that arrived, whatever share of it
            // the image counted (a NEE / MIS weight is an estimator's
            // choice); next-event rays do not propagate.
            const V3 cl = mul(v3(le[0], le[1], le[2]), T);
            for (uint32_t k = 0; k + 1 < depth; k++) {
                const size_t r = (size_t)k * ga.rec.stride + sample;
                if (ga.rec.code[r] == SVO_NO_NODE) continue;
                const float fr = ga.rec.tr[r], fg = ga.rec.tg[r], fb = ga.rec.tb[r];
                const V3 L = v3(fr == 0.0f ? 0.0f : cl.x / fr, fg == 0.0f ? 0.0f : cl.y / fg,
                                fb == 0.0f ? 0.0f : cl.z / fb);
                ga.rec.rad[r] = luminance(L);
                if (ga.rec.vtx) ga.rec.vtx[7 * r + 6] = ga.rec.rad[r];
            }
        }
        return false;
    }
    if (depth >= q.max_depth) return false;
    V3 c0, c1, c2;
    leaf_corners(q, hit.leaf[rid], hit.inst ? hit.inst[rid] : 0u, c0, c1, c2);
    const V3 d = v3(in.dx[rid], in.dy[rid], in.dz[rid]);
    const uint32_t li = hit.leaf[rid];
    const Surface s = q.tbn ? make_surface_tbn(c0, c1, c2, hit.ba[rid], hit.bb[rid], d,
                                               q.tbn + 3 * (size_t)__float_as_uint(q.leaves[3 * (size_t)li].w))
                            : make_surface(c0, c1, c2, hit.ba[rid], hit.bb[rid], d);
    const V3 wi = -normalize(d);
    Rng rng{in.rng[rid], pcg32_inc(q.sample_base + sample)};
    if (NEE && q.nee && q.n_lights &&
        (batch == MRH_BATCH_LAMBERT || (batch == MRH_BATCH_METAL && (1.0f - mats.metal[(size_t)id * 10]) < 0.99f))) {
        // PathTracerKC.cuh:493-575: one light sample per non-specular vertex
        V3 ldir;
        float dist2, pl;
        uint32_t linst;
        const uint32_t lleaf = sample_light(q, s.pos, rng, ldir, dist2, pl, linst);
        float pb;
        V3 f;
        if (batch == MRH_BATCH_LAMBERT) {
            const float* a = mats.lambert + (size_t)id * 3;
            f = lambert_eval(s, v3(a[0], a[1], a[2]), ldir, pb);
        } else {
            f = metal_eval(s, wi, mats.metal + (size_t)id * 10, ldir, pb);
        }
        float pn = pl;
        if (q.mis) {
            pn /= power_heuristic(pl, pb);
            if (isnan(pn)) pn = 0.0f;
        }
        V3 nf = mul(T, f);
        nf = (pn == 0.0f) ? v3(0.0f, 0.0f, 0.0f) : divs(nf, pn);
        // a light point on the vertex itself has no direction: no ray
        if (dist2 > 0.0f && !is_zero(nf)) {
            nee.o = nudge(s.pos, s.geo_n);
            nee.d = ldir;
            nee.T = nf;
            nee.dm = depth | DM_NEE;
            nee.rng = (uint64_t)lleaf | ((uint64_t)linst << 32);   // the light's leaf and instance
            nee.pdf = 0.0f;
            has_nee = true;
        }
    }
    BsdfOut o;
    o.medium = medium;
    bool guided = false;
    uint32_t gnode = 0;
    float gtotal = 0.0f;
    GuideBins gbins;
    V3 ns = v3(0.0f, 0.0f, 1.0f);
    uint32_t wfl = 0;
    if (GUIDE && guide && ga.wf_flag) {
        wfl = ga.wf_flag[rid];
        guided = wfl != 0u;
    } else if (GUIDE && guide &&
        (batch == MRH_BATCH_LAMBERT || (batch == MRH_BATCH_METAL && (1.0f - mats.metal[(size_t)id * 10]) < 0.99f))) {
        ns = rotate(conj(s.w2t), v3(0.0f, 0.0f, 1.0f));
        gnode = svo_lookup_guide(ga.svo,
                                 svo_code(s.pos, ga.svo.lox, ga.svo.loy, ga.svo.loz, ga.svo.inv_cell, ga.svo.depth),
                                 ga.min_samples);
        gbins = guide_bins(ga.svo, gnode);
        gtotal = guide_total(gbins, ns);
        guided = gtotal > 0.0f;
    }
    if (GUIDE && guided) {
        V3 dir;
        float pg = 0.0f, pb = 0.0f;
        V3 fcos;
        if (ga.wf_flag) {
            // the direction was drawn before the bounce's fields were
            // sampled (WFPGTracerKC.cuh:1490-1553); the BxDF is evaluated
            // there as the path work does (:389-430)
            const bool bx = wfl == 2u;
            const float* dd = (bx ? ga.wf_bdir : ga.wf_gdir) + 3 * (size_t)rid;
            dir = v3(dd[0], dd[1], dd[2]);
            pg = bx ? ga.wf_epdf[rid] : ga.wf_gpdf[rid];
            if (wfl == 3u) {
                // the BxDF's own sample failed: the path ends, as without the guide
                fcos = v3(0.0f, 0.0f, 0.0f);
                pg = 0.0f;
            } else if (batch == MRH_BATCH_LAMBERT) {
                const float* a = mats.lambert + (size_t)id * 3;
                fcos = lambert_eval(s, v3(a[0], a[1], a[2]), dir, pb);
            } else {
                fcos = metal_eval(s, wi, mats.metal + (size_t)id * 10, dir, pb);
            }
        } else if (rng.uniform() < ga.ratio) {
            const float u0 = rng.uniform(), u1 = rng.uniform(), u2 = rng.uniform();
            dir = guide_sample(gbins, ns, gtotal, u0, u1, u2, pg);
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
    } else if (batch == MRH_BATCH_REFLECT) {
        reflect_sample(s, wi, mats.reflect + (size_t)id * 4, o);
    } else {
        dielectric_sample(s, wi, mats.dielectric + (size_t)id * 4, id, medium, mats.dielectric, rng, o);
    }
    V3 t2 = mul(T, o.refl);
    t2 = (o.pdf == 0.0f) ? v3(0.0f, 0.0f, 0.0f) : divs(t2, o.pdf);
    const float avg = dot(t2, v3(0.333f, 0.333f, 0.333f));
    bool terminate = false;
    if (!o.specular && depth > q.rr_start) {
        // TracerFunctions.cuh:321-330
        const float pr = avg < 0.005f ? 0.005f : (avg > 1.0f ? 1.0f : avg);
        if (rng.uniform() >= pr) terminate = true;
        else t2 = scale(t2, 1.0f / pr);
    }
    if ((terminate && !o.specular) || is_zero(t2)) return false;
    if (LEARN && learn) {
        // path node of this vertex (PPGTracerKC.cuh:605-620)
        const size_t r = (size_t)(depth - 1u) * ga.rec.stride + sample;
        ga.rec.code[r] = svo_code(s.pos, ga.lox, ga.loy, ga.loz, ga.inv_cell, ga.depth);
        ga.rec.bin[r] = dir_to_bin(normalize(o.d));
        ga.rec.tr[r] = t2.x;
        ga.rec.tg[r] = t2.y;
        ga.rec.tb[r] = t2.z;
        ga.rec.rad[r] = 0.0f;
        if (ga.rec.vtx) {
            float* x = ga.rec.vtx + 7 * r;
            x[0] = s.pos.x; x[1] = s.pos.y; x[2] = s.pos.z;
            x[3] = wi.x; x[4] = wi.y; x[5] = wi.z;
            x[6] = 0.0f;
        }
    }
    nx.o = o.o;
    nx.d = o.d;
    nx.T = t2;
    nx.dm = ((depth + 1u) & 0xFFu) | (o.specular ? DM_SPEC : 0u) | (o.medium << DM_MEDIUM_SHIFT);
    nx.rng = rng.state;
    nx.pdf = o.pdf;
    return true;
}

// Appends the surviving rays of a wave into out at the counter *ctr
// (one atomic per wave); with_pdf also stores the ray's BxDF pdf.
__device__ __forceinline__ void append_ray(bool alive, const RayNext& nx, uint32_t sample, PathBuf& out,
                                           uint32_t* ctr, uint32_t* slot_out, uint32_t slot, bool with_pdf)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t m = __ballot(alive);
    if (!m) return;
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1u;
    uint32_t slot0 = 0;
    if (lane == leader) slot0 = atomicAdd(ctr, (uint32_t)__popcll(m));
    slot0 = __shfl(slot0, (int)leader, 64);
    if (alive) {
        const uint32_t k = slot0 + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        out.ox[k] = nx.o.x; out.oy[k] = nx.o.y; out.oz[k] = nx.o.z;
        out.dx[k] = nx.d.x; out.dy[k] = nx.d.y; out.dz[k] = nx.d.z;
        out.tr[k] = nx.T.x; out.tg[k] = nx.T.y; out.tb[k] = nx.T.z;
        out.sample[k] = sample;
        out.dm[k] = nx.dm;
        out.rng[k] = nx.rng;
        if (with_pdf) out.pdf[k] = nx.pdf;
        if (slot_out) slot_out[k] = slot;
    }
}

// append_ray with one atomic per block of 256 instead of one per wave: the
// four waves' survivor counts are scanned in LDS (s_wave: 4 words) and one
// lane claims the block's range.  All rays of a bounce append to one
// counter, so this cuts that address's atomics fourfold (a wave-level
// append left the shade step waiting on them).  Every thread of the block
// must call it.
__device__ __forceinline__ void append_ray_block(bool alive, const RayNext& nx, uint32_t sample, PathBuf& out,
                                                 uint32_t* ctr, bool with_pdf, uint32_t* s_wave,
                                                 uint32_t* slot_out = nullptr, uint32_t slot = 0u)
{
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t m = __ballot(alive);
    if (lane == 0) s_wave[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t c0 = s_wave[0], c1 = s_wave[1], c2 = s_wave[2], c3 = s_wave[3];
        const uint32_t total = c0 + c1 + c2 + c3;
        const uint32_t b = total ? atomicAdd(ctr, total) : 0u;
        s_wave[0] = b;
        s_wave[1] = b + c0;
        s_wave[2] = b + c0 + c1;
        s_wave[3] = b + c0 + c1 + c2;
    }
    __syncthreads();
    if (alive) {
        const uint32_t k = s_wave[wave] + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        out.ox[k] = nx.o.x; out.oy[k] = nx.o.y; out.oz[k] = nx.o.z;
        out.dx[k] = nx.d.x; out.dy[k] = nx.d.y; out.dz[k] = nx.d.z;
        out.tr[k] = nx.T.x; out.tg[k] = nx.T.y; out.tb[k] = nx.T.z;
        out.sample[k] = sample;
        out.dm[k] = nx.dm;
        out.rng[k] = nx.rng;
        if (with_pdf) out.pdf[k] = nx.pdf;
        if (slot_out) slot_out[k] = slot;
    }
    __syncthreads();
}

}  // namespace
}  // namespace mrh
