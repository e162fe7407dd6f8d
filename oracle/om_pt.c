/*
 * om_pt.c — oracle: camera rays, BxDFs, surface generation and the plain
 * path tracer, restated from the reference's device code.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Paths are traced depth-first, one sample at a time; the wavefront order of
 * the GPU path does not change any path's arithmetic because every path owns
 * its PCG32 stream, keyed by (seed, global sample id).  Per-path arithmetic is
 * FP32 in the same operation order as the HIP kernels; pixel accumulation is
 * FP64.
 */
#include "oracle.h"
#include "om.h"

#include <float.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define OM_PI_F     3.14159265358979323846f
#define OM_INV_PI_F 0.31830988618379067154f

/* ---------------------------- helpers -------------------------------- */

void om_sincos_2pi_f(float u, float* s, float* c) { om_sincos_2pi(u, s, c); }

/* RayLib/HemiDistribution.h:32-51 (HemiUniformCDF); cos/sin of 2*pi*xi1
 * evaluated by om_sincos_2pi.  pdf = (1/pi) * 0.5. */
static om_v3 hemi_uniform(float xi0, float xi1, float* pdf)
{
    float c0 = sqrtf(fmaxf(0.0f, 1.0f - xi0 * xi0));
    float s, c;
    om_sincos_2pi(xi1, &s, &c);
    *pdf = OM_INV_PI_F * 0.5f;
    return v3(c0 * c, c0 * s, xi0);
}

void om_hemi_uniform(float xi0, float xi1, float dir[3], float* pdf)
{
    om_v3 d = hemi_uniform(xi0, xi1, pdf);
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
}

/* RayLib/Ray.hpp:503-545 (Ray<float>::Nudge, Ray Tracing Gems I ch. 6),
 * curvature offset 0. */
static om_v3 nudge(om_v3 p, om_v3 n)
{
    const float origin = 1.0f / 32.0f;
    const float float_scale = 1.0f / 65536.0f;
    const float int_scale = 256.0f;
    float pv[3] = {p.x, p.y, p.z};
    float nv[3] = {n.x, n.y, n.z};
    float out[3];
    for (int i = 0; i < 3; i++) {
        int32_t ofi = (int32_t)(int_scale * nv[i]);
        int32_t pi;
        memcpy(&pi, &pv[i], 4);
        pi += (pv[i] < 0.0f) ? -ofi : ofi;
        float pf;
        memcpy(&pf, &pi, 4);
        out[i] = (fabsf(pv[i]) < origin) ? (pv[i] + float_scale * nv[i]) : pf;
    }
    return v3(out[0], out[1], out[2]);
}

void om_nudge(const float p[3], const float n[3], float out[3])
{
    om_v3 r = nudge(v3(p[0], p[1], p[2]), v3(n[0], n[1], n[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

void om_rot_between_z(const float n[3], float q[4])
{
    om_q r = q_rot_between_z(v3(n[0], n[1], n[2]));
    q[0] = r.w; q[1] = r.x; q[2] = r.y; q[3] = r.z;
}

void om_quat_rotate(const float q[4], const float v[3], float out[3])
{
    om_v3 r = q_rotate(q_make(q[0], q[1], q[2], q[3]), v3(v[0], v[1], v[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* TracerFunctions.cuh:8-30 */
float om_fresnel_dielectric(float cos_in, float ior_in, float ior_out)
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

/* TracerFunctions.cuh:33-79 (per colour channel) */
static float fresnel_conductor1(float cos_in, float eta, float k)
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
    float perp_t1 = a2b2 + cos2;
    float perp_t2 = a * (2.0f * cos_in);
    float perp = (perp_t1 - perp_t2) / (perp_t1 + perp_t2);
    float par_t1 = a2b2 * cos2 + sin2_4;
    float par_t2 = perp_t2 * sin2_2;
    float par = perp * (par_t1 - par_t2) / (par_t1 + par_t2);
    return (par + perp) * 0.5f;
}

void om_fresnel_conductor(float cos_in, const float eta[3], const float k[3], float out[3])
{
    for (int i = 0; i < 3; i++) out[i] = fresnel_conductor1(cos_in, eta[i], k[i]);
}

/* TracerFunctions.cuh:81-90 */
float om_dggx(float ndh, float alpha)
{
    float a2 = alpha * alpha;
    float denom = ndh * ndh * (a2 - 1.0f) + 1.0f;
    denom = denom * denom;
    denom *= OM_PI_F;
    return a2 / denom;
}

/* TracerFunctions.cuh:92-101 */
static float lambda_smith(om_v3 v, float alpha)
{
    om_v3 s = v3_mul(v, v);
    float a2 = alpha * alpha;
    float inner = a2 * (s.x + s.y) / s.z;
    float l = sqrtf(1.0f + inner) - 1.0f;
    l *= 0.5f;
    return l;
}

static float gsmith_single(om_v3 v, float alpha) { return 1.0f / (1.0f + lambda_smith(v, alpha)); }

static float gsmith_corr(om_v3 wo, om_v3 wi, float alpha)
{
    return 1.0f / (lambda_smith(wo, alpha) + lambda_smith(wi, alpha) + 1.0f);
}

float om_gsmith_single(const float v[3], float alpha) { return gsmith_single(v3(v[0], v[1], v[2]), alpha); }
float om_gsmith_correlated(const float wo[3], const float wi[3], float alpha)
{
    return gsmith_corr(v3(wo[0], wo[1], wo[2]), v3(wi[0], wi[1], wi[2]), alpha);
}

/* TracerFunctions.cuh:186-196 */
static float vndf_pdf(om_v3 v, om_v3 h, float alpha)
{
    float vdh = fmaxf(0.0f, v3_dot(h, v));
    float ndh = fmaxf(0.0f, h.z);
    float ndv = fmaxf(0.0f, v.z);
    float d = om_dggx(ndh, alpha);
    float g1 = gsmith_single(v, alpha);
    return (ndv == 0.0f) ? 0.0f : (vdh * d * g1 / ndv);
}

float om_vndf_pdf(const float v[3], const float h[3], float alpha)
{
    return vndf_pdf(v3(v[0], v[1], v[2]), v3(h[0], h[1], h[2]), alpha);
}

/* TracerFunctions.cuh:198-258 (Heitz 2018, isotropic) */
static om_v3 vndf_sample(om_v3 v, float a, float xi0, float xi1, float* pdf)
{
    om_v3 vh = v3_normalize(v3(a * v.x, a * v.y, v.z));
    float lensq = 0.0f;
    lensq += vh.x * vh.x;
    lensq += vh.y * vh.y;
    om_v3 t1v = (lensq > 0.0f) ? v3_divs(v3(-vh.y, vh.x, 0.0f), sqrtf(lensq)) : v3(1.0f, 0.0f, 0.0f);
    om_v3 t2v = v3_cross(vh, t1v);
    float r = sqrtf(xi0);
    float sp, cp;
    om_sincos_2pi(xi1, &sp, &cp);
    float t1 = r * cp;
    float t2 = r * sp;
    float s = 0.5f * (1.0f + vh.z);
    t2 = (1.0f - s) * sqrtf(1.0f - t1 * t1) + s * t2;
    float val = 1.0f - t1 * t1 - t2 * t2;
    float sq = sqrtf(fmaxf(0.0f, val));
    om_v3 nh = v3_add(v3_add(v3_scale(t1v, t1), v3_scale(t2v, t2)), v3_scale(vh, sq));
    om_v3 nm = v3(a * nh.x, a * nh.y, fmaxf(0.0f, nh.z));
    float len2 = v3_dot(nm, nm);
    if (len2 < 1.0e-5f) nm = v3(0.0f, 0.0f, 1.0f);
    else nm = v3_scale(nm, 1.0f / sqrtf(len2));
    *pdf = vndf_pdf(v, nm, a);
    return nm;
}

void om_vndf_sample(const float v[3], float alpha, float xi0, float xi1, float h[3], float* pdf)
{
    om_v3 r = vndf_sample(v3(v[0], v[1], v[2]), alpha, xi0, xi1, pdf);
    h[0] = r.x; h[1] = r.y; h[2] = r.z;
}

/* ---------------------------- camera --------------------------------- */

/* GPUCameraPinhole.cuh:140-170 */
void om_camera_setup(const float pos[3], const float gaze[3], const float up[3],
                     float fov_x, float near_d, float far_d, uint32_t w, uint32_t h,
                     om_camera* cam)
{
    float fov_y = (float)(2.0 * atan(tan((double)fov_x * 0.5) * (double)h / (double)w));
    om_v3 p = v3(pos[0], pos[1], pos[2]);
    om_v3 g = v3(gaze[0], gaze[1], gaze[2]);
    om_v3 u = v3(up[0], up[1], up[2]);
    float wh = tanf(fov_x * 0.5f) * near_d;
    float hh = tanf(fov_y * 0.5f) * near_d;
    om_v3 gd = v3_sub(g, p);
    om_v3 right = v3_normalize(v3_cross(gd, u));
    u = v3_normalize(v3_cross(right, gd));
    gd = v3_normalize(v3_cross(u, right));
    om_v3 bl = v3_add(v3_sub(v3_sub(p, v3_scale(right, wh)), v3_scale(u, hh)), v3_scale(gd, near_d));
    cam->position[0] = p.x; cam->position[1] = p.y; cam->position[2] = p.z;
    cam->bottom_left[0] = bl.x; cam->bottom_left[1] = bl.y; cam->bottom_left[2] = bl.z;
    cam->right[0] = right.x; cam->right[1] = right.y; cam->right[2] = right.z;
    cam->up[0] = u.x; cam->up[1] = u.y; cam->up[2] = u.z;
    cam->plane_size[0] = wh * 2.0f;
    cam->plane_size[1] = hh * 2.0f;
    cam->near_far[0] = near_d;
    cam->near_far[1] = far_d;
}

/* GenerationKernels.cuh:29-108 + GPUCameraPinhole.cuh:209-240: stratified
 * pinhole sample.  Sample ids run pixel-major: sample_id =
 * (y*W + x)*S^2 + (sy*S + sx). */
static void camera_ray_jittered(const om_camera* cam, uint32_t w, uint32_t h, uint32_t S,
                                uint32_t sample_id, float jx, float jy, om_v3* o, om_v3* d,
                                float* coord)
{
    uint32_t spp = S * S;
    uint32_t pix = sample_id / spp;
    uint32_t s = sample_id - pix * spp;
    uint32_t px = pix % w, py = pix / w;
    uint32_t sx = s % S, sy = s / S;
    float dx = cam->plane_size[0] / (float)(w * S);
    float dy = cam->plane_size[1] / (float)(h * S);
    float sdx = (float)(px * S + sx) * dx;
    float sdy = (float)(py * S + sy) * dy;
    sdx += jx * dx;
    sdy += jy * dy;
    om_v3 bl = v3(cam->bottom_left[0], cam->bottom_left[1], cam->bottom_left[2]);
    om_v3 r = v3(cam->right[0], cam->right[1], cam->right[2]);
    om_v3 u = v3(cam->up[0], cam->up[1], cam->up[2]);
    om_v3 pos = v3(cam->position[0], cam->position[1], cam->position[2]);
    om_v3 sp = v3_add(bl, v3_add(v3_scale(r, sdx), v3_scale(u, sdy)));
    *d = v3_normalize(v3_sub(sp, pos));
    *o = pos;
    if (coord) {
        /* GPUCameraPinhole.cuh:236-237 + GenerationKernels.cuh:77-78: image
         * coordinate (stratum + jitter) / strata * resolution, kept inside
         * its own pixel */
        float cx = ((float)(px * S + sx) + jx) / (float)(w * S) * (float)w;
        float cy = ((float)(py * S + sy) + jy) / (float)(h * S) * (float)h;
        cx = fminf(fmaxf(cx, (float)px), nextafterf((float)(px + 1), 0.0f));
        cy = fminf(fmaxf(cy, (float)py), nextafterf((float)(py + 1), 0.0f));
        coord[0] = cx;
        coord[1] = cy;
    }
}

/* The stratum jitter is the sample stream's first two uniforms. */
static void camera_ray(const om_camera* cam, uint32_t w, uint32_t h, uint32_t S,
                       uint32_t sample_id, om_rng* rng, om_v3* o, om_v3* d, float* coord)
{
    float jx = om_uniform(rng);
    float jy = om_uniform(rng);
    camera_ray_jittered(cam, w, h, S, sample_id, jx, jy, o, d, coord);
}

/* A camera ray with a given stratum jitter (jx, jy in [0, 1]; 1 reaches the
 * far edge of the stratum), for closed-form checks of the camera basis. */
void om_camera_ray_at(const om_camera* cam, uint32_t w, uint32_t h, uint32_t samples,
                      uint32_t sample_id, float jx, float jy, float o[3], float d[3])
{
    om_v3 oo, dd;
    camera_ray_jittered(cam, w, h, samples, sample_id, jx, jy, &oo, &dd, NULL);
    o[0] = oo.x; o[1] = oo.y; o[2] = oo.z;
    d[0] = dd.x; d[1] = dd.y; d[2] = dd.z;
}

/* The raw PCG32 stream of pcg32_srandom_r(initstate, initseq) (O'Neill 2014),
 * and the SplitMix64 finaliser, exported for known-answer checks. */
void om_pcg32_stream(uint64_t initstate, uint64_t initseq, uint32_t n, uint32_t* out)
{
    om_rng r;
    om_pcg32_seed(&r, initstate, initseq);
    for (uint32_t i = 0; i < n; ++i) out[i] = om_pcg32(&r);
}

uint64_t om_splitmix64(uint64_t z) { return om_mix64(z); }

/* The first n uniforms of a path's stream (seed, global sample id). */
void om_path_uniforms(uint32_t seed, uint32_t sample_id, uint32_t n, float* out)
{
    om_rng r;
    om_path_rng(&r, seed, sample_id);
    for (uint32_t i = 0; i < n; ++i) out[i] = om_uniform(&r);
}

void om_camera_ray(const om_camera* cam, uint32_t w, uint32_t h, uint32_t samples,
                   uint32_t seed, uint32_t sample_id, float o[3], float d[3],
                   uint64_t* rng_state)
{
    om_rng rng;
    om_path_rng(&rng, seed, sample_id);
    om_v3 oo, dd;
    camera_ray(cam, w, h, samples, sample_id, &rng, &oo, &dd, NULL);
    o[0] = oo.x; o[1] = oo.y; o[2] = oo.z;
    d[0] = dd.x; d[1] = dd.y; d[2] = dd.z;
    *rng_state = rng.state;
}

/* ---------------------------- path tracer ---------------------------- */

typedef struct {
    om_v3 pos;
    om_v3 geo_n;      /* geometric normal, flipped to the ray side */
    om_q  w2t;        /* world -> tangent rotation */
    int   back;
} om_surface;

/* GPUPrimitiveTriangle.h:626-670 (GenBasicSurface) for flat-shaded,
 * two-sided triangles: every corner carries the face frame
 * conj(RotationBetweenZAxis(n)), so the barycentric SLERP is that frame.
 * Back-side hits flip the geometric normal and pre-multiply the frame by the
 * 180-degree rotation about tangent X (TANGENT_ROT = (0,1,0,0)). */
static om_surface make_surface(const float* tri, float ba, float bb, om_v3 ray_dir)
{
    om_surface s;
    om_v3 p0 = v3(tri[0], tri[1], tri[2]);
    om_v3 p1 = v3(tri[3], tri[4], tri[5]);
    om_v3 p2 = v3(tri[6], tri[7], tri[8]);
    float c = 1.0f - ba - bb;
    s.pos = v3_add(v3_add(v3_scale(p0, ba), v3_scale(p1, bb)), v3_scale(p2, c));
    om_v3 n = v3_normalize(v3_cross(v3_sub(p1, p0), v3_sub(p2, p0)));
    s.w2t = q_conj(q_rot_between_z(n));
    s.back = v3_dot(n, ray_dir) > 0.0f;
    if (s.back) {
        n = v3_neg(n);
        s.w2t = q_mul(q_make(0.0f, 1.0f, 0.0f, 0.0f), s.w2t);
    }
    s.geo_n = n;
    return s;
}

/* ------------------------- smooth shading frames ----------------------- */
/* RayLib/Vector.hpp:877-886 (OrthogonalVector) */
static om_v3 orthogonal_vector(om_v3 v)
{
    if (fabsf(v.x) > fabsf(v.y)) return v3_divs(v3(-v.z, 0.0f, v.x), sqrtf(v.x * v.x + v.z * v.z));
    return v3_divs(v3(0.0f, v.z, -v.y), sqrtf(v.y * v.y + v.z * v.z));
}

static float q_dot(om_q a, om_q b)
{
    float d = 0.0f;
    d += a.w * b.w;
    d += a.x * b.x;
    d += a.y * b.y;
    d += a.z * b.z;
    return d;
}
static om_q q_scale(om_q a, float s) { return q_make(a.w * s, a.x * s, a.y * s, a.z * s); }
static om_q q_add(om_q a, om_q b) { return q_make(a.w + b.w, a.x + b.x, a.y + b.y, a.z + b.z); }
/* Vector::Normalize: times the reciprocal of the length */
static om_q q_normalize(om_q q) { return q_scale(q, 1.0f / sqrtf(q_dot(q, q))); }

/* TransformGen::Space (RayLib/Quaternion.hpp:370-436): Mike Day's
 * matrix-to-quaternion conversion of the basis (x, y, z), normalised and
 * conjugated; x is negated when Cross(x, y) - z exceeds 0.1 in every axis */
static om_q space_quat(om_v3 x, om_v3 y, om_v3 z)
{
    om_v3 c = v3_sub(v3_cross(x, y), z);
    if (fabsf(c.x) > 0.1f && fabsf(c.y) > 0.1f && fabsf(c.z) > 0.1f) x = v3_neg(x);
    float t;
    om_q q;
    if (z.z < 0.0f) {
        if (x.x > y.y) {
            t = 1.0f + x.x - y.y - z.z;
            q = q_make(y.z - z.y, t, x.y + y.x, z.x + x.z);
        } else {
            t = 1.0f - x.x + y.y - z.z;
            q = q_make(z.x - x.z, x.y + y.x, t, y.z + z.y);
        }
    } else {
        if (x.x < -y.y) {
            t = 1.0f - x.x - y.y + z.z;
            q = q_make(x.y - y.x, z.x + x.z, y.z + z.y, t);
        } else {
            t = 1.0f + x.x + y.y + z.z;
            q = q_make(t, y.z - z.y, z.x - x.z, x.y - y.x);
        }
    }
    q = q_scale(q, 0.5f / sqrtf(t));
    return q_conj(q_normalize(q));
}

/* GPUPrimitiveTriangle.cpp:286-316 for a vertex without texture
 * coordinates: tangent OrthogonalVector(n), Gram-Schmidt, bitangent n x t */
static om_q vertex_frame(om_v3 n)
{
    om_v3 t = orthogonal_vector(n);
    if (v3_dot(t, t) < 1.0e-5f) t = orthogonal_vector(n);
    t = v3_normalize(t);
    om_v3 tn = v3_normalize(v3_sub(t, v3_scale(n, v3_dot(n, t))));
    om_v3 b = v3_cross(n, tn);
    return space_quat(tn, b, n);
}

/* acos on [0, 1] and sin on [0, pi/2] from the shared atan and sin/cos */
static float acos_unit(float x) { return 2.0f * om_atan_unit_f(sqrtf((1.0f - x) / (1.0f + x))); }
static float sin_q(float x)
{
    float s, c;
    om_sincos_2pi(x * 0.15915494309189534f, &s, &c);
    return s;
}

/* Quat::SLerp (RayLib/Quaternion.hpp:256-281) */
static om_q q_slerp(om_q a, om_q b, float t)
{
    float cos_t = q_dot(a, b);
    float cos_f = (cos_t >= 0.0f) ? cos_t : -cos_t;
    float s0, s1;
    if (cos_f < 1.0f - 1.0e-5f) {
        float angle = acos_unit(cos_f);
        float sa = sin_q(angle);
        s0 = sin_q(angle * (1.0f - t)) / sa;
        s1 = sin_q(angle * t) / sa;
    } else {
        s0 = 1.0f - t;
        s1 = t;
    }
    s1 = (cos_t >= 0.0f) ? s1 : -s1;
    return q_add(q_scale(a, s0), q_scale(b, s1));
}

/* Quat::BarySLerp (RayLib/Quaternion.hpp:283-318) */
static om_q q_bary_slerp(om_q q0, om_q q1, om_q q2, float a, float b)
{
    float c = 1.0f - a - b;
    if (fabsf(a + b) < 1.0e-5f) return q2;
    float ab = a / (a + b);
    return q_slerp(q_slerp(q1, q0, ab), q2, c);
}

/* GPUPrimitiveTriangle.h:626-670 with per-vertex frames: the corners'
 * world-to-tangent rotations blended by BarySLerp and normalised, flipped by
 * TANGENT_ROT on the back side; position and geometric normal as the flat
 * surface */
static om_surface make_surface_tbn(const float* tri, float ba, float bb, om_v3 ray_dir, const float* tq)
{
    om_surface s = make_surface(tri, ba, bb, ray_dir);
    om_q q0 = q_make(tq[0], tq[1], tq[2], tq[3]);
    om_q q1 = q_make(tq[4], tq[5], tq[6], tq[7]);
    om_q q2 = q_make(tq[8], tq[9], tq[10], tq[11]);
    om_q w = q_normalize(q_bary_slerp(q0, q1, q2, ba, bb));
    if (s.back) w = q_mul(q_make(0.0f, 1.0f, 0.0f, 0.0f), w);
    s.w2t = w;
    return s;
}

void om_orthogonal_vector(const float v[3], float out[3])
{
    om_v3 r = orthogonal_vector(v3(v[0], v[1], v[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void om_space_quat(const float x[3], const float y[3], const float z[3], float q[4])
{
    om_q r = space_quat(v3(x[0], x[1], x[2]), v3(y[0], y[1], y[2]), v3(z[0], z[1], z[2]));
    q[0] = r.w; q[1] = r.x; q[2] = r.y; q[3] = r.z;
}
void om_qslerp(const float a[4], const float b[4], float t, float out[4])
{
    om_q r = q_slerp(q_make(a[0], a[1], a[2], a[3]), q_make(b[0], b[1], b[2], b[3]), t);
    out[0] = r.w; out[1] = r.x; out[2] = r.y; out[3] = r.z;
}
void om_bary_slerp(const float q0[4], const float q1[4], const float q2[4], float a, float b, float out[4])
{
    om_q r = q_bary_slerp(q_make(q0[0], q0[1], q0[2], q0[3]), q_make(q1[0], q1[1], q1[2], q1[3]),
                          q_make(q2[0], q2[1], q2[2], q2[3]), a, b);
    out[0] = r.w; out[1] = r.x; out[2] = r.y; out[3] = r.z;
}
/* per-corner rotations of n triangles' corner normals [n][9] (normalised
 * first) -> [n][3][4] (w, x, y, z) */
void om_corner_frames(const float* normals, uint32_t n, float* out)
{
    for (size_t k = 0; k < (size_t)n * 3; k++) {
        om_v3 nv = v3_normalize(v3(normals[3 * k], normals[3 * k + 1], normals[3 * k + 2]));
        om_q q = vertex_frame(nv);
        out[4 * k + 0] = q.w; out[4 * k + 1] = q.x; out[4 * k + 2] = q.y; out[4 * k + 3] = q.z;
    }
}
/* world-to-tangent rotation of a smooth surface point (tests) */
void om_smooth_frame(const float tri[9], const float tq[12], float ba, float bb, const float dir[3], float q[4],
                     float normal[3])
{
    om_surface s = make_surface_tbn(tri, ba, bb, v3(dir[0], dir[1], dir[2]), tq);
    q[0] = s.w2t.w; q[1] = s.w2t.x; q[2] = s.w2t.y; q[3] = s.w2t.z;
    om_v3 nv = q_rotate(q_conj(s.w2t), v3(0.0f, 0.0f, 1.0f));
    normal[0] = nv.x; normal[1] = nv.y; normal[2] = nv.z;
}

typedef struct {
    om_v3 o, d;
    om_v3 refl;
    float pdf;
    int   specular;
    uint32_t medium;
} om_bsdf_out;

/* SimpleMaterialsKC.cuh:16-60 (LambertConstFuncs::Sample) */
static void lambert_sample(const om_surface* s, om_v3 albedo, om_rng* rng, om_bsdf_out* out)
{
    om_v3 normal = q_rotate(q_conj(s->w2t), v3(0.0f, 0.0f, 1.0f));
    float xi0 = om_uniform(rng);
    float xi1 = om_uniform(rng);
    float pdf;
    om_v3 dir = v3_normalize(hemi_uniform(xi0, xi1, &pdf));
    om_q q = q_rot_between_z(normal);
    dir = q_rotate(q, dir);
    float ndl = fabsf(v3_dot(normal, dir));
    out->o = nudge(s->pos, s->geo_n);
    out->d = dir;
    out->refl = v3_scale(v3_scale(albedo, ndl), OM_INV_PI_F);
    out->pdf = pdf;
    out->specular = 0;
}

/* MetalGlassMaterialsKC.cuh:24-150 (MetalMatFuncs::Sample) */
static void metal_sample(const om_surface* s, om_v3 wi, const float* m, om_rng* rng, om_bsdf_out* out)
{
    float rough = m[0];
    float alpha = rough * rough;
    om_v3 V = q_rotate(s->w2t, wi);
    float xi0 = om_uniform(rng);
    float xi1 = om_uniform(rng);
    float pdf;
    om_v3 H = vndf_sample(V, alpha, xi0, xi1, &pdf);
    float vdh = fmaxf(0.0f, v3_dot(V, H));
    om_v3 L = v3_sub(v3_scale(H, 2.0f * vdh), V);
    pdf = (vdh == 0.0f) ? 0.0f : pdf / (4.0f * vdh);
    float ldh = fmaxf(0.0f, v3_dot(L, H));
    float ndh = fmaxf(0.0f, H.z);
    float ndv = fmaxf(0.0f, V.z);
    om_v3 F = v3(fresnel_conductor1(vdh, m[1], m[4]),
                 fresnel_conductor1(vdh, m[2], m[5]),
                 fresnel_conductor1(vdh, m[3], m[6]));
    F = v3_mul(F, v3(m[7], m[8], m[9]));
    float D = om_dggx(ndh, alpha);
    float G = gsmith_corr(V, L, alpha);
    G = (ldh == 0.0f) ? 0.0f : G;
    G = (vdh == 0.0f) ? 0.0f : G;
    /* D * F * G * 0.25 / NdV, evaluated left to right per channel */
    om_v3 spec = v3(((D * F.x) * G) * 0.25f, ((D * F.y) * G) * 0.25f, ((D * F.z) * G) * 0.25f);
    spec = v3_divs(spec, ndv);
    if (ndv == 0.0f) spec = v3(0.0f, 0.0f, 0.0f);
    if (isinf(D) || isnan(D)) {
        if (rough == 0.0f) L = v3_sub(v3_scale(v3(0.0f, 0.0f, 1.0f), 2.0f * V.z), V);
        spec = v3_scale(F, G);
        float gs = gsmith_single(V, alpha);
        if (gs != 0.0f) spec = v3_divs(spec, gs);
        else spec = v3(0.0f, 0.0f, 0.0f);
        pdf = 1.0f;
    }
    out->d = q_rotate(q_conj(s->w2t), L);
    out->o = nudge(s->pos, s->geo_n);
    out->refl = spec;
    out->pdf = pdf;
    /* MetalMatFuncs::Specularity = 1 - roughness; specular at >= 0.99
     * (TracerConstants.h:6) */
    out->specular = (1.0f - rough) >= 0.99f;
}

/* SimpleMaterialsKC.cuh:172-290 (RefractMatFuncs::Sample).  Media: 0 is the
 * vacuum (IOR 1); medium k+1 is the interior of dielectric material k. */
static void dielectric_sample(const om_surface* s, om_v3 wi, const float* m, uint32_t mat_id,
                              uint32_t medium, const float* dielectric, om_rng* rng,
                              om_bsdf_out* out)
{
    om_v3 albedo = v3(m[0], m[1], m[2]);
    float ior = m[3];
    om_v3 normal = q_rotate(q_conj(s->w2t), v3(0.0f, 0.0f, 1.0f));
    int entering = !s->back;
    float from = (medium == 0u) ? 1.0f : dielectric[(size_t)(medium - 1u) * 4 + 3];
    float to = entering ? ior : 1.0f;
    float f = om_fresnel_dielectric(fabsf(v3_dot(wi, normal)), from, to);
    float xi = om_uniform(rng);
    out->specular = 1;
    if (xi < f) {
        float k = 2.0f * v3_dot(wi, normal);
        out->d = v3_sub(v3_scale(normal, k), wi);
        out->o = nudge(s->pos, s->geo_n);
        out->pdf = f;
        out->medium = medium;
        out->refl = v3_scale(albedo, f);
    } else {
        out->medium = entering ? (mat_id + 1u) : 0u;
        out->pdf = 1.0f - f;
        float ratio = from / to;
        float cos_in = v3_dot(normal, wi);
        float sin_in2 = fmaxf(0.0f, 1.0f - cos_in * cos_in);
        float sin_out2 = ratio * ratio * sin_in2;
        float cos_out = sqrtf(fmaxf(0.0f, 1.0f - sin_out2));
        if (sin_out2 >= 1.0f) {
            out->pdf = 0.0f;
            out->refl = v3(0.0f, 0.0f, 0.0f);
            out->d = wi;
            out->o = s->pos;
            return;
        }
        float kk = ratio * cos_in - cos_out;
        out->d = v3_add(v3_scale(v3_neg(wi), ratio), v3_scale(normal, kk));
        out->o = nudge(s->pos, v3_neg(s->geo_n));
        out->refl = v3_scale(v3_scale(albedo, 1.0f), 1.0f - f);
    }
}

/* SimpleMaterialsKC.cuh:105-151 (ReflectMatFuncs::Sample): mirror about the
 * shading normal with the albedo at roughness 0, nothing otherwise */
static void reflect_sample(const om_surface* s, om_v3 wi, const float* m, om_bsdf_out* out)
{
    om_v3 normal = q_rotate(q_conj(s->w2t), v3(0.0f, 0.0f, 1.0f));
    out->specular = 1;
    out->pdf = 1.0f;
    out->o = s->pos;
    out->d = wi;
    if (m[3] == 0.0f) {
        float k = 2.0f * v3_dot(wi, normal);
        out->d = v3_sub(v3_scale(normal, k), wi);
        out->o = nudge(s->pos, s->geo_n);
        out->refl = v3(m[0], m[1], m[2]);
    } else {
        out->refl = v3(0.0f, 0.0f, 0.0f);
    }
}

/* Lambert f*cos and its sampler's own pdf (uniform hemisphere, 1/(2 pi) on
 * the normal's side); SimpleMaterialsKC.cuh:80-101 for the evaluation. */
static om_v3 lambert_eval(const om_surface* s, om_v3 albedo, om_v3 dir, float* pdf)
{
    om_v3 normal = q_rotate(q_conj(s->w2t), v3(0.0f, 0.0f, 1.0f));
    float ndl = v3_dot(normal, dir);
    *pdf = (ndl > 0.0f) ? OM_INV_PI_F * 0.5f : 0.0f;
    return v3_scale(v3_scale(albedo, fmaxf(ndl, 0.0f)), OM_INV_PI_F);
}

/* MetalMatFuncs::Evaluate / Pdf (MetalGlassMaterialsKC.cuh:152-220) */
static om_v3 metal_eval(const om_surface* s, om_v3 wi, const float* m, om_v3 dir, float* pdf)
{
    float rough = m[0];
    float alpha = rough * rough;
    om_v3 L = q_rotate(s->w2t, dir);
    om_v3 V = q_rotate(s->w2t, wi);
    om_v3 H = v3_normalize(v3_add(L, V));
    float ndh = fmaxf(0.0f, H.z);
    float vdh = fmaxf(0.0f, v3_dot(V, H));
    float ldh = fmaxf(0.0f, v3_dot(L, H));
    float ndv = fmaxf(0.0f, V.z);
    float pspec = vndf_pdf(V, H, alpha);
    float D = om_dggx(ndh, alpha);
    pspec = (isnan(D) || isinf(D)) ? 0.0f : pspec;
    *pdf = (vdh == 0.0f) ? 0.0f : pspec / (4.0f * vdh);
    D = (isnan(D) || isinf(D)) ? 0.0f : D;
    float G = gsmith_corr(V, L, alpha);
    G = (ldh == 0.0f) ? 0.0f : G;
    G = (vdh == 0.0f) ? 0.0f : G;
    om_v3 F = v3(fresnel_conductor1(vdh, m[1], m[4]),
                 fresnel_conductor1(vdh, m[2], m[5]),
                 fresnel_conductor1(vdh, m[3], m[6]));
    F = v3_mul(F, v3(m[7], m[8], m[9]));
    om_v3 spec = v3(((D * F.x) * G) * 0.25f, ((D * F.y) * G) * 0.25f, ((D * F.z) * G) * 0.25f);
    spec = v3_divs(spec, ndv);
    return (ndv == 0.0f) ? v3(0.0f, 0.0f, 0.0f) : spec;
}

/* ------------------------- next-event estimation ---------------------- */

/* closest hit of one ray over the scene's traversal tree(s); *inst
 * receives the instance (0 in a single-level scene) */
static uint32_t closest_hit(const om_scene* sc, om_v3 o, om_v3 d, float tmin, float* tmax, float b[2],
                            uint32_t* inst)
{
    float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    *inst = 0;
    if (sc->two_level)
        return om_trace2l_one(sc->two_level, sc->tris, sc->tree_width, oo, dd, tmin, tmax, b, 0, inst, NULL, NULL);
    return sc->nodes4 ? om_trace4_one(sc->tris, sc->n, sc->sorted_idx, sc->nodes4, sc->root4, oo, dd, tmin, tmax, b,
                                      0, NULL, NULL)
                      : om_trace_one(sc->tris, sc->n, sc->sorted_idx, sc->left, sc->right, sc->node_aabb, sc->root,
                                     oo, dd, tmin, tmax, b, 0, NULL, NULL);
}

/* World-space corners of primitive prim (of instance inst in a two-level
 * scene: the local corners taken to the world by the instance's
 * local-to-world matrix, GPUTransformSingle::LocalToWorld,
 * GPUTransformSingle.cuh:160-166) */
static void prim_corners(const om_scene* sc, uint32_t prim, uint32_t inst, float out[9])
{
    const float* t = sc->tris + (size_t)prim * 9;
    if (!sc->two_level) {
        for (int k = 0; k < 9; k++) out[k] = t[k];
        return;
    }
    const float* m = sc->two_level->l2w + (size_t)inst * 12;
    for (int c = 0; c < 3; c++) om_xf_point(m, t + 3 * c, out + 3 * c);
}

/* HitKey of a hit: the instance's override where it has one */
static uint32_t prim_key_of(const om_scene* sc, uint32_t prim, uint32_t inst)
{
    if (sc->two_level && sc->two_level->inst_key && sc->two_level->inst_key[inst])
        return sc->two_level->inst_key[inst];
    return sc->prim_key[prim];
}

/* TracerFunctions.cuh:302-309: power heuristic, exponent 2, one sample each */
static float power_heuristic(float pdf0, float pdf1)
{
    float w0 = 1.0f * pdf0;
    float w1 = 1.0f * pdf1;
    return (w0 * w0) / (w0 * w0 + w1 * w1);
}

/* TriFunctions::Area (GPUPrimitiveTriangle.h:321-340) */
static float tri_area(om_v3 p0, om_v3 p1, om_v3 p2)
{
    om_v3 c = v3_cross(v3_sub(p1, p0), v3_sub(p2, p0));
    return sqrtf(v3_dot(c, c)) * 0.5f;
}

/* GPULight<PGroup>::Pdf, hit form (GPULightPrimitive.cuh:245-256):
 * (1 / area) * t^2 / |n . -d| for a unit direction d reaching the light at t */
static float light_pdf_hit(const float* tri, float t, om_v3 d)
{
    om_v3 p0 = v3(tri[0], tri[1], tri[2]), p1 = v3(tri[3], tri[4], tri[5]), p2 = v3(tri[6], tri[7], tri[8]);
    float pdf = 1.0f / tri_area(p0, p1, p2);
    om_v3 n = v3_normalize(v3_cross(v3_sub(p1, p0), v3_sub(p2, p0)));
    float ndl = fabsf(v3_dot(n, v3_neg(d)));
    return pdf * t * t / ndl;
}

/* GPULightSamplerUniform::SampleLight (GPULightSamplerUniform.cuh:82-130):
 * uniform light index, then GPULight<PGroup>::Sample
 * (GPULightPrimitive.cuh:139-166) with the triangle's uniform point
 * (GPUPrimitiveTriangle.h:65-110, Osada 2002: a = 1 - sqrt(x0),
 * b = (1 - x1) sqrt(x0), c = x1 sqrt(x0)).  Returns the light's primitive;
 * *pdf = solid-angle pdf x 1 / light count. */
static uint32_t sample_light_xi(const om_scene* sc, om_v3 p, float r1, float xi0, float xi1, om_v3* dir,
                                float* dist2, float* pdf, uint32_t* inst)
{
    r1 *= (float)sc->n_lights;
    uint32_t idx = (uint32_t)r1;
    if (idx == sc->n_lights) idx--;
    uint32_t prim = sc->lights[idx];
    *inst = sc->light_inst ? sc->light_inst[idx] : 0u;
    float t[9];
    prim_corners(sc, prim, *inst, t);
    om_v3 p0 = v3(t[0], t[1], t[2]), p1 = v3(t[3], t[4], t[5]), p2 = v3(t[6], t[7], t[8]);
    float s1 = sqrtf(xi0);
    float s2 = xi1;
    float a = 1.0f - s1;
    float b = (1.0f - s2) * s1;
    float c = s1 * s2;
    float pd = 1.0f / tri_area(p0, p1, p2);
    om_v3 n = v3_normalize(v3_cross(v3_sub(p1, p0), v3_sub(p2, p0)));
    om_v3 lp = v3_add(v3_add(v3_scale(p0, a), v3_scale(p1, b)), v3_scale(p2, c));
    om_v3 dv = v3_sub(lp, p);
    float d2 = v3_dot(dv, dv);
    float dist = sqrtf(d2);
    dv = v3_scale(dv, 1.0f / dist);
    float ndl = fabsf(v3_dot(n, v3_neg(dv)));
    pd *= d2 / ndl;
    pd *= 1.0f / (float)sc->n_lights;
    *dir = dv;
    *dist2 = d2;
    *pdf = pd;
    return prim;
}

static uint32_t sample_light(const om_scene* sc, om_v3 p, om_rng* rng, om_v3* dir, float* dist2, float* pdf,
                             uint32_t* inst)
{
    float r1 = om_uniform(rng);
    float xi0 = om_uniform(rng);
    float xi1 = om_uniform(rng);
    return sample_light_xi(sc, p, r1, xi0, xi1, dir, dist2, pdf, inst);
}

uint32_t om_sample_light(const om_scene* sc, const float p[3], const float xi[3], float dir[3], float* dist2,
                         float* pdf)
{
    om_v3 d;
    uint32_t inst;
    uint32_t prim = sample_light_xi(sc, v3(p[0], p[1], p[2]), xi[0], xi[1], xi[2], &d, dist2, pdf, &inst);
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
    return prim;
}

float om_light_pdf_hit(const float tri[9], float t, const float d[3])
{
    return light_pdf_hit(tri, t, v3(d[0], d[1], d[2]));
}

float om_power_heuristic(float pdf0, float pdf1) { return power_heuristic(pdf0, pdf1); }

/* Guide state of one frame of the guided tracer */
typedef struct {
    const om_svo* svo;       /* guide to sample, NULL = plain path tracing */
    float ratio;             /* probability of sampling the guide */
    uint32_t min_samples;    /* guide lookup threshold */
    int learn;               /* record path vertices */
    float lox, loy, loz, inv_cell;
    uint32_t depth;
} om_guide_frame;

/* path-vertex records of one thread, appended as paths finish */
typedef struct {
    uint32_t* code;
    uint32_t* bin;
    float* rad;
    size_t n, cap;
} om_rec_list;

static void rec_push(om_rec_list* r, uint32_t code, uint32_t bin, float rad)
{
    if (r->n == r->cap) {
        r->cap = r->cap ? r->cap * 2 : 4096;
        r->code = (uint32_t*)realloc(r->code, r->cap * 4);
        r->bin = (uint32_t*)realloc(r->bin, r->cap * 4);
        r->rad = (float*)realloc(r->rad, r->cap * 4);
    }
    r->code[r->n] = code;
    r->bin[r->n] = bin;
    r->rad[r->n] = rad;
    r->n++;
}

typedef struct {
    const om_scene* s;
    const om_camera* cam;
    const om_render_params* p;
    double* img;
    uint32_t tid;
    uint64_t segments;
    uint32_t sample_base;
    const om_guide_frame* g;
    om_rec_list rec;
    float* values;     /* per-sample radiance (om_render_samples) or NULL */
    float* coords;
} om_job;

/* One path, depth first.  With g->svo set, non-specular Lambert / metal
 * vertices draw from the one-sample mixture of guide and BxDF
 * (WFPGTracerKC.cuh:1480-1553: pdf = r p_guide + (1 - r) p_bxdf); with
 * g->learn, every continuing vertex leaves a record (cell, direction bin,
 * throughput) whose radiance the emitter hit fills in (PathNode.cuh:152-190). */
/* pixel (FP64) and, when set, the sample's own FP32 sum take contribution c */
static void contribute(double* pixel, float* sacc, om_v3 c)
{
    if (pixel) {
        pixel[0] += (double)c.x;
        pixel[1] += (double)c.y;
        pixel[2] += (double)c.z;
    }
    if (sacc) {
        sacc[0] += c.x;
        sacc[1] += c.y;
        sacc[2] += c.z;
    }
}

static void trace_path(const om_scene* sc, const om_camera* cam, const om_render_params* p,
                       uint32_t sample_base, uint32_t sample_id, const om_guide_frame* g,
                       double* pixel, uint64_t* segments, om_rec_list* recs, float* sacc, float* coord)
{
    om_rng rng;
    om_path_rng(&rng, p->seed, sample_base + sample_id);
    om_v3 o, d;
    camera_ray(cam, p->width, p->height, p->samples, sample_id, &rng, &o, &d, coord);
    om_v3 T = v3(1.0f, 1.0f, 1.0f);
    uint32_t depth = 1, medium = 0;
    float tmin = cam->near_far[0], tmax = cam->near_far[1];
    const int learn = g && g->learn;
    const int nee = p->nee != 0, mis = p->nee && p->mis;
    int spec_ray = 0;          /* the ray left a specular vertex (SPECULAR_PATH_RAY) */
    float prev_pdf = 0.0f;     /* BxDF pdf of the ray's direction (prevPDF) */
    uint32_t rc[64], rb[64];
    om_v3 rt[64];
    float rr[64];
    uint32_t nrec = 0;

    for (uint32_t bounce = 0; bounce < p->max_depth; bounce++) {
        (*segments)++;
        float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
        float b[2];
        uint32_t inst = 0;
        uint32_t prim = sc->two_level
                            ? om_trace2l_one(sc->two_level, sc->tris, sc->tree_width, oo, dd, tmin, &tmax, b, 0, &inst,
                                             NULL, NULL)
                        : sc->nodes4
                            ? om_trace4_one(sc->tris, sc->n, sc->sorted_idx, sc->nodes4, sc->root4, oo, dd, tmin,
                                            &tmax, b, 0, NULL, NULL)
                            : om_trace_one(sc->tris, sc->n, sc->sorted_idx, sc->left, sc->right,
                                           sc->node_aabb, sc->root, oo, dd, tmin, &tmax, b, 0, NULL, NULL);
        if (prim == OM_INVALID) break;
        float tri[9];
        prim_corners(sc, prim, inst, tri);
        uint32_t key = prim_key_of(sc, prim, inst);
        uint32_t batch = key >> 24, id = key & 0xFFFFFFu;
        if (batch == OM_BATCH_EMISSIVE) {
            /* PathTracerKC.cuh:337-426: Le * throughput.  With NEE, a path
             * ray that left a non-specular vertex counts only with MIS, and
             * then with the BxDF strategy's power-heuristic weight; camera
             * and specular rays count in full. */
            const float* le = sc->emissive + (size_t)id * 3;
            om_v3 c = v3(le[0] * T.x, le[1] * T.y, le[2] * T.z);
            int counts = 1;
            if (nee && depth > 1 && !spec_ray) {
                if (!mis) {
                    counts = 0;
                } else {
                    float pc = light_pdf_hit(tri, tmax, v3_normalize(d));
                    float pm = 1.0f / (float)sc->n_lights;
                    c = v3_scale(c, power_heuristic(prev_pdf, pc * pm));
                }
            }
            if (counts) contribute(pixel, sacc, c);
            /* the guide learns the radiance that arrived, whatever share of
             * it the image counted; next-event rays do not propagate */
            if (learn) {
                om_v3 c = v3_mul(v3(le[0], le[1], le[2]), T);
                for (uint32_t k = 0; k < nrec; k++) {
                    om_v3 f = rt[k];
                    om_v3 L = v3(f.x == 0.0f ? 0.0f : c.x / f.x, f.y == 0.0f ? 0.0f : c.y / f.y,
                                 f.z == 0.0f ? 0.0f : c.z / f.z);
                    rr[k] = om_luminance(L);
                }
            }
            break;
        }
        if (batch == OM_BATCH_NULL) break;
        if (bounce + 1 == p->max_depth) break;   /* the last segment only collects emission */
        om_surface s = sc->tbn ? make_surface_tbn(tri, b[0], b[1], d, sc->tbn + (size_t)prim * 12)
                               : make_surface(tri, b[0], b[1], d);
        om_v3 wi = v3_neg(v3_normalize(d));
        if (nee && sc->n_lights &&
            (batch == OM_BATCH_LAMBERT || (batch == OM_BATCH_METAL && (1.0f - sc->metal[(size_t)id * 10]) < 0.99f))) {
            /* PathTracerKC.cuh:493-575: one light sample per non-specular
             * vertex, traced as a closest-hit ray; it counts only where it
             * hits the light it was sent to (isCorrectNEERay) */
            om_v3 ldir;
            float dist2, pl, pb;
            uint32_t linst;
            uint32_t lprim = sample_light(sc, s.pos, &rng, &ldir, &dist2, &pl, &linst);
            om_v3 f;
            if (batch == OM_BATCH_LAMBERT) {
                const float* a = sc->lambert + (size_t)id * 3;
                f = lambert_eval(&s, v3(a[0], a[1], a[2]), ldir, &pb);
            } else {
                f = metal_eval(&s, wi, sc->metal + (size_t)id * 10, ldir, &pb);
            }
            float pn = pl;
            if (mis) {
                pn /= power_heuristic(pl, pb);
                if (isnan(pn)) pn = 0.0f;
            }
            om_v3 nf = v3_mul(T, f);
            nf = (pn == 0.0f) ? v3(0.0f, 0.0f, 0.0f) : v3_divs(nf, pn);
            if (dist2 > 0.0f && !v3_is_zero(nf)) {
                (*segments)++;
                float lt = FLT_MAX, lb[2];
                uint32_t hinst;
                uint32_t hp = closest_hit(sc, nudge(s.pos, s.geo_n), ldir, 0.0f, &lt, lb, &hinst);
                if (hp == lprim && hinst == linst) {
                    const float* le = sc->emissive + (size_t)(prim_key_of(sc, hp, hinst) & 0xFFFFFFu) * 3;
                    contribute(pixel, sacc, v3(le[0] * nf.x, le[1] * nf.y, le[2] * nf.z));
                }
            }
        }
        om_bsdf_out out;
        out.medium = medium;
        int guided = 0;
        uint32_t gnode = 0;
        float gtotal = 0.0f;
        om_v3 ns = v3(0.0f, 0.0f, 1.0f);
        if (g && g->svo && (batch == OM_BATCH_LAMBERT ||
                            (batch == OM_BATCH_METAL && (1.0f - sc->metal[(size_t)id * 10]) < 0.99f))) {
            const om_svo* sv = g->svo;
            ns = q_rotate(q_conj(s.w2t), v3(0.0f, 0.0f, 1.0f));
            float inv_cell = (float)(1u << sv->depth) / sv->cube[3];
            gnode = om_svo_lookup_guide(sv, om_svo_code_v(s.pos, sv->cube[0], sv->cube[1], sv->cube[2],
                                                          inv_cell, sv->depth), g->min_samples);
            gtotal = om_guide_total_v(sv, gnode, ns);
            guided = gtotal > 0.0f;
        }
        if (guided) {
            float xi = om_uniform(&rng);
            om_v3 dir, fcos;
            float pg = 0.0f, pb = 0.0f;
            if (xi < g->ratio) {
                float u0 = om_uniform(&rng), u1 = om_uniform(&rng), u2 = om_uniform(&rng);
                dir = om_guide_sample_v(g->svo, gnode, ns, gtotal, u0, u1, u2, &pg);
                if (batch == OM_BATCH_LAMBERT) {
                    const float* a = sc->lambert + (size_t)id * 3;
                    fcos = lambert_eval(&s, v3(a[0], a[1], a[2]), dir, &pb);
                } else {
                    fcos = metal_eval(&s, wi, sc->metal + (size_t)id * 10, dir, &pb);
                }
            } else {
                if (batch == OM_BATCH_LAMBERT) {
                    const float* a = sc->lambert + (size_t)id * 3;
                    lambert_sample(&s, v3(a[0], a[1], a[2]), &rng, &out);
                } else {
                    metal_sample(&s, wi, sc->metal + (size_t)id * 10, &rng, &out);
                }
                dir = out.d;
                fcos = out.refl;
                pb = out.pdf;
                pg = om_guide_pdf_v(g->svo, gnode, ns, gtotal, dir);
            }
            out.d = dir;
            out.o = nudge(s.pos, s.geo_n);
            out.refl = fcos;
            out.pdf = g->ratio * pg + (1.0f - g->ratio) * pb;
            out.specular = 0;
            out.medium = medium;
        } else if (batch == OM_BATCH_LAMBERT) {
            const float* a = sc->lambert + (size_t)id * 3;
            lambert_sample(&s, v3(a[0], a[1], a[2]), &rng, &out);
        } else if (batch == OM_BATCH_METAL) {
            metal_sample(&s, wi, sc->metal + (size_t)id * 10, &rng, &out);
        } else if (batch == OM_BATCH_REFLECT) {
            reflect_sample(&s, wi, sc->reflect + (size_t)id * 4, &out);
        } else {
            dielectric_sample(&s, wi, sc->dielectric + (size_t)id * 4, id, medium, sc->dielectric,
                              &rng, &out);
        }
        /* PathTracerKC.cuh:600-647 */
        om_v3 nT = v3_mul(T, out.refl);
        nT = (out.pdf == 0.0f) ? v3(0.0f, 0.0f, 0.0f) : v3_divs(nT, out.pdf);
        float avg = v3_dot(nT, v3(0.333f, 0.333f, 0.333f));
        int terminate = 0;
        if (!out.specular && depth > p->rr_start) {
            /* TracerFunctions.cuh:321-330 */
            float pr = avg < 0.005f ? 0.005f : (avg > 1.0f ? 1.0f : avg);
            if (om_uniform(&rng) >= pr) terminate = 1;
            else nT = v3_scale(nT, 1.0f / pr);
        }
        if ((terminate && !out.specular) || v3_is_zero(nT)) break;
        if (learn) {
            rc[nrec] = om_svo_code_v(s.pos, g->lox, g->loy, g->loz, g->inv_cell, g->depth);
            rb[nrec] = om_dir_to_bin_v(v3_normalize(out.d));
            rt[nrec] = nT;
            rr[nrec] = 0.0f;
            nrec++;
        }
        T = nT;
        o = out.o;
        d = out.d;
        medium = out.medium;
        spec_ray = out.specular;
        prev_pdf = out.pdf;
        depth++;
        tmin = 0.0f;
        tmax = FLT_MAX;
    }
    for (uint32_t k = 0; k < nrec; k++) rec_push(recs, rc[k], rb[k], rr[k]);
}

static void* render_worker(void* arg)
{
    om_job* job = (om_job*)arg;
    const om_render_params* p = job->p;
    uint32_t spp = p->samples * p->samples;
    uint32_t step = p->row_step ? p->row_step : 1u;
    uint32_t nthreads = p->threads ? p->threads : 1u;
    uint32_t row_index = 0;
    for (uint32_t y = p->row_offset; y < p->height; y += step, row_index++) {
        if (row_index % nthreads != job->tid) continue;
        for (uint32_t x = 0; x < p->width; x++) {
            uint32_t pix = y * p->width + x;
            double* px = job->img ? job->img + (size_t)pix * 3 : NULL;
            for (uint32_t s = 0; s < spp; s++) {
                size_t i = (size_t)pix * spp + s;
                trace_path(job->s, job->cam, p, job->sample_base, (uint32_t)i, job->g, px, &job->segments,
                           &job->rec, job->values ? job->values + 3 * i : NULL,
                           job->coords ? job->coords + 2 * i : NULL);
            }
        }
    }
    return NULL;
}

/* one frame over the selected rows; merges the threads' records into *recs */
static uint64_t render_frame_to(const om_scene* s, const om_camera* cam, const om_render_params* p,
                                uint32_t sample_base, const om_guide_frame* g, double* img, om_rec_list* recs,
                                float* values, float* coords)
{
    uint32_t nthreads = p->threads ? p->threads : 1u;
    om_job* jobs = (om_job*)calloc(nthreads, sizeof(om_job));
    pthread_t* th = (pthread_t*)calloc(nthreads, sizeof(pthread_t));
    for (uint32_t t = 0; t < nthreads; t++) {
        jobs[t].s = s; jobs[t].cam = cam; jobs[t].p = p; jobs[t].img = img; jobs[t].tid = t;
        jobs[t].sample_base = sample_base;
        jobs[t].g = g;
        jobs[t].values = values;
        jobs[t].coords = coords;
        if (nthreads > 1) pthread_create(&th[t], NULL, render_worker, &jobs[t]);
    }
    if (nthreads == 1) render_worker(&jobs[0]);
    uint64_t seg = 0;
    for (uint32_t t = 0; t < nthreads; t++) {
        if (nthreads > 1) pthread_join(th[t], NULL);
        seg += jobs[t].segments;
        om_rec_list* r = &jobs[t].rec;
        for (size_t i = 0; i < r->n; i++)
            if (recs) rec_push(recs, r->code[i], r->bin[i], r->rad[i]);
        free(r->code);
        free(r->bin);
        free(r->rad);
    }
    free(jobs);
    free(th);
    return seg;
}

static uint64_t render_frame(const om_scene* s, const om_camera* cam, const om_render_params* p,
                             uint32_t sample_base, const om_guide_frame* g, double* img, om_rec_list* recs)
{
    return render_frame_to(s, cam, p, sample_base, g, img, recs, NULL, NULL);
}

uint64_t om_render_samples(const om_scene* s, const om_camera* cam, const om_render_params* p, uint32_t sample_base,
                           float* values, float* coords)
{
    om_render_params q = *p;
    q.row_step = 1;
    q.row_offset = 0;
    size_t n = (size_t)p->width * p->height * p->samples * p->samples;
    memset(values, 0, n * 3 * sizeof(float));
    return render_frame_to(s, cam, &q, sample_base, NULL, NULL, NULL, values, coords);
}

uint64_t om_render(const om_scene* s, const om_camera* cam, const om_render_params* p, double* img)
{
    return render_frame(s, cam, p, 0u, NULL, img, NULL);
}

/* WFPGTracer::Render / Finalize (WFPGTracer.cu:1590-1790) restated over the
 * structure of the HIP path: `iterations` frames of samples^2 spp, frame i
 * keyed by sample ids i*W*H*spp + local id; every frame but the last records
 * its vertices; after it the octree is rebuilt over those cells (cube = scene
 * bounds padded by one voxel), splatted and filtered, and guides the next
 * frame.  Renders every row (the guide depends on all paths).  img receives
 * the sum of all frames' contributions. */
uint64_t om_render_guided(const om_scene* s, const om_camera* cam, const om_render_params* p,
                          const om_guide_params* gp, double* img, om_svo* last)
{
    om_render_params q = *p;
    q.row_step = 1;
    q.row_offset = 0;
    /* scene bounds: min / max over all triangle vertices; in a two-level
     * scene over the instances' world boxes (the top tree's extent) */
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (s->two_level) {
        const om_two_level* tl = s->two_level;
        for (uint32_t k = 0; k < tl->n_inst; k++) {
            uint32_t m = tl->inst_mesh[k];
            float box[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY}, wb[6];
            for (uint32_t i = tl->mesh_first[m]; i < tl->mesh_first[m + 1]; i++)
                for (int v = 0; v < 3; v++)
                    for (int a = 0; a < 3; a++) {
                        float x = s->tris[(size_t)i * 9 + 3 * v + a];
                        box[a] = fminf(box[a], x);
                        box[3 + a] = fmaxf(box[3 + a], x);
                    }
            om_world_box(tl->l2w + (size_t)k * 12, box, wb);
            for (int a = 0; a < 3; a++) {
                lo[a] = fminf(lo[a], wb[a]);
                hi[a] = fmaxf(hi[a], wb[3 + a]);
            }
        }
    } else {
        for (uint32_t i = 0; i < s->n; i++)
            for (int v = 0; v < 3; v++)
                for (int a = 0; a < 3; a++) {
                    float x = s->tris[(size_t)i * 9 + 3 * v + a];
                    lo[a] = fminf(lo[a], x);
                    hi[a] = fmaxf(hi[a], x);
                }
    }
    float span = 0.0f;
    for (int a = 0; a < 3; a++) span = fmaxf(span, hi[a] - lo[a]);
    if (!(span > 0.0f)) span = 1.0f;
    float cells = (float)(1u << gp->svo_depth);
    float pad = span / (cells - 2.0f > 1.0f ? cells - 2.0f : 1.0f);
    float cube[4] = {lo[0] - pad, lo[1] - pad, lo[2] - pad, span + 2.0f * pad};
    uint32_t total = p->width * p->height * p->samples * p->samples;

    om_svo svo;
    memset(&svo, 0, sizeof(svo));
    int have = 0;
    uint64_t seg = 0;
    for (uint32_t it = 0; it < gp->iterations; it++) {
        om_guide_frame g;
        memset(&g, 0, sizeof(g));
        g.svo = have ? &svo : NULL;
        g.ratio = gp->guide_ratio;
        g.min_samples = gp->min_samples;
        g.learn = gp->iterations > 1 && it + 1 < gp->iterations;
        g.lox = cube[0];
        g.loy = cube[1];
        g.loz = cube[2];
        g.inv_cell = cells / cube[3];
        g.depth = gp->svo_depth;
        om_rec_list recs;
        memset(&recs, 0, sizeof(recs));
        seg += render_frame(s, cam, &q, it * total, &g, img, &recs);
        if (g.learn && recs.n > 0) {
            om_svo next;
            om_svo_build(&next, recs.code, (uint32_t)recs.n, cube, gp->svo_depth);
            om_svo_splat(&next, recs.code, recs.bin, recs.rad, (uint32_t)recs.n);
            om_svo_filter(&next);
            om_svo_free(&svo);
            svo = next;
            have = 1;
        }
        free(recs.code);
        free(recs.bin);
        free(recs.rad);
    }
    if (last) *last = svo;
    else om_svo_free(&svo);
    return seg;
}

/* ---------------------------- exported helpers ----------------------- */

/* RayLib/Ray.hpp:304-309: 2 (wi.n) n - wi */
void om_reflect(const float wi[3], const float n[3], float out[3])
{
    om_v3 w = v3(wi[0], wi[1], wi[2]), nn = v3(n[0], n[1], n[2]);
    om_v3 r = v3_sub(v3_scale(nn, 2.0f * v3_dot(w, nn)), w);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* RayLib/Ray.hpp:322-355 (wi points away from the surface, n on its side) */
int om_refract(const float wi[3], const float n[3], float from, float to, float out[3])
{
    om_v3 w = v3(wi[0], wi[1], wi[2]), nn = v3(n[0], n[1], n[2]);
    float ratio = from / to;
    float cos_in = v3_dot(nn, w);
    float sin_in2 = fmaxf(0.0f, 1.0f - cos_in * cos_in);
    float sin_out2 = ratio * ratio * sin_in2;
    float cos_out = sqrtf(fmaxf(0.0f, 1.0f - sin_out2));
    if (sin_out2 >= 1.0f) return 0;
    om_v3 r = v3_add(v3_scale(v3_neg(w), ratio), v3_scale(nn, ratio * cos_in - cos_out));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
    return 1;
}

/* RayLib/Triangle.h:203-209 */
void om_triangle_normal(const float tri[9], float out[3])
{
    om_v3 p0 = v3(tri[0], tri[1], tri[2]), p1 = v3(tri[3], tri[4], tri[5]), p2 = v3(tri[6], tri[7], tri[8]);
    om_v3 n = v3_normalize(v3_cross(v3_sub(p1, p0), v3_sub(p2, p0)));
    out[0] = n.x; out[1] = n.y; out[2] = n.z;
}

/* RayLib/AABB.hpp:79-82 */
void om_aabb_centroid(const float box[6], float out[3])
{
    for (int a = 0; a < 3; a++) out[a] = box[a] + (box[3 + a] - box[a]) * 0.5f;
}
