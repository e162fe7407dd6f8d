/*
 * om.h — inline scalar math for the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (meturay_amd/) may
 * include, link or call anything under oracle/.  The oracle is the checker
 * the GPU path is compared against; it is never the thing measured or
 * shipped.
 *
 * Float semantics: every routine here is written as plain IEEE-754 binary32
 * operations in a fixed order and the oracle is compiled with
 * -ffp-contract=off, so that the same sequence of operations executed by the
 * HIP kernels (also built without contraction) produces identical bits.
 * Constants carry the `f` suffix so no silent promotion to double happens.
 */
#ifndef OM_H
#define OM_H

#include <math.h>
#include <stdint.h>
#include <string.h>

typedef struct { float x, y, z; } om_v3;
typedef struct { float w, x, y, z; } om_q;   /* quaternion, scalar first */

static inline om_v3 v3(float x, float y, float z) { om_v3 r = {x, y, z}; return r; }
static inline om_v3 v3_add(om_v3 a, om_v3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline om_v3 v3_sub(om_v3 a, om_v3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline om_v3 v3_mul(om_v3 a, om_v3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline om_v3 v3_scale(om_v3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline om_v3 v3_neg(om_v3 a) { return v3(-a.x, -a.y, -a.z); }
static inline om_v3 v3_divs(om_v3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }

/* RayLib/Vector.hpp:428-437 — accumulation starts from 0 and adds x, y, z in
 * order: ((0 + x0*y0) + x1*y1) + x2*y2. */
static inline float v3_dot(om_v3 a, om_v3 b)
{
    float d = 0.0f;
    d += a.x * b.x;
    d += a.y * b.y;
    d += a.z * b.z;
    return d;
}

/* RayLib/Vector.hpp:867-873 */
static inline om_v3 v3_cross(om_v3 a, om_v3 b)
{
    return v3(a.y * b.z - a.z * b.y,
              a.z * b.x - a.x * b.z,
              a.x * b.y - a.y * b.x);
}

/* RayLib/Vector.hpp:504-529: inv = 1 / sqrt(len^2); v * inv */
static inline om_v3 v3_normalize(om_v3 a)
{
    float inv = 1.0f / sqrtf(v3_dot(a, a));
    return v3_scale(a, inv);
}

static inline int v3_is_zero(om_v3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }

/* ---------------- quaternions (RayLib/Quaternion.hpp) ---------------- */
static inline om_q q_make(float w, float x, float y, float z) { om_q q = {w, x, y, z}; return q; }
static inline om_q q_conj(om_q q) { return q_make(q.w, -q.x, -q.y, -q.z); }

/* Hamilton product, Quaternion.hpp:76-87 (term order as written there). */
static inline om_q q_mul(om_q a, om_q b)
{
    return q_make(a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z,
                  a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
                  a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
                  a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w);
}

/* Quaternion.hpp:229-238: q * (v * conj(q)) */
static inline om_v3 q_rotate(om_q q, om_v3 v)
{
    om_q vq = q_make(0.0f, v.x, v.y, v.z);
    om_q r = q_mul(q, q_mul(vq, q_conj(q)));
    return v3(r.x, r.y, r.z);
}

/* Quaternion.hpp:333-361: rotation taking +Z onto b (b unit length). */
static inline om_q q_rot_between_z(om_v3 b)
{
    om_v3 zc = v3(-b.y, b.x, 0.0f);
    float zd = b.z;
    float s = sqrtf((1.0f - zd) * 0.5f);
    float c = sqrtf((zd + 1.0f) * 0.5f);
    zc = v3_normalize(zc);
    if (fabsf(zd + 1.0f) < 1.0e-4f) return q_make(0.0f, 0.0f, 1.0f, 0.0f);
    if (fabsf(zd - 1.0f) < 1.0e-4f) return q_make(1.0f, 0.0f, 0.0f, 0.0f);
    return q_make(c, zc.x * s, zc.y * s, zc.z * s);
}

/* ---------------- PCG32 (O'Neill 2014, pcg32_srandom_r / pcg32_random_r) */
typedef struct { uint64_t state, inc; } om_rng;

static inline uint32_t om_pcg32(om_rng* r)
{
    uint64_t old = r->state;
    r->state = old * 6364136223846793005ULL + r->inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

static inline void om_pcg32_seed(om_rng* r, uint64_t initstate, uint64_t initseq)
{
    r->state = 0u;
    r->inc = (initseq << 1u) | 1u;
    om_pcg32(r);
    r->state += initstate;
    om_pcg32(r);
}

/* uniform float in [0,1): top 24 bits */
static inline float om_uniform(om_rng* r) { return (float)(om_pcg32(r) >> 8) * 0x1.0p-24f; }

/* SplitMix64 finaliser, used to derive a path's PCG32 initial state from
 * (seed, global sample id). */
static inline uint64_t om_mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* Path stream key: one PCG32 stream per (request seed, global sample id). */
static inline void om_path_rng(om_rng* r, uint32_t seed, uint32_t sample_id)
{
    om_pcg32_seed(r, om_mix64(((uint64_t)seed << 32) ^ (uint64_t)sample_id), (uint64_t)sample_id);
}

/* ---------------- sin/cos of 2*pi*u -----------------------------------
 * Octant reduction z = 4u (exact), j = rint(z), r = z - j in [-1/2, 1/2]
 * (exact), theta = r*pi/2; Taylor polynomials of sin(pi/2 r) through r^9
 * and cos(pi/2 r) through r^8 (truncation error < 2e-9 on |theta|<=pi/4),
 * evaluated by Horner without contraction, then rotated by the quadrant.
 */
#define OM_S1  1.57079632679489662f
#define OM_S3 -0.64596409750624625f
#define OM_S5  0.07969262624616704f
#define OM_S7 -0.00468175413531868f
#define OM_S9  0.00016044118478736f
#define OM_C2 -1.23370055013616983f
#define OM_C4  0.25366950790104802f
#define OM_C6 -0.02086348076335296f
#define OM_C8  0.00091926027483943f

static inline void om_sincos_2pi(float u, float* s, float* c)
{
    float z = 4.0f * u;
    float j = rintf(z);
    float r = z - j;
    int q = ((int)j) & 3;
    float r2 = r * r;
    float sp = r * (OM_S1 + r2 * (OM_S3 + r2 * (OM_S5 + r2 * (OM_S7 + r2 * OM_S9))));
    float cp = 1.0f + r2 * (OM_C2 + r2 * (OM_C4 + r2 * (OM_C6 + r2 * OM_C8)));
    switch (q) {
    case 0: *s = sp;  *c = cp;  break;
    case 1: *s = cp;  *c = -sp; break;
    case 2: *s = -sp; *c = -cp; break;
    default: *s = -cp; *c = sp; break;
    }
}

/* ---------------- guide direction map and voxel grid -------------------
 * Equal-area concentric octahedral map (Clarberg 2008; the reference's
 * RayLib/CoordinateConversion.h:112-180 uses the same map for its radiance
 * fields), 8x8 bins, and 30-bit Morton voxel codes of a cube grid. */
#define OM_GUIDE_RES  8u
#define OM_GUIDE_BINS 64u
#define OM_SVO_NO_NODE 0xFFFFFFFFu
#define OM_GUIDE_EPS 1.0e-4f
#define OM_GUIDE_FIX_ONE 65536.0f
#define OM_GUIDE_FIX_MAX 32767.0f

uint32_t om_morton_expand10(uint32_t v);

/* 48.16 fixed point of a splatted radiance sample, clamped to [0, 32767] */
static inline uint64_t om_guide_fix(float v)
{
    return (v > 0.0f) ? (uint64_t)(fminf(v, OM_GUIDE_FIX_MAX) * OM_GUIDE_FIX_ONE) : 0ull;
}

static inline om_v3 om_octa_to_dir_v(float u, float v)
{
    float a = 2.0f * u - 1.0f, b = 2.0f * v - 1.0f;
    float ua = fabsf(a), va = fabsf(b);
    float d = 1.0f - (ua + va);
    float r = 1.0f - fabsf(d);
    float phin = (r == 0.0f) ? 0.0f : ((va - ua) / r + 1.0f) * 0.125f;
    float z = copysignf(1.0f - r * r, d);
    float s, c;
    om_sincos_2pi(phin, &s, &c);
    float k = r * sqrtf(fmaxf(0.0f, 2.0f - r * r));
    return v3(copysignf(k * c, a), copysignf(k * s, b), z);
}

/* atan(t), t in [0,1]: two half-angle steps t <- t / (1 + sqrt(1 + t^2))
 * (each halves the angle) bring t below tan(pi/16); Taylor series to t^13. */
static inline float om_atan_unit_f(float t)
{
    t = t / (1.0f + sqrtf(1.0f + t * t));
    t = t / (1.0f + sqrtf(1.0f + t * t));
    float t2 = t * t;
    float p = 0.07692307692307693f;
    p = p * t2 - 0.09090909090909091f;
    p = p * t2 + 0.1111111111111111f;
    p = p * t2 - 0.14285714285714285f;
    p = p * t2 + 0.2f;
    p = p * t2 - 0.3333333333333333f;
    p = p * t2 + 1.0f;
    return 4.0f * (t * p);
}

/* atan2(y, x) / (pi/2) for x, y >= 0 */
static inline float om_quadrant_angle(float y, float x)
{
    if (y == 0.0f && x == 0.0f) return 0.0f;
    if (y <= x) return om_atan_unit_f(y / x) * 0.63661977236758134f;
    return 1.0f - om_atan_unit_f(x / y) * 0.63661977236758134f;
}

static inline void om_dir_to_octa_v(om_v3 w, float* u, float* v)
{
    float r = sqrtf(fmaxf(0.0f, 1.0f - fabsf(w.z)));
    float t = om_quadrant_angle(fabsf(w.y), fabsf(w.x));
    float b = r * t, a = r - b;
    if (w.z < 0.0f) {
        float a2 = 1.0f - b, b2 = 1.0f - a;
        a = a2;
        b = b2;
    }
    *u = (copysignf(a, w.x) + 1.0f) * 0.5f;
    *v = (copysignf(b, w.y) + 1.0f) * 0.5f;
}

static inline uint32_t om_dir_to_bin_v(om_v3 w)
{
    float u, v;
    om_dir_to_octa_v(w, &u, &v);
    uint32_t ix = (uint32_t)(u * (float)OM_GUIDE_RES), iy = (uint32_t)(v * (float)OM_GUIDE_RES);
    ix = ix > OM_GUIDE_RES - 1 ? OM_GUIDE_RES - 1 : ix;
    iy = iy > OM_GUIDE_RES - 1 ? OM_GUIDE_RES - 1 : iy;
    return iy * OM_GUIDE_RES + ix;
}

/* Morton code of the cell holding p at `depth` of the grid with origin lo and
 * inv_cell cells per unit; points outside clamp to the border cells. */
static inline uint32_t om_svo_code_v(om_v3 p, float lox, float loy, float loz, float inv_cell,
                                     uint32_t depth)
{
    uint32_t qmax = (1u << depth) - 1u;
    float f[3] = {(p.x - lox) * inv_cell, (p.y - loy) * inv_cell, (p.z - loz) * inv_cell};
    uint32_t q[3];
    for (int a = 0; a < 3; a++) {
        float g = floorf(f[a]);
        q[a] = (g <= 0.0f) ? 0u : (g >= (float)qmax ? qmax : (uint32_t)g);
    }
    return om_morton_expand10(q[0]) | (om_morton_expand10(q[1]) << 1) | (om_morton_expand10(q[2]) << 2);
}

/* RayLib/ColorConversion.h:15-22 (its coefficients, as written there) */
static inline float om_luminance(om_v3 c) { return (0.2120f * c.x + 0.7150f * c.y) + 0.0722f * c.z; }

static inline uint32_t om_f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float om_u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* single-ray traversal shared by the closest/any-hit entry points and the
 * path tracer (om_geom.c) */
uint32_t om_trace_one(const float* tris, uint32_t n, const uint32_t* sorted_idx,
                      const uint32_t* left, const uint32_t* right, const float* node_aabb,
                      uint32_t root, const float o[3], const float d[3], float tmin,
                      float* tmax_io, float bary_out[2], int any_hit,
                      uint32_t* visits, uint32_t* tests);

uint32_t om_trace4_one(const float* tris, uint32_t n, const uint32_t* sorted_idx, const float* nodes4,
                       uint32_t root4, const float o[3], const float d[3], float tmin, float* tmax_io,
                       float bary_out[2], int any_hit, uint32_t* visits, uint32_t* tests);

/* guide evaluation shared by the guided path tracer (om_guide.c) */
struct om_svo;
float    om_guide_total_v(const struct om_svo* s, uint32_t node, om_v3 n);
om_v3    om_guide_sample_v(const struct om_svo* s, uint32_t node, om_v3 n, float total, float xi0, float xi1,
                           float xi2, float* pdf);
float    om_guide_pdf_v(const struct om_svo* s, uint32_t node, om_v3 n, float total, om_v3 d);

#endif
