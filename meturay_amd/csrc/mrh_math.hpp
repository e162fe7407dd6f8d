// mrh_math.hpp — scalar math shared by the HIP kernels.
//
// Everything is binary32 in a fixed operation order; the library is built
// with -ffp-contract=off so no FMA is formed and results do not depend on
// instruction selection.  Division and sqrt are the correctly rounded
// forms (hipcc's default for f32).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MRH_HD __host__ __device__ __forceinline__

namespace mrh {

struct V3 { float x, y, z; };
struct Q4 { float w, x, y, z; };   // scalar first

MRH_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
MRH_HD V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
MRH_HD V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
MRH_HD V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
MRH_HD V3 mul(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
MRH_HD V3 scale(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
MRH_HD V3 divs(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
MRH_HD bool is_zero(V3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }

// Vector::Dot order (RayLib/Vector.hpp:428-437): ((0 + x) + y) + z
MRH_HD float dot(V3 a, V3 b)
{
    float d = 0.0f;
    d += a.x * b.x;
    d += a.y * b.y;
    d += a.z * b.z;
    return d;
}

// RayLib/Vector.hpp:867-873
MRH_HD V3 cross(V3 a, V3 b)
{
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// RayLib/Vector.hpp:519-529
MRH_HD V3 normalize(V3 a) { return scale(a, 1.0f / sqrtf(dot(a, a))); }

MRH_HD Q4 qmake(float w, float x, float y, float z) { return Q4{w, x, y, z}; }
MRH_HD Q4 conj(Q4 q) { return qmake(q.w, -q.x, -q.y, -q.z); }

// RayLib/Quaternion.hpp:76-87
MRH_HD Q4 qmul(Q4 a, Q4 b)
{
    return qmake(a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z,
                 a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
                 a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
                 a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w);
}

// RayLib/Quaternion.hpp:229-238
MRH_HD V3 rotate(Q4 q, V3 v)
{
    Q4 r = qmul(q, qmul(qmake(0.0f, v.x, v.y, v.z), conj(q)));
    return v3(r.x, r.y, r.z);
}

// RayLib/Quaternion.hpp:333-361 (rotation taking +Z to unit vector b)
MRH_HD Q4 rot_between_z(V3 b)
{
    V3 zc = v3(-b.y, b.x, 0.0f);
    float zd = b.z;
    float s = sqrtf((1.0f - zd) * 0.5f);
    float c = sqrtf((zd + 1.0f) * 0.5f);
    zc = normalize(zc);
    if (fabsf(zd + 1.0f) < 1.0e-4f) return qmake(0.0f, 0.0f, 1.0f, 0.0f);
    if (fabsf(zd - 1.0f) < 1.0e-4f) return qmake(1.0f, 0.0f, 0.0f, 0.0f);
    return qmake(c, zc.x * s, zc.y * s, zc.z * s);
}

// ---------------------------------------------------------------- PCG32
// O'Neill 2014 (pcg32_random_r / pcg32_srandom_r).  A path keeps only the
// 64-bit state; the stream increment is rebuilt from its sample id.
MRH_HD uint32_t pcg32_next(uint64_t& state, uint64_t inc)
{
    uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

MRH_HD uint64_t pcg32_inc(uint32_t sample_id) { return ((uint64_t)sample_id << 1u) | 1u; }

MRH_HD uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Initial state of the stream keyed by (seed, global sample id)
MRH_HD uint64_t path_rng_init(uint32_t seed, uint32_t sample_id)
{
    uint64_t inc = pcg32_inc(sample_id);
    uint64_t st = 0u;
    pcg32_next(st, inc);
    st += mix64(((uint64_t)seed << 32) ^ (uint64_t)sample_id);
    pcg32_next(st, inc);
    return st;
}

struct Rng {
    uint64_t state, inc;
    MRH_HD float uniform() { return (float)(pcg32_next(state, inc) >> 8) * 0x1.0p-24f; }
};

// ------------------------------------------------------------ sin/cos(2 pi u)
// Quadrant reduction z = 4u, j = rint(z), r = z - j (both exact), then
// Taylor polynomials of sin/cos(pi/2 * r) through r^9 / r^8 (truncation
// < 2e-9 for |r| <= 1/2), Horner form without contraction.
MRH_HD void sincos_2pi(float u, float& s, float& c)
{
    const float S1 = 1.57079632679489662f, S3 = -0.64596409750624625f,
                S5 = 0.07969262624616704f, S7 = -0.00468175413531868f,
                S9 = 0.00016044118478736f;
    const float C2 = -1.23370055013616983f, C4 = 0.25366950790104802f,
                C6 = -0.02086348076335296f, C8 = 0.00091926027483943f;
    float z = 4.0f * u;
    float j = rintf(z);
    float r = z - j;
    int q = ((int)j) & 3;
    float r2 = r * r;
    float sp = r * (S1 + r2 * (S3 + r2 * (S5 + r2 * (S7 + r2 * S9))));
    float cp = 1.0f + r2 * (C2 + r2 * (C4 + r2 * (C6 + r2 * C8)));
    if (q == 0) { s = sp; c = cp; }
    else if (q == 1) { s = cp; c = -sp; }
    else if (q == 2) { s = -sp; c = -cp; }
    else { s = -cp; c = sp; }
}

// atan(t) for t in [0, 1]: two half-angle reductions,
// atan(t) = 2 atan(t / (1 + sqrt(1 + t^2))), bring t below tan(pi/16),
// where the Taylor series to t^13 is exact to about 2e-12.  Written out so
// that host and device round identically.
MRH_HD float atan_unit(float t)
{
    t = t / (1.0f + sqrtf(1.0f + t * t));
    t = t / (1.0f + sqrtf(1.0f + t * t));
    const float t2 = t * t;
    float p = 0.07692307692307693f;                  //  1/13
    p = p * t2 - 0.09090909090909091f;               // -1/11
    p = p * t2 + 0.1111111111111111f;                //  1/9
    p = p * t2 - 0.14285714285714285f;               // -1/7
    p = p * t2 + 0.2f;                               //  1/5
    p = p * t2 - 0.3333333333333333f;                // -1/3
    p = p * t2 + 1.0f;
    return 4.0f * (t * p);
}

// ------------------------------------------------------------ shading frames
// Per-vertex tangent frames of smooth-shaded triangles and their barycentric
// interpolation (GPUPrimitiveTriangle.cpp:286-316 at load,
// GPUPrimitiveTriangle.h:626-670 at shading).  acos and sin are evaluated
// with atan_unit and sincos_2pi, so host and device round identically.

// RayLib/Vector.hpp:877-886 (pbrt's CoordinateSystem)
MRH_HD V3 orthogonal_vector(V3 v)
{
    if (fabsf(v.x) > fabsf(v.y)) return divs(v3(-v.z, 0.0f, v.x), sqrtf(v.x * v.x + v.z * v.z));
    return divs(v3(0.0f, v.z, -v.y), sqrtf(v.y * v.y + v.z * v.z));
}

// Vector<4>::Dot order, component-wise scale and sum, Vector::Normalize
MRH_HD float qdot(Q4 a, Q4 b)
{
    float d = 0.0f;
    d += a.w * b.w;
    d += a.x * b.x;
    d += a.y * b.y;
    d += a.z * b.z;
    return d;
}
MRH_HD Q4 qscale(Q4 a, float s) { return qmake(a.w * s, a.x * s, a.y * s, a.z * s); }
MRH_HD Q4 qadd(Q4 a, Q4 b) { return qmake(a.w + b.w, a.x + b.x, a.y + b.y, a.z + b.z); }
MRH_HD Q4 qnormalize(Q4 q) { return qscale(q, 1.0f / sqrtf(qdot(q, q))); }

// TransformGen::Space (RayLib/Quaternion.hpp:370-436): the rotation taking
// the orthonormal basis (x, y, z) to the standard axes (Mike Day's
// matrix-to-quaternion conversion, normalised, conjugated)
MRH_HD Q4 space_quat(V3 x, V3 y, V3 z)
{
    const V3 c = cross(x, y) - z;
    if (fabsf(c.x) > 0.1f && fabsf(c.y) > 0.1f && fabsf(c.z) > 0.1f) x = -x;
    float t;
    Q4 q;
    if (z.z < 0.0f) {
        if (x.x > y.y) {
            t = 1.0f + x.x - y.y - z.z;
            q = qmake(y.z - z.y, t, x.y + y.x, z.x + x.z);
        } else {
            t = 1.0f - x.x + y.y - z.z;
            q = qmake(z.x - x.z, x.y + y.x, t, y.z + z.y);
        }
    } else {
        if (x.x < -y.y) {
            t = 1.0f - x.x - y.y + z.z;
            q = qmake(x.y - y.x, z.x + x.z, y.z + z.y, t);
        } else {
            t = 1.0f + x.x + y.y + z.z;
            q = qmake(t, y.z - z.y, z.x - x.z, x.y - y.x);
        }
    }
    q = qscale(q, 0.5f / sqrtf(t));
    return conj(qnormalize(q));
}

// World-to-tangent rotation of a vertex with unit shading normal n: the
// tangent is OrthogonalVector(n) (the reference's tangent for a vertex
// without texture coordinates), made orthonormal to n (Gram-Schmidt), the
// bitangent n x t (GPUPrimitiveTriangle.cpp:286-316)
MRH_HD Q4 vertex_frame(V3 n)
{
    V3 t = orthogonal_vector(n);
    if (dot(t, t) < 1.0e-5f) t = orthogonal_vector(n);
    t = normalize(t);
    const V3 tn = normalize(t - scale(n, dot(n, t)));
    const V3 b = cross(n, tn);
    return space_quat(tn, b, n);
}

// acos(x) for x in [0, 1] as 2 atan(sqrt((1 - x) / (1 + x))); sin(x) for
// x in [0, pi/2] through sincos_2pi
MRH_HD float acos_unit(float x) { return 2.0f * atan_unit(sqrtf((1.0f - x) / (1.0f + x))); }
MRH_HD float sin_q(float x)
{
    float s, c;
    sincos_2pi(x * 0.15915494309189534f, s, c);
    return s;
}

// Quat::SLerp (RayLib/Quaternion.hpp:256-281): the shorter arc, a plain
// lerp when the quaternions are within 1e-5 of parallel; not normalised
MRH_HD Q4 qslerp(Q4 a, Q4 b, float t)
{
    const float cos_t = qdot(a, b);
    const float cos_f = (cos_t >= 0.0f) ? cos_t : -cos_t;
    float s0, s1;
    if (cos_f < 1.0f - 1.0e-5f) {
        const float angle = acos_unit(cos_f);
        const float sa = sin_q(angle);
        s0 = sin_q(angle * (1.0f - t)) / sa;
        s1 = sin_q(angle * t) / sa;
    } else {
        s0 = 1.0f - t;
        s1 = t;
    }
    s1 = (cos_t >= 0.0f) ? s1 : -s1;
    return qadd(qscale(a, s0), qscale(b, s1));
}

// Quat::BarySLerp (RayLib/Quaternion.hpp:283-318): a, b = weights of q0, q1
MRH_HD Q4 bary_slerp(Q4 q0, Q4 q1, Q4 q2, float a, float b)
{
    const float c = 1.0f - a - b;
    if (fabsf(a + b) < 1.0e-5f) return q2;
    const float ab = a / (a + b);
    return qslerp(qslerp(q1, q0, ab), q2, c);
}

// ------------------------------------------------------------ Morton codes
// RayLib/MortonCode.h:45-70 (Expand3D, 32- and 64-bit)
MRH_HD uint32_t expand10(uint32_t x)
{
    x &= 0x000003ffu;
    x = (x ^ (x << 16)) & 0xff0000ffu;
    x = (x ^ (x << 8)) & 0x0300f00fu;
    x = (x ^ (x << 4)) & 0x030c30c3u;
    x = (x ^ (x << 2)) & 0x09249249u;
    return x;
}

MRH_HD uint64_t expand21(uint32_t v)
{
    uint64_t x = v & 0x1fffffu;
    x = (x | x << 32) & 0x001f00000000ffffull;
    x = (x | x << 16) & 0x001f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

// ------------------------------------------------------------ ray tests
// RayLib/Ray.hpp:143-185, Moller-Trumbore without culling.  On a hit
// returns t and the weights (1-u-v, u) of corners 0 and 1.
MRH_HD bool intersect_tri(V3 o, V3 d, V3 p0, V3 p1, V3 p2, float& t, float& ba, float& bb)
{
    V3 e0 = p1 - p0;
    V3 e1 = p2 - p0;
    V3 p = cross(d, e1);
    float det = dot(e0, p);
    if (fabsf(det) < 1.0e-7f) return false;
    float inv_det = 1.0f / det;
    V3 tv = o - p0;
    float u = dot(tv, p) * inv_det;
    if (u < 0.0f || u > 1.0f) return false;
    V3 q = cross(tv, e0);
    float v = dot(d, q) * inv_det;
    if (v < 0.0f || (v + u) > 1.0f) return false;
    float tt = dot(e1, q) * inv_det;
    if (tt <= 1.0e-7f) return false;
    t = tt;
    ba = 1.0f - u - v;
    bb = u;
    return true;
}

// RayLib/Ray.hpp:248-270 slab test with IEEE minNum/maxNum.
MRH_HD bool intersect_box(V3 o, V3 inv_d, float bx0, float by0, float bz0, float bx1,
                          float by1, float bz1, float tmin, float tmax, float& t_enter)
{
    float t0 = (bx0 - o.x) * inv_d.x, t1 = (bx1 - o.x) * inv_d.x;
    float te = fmaxf(tmin, fminf(t0, t1));
    float tx = fminf(tmax, fmaxf(t0, t1));
    t0 = (by0 - o.y) * inv_d.y; t1 = (by1 - o.y) * inv_d.y;
    te = fmaxf(te, fminf(t0, t1));
    tx = fminf(tx, fmaxf(t0, t1));
    t0 = (bz0 - o.z) * inv_d.z; t1 = (bz1 - o.z) * inv_d.z;
    te = fmaxf(te, fminf(t0, t1));
    tx = fminf(tx, fmaxf(t0, t1));
    t_enter = te;
    return tx >= te;
}

// RayLib/Ray.hpp:503-545 (Ray Tracing Gems I ch. 6), curvature offset 0
MRH_HD float nudge1(float p, float n)
{
    int32_t ofi = (int32_t)(256.0f * n);
    int32_t pi = __builtin_bit_cast(int32_t, p);
    pi += (p < 0.0f) ? -ofi : ofi;
    float pf = __builtin_bit_cast(float, pi);
    return (fabsf(p) < (1.0f / 32.0f)) ? (p + (1.0f / 65536.0f) * n) : pf;
}

MRH_HD V3 nudge(V3 p, V3 n) { return v3(nudge1(p.x, n.x), nudge1(p.y, n.y), nudge1(p.z, n.z)); }

}  // namespace mrh
