// ref_wrap.cpp — extern "C" entry points over the reference's OWN
// host-compilable RayLib headers, used only to write golden vectors
// (tests/golden/gen_golden.py) and to cross-check the oracle.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile (target `ref`)
// against the reference checkout; nothing from the reference is copied here.
// The headers compile unchanged with the host compiler (METU_CUDA undefined
// makes __host__/__device__ empty, RayLib/CudaCheck.h).
//
// The headers call an unqualified abs() on floats (Ray.hpp:157,
// Quaternion.hpp:348,354), which CUDA resolves to its float overload.  With
// only <cstdlib> in scope the host compiler picks int abs(int) and silently
// truncates.  The standard C++ <math.h> puts the float overloads of abs in the
// global namespace first, so the host build computes what the device code
// does.
#include <math.h>
#include "RayLib/MortonCode.h"
#include "RayLib/Ray.h"
#include "RayLib/AABB.h"
#include "RayLib/HemiDistribution.h"
#include "RayLib/Quaternion.h"
#include "RayLib/Triangle.h"
#include "RayLib/Matrix.h"
#include "RayLib/CoordinateConversion.h"

#include <cstdint>

extern "C" {

// RayLib/MortonCode.h:152-170
uint32_t ref_compose30(uint32_t x, uint32_t y, uint32_t z)
{ return MortonCode::Compose3D<uint32_t>(Vector3ui(x, y, z)); }
uint64_t ref_compose63(uint32_t x, uint32_t y, uint32_t z)
{ return MortonCode::Compose3D<uint64_t>(Vector3ui(x, y, z)); }
void ref_decompose30(uint32_t c, uint32_t out[3])
{ Vector3ui v = MortonCode::Decompose3D<uint32_t>(c); out[0] = v[0]; out[1] = v[1]; out[2] = v[2]; }
void ref_decompose63(uint64_t c, uint32_t out[3])
{ Vector3ui v = MortonCode::Decompose3D<uint64_t>(c); out[0] = v[0]; out[1] = v[1]; out[2] = v[2]; }

// RayLib/Ray.hpp:143-185
int ref_intersect_triangle(const float o[3], const float d[3], const float tri[9],
                           float* t, float bary[3])
{
    RayF r(Vector3f(d[0], d[1], d[2]), Vector3f(o[0], o[1], o[2]));
    Vector3f b(0.0f, 0.0f, 0.0f); float tt = 0.0f;   // misses leave b as given
    bool h = r.IntersectsTriangle(b, tt, Vector3f(tri[0], tri[1], tri[2]),
                                  Vector3f(tri[3], tri[4], tri[5]),
                                  Vector3f(tri[6], tri[7], tri[8]), false);
    *t = tt; bary[0] = b[0]; bary[1] = b[1]; bary[2] = b[2];
    return h ? 1 : 0;
}

// RayLib/Ray.hpp:248-270
int ref_intersect_aabb(const float o[3], const float d[3], const float box[6],
                       float tmin, float tmax, float tout[2])
{
    RayF r(Vector3f(d[0], d[1], d[2]), Vector3f(o[0], o[1], o[2]));
    Vector2f to;
    bool h = r.IntersectsAABB(to, Vector3f(box[0], box[1], box[2]),
                              Vector3f(box[3], box[4], box[5]), Vector2f(tmin, tmax));
    tout[0] = to[0]; tout[1] = to[1];
    return h ? 1 : 0;
}

// RayLib/Ray.hpp:503-545
void ref_nudge(const float p[3], const float d[3], const float n[3], float out[3])
{
    RayF r(Vector3f(d[0], d[1], d[2]), Vector3f(p[0], p[1], p[2]));
    RayF o = r.Nudge(Vector3f(n[0], n[1], n[2]), 0.0f);
    Vector3f q = o.getPosition();
    out[0] = q[0]; out[1] = q[1]; out[2] = q[2];
}

// RayLib/Ray.hpp:304-309
void ref_reflect(const float wi[3], const float n[3], float out[3])
{
    RayF r(Vector3f(wi[0], wi[1], wi[2]), Vector3f(0.0f));
    Vector3f v = r.Reflect(Vector3f(n[0], n[1], n[2])).getDirection();
    out[0] = v[0]; out[1] = v[1]; out[2] = v[2];
}

// Ray::Refract (RayLib/Ray.hpp:322-355) calls an unqualified max() that only
// CUDA's headers declare, so it is not host-compilable; it is pinned by
// Snell's-law closed forms in the tests instead.

// RayLib/HemiDistribution.h:8-51
void ref_hemi_uniform(float xi0, float xi1, float dir[3], float* pdf)
{
    Vector3f v = HemiDistribution::HemiUniformCDF(Vector2f(xi0, xi1), *pdf);
    dir[0] = v[0]; dir[1] = v[1]; dir[2] = v[2];
}
void ref_hemi_cosine(float xi0, float xi1, float dir[3], float* pdf)
{
    Vector3f v = HemiDistribution::HemiCosineCDF(Vector2f(xi0, xi1), *pdf);
    dir[0] = v[0]; dir[1] = v[1]; dir[2] = v[2];
}

// RayLib/Quaternion.hpp:229-238, 333-361
void ref_rot_between_z(const float n[3], float q[4])
{
    QuatF r = Quat::RotationBetweenZAxis(Vector3f(n[0], n[1], n[2]));
    q[0] = r[0]; q[1] = r[1]; q[2] = r[2]; q[3] = r[3];
}
void ref_quat_rotate(const float q[4], const float v[3], float out[3])
{
    QuatF r(q[0], q[1], q[2], q[3]);
    Vector3f o = r.ApplyRotation(Vector3f(v[0], v[1], v[2]));
    out[0] = o[0]; out[1] = o[1]; out[2] = o[2];
}
void ref_quat_mul(const float a[4], const float b[4], float out[4])
{
    QuatF r = QuatF(a[0], a[1], a[2], a[3]) * QuatF(b[0], b[1], b[2], b[3]);
    out[0] = r[0]; out[1] = r[1]; out[2] = r[2]; out[3] = r[3];
}

// RayLib/Triangle.h:203-209, RayLib/AABB.hpp:79-82
void ref_triangle_normal(const float tri[9], float out[3])
{
    Vector3f p[3] = {Vector3f(tri[0], tri[1], tri[2]), Vector3f(tri[3], tri[4], tri[5]),
                     Vector3f(tri[6], tri[7], tri[8])};
    Vector3f n = Triangle::Normal(p);
    out[0] = n[0]; out[1] = n[1]; out[2] = n[2];
}
void ref_triangle_aabb(const float tri[9], float out[6])
{
    AABB3f b = Triangle::BoundingBox(Vector3f(tri[0], tri[1], tri[2]), Vector3f(tri[3], tri[4], tri[5]),
                                     Vector3f(tri[6], tri[7], tri[8]));
    out[0] = b.Min()[0]; out[1] = b.Min()[1]; out[2] = b.Min()[2];
    out[3] = b.Max()[0]; out[4] = b.Max()[1]; out[5] = b.Max()[2];
}
void ref_aabb_centroid(const float box[6], float out[3])
{
    AABB3f b(Vector3f(box[0], box[1], box[2]), Vector3f(box[3], box[4], box[5]));
    Vector3f c = b.Centroid();
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2];
}

// RayLib/Ray.hpp:461-465 (Ray::Transform by a 4x4 matrix) with
// Matrix<4>::operator*(Vector<4>) (RayLib/Matrix.hpp:268-289); m12 is the
// row-major top 3x4 of an affine matrix, the last row 0 0 0 1.
void ref_ray_transform(const float m12[12], const float o[3], const float d[3], float out_o[3],
                       float out_d[3])
{
    Matrix4x4 m;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 4; c++) m(r, c) = m12[4 * r + c];
    m(3, 0) = 0.0f; m(3, 1) = 0.0f; m(3, 2) = 0.0f; m(3, 3) = 1.0f;
    RayF ray(Vector3f(d[0], d[1], d[2]), Vector3f(o[0], o[1], o[2]));
    RayF t = ray.Transform(m);
    out_o[0] = t.getPosition()[0]; out_o[1] = t.getPosition()[1]; out_o[2] = t.getPosition()[2];
    out_d[0] = t.getDirection()[0]; out_d[1] = t.getDirection()[1]; out_d[2] = t.getDirection()[2];
}

// RayLib/Vector.hpp:877-886
void ref_orthogonal_vector(const float v[3], float out[3])
{
    Vector3f r = OrthogonalVector(Vector3f(v[0], v[1], v[2]));
    out[0] = r[0]; out[1] = r[1]; out[2] = r[2];
}

// RayLib/Quaternion.hpp:370-436 (TransformGen::Space), q as (w, x, y, z)
void ref_space_quat(const float x[3], const float y[3], const float z[3], float q[4])
{
    QuatF r;
    TransformGen::Space(r, Vector3f(x[0], x[1], x[2]), Vector3f(y[0], y[1], y[2]), Vector3f(z[0], z[1], z[2]));
    q[0] = r[0]; q[1] = r[1]; q[2] = r[2]; q[3] = r[3];
}

// RayLib/Quaternion.hpp:256-318 (Quat::SLerp, Quat::BarySLerp)
void ref_slerp(const float a[4], const float b[4], float t, float out[4])
{
    QuatF r = Quat::SLerp(QuatF(a[0], a[1], a[2], a[3]), QuatF(b[0], b[1], b[2], b[3]), t);
    out[0] = r[0]; out[1] = r[1]; out[2] = r[2]; out[3] = r[3];
}
void ref_bary_slerp(const float q0[4], const float q1[4], const float q2[4], float a, float b, float out[4])
{
    QuatF r = Quat::BarySLerp(QuatF(q0[0], q0[1], q0[2], q0[3]), QuatF(q1[0], q1[1], q1[2], q1[3]),
                              QuatF(q2[0], q2[1], q2[2], q2[3]), a, b);
    out[0] = r[0]; out[1] = r[1]; out[2] = r[2]; out[3] = r[3];
}

// RayLib/CoordinateConversion.h:112-235 (concentric octahedral map)
void ref_octa_to_dir(float u, float v, float out[3])
{
    Vector3f d = Utility::CocentricOctohedralToDirection(Vector2f(u, v));
    out[0] = d[0]; out[1] = d[1]; out[2] = d[2];
}
void ref_dir_to_octa(const float d[3], float out[2])
{
    Vector2f st = Utility::DirectionToCocentricOctohedral(Vector3f(d[0], d[1], d[2]));
    out[0] = st[0]; out[1] = st[1];
}
void ref_octa_wrap_int(int32_t x, int32_t y, int32_t X, int32_t Y, int32_t out[2])
{
    Vector2i w = Utility::CocentricOctohedralWrapInt(Vector2i(x, y), Vector2i(X, Y));
    out[0] = w[0]; out[1] = w[1];
}

}  // extern "C"
