/*
 * om_vox.c — oracle for the reference's triangle voxeliser, the geometry
 * source of its path-guiding octree (TriFunctions::Voxelize,
 * TracerCUDA/GPUPrimitiveTriangle.h:384-620, driven by KCGetVoxelCount /
 * KCVoxelizePrims, GPUAcceleratorCommonKC.cuh:439-542, and
 * AnisoSVOctreeCPU::Constrcut, AnisoSVO.cu:657-760).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Per triangle: project onto the plane of the dominant axis of its normal
 * (the octree cube mapped to [-1,1]^3, then turned so that the dominant
 * axis faces +z), rasterise conservatively at the leaf resolution R (each
 * edge pushed out by half a pixel along its normal's L1 direction), and for
 * every covered pixel emit the leaf voxel that holds the triangle's plane
 * point under the pixel centre (barycentrics of the original triangle,
 * unclamped; voxel index clamped to [0, R-1]).  Records come out triangle by
 * triangle, pixels row by row.
 *
 * The reference turns the projected points with a float quaternion of angle
 * pi/2 or pi; here the turn is the exact axis permutation it approximates
 * (the pixel grid is symmetric under the sign flips, so only which two axes
 * span the raster matters).  The two agree except within float rounding of a
 * pixel edge.  Products are not fused.
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"
#include "om.h"

static uint32_t vox_expand10(uint32_t v)
{
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

typedef struct {
    float p[3][3];             /* world positions */
    float q[3][2];             /* projected 2D positions */
    float c0[2], ec0[2], ec1[2], denom_c;   /* conservative triangle */
    float e0[2], e1[2], denom;              /* original triangle */
    int32_t x0, x1, y0, y1;                 /* pixel rectangle [x0,x1) x [y0,y1) */
} om_vox_setup;

/* GPUPrimitiveTriangle.h:431-540 */
static void vox_setup(const float* tri, const float cube[4], uint32_t R, om_vox_setup* s)
{
    for (int k = 0; k < 3; k++)
        for (int a = 0; a < 3; a++) s->p[k][a] = tri[3 * k + a];
    /* Triangle::Normal: normalised cross of the edges (RayLib/Triangle.h:203-210) */
    const float ea[3] = {s->p[1][0] - s->p[0][0], s->p[1][1] - s->p[0][1], s->p[1][2] - s->p[0][2]};
    const float eb[3] = {s->p[2][0] - s->p[0][0], s->p[2][1] - s->p[0][1], s->p[2][2] - s->p[0][2]};
    float n[3] = {ea[1] * eb[2] - ea[2] * eb[1], ea[2] * eb[0] - ea[0] * eb[2], ea[0] * eb[1] - ea[1] * eb[0]};
    const float inv = 1.0f / sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    n[0] *= inv;
    n[1] *= inv;
    n[2] *= inv;
    /* dominant axis: the first largest |component| (Vector::Max) */
    int dom = 0;
    float best = fabsf(n[0]);
    if (fabsf(n[1]) > best) { dom = 1; best = fabsf(n[1]); }
    if (fabsf(n[2]) > best) dom = 2;
    const int neg = signbit(n[dom]) ? 1 : 0;
    /* cube -> [-1,1]^3 (TransformGen::Ortogonal, RayLib/Matrix.hpp:1021-1039) */
    for (int k = 0; k < 3; k++) {
        float ndc[3];
        for (int a = 0; a < 3; a++) {
            const float lo = cube[a], hi = cube[a] + cube[3];
            const float sc = 2.0f / (hi - lo), tr = -((hi + lo) / (hi - lo));
            ndc[a] = sc * s->p[k][a] + tr;
        }
        /* turn the dominant axis to +z; keep the raster axes */
        float u, v;
        if (dom == 0) { u = neg ? ndc[2] : -ndc[2]; v = ndc[1]; }
        else if (dom == 1) { u = ndc[0]; v = neg ? ndc[2] : -ndc[2]; }
        else { u = neg ? -ndc[0] : ndc[0]; v = ndc[1]; }
        s->q[k][0] = u;
        s->q[k][1] = v;
    }
    float mn[2] = {s->q[0][0], s->q[0][1]}, mx[2] = {s->q[0][0], s->q[0][1]};
    for (int k = 1; k < 3; k++)
        for (int a = 0; a < 2; a++) {
            mn[a] = fminf(mn[a], s->q[k][a]);
            mx[a] = fmaxf(mx[a], s->q[k][a]);
        }
    const float Rf = (float)R;
    int32_t r[4] = {(int32_t)floorf((0.5f + 0.5f * mn[0]) * Rf), (int32_t)ceilf((0.5f + 0.5f * mx[0]) * Rf),
                    (int32_t)floorf((0.5f + 0.5f * mn[1]) * Rf), (int32_t)ceilf((0.5f + 0.5f * mx[1]) * Rf)};
    for (int i = 0; i < 4; i++) r[i] = r[i] < 0 ? 0 : (r[i] > (int32_t)R ? (int32_t)R : r[i]);
    s->x0 = r[0];
    s->x1 = r[1];
    s->y0 = r[2];
    s->y1 = r[3];
    /* conservative edges (GPU Gems 2 ch. 42): line through q_k along the
     * edge, moved out by half a pixel */
    const float hp = 1.0f / Rf;
    float pl[3][3];
    for (int k = 0; k < 3; k++) {
        const float* a = s->q[k];
        const float* b = s->q[(k + 1) % 3];
        const float dx = b[0] - a[0], dy = b[1] - a[1];
        pl[k][0] = dy;
        pl[k][1] = -dx;
        pl[k][2] = dx * a[1] - dy * a[0];
        pl[k][2] -= hp * fabsf(pl[k][0]) + hp * fabsf(pl[k][1]);
    }
    /* corners: crossings of adjacent moved lines */
    const int pair[3][2] = {{0, 2}, {0, 1}, {1, 2}};
    float cc[3][2];
    for (int k = 0; k < 3; k++) {
        const float* A = pl[pair[k][0]];
        const float* B = pl[pair[k][1]];
        const float X = A[1] * B[2] - A[2] * B[1], Y = A[2] * B[0] - A[0] * B[2], Z = A[0] * B[1] - A[1] * B[0];
        cc[k][0] = X / Z;
        cc[k][1] = Y / Z;
    }
    s->c0[0] = cc[0][0];
    s->c0[1] = cc[0][1];
    s->ec0[0] = cc[1][0] - cc[0][0];
    s->ec0[1] = cc[1][1] - cc[0][1];
    s->ec1[0] = cc[2][0] - cc[0][0];
    s->ec1[1] = cc[2][1] - cc[0][1];
    s->denom_c = 1.0f / (s->ec0[0] * s->ec1[1] - s->ec1[0] * s->ec0[1]);
    s->e0[0] = s->q[1][0] - s->q[0][0];
    s->e0[1] = s->q[1][1] - s->q[0][1];
    s->e1[0] = s->q[2][0] - s->q[0][0];
    s->e1[1] = s->q[2][1] - s->q[0][1];
    s->denom = 1.0f / (s->e0[0] * s->e1[1] - s->e1[0] * s->e0[1]);
}

/* one pixel: 1 and the voxel code when the conservative triangle covers its
 * centre (GPUPrimitiveTriangle.h:548-612) */
static int vox_pixel(const om_vox_setup* s, const float cube[4], uint32_t R, int32_t x, int32_t y, uint32_t* code)
{
    const float dp = 2.0f / (float)R;
    const float px = ((float)x + 0.5f) * dp - 1.0f, py = ((float)y + 0.5f) * dp - 1.0f;
    const float g0 = px - s->c0[0], g1 = py - s->c0[1];
    const float b = (g0 * s->ec1[1] - s->ec1[0] * g1) * s->denom_c;
    const float c = (s->ec0[0] * g1 - g0 * s->ec0[1]) * s->denom_c;
    const float a = 1.0f - b - c;
    if (!(b >= 0.0f && b <= 1.0f && c >= 0.0f && c <= 1.0f && a >= 0.0f && a <= 1.0f)) return 0;
    const float h0 = px - s->q[0][0], h1 = py - s->q[0][1];
    const float B = (h0 * s->e1[1] - s->e1[0] * h1) * s->denom;
    const float C = (s->e0[0] * h1 - h0 * s->e0[1]) * s->denom;
    const float A = 1.0f - B - C;
    uint32_t q[3];
    for (int k = 0; k < 3; k++) {
        const float pos = s->p[0][k] * A + s->p[1][k] * B + s->p[2][k] * C;
        const float f = ((pos - cube[k]) / cube[3]) * (float)R;
        q[k] = !(f > 0.0f) ? 0u : (f >= (float)(R - 1) ? R - 1 : (uint32_t)f);
    }
    *code = vox_expand10(q[0]) | (vox_expand10(q[1]) << 1) | (vox_expand10(q[2]) << 2);
    return 1;
}

/* Voxel records of n triangles (float[n][9]) in the cube (lo xyz, size) at
 * depth (R = 2^depth <= 1024): returns the record count; codes and prims
 * (capacity cap, may be NULL) receive the first min(count, cap) records and
 * their triangle indices. */
uint64_t om_voxelize(const float* tris, uint32_t n, const float cube[4], uint32_t depth, uint32_t* codes,
                     uint32_t* prims, uint64_t cap)
{
    const uint32_t R = 1u << depth;
    uint64_t w = 0;
    for (uint32_t t = 0; t < n; t++) {
        om_vox_setup s;
        vox_setup(tris + 9 * (size_t)t, cube, R, &s);
        for (int32_t y = s.y0; y < s.y1; y++)
            for (int32_t x = s.x0; x < s.x1; x++) {
                uint32_t code;
                if (vox_pixel(&s, cube, R, x, y, &code)) {
                    if (codes && w < cap) codes[w] = code;
                    if (prims && w < cap) prims[w] = t;
                    w++;
                }
            }
    }
    return w;
}

/* ---- Cone trace through the guiding octree (AnisoSVOctreeGPU::ConeTraceRay,
 * TracerCUDA/AnisoSVO.cuh:767-1032), on the octree of om_svo_build (levels
 * root first; a node's children are first_child + rank in child_mask).
 *
 * The ray is taken to the cube's unit space and mirrored so that every
 * direction component is positive (the child index is mirrored back when a
 * child is looked up); a Morton bit stack of the child indices taken on the
 * way down and the parent links replace a node stack.  Descending into an
 * existing child stops at a leaf, at a level whose voxel the cone's disc
 * (diameter^2 = 4 t^2 aperture / pi at the current t) is wider than, or at
 * `max_query_level_offset` levels above the leaves; a missing child is
 * crossed without a node and the walk moves on past its far face.  Returns
 * the t reached; *node = the node stopped at (OM_SVO_NO_NODE: the ray left
 * the cube), *is_leaf whether it is a leaf.  A step limit (the reference has
 * none) ends a walk that does not finish within OM_CONE_MAX_STEPS steps as a
 * miss. */
#define OM_CONE_MAX_STEPS (1u << 16)

static float neg_exp2i(uint32_t p) { return p == 0 ? 1.0f : ldexpf(1.0f, -(int)p); }

float om_svo_cone_trace(const om_svo* s, const float o[3], const float dir[3], float tmin, float aperture,
                        uint32_t max_query_level_offset, uint32_t* node_out, int* is_leaf)
{
    const float EPS = 1.0e-5f;
    *node_out = OM_SVO_NO_NODE;
    *is_leaf = 0;
    float d[3], p[3], coeff[3], bias[3];
    const float scale = 1.0f / s->cube[3];
    uint32_t mirror = 0;
    for (int a = 0; a < 3; a++) {
        d[a] = (fabsf(dir[a]) < EPS) ? copysignf(EPS, dir[a]) : dir[a];
        d[a] = d[a] * scale;
        p[a] = (o[a] - s->cube[a]) * scale;
        if (d[a] < 0.0f) {
            p[a] = 1.0f - p[a];
            mirror |= 1u << a;
        }
        coeff[a] = 1.0f / fabsf(d[a]);
        bias[a] = p[a] * coeff[a];
    }
    const float te[3] = {-bias[0], -bias[1], -bias[2]};
    const float tsvo = fmaxf(fmaxf(te[0], te[1]), te[2]);
    if (!(te[0] < tmin && te[1] < tmin && te[2] < tmin)) tmin = nextafterf(tsvo, INFINITY);
    const float leaf = s->cube[3] / (float)(1u << s->depth);
    uint64_t stack = 0;
    uint32_t cnt = 0;
    uint32_t nid = 0;
    float corner[3] = {0.0f, 0.0f, 0.0f};
    int empty = 0;
    for (uint32_t step = 0; nid != OM_SVO_NO_NODE; step++) {
        if (step >= OM_CONE_MAX_STEPS) {
            *node_out = OM_SVO_NO_NODE;
            *is_leaf = 0;
            return tmin;
        }
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
                if (cnt == 0) break;                    /* above the root: left the cube */
                nid = s->parent[nid];
            }
            cnt--;
            empty = 0;
            continue;
        }
        uint32_t child = 0;
        for (int a = 0; a < 3; a++)
            if ((corner[a] + cs) * coeff[a] - bias[a] < tmin) child |= 1u << a;
        const uint32_t mir = child ^ mirror;
        const uint32_t mask = s->child_mask[nid];
        const int has = (mask >> mir) & 1u;
        if (has) {
            nid = s->first_child[nid] + (uint32_t)__builtin_popcount(mask & ((1u << mir) - 1u));
            const uint32_t level = cnt + 1;
            const float vsz = leaf * (float)(1u << (s->depth - level));
            const float diam2 = 4.0f * ((tmin * tmin * aperture) * 0.318309886183790671538f);
            const int child_leaf = (level == s->depth);
            const int term = (diam2 > vsz * vsz) || ((s->depth - level) <= max_query_level_offset);
            if (child_leaf || term) {
                *is_leaf = child_leaf;
                *node_out = nid;
                return tmin;
            }
        }
        empty = !has;
        for (int a = 0; a < 3; a++)
            if ((child >> a) & 1u) corner[a] += cs;
        stack = (stack << 3) | child;
        cnt++;
    }
    *node_out = OM_SVO_NO_NODE;
    return tmin;
}

/* rays: n x 7 floats (origin, direction, tmin) */
void om_svo_cone_trace_batch(const om_svo* s, const float* rays, uint32_t n, float aperture, uint32_t offset,
                             uint32_t* nodes, float* t, uint32_t* leaf)
{
    for (uint32_t i = 0; i < n; i++) {
        const float* r = rays + 7 * (size_t)i;
        int lf;
        t[i] = om_svo_cone_trace(s, r, r + 3, r[6], aperture, offset, &nodes[i], &lf);
        leaf[i] = (uint32_t)lf;
    }
}

/* GenerateRadianceField's trace and radiance query (WFPGTracerKC.cuh:
 * 1252-1318, RecursiveConeTrace.cuh:219-247, AnisoSVO.cuh:433-489,
 * 1035-1053) for every texel of every bin: the jittered concentric-
 * octahedral direction (y up), a cone of 4 pi / X^2 (OctohedralConeAperture,
 * WFPGTracer.cu:134-141) from the bin's origin and tmin, and the stopped
 * node's front or back radiance by its normal (boundary radiance on a
 * miss).  origins: [bin][4]; jitter: [bin][2]; irr: [node][2]; normals:
 * [node][3]; fields / nodes / t: [bin][Y][X] (nodes, t may be NULL). */
void om_wfpg_trace_fields(const om_svo* s, uint32_t X, uint32_t Y, uint32_t bins, const float* origins,
                          const float* jitter, const float* irr, const float* normals, float boundary,
                          uint32_t offset, float* fields, uint32_t* nodes, float* t)
{
    const float aperture = (4.0f * 3.14159265358979323846f) / ((float)X * (float)X);
    for (uint32_t b = 0; b < bins; b++)
        for (uint32_t y = 0; y < Y; y++)
            for (uint32_t x = 0; x < X; x++) {
                const size_t i = ((size_t)b * Y + y) * X + x;
                const float uv[2] = {((float)x + jitter[2 * b]) / (float)X, ((float)y + jitter[2 * b + 1]) / (float)Y};
                float z[3];
                om_octa_to_dirs(uv, 1, z);
                const float d[3] = {z[1], z[2], z[0]};
                uint32_t node;
                int leaf;
                const float tt = om_svo_cone_trace(s, origins + 4 * (size_t)b, d, origins[4 * (size_t)b + 3], aperture,
                                                   offset, &node, &leaf);
                float rad = boundary;
                if (node != OM_SVO_NO_NODE) {
                    const float* nrm = normals + 3 * (size_t)node;
                    const int front = (nrm[0] * -d[0] + nrm[1] * -d[1] + nrm[2] * -d[2]) >= 0.0f;
                    rad = front ? irr[2 * (size_t)node] : irr[2 * (size_t)node + 1];
                }
                fields[i] = rad;
                if (nodes) nodes[i] = node;
                if (t) t[i] = tt;
            }
}

/* ---- Node payload of the guiding octree (VoxelPayload, AnisoSVO.cuh:23-115):
 * per node a normal and the average radiance leaving its front and back
 * side, as the reference forms them:
 * - leaf normals: a 2-means clustering (4 rounds, seeded with the first
 *   record's normal and its negation; a normal joins the mean it is FARTHER
 *   from, the reference's rule, which only swaps the two labels each round)
 *   of the unit normals of the leaf's voxel records' triangles, in record
 *   order; the normal is the average of the larger cluster (its length < 1
 *   says how much the normals spread) (KCReduceVoxelPayload,
 *   AnisoSVOKC.cuh:8-230);
 * - internal normals: the same clustering over the children's average
 *   normals (KCAverageNormals, AnisoSVO.cu:419-520);
 * - radiance: each path vertex (position, wo toward the previous vertex,
 *   luminance) goes to the nearest existing leaf (its own cell, else the
 *   neighbours nearest first, NearestNodeIndex, AnisoSVO.cuh:1168-1240) and
 *   adds to the side its wo faces (WriteLeafRadiance, :491-536); sums are
 *   48.16 fixed point so that they do not depend on the order of the adds;
 * - leaf averages, then node averages bottom-up over the children's valid
 *   sides (a child whose normal opposes the node's swaps its sides), then
 *   sides without samples take the parent's value top-down
 *   (NormalizeAndFilterRadiance, AnisoSVO.cu:1415-1483, 530-656).  A side
 *   with no sample anywhere above keeps the marker 65504 (the reference's
 *   half maximum).
 * vertices: [n][7] (position, wo, luminance).  Outputs [nodes][2] and
 * [nodes][4] (unit normal xyz, length of the average normal). */
#define OM_PAYLOAD_INVALID 65504.0f

static void kmeans2(const float (*v)[3], uint32_t n, float out[4])
{
    /* the first non-zero vector seeds the means */
    uint32_t first = n;
    for (uint32_t i = 0; i < n; i++)
        if (!(v[i][0] == 0.0f && v[i][1] == 0.0f && v[i][2] == 0.0f)) { first = i; break; }
    if (first == n) {
        out[0] = out[1] = out[2] = out[3] = 0.0f;
        return;
    }
    float m[2][3] = {{v[first][0], v[first][1], v[first][2]}, {-v[first][0], -v[first][1], -v[first][2]}};
    uint32_t cnt[2] = {0, 0};
    for (int it = 0; it < 4; it++) {
        for (int k = 0; k < 2; k++) {
            const float inv = 1.0f / sqrtf(m[k][0] * m[k][0] + m[k][1] * m[k][1] + m[k][2] * m[k][2]);
            m[k][0] *= inv;
            m[k][1] *= inv;
            m[k][2] *= inv;
        }
        float acc[2][3] = {{0, 0, 0}, {0, 0, 0}};
        cnt[0] = cnt[1] = 0;
        for (uint32_t i = 0; i < n; i++) {
            if (v[i][0] == 0.0f && v[i][1] == 0.0f && v[i][2] == 0.0f) continue;
            const float d0 = 1.0f - (m[0][0] * v[i][0] + m[0][1] * v[i][1] + m[0][2] * v[i][2]);
            const float d1 = 1.0f - (m[1][0] * v[i][0] + m[1][1] * v[i][1] + m[1][2] * v[i][2]);
            const int k = (d0 >= d1) ? 0 : 1;
            acc[k][0] += v[i][0];
            acc[k][1] += v[i][1];
            acc[k][2] += v[i][2];
            cnt[k]++;
        }
        for (int k = 0; k < 2; k++)
            for (int a = 0; a < 3; a++) m[k][a] = acc[k][a];
    }
    const int k = (cnt[0] >= cnt[1]) ? 0 : 1;
    const float r = 1.0f / (float)cnt[k];
    const float mean[3] = {m[k][0] * r, m[k][1] * r, m[k][2] * r};
    const float len = sqrtf(mean[0] * mean[0] + mean[1] * mean[1] + mean[2] * mean[2]);
    const float il = 1.0f / len;
    out[0] = mean[0] * il;
    out[1] = mean[1] * il;
    out[2] = mean[2] * il;
    out[3] = len;
}

/* deepest node on the path of `code` that is a leaf, or OM_SVO_NO_NODE */
static uint32_t leaf_of(const om_svo* s, uint32_t code)
{
    uint32_t node = 0;
    for (uint32_t l = 0; l < s->depth; l++) {
        const uint32_t oct = (code >> (3u * (s->depth - 1u - l))) & 7u;
        const uint32_t m = s->child_mask[node];
        if (!((m >> oct) & 1u)) return OM_SVO_NO_NODE;
        node = s->first_child[node] + (uint32_t)__builtin_popcount(m & ((1u << oct) - 1u));
    }
    return node;
}

/* NearestNodeIndex (AnisoSVO.cuh:1168-1240): the point's own leaf cell, then
 * its neighbours towards the nearest faces first; cells outside the grid are
 * skipped */
static uint32_t nearest_leaf(const om_svo* s, const float p[3])
{
    const uint32_t R = 1u << s->depth;
    for (int a = 0; a < 3; a++)
        if (p[a] < s->cube[a] || p[a] > s->cube[a] + s->cube[3]) return OM_SVO_NO_NODE;
    const float leaf = s->cube[3] / (float)R;
    float fi[3], fr[3], dist[3];
    int inc[3];
    for (int a = 0; a < 3; a++) {
        const float l = (p[a] - s->cube[a]) / leaf;
        fi[a] = floorf(l);
        fr[a] = l - fi[a];
        inc[a] = (fr[a] < 0.5f) ? -1 : 1;
        dist[a] = 0.5f - fabsf(fr[a] - 0.5f);
    }
    int m0 = 0;
    if (dist[1] < dist[m0]) m0 = 1;
    if (dist[2] < dist[m0]) m0 = 2;
    const int o1 = (m0 + 1) % 3, o2 = (m0 + 2) % 3;
    const int m1 = (dist[o2] < dist[o1]) ? o2 : o1;
    const int m2 = 3 - m0 - m1;
    for (uint32_t i = 0; i < 8; i++) {
        int sh[3];
        sh[m0] = (int)(i & 1u);
        sh[m1] = (int)((i >> 1) & 1u);
        sh[m2] = (int)((i >> 2) & 1u);
        int64_t q[3];
        int ok = 1;
        for (int a = 0; a < 3; a++) {
            q[a] = (int64_t)fi[a] + sh[a] * inc[a];
            if (q[a] < 0 || q[a] >= (int64_t)R) ok = 0;
        }
        if (!ok) continue;
        const uint32_t code = vox_expand10((uint32_t)q[0]) | (vox_expand10((uint32_t)q[1]) << 1) |
                              (vox_expand10((uint32_t)q[2]) << 2);
        const uint32_t leafn = leaf_of(s, code);
        if (leafn != OM_SVO_NO_NODE) return leafn;
    }
    return OM_SVO_NO_NODE;
}

void om_svo_payload(const om_svo* s, const float* tris, const uint32_t* codes, const uint32_t* prims, uint32_t n_rec,
                    const float* vertices, uint32_t n_vert, float* irr, float* normals)
{
    const uint32_t nodes = s->nodes, D = s->depth;
    const uint32_t leaf0 = s->level_offset[D];
    float (*nv)[4] = (float (*)[4])normals;
    /* leaf normals: records grouped by leaf in record order (counting sort) */
    uint32_t* cnt = (uint32_t*)calloc(nodes + 1, sizeof(uint32_t));
    uint32_t* leafof = (uint32_t*)malloc((size_t)(n_rec ? n_rec : 1) * sizeof(uint32_t));
    for (uint32_t r = 0; r < n_rec; r++) {
        leafof[r] = leaf_of(s, codes[r]);
        if (leafof[r] != OM_SVO_NO_NODE) cnt[leafof[r] + 1]++;
    }
    for (uint32_t i = 0; i < nodes; i++) cnt[i + 1] += cnt[i];
    uint32_t* pos = (uint32_t*)malloc((size_t)(nodes + 1) * sizeof(uint32_t));
    for (uint32_t i = 0; i <= nodes; i++) pos[i] = cnt[i];
    uint32_t* order = (uint32_t*)malloc((size_t)(n_rec ? n_rec : 1) * sizeof(uint32_t));
    for (uint32_t r = 0; r < n_rec; r++)
        if (leafof[r] != OM_SVO_NO_NODE) order[pos[leafof[r]]++] = r;
    float (*buf)[3] = (float (*)[3])malloc((size_t)(n_rec ? n_rec : 8) * 3 * sizeof(float));
    for (uint32_t lf = leaf0; lf < nodes; lf++) {
        const uint32_t b = cnt[lf], e = cnt[lf + 1];
        for (uint32_t k = b; k < e; k++) {
            const float* t = tris + 9 * (size_t)prims[order[k]];
            const float ea[3] = {t[3] - t[0], t[4] - t[1], t[5] - t[2]};
            const float eb[3] = {t[6] - t[0], t[7] - t[1], t[8] - t[2]};
            float nn[3] = {ea[1] * eb[2] - ea[2] * eb[1], ea[2] * eb[0] - ea[0] * eb[2], ea[0] * eb[1] - ea[1] * eb[0]};
            const float inv = 1.0f / sqrtf(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
            buf[k - b][0] = nn[0] * inv;
            buf[k - b][1] = nn[1] * inv;
            buf[k - b][2] = nn[2] * inv;
        }
        kmeans2((const float (*)[3])buf, e - b, nv[lf]);
    }
    /* internal normals bottom-up over the children's average normals */
    for (int l = (int)D - 1; l >= 0; l--)
        for (uint32_t nd = s->level_offset[l]; nd < s->level_offset[l + 1]; nd++) {
            const uint32_t fc = s->first_child[nd], k = (uint32_t)__builtin_popcount(s->child_mask[nd]);
            float ch[8][3];
            for (uint32_t i = 0; i < k; i++)
                for (int a = 0; a < 3; a++) ch[i][a] = nv[fc + i][a] * nv[fc + i][3];
            kmeans2((const float (*)[3])ch, k, nv[nd]);
        }
    /* splat in 48.16 fixed point */
    uint64_t* tot = (uint64_t*)calloc((size_t)nodes * 2, sizeof(uint64_t));
    uint32_t* num = (uint32_t*)calloc((size_t)nodes * 2, sizeof(uint32_t));
    for (uint32_t v = 0; v < n_vert; v++) {
        const float* x = vertices + 7 * (size_t)v;
        const uint32_t lf = nearest_leaf(s, x);
        if (lf == OM_SVO_NO_NODE) continue;
        const float nd = nv[lf][0] * x[3] + nv[lf][1] * x[4] + nv[lf][2] * x[5];
        const int side = (nd >= 0.0f) ? 0 : 1;
        const float lum = x[6];
        const uint64_t f = (lum > 0.0f) ? (uint64_t)(fminf(lum, 2147483647.0f) * 65536.0f) : 0ull;
        tot[2 * (size_t)lf + side] += f;
        num[2 * (size_t)lf + side]++;
    }
    for (uint32_t lf = leaf0; lf < nodes; lf++)
        for (int j = 0; j < 2; j++) {
            const uint32_t c = num[2 * (size_t)lf + j];
            irr[2 * (size_t)lf + j] = c ? ((float)tot[2 * (size_t)lf + j] * (1.0f / 65536.0f)) / (float)c
                                        : OM_PAYLOAD_INVALID;
        }
    /* bottom-up over the valid sides, swapping a child whose normal opposes */
    for (int l = (int)D - 1; l >= 0; l--)
        for (uint32_t nd = s->level_offset[l]; nd < s->level_offset[l + 1]; nd++) {
            const uint32_t fc = s->first_child[nd], k = (uint32_t)__builtin_popcount(s->child_mask[nd]);
            float sum[2] = {0.0f, 0.0f};
            uint32_t c[2] = {0, 0};
            for (uint32_t i = 0; i < k; i++) {
                const float* cn = nv[fc + i];
                const int swap = (cn[0] * nv[nd][0] + cn[1] * nv[nd][1] + cn[2] * nv[nd][2]) < 0.0f;
                for (int j = 0; j < 2; j++) {
                    const float val = irr[2 * (size_t)(fc + i) + (swap ? 1 - j : j)];
                    if (val != OM_PAYLOAD_INVALID) {
                        sum[j] += val;
                        c[j]++;
                    }
                }
            }
            for (int j = 0; j < 2; j++) irr[2 * (size_t)nd + j] = c[j] ? sum[j] * (1.0f / (float)c[j]) : OM_PAYLOAD_INVALID;
        }
    /* top-down: sides without samples take the parent's */
    for (uint32_t l = 1; l <= D; l++)
        for (uint32_t nd = s->level_offset[l]; nd < s->level_offset[l + 1]; nd++) {
            const uint32_t p = s->parent[nd];
            for (int j = 0; j < 2; j++)
                if (irr[2 * (size_t)nd + j] == OM_PAYLOAD_INVALID) irr[2 * (size_t)nd + j] = irr[2 * (size_t)p + j];
        }
    free(cnt);
    free(leafof);
    free(pos);
    free(order);
    free(buf);
    free(tot);
    free(num);
}
