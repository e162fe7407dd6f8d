/*
 * oracle.h — CPU restatement of MRay's wavefront path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg are the only callers.  The product never links this.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the reference checkout).  Layouts:
 *   tris   : float[n][9]  (v0.xyz, v1.xyz, v2.xyz), primitive order
 *   aabb   : float[n][6]  (min.xyz, max.xyz)
 *   rays   : 7 planes of float[pitch]: ox, oy, oz, dx, dy, dz, tmax
 *   hits   : prim u32 (0xFFFFFFFF = miss), t f32, bary a/b f32 (2 planes)
 *   LBVH   : leaves are nodes [0,n), internal node p is node n+p and its
 *            split lies between sorted leaves p and p+1;
 *            left/right[p] and parent[] hold global node indices,
 *            parent[root] = 0xFFFFFFFF.
 */
#ifndef MRH_ORACLE_H
#define MRH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OM_INVALID 0xFFFFFFFFu

/* ---- Morton codes (RayLib/MortonCode.h:43-158, LinearBVH.hpp:119-166,400-414) */
uint32_t om_morton_expand10(uint32_t v);
uint64_t om_morton_expand21(uint32_t v);
uint32_t om_morton_compose30(uint32_t x, uint32_t y, uint32_t z);
uint64_t om_morton_compose63(uint32_t x, uint32_t y, uint32_t z);
void     om_tri_aabbs(const float* tris, uint32_t n, float* aabbs);
void     om_extent(const float* aabbs, uint32_t n, float ext[6]);
float    om_morton_delta(const float ext[6], uint32_t bits_per_dim);
void     om_quantise(const float aabb[6], const float ext[6], float delta,
                     uint32_t bits_per_dim, uint32_t q[3]);
/* bits = 30 writes uint32 codes, bits = 63 writes uint64 codes */
void     om_morton_codes(const float* tris, uint32_t n, uint32_t bits, void* codes);

/* ---- Sorting / partition (RayMemory.cu:187-312, LinearBVH.hpp:489-492) ---- */
/* stable ascending sort of (code, index); writes sorted codes and indices */
void om_sort_pairs_u32(const uint32_t* keys, uint32_t n, uint32_t* sorted_keys, uint32_t* sorted_idx);
void om_sort_pairs_u64(const uint64_t* keys, uint32_t n, uint64_t* sorted_keys, uint32_t* sorted_idx);
/* Stable DESCENDING sort by key bits [begin_bit, end_bit); order[] receives
 * the input positions.  offsets[] (2^(end-begin) + 1 entries) receives, for
 * every digit d, the first sorted position holding digit d; entries are in
 * descending digit order, so offsets[0] is digit max .. ; offsets[last] = n. */
void om_partition_desc(const uint32_t* keys, uint32_t n, uint32_t begin_bit, uint32_t end_bit,
                       uint32_t* order, uint32_t* offsets);

/* ---- LBVH (LinearBVH.hpp:30-100, 183-396, 423-596) ---- */
/* Returns 0 on success.  sorted_codes may be NULL.  Arrays sized for n >= 1:
 * sorted_idx[n], left[n-1], right[n-1], parent[2n-1], node_aabb[(2n-1)*6]. */
int om_lbvh_build(const float* tris, uint32_t n, uint32_t bits,
                  void* sorted_codes, uint32_t* sorted_idx,
                  uint32_t* left, uint32_t* right, uint32_t* parent,
                  float* node_aabb, uint32_t* root);

/* ---- Ray casting (Ray.hpp:143-283, GPUAcceleratorBVHKC.cuh:229-346,
 *      GPUAcceleratorCommonKC.cuh:33-78) ---- */
int  om_intersect_triangle(const float o[3], const float d[3], const float tri[9],
                           float* t, float bary[3]);
int  om_intersect_aabb(const float o[3], const float inv_d[3], const float box[6],
                       float tmin, float tmax, float* t_enter);
/* box test of the 4-wide walks: planes ordered by the direction's sign,
 * distances fma(plane, inv, -o * inv) */
int  om_slab_box_test(const float o[3], const float d[3], const float box[6], float tmin, float tmax,
                      float* t_enter);
/* stats (optional, may be NULL): 2 u32 per ray = internal-node visits, triangle tests */
void om_trace_closest(const float* tris, uint32_t n, const uint32_t* sorted_idx,
                      const uint32_t* left, const uint32_t* right, const float* node_aabb,
                      uint32_t root, const float* rays, uint32_t nrays, uint32_t pitch,
                      float tmin, uint32_t* hit_prim, float* hit_t, float* hit_bary,
                      uint32_t* stats);
/* any-hit: hit_prim receives 1 if any triangle with t in [tmin, tmax) exists, else 0 */
void om_trace_any(const float* tris, uint32_t n, const uint32_t* sorted_idx,
                  const uint32_t* left, const uint32_t* right, const float* node_aabb,
                  uint32_t root, const float* rays, uint32_t nrays, uint32_t pitch,
                  float tmin, uint32_t* hit_any);
/* 4-wide traversal tree collapsed from the LBVH (om_geom.c); nodes4 is
 * float[n-1][32]; returns the root code. */
uint32_t om_bvh4_build(uint32_t n, const uint32_t* left, const uint32_t* right, const float* node_aabb,
                       uint32_t root, float* nodes4);
void om_trace4_closest(const float* tris, uint32_t n, const uint32_t* sorted_idx, const float* nodes4,
                       uint32_t root4, const float* rays, uint32_t nrays, uint32_t pitch, float tmin,
                       uint32_t* hit_prim, float* hit_t, float* hit_bary, uint32_t* stats);
void om_trace4_any(const float* tris, uint32_t n, const uint32_t* sorted_idx, const float* nodes4,
                   uint32_t root4, const float* rays, uint32_t nrays, uint32_t pitch, float tmin,
                   uint32_t* hit_any);
void om_trace_brute(const float* tris, uint32_t n, const float* rays, uint32_t nrays,
                    uint32_t pitch, float tmin, uint32_t* hit_prim, float* hit_t,
                    float* hit_bary);

/* ---- Two-level scenes (GPUAcceleratorBVHKC.cuh:274-345 and 513-652,
 *      GPUTransformSingle.cuh:141-197, Ray.hpp:461-465, Matrix.hpp:268-289) ----
 * Meshes are ranges of one triangle array in local space, each with its own
 * LBVH; instances place a mesh by a row-major 3x4 local-to-world matrix; a
 * top-level LBVH is built over the instances' world boxes given as the
 * degenerate triangles (min, max, max). */
typedef struct {
    uint32_t        n_inst;
    const float*    w2l;             /* [n_inst][12] world-to-local (om_affine_inverse of l2w) */
    const float*    l2w;             /* [n_inst][12] local-to-world */
    const uint32_t* inst_mesh;       /* [n_inst] */
    const uint32_t* inst_key;        /* [n_inst] HitKey override (0 = the mesh's keys), or NULL */
    const uint32_t* top_sorted_idx;  /* top tree: sorted leaf -> instance */
    const float*    top_nodes4;
    uint32_t        top_root4;
    uint32_t        n_meshes;
    const uint32_t* mesh_first;      /* [n_meshes + 1] */
    const uint32_t* mesh_sorted_idx; /* [n]: mesh m's sorted leaves at [first[m], first[m+1]), mesh-local */
    const float*    mesh_nodes4;     /* concatenated [sum(n_m - 1)][32] */
    const uint32_t* mesh_left;       /* concatenated [sum(n_m - 1)], mesh-local node numbers */
    const uint32_t* mesh_right;
    const float*    mesh_node_aabb;  /* concatenated [sum(2 n_m - 1)][6] */
    const uint32_t* mesh_root;       /* [n_meshes] binary root (mesh-local node number) */
    const uint32_t* mesh_root4;      /* [n_meshes] */
    const uint32_t* mesh_ibase;      /* [n_meshes] first internal node in nodes4 / left / right */
    const uint32_t* mesh_abase;      /* [n_meshes] first node in node_aabb */
} om_two_level;

void     om_xf_ray(const float m12[12], const float o[3], const float d[3], float out_o[3], float out_d[3]);
void     om_xf_point(const float m12[12], const float p[3], float out[3]);
int      om_affine_inverse(const float l2w[12], float w2l[12]);
void     om_world_box(const float l2w[12], const float box[6], float out[6]);
uint32_t om_trace2l_one(const om_two_level* tl, const float* tris, uint32_t width, const float o[3],
                        const float d[3], float tmin, float* tmax_io, float bary_out[2], int any_hit,
                        uint32_t* inst_out, uint32_t* visits, uint32_t* tests);
void     om_trace2l_closest(const om_two_level* tl, const float* tris, uint32_t width, const float* rays,
                            uint32_t nrays, uint32_t pitch, float tmin, uint32_t* hit_inst, uint32_t* hit_prim,
                            float* hit_t, float* hit_bary, uint32_t* stats);

/* ---- Sampling helpers (RayLib/HemiDistribution.h, Ray.hpp:503-560,
 *      TracerFunctions.cuh) ---- */
void  om_sincos_2pi_f(float u, float* s, float* c);
void  om_reflect(const float wi[3], const float n[3], float out[3]);
int   om_refract(const float wi[3], const float n[3], float from, float to, float out[3]);
void  om_triangle_normal(const float tri[9], float out[3]);
void  om_aabb_centroid(const float box[6], float out[3]);
void  om_hemi_uniform(float xi0, float xi1, float dir[3], float* pdf);
void  om_nudge(const float p[3], const float n[3], float out[3]);
void  om_rot_between_z(const float n[3], float q[4]);
void  om_quat_rotate(const float q[4], const float v[3], float out[3]);
float om_fresnel_dielectric(float cos_in, float ior_in, float ior_out);
void  om_fresnel_conductor(float cos_in, const float eta[3], const float k[3], float out[3]);
float om_dggx(float ndh, float alpha);
float om_gsmith_single(const float v[3], float alpha);
float om_gsmith_correlated(const float wo[3], const float wi[3], float alpha);
float om_vndf_pdf(const float v[3], const float h[3], float alpha);
void  om_vndf_sample(const float v[3], float alpha, float xi0, float xi1, float h[3], float* pdf);

/* ---- Scene + path tracer (PathTracer.cu:144-291, PathTracerKC.cuh:337-648,
 *      SimpleMaterialsKC.cuh:16-303, MetalGlassMaterialsKC.cuh:11-222,
 *      GenerationKernels.cuh:29-108, GPUCameraPinhole.cuh:140-240) ---- */

/* Work batches of the 32-bit HitKey (batch:8 | id:24, RayLib/HitStructs.h:42-77) */
#define OM_BATCH_NULL       0u
#define OM_BATCH_LAMBERT    1u
#define OM_BATCH_METAL      2u
#define OM_BATCH_DIELECTRIC 3u
#define OM_BATCH_EMISSIVE   4u
#define OM_BATCH_REFLECT    5u

typedef struct {
    const float*    tris;          /* n x 9 */
    const uint32_t* prim_key;      /* n keys */
    uint32_t        n;
    const float*    lambert;       /* albedo rgb per id */
    const float*    metal;         /* per id: roughness, eta rgb, k rgb, specular rgb (10) */
    const float*    dielectric;    /* per id: albedo rgb, ior (4) */
    const float*    emissive;      /* radiance rgb per id */
    /* LBVH over tris */
    const uint32_t* sorted_idx;
    const uint32_t* left;
    const uint32_t* right;
    const float*    node_aabb;
    uint32_t        root;
    /* 4-wide traversal tree (om_bvh4_build); NULL = walk the binary LBVH */
    const float*    nodes4;
    uint32_t        root4;
    /* light list: primitive index of every emissive triangle, in primitive
     * order (one primitive-backed light per triangle) */
    const uint32_t* lights;
    uint32_t        n_lights;
    const float*    reflect;       /* per id: albedo rgb, roughness (4) */
    /* two-level scene (NULL: single-level): tris are the meshes' local
     * triangles and the trees above are unused; lights are scene primitives
     * with their instances in light_inst (instance-major order) */
    const om_two_level* two_level;
    uint32_t        tree_width;    /* two-level: the meshes' tree width, 4 or 2 */
    const uint32_t* light_inst;
    /* smooth shading (NULL: flat): per-corner world-to-tangent rotations,
     * [n][3][4] (w, x, y, z) in primitive order (om_corner_frames) */
    const float*    tbn;
} om_scene;

typedef struct {
    float position[3];
    float bottom_left[3];
    float right[3];
    float up[3];
    float plane_size[2];
    float near_far[2];
} om_camera;

/* GPUCameraPinhole.cuh:140-170 */
void om_camera_setup(const float pos[3], const float gaze[3], const float up[3],
                     float fov_x, float near_d, float far_d, uint32_t w, uint32_t h,
                     om_camera* cam);

typedef struct {
    uint32_t width, height;
    uint32_t samples;        /* strata per axis: spp = samples^2 */
    uint32_t max_depth;
    uint32_t rr_start;
    uint32_t seed;
    uint32_t row_step;       /* render rows y with y % row_step == row_offset */
    uint32_t row_offset;
    uint32_t threads;
    uint32_t nee;            /* next-event estimation ("NextEventEstimation") */
    uint32_t mis;            /* power-heuristic MIS of NEE and BxDF ("DirectLightMIS"), with nee only */
} om_render_params;

/* Next-event estimation pieces (GPULightSamplerUniform.cuh:82-130,
 * GPULightPrimitive.cuh:139-166 / 245-256, TracerFunctions.cuh:302-309):
 * light sample from xi = (light choice, two barycentric numbers), returning
 * the light's primitive, unit direction, squared distance and pdf (solid
 * angle / light count); the solid-angle pdf of hitting triangle tri at
 * distance t along unit d; the power-2 heuristic weight of pdf0. */
uint32_t om_sample_light(const om_scene* sc, const float p[3], const float xi[3], float dir[3], float* dist2,
                         float* pdf);
float    om_light_pdf_hit(const float tri[9], float t, const float d[3]);
float    om_power_heuristic(float pdf0, float pdf1);

/* Reconstruction filters (om_filter.c) */
#define OM_FILTER_BOX      0u
#define OM_FILTER_TENT     1u
#define OM_FILTER_GAUSSIAN 2u
#define OM_FILTER_MITCHELL 3u
typedef struct {
    uint32_t type;
    float    radius, alpha, b, c;
} om_filter_desc;
/* FilterRadiusToPixelWH */
int   om_filter_pixel_wh(float radius);
/* filter weight at pixel-centre offset (dx, dy) from the sample */
float om_filter_weight(const om_filter_desc* d, float dx, float dy);
/* arbitrary samples (values n x 3, coords n x 2), accumulated (+=) into
 * weighted (w*h x 3) and weights (w*h) */
void  om_filter_samples(const om_filter_desc* d, const float* values, const float* coords, uint32_t n,
                        uint32_t w, uint32_t h, float* weighted, float* weights);
/* samples grouped per pixel: pixel q owns samples q*spp .. q*spp + spp - 1 */
void  om_filter_samples_grouped(const om_filter_desc* d, const float* values, const float* coords,
                                uint32_t spp, uint32_t w, uint32_t h, float* weighted, float* weights);

/* Renders the selected rows; img is double[h*w*3] (zeroed by the caller),
 * accumulating the per-sample FP32 contributions in FP64.  Returns the number
 * of ray segments traced. */
uint64_t om_render(const om_scene* s, const om_camera* cam, const om_render_params* p,
                   double* img);
/* All rows; every sample's radiance (FP32, contributions summed in path
 * order) into values (W*H*spp x 3) and its image coordinate into coords
 * (W*H*spp x 2), kept inside its own pixel; global sample ids start at
 * sample_base (a tracer's later frames).  Returns the segments traced. */
uint64_t om_render_samples(const om_scene* s, const om_camera* cam, const om_render_params* p, uint32_t sample_base,
                           float* values, float* coords);

/* ---- Path guiding (AnisoSVO.cu:225-263, 657-1188, 1415-1483;
 *      WFPGTracerKC.cuh:202-520, 1480-1766; WFPGTracer.cu:1590-1790) ---- */
typedef struct om_svo {
    uint32_t depth, nodes;
    uint32_t level_count[11], level_offset[12];
    float    cube[4];                 /* lo xyz, size */
    uint32_t *codes, *first_child, *child_mask, *parent;
    uint64_t *acc;                    /* [nodes][64] 48.16 fixed-point splat sums */
    uint32_t *count;                  /* [nodes][64] */
    float    *sum;                    /* [nodes][64] means before the blur (after filter) */
    float    *radiance;               /* [nodes][64] filtered */
    uint32_t *node_count;             /* [nodes] samples in the subtree (after filter) */
    float    centers[64][3];
} om_svo;
/* 0 on success; codes are 30-bit cell codes at `depth` (1..10) */
int      om_svo_build(om_svo* s, const uint32_t* codes, uint32_t n, const float cube[4], uint32_t depth);
void     om_svo_free(om_svo* s);
uint32_t om_svo_lookup(const om_svo* s, uint32_t code);
uint32_t om_svo_lookup_guide(const om_svo* s, uint32_t code, uint32_t min_count);
void     om_svo_splat(om_svo* s, const uint32_t* codes, const uint32_t* bins, const float* values,
                      uint32_t n);
void     om_svo_filter(om_svo* s);
/* ray binning by octree node (BinRayCount collapse) and the partition of
 * the rays by bin; order / part_* may be NULL */
void     om_svo_bin_rays(const om_svo* s, const uint32_t* codes, uint32_t n, uint32_t min_level, uint32_t min_rays,
                         uint32_t* bins, uint32_t* order, uint32_t* part_offsets, uint32_t* part_bins,
                         uint32_t* part_count);
/* pts / dirs / normals / xi: n x 3 floats; uv: n x 2 */
void     om_svo_codes(const float* pts, uint32_t n, const float cube[4], uint32_t depth, uint32_t* codes);
void     om_dir_to_bins(const float* dirs, uint32_t n, uint32_t* bins);
void     om_octa_to_dirs(const float* uv, uint32_t n, float* dirs);
void     om_dir_to_octas(const float* dirs, uint32_t n, float* uv);
float    om_atan_unit(float t);
void     om_guide_sample(const om_svo* s, const uint32_t* nodes, const float* normals, const float* xi,
                         uint32_t n, float* dirs, float* pdf);
void     om_guide_pdf(const om_svo* s, const uint32_t* nodes, const float* normals, const float* dirs,
                      uint32_t n, float* pdf);
typedef struct {
    uint32_t iterations, svo_depth;
    float    guide_ratio;
    uint32_t min_samples;             /* guide lookup stops above sparser nodes */
} om_guide_params;
/* All rows (row_step / row_offset ignored); img = sum over all frames; *last
 * (may be NULL) receives the final guide (free with om_svo_free). */
uint64_t om_render_guided(const om_scene* s, const om_camera* cam, const om_render_params* p,
                          const om_guide_params* g, double* img, om_svo* last);
/* Camera ray of one sample (for raygen parity tests): writes o, d, and the
 * path RNG state after the two jitter draws. */
void om_camera_ray(const om_camera* cam, uint32_t w, uint32_t h, uint32_t samples,
                   uint32_t seed, uint32_t sample_id, float o[3], float d[3],
                   uint64_t* rng_state);
void om_camera_ray_at(const om_camera* cam, uint32_t w, uint32_t h, uint32_t samples,
                      uint32_t sample_id, float jx, float jy, float o[3], float d[3]);
void om_pcg32_stream(uint64_t initstate, uint64_t initseq, uint32_t n, uint32_t* out);
uint64_t om_splitmix64(uint64_t z);
void om_path_uniforms(uint32_t seed, uint32_t sample_id, uint32_t n, float* out);

/* smooth shading frames (GPUPrimitiveTriangle.cpp:286-316, RayLib/Quaternion.hpp:256-436) */
void om_orthogonal_vector(const float v[3], float out[3]);
void om_space_quat(const float x[3], const float y[3], const float z[3], float q[4]);
void om_qslerp(const float a[4], const float b[4], float t, float out[4]);
void om_bary_slerp(const float q0[4], const float q1[4], const float q2[4], float a, float b, float out[4]);
void om_corner_frames(const float* normals, uint32_t n, float* out);
void om_smooth_frame(const float tri[9], const float tq[12], float ba, float bb, const float dir[3], float q[4],
                     float normal[3]);

/* ---- Per-bin radiance fields of the reference's wavefront path guiding
 * (WFPGTracerKC.cuh:1252-1545): filter, piecewise-constant distribution,
 * one guided sample per ray.  Fields are [bin][y][x]; cdf arrays hold the
 * end of each texel; uv / xi are 2 floats per ray, dirs 3 (y up). */
typedef struct {
    uint32_t        field_x, field_y, bin_count, ray_count;
    float           filter_alpha;
    const float*    fields;
    const uint32_t* bin_offsets;      /* bin_count + 1 */
    const uint32_t* ray_ids;
    const float*    xi;
    const float*    eval_uv;          /* may be NULL */
    float*          out_uv;
    float*          out_dir;          /* may be NULL */
    float*          out_pdf;
    float*          out_eval_pdf;     /* written iff eval_uv */
    float*          out_filtered;     /* may be NULL */
    const float*    bxdf_field;       /* product mode when not NULL: [ray][64] */
    uint32_t        seed;             /* product mode's random stream */
} om_wfpg_desc;
void  om_wfpg_gauss_weights(float alpha, float w[11]);
void  om_octa_wrap_int(int32_t x, int32_t y, int32_t X, int32_t Y, int32_t out[2]);
void  om_wfpg_filter(const float* field, uint32_t X, uint32_t Y, const float w[11], float* out);
void  om_wfpg_pwc(const float* f, uint32_t X, uint32_t Y, float* pdfx, float* cdfx, float* pdfy, float* cdfy);
void  om_wfpg_pwc_sample(const float* cdfx, const float* cdfy, uint32_t X, uint32_t Y, float xi0, float xi1,
                         float uv[2]);
float om_wfpg_pwc_pdf(const float* pdfx, const float* pdfy, uint32_t X, uint32_t Y, float u, float v);
int   om_wfpg_sample_fields(const om_wfpg_desc* d);
float om_wfpg_uniform(uint32_t seed, uint32_t ray, uint32_t k);

/* ---- Triangle voxeliser of the guiding octree (GPUPrimitiveTriangle.h:
 * 384-620): records of every conservatively covered pixel, triangle by
 * triangle, row by row; codes are cell codes at `depth` of the cube. */
uint64_t om_voxelize(const float* tris, uint32_t n, const float cube[4], uint32_t depth, uint32_t* codes,
                     uint32_t* prims, uint64_t cap);
/* Cone trace through an om_svo_build octree (AnisoSVO.cuh:767-1032):
 * returns t; *node = node stopped at (0xFFFFFFFF: left the cube).
 * rays for the batch: n x 7 floats (origin, direction, tmin). */
float    om_svo_cone_trace(const om_svo* s, const float o[3], const float dir[3], float tmin, float aperture,
                           uint32_t max_query_level_offset, uint32_t* node, int* is_leaf);
void     om_svo_cone_trace_batch(const om_svo* s, const float* rays, uint32_t n, float aperture, uint32_t offset,
                                 uint32_t* nodes, float* t, uint32_t* leaf);
/* node payload of the guiding octree (AnisoSVO.cu:225-263, 419-656,
 * 1415-1483; AnisoSVOKC.cuh:8-230): vertices [n][7] (position, wo,
 * luminance); irr [nodes][2] (front, back); normals [nodes][4] (unit xyz,
 * length of the average normal) */
void     om_svo_payload(const om_svo* s, const float* tris, const uint32_t* codes, const uint32_t* prims,
                        uint32_t n_rec, const float* vertices, uint32_t n_vert, float* irr, float* normals);
/* cone-traced radiance fields of bins (WFPGTracerKC.cuh:1252-1318) */
void     om_wfpg_trace_fields(const om_svo* s, uint32_t X, uint32_t Y, uint32_t bins, const float* origins,
                              const float* jitter, const float* irr, const float* normals, float boundary,
                              uint32_t offset, float* fields, uint32_t* nodes, float* t);

#ifdef __cplusplus
}
#endif

#endif
