/*
 * mrh.h — C-ABI of the MI355X wavefront path-tracing core.
 *
 * Plain pointers and sizes only.  "d_" arguments are device pointers in the
 * context's device memory (the caller owns them; the library never frees
 * caller memory).  Every call is asynchronous on the context's stream unless
 * it says otherwise; mrh_sync() waits for it.  Every entry point returns an
 * mrh_status; on failure mrh_last_error() holds the message.  No C++
 * exceptions cross this boundary.
 *
 * Each declaration cites the reference interface it replaces
 * (paths relative to the reference checkout).
 */
#ifndef MRH_H
#define MRH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    MRH_OK = 0,
    MRH_ERR_INVALID = 1,      /* bad argument (TracerError::..., SceneError::...) */
    MRH_ERR_HIP = 2,          /* HIP runtime error */
    MRH_ERR_OOM = 3,          /* device allocation failed (TracerException on OOM, CudaCheck.h:32-42) */
    MRH_ERR_UNSUPPORTED = 4,  /* option or type not implemented */
    MRH_ERR_BUDGET = 5,       /* request exceeds the memory budget */
    MRH_ERR_CANCELLED = 6,    /* serving: the request was cancelled (mrh_server_cancel) */
    MRH_ERR_TIMEOUT = 7       /* serving: the request outlived mrh_server_params.timeout_ms */
} mrh_status;

typedef struct mrh_context mrh_context;
typedef struct mrh_lbvh mrh_lbvh;
typedef struct mrh_scene mrh_scene;
typedef struct mrh_tracer mrh_tracer;

/* ------------------------------------------------------------------ */
/* Context: device, stream, scratch memory, per-kernel timers.         */
/* Replaces CudaSystem::Initialize + CudaGPU (TracerCUDA/CudaSystem.h:97-118,
 * CudaSystem.cpp:109-122) as created by TracerSystemCUDA::Initialize
 * (TracerSystemCUDA.cpp:19).                                           */
/* ------------------------------------------------------------------ */
int32_t     mrh_context_create(int32_t device, mrh_context** out);
int32_t     mrh_context_destroy(mrh_context* ctx);
const char* mrh_last_error(const mrh_context* ctx);
int32_t     mrh_sync(mrh_context* ctx);
/* hipStream_t of the context, returned as void* */
int32_t     mrh_stream(mrh_context* ctx, void** stream);
/* Copy between any two device/host pointers on the context stream
 * (hipMemcpyDefault), then wait for it.  Utility for hosts that keep their
 * buffers outside any framework. */
int32_t     mrh_memcpy(mrh_context* ctx, void* dst, const void* src, size_t bytes);
/* Device memory on the context's device for hosts without an allocator of
 * their own (DeviceMemory, TracerCUDA/DeviceMemory.cpp; no managed memory). */
int32_t     mrh_device_alloc(mrh_context* ctx, void** out, size_t bytes);
int32_t     mrh_device_free(mrh_context* ctx, void* ptr);
/* Per-kernel hipEvent timing on the context stream (the reference's
 * cudaEvent timers, WFPGTracer.cu:592-596).  When enabled every launch is
 * bracketed by events; mrh_profile_get returns total ms and launches for a
 * kernel name, summed since the last mrh_profile_reset. */
int32_t     mrh_profile_enable(mrh_context* ctx, int32_t on);
int32_t     mrh_profile_reset(mrh_context* ctx);
int32_t     mrh_profile_get(mrh_context* ctx, const char* kernel, double* total_ms,
                            uint64_t* launches);
/* Launch-shape knobs of the context, set by the autotuner (no counterpart in
 * the reference, whose block counts are fixed in CudaSystem.cpp:109-122):
 *   "trace_lds_entries"   traversal stack entries kept in LDS per thread
 *                         (0, 8, 12, 16, 24; 0 = whole stack in scratch)
 *   "trace_blocks_per_cu" persistent traversal grid, blocks of 256 per CU
 *   "trace_width"         traversal tree: 2 = the binary LBVH, 4 = the 4-wide
 *                         tree collapsed from it (trees deeper than 84 use 2)
 *   "shade_partition"     material partition of a bounce: 0 = one global
 *                         radix pass over the HitKey batch byte (the rays are
 *                         shaded through the sorted order), 1 = each block
 *                         partitions its own tile of rays in LDS
 *   "trace_block"         threads per block of the frame renders' traversal
 *                         kernel (128, 256, 512; two-level scenes and the
 *                         serving engine use 256)
 *   "shade_tile"          rays per tile of shade_partition 1 (512, 1024,
 *                         2048)
 *   "fail_alloc_after"    failure-handling tests: k > 0 makes the k-th
 *                         following device allocation of this context fail
 *                         with MRH_ERR_OOM (0 = off)
 * mrh_tune_get reads a knob back. */
int32_t     mrh_tune(mrh_context* ctx, const char* knob, int64_t value);
int32_t     mrh_tune_get(mrh_context* ctx, const char* knob, int64_t* value);

/* ------------------------------------------------------------------ */
/* Morton codes.  Replaces KCGenerateAABBs + ReduceArrayGPU + FindMortonDelta
 * + KCGenMortonCodes (TracerCUDA/LinearBVH.hpp:103-166, 400-414, 446-477).
 * d_tris: float[n][9] triangle corners.  bits = 30 writes uint32 codes,
 * bits = 63 writes uint64 codes.  Quantisation: centroid = min + 0.5*span,
 * delta = float(max span * 2^-b) (b = 10 or 21), IEEE division, clamp to
 * 2^b - 1.                                                              */
/* ------------------------------------------------------------------ */
int32_t mrh_morton_codes(mrh_context* ctx, const float* d_tris, uint32_t n, uint32_t bits,
                         void* d_codes);

/* ------------------------------------------------------------------ */
/* Onesweep LSD radix sort of (key, value) pairs.  Replaces
 * RadixSortValKeyGPU (TracerCUDA/ParallelRadixSort.cuh) and the two
 * cub::DeviceRadixSort::SortPairsDescending passes of RayMemory::SortKeys
 * (TracerCUDA/RayMemory.cu:187-228).  Stable.  key_bytes = 4 or 8; bits
 * [begin_bit, end_bit) are sorted; d_vals_in = NULL means values 0..n-1.
 * Input and output ranges must not overlap.                             */
/* ------------------------------------------------------------------ */
int32_t mrh_sort_pairs(mrh_context* ctx, const void* d_keys_in, const uint32_t* d_vals_in,
                       void* d_keys_out, uint32_t* d_vals_out, uint32_t n, uint32_t key_bytes,
                       uint32_t begin_bit, uint32_t end_bit, int32_t descending);

/* Partition ray ids by HitKey.  Replaces RayMemory::SortKeys + Partition
 * (TracerCUDA/RayMemory.cu:187-312) and RayCaster::PartitionRaysWRTWork
 * (TracerCUDA/RayCaster.cu:21-37): a stable DESCENDING sort of (key, id)
 * over bits [sort_begin_bit, end_bit), d_order receives the ids.
 * d_offsets receives 2^(end_bit - part_begin_bit) + 1 entries: entry j is
 * the first sorted position whose bits [part_begin_bit, end_bit) equal
 * (2^w - 1 - j), the last entry is n; so batch b's rays are
 * d_order[d_offsets[2^w-1-b] .. d_offsets[2^w-b]), the NullBatch (0) tail
 * last.  end_bit - part_begin_bit <= 8 and part_begin_bit >= sort_begin_bit.
 * No host round trip (the reference copies the split list to a host
 * std::set every bounce, RayMemory.cu:275-297).                          */
int32_t mrh_partition_by_key(mrh_context* ctx, const uint32_t* d_keys, uint32_t n,
                             uint32_t sort_begin_bit, uint32_t part_begin_bit, uint32_t end_bit,
                             uint32_t* d_order, uint32_t* d_offsets);

/* mrh_partition_by_key that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, hosts that forbid allocation).  Same
 * arguments, results and errors as mrh_partition_by_key, bit for bit.  Its
 * workspace (look-back status, histograms, ping-pong keys and values) is a
 * fixed 68 MiB array in device memory, one per device, held until the
 * process ends; n above MRH_PARTITION_NOALLOC_MAX_KEYS fails with
 * MRH_ERR_INVALID.  Every call on a device shares that device's workspace:
 * calls that overlap on one device (two contexts, two streams, two threads)
 * race and must be serialised by the caller.  Records no profiler events. */
#define MRH_PARTITION_NOALLOC_MAX_KEYS (1u << 22)
int32_t mrh_partition_by_key_noalloc(mrh_context* ctx, const uint32_t* d_keys, uint32_t n,
                                     uint32_t sort_begin_bit, uint32_t part_begin_bit,
                                     uint32_t end_bit, uint32_t* d_order, uint32_t* d_offsets);

/* ------------------------------------------------------------------ */
/* LBVH.  Replaces LinearBVHCPU::Construct (TracerCUDA/LinearBVH.hpp:423-596)
 * and, as the ray-casting accelerator, GPUAccBVHGroup::ConstructAccelerator
 * (TracerCUDA/GPUAcceleratorBVH.hpp:353-640).  Topology in the reference's
 * numbering: leaves are nodes [0,n) in sorted order, internal node p is
 * node n+p with its split between sorted leaves p and p+1, the parent is
 * chosen by ChooseParentAndUpdate (LinearBVH.hpp:43-100); duplicate codes are
 * ordered by the (code, index) tie-break instead of aborting
 * (LinearBVH.hpp:506-521).                                                */
/* ------------------------------------------------------------------ */
typedef struct {
    uint32_t    n;            /* leaves */
    uint32_t    bits;         /* 30 or 63 */
    uint32_t    root;         /* global node index of the root */
    uint32_t    max_depth;    /* deepest leaf (root = depth 0) */
    const void* d_sorted_codes;   /* uint32[n] or uint64[n] */
    const uint32_t* d_sorted_idx; /* primitive of sorted leaf i */
    const uint32_t* d_left;       /* [n-1] global node index */
    const uint32_t* d_right;      /* [n-1] */
    const uint32_t* d_parent;     /* [2n-1]; root holds 0xFFFFFFFF */
    const float*    d_node_aabb;  /* [(2n-1)][6] min.xyz max.xyz */
    /* 4-wide traversal tree collapsed from the LBVH: node p (an internal
     * node at even depth) = float[32]: min x[4], min y[4], min z[4],
     * max x[4], max y[4], max z[4], child code[4] (0x80000000 | sorted
     * leaf, internal index, 0xFFFFFFFF empty), pad[4]; slots are the
     * children with internal children replaced by their two children. */
    const float*    d_nodes4;     /* [n-1][32]; odd-depth entries unused */
    uint32_t        root4;        /* code of the 4-wide root */
    uint32_t        reserved;
} mrh_lbvh_view;

/* d_tris is read during the call only; the caller may free it after the
 * call returns.  Synchronous (returns after the build). */
int32_t mrh_lbvh_build(mrh_context* ctx, const float* d_tris, uint32_t n, uint32_t bits,
                       mrh_lbvh** out);
int32_t mrh_lbvh_view_get(const mrh_lbvh* bvh, mrh_lbvh_view* out);
int32_t mrh_lbvh_destroy(mrh_lbvh* bvh);
/* Rebuilds in place from new corners of the same count (timed builds);
 * d_tris as for mrh_lbvh_build (NULL is invalid). */
int32_t mrh_lbvh_rebuild(mrh_context* ctx, mrh_lbvh* bvh, const float* d_tris);

/* ------------------------------------------------------------------ */
/* Ray casting.  Replaces GPUAccBVHGroup::Hit -> KCIntersectBVH /
 * KCIntersectBVHStackless + AcceptHit (TracerCUDA/GPUAcceleratorBVHKC.cuh:
 * 229-511, GPUAcceleratorCommonKC.cuh:33-78) and GPUBaseAcceleratorI::Hit
 * (GPUAcceleratorI.h:149-166).
 * d_rays: 7 planes of float[pitch]: ox, oy, oz, dx, dy, dz, tmax.
 * Outputs: d_prim (primitive index, 0xFFFFFFFF = miss), d_t, d_bary
 * (2 planes of float[n]: weights of corners 0 and 1).  Closest hit accepts
 * t in [tmin, tmax).  d_stats (optional): 2 uint32 per ray = internal-node
 * visits and triangle tests.                                             */
/* ------------------------------------------------------------------ */
int32_t mrh_trace_closest(mrh_context* ctx, const mrh_lbvh* bvh, const float* d_rays,
                          uint32_t n, uint32_t pitch, float tmin, uint32_t* d_prim,
                          float* d_t, float* d_bary, uint32_t* d_stats);
/* any hit (shadow rays): d_hit[i] = 1 if some t in [tmin, tmax) exists */
int32_t mrh_trace_any(mrh_context* ctx, const mrh_lbvh* bvh, const float* d_rays, uint32_t n,
                      uint32_t pitch, float tmin, uint32_t* d_hit);

/* ------------------------------------------------------------------ */
/* Scenes and the plain wavefront path tracer.                          */
/* Replaces GPUSceneI + the Triangle primitive group, the LambertC / Metal /
 * Refract material groups and primitive area lights
 * (TracerCUDA/GPUPrimitiveTriangle.h, SimpleMaterialsKC.cuh:16-303,
 * MetalGlassMaterialsKC.cuh:11-222, GPULightP.cuh:133-140), and
 * PathTracer::GenerateWork/Render/Finalize (TracerCUDA/PathTracer.cu:144-313). */
/* ------------------------------------------------------------------ */

/* HitKey work batches (RayLib/HitStructs.h:42-77: key = batch:8 | id:24) */
#define MRH_BATCH_NULL       0u
#define MRH_BATCH_LAMBERT    1u
#define MRH_BATCH_METAL      2u
#define MRH_BATCH_DIELECTRIC 3u
#define MRH_BATCH_EMISSIVE   4u
#define MRH_BATCH_REFLECT    5u   /* "Reflect": perfect mirror with albedo */
#define MRH_KEY(batch, id)   (((uint32_t)(batch) << 24) | ((uint32_t)(id) & 0xFFFFFFu))

typedef struct {
    const float*    tris;        /* host float[n][9] */
    const uint32_t* prim_key;    /* host uint32[n] HitKey per triangle */
    uint32_t        n;
    const float*    lambert;     /* host float[n_lambert][3] albedo */
    uint32_t        n_lambert;
    const float*    metal;       /* host float[n_metal][10]: roughness, eta rgb, k rgb, specular rgb */
    uint32_t        n_metal;
    const float*    dielectric;  /* host float[n_dielectric][4]: albedo rgb, ior */
    uint32_t        n_dielectric;
    const float*    emissive;    /* host float[n_emissive][3] radiance */
    uint32_t        n_emissive;
    uint32_t        morton_bits; /* 30 or 63 */
    const float*    reflect;     /* host float[n_reflect][4]: albedo rgb, roughness (mirror at 0) */
    uint32_t        n_reflect;
    /* NULL: a single-level scene, tris in world space.  Otherwise a
     * two-level scene: tris / prim_key are the meshes' triangles in their
     * local spaces, and the instances place them in the world. */
    const struct mrh_instancing_desc* instancing;
    /* NULL: flat shading (every corner carries the face frame).  Otherwise
     * host float[n][9]: a shading normal per triangle corner (any nonzero
     * length; normalised).  Each corner gets the world-to-tangent rotation
     * of its normal with the tangent OrthogonalVector(n), Gram-Schmidt and
     * bitangent n x t (GPUPrimitiveTriangle.cpp:286-316, TransformGen::Space);
     * shading blends the three with Quat::BarySLerp
     * (GPUPrimitiveTriangle.h:626-670).  Single-level scenes only. */
    const float* normals;
} mrh_scene_desc;

/* Two-level scenes.  Replaces the base accelerator over per-surface
 * accelerators (GPUBaseAcceleratorBVH: KCIntersectBaseBVH,
 * TracerCUDA/GPUAcceleratorBVHKC.cuh:513-652) and the inner traversal of
 * a transformed surface (GPUAcceleratorBVHKC.cuh:274-345, the ray taken to
 * local space by GPUTransformSingle::WorldToLocal = Ray::Transform(inverse),
 * GPUTransformSingle.cuh:141-147, RayLib/Ray.hpp:461-465).  Every mesh gets
 * its own LBVH; a top-level LBVH is built over the instances' world boxes
 * (the mesh root box's eight corners taken to the world,
 * GPUTransformSingle.cuh:180-197).  One trace launch walks the top tree and,
 * at an instance, the mesh's tree with the ray in local space; the
 * direction is not renormalised, so t is the world ray's parameter.  The
 * world-to-local matrix is the inverse of xform computed in double and
 * rounded to float. */
typedef struct mrh_instancing_desc {
    const uint32_t* mesh_first;  /* host uint32[n_meshes + 1]: mesh m owns triangles
                                    [mesh_first[m], mesh_first[m + 1]); mesh_first[0] = 0,
                                    mesh_first[n_meshes] = n, no mesh is empty */
    uint32_t        n_meshes;
    const float*    xform;       /* host float[n_instances][12]: local-to-world, row-major 3x4
                                    (the last row of the 4x4 is 0 0 0 1) */
    const uint32_t* inst_mesh;   /* host uint32[n_instances] */
    const uint32_t* inst_key;    /* host uint32[n_instances] or NULL: a HitKey for all of the
                                    instance's triangles, 0 = the mesh's own keys (the
                                    per-surface material of a scene file) */
    uint32_t        n_instances;
} mrh_instancing_desc;

typedef struct {
    float    position[3];
    float    gaze[3];        /* point looked at */
    float    up[3];
    float    fov_x;          /* radians; fov_y follows the aspect ratio */
    float    near_plane;
    float    far_plane;
} mrh_camera_desc;

typedef struct {
    uint32_t width, height;
    uint32_t samples;        /* strata per axis ("Samples"), spp = samples^2 */
    uint32_t max_depth;      /* "MaxDepth" */
    uint32_t rr_start;       /* "RussianRouletteStart" */
    uint32_t nee;            /* "NextEventEstimation": 1 = one light sample per
                                non-specular vertex (uniform light choice over the
                                emissive triangles, uniform point on it) */
    uint32_t seed;
    uint32_t count_traversal; /* 1: also count node visits / triangle tests (slower) */
    uint32_t direct_light_mis; /* "DirectLightMIS": 1 = power-heuristic MIS of the
                                NEE and BxDF strategies (used with nee = 1 only) */
} mrh_render_params;

typedef struct {
    uint64_t segments;        /* rays traced over all bounces */
    uint32_t bounce_rays[32]; /* live rays per bounce (first 32 bounces) */
    uint32_t bounces;
    double   render_ms;       /* wall time of the bounce loop on the stream */
    uint64_t node_visits;     /* internal-node visits (count_traversal = 1) */
    uint64_t tri_tests;       /* triangle tests (count_traversal = 1) */
} mrh_render_stats;

/* Uploads the scene and builds its LBVH (synchronous). */
int32_t mrh_scene_create(mrh_context* ctx, const mrh_scene_desc* desc, mrh_scene** out);
int32_t mrh_scene_destroy(mrh_scene* scene);
/* The scene's LBVH (MRH_ERR_UNSUPPORTED for a two-level scene: use
 * mrh_scene_trace_closest) */
int32_t mrh_scene_bvh(const mrh_scene* scene, const mrh_lbvh** out);
/* Device bytes held by the scene (TracerSystemI/GPUTracerI::TotalGPUMemoryUsed) */
int32_t mrh_scene_bytes(const mrh_scene* scene, uint64_t* bytes);
/* Closest hit against a scene, single- or two-level (GPUBaseAcceleratorI::Hit
 * followed by the inner accelerators' Hit, GPUAcceleratorI.h:149-166).  Rays
 * and outputs as mrh_trace_closest; d_prim is the index into the scene's
 * tris, d_inst (may be NULL) the instance (0 in a single-level scene,
 * 0xFFFFFFFF on a miss).  d_stats counts the top tree's internal nodes and
 * the meshes' internal nodes together. */
int32_t mrh_scene_trace_closest(mrh_context* ctx, const mrh_scene* scene, const float* d_rays, uint32_t n,
                                uint32_t pitch, float tmin, uint32_t* d_inst, uint32_t* d_prim, float* d_t,
                                float* d_bary, uint32_t* d_stats);

/* Renders one frame: d_image (device float[height][width][3], bottom row
 * first) receives the per-pixel mean radiance.  Synchronous; stats may be NULL. */
int32_t mrh_render(mrh_context* ctx, mrh_scene* scene, const mrh_camera_desc* cam,
                   const mrh_render_params* params, float* d_image, mrh_render_stats* stats);

/* Device bytes mrh_render needs for a frame of this size */
int32_t mrh_render_bytes(const mrh_render_params* params, uint64_t* bytes);

/* ------------------------------------------------------------------ */
/* Reconstruction filters (GPUReconFilterI and its Box / Tent / Gaussian /
 * Mitchell implementations, GPUReconFilter.h:59-260, GPUReconFilterKC.cuh,
 * options "type" / "radius" / "alpha" / "b" / "c", TracerSystemCUDA.cpp:
 * 109-140).  A sample at image coordinate c (pixel units, pixel (x, y)
 * covers [x, x+1) x [y, y+1)) reaches the pixels of the conservative window
 * of FilterRadiusToPixelWH(radius) whose centres lie within `radius` of c
 * (radius 0: its own pixel only), with weight filter(centre - c).  A pixel
 * holds sum(w * value) and sum(w); the image is their ratio (0 where the
 * weights sum to 0).  Within a pixel, contributions are summed in a fixed
 * order (neighbour rows, then columns, then sample order), so results are
 * reproducible.                                                          */
/* ------------------------------------------------------------------ */
typedef enum {
    MRH_FILTER_BOX = 0,       /* "Box": 1 inside |dx|, |dy| < radius (1 everywhere at radius 0) */
    MRH_FILTER_TENT = 1,      /* "Tent": (1 - |dx|/r)(1 - |dy|/r), clamped at 0 */
    MRH_FILTER_GAUSSIAN = 2,  /* "Gaussian": (e^(-a dx^2) - e^(-a r^2))+ (same in y) */
    MRH_FILTER_MITCHELL = 3   /* "Mitchell": Mitchell-Netravali (b, c) of 2|dx|/r times of 2|dy|/r */
} mrh_filter_type;

typedef struct {
    uint32_t type;            /* mrh_filter_type */
    float    radius;          /* pixels, >= 0; Tent, Gaussian and Mitchell need > 0 */
    float    alpha;           /* Gaussian "alpha" */
    float    b, c;            /* Mitchell "b", "c" */
} mrh_filter_desc;

/* GPUReconFilterI::FilterToImg: n samples (values float[n][3], image
 * coordinates float[n][2]) are filtered into a w x h image, accumulating
 * (+=) d_weighted (float[h][w][3], sum of w * value) and d_weights
 * (float[h][w], sum of w).  Samples outside [0, w) x [0, h) are ignored.
 * Synchronous. */
int32_t mrh_filter_samples(mrh_context* ctx, const mrh_filter_desc* filter, const float* d_values,
                           const float* d_coords, uint32_t n, uint32_t width, uint32_t height,
                           float* d_weighted, float* d_weights);

/* d_image = d_weighted / d_weights per pixel (0 where the weight is 0) */
int32_t mrh_filter_resolve(mrh_context* ctx, const float* d_weighted, const float* d_weights,
                           uint32_t width, uint32_t height, float* d_image);

/* Scene files (host only; no device needed).  Reads the reference's JSON
 * scene format (GPUSceneJson.cpp:690-885: Cameras, Lights, Mediums,
 * Transforms, Materials, Primitives, Surfaces, LightSurfaces, CameraSurfaces;
 * // and slash-star comments allowed) with Wavefront OBJ meshes (paths relative
 * to the scene file; innerIndex = mesh of the file, meshes split at o / g /
 * usemtl) into world-space triangles, HitKeys and material tables
 * (mrh_scene_file_desc, ready for mrh_scene_create) and Pinhole cameras
 * (CameraSurfaces order).  Supported: Null / Primitive lights, Vacuum /
 * Homogeneous media, Identity / Single transforms, LambertC or Lambert with
 * a constant albedo, Reflect, Refract, Metal, inline and OBJ triangles.
 * Vertex normals (OBJ vn, the "normal" lists of nodeTriangle and
 * nodeTriangleIndexed) become mrh_scene_desc.normals in world space (inverse
 * transpose, renormalised); triangles without normals in a scene with some
 * get their face normal, and the instanced mode below shades flat, noting
 * it in the warnings.  Anything else fails with MRH_ERR_UNSUPPORTED (malformed files with
 * MRH_ERR_INVALID), the reason in msg; with MRH_SCENE_FILE_SUBSTITUTE it is
 * replaced (a Lambert for other materials, a tessellated sphere for analytic
 * ones) or skipped instead and listed in mrh_scene_file_warnings. */
typedef struct mrh_scene_file mrh_scene_file;
#define MRH_SCENE_FILE_STRICT     0u
#define MRH_SCENE_FILE_SUBSTITUTE 1u
/* Instanced (may be or-ed with the above): every primitive a surface or
 * light surface uses becomes one mesh in local space and every surface
 * primitive an instance with the surface's transform and material key, so
 * the scene is two-level (mrh_scene_desc.instancing) instead of baked into
 * world-space triangles. */
#define MRH_SCENE_FILE_INSTANCED  2u
int32_t mrh_scene_file_load(const char* path, uint32_t flags, mrh_scene_file** out, char* msg, uint32_t msg_len);
/* pointers into the file object's arrays (valid until it is destroyed) */
int32_t mrh_scene_file_desc(const mrh_scene_file* file, mrh_scene_desc* out);
int32_t mrh_scene_file_cameras(const mrh_scene_file* file, mrh_camera_desc* out, uint32_t max, uint32_t* count);
const char* mrh_scene_file_warnings(const mrh_scene_file* file);
int32_t mrh_scene_file_destroy(mrh_scene_file* file);
/* meshes of an OBJ file and the triangle count of the first `max` */
int32_t mrh_obj_mesh_counts(const char* path, uint32_t* meshes, uint32_t* triangles, uint32_t max);

/* Image files (host only; no device needed).  ImageIO::WriteImage with
 * ImageType::EXR (ImageIO/ImageIOI.h:83-91) writes an OpenEXR 2.0 scanline
 * file, uncompressed, FP32 channels R, G, B; PFM is the portable float
 * map ("PF", little endian).  rgb: host float[height][width][3], bottom row
 * first (the layout every image of this library has). */
typedef enum { MRH_IMAGE_EXR = 0, MRH_IMAGE_PFM = 1 } mrh_image_type;
int32_t mrh_image_write(const char* path, uint32_t type, uint32_t width, uint32_t height, const float* rgb);

/* mrh_render with a reconstruction filter (RayTracer::Finalize ->
 * FilterToImg, RayTracer.cu:81-95): every sample's radiance and image
 * coordinate are kept and filtered into d_image.  With Box radius 0 the
 * image equals mrh_render's per-pixel mean. */
int32_t mrh_render_filtered(mrh_context* ctx, mrh_scene* scene, const mrh_camera_desc* cam,
                            const mrh_render_params* params, const mrh_filter_desc* filter, float* d_image,
                            mrh_render_stats* stats);

/* ------------------------------------------------------------------ */
/* Path-guiding sparse voxel octree.  Replaces the WFPG tracer's SVO
 * (AnisoSVOctreeCPU::Constrcut AnisoSVO.cu:657-1188, KCAccumulateRadianceToLeaf
 * AnisoSVO.cu:225-263, NormalizeAndFilterRadiance AnisoSVO.cu:1415-1483) and
 * its radiance-field sampling (KCGenAndSampleDistributionProduct,
 * WFPGTracerKC.cuh:1562-1766) with the structure of this library: leaves are
 * the occupied cells of a 2^depth grid (depth <= 10) over a cube, every node
 * holds 8x8 FP32 radiance bins over the equal-area concentric octahedral map.
 * Codes are 30-bit Morton codes of cells at `depth` (cube = lo.xyz, size).  */
/* ------------------------------------------------------------------ */
typedef struct mrh_svo mrh_svo;
typedef struct {
    uint32_t depth, nodes;
    uint32_t level_count[11];      /* nodes per level, root = level 0 */
    uint32_t level_offset[12];     /* first node of each level */
    float    cube[4];
    const uint32_t* d_codes;       /* [nodes] level code (cell code >> 3*(depth-level)) */
    const uint32_t* d_first_child; /* [nodes] 0xFFFFFFFF at leaves */
    const uint32_t* d_child_mask;  /* [nodes] 8-bit occupancy */
    const uint32_t* d_parent;      /* [nodes] 0xFFFFFFFF at the root */
    const float*    d_sum;         /* [nodes][64] after filter: per-bin means before blur */
    const uint32_t* d_count;       /* [nodes][64] samples per bin */
    const float*    d_radiance;    /* [nodes][64] filtered radiance (3x3 blur + 1e-4 floor) */
    const uint32_t* d_node_count;  /* [nodes] samples in the subtree (after filter) */
} mrh_svo_view;

int32_t mrh_svo_create(mrh_context* ctx, mrh_svo** out);
int32_t mrh_svo_destroy(mrh_svo* svo);
/* d_points: 3 planes of float[pitch] */
int32_t mrh_svo_codes(mrh_context* ctx, const float* d_points, uint32_t n, uint32_t pitch,
                      const float cube[4], uint32_t depth, uint32_t* d_codes);
/* Builds the octree over the cells of d_codes and clears its radiance (synchronous). */
int32_t mrh_svo_build(mrh_context* ctx, mrh_svo* svo, const uint32_t* d_codes, uint32_t n,
                      const float cube[4], uint32_t depth);
/* Adds value i to bin d_bins[i] of the deepest node holding d_codes[i]
 * (codes 0xFFFFFFFF are skipped). */
int32_t mrh_svo_splat(mrh_context* ctx, mrh_svo* svo, const uint32_t* d_codes, const uint32_t* d_bins,
                      const float* d_values, uint32_t n);
int32_t mrh_svo_filter(mrh_context* ctx, mrh_svo* svo);
/* Deepest existing node on each code's root-to-cell path; with min_count > 0
 * the deepest whose subtree holds at least min_count samples (after filter). */
int32_t mrh_svo_lookup(mrh_context* ctx, const mrh_svo* svo, const uint32_t* d_codes, uint32_t n,
                       uint32_t min_count, uint32_t* d_nodes);
/* Product sampling: bin ~ radiance x max(cos(n, bin centre), 0), then a
 * uniform point of the bin; d_normals, d_xi (3 planes), d_dirs 3 planes of
 * float[n]; d_pdf solid-angle pdf (0 when the product vanishes). */
int32_t mrh_svo_sample(mrh_context* ctx, const mrh_svo* svo, const uint32_t* d_nodes,
                       const float* d_normals, const float* d_xi, uint32_t n, float* d_dirs,
                       float* d_pdf);
int32_t mrh_svo_pdf(mrh_context* ctx, const mrh_svo* svo, const uint32_t* d_nodes,
                    const float* d_normals, const float* d_dirs, uint32_t n, float* d_pdf);
int32_t mrh_svo_view_get(const mrh_svo* svo, mrh_svo_view* out);
/* Ray binning by octree node.  Replaces WFPGTracer::AccumulateRayHitsToSVO
 * (WFPGTracer.cu:488-588: leaf of each ray's hit, sort, run-length encode,
 * leaf ray counts), AnisoSVOctreeCPU::CollapseRayCounts (AnisoSVO.cu:265-350,
 * 1485-1529), KCCheckReducedSVOBins (WFPGTracerKC.cuh:1055-1078) and
 * RayCaster::PartitionRaysWRTCustomData (the partition of the rays by bin).
 * d_codes: cell codes (mrh_svo_codes at the octree's depth) of the rays'
 * hit points, 0xFFFFFFFF for rays that take no bin (next-event rays,
 * misses).  Each ray starts at the deepest existing node of its cell; from
 * the leaf level up to min_level (clamped to the depth, "MinRayBinLevel") a
 * node is marked when it is at min_level or holds at least bin_ray_count
 * rays ("BinRayCount", 1..MRH_BIN_RAY_COUNT_MAX), and hands its rays to its
 * parent otherwise.  d_bins[i] = the first marked node up from ray i's
 * start node (the start node if none is marked), 0xFFFFFFFF for rays
 * without a code.  With d_order: the ray ids sorted by bin (ascending node,
 * stable), binned rays first; d_part_offsets[k] / d_part_bins[k] = first
 * sorted position and node of the k-th bin, d_part_offsets[count] = the
 * number of binned rays, *d_part_count = count (device word; n + 1, n and 1
 * words).  Everything stays on the device. */
#define MRH_BIN_RAY_COUNT_MAX 4096u
int32_t mrh_svo_bin_rays(mrh_context* ctx, mrh_svo* svo, const uint32_t* d_codes, uint32_t n, uint32_t min_level,
                         uint32_t bin_ray_count, uint32_t* d_bins, uint32_t* d_order, uint32_t* d_part_offsets,
                         uint32_t* d_part_bins, uint32_t* d_part_count);

/* Geometry of the reference's guiding octree: its triangle voxeliser
 * (TriFunctions::Voxelize, GPUPrimitiveTriangle.h:384-620, with
 * EachPrimVoxelCount / VoxelizeSurfaces, GPUAcceleratorCommonKC.cuh:439-542,
 * as AnisoSVOctreeCPU::Constrcut calls them, AnisoSVO.cu:657-760).  Each
 * triangle (d_tris: n x 9 floats, world space) is projected onto the plane
 * of its normal's dominant axis with the cube (lo xyz, size) mapped to
 * [-1,1]^3, rasterised conservatively at the leaf resolution 2^depth (edges
 * pushed out by half a pixel), and every covered pixel gives one record: the
 * cell code (as mrh_svo_codes writes it at `depth`) of the leaf holding the
 * triangle's plane point under the pixel centre.  Records come triangle by
 * triangle, pixels row by row; shared edges give duplicates, which
 * mrh_svo_build folds.  *out_count = the number of records; they are written
 * to d_codes only when it holds them all (count <= capacity), so a call with
 * d_codes NULL sizes the buffer; d_prims (may be NULL) receives each record's
 * triangle index (the payload's normals come from it).  The reference pads the scene bounds by one
 * leaf on every side of its largest span to get the cube
 * (AnisoSVO.cu:670-678).  Synchronises the context's stream. */
int32_t mrh_svo_voxelize(mrh_context* ctx, const float* d_tris, uint32_t n, const float cube[4], uint32_t depth,
                         uint32_t* d_codes, uint32_t* d_prims, uint32_t capacity, uint32_t* out_count);

/* Cone trace through the octree.  Replaces AnisoSVOctreeGPU::ConeTraceRay
 * (AnisoSVO.cuh:767-1032; as GenerateRadianceField's batched trace calls it,
 * WFPGTracerKC.cuh:1282-1310).  d_rays: n x 7 floats (origin, direction,
 * tmin), world space.  Each ray walks the octree front to back and stops on
 * entering an existing node that is a leaf, whose voxel the cone's disc is
 * narrower than (diameter^2 = 4 t^2 aperture / pi > voxel size^2;
 * aperture = the cone's solid angle), or that lies max_query_level_offset or
 * fewer levels above the leaves.  d_nodes[i] = that node's global index
 * (0xFFFFFFFF when the ray leaves the cube first), d_t[i] = the t reached
 * (the node's entry), d_is_leaf[i] (may be NULL) = 1 for a leaf.  Directions
 * closer than 1e-5 to an axis plane are pushed to 1e-5, as the reference
 * does.  A walk longer than 65536 steps ends as a miss (the reference has no
 * bound). */
int32_t mrh_svo_cone_trace(mrh_context* ctx, const mrh_svo* svo, const float* d_rays, uint32_t n, float aperture,
                           uint32_t max_query_level_offset, uint32_t* d_nodes, float* d_t, uint32_t* d_is_leaf);

/* The octree's node payload for the field trace.  Replaces the reference's
 * VoxelPayload set-up (AnisoSVO.cuh:23-115): leaf normals from its voxel
 * records (KCReduceVoxelPayload, AnisoSVOKC.cuh:8-230), internal normals
 * bottom-up (KCAverageNormals, AnisoSVO.cu:419-520), the path vertices'
 * radiance into the nearest leaf's front or back side
 * (KCAccumulateRadianceToLeaf, :225-263, NearestNodeIndex, AnisoSVO.cuh:
 * 1168-1240, WriteLeafRadiance, :491-536) and NormalizeAndFilterRadiance
 * (:1415-1483, 530-656).  Normals: a 2-means clustering (4 rounds, seeded
 * with the first normal and its negation) of the unit normals of a leaf's
 * records' triangles, in record order, and of the children's average
 * normals for a node; a node's normal is the larger cluster's average
 * (node_normals: its direction; node_normal_length, optional: its length,
 * below 1 where the normals spread).  Radiance: each vertex (7 floats:
 * position, wo toward the previous path vertex, luminance) adds to the side
 * of the nearest existing leaf (its own cell, else the neighbour cells
 * nearest first) that wo faces; sums are 48.16 fixed point, so the result does
 * not depend on the order of the adds (the reference adds floats).  Leaf
 * sides average their samples; node sides average their children's sides
 * that have samples (a child whose normal opposes the node's swaps its two
 * sides); a side still without samples takes its parent's value, top down;
 * one without samples anywhere up to the root keeps 65504 (the reference's
 * half-float maximum).  The octree must have been built from the records'
 * codes (mrh_svo_voxelize with d_prims). */
typedef struct {
    const float*    tris;               /* [n_tris][9]: the voxelised triangles */
    const uint32_t* codes;              /* [n_records] voxel records (mrh_svo_voxelize) */
    const uint32_t* prims;              /* [n_records] their triangles */
    uint32_t        n_records;
    const float*    vertices;           /* [n_vertices][7] path vertices */
    uint32_t        n_vertices;
    float*          node_irradiance;    /* [nodes][2] front, back */
    float*          node_normals;       /* [nodes][3] */
    float*          node_normal_length; /* optional [nodes] */
    const uint32_t* vertex_mask;        /* optional [n_vertices]: a vertex whose entry is
                                           0xFFFFFFFF is skipped (path-vertex slots left empty) */
} mrh_svo_payload_desc;
int32_t mrh_svo_payload(mrh_context* ctx, const mrh_svo* svo, const mrh_svo_payload_desc* desc);

/* The radiance fields of the bins, by cone tracing.  Replaces
 * GenerateRadianceField's trace and radiance query (WFPGTracerKC.cuh:
 * 1252-1318: BatchConeTracer::BatchedConeTraceRay, RecursiveConeTrace.cuh:
 * 219-247; AnisoSVOctreeGPU::ReadRadiance, AnisoSVO.cuh:1035-1053;
 * VoxelPayload::ReadRadiance, :433-489).  Texel (x, y) of bin b looks along
 * the concentric-octahedral direction of ((x, y) + jitter[b]) / (X, Y),
 * y up, with a cone of the texel's solid angle (4 pi / field_x^2, as
 * OctohedralConeAperture forms it, WFPGTracer.cu:134-141) from
 * bin_origins[b] (xyz, tmin: the reference takes a random ray of the bin,
 * its hit point pushed off the surface, and tmin = 0.15 leaf + 1e-5,
 * WFPGTracerKC.cuh:1132-1192).  The texel's radiance is the stopped node's
 * node_irradiance[node][0] when the cone arrives on the side its
 * node_normals[node] faces (normal . -dir >= 0), [1] otherwise, and
 * boundary_radiance when the cone leaves the cube (a constant stands for
 * the reference's boundary light).  out_fields feeds
 * mrh_wfpg_sample_fields.  The node payload (VoxelPayload's average
 * front/back radiance and normal per node) is the caller's. */
typedef struct {
    uint32_t       field_x, field_y;       /* 8, 16, 32 or 64 each */
    uint32_t       bin_count;
    uint32_t       max_query_level_offset; /* 0 in the reference's field trace */
    float          boundary_radiance;
    const float*   bin_origins;            /* [bin_count][4] */
    const float*   bin_jitter;             /* [bin_count][2] in [0,1) */
    const float*   node_irradiance;        /* [nodes][2] front, back */
    const float*   node_normals;           /* [nodes][3] */
    float*         out_fields;             /* [bin_count][field_y][field_x] */
    uint32_t*      out_nodes;              /* optional, same shape */
    float*         out_t;                  /* optional, same shape */
} mrh_wfpg_trace_desc;
int32_t mrh_wfpg_trace_fields(mrh_context* ctx, const mrh_svo* svo, const mrh_wfpg_trace_desc* desc);

/* Per-bin radiance fields of the reference's wavefront path guiding: the
 * field-to-sample half of KCGenAndSampleDistribution (WFPGTracerKC.cuh:
 * 1433-1545, with GenerateRadianceField's radiance query and filter,
 * :1312-1340).  Each bin brings a field_y x field_x field of incoming
 * radiance over the concentric octahedral map (row v, column u; the reference
 * fills it by cone tracing its voxel octree, :1282-1310).  Per bin: texels of
 * exactly 0 become 1e-4; the 11-tap Gaussian of variance filter_alpha
 * ("RFieldFilterAlpha", StaticGaussianFilter1D<11>, BlockTextureFilter.cuh:
 * 84-121) runs along u then v, folding taps outside the map back with the
 * map's edge mirroring (Utility::CocentricOctohedralWrapInt,
 * CoordinateConversion.h:213-235); a 2D piecewise-constant distribution is
 * built over the result (BlockPWCDistribution2D, GPUBlockPWCDistribution.cuh:
 * 107-200); then every ray of the bin draws (u,v) from it with its two
 * uniforms xi (BlockPWCDistribution2D::Sample, :202-275) and gets the guide
 * pdf of that sample in solid angle (pdf of the texel / 4 pi, :277-300 and
 * WFPGTracerKC.cuh:1511-1512).  With eval_uv each ray also gets the guide pdf
 * at that (u,v): the pdf the balance heuristic needs for a BxDF-sampled
 * direction (:1518-1533; the BxDF half of that MIS stays with the material
 * code).  out_dir receives the sampled direction with y up, as the reference
 * hands it to the path (NormProjectionFunc, :1420-1424).  Bins are the
 * partition that mrh_svo_bin_rays writes: rays ray_ids[bin_offsets[b] ..
 * bin_offsets[b+1]) use field b; per-ray arrays are indexed by ray id, ids
 * >= ray_count are skipped, rays of no bin are left untouched.  Results do not
 * depend on the launch shape: each ray brings its own uniforms.  All pointers
 * are device pointers, read or written during the call only. */
typedef struct {
    uint32_t        field_x, field_y;  /* texels along u and v: 8, 16, 32 or 64 each */
    uint32_t        bin_count;
    uint32_t        ray_count;         /* rays the per-ray arrays hold */
    float           filter_alpha;      /* Gaussian variance, finite and > 0 (reference default 1) */
    const float*    fields;            /* [bin_count][field_y][field_x] */
    const uint32_t* bin_offsets;       /* [bin_count + 1], non-decreasing */
    const uint32_t* ray_ids;           /* [bin_offsets[bin_count]] */
    const float*    xi;                /* [ray_count][2] */
    const float*    eval_uv;           /* optional [ray_count][2] */
    float*          out_uv;            /* [ray_count][2] */
    float*          out_dir;           /* optional [ray_count][3] */
    float*          out_pdf;           /* [ray_count] */
    float*          out_eval_pdf;      /* [ray_count], given iff eval_uv is */
    float*          out_filtered;      /* optional [bin_count][field_y][field_x] */
    /* Product mode (ProductSampler, ProductSampler.cuh:16-725, pure guiding as
     * KCGenAndSampleDistributionProduct uses it, WFPGTracerKC.cuh:1562-1766)
     * when bxdf_field is given: per ray the luminance of its BxDF towards the
     * 8x8 cell directions (cell (i, j) = row i, column j at index 8 i + j; the
     * reference evaluates them at the cell centres' octahedral directions,
     * y up, tilted by 4 pi / 64 radians about cross(direction, normal),
     * WFPGTracerKC.cuh:1602-1615).  The filtered field is reduced to 8x8
     * cells (average, maximum); each ray draws a cell from max(1e-3, bxdf *
     * cell average) and a texel inside it by rejection (24 rounds of 64
     * tries, accepted with probability texel / cell maximum); xi is not read;
     * every draw is a hash of (seed, ray id, draw index).  out_pdf and
     * out_eval_pdf are the product pdfs (solid angle). */
    const float*    bxdf_field;        /* optional [ray_count][64] */
    uint32_t        seed;
} mrh_wfpg_fields_desc;
int32_t mrh_wfpg_sample_fields(mrh_context* ctx, const mrh_wfpg_fields_desc* desc);

/* Guided path tracing: replaces WFPGTracer::Render / Finalize
 * (WFPGTracer.cu:1590-1790) and the WFPG path work (WFPGTracerKC.cuh:202-520,
 * options WFPGTracer.h:25-96).  `iterations` frames of samples^2 spp each;
 * every frame but the last records its path vertices (PathNode.cuh:152-190,
 * PPGTracerKC.cuh:605-620); after it the octree is rebuilt over the vertices'
 * cells (depth svo_depth, cube = scene bounds padded by one voxel), their
 * incident radiance is splatted and the tree filtered; the next frame draws
 * each non-specular bounce from guide (probability guide_ratio) or BxDF and
 * weights it by the mixture pdf.  d_image = mean of all frames.  `svo` may be
 * NULL; otherwise it keeps the last guide for inspection.  Deterministic:
 * splat sums are fixed point, so equal inputs give equal guides and paths.
 * With params.nee (and direct_light_mis) each non-specular vertex also sends
 * a next-event ray as in mrh_render (WFPGTracerKC.cuh:294-373, BxDF pdf in
 * the NEE weight, the mixture pdf as a path ray's prevPDF); the guide learns
 * the radiance that path rays bring back, unweighted, and next-event rays do
 * not propagate into it. */
typedef struct {
    uint32_t iterations;     /* frames, 1..64 (1 = plain path tracing) */
    uint32_t svo_depth;      /* "OctreeLevel", 1..10 */
    float    guide_ratio;    /* "MISRatio": probability of sampling the guide */
    uint32_t min_samples;    /* guide lookup stops above nodes with fewer samples ("BinRayCount") */
} mrh_guide_params;

int32_t mrh_render_guided(mrh_context* ctx, mrh_scene* scene, const mrh_camera_desc* cam,
                          const mrh_render_params* params, const mrh_guide_params* guide, mrh_svo* svo,
                          float* d_image, mrh_render_stats* stats);

/* Guided render with the reference's own guide (WFPGTracer: Render,
 * GenerateGuidedDirections and Finalize, WFPGTracer.cu:590-760, 1590-1790;
 * the path work's use of the guided direction, WFPGTracerKC.cuh:380-430).
 * Replaces the octree guide of mrh_render_guided with the stages of the
 * voxelised octree: the scene's triangles are voxelised at svo_depth
 * (mrh_svo_voxelize, mrh_svo_build), `iterations` frames follow as in
 * mrh_render_guided, every frame but the last records its path vertices and
 * forms the octree's node payload from them when it ends (mrh_svo_payload),
 * and every later frame, at every bounce: draws the one-sample mixture
 * choice for each non-specular hit (guide with probability guide_ratio,
 * else the BxDF's own sample), bins the rays by octree node
 * (mrh_svo_bin_rays with min_bin_level / bin_ray_count), cone-traces one
 * field per bin from the hit point of its ray of the smallest path id
 * (mrh_wfpg_trace_fields:
 * 64^2 texels at the first bounce, 32^2, 16^2, then 8^2), filters and
 * samples it (mrh_wfpg_sample_fields with filter_alpha, the guide pdf at
 * the BxDF's direction for the balance heuristic), and weights the path by
 * f cos / (r p_guide + (1 - r) p_bxdf).  Each recorded path vertex takes
 * 64 B of device memory.  Single-level scenes only
 * (MRH_ERR_UNSUPPORTED otherwise).  NEE and MIS as in mrh_render_guided.
 * Differences from the reference: the bin origin is the bin's ray of the
 * smallest path id, not a random one (so the render does not depend on the
 * order of a bounce's rays); the payload is formed from the last learning
 * frame's vertices.  With product = 1 (the reference's default, productPG
 * with purePG off: KCGenAndSampleDistributionProduct and
 * ProductSampler::SampleMIS, WFPGTracerKC.cuh:1562-1766) the guide's half of
 * the mixture samples the product of the bin's field and the ray's BxDF:
 * one wave per ray evaluates the BxDF's luminance (f cos) at the 8x8
 * cell-centre directions tilted toward the shading normal by 4 pi / 64
 * radians (the reference's BxDFProjFunc, :1604-1629), the field stage's
 * product sampler draws the guide direction and gives the product pdf at
 * the BxDF's direction, and the weight is f cos / (r p_product + (1 - r)
 * p_bxdf).  The product sampler's draws are hashed by the ray's place in the
 * bounce, so product renders are not reproducible path for path. */
typedef struct {
    uint32_t iterations;        /* frames, 1..64 (1 = plain path tracing) */
    uint32_t svo_depth;         /* "OctreeLevel", 1..10 */
    float    guide_ratio;       /* "MISRatio" */
    uint32_t min_bin_level;     /* "MinRayBinLevel", 0..10 */
    uint32_t bin_ray_count;     /* "BinRayCount", 1..MRH_BIN_RAY_COUNT_MAX */
    float    filter_alpha;      /* "RFieldFilterAlpha", > 0 */
    float    boundary_radiance; /* radiance of cones that leave the octree, >= 0 */
    uint32_t product;           /* "ProductPG": 1 = the guide samples the product of the field and
                                   the BxDF (the field stage's product mode), 0 = the field alone;
                                   either way mixed with the BxDF by guide_ratio ("PurePG" =
                                   guide_ratio 1) */
} mrh_wfpg_params;
typedef struct {
    uint64_t bins;              /* bins over every guided bounce */
    uint64_t binned_rays;       /* rays sampled from a field, over every guided bounce */
    uint32_t guided_bounces;    /* bounces whose fields were traced */
    uint32_t octree_nodes;      /* nodes of the voxel octree */
    uint32_t voxel_records;     /* voxel records of the scene's triangles */
} mrh_wfpg_stats;
int32_t mrh_render_wfpg(mrh_context* ctx, mrh_scene* scene, const mrh_camera_desc* cam,
                        const mrh_render_params* params, const mrh_wfpg_params* wfpg, float* d_image,
                        mrh_render_stats* stats, mrh_wfpg_stats* wstats);

/* Autotuner (no counterpart in the reference): coordinate descent over the
 * launch shape of frame renders — traversal tree width, traversal stack
 * entries in LDS, blocks per CU of the traversal grid, traversal block size,
 * rays per partition tile, each axis swept with the others at their best so
 * far.  Each candidate renders the given frame `reps` times (at least 3) and
 * scores the median; the lowest is left set on the context (mrh_tune) and
 * every trial is listed.  d_image is scratch for the trial frames. */
typedef struct {
    uint32_t best_lds, best_bpc;
    double   best_ms;
    uint32_t trials;
    uint32_t trial_lds[32], trial_bpc[32];
    double   trial_ms[32];
    uint32_t best_width;
    uint32_t trial_width[32];
    uint32_t best_block, best_tile;
    uint32_t trial_block[32], trial_tile[32];
    uint32_t reps;               /* renders per candidate actually used */
} mrh_tune_result;
int32_t mrh_autotune(mrh_context* ctx, mrh_scene* scene, const mrh_camera_desc* cam,
                     const mrh_render_params* params, float* d_image, uint32_t reps,
                     mrh_tune_result* out);

/* Incremental tracer over a resident scene, the GPUTracerI protocol
 * (RayLib/GPUTracerI.h:30-72): mrh_tracer_create ~ GenerateTracer +
 * Initialize + SetOptions; mrh_tracer_generate_work ~ GenerateWork(camera)
 * (raygen of one frame); mrh_tracer_render ~ Render(): one bounce per call,
 * *more = 0 once MaxDepth bounces are done (PathTracer.cu:144-291);
 * mrh_tracer_finalize ~ Finalize(): writes the mean of every finished frame
 * to d_image (progressive accumulation; frame k uses global sample ids
 * k*W*H*spp + i); mrh_tracer_reset_image ~ ResetImage(). */
int32_t mrh_tracer_create(mrh_context* ctx, mrh_scene* scene, const mrh_render_params* params,
                          mrh_tracer** out);
int32_t mrh_tracer_generate_work(mrh_tracer* tracer, const mrh_camera_desc* cam);
int32_t mrh_tracer_render(mrh_tracer* tracer, int32_t* more);
int32_t mrh_tracer_finalize(mrh_tracer* tracer, float* d_image);
int32_t mrh_tracer_reset_image(mrh_tracer* tracer);
int32_t mrh_tracer_stats(const mrh_tracer* tracer, mrh_render_stats* out, uint32_t* frames);
/* Turns the tracer into the guided tracer (WFPGTracer, WFPGTracer.h:25-96,
 * WFPGTracer.cu:1590-1790): from the next frame on, frame k (counted from
 * this call) records its path vertices while k + 1 < guide->iterations and
 * the octree is rebuilt from them when the frame ends; every frame after
 * the first samples from the mixture of the latest guide and the BxDF.
 * `iterations` frames reproduce mrh_render_guided.  svo may be NULL. */
int32_t mrh_tracer_set_guide(mrh_tracer* tracer, const mrh_guide_params* guide, mrh_svo* svo);
/* mrh_tracer_set_guide with the reference's own guide, the field-traced one
 * of mrh_render_wfpg (WFPGTracer, WFPGTracer.cu:590-760, 1590-1790): the
 * scene is voxelised now; frame k (counted from this call) records its path
 * vertices while k + 1 < wfpg->iterations and forms the octree's payload
 * from them when it ends; every frame after the first guides its
 * non-specular hits with per-bin cone-traced fields.  `iterations` frames
 * reproduce mrh_render_wfpg.  Replaces any guide set before. */
int32_t mrh_tracer_set_wfpg(mrh_tracer* tracer, const mrh_wfpg_params* wfpg);
/* GPUTracerI::AttachReconFilter (RayLib/GPUTracerI.h:69): frames from the
 * next GenerateWork on are filtered (filter NULL: per-pixel mean again);
 * resets the accumulated image. */
int32_t mrh_tracer_set_filter(mrh_tracer* tracer, const mrh_filter_desc* filter);
int32_t mrh_tracer_destroy(mrh_tracer* tracer);

/* ------------------------------------------------------------------ */
/* Synthetic workloads (host only; no device needed).                   */
/* ------------------------------------------------------------------ */
/* Random triangles: centroids U[0,1)^3, corners offset by U(-0.01,0.01)^3;
 * rays (7 planes) with origins U[0,1)^3 and directions uniform on S^2,
 * tmax = FLT_MAX.  Either output may be NULL. */
int32_t mrh_gen_random_tris(uint32_t n, uint32_t seed, float* tris);
int32_t mrh_gen_random_rays(uint32_t n, uint32_t seed, float* rays);

/* ------------------------------------------------------------------ */
/* Serving engine (no counterpart in the reference, whose TracerThread
 * renders one scene with one tracer at a time, TracerThread.cpp,
 * PathTracer.cu:144-291).  Requests for resident scenes are queued,
 * admitted into <= 64 slots while the HBM budget allows, and traced
 * together: one wavefront per step over a shared queue of paths from every
 * admitted request (continuous batching), with the per-step launches
 * replayed from hipGraphs.  Each request's paths are the ones mrh_render /
 * mrh_render_guided trace for it (same streams, same arithmetic), so its
 * image matches a solo render up to the order of the per-pixel float adds.
 * The server works on a private context (mrh_server_context) on the same
 * device; the scenes and images of its requests must stay alive until the
 * request completes.                                                   */
/* ------------------------------------------------------------------ */
typedef struct mrh_server mrh_server;

typedef struct {
    uint64_t hbm_budget;     /* device bytes the server may hold: its queues (from creation on),
                                the scenes of admitted requests (each once) and every admitted
                                request's image, learn records and guide */
    uint32_t wave_paths;     /* paths in flight per step; each may also have a
                                next-event ray in flight (queues hold 2x rays) */
    uint32_t max_slots;      /* concurrent requests, 1..64 */
    uint32_t use_graphs;     /* replay each step's launches from a captured hipGraph */
    uint32_t timeout_ms;     /* 0 = none; otherwise a request not done this long after its
                                submission ends with MRH_ERR_TIMEOUT */
    uint32_t policy;         /* order in which admitted requests inject new paths while the
                                queue has room: MRH_SERVE_SMALLEST_FIRST (0, least remaining
                                work first: lowest mean latency) or MRH_SERVE_ARRIVAL (1, in
                                submission order: the oldest request goes first) */
} mrh_server_params;
#define MRH_SERVE_SMALLEST_FIRST 0u
#define MRH_SERVE_ARRIVAL        1u

typedef struct {
    mrh_scene*        scene;
    mrh_camera_desc   camera;
    mrh_render_params params;   /* samples = strata per axis of each iteration */
    mrh_guide_params  guide;    /* iterations <= 1: plain path tracing */
    float*            d_image;  /* width*height*3 floats, written when the request completes */
    uint64_t          user_tag;
} mrh_request;

typedef struct {
    uint64_t ticket, user_tag;
    int32_t  status;            /* MRH_OK or the error that ended the request */
    uint32_t reserved;
    uint64_t segments;          /* rays traced for the request */
    double   submit_s, admit_s, done_s;   /* CLOCK_MONOTONIC seconds */
} mrh_request_done;

typedef struct {
    uint64_t steps, segments, budget_used;
    uint32_t active, queued;
    uint32_t live_paths;        /* rays in flight: paths and their next-event rays */
    uint32_t reserved;
    uint64_t fixed_bytes;       /* of budget_used: queues and step buffers */
    uint64_t scene_bytes;       /* of budget_used: scenes of admitted requests */
} mrh_server_info;

int32_t mrh_server_create(mrh_context* ctx, const mrh_server_params* params, mrh_server** out);
int32_t mrh_server_context(mrh_server* server, mrh_context** out);
/* Queues a request (MRH_ERR_BUDGET if it can never fit the budget).  A
 * request that fails on its own (its allocations, its learn pass) ends alone
 * with that status in mrh_server_poll; the others go on. */
int32_t mrh_server_submit(mrh_server* server, const mrh_request* request, uint64_t* ticket);
/* Cancels a queued or admitted request: a queued one is reported at once, an
 * admitted one takes no new samples, its paths in flight end at their next
 * shading, and it is reported when the last has ended (a step or two later).
 * Either way mrh_server_poll returns it with MRH_ERR_CANCELLED and its image
 * is undefined.  MRH_ERR_INVALID for a ticket that is neither. */
int32_t mrh_server_cancel(mrh_server* server, uint64_t ticket);
/* Admits what fits, traces one wavefront step, finishes what ended
 * (synchronous).  active / queued may be NULL. */
int32_t mrh_server_step(mrh_server* server, uint32_t* active, uint32_t* queued);
/* Moves up to max completed requests into out (oldest first). */
int32_t mrh_server_poll(mrh_server* server, mrh_request_done* out, uint32_t max, uint32_t* n);
int32_t mrh_server_stats(const mrh_server* server, mrh_server_info* out);
int32_t mrh_server_destroy(mrh_server* server);

/* Scene generators fill caller buffers sized by a first call with NULL
 * buffers (counts written to the out-parameters). */
typedef struct {
    float*    tris;
    uint32_t* prim_key;
    float*    lambert;
    float*    metal;
    float*    dielectric;
    float*    emissive;
    uint32_t  n, n_lambert, n_metal, n_dielectric, n_emissive;
    mrh_camera_desc camera;
} mrh_scene_buffers;

/* Closed box of 12 triangles (walls albedo 0.75/0.25), icosphere of
 * 20*4^subdiv triangles (radius 0.3, centred) and a 2-triangle ceiling
 * light of radiance 15. */
int32_t mrh_gen_cornell(uint32_t subdiv, mrh_scene_buffers* out);
/* grid x grid displaced UV spheres (stacks x slices each), materials drawn
 * from {Lambert, GGX conductor, dielectric} with 5% emissive spheres. */
int32_t mrh_gen_sphere_grid(uint32_t grid, uint32_t stacks, uint32_t slices, uint32_t seed,
                            mrh_scene_buffers* out);

#ifdef __cplusplus
}
#endif
#endif
