/*
 * om_guide.c — oracle: the path-guiding sparse voxel octree and its
 * product sampler.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Structure (the reference builds its SVO by sort + run-length + per-level
 * unique prefixes, AnisoSVO.cu:657-1188): the leaves are the distinct cells
 * of the input codes; level l holds the distinct prefixes code >> 3(depth-l);
 * nodes are numbered root first, level by level, ascending code within a
 * level; a node's children are contiguous, so first_child + rank among the
 * set bits of child_mask addresses any child.
 * Splat (KCAccumulateRadianceToLeaf, AnisoSVO.cu:225-263): every sample lands
 * in the deepest existing node on its cell's path, bin = direction bin, sums
 * in 48.16 fixed point plus a count.
 * Filter (NormalizeAndFilterRadiance, AnisoSVO.cu:1415-1483): sums and counts
 * move up (each parent adds its children, in child order), bins become means,
 * empty bins take the parent's mean top-down, then a clamped 3x3 box blur of
 * the 8x8 bin grid and a 1e-4 floor.
 * Product sampler (the role of KCGenAndSampleDistributionProduct,
 * WFPGTracerKC.cuh:1562-1766): bin weight = radiance x max(n . centre, 0),
 * pick a bin by a linear scan of the running sum, a uniform point in it.
 */
#include "oracle.h"
#include "om.h"

#include <stdlib.h>
#include <string.h>

static int cmp_u32(const void* a, const void* b)
{
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return (x > y) - (x < y);
}

void om_svo_free(om_svo* s)
{
    free(s->codes);
    free(s->first_child);
    free(s->child_mask);
    free(s->parent);
    free(s->acc);
    free(s->count);
    free(s->sum);
    free(s->radiance);
    free(s->node_count);
    memset(s, 0, sizeof(*s));
}

int om_svo_build(om_svo* s, const uint32_t* codes, uint32_t n, const float cube[4], uint32_t depth)
{
    memset(s, 0, sizeof(*s));
    if (n == 0 || depth < 1 || depth > 10) return -1;
    s->depth = depth;
    for (int i = 0; i < 4; i++) s->cube[i] = cube[i];
    uint32_t* leaf = (uint32_t*)malloc((size_t)n * 4);
    memcpy(leaf, codes, (size_t)n * 4);
    qsort(leaf, n, 4, cmp_u32);
    /* distinct prefixes per level */
    uint32_t* lvl[11] = {0};
    uint32_t cnt[11] = {0};
    for (int l = (int)depth; l >= 0; l--) {
        uint32_t shift = 3u * (depth - (uint32_t)l);
        lvl[l] = (uint32_t*)malloc((size_t)n * 4);
        uint32_t k = 0;
        for (uint32_t i = 0; i < n; i++) {
            uint32_t c = leaf[i] >> shift;
            if (k == 0 || lvl[l][k - 1] != c) lvl[l][k++] = c;
        }
        cnt[l] = k;
    }
    free(leaf);
    s->level_offset[0] = 0;
    for (uint32_t l = 0; l <= depth; l++) {
        s->level_count[l] = cnt[l];
        s->level_offset[l + 1] = s->level_offset[l] + cnt[l];
    }
    for (uint32_t l = depth + 1; l < 11; l++) s->level_offset[l + 1] = s->level_offset[l];
    uint32_t N = s->level_offset[depth + 1];
    s->nodes = N;
    s->codes = (uint32_t*)malloc((size_t)N * 4);
    s->first_child = (uint32_t*)malloc((size_t)N * 4);
    s->child_mask = (uint32_t*)calloc(N, 4);
    s->parent = (uint32_t*)malloc((size_t)N * 4);
    s->acc = (uint64_t*)calloc((size_t)N * OM_GUIDE_BINS, 8);
    s->count = (uint32_t*)calloc((size_t)N * OM_GUIDE_BINS, 4);
    s->sum = (float*)calloc((size_t)N * OM_GUIDE_BINS, 4);
    s->radiance = (float*)malloc((size_t)N * OM_GUIDE_BINS * 4);
    s->node_count = (uint32_t*)calloc(N, 4);
    for (size_t i = 0; i < (size_t)N * OM_GUIDE_BINS; i++) s->radiance[i] = OM_GUIDE_EPS;
    for (uint32_t l = 0; l <= depth; l++) {
        uint32_t off = s->level_offset[l];
        uint32_t poff = l ? s->level_offset[l - 1] : 0u;
        uint32_t pi = 0;   /* parent cursor: parents are the distinct prefixes, in order */
        for (uint32_t i = 0; i < cnt[l]; i++) {
            uint32_t g = off + i, c = lvl[l][i];
            s->codes[g] = c;
            s->first_child[g] = OM_SVO_NO_NODE;
            if (l == 0) {
                s->parent[g] = OM_SVO_NO_NODE;
                continue;
            }
            while (lvl[l - 1][pi] != (c >> 3)) pi++;
            uint32_t p = poff + pi;
            s->parent[g] = p;
            s->child_mask[p] |= 1u << (c & 7u);
            if (s->first_child[p] == OM_SVO_NO_NODE) s->first_child[p] = g;
        }
    }
    for (uint32_t l = 0; l <= depth; l++) free(lvl[l]);
    for (uint32_t b = 0; b < OM_GUIDE_BINS; b++) {
        om_v3 c = om_octa_to_dir_v(((float)(b % OM_GUIDE_RES) + 0.5f) / (float)OM_GUIDE_RES,
                                   ((float)(b / OM_GUIDE_RES) + 0.5f) / (float)OM_GUIDE_RES);
        s->centers[b][0] = c.x;
        s->centers[b][1] = c.y;
        s->centers[b][2] = c.z;
    }
    return s->level_count[0] == 1 ? 0 : -2;
}

uint32_t om_svo_lookup(const om_svo* s, uint32_t code)
{
    uint32_t node = 0;
    for (uint32_t l = 0; l < s->depth; l++) {
        uint32_t m = s->child_mask[node];
        uint32_t oct = (code >> (3u * (s->depth - 1u - l))) & 7u;
        if (!((m >> oct) & 1u)) break;
        node = s->first_child[node] + (uint32_t)__builtin_popcount(m & ((1u << oct) - 1u));
    }
    return node;
}

/* Guide lookup: the deepest node on the path whose filtered data rests on at
 * least min_count samples (the role of the reference's collapse of sparse
 * bins to their ancestors, AnisoSVO.cu:1485-1529, BinRayCount).  The root is
 * returned when no child qualifies. */
uint32_t om_svo_lookup_guide(const om_svo* s, uint32_t code, uint32_t min_count)
{
    uint32_t node = 0;
    for (uint32_t l = 0; l < s->depth; l++) {
        uint32_t m = s->child_mask[node];
        uint32_t oct = (code >> (3u * (s->depth - 1u - l))) & 7u;
        if (!((m >> oct) & 1u)) break;
        uint32_t child = s->first_child[node] + (uint32_t)__builtin_popcount(m & ((1u << oct) - 1u));
        if (s->node_count[child] < min_count) break;
        node = child;
    }
    return node;
}

/* Ray binning (WFPGTracer::AccumulateRayHitsToSVO, WFPGTracer.cu:488-588;
 * AnisoSVOctreeCPU::CollapseRayCounts + KCCollapseRayCounts(Leaf),
 * AnisoSVO.cu:265-350, 1485-1529; KCCheckReducedSVOBins +
 * AnisoSVOctreeGPU::FindMarkedBin, WFPGTracerKC.cuh:1055-1078,
 * AnisoSVO.cuh:1244-1262; RayCaster::PartitionRaysWRTCustomData).
 * Each ray starts at the deepest existing node of its cell (the reference's
 * leaf); the ray counts of the nodes are collapsed bottom-up from the leaf
 * level to min_level: a node at min_level is marked, a node with at least
 * min_rays rays is marked, any other node hands its rays to its parent.  A
 * ray's bin is the first marked node on the way up from its start node (its
 * start node when none is marked: a start node above min_level).  Rays with
 * code OM_SVO_NO_NODE get no bin.  order: ray ids stably sorted by bin,
 * binned rays first; part_offsets[k] / part_bins[k] the first sorted
 * position and node of bin run k, part_offsets[count] the end of the binned
 * rays. */
void om_svo_bin_rays(const om_svo* s, const uint32_t* codes, uint32_t n, uint32_t min_level, uint32_t min_rays,
                     uint32_t* bins, uint32_t* order, uint32_t* part_offsets, uint32_t* part_bins,
                     uint32_t* part_count)
{
    const uint32_t N = s->nodes;
    uint32_t* start = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t* cnt = (uint32_t*)calloc(N ? N : 1, sizeof(uint32_t));
    uint8_t* marked = (uint8_t*)calloc(N ? N : 1, 1);
    if (min_level > s->depth) min_level = s->depth;   /* WFPGTracer.cu:612-613 */
    /* start nodes and their ray counts (sort + run-length encode) */
    for (uint32_t i = 0; i < n; i++) {
        start[i] = (codes[i] == OM_SVO_NO_NODE) ? OM_SVO_NO_NODE : om_svo_lookup(s, codes[i]);
        if (start[i] != OM_SVO_NO_NODE) cnt[start[i]]++;
    }
    /* collapse, leaf level first */
    for (int32_t l = (int32_t)s->depth; l >= (int32_t)min_level; l--) {
        for (uint32_t k = 0; k < s->level_count[l]; k++) {
            const uint32_t v = s->level_offset[l] + k;
            if ((uint32_t)l == min_level) {
                marked[v] = 1;
                continue;
            }
            if (cnt[v] < min_rays) {
                if (s->parent[v] != OM_SVO_NO_NODE) cnt[s->parent[v]] += cnt[v];
            } else {
                marked[v] = 1;
            }
        }
    }
    /* each ray climbs to its first marked node */
    for (uint32_t i = 0; i < n; i++) {
        uint32_t v = start[i];
        if (v != OM_SVO_NO_NODE) {
            uint32_t u = v;
            while (!marked[u] && s->parent[u] != OM_SVO_NO_NODE) u = s->parent[u];
            if (!marked[u]) u = v;
            v = u;
        }
        bins[i] = v;
    }
    /* partition by bin: stable counting sort, binned rays first */
    if (order) {
        uint32_t* first = (uint32_t*)calloc((size_t)N + 2, sizeof(uint32_t));
        for (uint32_t i = 0; i < n; i++) first[(bins[i] == OM_SVO_NO_NODE ? N : bins[i]) + 1]++;
        for (uint32_t v = 0; v <= N; v++) first[v + 1] += first[v];
        uint32_t runs = 0;
        for (uint32_t v = 0; v < N; v++)
            if (first[v + 1] > first[v]) {
                if (part_offsets) part_offsets[runs] = first[v];
                if (part_bins) part_bins[runs] = v;
                runs++;
            }
        if (part_offsets) part_offsets[runs] = first[N];
        if (part_count) *part_count = runs;
        for (uint32_t i = 0; i < n; i++) order[first[bins[i] == OM_SVO_NO_NODE ? N : bins[i]]++] = i;
        free(first);
    }
    free(start);
    free(cnt);
    free(marked);
}

void om_svo_splat(om_svo* s, const uint32_t* codes, const uint32_t* bins, const float* values, uint32_t n)
{
    for (uint32_t i = 0; i < n; i++) {
        if (codes[i] == OM_SVO_NO_NODE) continue;
        uint32_t node = om_svo_lookup(s, codes[i]);
        size_t k = (size_t)node * OM_GUIDE_BINS + (bins[i] & (OM_GUIDE_BINS - 1u));
        s->acc[k] += om_guide_fix(values[i]);
        s->count[k] += 1u;
    }
}

void om_svo_filter(om_svo* s)
{
    size_t nb = (size_t)s->nodes * OM_GUIDE_BINS;
    for (size_t i = 0; i < nb; i++) s->sum[i] = (float)s->acc[i] * (1.0f / OM_GUIDE_FIX_ONE);
    /* bottom-up: a node adds its children's sums, in child order */
    for (int l = (int)s->depth - 1; l >= 0; l--) {
        for (uint32_t i = 0; i < s->level_count[l]; i++) {
            uint32_t node = s->level_offset[l] + i;
            uint32_t m = s->child_mask[node];
            if (!m) continue;
            uint32_t fc = s->first_child[node], k = (uint32_t)__builtin_popcount(m);
            for (uint32_t b = 0; b < OM_GUIDE_BINS; b++) {
                float sm = s->sum[(size_t)node * OM_GUIDE_BINS + b];
                uint32_t c = s->count[(size_t)node * OM_GUIDE_BINS + b];
                for (uint32_t j = 0; j < k; j++) {
                    sm += s->sum[(size_t)(fc + j) * OM_GUIDE_BINS + b];
                    c += s->count[(size_t)(fc + j) * OM_GUIDE_BINS + b];
                }
                s->sum[(size_t)node * OM_GUIDE_BINS + b] = sm;
                s->count[(size_t)node * OM_GUIDE_BINS + b] = c;
            }
        }
    }
    /* samples under each node */
    for (uint32_t node = 0; node < s->nodes; node++) {
        uint32_t c = 0;
        for (uint32_t b = 0; b < OM_GUIDE_BINS; b++) c += s->count[(size_t)node * OM_GUIDE_BINS + b];
        s->node_count[node] = c;
    }
    /* top-down: means; empty bins inherit the parent's mean (root: 0) */
    for (uint32_t l = 0; l <= s->depth; l++) {
        for (uint32_t i = 0; i < s->level_count[l]; i++) {
            uint32_t node = s->level_offset[l] + i;
            for (uint32_t b = 0; b < OM_GUIDE_BINS; b++) {
                size_t k = (size_t)node * OM_GUIDE_BINS + b;
                uint32_t c = s->count[k];
                if (c) {
                    s->sum[k] = s->sum[k] / (float)c;
                } else {
                    uint32_t p = s->parent[node];
                    s->sum[k] = (p == OM_SVO_NO_NODE) ? 0.0f : s->sum[(size_t)p * OM_GUIDE_BINS + b];
                }
            }
        }
    }
    /* 3x3 box blur over the bin grid, clamped at the borders, + floor */
    for (uint32_t node = 0; node < s->nodes; node++) {
        for (uint32_t b = 0; b < OM_GUIDE_BINS; b++) {
            int ix = (int)(b % OM_GUIDE_RES), iy = (int)(b / OM_GUIDE_RES);
            float sm = 0.0f;
            int k = 0;
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    int x = ix + dx, y = iy + dy;
                    if (x < 0 || y < 0 || x >= (int)OM_GUIDE_RES || y >= (int)OM_GUIDE_RES) continue;
                    sm += s->sum[(size_t)node * OM_GUIDE_BINS + (size_t)(y * (int)OM_GUIDE_RES + x)];
                    k++;
                }
            s->radiance[(size_t)node * OM_GUIDE_BINS + b] = sm / (float)k + OM_GUIDE_EPS;
        }
    }
}

static float guide_weight(const om_svo* s, const float* rad, uint32_t b, om_v3 n)
{
    om_v3 c = v3(s->centers[b][0], s->centers[b][1], s->centers[b][2]);
    return rad[b] * fmaxf(v3_dot(n, c), 0.0f);
}

float om_guide_total_v(const om_svo* s, uint32_t node, om_v3 n)
{
    const float* rad = s->radiance + (size_t)node * OM_GUIDE_BINS;
    float t = 0.0f;
    for (uint32_t b = 0; b < OM_GUIDE_BINS; b++) t += guide_weight(s, rad, b, n);
    return t;
}

#define OM_GUIDE_BIN_SR_INV ((float)OM_GUIDE_BINS * 0.07957747154594767f)

om_v3 om_guide_sample_v(const om_svo* s, uint32_t node, om_v3 n, float total, float xi0, float xi1,
                        float xi2, float* pdf)
{
    const float* rad = s->radiance + (size_t)node * OM_GUIDE_BINS;
    if (!(total > 0.0f)) {
        *pdf = 0.0f;
        return v3(0.0f, 0.0f, 1.0f);
    }
    float target = xi0 * total;
    float acc = 0.0f, wp = 0.0f;
    uint32_t pick = 0;
    for (uint32_t b = 0; b < OM_GUIDE_BINS; b++) {
        float w = guide_weight(s, rad, b, n);
        if (w > 0.0f) {
            pick = b;
            wp = w;
            acc += w;
            if (acc > target) break;
        }
    }
    float u = ((float)(pick % OM_GUIDE_RES) + xi1) / (float)OM_GUIDE_RES;
    float v = ((float)(pick / OM_GUIDE_RES) + xi2) / (float)OM_GUIDE_RES;
    *pdf = wp / total * OM_GUIDE_BIN_SR_INV;
    return om_octa_to_dir_v(u, v);
}

float om_guide_pdf_v(const om_svo* s, uint32_t node, om_v3 n, float total, om_v3 d)
{
    if (!(total > 0.0f)) return 0.0f;
    float w = guide_weight(s, s->radiance + (size_t)node * OM_GUIDE_BINS, om_dir_to_bin_v(d), n);
    return w / total * OM_GUIDE_BIN_SR_INV;
}

/* ---------------------------- array entry points ----------------------- */

void om_svo_codes(const float* pts, uint32_t n, const float cube[4], uint32_t depth, uint32_t* codes)
{
    float inv_cell = (float)(1u << depth) / cube[3];
    for (uint32_t i = 0; i < n; i++)
        codes[i] = om_svo_code_v(v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]), cube[0], cube[1], cube[2],
                                 inv_cell, depth);
}

void om_dir_to_bins(const float* dirs, uint32_t n, uint32_t* bins)
{
    for (uint32_t i = 0; i < n; i++) bins[i] = om_dir_to_bin_v(v3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
}

void om_octa_to_dirs(const float* uv, uint32_t n, float* dirs)
{
    for (uint32_t i = 0; i < n; i++) {
        om_v3 d = om_octa_to_dir_v(uv[2 * i], uv[2 * i + 1]);
        dirs[3 * i] = d.x;
        dirs[3 * i + 1] = d.y;
        dirs[3 * i + 2] = d.z;
    }
}

void om_dir_to_octas(const float* dirs, uint32_t n, float* uv)
{
    for (uint32_t i = 0; i < n; i++)
        om_dir_to_octa_v(v3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]), &uv[2 * i], &uv[2 * i + 1]);
}

float om_atan_unit(float t) { return om_atan_unit_f(t); }

void om_guide_sample(const om_svo* s, const uint32_t* nodes, const float* normals, const float* xi, uint32_t n,
                     float* dirs, float* pdf)
{
    for (uint32_t i = 0; i < n; i++) {
        om_v3 nn = v3(normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]);
        float t = om_guide_total_v(s, nodes[i], nn);
        om_v3 d = om_guide_sample_v(s, nodes[i], nn, t, xi[3 * i], xi[3 * i + 1], xi[3 * i + 2], &pdf[i]);
        dirs[3 * i] = d.x;
        dirs[3 * i + 1] = d.y;
        dirs[3 * i + 2] = d.z;
    }
}

void om_guide_pdf(const om_svo* s, const uint32_t* nodes, const float* normals, const float* dirs, uint32_t n,
                  float* pdf)
{
    for (uint32_t i = 0; i < n; i++) {
        om_v3 nn = v3(normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]);
        float t = om_guide_total_v(s, nodes[i], nn);
        pdf[i] = om_guide_pdf_v(s, nodes[i], nn, t, v3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
    }
}
