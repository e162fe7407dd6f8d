// mrh_filter.hip — reconstruction filters (GPUReconFilter::FilterToImgInternal,
// GPUReconFilter.h:59-260, and its kernels, GPUReconFilterKC.cuh).
//
// The reference scatters: every sample writes (pixel, sample) pairs for the
// pixels of its window, the pairs are radix-sorted by pixel, split into
// segments and each segment reduced by a block or a warp.  Here the work is
// a gather: samples are grouped by the pixel they lie in (the path tracer's
// samples already are; arbitrary samples are sorted once by pixel), and one
// thread per output pixel walks the samples of the neighbouring pixels that
// can reach it, applies the reference's acceptance test (window of
// FilterRadiusToPixelWH + centre within the radius) and sums w * value and
// w in a fixed order.  No pair buffer of pixels-per-sample x samples entries
// and no atomics; each sample's value is read by the few pixels around it,
// out of the L2.
#include "mrh_filter.hpp"

#include <cmath>

namespace mrh {
namespace {

// FilterRadiusToPixelWH (GPUReconFilterKC.cuh:8-27): side of the square of
// pixels a sample can reach; every further 0.5 of radius adds a pixel.
int32_t radius_to_pixel_wh(float r)
{
    uint32_t result = 1;
    if (r == 0.0f) return (int32_t)result;
    uint32_t quot = (uint32_t)(r / 0.5f);
    const float rem = fmodf(r, 0.5f);
    if (rem == 0.0f) quot -= 1;
    result += quot + 1;
    return (int32_t)result;
}

// Weight of a sample at |pixel centre - sample| = (dx, dy): the operator()
// of GPUBoxFilterFunctor, GPUTentFilterFunctor, GPUGaussianFilterFunctor
// and GPUMitchellFilterFunctor (GPUReconFilter{Box,Tent,Gaussian,Mitchell}.h).
template <uint32_t T>
__device__ __forceinline__ float filter_weight(const FilterDev& f, float dx, float dy)
{
    if constexpr (T == MRH_FILTER_BOX) {
        if (f.radius == 0.0f) return 1.0f;
        return (dx < f.radius && dy < f.radius) ? 1.0f : 0.0f;
    } else if constexpr (T == MRH_FILTER_TENT) {
        float lx = fmaxf(0.0f, f.radius - dx);
        float ly = fmaxf(0.0f, f.radius - dy);
        lx *= f.recip;
        ly *= f.recip;
        return lx * ly;
    } else if constexpr (T == MRH_FILTER_GAUSSIAN) {
        const float gx = fmaxf(0.0f, expf(f.neg_alpha * dx * dx) - f.gexp);
        const float gy = fmaxf(0.0f, expf(f.neg_alpha * dy * dy) - f.gexp);
        return gx * gy;
    } else {
        auto m1 = [&](float x) {
            x = fabsf(x);
            const float x2 = x * x;
            const float x3 = x2 * x;
            if (!(x < 2.0f)) return 0.0f;
            const bool in = x < 1.0f;
            const float c0 = in ? f.m01[0] : f.m12[0], c1 = in ? f.m01[1] : f.m12[1];
            const float c2 = in ? f.m01[2] : f.m12[2], c3 = in ? f.m01[3] : f.m12[3];
            return c0 * x3 + c1 * x2 + c2 * x + c3;
        };
        return m1(2.0f * dx * f.recip) * m1(2.0f * dy * f.recip);
    }
}

// One thread per output pixel (see the file comment).  For pixel p the
// accepted samples are those whose window (KCExpandSamplesToPixels,
// GPUReconFilterKC.cuh:29-128: [lo, hi) around the sample's pixel, shifted
// by -1 for even widths when the sample sits in the lower half of its
// pixel) holds p and, unless the window is one pixel, whose distance to p's
// centre is at most the radius.
template <uint32_t T>
__global__ __launch_bounds__(256) void k_filter_gather(FilterDev f, const float* __restrict__ values,
                                                       const float2* __restrict__ coords,
                                                       const uint32_t* __restrict__ perm,
                                                       const uint32_t* __restrict__ offsets, uint32_t spp,
                                                       uint32_t w, uint32_t h, float* __restrict__ weighted,
                                                       float* __restrict__ weights)
{
    const uint32_t n = w * h;
    // offsets between p and the pixel a sample lies in that can be accepted
    const int32_t dmin = (f.range & 1) ? f.lo : f.lo - 1;
    const int32_t dmax = f.hi - 1;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const int32_t px = (int32_t)(p % w), py = (int32_t)(p / w);
        const float pcx = (float)px + 0.5f, pcy = (float)py + 0.5f;
        float ar = 0.0f, ag = 0.0f, ab = 0.0f, ws = 0.0f;
        for (int32_t sy = py - dmax; sy <= py - dmin; sy++) {
            if (sy < 0 || sy >= (int32_t)h) continue;
            for (int32_t sx = px - dmax; sx <= px - dmin; sx++) {
                if (sx < 0 || sx >= (int32_t)w) continue;
                const uint32_t q = (uint32_t)sy * w + (uint32_t)sx;
                const uint32_t b = perm ? offsets[q] : q * spp;
                const uint32_t e = perm ? offsets[q + 1] : b + spp;
                for (uint32_t k = b; k < e; k++) {
                    const uint32_t s = perm ? perm[k] : k;
                    const float2 c = coords[s];
                    const float ix = truncf(c.x), iy = truncf(c.y);
                    const float rx = c.x - ix, ry = c.y - iy;
                    const int32_t dxi = px - (int32_t)ix, dyi = py - (int32_t)iy;
                    int32_t lox = f.lo, hix = f.hi, loy = f.lo, hiy = f.hi;
                    if ((f.range & 1) == 0) {
                        if (rx < 0.5f) { lox--; hix--; }
                        if (ry < 0.5f) { loy--; hiy--; }
                    }
                    if (dxi < lox || dxi >= hix || dyi < loy || dyi >= hiy) continue;
                    if (f.range != 1) {
                        const float vx = c.x - ((ix + (float)dxi) + 0.5f);
                        const float vy = c.y - ((iy + (float)dyi) + 0.5f);
                        float lsq = 0.0f;
                        lsq += vx * vx;
                        lsq += vy * vy;
                        if (!(lsq <= f.rsq)) continue;
                    }
                    const float wgt = filter_weight<T>(f, fabsf(pcx - c.x), fabsf(pcy - c.y)) * 1.0f;
                    ar += wgt * values[3 * (size_t)s + 0];
                    ag += wgt * values[3 * (size_t)s + 1];
                    ab += wgt * values[3 * (size_t)s + 2];
                    ws += wgt;
                }
            }
        }
        weighted[3 * (size_t)p + 0] += ar;
        weighted[3 * (size_t)p + 1] += ag;
        weighted[3 * (size_t)p + 2] += ab;
        weights[p] += ws;
    }
}

// pixel key of every sample (w * h for samples outside the image)
__global__ void k_filter_keys(const float2* __restrict__ coords, uint32_t n, uint32_t w, uint32_t h,
                              uint32_t* __restrict__ keys)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float2 c = coords[i];
        uint32_t k = w * h;
        if (c.x >= 0.0f && c.x < (float)w && c.y >= 0.0f && c.y < (float)h)
            k = (uint32_t)truncf(c.y) * w + (uint32_t)truncf(c.x);
        keys[i] = k;
    }
}

// offsets[p] = first sorted index whose key is >= p, for p in [0, w*h]
__global__ void k_filter_offsets(const uint32_t* __restrict__ sorted, uint32_t n, uint32_t pixels,
                                 uint32_t* __restrict__ offsets)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += gridDim.x * blockDim.x) {
        const uint32_t first = (i == 0) ? 0u : sorted[i - 1] + 1u;
        const uint32_t last = (i == n) ? pixels : min(sorted[i], pixels);
        for (uint32_t p = first; p <= last; p++) offsets[p] = i;
    }
}

__global__ void k_filter_resolve(const float* __restrict__ weighted, const float* __restrict__ weights,
                                 size_t pixels, float* __restrict__ out)
{
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < pixels; p += (size_t)gridDim.x * blockDim.x) {
        const float ws = weights[p];
        for (int c = 0; c < 3; c++) out[3 * p + c] = (ws == 0.0f) ? 0.0f : weighted[3 * p + c] / ws;
    }
}

}  // namespace

int32_t filter_make(mrh_context* ctx, const mrh_filter_desc* d, FilterDev* out)
{
    FilterDev f = {};
    f.type = d->type;
    if (d->type > MRH_FILTER_MITCHELL) return fail(ctx, MRH_ERR_INVALID, "filter: unknown type %u", d->type);
    if (!(d->radius >= 0.0f) || !std::isfinite(d->radius) || d->radius > 16.0f)
        return fail(ctx, MRH_ERR_INVALID, "filter: radius must be in [0, 16]");
    if (d->type != MRH_FILTER_BOX && d->radius == 0.0f)
        return fail(ctx, MRH_ERR_INVALID, "filter: Tent, Gaussian and Mitchell need a radius above 0");
    f.radius = d->radius;
    f.rsq = d->radius * d->radius;
    f.recip = (d->radius > 0.0f) ? 1.0f / d->radius : 0.0f;
    f.range = radius_to_pixel_wh(d->radius);
    f.lo = -(f.range - 1) / 2;
    f.hi = (f.range + 2) / 2;
    if (d->type == MRH_FILTER_GAUSSIAN) {
        if (!(d->alpha > 0.0f) || !std::isfinite(d->alpha))
            return fail(ctx, MRH_ERR_INVALID, "filter: Gaussian alpha must be positive");
        // GPUGaussianFilterFunctor(radius, alpha): exponent = Gauss(radius)
        f.neg_alpha = -d->alpha;
        f.gexp = expf(f.neg_alpha * d->radius * d->radius);
    }
    if (d->type == MRH_FILTER_MITCHELL) {
        if (!std::isfinite(d->b) || !std::isfinite(d->c))
            return fail(ctx, MRH_ERR_INVALID, "filter: Mitchell b and c must be finite");
        // GPUMitchellFilterFunctor(radius, b, c) (GPUReconFilterMitchell.h:70-86),
        // with the published constant term 6 - 2b of Mitchell & Netravali
        // 1988 on [0, 1); the reference's 6 + 2b makes the kernel jump at
        // |x| = 1
        const float t = 1.0f / 6.0f, b = d->b, c = d->c;
        f.m01[0] = t * (12 - 9 * b - 6 * c);
        f.m01[1] = t * (-18 + 12 * b + 6 * c);
        f.m01[2] = 0.0f;
        f.m01[3] = t * (6 - 2 * b);
        f.m12[0] = t * (-1 * b - 6 * c);
        f.m12[1] = t * (6 * b + 30 * c);
        f.m12[2] = t * (-12 * b - 48 * c);
        f.m12[3] = t * (8 * b + 24 * c);
    }
    *out = f;
    return MRH_OK;
}

int32_t filter_gather(mrh_context* ctx, const FilterDev& f, const float* values, const float2* coords,
                      const uint32_t* perm, const uint32_t* offsets, uint32_t spp, uint32_t w, uint32_t h,
                      float* weighted, float* weights)
{
    const uint32_t grid = stride_grid(ctx, (uint64_t)w * h, 256);
    switch (f.type) {
    case MRH_FILTER_BOX:
        return launch(ctx, "filter", dim3(grid), dim3(256), 0, k_filter_gather<MRH_FILTER_BOX>, f, values, coords,
                      perm, offsets, spp, w, h, weighted, weights);
    case MRH_FILTER_TENT:
        return launch(ctx, "filter", dim3(grid), dim3(256), 0, k_filter_gather<MRH_FILTER_TENT>, f, values, coords,
                      perm, offsets, spp, w, h, weighted, weights);
    case MRH_FILTER_GAUSSIAN:
        return launch(ctx, "filter", dim3(grid), dim3(256), 0, k_filter_gather<MRH_FILTER_GAUSSIAN>, f, values,
                      coords, perm, offsets, spp, w, h, weighted, weights);
    default:
        return launch(ctx, "filter", dim3(grid), dim3(256), 0, k_filter_gather<MRH_FILTER_MITCHELL>, f, values,
                      coords, perm, offsets, spp, w, h, weighted, weights);
    }
}

int32_t filter_resolve(mrh_context* ctx, const float* weighted, const float* weights, size_t pixels, float* out)
{
    return launch(ctx, "filter_resolve", dim3(stride_grid(ctx, pixels, 256)), dim3(256), 0, k_filter_resolve,
                  weighted, weights, pixels, out);
}

}  // namespace mrh

using namespace mrh;

extern "C" {

int32_t mrh_filter_samples(mrh_context* ctx, const mrh_filter_desc* filter, const float* d_values,
                           const float* d_coords, uint32_t n, uint32_t width, uint32_t height,
                           float* d_weighted, float* d_weights)
{
    if (!ctx || !filter || !d_weighted || !d_weights) return MRH_ERR_INVALID;
    if (n && (!d_values || !d_coords)) return fail(ctx, MRH_ERR_INVALID, "mrh_filter_samples: null samples");
    if (width == 0 || height == 0 || (uint64_t)width * height >= (1ull << 31))
        return fail(ctx, MRH_ERR_INVALID, "mrh_filter_samples: bad image size");
    if (n >= (1u << 31)) return fail(ctx, MRH_ERR_INVALID, "mrh_filter_samples: too many samples");
    FilterDev f;
    MRH_TRY(filter_make(ctx, filter, &f));
    const uint32_t pixels = width * height;
    // keys, sorted keys, sample order and pixel offsets
    const size_t bytes = ((size_t)n * 12 + (size_t)(pixels + 1) * 4 + 1024);
    void* mem = nullptr;
    MRH_TRY(scratch(ctx, SLOT_FILTER, bytes, &mem));
    uint32_t* keys = (uint32_t*)mem;
    uint32_t* sorted = keys + n;
    uint32_t* perm = sorted + n;
    uint32_t* offsets = perm + n;
    const float2* coords = (const float2*)d_coords;
    if (n) {
        MRH_TRY(launch(ctx, "filter_keys", dim3(stride_grid(ctx, n, 256)), dim3(256), 0, k_filter_keys, coords, n,
                       width, height, keys));
        uint32_t bits = 1;
        while (bits < 32 && (1ull << bits) <= pixels) bits++;
        SortArgs sa;
        sa.keys_in = keys;
        sa.keys_out = sorted;
        sa.vals_out = perm;
        sa.n_max = n;
        sa.key_bytes = 4;
        sa.begin_bit = 0;
        sa.end_bit = bits;
        MRH_TRY(radix_sort(ctx, sa));
    }
    MRH_TRY(launch(ctx, "filter_offsets", dim3(stride_grid(ctx, (uint64_t)n + 1, 256)), dim3(256), 0,
                   k_filter_offsets, (const uint32_t*)sorted, n, pixels, offsets));
    MRH_TRY(filter_gather(ctx, f, d_values, coords, perm, offsets, 0u, width, height, d_weighted, d_weights));
    MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return MRH_OK;
}

int32_t mrh_filter_resolve(mrh_context* ctx, const float* d_weighted, const float* d_weights, uint32_t width,
                           uint32_t height, float* d_image)
{
    if (!ctx || !d_weighted || !d_weights || !d_image) return MRH_ERR_INVALID;
    MRH_TRY(filter_resolve(ctx, d_weighted, d_weights, (size_t)width * height, d_image));
    MRH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return MRH_OK;
}

}  // extern "C"
