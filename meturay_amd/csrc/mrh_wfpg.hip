// mrh_wfpg.hip — per-bin radiance fields of the reference's wavefront path
// guiding: empty-texel raise, Gaussian filter with concentric-octahedral
// edge folding, 2D piecewise-constant distribution and one guided sample
// per ray (KCGenAndSampleDistribution, WFPGTracerKC.cuh:1433-1545).
//
// One 256-thread workgroup per bin (grid-stride over bins).  The field lives
// in LDS for the whole bin: rows padded to 65 floats so that the
// one-lane-per-row prefix walks hit 64 distinct banks.  Float operation
// order is fixed and launch-independent: taps accumulate with fma from -5 to
// 5, prefix sums run left to right, and every ray brings its own uniforms.
#include "mrh_guide.hpp"

#include <cmath>

namespace mrh {
namespace {

constexpr uint32_t WFPG_BLOCK = 256;
constexpr uint32_t WFPG_MAX = 64;           // texels along a field side
constexpr uint32_t WFPG_LS = WFPG_MAX + 1;  // LDS row stride
constexpr int WFPG_HALF = 5;                // 11 taps

struct WfpgTaps {
    float w[2 * WFPG_HALF + 1];
};

// StaticGaussianFilter1D<11> (BlockTextureFilter.cuh:84-121): g(t) =
// exp(-t^2 / (2 alpha)) / (sqrt(pi) sqrt(2) sqrt(alpha)), normalised by the
// float sum of the 11 samples.  Formed once on the host.
WfpgTaps gauss_taps(float alpha)
{
    const float W = 1.0f / 1.77245385090551602729f / 1.41421356237309504880f;
    float g[2 * WFPG_HALF + 1];
    for (int i = -WFPG_HALF; i <= WFPG_HALF; i++) {
        const float t = (float)i;
        g[i + WFPG_HALF] = W / std::sqrt(alpha) * std::exp(-(t * t) * 0.5f / alpha);
    }
    float total = 0.0f;
    for (int i = 0; i < 2 * WFPG_HALF + 1; i++) total += g[i];
    WfpgTaps taps;
    for (int i = 0; i < 2 * WFPG_HALF + 1; i++) taps.w[i] = g[i] / total;
    return taps;
}

// Utility::CocentricOctohedralWrapInt (CoordinateConversion.h:213-235):
// fold a texel index back into [0,X) x [0,Y); an odd fold count in exactly
// one axis mirrors both coordinates.
__device__ inline void octa_fold(int x, int y, int X, int Y, int& ox, int& oy)
{
    const int sx = (x < 0) ? -2 * X - x : x;
    const int sy = (y < 0) ? -2 * Y - y : y;
    const int ix = sx / X, iy = sy / Y;
    int fx = abs(sx % X), fy = abs(sy % Y);
    if ((ix & 1) ^ (iy & 1)) {
        fx = (X - 1) - fx;
        fy = (Y - 1) - fy;
    }
    ox = fx;
    oy = fy;
}

// Continuous index of xi in one axis of the distribution
// (BlockPWCDistribution2D::Sample with BinarySearchInBetween): c holds the
// cdf at the END of each texel.  Found interval k (c[k] <= xi < c[k+1]) ->
// k + frac + 1; below c[0] -> xi / c[0]; at or above the last entry ->
// size - 1e-5.
__device__ inline float pwc_axis(const float* c, int size, float xi)
{
    int lo = 0, hi = size - 2;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const float cm = c[mid], cn = c[mid + 1];
        if (cm <= xi && cn > xi) return ((float)mid + (xi - cm) / (cn - cm)) + 1.0f;
        if (xi < cm) hi = mid - 1;
        else lo = mid + 1;
    }
    if (xi < c[size - 1]) return xi / c[0];
    return (float)size - 1.0e-5f;
}

// pdf of the texel under (u,v) times 1/(4 pi) (solid angle of the map);
// coordinates outside [0,1) (or NaN) clamp to the border texels
__device__ inline float pwc_pdf(const float* s_pdf, const float* s_py, uint32_t X, uint32_t Y, float u, float v)
{
    const float fx = u * (float)X, fy = v * (float)Y;
    uint32_t ix = (fx > 0.0f) ? (uint32_t)fminf(fx, 65536.0f) : 0u;
    uint32_t iy = (fy > 0.0f) ? (uint32_t)fminf(fy, 65536.0f) : 0u;
    ix = ix >= X ? X - 1 : ix;
    iy = iy >= Y ? Y - 1 : iy;
    const float p = s_pdf[iy * WFPG_LS + ix] * s_py[iy];
    return p * (0.25f * 0.318309886183790671538f);
}

__global__ __launch_bounds__(WFPG_BLOCK) void k_wfpg_fields(
    const float* __restrict__ fields, uint32_t X, uint32_t Y, uint32_t lx, uint32_t bins, WfpgTaps taps,
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ ray_ids, uint32_t ray_count,
    const float2* __restrict__ xi, const float2* __restrict__ eval_uv, float2* __restrict__ out_uv,
    float* __restrict__ out_dir, float* __restrict__ out_pdf, float* __restrict__ out_eval_pdf,
    float* __restrict__ out_filtered)
{
    __shared__ float s_a[WFPG_MAX * WFPG_LS];  // field -> filtered field -> row pdfs
    __shared__ float s_b[WFPG_MAX * WFPG_LS];  // u-pass result -> row cdfs
    __shared__ float s_row[WFPG_MAX];          // row totals -> marginal pdf
    __shared__ float s_cy[WFPG_MAX];           // marginal cdf
    const uint32_t n = X * Y, tid = threadIdx.x, xm = X - 1;
    const int Xi = (int)X, Yi = (int)Y;
    const float dX = 1.0f / (float)X, dY = 1.0f / (float)Y;
    for (uint32_t bin = blockIdx.x; bin < bins; bin += gridDim.x) {
        const float* src = fields + (size_t)bin * n;
        // radiance query result: an empty texel keeps a small weight
        // (WFPGTracerKC.cuh:1316-1317)
        for (uint32_t p = tid; p < n; p += WFPG_BLOCK) {
            const float v = src[p];
            s_a[(p >> lx) * WFPG_LS + (p & xm)] = (v == 0.0f) ? 1.0e-4f : v;
        }
        __syncthreads();
        // filter along u (BlockTextureFilter.cuh:157-174)
        for (uint32_t p = tid; p < n; p += WFPG_BLOCK) {
            const int x = (int)(p & xm), y = (int)(p >> lx);
            float acc = 0.0f;
#pragma unroll
            for (int j = -WFPG_HALF; j <= WFPG_HALF; j++) {
                int qx = x + j, qy = y;
                if (qx < 0 || qx >= Xi) octa_fold(qx, y, Xi, Yi, qx, qy);
                acc = fmaf(s_a[qy * WFPG_LS + qx], taps.w[j + WFPG_HALF], acc);
            }
            s_b[y * WFPG_LS + x] = acc;
        }
        __syncthreads();
        // filter along v (:178-195)
        for (uint32_t p = tid; p < n; p += WFPG_BLOCK) {
            const int x = (int)(p & xm), y = (int)(p >> lx);
            float acc = 0.0f;
#pragma unroll
            for (int j = -WFPG_HALF; j <= WFPG_HALF; j++) {
                int qx = x, qy = y + j;
                if (qy < 0 || qy >= Yi) octa_fold(x, qy, Xi, Yi, qx, qy);
                acc = fmaf(s_b[qy * WFPG_LS + qx], taps.w[j + WFPG_HALF], acc);
            }
            s_a[y * WFPG_LS + x] = acc;
            if (out_filtered) out_filtered[(size_t)bin * n + p] = acc;
        }
        __syncthreads();
        // conditional distributions, one lane per row
        // (GPUBlockPWCDistribution.cuh:107-156)
        if (tid < Y) {
            float* ra = s_a + tid * WFPG_LS;
            float* rb = s_b + tid * WFPG_LS;
            float s = 0.0f;
            for (uint32_t x = 0; x < X; x++) {
                s += ra[x];
                rb[x] = s;
            }
            s_row[tid] = s;
            const float sdx = s * dX;
            if (sdx != 0.0f) {
                const float inv = 1.0f / sdx;
                for (uint32_t x = 0; x < X; x++) {
                    ra[x] = ra[x] * inv;
                    rb[x] = (rb[x] * dX) * inv;
                }
            } else {
                for (uint32_t x = 0; x < X; x++) rb[x] = rb[x] * dX;
            }
        }
        __syncthreads();
        // marginal over rows (:158-199)
        if (tid == 0) {
            float c = 0.0f;
            for (uint32_t y = 0; y < Y; y++) {
                c += s_row[y] * dY;
                s_cy[y] = c;
            }
            const float inv = (c != 0.0f) ? 1.0f / c : 1.0f;
            for (uint32_t y = 0; y < Y; y++) {
                float p = s_row[y] * dY, cc = s_cy[y];
                if (c != 0.0f) {
                    p = p * inv;
                    cc = cc * inv;
                }
                s_row[y] = p * (float)Y;
                s_cy[y] = cc;
            }
        }
        __syncthreads();
        // one guided sample per ray of the bin (WFPGTracerKC.cuh:1466-1533)
        const uint32_t lo = offsets[bin], hi = offsets[bin + 1];
        for (uint32_t k = lo + tid; k < hi; k += WFPG_BLOCK) {
            const uint32_t r = ray_ids[k];
            if (r >= ray_count) continue;
            const float2 x2 = xi[r];
            float u, v;
            if (s_cy[Y - 1] == 0.0f) {
                u = x2.x;
                v = x2.y;
            } else {
                const float iy = pwc_axis(s_cy, Yi, x2.y);
                int row = (int)iy;
                row = row >= Yi ? Yi - 1 : (row < 0 ? 0 : row);
                const float ix = pwc_axis(s_b + row * WFPG_LS, Xi, x2.x);
                u = ix * dX;
                v = iy * dY;
            }
            out_uv[r] = make_float2(u, v);
            out_pdf[r] = pwc_pdf(s_a, s_row, X, Y, u, v);
            if (out_dir) {
                const V3 d = octa_to_dir(u, v);   // z up; the path uses y up
                out_dir[3 * (size_t)r] = d.y;
                out_dir[3 * (size_t)r + 1] = d.z;
                out_dir[3 * (size_t)r + 2] = d.x;
            }
            if (eval_uv) {
                const float2 e = eval_uv[r];
                out_eval_pdf[r] = pwc_pdf(s_a, s_row, X, Y, e.x, e.y);
            }
        }
        __syncthreads();
    }
}

bool field_side_ok(uint32_t s) { return s == 8 || s == 16 || s == 32 || s == 64; }

}  // namespace
}  // namespace mrh

using namespace mrh;

int32_t mrh_wfpg_sample_fields(mrh_context* ctx, const mrh_wfpg_fields_desc* d)
{
    if (!ctx || !d) return MRH_ERR_INVALID;
    if (!field_side_ok(d->field_x) || !field_side_ok(d->field_y))
        return fail(ctx, MRH_ERR_INVALID, "mrh_wfpg_sample_fields: field sides must be 8, 16, 32 or 64");
    if (!(d->filter_alpha > 0.0f) || !std::isfinite(d->filter_alpha))
        return fail(ctx, MRH_ERR_INVALID, "mrh_wfpg_sample_fields: filter_alpha must be finite and > 0");
    if ((d->eval_uv == nullptr) != (d->out_eval_pdf == nullptr))
        return fail(ctx, MRH_ERR_INVALID, "mrh_wfpg_sample_fields: eval_uv and out_eval_pdf go together");
    if (d->bin_count == 0) return MRH_OK;
    if (!d->fields || !d->bin_offsets || !d->ray_ids)
        return fail(ctx, MRH_ERR_INVALID, "mrh_wfpg_sample_fields: null fields, bin_offsets or ray_ids");
    if (d->ray_count && (!d->xi || !d->out_uv || !d->out_pdf))
        return fail(ctx, MRH_ERR_INVALID, "mrh_wfpg_sample_fields: null xi, out_uv or out_pdf");
    const uint32_t lx = (uint32_t)__builtin_ctz(d->field_x);
    const uint32_t grid = std::min<uint64_t>(d->bin_count, (uint64_t)ctx->num_cus * 4u);
    return launch(ctx, "wfpg_fields", dim3(grid), dim3(WFPG_BLOCK), 0, k_wfpg_fields, d->fields, d->field_x,
                  d->field_y, lx, d->bin_count, gauss_taps(d->filter_alpha), d->bin_offsets, d->ray_ids,
                  d->ray_count, (const float2*)d->xi, (const float2*)d->eval_uv, (float2*)d->out_uv, d->out_dir,
                  d->out_pdf, d->out_eval_pdf, d->out_filtered);
}
