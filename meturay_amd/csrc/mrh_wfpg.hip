// mrh_wfpg.hip — per-bin radiance fields of the reference's wavefront path
// guiding: empty-texel raise, Gaussian filter with concentric-octahedral
// edge folding, 2D piecewise-constant distribution and one guided sample
// per ray (KCGenAndSampleDistribution, WFPGTracerKC.cuh:1433-1545).
//
// One 256-thread workgroup per bin (grid-stride over bins).  The field lives
// in LDS for the whole bin, rows padded by one float.  Float operation order
// is fixed and launch-independent: taps accumulate with fma from -5 to 5,
// prefix sums run left to right, and every ray brings its own uniforms.
#include "mrh_guide.hpp"

#include <cmath>

namespace mrh {
namespace {

constexpr int WFPG_BLOCK = 256;
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
// size - 1e-5.  The interval is found as the count of entries <= xi, in
// log2(size) fixed steps: for a non-decreasing cdf (non-negative radiance)
// that is the one interval the reference's bisection stops at.
template <int SIZE>
__device__ inline float pwc_axis(const float* c, float xi)
{
    int pos = 0;
#pragma unroll
    for (int step = SIZE / 2; step >= 1; step >>= 1)
        pos = (c[pos + step - 1] <= xi) ? pos + step : pos;
    if (pos == SIZE - 1 && c[SIZE - 1] <= xi) pos = SIZE;   // the steps reach at most SIZE - 1
    if (pos == 0) return xi / c[0];
    if (pos == SIZE) return (float)SIZE - 1.0e-5f;
    const float cm = c[pos - 1], cn = c[pos];
    return ((float)(pos - 1) + (xi - cm) / (cn - cm)) + 1.0f;
}

// pdf of the texel under (u,v) times 1/(4 pi) (solid angle of the map);
// coordinates outside [0,1) (or NaN) clamp to the border texels
template <int X, int Y, int LS>
__device__ inline float pwc_pdf(const float* s_pdf, const float* s_py, float u, float v)
{
    const float fx = u * (float)X, fy = v * (float)Y;
    uint32_t ix = (fx > 0.0f) ? (uint32_t)fminf(fx, 65536.0f) : 0u;
    uint32_t iy = (fy > 0.0f) ? (uint32_t)fminf(fy, 65536.0f) : 0u;
    ix = ix >= (uint32_t)X ? X - 1 : ix;
    iy = iy >= (uint32_t)Y ? Y - 1 : iy;
    const float p = s_pdf[iy * LS + ix] * s_py[iy];
    return p * (0.25f * 0.318309886183790671538f);
}

// ---- 8x8 product sampling (ProductSampler, ProductSampler.cuh:16-725, pure
// guiding as KCGenAndSampleDistributionProduct uses it, WFPGTracerKC.cuh:
// 1562-1766): one wave per ray, lane c = cell c of the 8x8 grid.  Every
// random draw is an integer hash of (seed, ray, draw index), so the result
// does not depend on the launch; a rejection round has 64 tries (the wave's
// lanes).  Broadcasts from a wave-uniform lane use readlane, the row
// prefixes (a different source lane per lane) the LDS crossbar.
constexpr int PS = 8;

// lowbias32 (C. Wellons' integer hash): a bijection of 32-bit words
__device__ inline uint32_t wf_hash(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// the ray's key: every draw k of the ray is hash(key ^ (k * odd + c))
__device__ inline uint32_t wf_key(uint32_t seed, uint32_t ray) { return wf_hash((seed * 0x9E3779B9u) ^ wf_hash(ray + 0x632BE5ABu)); }

__device__ inline float wf_uniform(uint32_t key, uint32_t k)
{
    return (float)(wf_hash(key ^ (k * 0x85EBCA6Bu + 0x27D4EB2Fu)) >> 8) * (1.0f / 16777216.0f);
}

// lane broadcasts: any source lane (LDS crossbar), or a wave-uniform one
__device__ inline float lane_f(float v, int lane) { return __shfl(v, lane, 64); }
__device__ inline float lane_u(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

template <int X, int Y, int LS>
__device__ void product_rays(const float* s_f, const float* s_cavg, const float* s_cmax, uint32_t lo, uint32_t hi,
                             const uint32_t* __restrict__ ray_ids, uint32_t ray_count, const float* __restrict__ bxdf,
                             uint32_t seed, const float2* __restrict__ eval_uv, float2* __restrict__ out_uv,
                             float* __restrict__ out_dir, float* __restrict__ out_pdf, float* __restrict__ out_eval_pdf)
{
    constexpr int DX = X / PS, DY = Y / PS;
    const int lane = (int)(threadIdx.x & 63u), wave = (int)(threadIdx.x >> 6);
    const int cx = lane % PS, cy = lane / PS;
    for (uint32_t k = lo + (uint32_t)wave; k < hi; k += WFPG_BLOCK / 64) {
        const uint32_t r = (uint32_t)uni((int)ray_ids[k]);
        if (r >= ray_count) continue;
        const uint32_t key = (uint32_t)uni((int)wf_key(seed, r));
        // product field and its distribution (GeneratePDFAndCDF, :268-330)
        const float prod = fmaxf(1.0e-3f, bxdf[64 * (size_t)r + lane] * s_cavg[lane]);
        float acc = 0.0f, mycdf = 0.0f;
#pragma unroll
        for (int q = 0; q < PS; q++) {
            acc += lane_f(prod, cy * PS + q);
            if (q == cx) mycdf = acc;
        }
        const float total = acc, rec = 1.0f / total;
        const float cdfx = mycdf * rec, pdfx = prod * (float)PS * rec;
        float cyv[PS], rowt[PS];
        float cacc = 0.0f;
#pragma unroll
        for (int q = 0; q < PS; q++) {
            rowt[q] = lane_u(total, q * PS);
            cacc += rowt[q];
            cyv[q] = cacc;
        }
        const float crec = 1.0f / cacc;
#pragma unroll
        for (int q = 0; q < PS; q++) cyv[q] *= crec;
        // cell (Sample, :333-420)
        const float xi0 = wf_uniform(key, 0u), xi1 = wf_uniform(key, 1u);
        int row = 0;
#pragma unroll
        for (int q = 0; q < PS; q++) row += (xi0 > cyv[q]) ? 1 : 0;
        row = uni(row > PS - 1 ? PS - 1 : row);
        int col = 0;
#pragma unroll
        for (int q = 0; q < PS; q++) col += (xi1 > lane_u(cdfx, row * PS + q)) ? 1 : 0;
        col = uni(col > PS - 1 ? PS - 1 : col);
        const int cell = row * PS + col;
        float pY = 0.0f;
#pragma unroll
        for (int q = 0; q < PS; q++)
            if (q == row) pY = rowt[q] * ((float)PS * crec);
        const float pX = lane_u(pdfx, cell);
        // texel inside the cell by rejection (:421-460)
        float u0 = 0.0f, u1 = 0.0f, tex = 0.0f;
        bool found = false;
        for (uint32_t it = 0; it < 24u; it++) {
            const uint32_t kk = 2u + 3u * (it * 64u + (uint32_t)lane);
            const float a0 = wf_uniform(key, kk), a1 = wf_uniform(key, kk + 1u);
            const float ac = wf_uniform(key, kk + 2u);
            const int ix = (int)(a0 * (float)DX), iy = (int)(a1 * (float)DY);
            const float val = s_f[(row * DY + iy) * LS + col * DX + ix];
            const uint64_t m = __ballot(ac <= val / s_cmax[cell]);
            const int pick = uni(m ? (int)__builtin_ctzll(m) : 0);
            u0 = lane_u(a0, pick);
            u1 = lane_u(a1, pick);
            tex = lane_u(val, pick);
            if (m) {
                found = true;
                break;
            }
        }
        const float inner = found ? tex / s_cavg[cell] : 1.0f;
        const float pdf = ((pY * pX) * inner) * (0.25f * 0.318309886183790671538f);
        const float u = ((float)col * (float)DX + u0 * (float)DX) * (1.0f / (float)X);
        const float v = ((float)row * (float)DY + u1 * (float)DY) * (1.0f / (float)Y);
        // pdf at eval_uv (Pdf, :463-490)
        float epdf = 0.0f;
        if (eval_uv) {
            const float2 e = eval_uv[r];
            const float ox = e.x * (float)PS, oy = e.y * (float)PS;
            const float fx = floorf(ox), fy = floorf(oy);
            int ex = (int)fx, ey = (int)fy;
            int ix = (int)floorf((ox - fx) * (float)DX), iy = (int)floorf((oy - fy) * (float)DY);
            ex = ex < 0 ? 0 : (ex > PS - 1 ? PS - 1 : ex);
            ey = ey < 0 ? 0 : (ey > PS - 1 ? PS - 1 : ey);
            ix = ix < 0 ? 0 : (ix > DX - 1 ? DX - 1 : ix);
            iy = iy < 0 ? 0 : (iy > DY - 1 ? DY - 1 : iy);
            const float val = s_f[(ey * DY + iy) * LS + ex * DX + ix];
            const float in2 = val / s_cavg[ey * PS + ex];
            const float ex_pdfx = lane_u(pdfx, uni(ey * PS + ex));
            float ex_pdfy = 0.0f;
#pragma unroll
            for (int q = 0; q < PS; q++)
                if (q == ey) ex_pdfy = rowt[q] * ((float)PS * crec);
            epdf = (ex_pdfx * ex_pdfy) * in2 * (0.25f * 0.318309886183790671538f);
        }
        if (lane == 0) {
            out_uv[r] = make_float2(u, v);
            out_pdf[r] = pdf;
            if (out_dir) {
                const V3 dz = octa_to_dir(u, v);
                out_dir[3 * (size_t)r] = dz.y;
                out_dir[3 * (size_t)r + 1] = dz.z;
                out_dir[3 * (size_t)r + 2] = dz.x;
            }
            if (eval_uv) out_eval_pdf[r] = epdf;
        }
    }
}

// One bin per workgroup iteration; X, Y compile-time so that texel
// indexing is shifts and each row's prefix runs in registers.  Four waves
// per SIMD (128 registers): a 64x64 field's LDS holds the CU to four
// workgroups anyway.
template <int X, int Y, bool PRODUCT>
__global__ __launch_bounds__(WFPG_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_wfpg_fields(
    const float* __restrict__ fields, uint32_t bins, WfpgTaps taps, const uint32_t* __restrict__ offsets,
    const uint32_t* __restrict__ ray_ids, uint32_t ray_count, const float2* __restrict__ xi,
    const float2* __restrict__ eval_uv, float2* __restrict__ out_uv, float* __restrict__ out_dir,
    float* __restrict__ out_pdf, float* __restrict__ out_eval_pdf, float* __restrict__ out_filtered,
    const float* __restrict__ bxdf, uint32_t seed)
{
    constexpr int N = X * Y;
    constexpr int LS = X + 1;            // padded row stride: lane-per-row walks hit distinct banks
    __shared__ float s_a[Y * LS];        // field -> filtered field -> row pdfs
    __shared__ float s_b[Y * LS];        // u-pass result -> row cdfs
    __shared__ float s_row[Y];           // marginal pdf
    __shared__ float s_cy[Y];            // marginal cdf
    __shared__ float s_cavg[PRODUCT ? PS * PS : 1], s_cmax[PRODUCT ? PS * PS : 1];   // 8x8 cells
    constexpr int PER = (N + WFPG_BLOCK - 1) / WFPG_BLOCK;   // texels per thread
    // filter runs: K consecutive texels from one window of K + 10
    constexpr int K = (N / WFPG_BLOCK > 8) ? 8 : ((N / WFPG_BLOCK > 1) ? N / WFPG_BLOCK : 1);
    constexpr int RUNS = N / K;
    static_assert(X % K == 0 && Y % K == 0, "filter runs tile the rows and columns");
    const int tid = (int)threadIdx.x;
    const float dX = 1.0f / (float)X, dY = 1.0f / (float)Y;
    // the next bin's field is loaded into registers while the current bin's
    // rays are sampled
    float pre[PER];
    auto fetch = [&](uint32_t b) {
        const float* src = fields + (size_t)b * N;
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int p = i * WFPG_BLOCK + tid;
            pre[i] = (p < N) ? src[p] : 0.0f;
        }
    };
    if (blockIdx.x < bins) fetch(blockIdx.x);
    for (uint32_t bin = blockIdx.x; bin < bins; bin += gridDim.x) {
        // this thread's first ray of the bin, loaded before the field work
        const uint32_t lo = offsets[bin], hi = offsets[bin + 1];
        uint32_t r0 = 0xFFFFFFFFu;
        float2 x0 = make_float2(0.0f, 0.0f);
        if (lo + (uint32_t)tid < hi) {
            r0 = ray_ids[lo + tid];
            if (!PRODUCT && r0 < ray_count) x0 = xi[r0];
        }
        // radiance query result: an empty texel keeps a small weight
        // (WFPGTracerKC.cuh:1316-1317)
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int p = i * WFPG_BLOCK + tid;
            if (p < N) s_a[(p / X) * LS + (p % X)] = (pre[i] == 0.0f) ? 1.0e-4f : pre[i];
        }
        __syncthreads();
        // filter along u (BlockTextureFilter.cuh:157-174): runs of K texels
        // of one row
        for (int q = tid; q < RUNS; q += WFPG_BLOCK) {
            const int y = q / (X / K), x0 = (q % (X / K)) * K;
            float w[K + 2 * WFPG_HALF];
#pragma unroll
            for (int i = 0; i < K + 2 * WFPG_HALF; i++) {
                int qx = x0 - WFPG_HALF + i, qy = y;
                if (qx < 0 || qx >= X) octa_fold(qx, y, X, Y, qx, qy);
                w[i] = s_a[qy * LS + qx];
            }
#pragma unroll
            for (int o = 0; o < K; o++) {
                float acc = 0.0f;
#pragma unroll
                for (int j = 0; j < 2 * WFPG_HALF + 1; j++) acc = fmaf(w[o + j], taps.w[j], acc);
                s_b[y * LS + x0 + o] = acc;
            }
        }
        __syncthreads();
        // filter along v (:178-195): runs of K texels of one column
        for (int q = tid; q < RUNS; q += WFPG_BLOCK) {
            const int x = q % X, y0 = (q / X) * K;
            float w[K + 2 * WFPG_HALF];
#pragma unroll
            for (int i = 0; i < K + 2 * WFPG_HALF; i++) {
                int qx = x, qy = y0 - WFPG_HALF + i;
                if (qy < 0 || qy >= Y) octa_fold(x, qy, X, Y, qx, qy);
                w[i] = s_b[qy * LS + qx];
            }
#pragma unroll
            for (int o = 0; o < K; o++) {
                float acc = 0.0f;
#pragma unroll
                for (int j = 0; j < 2 * WFPG_HALF + 1; j++) acc = fmaf(w[o + j], taps.w[j], acc);
                s_a[(y0 + o) * LS + x] = acc;
                if (out_filtered) out_filtered[(size_t)bin * N + (y0 + o) * X + x] = acc;
            }
        }
        __syncthreads();
        if constexpr (!PRODUCT) {
            // conditional distributions, one lane of the first wave per row
            // (GPUBlockPWCDistribution.cuh:107-156): the row total first, then
            // the same left-to-right prefix again while the row is normalised
            float m = 0.0f;                  // this lane's row total * dY
            if (tid < Y) {
                float* ra = s_a + tid * LS;
                float* rb = s_b + tid * LS;
                float s = 0.0f;
    #pragma unroll 16
                for (int x = 0; x < X; x++) s += ra[x];
                const float sdx = s * dX;
                const bool norm = (sdx != 0.0f);
                const float inv = norm ? 1.0f / sdx : 1.0f;
                float c = 0.0f;
    #pragma unroll 16
                for (int x = 0; x < X; x++) {
                    const float v = ra[x];
                    c += v;
                    if (norm) {
                        ra[x] = v * inv;
                        rb[x] = (c * dX) * inv;
                    } else {
                        rb[x] = c * dX;
                    }
                }
                m = s * dY;
            }
            // marginal over rows (:158-199) in the first wave: lane y adds the
            // row terms 0..y in order, so it holds exactly the left-to-right prefix
            if (tid < 64) {
                float c = 0.0f;
    #pragma unroll
                for (int k = 0; k < Y; k++) {
                    const float mk = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m), k));
                    if (k <= tid) c += mk;
                }
                const float T = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c), Y - 1));
                if (tid < Y) {
                    float p = m, cc = c;
                    if (T != 0.0f) {
                        const float inv = 1.0f / T;
                        p = p * inv;
                        cc = cc * inv;
                    }
                    s_row[tid] = p * (float)Y;
                    s_cy[tid] = cc;
                }
            }
        } else {
            // the field reduced to 8x8 cells: average and maximum, row by row
            // (ProductSampler constructor, ProductSampler.cuh:176-228)
            if (tid < PS * PS) {
                constexpr int DX = X / PS, DY = Y / PS;
                const int cx = tid % PS, cy = tid / PS;
                float sum = 0.0f, mx = 0.0f;
                for (int y = 0; y < DY; y++)
                    for (int x = 0; x < DX; x++) {
                        const float v = s_a[(cy * DY + y) * LS + cx * DX + x];
                        sum += v;
                        mx = fmaxf(mx, v);
                    }
                s_cavg[tid] = sum / (float)(DX * DY);
                s_cmax[tid] = mx;
            }
        }
        __syncthreads();
        if (bin + gridDim.x < bins) fetch(bin + gridDim.x);
        if constexpr (!PRODUCT) {
            // one guided sample per ray of the bin (WFPGTracerKC.cuh:1466-1533)
            for (uint32_t k = lo + tid; k < hi; k += WFPG_BLOCK) {
                const bool first = (k == lo + (uint32_t)tid);
                const uint32_t r = first ? r0 : ray_ids[k];
                if (r >= ray_count) continue;
                const float2 x2 = first ? x0 : xi[r];
                float u, v;
                if (s_cy[Y - 1] == 0.0f) {
                    u = x2.x;
                    v = x2.y;
                } else {
                    const float iy = pwc_axis<Y>(s_cy, x2.y);
                    int row = (int)iy;
                    row = row >= Y ? Y - 1 : (row < 0 ? 0 : row);
                    const float ix = pwc_axis<X>(s_b + row * LS, x2.x);
                    u = ix * dX;
                    v = iy * dY;
                }
                out_uv[r] = make_float2(u, v);
                out_pdf[r] = pwc_pdf<X, Y, LS>(s_a, s_row, u, v);
                if (out_dir) {
                    const V3 d = octa_to_dir(u, v);   // z up; the path uses y up
                    out_dir[3 * (size_t)r] = d.y;
                    out_dir[3 * (size_t)r + 1] = d.z;
                    out_dir[3 * (size_t)r + 2] = d.x;
                }
                if (eval_uv) {
                    const float2 e = eval_uv[r];
                    out_eval_pdf[r] = pwc_pdf<X, Y, LS>(s_a, s_row, e.x, e.y);
                }
            }
        } else {
            product_rays<X, Y, LS>(s_a, s_cavg, s_cmax, lo, hi, ray_ids, ray_count, bxdf, seed, eval_uv, out_uv,
                                   out_dir, out_pdf, out_eval_pdf);
        }
        __syncthreads();
    }
}

// Workgroups per CU the kernel's LDS leaves room for (at most 8).
template <int X, int Y>
constexpr uint32_t wfpg_blocks_per_cu()
{
    const uint32_t lds = (uint32_t)(2 * Y * (X + 1) + 2 * Y) * 4u;
    const uint32_t b = (160u * 1024u) / lds;
    return b > 8u ? 8u : b;
}

template <int X, int Y>
int32_t launch_wfpg(mrh_context* ctx, const mrh_wfpg_fields_desc* d)
{
    const uint32_t grid = (uint32_t)std::min<uint64_t>(d->bin_count,
                                                        (uint64_t)ctx->num_cus * wfpg_blocks_per_cu<X, Y>());
    if (d->bxdf_field)
        return launch(ctx, "wfpg_fields_product", dim3(grid), dim3(WFPG_BLOCK), 0, k_wfpg_fields<X, Y, true>,
                      d->fields, d->bin_count, gauss_taps(d->filter_alpha), d->bin_offsets, d->ray_ids, d->ray_count,
                      (const float2*)d->xi, (const float2*)d->eval_uv, (float2*)d->out_uv, d->out_dir, d->out_pdf,
                      d->out_eval_pdf, d->out_filtered, d->bxdf_field, d->seed);
    return launch(ctx, "wfpg_fields", dim3(grid), dim3(WFPG_BLOCK), 0, k_wfpg_fields<X, Y, false>, d->fields,
                  d->bin_count, gauss_taps(d->filter_alpha), d->bin_offsets, d->ray_ids, d->ray_count,
                  (const float2*)d->xi, (const float2*)d->eval_uv, (float2*)d->out_uv, d->out_dir, d->out_pdf,
                  d->out_eval_pdf, d->out_filtered, (const float*)nullptr, 0u);
}

template <int X>
int32_t launch_wfpg_x(mrh_context* ctx, const mrh_wfpg_fields_desc* d)
{
    switch (d->field_y) {
    case 8: return launch_wfpg<X, 8>(ctx, d);
    case 16: return launch_wfpg<X, 16>(ctx, d);
    case 32: return launch_wfpg<X, 32>(ctx, d);
    default: return launch_wfpg<X, 64>(ctx, d);
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
    if (d->ray_count && ((!d->xi && !d->bxdf_field) || !d->out_uv || !d->out_pdf))
        return fail(ctx, MRH_ERR_INVALID, "mrh_wfpg_sample_fields: null xi, out_uv or out_pdf");
    switch (d->field_x) {
    case 8: return launch_wfpg_x<8>(ctx, d);
    case 16: return launch_wfpg_x<16>(ctx, d);
    case 32: return launch_wfpg_x<32>(ctx, d);
    default: return launch_wfpg_x<64>(ctx, d);
    }
}
