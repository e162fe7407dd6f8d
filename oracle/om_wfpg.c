/*
 * om_wfpg.c — oracle for the per-bin radiance-field stage of the reference's
 * wavefront path guiding (KCGenAndSampleDistribution,
 * TracerCUDA/WFPGTracerKC.cuh:1370-1545 with GenerateRadianceField :1252-1345):
 * raise empty texels, Gaussian-filter the field with concentric-octahedral
 * wrapping, build the 2D piecewise-constant distribution, draw one guided
 * direction per ray of the bin and evaluate the guide pdf.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Float operation order is written out so that the GPU kernel can follow it
 * exactly: the filter accumulates with fmaf over taps -5..5 (the reference's
 * `dataOut += tex * w` under nvcc's default contraction), and every prefix sum
 * adds left to right.  The reference's prefix sums are CUB block scans whose
 * association differs; its own test bounds that difference by 1e-4
 * (Test_TracerCUDA/T_BlockPWCDistribution.cu:330-345).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define OM_WFPG_TAPS 11
#define OM_WFPG_HALF 5

/* StaticGaussianFilter1D<11> (TracerCUDA/BlockTextureFilter.cuh:84-121):
 * g(t) = 1 / (sqrt(pi) sqrt(2) sqrt(alpha)) * exp(-t^2 / (2 alpha)),
 * weight(i) = g(i) / sum_{k=-5..5} g(k).  The weights are formed in float as
 * the reference's host constructor forms its total. */
void om_wfpg_gauss_weights(float alpha, float w[11])
{
    const float W = 1.0f / 1.77245385090551602729f / 1.41421356237309504880f;
    float g[OM_WFPG_TAPS];
    for (int i = -OM_WFPG_HALF; i <= OM_WFPG_HALF; i++) {
        const float t = (float)i;
        g[i + OM_WFPG_HALF] = W / sqrtf(alpha) * expf(-(t * t) * 0.5f / alpha);
    }
    float total = 0.0f;
    for (int i = 0; i < OM_WFPG_TAPS; i++) total += g[i];
    for (int i = 0; i < OM_WFPG_TAPS; i++) w[i] = g[i] / total;
}

/* Utility::CocentricOctohedralWrapInt (RayLib/CoordinateConversion.h:213-235):
 * a texel index outside [0,X) x [0,Y) folds back across the map's edges; an
 * odd number of folds in exactly one axis mirrors both coordinates. */
void om_octa_wrap_int(int32_t x, int32_t y, int32_t X, int32_t Y, int32_t out[2])
{
    const int32_t sx = (x < 0) ? -2 * X - x : x;
    const int32_t sy = (y < 0) ? -2 * Y - y : y;
    const int32_t ix = sx / X, iy = sy / Y;
    int32_t fx = abs(sx % X), fy = abs(sy % Y);
    if ((ix & 1) ^ (iy & 1)) {
        fx = (X - 1) - fx;
        fy = (Y - 1) - fy;
    }
    out[0] = fx;
    out[1] = fy;
}

/* BlockTextureFilter2D::Filter (TracerCUDA/BlockTextureFilter.cuh:133-197):
 * an X pass then a Y pass over [Y][X] texels; a tap outside the map along the
 * pass axis is wrapped by om_octa_wrap_int. */
void om_wfpg_filter(const float* field, uint32_t X, uint32_t Y, const float w[11], float* out)
{
    const uint32_t n = X * Y;
    float* tmp = (float*)malloc((size_t)n * sizeof(float));
    for (uint32_t p = 0; p < n; p++) {
        const int32_t x = (int32_t)(p % X), y = (int32_t)(p / X);
        float acc = 0.0f;
        for (int32_t j = -OM_WFPG_HALF; j <= OM_WFPG_HALF; j++) {
            int32_t q[2] = {x + j, y};
            if (q[0] >= (int32_t)X || q[0] < 0) om_octa_wrap_int(x + j, y, (int32_t)X, (int32_t)Y, q);
            acc = fmaf(field[(size_t)q[1] * X + (uint32_t)q[0]], w[j + OM_WFPG_HALF], acc);
        }
        tmp[p] = acc;
    }
    for (uint32_t p = 0; p < n; p++) {
        const int32_t x = (int32_t)(p % X), y = (int32_t)(p / X);
        float acc = 0.0f;
        for (int32_t j = -OM_WFPG_HALF; j <= OM_WFPG_HALF; j++) {
            int32_t q[2] = {x, y + j};
            if (q[1] >= (int32_t)Y || q[1] < 0) om_octa_wrap_int(x, y + j, (int32_t)X, (int32_t)Y, q);
            acc = fmaf(tmp[(size_t)q[1] * X + (uint32_t)q[0]], w[j + OM_WFPG_HALF], acc);
        }
        out[p] = acc;
    }
    free(tmp);
}

/* BlockPWCDistribution2D constructor (TracerCUDA/GPUBlockPWCDistribution.cuh:
 * 107-200).  Per row: raw inclusive prefix s, total S; cdf = (s * dX) *
 * (1 / (S * dX)), pdf = f * (1 / (S * dX)) (left as f when S = 0).  Marginal
 * over rows: m = S * dY, inclusive prefix c, total T; cdfY = c * (1 / T),
 * pdfY = (m * (1 / T)) * Y.  cdf arrays hold the END of each texel (the
 * leading 0 is implicit), as the reference keeps them in shared memory. */
void om_wfpg_pwc(const float* f, uint32_t X, uint32_t Y, float* pdfx, float* cdfx, float* pdfy, float* cdfy)
{
    const float dX = 1.0f / (float)X, dY = 1.0f / (float)Y;
    float* rows = (float*)malloc((size_t)Y * sizeof(float));
    for (uint32_t y = 0; y < Y; y++) {
        const float* r = f + (size_t)y * X;
        float s = 0.0f;
        for (uint32_t x = 0; x < X; x++) {
            s += r[x];
            cdfx[(size_t)y * X + x] = s;
        }
        rows[y] = s;
        const float sdx = s * dX;
        for (uint32_t x = 0; x < X; x++) {
            float c = cdfx[(size_t)y * X + x] * dX, p = r[x];
            if (sdx != 0.0f) {
                p *= (1.0f / sdx);
                c *= (1.0f / sdx);
            }
            pdfx[(size_t)y * X + x] = p;
            cdfx[(size_t)y * X + x] = c;
        }
    }
    float c = 0.0f;
    for (uint32_t y = 0; y < Y; y++) {
        c += rows[y] * dY;
        cdfy[y] = c;
    }
    const float T = c;
    for (uint32_t y = 0; y < Y; y++) {
        float p = rows[y] * dY, cc = cdfy[y];
        if (T != 0.0f) {
            p *= (1.0f / T);
            cc *= (1.0f / T);
        }
        pdfy[y] = p * (float)Y;
        cdfy[y] = cc;
    }
    free(rows);
}

/* One axis of BlockPWCDistribution2D::Sample (GPUBlockPWCDistribution.cuh:
 * 202-275 with GPUFunctions::BinarySearchInBetween, BinarySearch.cuh:13-39):
 * the continuous index of xi in a cdf list whose entries end the texels.
 * The search looks for k in [0, size-2] with c[k] <= xi < c[k+1]; the
 * reference's search also reads c[size] when xi >= c[size-1], a case whose
 * result it then replaces by the clamp below, so the two agree. */
static float om_pwc_axis(const float* c, uint32_t size, float xi)
{
    int32_t lo = 0, hi = (int32_t)size - 2;
    while (lo <= hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (c[mid] <= xi && c[mid + 1] > xi) return ((float)mid + (xi - c[mid]) / (c[mid + 1] - c[mid])) + 1.0f;
        if (xi < c[mid]) hi = mid - 1;
        else lo = mid + 1;
    }
    if (xi < c[size - 1]) return xi / c[0];
    return (float)size - 1.0e-5f;
}

void om_wfpg_pwc_sample(const float* cdfx, const float* cdfy, uint32_t X, uint32_t Y, float xi0, float xi1,
                        float uv[2])
{
    if (cdfy[Y - 1] == 0.0f) {           /* empty distribution: uniform */
        uv[0] = xi0;
        uv[1] = xi1;
        return;
    }
    const float iy = om_pwc_axis(cdfy, Y, xi1);
    int32_t row = (int32_t)iy;
    row = (row >= (int32_t)Y) ? (int32_t)Y - 1 : (row < 0 ? 0 : row);
    const float ix = om_pwc_axis(cdfx + (size_t)row * X, X, xi0);
    uv[0] = ix * (1.0f / (float)X);
    uv[1] = iy * (1.0f / (float)Y);
}

/* BlockPWCDistribution2D::Pdf (GPUBlockPWCDistribution.cuh:277-300) at
 * (u*X, v*Y), times 1/(4 pi) for the octahedral map's solid-angle Jacobian
 * (WFPGTracerKC.cuh:1511-1512).  Indices clamp to the border texels (the
 * reference reads past the row for u = 1). */
float om_wfpg_pwc_pdf(const float* pdfx, const float* pdfy, uint32_t X, uint32_t Y, float u, float v)
{
    const float fx = u * (float)X, fy = v * (float)Y;
    uint32_t ix = (fx > 0.0f) ? (uint32_t)fminf(fx, 65536.0f) : 0u;
    uint32_t iy = (fy > 0.0f) ? (uint32_t)fminf(fy, 65536.0f) : 0u;
    if (ix >= X) ix = X - 1;
    if (iy >= Y) iy = Y - 1;
    const float p = pdfx[(size_t)iy * X + ix] * pdfy[iy];
    return p * (0.25f * 0.318309886183790671538f);
}

/* ---- 8x8 product sampling (ProductSampler, TracerCUDA/ProductSampler.cuh:
 * 16-725, as KCGenAndSampleDistributionProduct uses it in pure guiding,
 * WFPGTracerKC.cuh:1562-1766).  The filtered field is reduced to 8x8 cells
 * (average and maximum over each cell's texels, row by row); per ray the
 * product max(1e-3, bxdf[cell] * average[cell]) gets a piecewise-constant
 * distribution over the cells (rows: prefix left to right; the marginal over
 * the row totals), a cell is drawn (row: first with xi0 <= cdf, column: first
 * with xi1 <= cdf), and a texel inside it by rejection: up to 24 rounds of 64
 * tries (the wave's lanes), each a uniform point of the cell accepted with
 * probability texel / cell maximum, the lowest accepting try winning; with
 * none, try 0 of the last round is kept as a uniform inner sample.  pdf =
 * pdfY * pdfX * (texel / cell average, 1 for a uniform inner sample) / 4 pi.
 * The reference's rounds are 32-lane warps and its uniforms come from a
 * per-thread generator; here a round has 64 tries and every draw is an
 * integer hash of (seed, ray, draw index). */
/* lowbias32 (C. Wellons' integer hash); draw k of a ray is
 * hash(key ^ (k * 0x85EBCA6B + 0x27D4EB2F)), key = hash((seed * 0x9E3779B9)
 * ^ hash(ray + 0x632BE5AB)), its top 24 bits over 2^24 */
static uint32_t wf_hash(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

float om_wfpg_uniform(uint32_t seed, uint32_t ray, uint32_t k)
{
    const uint32_t key = wf_hash((seed * 0x9E3779B9u) ^ wf_hash(ray + 0x632BE5ABu));
    return (float)(wf_hash(key ^ (k * 0x85EBCA6Bu + 0x27D4EB2Fu)) >> 8) * (1.0f / 16777216.0f);
}

#define OM_PS 8

static void om_small_field(const float* f, uint32_t X, uint32_t Y, float* avg, float* mx)
{
    const uint32_t DX = X / OM_PS, DY = Y / OM_PS;
    for (uint32_t c = 0; c < OM_PS * OM_PS; c++) {
        const uint32_t cx = c % OM_PS, cy = c / OM_PS;
        float s = 0.0f, m = 0.0f;
        for (uint32_t y = 0; y < DY; y++)
            for (uint32_t x = 0; x < DX; x++) {
                const float v = f[(size_t)(cy * DY + y) * X + cx * DX + x];
                s += v;
                m = fmaxf(m, v);
            }
        avg[c] = s / (float)(DX * DY);
        mx[c] = m;
    }
}

/* the product distribution of one ray: pdfx/cdfx [64], pdfy/cdfy [8] */
static void om_product_pwc(const float* avg, const float* bxdf, float* pdfx, float* cdfx, float* pdfy, float* cdfy)
{
    float prod[OM_PS * OM_PS], rows[OM_PS];
    for (int c = 0; c < OM_PS * OM_PS; c++) prod[c] = fmaxf(1.0e-3f, bxdf[c] * avg[c]);
    for (int r = 0; r < OM_PS; r++) {
        float acc = 0.0f;
        for (int x = 0; x < OM_PS; x++) {
            acc += prod[r * OM_PS + x];
            cdfx[r * OM_PS + x] = acc;
        }
        const float tot = acc, rec = 1.0f / tot;
        rows[r] = tot;
        for (int x = 0; x < OM_PS; x++) {
            cdfx[r * OM_PS + x] *= rec;
            pdfx[r * OM_PS + x] = prod[r * OM_PS + x] * (float)OM_PS * rec;
        }
    }
    float acc = 0.0f;
    for (int r = 0; r < OM_PS; r++) {
        acc += rows[r];
        cdfy[r] = acc;
    }
    const float rec = 1.0f / acc;
    for (int r = 0; r < OM_PS; r++) {
        cdfy[r] *= rec;
        pdfy[r] = rows[r] * ((float)OM_PS * rec);
    }
}

static int om_first_at_or_above(const float* c, float xi)
{
    int k = 0;
    for (int i = 0; i < OM_PS; i++) k += (xi > c[i]) ? 1 : 0;
    return k > OM_PS - 1 ? OM_PS - 1 : k;
}

static float om_product_pdf(const float* f, uint32_t X, uint32_t Y, const float* avg, const float* pdfx,
                            const float* pdfy, float u, float v)
{
    const uint32_t DX = X / OM_PS, DY = Y / OM_PS;
    const float ox = u * (float)OM_PS, oy = v * (float)OM_PS;
    const float fx = floorf(ox), fy = floorf(oy);
    int cx = (int)fx, cy = (int)fy;
    int ix = (int)floorf((ox - fx) * (float)DX), iy = (int)floorf((oy - fy) * (float)DY);
    cx = cx < 0 ? 0 : (cx > OM_PS - 1 ? OM_PS - 1 : cx);
    cy = cy < 0 ? 0 : (cy > OM_PS - 1 ? OM_PS - 1 : cy);
    ix = ix < 0 ? 0 : (ix > (int)DX - 1 ? (int)DX - 1 : ix);
    iy = iy < 0 ? 0 : (iy > (int)DY - 1 ? (int)DY - 1 : iy);
    const float val = f[(size_t)(cy * (int)DY + iy) * X + cx * (int)DX + ix];
    const float inner = val / avg[cy * OM_PS + cx];
    return (pdfx[cy * OM_PS + cx] * pdfy[cy]) * inner * (0.25f * 0.318309886183790671538f);
}

static void om_product_sample(const float* f, uint32_t X, uint32_t Y, const float* avg, const float* mx,
                              const float* pdfx, const float* cdfx, const float* pdfy, const float* cdfy,
                              uint32_t seed, uint32_t ray, float uv[2], float* pdf)
{
    const uint32_t DX = X / OM_PS, DY = Y / OM_PS;
    const float xi0 = om_wfpg_uniform(seed, ray, 0), xi1 = om_wfpg_uniform(seed, ray, 1);
    const int row = om_first_at_or_above(cdfy, xi0);
    const int col = om_first_at_or_above(cdfx + row * OM_PS, xi1);
    const int cell = row * OM_PS + col;
    float u0 = 0.0f, u1 = 0.0f, tex = 0.0f;
    int found = 0;
    for (uint32_t it = 0; it < 24 && !found; it++)
        for (uint32_t lane = 0; lane < 64; lane++) {
            const uint32_t k = 2u + 3u * (it * 64u + lane);
            const float a0 = om_wfpg_uniform(seed, ray, k), a1 = om_wfpg_uniform(seed, ray, k + 1);
            const float acc = om_wfpg_uniform(seed, ray, k + 2);
            const int ix = (int)(a0 * (float)DX), iy = (int)(a1 * (float)DY);
            const float val = f[(size_t)(row * (int)DY + iy) * X + col * (int)DX + ix];
            if (lane == 0) {
                u0 = a0;
                u1 = a1;
                tex = val;
            }
            if (acc <= val / mx[cell]) {
                u0 = a0;
                u1 = a1;
                tex = val;
                found = 1;
                break;
            }
        }
    const float inner = found ? tex / avg[cell] : 1.0f;
    *pdf = ((pdfy[row] * pdfx[cell]) * inner) * (0.25f * 0.318309886183790671538f);
    uv[0] = ((float)col * (float)DX + u0 * (float)DX) * (1.0f / (float)X);
    uv[1] = ((float)row * (float)DY + u1 * (float)DY) * (1.0f / (float)Y);
}

/* The whole stage for every bin (KCGenAndSampleDistribution,
 * WFPGTracerKC.cuh:1433-1545, radiance query :1312-1318). */
int om_wfpg_sample_fields(const om_wfpg_desc* d)
{
    const uint32_t X = d->field_x, Y = d->field_y, n = X * Y;
    float w[OM_WFPG_TAPS];
    om_wfpg_gauss_weights(d->filter_alpha, w);
    float* raw = (float*)malloc((size_t)n * sizeof(float));
    float* filt = (float*)malloc((size_t)n * sizeof(float));
    float* pdfx = (float*)malloc((size_t)n * sizeof(float));
    float* cdfx = (float*)malloc((size_t)n * sizeof(float));
    float pdfy[64], cdfy[64];
    for (uint32_t b = 0; b < d->bin_count; b++) {
        const float* src = d->fields + (size_t)b * n;
        for (uint32_t p = 0; p < n; p++) raw[p] = (src[p] == 0.0f) ? 1.0e-4f : src[p];
        om_wfpg_filter(raw, X, Y, w, filt);
        if (d->out_filtered) memcpy(d->out_filtered + (size_t)b * n, filt, (size_t)n * sizeof(float));
        float savg[64], smax[64], ppx[64], pcx[64], ppy[8], pcy[8];
        if (d->bxdf_field) om_small_field(filt, X, Y, savg, smax);
        else om_wfpg_pwc(filt, X, Y, pdfx, cdfx, pdfy, cdfy);
        const uint32_t lo = d->bin_offsets[b], hi = d->bin_offsets[b + 1];
        for (uint32_t k = lo; k < hi; k++) {
            const uint32_t r = d->ray_ids[k];
            if (r >= d->ray_count) continue;
            float uv[2];
            if (d->bxdf_field) {
                om_product_pwc(savg, d->bxdf_field + 64 * (size_t)r, ppx, pcx, ppy, pcy);
                float pdf;
                om_product_sample(filt, X, Y, savg, smax, ppx, pcx, ppy, pcy, d->seed, r, uv, &pdf);
                d->out_uv[2 * (size_t)r] = uv[0];
                d->out_uv[2 * (size_t)r + 1] = uv[1];
                d->out_pdf[r] = pdf;
                if (d->eval_uv)
                    d->out_eval_pdf[r] = om_product_pdf(filt, X, Y, savg, ppx, ppy, d->eval_uv[2 * (size_t)r],
                                                        d->eval_uv[2 * (size_t)r + 1]);
            } else {
                om_wfpg_pwc_sample(cdfx, cdfy, X, Y, d->xi[2 * (size_t)r], d->xi[2 * (size_t)r + 1], uv);
                d->out_uv[2 * (size_t)r] = uv[0];
                d->out_uv[2 * (size_t)r + 1] = uv[1];
                d->out_pdf[r] = om_wfpg_pwc_pdf(pdfx, pdfy, X, Y, uv[0], uv[1]);
                if (d->eval_uv)
                    d->out_eval_pdf[r] = om_wfpg_pwc_pdf(pdfx, pdfy, X, Y, d->eval_uv[2 * (size_t)r],
                                                         d->eval_uv[2 * (size_t)r + 1]);
            }
            if (d->out_dir) {
                float z[3];
                om_octa_to_dirs(uv, 1, z);
                d->out_dir[3 * (size_t)r] = z[1];     /* y-up: (y, z, x) of the z-up map */
                d->out_dir[3 * (size_t)r + 1] = z[2];
                d->out_dir[3 * (size_t)r + 2] = z[0];
            }
        }
    }
    free(raw);
    free(filt);
    free(pdfx);
    free(cdfx);
    return 0;
}
