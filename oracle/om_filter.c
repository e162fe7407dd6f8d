/*
 * om_filter.c — oracle: reconstruction filters (GPUReconFilter::FilterToImg,
 * GPUReconFilter.h:59-260, GPUReconFilterKC.cuh, and the Box / Tent /
 * Gaussian / Mitchell functors of GPUReconFilter{Box,Tent,Gaussian,
 * Mitchell}.h), restated as a gather with the HIP kernel's summation order.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * For output pixel p, the pixels q whose samples can reach p are visited in
 * row-major order and each q's samples in increasing sample index; a sample
 * is accepted when p lies in its window (FilterRadiusToPixelWH, shifted for
 * even widths when the sample sits in the lower half of its pixel) and, for
 * windows wider than one pixel, p's centre lies within the radius.  All
 * arithmetic is FP32 in the kernel's order.
 */
#include "oracle.h"
#include "om.h"

#include <stdlib.h>
#include <string.h>

/* GPUReconFilterKC.cuh:8-27 */
int om_filter_pixel_wh(float r)
{
    uint32_t result = 1;
    if (r == 0.0f) return (int)result;
    uint32_t quot = (uint32_t)(r / 0.5f);
    float rem = fmodf(r, 0.5f);
    if (rem == 0.0f) quot -= 1;
    result += quot + 1;
    return (int)result;
}

typedef struct {
    uint32_t type;
    int range, lo, hi;
    float radius, rsq, recip, neg_alpha, gexp;
    float m01[4], m12[4];
} om_filt;

static void filt_make(const om_filter_desc* d, om_filt* f)
{
    memset(f, 0, sizeof(*f));
    f->type = d->type;
    f->radius = d->radius;
    f->rsq = d->radius * d->radius;
    f->recip = (d->radius > 0.0f) ? 1.0f / d->radius : 0.0f;
    f->range = om_filter_pixel_wh(d->radius);
    f->lo = -(f->range - 1) / 2;
    f->hi = (f->range + 2) / 2;
    if (d->type == OM_FILTER_GAUSSIAN) {
        /* GPUGaussianFilterFunctor: exponent = Gauss(radius) */
        f->neg_alpha = -d->alpha;
        f->gexp = expf(f->neg_alpha * d->radius * d->radius);
    }
    if (d->type == OM_FILTER_MITCHELL) {
        /* Mitchell & Netravali 1988 (B, C) cubic; constant term 6 - 2B on
         * [0, 1) (the reference writes 6 + 2B, GPUReconFilterMitchell.h:80) */
        float t = 1.0f / 6.0f, b = d->b, c = d->c;
        f->m01[0] = t * (12 - 9 * b - 6 * c);
        f->m01[1] = t * (-18 + 12 * b + 6 * c);
        f->m01[2] = 0.0f;
        f->m01[3] = t * (6 - 2 * b);
        f->m12[0] = t * (-1 * b - 6 * c);
        f->m12[1] = t * (6 * b + 30 * c);
        f->m12[2] = t * (-12 * b - 48 * c);
        f->m12[3] = t * (8 * b + 24 * c);
    }
}

static float mitchell1(const om_filt* f, float x)
{
    x = fabsf(x);
    float x2 = x * x;
    float x3 = x2 * x;
    if (!(x < 2.0f)) return 0.0f;
    const float* c = (x < 1.0f) ? f->m01 : f->m12;
    return c[0] * x3 + c[1] * x2 + c[2] * x + c[3];
}

/* the functors' operator() at |pixel centre - sample| = (dx, dy) */
static float filt_weight(const om_filt* f, float dx, float dy)
{
    switch (f->type) {
    case OM_FILTER_BOX:
        if (f->radius == 0.0f) return 1.0f;
        return (dx < f->radius && dy < f->radius) ? 1.0f : 0.0f;
    case OM_FILTER_TENT: {
        float lx = fmaxf(0.0f, f->radius - dx);
        float ly = fmaxf(0.0f, f->radius - dy);
        lx *= f->recip;
        ly *= f->recip;
        return lx * ly;
    }
    case OM_FILTER_GAUSSIAN: {
        float gx = fmaxf(0.0f, expf(f->neg_alpha * dx * dx) - f->gexp);
        float gy = fmaxf(0.0f, expf(f->neg_alpha * dy * dy) - f->gexp);
        return gx * gy;
    }
    default:
        return mitchell1(f, 2.0f * dx * f->recip) * mitchell1(f, 2.0f * dy * f->recip);
    }
}

float om_filter_weight(const om_filter_desc* d, float dx, float dy)
{
    om_filt f;
    filt_make(d, &f);
    return filt_weight(&f, fabsf(dx), fabsf(dy));
}

/* Gather over samples grouped by pixel: samples of pixel q are
 * idx[off[q] .. off[q+1]) (idx NULL: off NULL too, samples q*spp ..). */
static void gather(const om_filt* f, const float* values, const float* coords, const uint32_t* idx,
                   const uint32_t* off, uint32_t spp, uint32_t w, uint32_t h, float* weighted, float* weights)
{
    int dmin = (f->range & 1) ? f->lo : f->lo - 1;
    int dmax = f->hi - 1;
    for (uint32_t p = 0; p < w * h; p++) {
        int px = (int)(p % w), py = (int)(p / w);
        float pcx = (float)px + 0.5f, pcy = (float)py + 0.5f;
        float ar = 0.0f, ag = 0.0f, ab = 0.0f, ws = 0.0f;
        for (int sy = py - dmax; sy <= py - dmin; sy++) {
            if (sy < 0 || sy >= (int)h) continue;
            for (int sx = px - dmax; sx <= px - dmin; sx++) {
                if (sx < 0 || sx >= (int)w) continue;
                uint32_t q = (uint32_t)sy * w + (uint32_t)sx;
                uint32_t b = idx ? off[q] : q * spp;
                uint32_t e = idx ? off[q + 1] : b + spp;
                for (uint32_t k = b; k < e; k++) {
                    uint32_t s = idx ? idx[k] : k;
                    float cx = coords[2 * (size_t)s], cy = coords[2 * (size_t)s + 1];
                    float ix = truncf(cx), iy = truncf(cy);
                    float rx = cx - ix, ry = cy - iy;
                    int dxi = px - (int)ix, dyi = py - (int)iy;
                    int lox = f->lo, hix = f->hi, loy = f->lo, hiy = f->hi;
                    if ((f->range & 1) == 0) {
                        if (rx < 0.5f) { lox--; hix--; }
                        if (ry < 0.5f) { loy--; hiy--; }
                    }
                    if (dxi < lox || dxi >= hix || dyi < loy || dyi >= hiy) continue;
                    if (f->range != 1) {
                        float vx = cx - ((ix + (float)dxi) + 0.5f);
                        float vy = cy - ((iy + (float)dyi) + 0.5f);
                        float lsq = 0.0f;
                        lsq += vx * vx;
                        lsq += vy * vy;
                        if (!(lsq <= f->rsq)) continue;
                    }
                    float wgt = filt_weight(f, fabsf(pcx - cx), fabsf(pcy - cy)) * 1.0f;
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

void om_filter_samples(const om_filter_desc* d, const float* values, const float* coords, uint32_t n,
                       uint32_t w, uint32_t h, float* weighted, float* weights)
{
    om_filt f;
    filt_make(d, &f);
    uint32_t pixels = w * h;
    /* stable counting sort by pixel (samples outside the image dropped) */
    uint32_t* key = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t* off = (uint32_t*)calloc(pixels + 2, sizeof(uint32_t));
    uint32_t* idx = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; i++) {
        float cx = coords[2 * (size_t)i], cy = coords[2 * (size_t)i + 1];
        uint32_t k = pixels;
        if (cx >= 0.0f && cx < (float)w && cy >= 0.0f && cy < (float)h)
            k = (uint32_t)truncf(cy) * w + (uint32_t)truncf(cx);
        key[i] = k;
        off[k + 1]++;
    }
    for (uint32_t p = 0; p <= pixels; p++) off[p + 1] += off[p];
    uint32_t* fill = (uint32_t*)malloc(sizeof(uint32_t) * (pixels + 1));
    memcpy(fill, off, sizeof(uint32_t) * (pixels + 1));
    for (uint32_t i = 0; i < n; i++) idx[fill[key[i]]++] = i;
    gather(&f, values, coords, idx, off, 0, w, h, weighted, weights);
    free(fill);
    free(idx);
    free(off);
    free(key);
}

void om_filter_samples_grouped(const om_filter_desc* d, const float* values, const float* coords, uint32_t spp,
                               uint32_t w, uint32_t h, float* weighted, float* weights)
{
    om_filt f;
    filt_make(d, &f);
    gather(&f, values, coords, NULL, NULL, spp, w, h, weighted, weights);
}
