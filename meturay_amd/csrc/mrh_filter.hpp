// mrh_filter.hpp — reconstruction filters: the filter constants the kernels
// take and the host entry points shared by the path tracer.
#pragma once

#include "mrh_internal.hpp"

namespace mrh {

// One filter with its constants, computed on the host (the reference builds
// its filter functors on the host too).
struct FilterDev {
    uint32_t type;                // mrh_filter_type
    int32_t range, lo, hi;        // FilterRadiusToPixelWH(radius) and its window [lo, hi)
    float radius, rsq, recip;     // radius, radius^2, 1 / radius
    float neg_alpha, gexp;        // Gaussian: -alpha, e^(-alpha r^2)
    float m01[4], m12[4];         // Mitchell cubic coefficients on [0, 1) and [1, 2)
};

// Validates a filter description and computes its constants.
int32_t filter_make(mrh_context* ctx, const mrh_filter_desc* d, FilterDev* out);

// Gathers into weighted (+= sum w v, [pixels][3]) and weights (+= sum w).
// Samples of pixel q are perm[offsets[q] .. offsets[q+1]) when perm is set,
// else samples q*spp .. q*spp + spp - 1 (the path tracer's layout, every
// sample's coordinate inside its own pixel).
int32_t filter_gather(mrh_context* ctx, const FilterDev& f, const float* values, const float2* coords,
                      const uint32_t* perm, const uint32_t* offsets, uint32_t spp, uint32_t w, uint32_t h,
                      float* weighted, float* weights);

// out = weighted / weights per pixel (0 where the weight is 0)
int32_t filter_resolve(mrh_context* ctx, const float* weighted, const float* weights, size_t pixels,
                       float* out);

}  // namespace mrh
