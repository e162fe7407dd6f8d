// mrh_autotune.cpp — picks the traversal launch shape by timing real frames.
//
// Coordinate descent: the traversal tree width first, then the LDS share of
// the traversal stack (at the current grid size), then the persistent grid
// size at the best LDS depth.
// Each candidate renders the caller's frame `reps` times; the fastest median
// wins and is left set on the context.
#include "mrh.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

namespace {

int32_t time_frame(mrh_context* ctx, mrh_scene* scene, const mrh_camera_desc* cam,
                   const mrh_render_params* p, float* d_image, uint32_t reps, double* ms)
{
    std::vector<double> t;
    for (uint32_t r = 0; r < reps; r++) {
        auto a = std::chrono::steady_clock::now();
        int32_t rc = mrh_render(ctx, scene, cam, p, d_image, nullptr);
        if (rc != MRH_OK) return rc;
        t.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count());
    }
    std::sort(t.begin(), t.end());
    *ms = t[t.size() / 2];
    return MRH_OK;
}

}  // namespace

extern "C" int32_t mrh_autotune(mrh_context* ctx, mrh_scene* scene, const mrh_camera_desc* cam,
                                const mrh_render_params* params, float* d_image, uint32_t reps,
                                mrh_tune_result* out)
{
    if (!ctx || !scene || !cam || !params || !d_image || !out) return MRH_ERR_INVALID;
    if (reps == 0) reps = 1;
    std::memset(out, 0, sizeof(*out));
    const int64_t lds_set[] = {0, 8, 12, 16, 24};
    const int64_t bpc_set[] = {4, 8, 16, 32};
    const int64_t width_set[] = {2, 4};
    int64_t best_lds = 16, best_bpc = 8, best_w = 4;
    mrh_tune_get(ctx, "trace_lds_entries", &best_lds);
    mrh_tune_get(ctx, "trace_blocks_per_cu", &best_bpc);
    mrh_tune_get(ctx, "trace_width", &best_w);
    double best = 1e300;
    // warm the scratch slots once
    double ms = 0.0;
    int32_t rc = time_frame(ctx, scene, cam, params, d_image, 1, &ms);
    if (rc != MRH_OK) return rc;
    for (int64_t w : width_set) {
        if ((rc = mrh_tune(ctx, "trace_width", w)) != MRH_OK) return rc;
        if ((rc = time_frame(ctx, scene, cam, params, d_image, reps, &ms)) != MRH_OK) return rc;
        if (out->trials < 32) {
            out->trial_width[out->trials] = (uint32_t)w;
            out->trial_lds[out->trials] = (uint32_t)best_lds;
            out->trial_bpc[out->trials] = (uint32_t)best_bpc;
            out->trial_ms[out->trials] = ms;
            out->trials++;
        }
        if (ms < best) { best = ms; best_w = w; }
    }
    mrh_tune(ctx, "trace_width", best_w);
    for (int64_t lds : lds_set) {
        if ((rc = mrh_tune(ctx, "trace_lds_entries", lds)) != MRH_OK) return rc;
        if ((rc = time_frame(ctx, scene, cam, params, d_image, reps, &ms)) != MRH_OK) return rc;
        if (out->trials < 32) {
            out->trial_width[out->trials] = (uint32_t)best_w;
            out->trial_lds[out->trials] = (uint32_t)lds;
            out->trial_bpc[out->trials] = (uint32_t)best_bpc;
            out->trial_ms[out->trials] = ms;
            out->trials++;
        }
        if (ms < best) { best = ms; best_lds = lds; }
    }
    mrh_tune(ctx, "trace_lds_entries", best_lds);
    for (int64_t bpc : bpc_set) {
        if ((rc = mrh_tune(ctx, "trace_blocks_per_cu", bpc)) != MRH_OK) return rc;
        if ((rc = time_frame(ctx, scene, cam, params, d_image, reps, &ms)) != MRH_OK) return rc;
        if (out->trials < 32) {
            out->trial_width[out->trials] = (uint32_t)best_w;
            out->trial_lds[out->trials] = (uint32_t)best_lds;
            out->trial_bpc[out->trials] = (uint32_t)bpc;
            out->trial_ms[out->trials] = ms;
            out->trials++;
        }
        if (ms < best) { best = ms; best_bpc = bpc; }
    }
    mrh_tune(ctx, "trace_blocks_per_cu", best_bpc);
    out->best_width = (uint32_t)best_w;
    out->best_lds = (uint32_t)best_lds;
    out->best_bpc = (uint32_t)best_bpc;
    out->best_ms = best;
    return MRH_OK;
}
