// mrh_autotune.cpp — picks the launch shape of frame renders by timing real
// frames.
//
// Coordinate descent over five axes, each swept with the others held at
// their current best: the traversal tree width, the LDS share of the
// traversal stack, the persistent traversal grid (blocks per CU), the
// traversal block size, then the rays per tile of the per-block material
// partition.  Each candidate renders the caller's frame `reps` times (at
// least 3) and is scored by the median; the lowest median over all trials
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
    const size_t m = t.size() / 2;
    *ms = (t.size() % 2) ? t[m] : 0.5 * (t[m - 1] + t[m]);
    return MRH_OK;
}

struct Axis {
    const char* knob;
    std::vector<int64_t> values;
};

}  // namespace

extern "C" int32_t mrh_autotune(mrh_context* ctx, mrh_scene* scene, const mrh_camera_desc* cam,
                                const mrh_render_params* params, float* d_image, uint32_t reps,
                                mrh_tune_result* out)
{
    if (!ctx || !scene || !cam || !params || !d_image || !out) return MRH_ERR_INVALID;
    reps = std::max(reps, 3u);
    std::memset(out, 0, sizeof(*out));
    const Axis axes[] = {
        {"trace_width", {2, 4}},
        {"trace_lds_entries", {0, 8, 12, 16, 24}},
        {"trace_blocks_per_cu", {4, 8, 16, 32}},
        {"trace_block", {128, 256, 512}},
        {"shade_tile", {512, 1024, 2048}},
    };
    constexpr int NAX = sizeof(axes) / sizeof(axes[0]);
    int64_t cur[NAX];
    for (int a = 0; a < NAX; a++) {
        int32_t rc = mrh_tune_get(ctx, axes[a].knob, &cur[a]);
        if (rc != MRH_OK) return rc;
    }
    // warm the scratch slots once
    double ms = 0.0, best = 1e300;
    int32_t rc = time_frame(ctx, scene, cam, params, d_image, 1, &ms);
    if (rc != MRH_OK) return rc;
    for (int a = 0; a < NAX; a++) {
        int64_t pick = cur[a];
        for (int64_t v : axes[a].values) {
            if ((rc = mrh_tune(ctx, axes[a].knob, v)) != MRH_OK) return rc;
            if ((rc = time_frame(ctx, scene, cam, params, d_image, reps, &ms)) != MRH_OK) return rc;
            if (out->trials < 32) {
                const uint32_t k = out->trials++;
                cur[a] = v;
                out->trial_width[k] = (uint32_t)cur[0];
                out->trial_lds[k] = (uint32_t)cur[1];
                out->trial_bpc[k] = (uint32_t)cur[2];
                out->trial_block[k] = (uint32_t)cur[3];
                out->trial_tile[k] = (uint32_t)cur[4];
                out->trial_ms[k] = ms;
            }
            if (ms < best) {
                best = ms;
                pick = v;
            }
        }
        cur[a] = pick;
        if ((rc = mrh_tune(ctx, axes[a].knob, pick)) != MRH_OK) return rc;
    }
    out->best_width = (uint32_t)cur[0];
    out->best_lds = (uint32_t)cur[1];
    out->best_bpc = (uint32_t)cur[2];
    out->best_block = (uint32_t)cur[3];
    out->best_tile = (uint32_t)cur[4];
    out->best_ms = best;
    out->reps = reps;
    return MRH_OK;
}
