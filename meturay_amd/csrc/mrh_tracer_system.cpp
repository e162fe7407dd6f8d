// mrh_tracer_system.cpp — the TracerSystemI / GPUTracerI implementation over
// the C-ABI (host code only; every GPU operation goes through mrh.h).
#include "mrh_tracer.hpp"

#include <chrono>
#include <cstring>

namespace mrh {

TracerError TracerError::FromStatus(int32_t status, const char* message)
{
    std::string m = message ? message : "";
    switch (status) {
    case MRH_OK: return TracerError(OK);
    case MRH_ERR_OOM: return TracerError(GPU_OUT_OF_MEMORY, m);
    case MRH_ERR_UNSUPPORTED: return TracerError(UNABLE_TO_INITIALIZE_TRACER, m);
    case MRH_ERR_BUDGET: return TracerError(GPU_OUT_OF_MEMORY, m);
    default: return TracerError(TRACER_INTERNAL_ERROR, m);
    }
}

namespace {

class SceneMRH final : public GPUSceneI {
public:
    SceneMRH(mrh_scene* h, size_t prims, std::vector<mrh_camera_desc> cams)
        : h(h), prims(prims), cams(std::move(cams)) {}
    ~SceneMRH() override { mrh_scene_destroy(h); }
    size_t PrimitiveCount() const override { return prims; }
    uint32_t CameraCount() const override { return (uint32_t)cams.size(); }
    const mrh_camera_desc& Camera(uint32_t id) const override { return cams.at(id); }
    size_t UsedGPUMemory() const override
    {
        uint64_t b = 0;
        mrh_scene_bytes(h, &b);
        return (size_t)b;
    }
    mrh_scene* Handle() const override { return h; }

private:
    mrh_scene* h;
    size_t prims;
    std::vector<mrh_camera_desc> cams;
};

// GPUReconFilter{Box,Tent,Gaussian,Mitchell} (TracerCUDA/GPUReconFilter*.h)
class ReconFilterMRH final : public GPUReconFilterI {
public:
    ReconFilterMRH(const char* type, const mrh_filter_desc& d) : type(type), desc(d) {}
    const char* Type() const override { return type; }
    const mrh_filter_desc& Desc() const override { return desc; }

private:
    const char* type;
    mrh_filter_desc desc;
};

// PathTracer (TracerCUDA/PathTracer.cu) and, with `guided`, WFPGTracer
// (TracerCUDA/WFPGTracer.cu) over mrh_tracer
class PathTracerMRH final : public GPUTracerI {
public:
    // "NextEventEstimation" defaults to on, as in PathTracer.h:26-33 and
    // WFPGTracer.h:62.
    PathTracerMRH(mrh_context* ctx, SceneMRH* scene, const TracerParameters& p, bool guided)
        : ctx(ctx), scene(scene), params(p), nee(true), guided(guided) {}
    ~PathTracerMRH() override { Release(); }

    void AttachTracerCallbacks(TracerCallbacksI& cb) override { callbacks = &cb; }

    void AttachReconFilter(GPUReconFilterI* f) override
    {
        filter = f;
        if (tracer) {
            int32_t rc = mrh_tracer_set_filter(tracer, f ? &f->Desc() : nullptr);
            if (rc != MRH_OK) Report(TracerError::FromStatus(rc, mrh_last_error(ctx)));
        }
    }

    TracerError Initialize() override
    {
        Release();
        mrh_render_params rp = {};
        rp.width = width;
        rp.height = height;
        rp.samples = samples;
        rp.max_depth = maxDepth;
        rp.rr_start = rrStart;
        rp.nee = nee ? 1u : 0u;
        rp.direct_light_mis = directLightMIS ? 1u : 0u;
        rp.seed = params.seed;
        int32_t rc = mrh_tracer_create(ctx, scene->Handle(), &rp, &tracer);
        if (rc != MRH_OK) return Report(TracerError::FromStatus(rc, mrh_last_error(ctx)));
        if (filter) {
            rc = mrh_tracer_set_filter(tracer, &filter->Desc());
            if (rc != MRH_OK) return Report(TracerError::FromStatus(rc, mrh_last_error(ctx)));
        }
        if (guided) {
            mrh_guide_params gp = {};
            gp.iterations = learnIterations;
            gp.svo_depth = octreeLevel;
            gp.guide_ratio = (float)misRatio;
            gp.min_samples = binRayCount;
            rc = mrh_tracer_set_guide(tracer, &gp, nullptr);
            if (rc != MRH_OK) return Report(TracerError::FromStatus(rc, mrh_last_error(ctx)));
        }
        void* img = nullptr;
        if (mrh_device_alloc(ctx, &img, (size_t)width * height * 3 * sizeof(float)) != MRH_OK)
            return Report(TracerError(TracerError::GPU_OUT_OF_MEMORY, "image buffer"));
        dImage = (float*)img;
        return TracerError::OK;
    }

    // PathTracer.h:17-33 option names; "Samples" = strata per axis.  The
    // guided tracer adds WFPGTracer.h:25-96's "OctreeLevel", "BinRayCount"
    // (guide lookups stop above octree nodes with fewer samples) and
    // "BXDF-GuideMISRatio", plus "LearnIterations": the number of frames
    // that learn (the octree of frame k guides frame k + 1).
    TracerError SetOptions(const Options& o) override
    {
        int64_t v;
        bool b;
        double dv;
        if (o.Has("OctreeLevel")) {
            TracerError e = o.GetInt(v, "OctreeLevel");
            if (e != TracerError::OK) return Report(e);
            octreeLevel = (uint32_t)v;
        }
        if (o.Has("BinRayCount")) {
            TracerError e = o.GetInt(v, "BinRayCount");
            if (e != TracerError::OK) return Report(e);
            binRayCount = (uint32_t)v;
        }
        if (o.Has("LearnIterations")) {
            TracerError e = o.GetInt(v, "LearnIterations");
            if (e != TracerError::OK) return Report(e);
            learnIterations = (uint32_t)v;
        }
        if (o.Has("BXDF-GuideMISRatio")) {
            TracerError e = o.GetDouble(dv, "BXDF-GuideMISRatio");
            if (e != TracerError::OK) return Report(e);
            misRatio = dv;
        }
        if (o.Has("Samples")) {
            TracerError e = o.GetInt(v, "Samples");
            if (e != TracerError::OK) return Report(e);
            samples = (uint32_t)v;
        }
        if (o.Has("MaxDepth")) {
            TracerError e = o.GetInt(v, "MaxDepth");
            if (e != TracerError::OK) return Report(e);
            maxDepth = (uint32_t)v;
        }
        if (o.Has("RussianRouletteStart")) {
            TracerError e = o.GetInt(v, "RussianRouletteStart");
            if (e != TracerError::OK) return Report(e);
            rrStart = (uint32_t)v;
        }
        if (o.Has("NextEventEstimation")) {
            TracerError e = o.GetBool(b, "NextEventEstimation");
            if (e != TracerError::OK) return Report(e);
            nee = b;
        }
        if (o.Has("DirectLightMIS")) {
            TracerError e = o.GetBool(b, "DirectLightMIS");
            if (e != TracerError::OK) return Report(e);
            directLightMIS = b;
        }
        if (o.Has("NEESampler")) {
            // LightSamplerCommon::StringToLightSamplerType (Tracers.cpp:10-24):
            // "Uniform" is the one light sampler
            std::string name;
            TracerError e = o.GetString(name, "NEESampler");
            if (e != TracerError::OK) return Report(e);
            if (name != "Uniform")
                return Report(TracerError(TracerError::UNABLE_TO_INITIALIZE_TRACER, "unknown NEESampler " + name));
        }
        options = o;
        return TracerError::OK;
    }

    void SetParameters(const TracerParameters& p) override { params = p; }
    void AskOptions() override {}
    void AskParameters() override {}

    void GenerateWork(uint32_t cameraId) override
    {
        if (cameraId >= scene->CameraCount()) {
            Report(TracerError(TracerError::UNABLE_TO_GENERATE_WORK, "camera id out of range"));
            return;
        }
        GenerateWork(scene->Camera(cameraId));
    }

    void GenerateWork(const mrh_camera_desc& cam) override
    {
        if (!tracer && Initialize() != TracerError::OK) return;
        frameStart = std::chrono::steady_clock::now();
        int32_t rc = mrh_tracer_generate_work(tracer, &cam);
        if (rc != MRH_OK) Report(TracerError::FromStatus(rc, mrh_last_error(ctx)));
    }

    bool Render() override
    {
        if (!tracer) return false;
        int32_t more = 0;
        int32_t rc = mrh_tracer_render(tracer, &more);
        if (rc != MRH_OK) {
            Report(TracerError::FromStatus(rc, mrh_last_error(ctx)));
            return false;
        }
        return more != 0;
    }

    void Finalize() override
    {
        if (!tracer) return;
        int32_t rc = mrh_tracer_finalize(tracer, dImage);
        if (rc != MRH_OK) {
            Report(TracerError::FromStatus(rc, mrh_last_error(ctx)));
            return;
        }
        double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - frameStart).count();
        if (!callbacks) return;
        std::vector<float> rgb;
        if (GetImage(rgb) != TracerError::OK) return;
        uint32_t frames = 0;
        mrh_tracer_stats(tracer, nullptr, &frames);
        callbacks->SendImage(rgb.data(), width, height, frames);
        const double paths = (double)width * height * samples * samples;
        callbacks->SendAnalyticData(paths / (ms * 1e3), ms);
    }

    void ResizeImage(uint32_t w, uint32_t h) override
    {
        width = w;
        height = h;
        if (tracer) Initialize();
    }

    void ResetImage() override
    {
        if (tracer) mrh_tracer_reset_image(tracer);
    }

    size_t TotalGPUMemoryUsed() const override
    {
        mrh_render_params rp = {};
        rp.width = width;
        rp.height = height;
        rp.samples = samples;
        uint64_t b = 0;
        mrh_render_bytes(&rp, &b);
        return (size_t)b + scene->UsedGPUMemory();
    }

    TracerError GetImage(std::vector<float>& rgb) const override
    {
        if (!dImage) return TracerError(TracerError::NO_LOGIC_SET, "no image yet");
        rgb.resize((size_t)width * height * 3);
        int32_t rc = mrh_memcpy(ctx, rgb.data(), dImage, rgb.size() * sizeof(float));
        return TracerError::FromStatus(rc, rc ? mrh_last_error(ctx) : nullptr);
    }

private:
    void Release()
    {
        if (tracer) mrh_tracer_destroy(tracer);
        tracer = nullptr;
        if (dImage) mrh_device_free(ctx, dImage);
        dImage = nullptr;
    }

    TracerError Report(TracerError e)
    {
        if (e != TracerError::OK && callbacks) callbacks->SendError(e);
        return e;
    }

    mrh_context* ctx;
    SceneMRH* scene;
    TracerParameters params;
    Options options;
    TracerCallbacksI* callbacks = nullptr;
    GPUReconFilterI* filter = nullptr;
    mrh_tracer* tracer = nullptr;
    float* dImage = nullptr;
    uint32_t width = 256, height = 256, samples = 1, maxDepth = 10, rrStart = 3;
    bool nee = true;
    bool guided = false;
    bool directLightMIS = false;
    uint32_t octreeLevel = 10, binRayCount = 512, learnIterations = 4;
    double misRatio = 0.5;
    std::chrono::steady_clock::time_point frameStart;
};

class TracerSystemMRH final : public TracerSystemI {
public:
    ~TracerSystemMRH() override
    {
        ClearScene();
        if (ctx) mrh_context_destroy(ctx);
    }

    TracerError Initialize(int32_t device) override
    {
        if (ctx) return TracerError::OK;
        int32_t rc = mrh_context_create(device, &ctx);
        if (rc != MRH_OK) return TracerError(TracerError::UNABLE_TO_INITIALIZE_TRACER, "no usable GPU");
        return TracerError::OK;
    }

    void ClearScene() override { scene.reset(); }

    SceneError GenerateScene(GPUSceneI*& out, const SceneDescription& d) override
    {
        out = nullptr;
        if (!ctx) return SceneError(SceneError::INTERNAL_ERROR, "tracer system not initialized");
        if (d.accelerator != "LBVH" && d.accelerator != "BasicBVH")
            return SceneError(SceneError::ACCELERATOR_TYPE_UNKNOWN, d.accelerator);
        const size_t n = d.hitKeys.size();
        if (n == 0 || d.triangles.size() != n * 9) return SceneError(SceneError::EMPTY_SCENE);
        mrh_scene_desc sd = {};
        sd.tris = d.triangles.data();
        sd.prim_key = d.hitKeys.data();
        sd.n = (uint32_t)n;
        sd.lambert = d.lambert.data();
        sd.n_lambert = (uint32_t)(d.lambert.size() / 3);
        sd.metal = d.metal.data();
        sd.n_metal = (uint32_t)(d.metal.size() / 10);
        sd.dielectric = d.dielectric.data();
        sd.n_dielectric = (uint32_t)(d.dielectric.size() / 4);
        sd.emissive = d.emissive.data();
        sd.n_emissive = (uint32_t)(d.emissive.size() / 3);
        sd.reflect = d.reflect.data();
        sd.n_reflect = (uint32_t)(d.reflect.size() / 4);
        sd.morton_bits = d.mortonBits;
        ClearScene();
        mrh_scene* h = nullptr;
        int32_t rc = mrh_scene_create(ctx, &sd, &h);
        if (rc == MRH_ERR_INVALID) return SceneError(SceneError::INVALID_HIT_KEY, mrh_last_error(ctx));
        if (rc == MRH_ERR_OOM) return SceneError(SceneError::GPU_OUT_OF_MEMORY, mrh_last_error(ctx));
        if (rc != MRH_OK) return SceneError(SceneError::INTERNAL_ERROR, mrh_last_error(ctx));
        scene = std::make_unique<SceneMRH>(h, n, d.cameras);
        out = scene.get();   // non-owning, as TracerSystemCUDA::GenerateScene
        return SceneError::OK;
    }

    SceneError GenerateScene(GPUSceneI*& out, const std::u8string& path) override
    {
        out = nullptr;
        const std::string p(path.begin(), path.end());
        const size_t dot = p.find_last_of('.');
        if (dot == std::string::npos || p.substr(dot) != ".json") return SceneError(SceneError::NO_LOADER, p);
        char msg[1024];
        mrh_scene_file* f = nullptr;
        const int32_t rc = mrh_scene_file_load(p.c_str(), MRH_SCENE_FILE_STRICT, &f, msg, sizeof(msg));
        if (rc != MRH_OK)
            return SceneError(rc == MRH_ERR_UNSUPPORTED ? SceneError::UNSUPPORTED_FEATURE : SceneError::SCENE_FILE_ERROR,
                              msg);
        mrh_scene_desc sd = {};
        mrh_scene_file_desc(f, &sd);
        SceneDescription d;
        d.triangles.assign(sd.tris, sd.tris + (size_t)sd.n * 9);
        d.hitKeys.assign(sd.prim_key, sd.prim_key + sd.n);
        d.lambert.assign(sd.lambert, sd.lambert + (size_t)sd.n_lambert * 3);
        d.metal.assign(sd.metal, sd.metal + (size_t)sd.n_metal * 10);
        d.dielectric.assign(sd.dielectric, sd.dielectric + (size_t)sd.n_dielectric * 4);
        d.emissive.assign(sd.emissive, sd.emissive + (size_t)sd.n_emissive * 3);
        d.reflect.assign(sd.reflect, sd.reflect + (size_t)sd.n_reflect * 4);
        uint32_t nc = 0;
        mrh_scene_file_cameras(f, nullptr, 0, &nc);
        d.cameras.resize(nc);
        mrh_scene_file_cameras(f, d.cameras.data(), nc, &nc);
        mrh_scene_file_destroy(f);
        return GenerateScene(out, d);
    }

    TracerError GenerateReconFilter(GPUReconFilterI*& out, const Options& o) override
    {
        out = nullptr;
        std::string type;
        if (o.GetString(type, GPUReconFilterI::TYPE_OPTION_NAME) != TracerError::OK)
            return TracerError(TracerError::UNABLE_TO_INITIALIZE_TRACER, "Filter Options has to have \"type\" field");
        double radius = 0.0;
        if (o.GetNumber(radius, GPUReconFilterI::RADIUS_OPTION_NAME) != TracerError::OK)
            return TracerError(TracerError::UNABLE_TO_INITIALIZE_TRACER,
                               "Filter Options has to have \"radius\" field");
        static const char* names[] = {"Box", "Tent", "Gaussian", "Mitchell"};
        mrh_filter_desc d = {};
        uint32_t k = 0;
        while (k < 4 && type != names[k]) k++;
        if (k == 4) return TracerError(TracerError::UNABLE_TO_INITIALIZE_TRACER, "unknown filter type " + type);
        d.type = k;
        d.radius = (float)radius;
        double v = 0.0;
        if (k == MRH_FILTER_GAUSSIAN) {
            if (o.GetNumber(v, "alpha") != TracerError::OK)
                return TracerError(TracerError::UNABLE_TO_INITIALIZE_TRACER, "Gaussian Filter parameter \"alpha\" not found");
            d.alpha = (float)v;
        }
        if (k == MRH_FILTER_MITCHELL) {
            if (o.GetNumber(v, "b") != TracerError::OK)
                return TracerError(TracerError::UNABLE_TO_INITIALIZE_TRACER, "Mitchell Filter parameter \"b\" not found");
            d.b = (float)v;
            if (o.GetNumber(v, "c") != TracerError::OK)
                return TracerError(TracerError::UNABLE_TO_INITIALIZE_TRACER, "Mitchell Filter parameter \"c\" not found");
            d.c = (float)v;
        }
        // the checks of filter_make, so a bad filter fails here
        if (!(d.radius >= 0.0f) || d.radius > 16.0f || (k != MRH_FILTER_BOX && d.radius == 0.0f))
            return TracerError(TracerError::UNABLE_TO_INITIALIZE_TRACER, "bad filter radius");
        if (k == MRH_FILTER_GAUSSIAN && !(d.alpha > 0.0f))
            return TracerError(TracerError::UNABLE_TO_INITIALIZE_TRACER, "Gaussian alpha must be positive");
        reconFilter = std::make_unique<ReconFilterMRH>(names[k], d);
        out = reconFilter.get();
        return TracerError::OK;
    }

    TracerError GenerateTracer(GPUTracerPtr& tracer, const TracerParameters& params, const Options& opts,
                               const std::string& type) override
    {
        if (!ctx || !scene) return TracerError(TracerError::NO_LOGIC_SET, "no scene");
        if (type != "PathTracer" && type != "WFPGTracer") return TracerError(TracerError::NO_LOGIC_FOR_TRACER, type);
        auto t = std::make_unique<PathTracerMRH>(ctx, scene.get(), params, type == "WFPGTracer");
        TracerError e = t->SetOptions(opts);
        if (e != TracerError::OK) return e;
        tracer = std::move(t);
        return TracerError::OK;
    }

private:
    mrh_context* ctx = nullptr;
    std::unique_ptr<SceneMRH> scene;
    std::unique_ptr<ReconFilterMRH> reconFilter;
};

}  // namespace
}  // namespace mrh

extern "C" {

mrh::TracerSystemI* GenerateTracerSystem() { return new mrh::TracerSystemMRH(); }

void DeleteTracerSystem(mrh::TracerSystemI* system) { delete system; }

}  // extern "C"
