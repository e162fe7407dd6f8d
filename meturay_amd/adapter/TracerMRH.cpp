// TracerMRH.cpp — the drop-in for TracerCUDA: MRay's own plugin interfaces,
// compiled from the reference's headers, implemented over the C-ABI of
// libmrh (include/mrh.h).  MRay loads it exactly as it loads TracerCUDA
// (tConf "TracerDLL": {"name": "TracerMRH", "entry": "GenerateTracerSystem",
// "exit": "DeleteTracerSystem"}, RayLib/SharedLib.h:52-84).
//
// Implements (same vtables, same argument meaning, same error values):
//   TracerSystemI   RayLib/TracerSystemI.h:34-51     (TracerSystemCUDA.cpp:19-141)
//   GPUTracerI      RayLib/GPUTracerI.h:30-72        (GPUTracer.cu:380-520, PathTracer.cu:120-142,
//                                                     WFPGTracer.cu:1355-1435)
//   GPUSceneI       RayLib/GPUSceneI.h:22-67         (GPUSceneJson: file read in LoadScene)
//   factory         TracerCUDA/EntryPoint.h:12-16
//
// What MRay's TracerThread (RayLib/TracerThread.cpp:44-183) calls behaves as
// TracerCUDA's: GenerateScene records the path and LoadScene reads it
// (mrh_scene_file_load: JSON + OBJ; a feature this library cannot render is
// SceneError::NO_LOGIC_FOR_MATERIAL / _PRIMITIVE / _LIGHT / ...);
// GenerateTracer("PathTracer" | "WFPGTracer") reads the reference's option
// set, every option required with its reference type; Finalize sends the
// image portion through TracerCallbacksI::SendImage in the visor's layout
// (pixels in the requested float format, then one float sample count per
// pixel, ImageMemory.cu:157-190).
//
// Differences a caller can see: the GPUSceneI group accessors return empty
// containers (TracerCUDA's device-side groups have no counterpart; MRay's
// host code does not call them); GenerateWork(const GPUCameraI&) reports
// UNABLE_TO_GENERATE_WORK (GPUCameraI is a CUDA device object); surface
// loader DLLs are accepted and not used (OBJ meshes are read directly);
// GPUReconFilterI pointers handed out are this library's own filter objects.
#include "RayLib/AnalyticData.h"
#include "RayLib/GPUTracerI.h"
#include "RayLib/Options.h"
#include "RayLib/SceneError.h"
#include "RayLib/System.h"
#include "RayLib/TracerCallbacksI.h"
#include "RayLib/TracerError.h"
#include "RayLib/TracerSystemI.h"
#include "RayLib/VisorTransform.h"

#include "mrh.h"

#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace {

// mrh_status -> TracerError (the C-ABI message as the extra string)
TracerError FromStatus(int32_t rc, mrh_context* ctx, TracerError::Type otherwise)
{
    if (rc == MRH_OK) return TracerError::OK;
    const std::string msg = ctx ? mrh_last_error(ctx) : "";
    if (rc == MRH_ERR_OOM || rc == MRH_ERR_BUDGET) return TracerError(TracerError::GPU_OUT_OF_MEMORY, msg);
    return TracerError(otherwise, msg);
}

// Reconstruction filter made by GenerateReconFilter (GPUReconFilterBox /
// Tent / Gaussian / Mitchell); handed to MRay as an opaque GPUReconFilterI*.
struct FilterMRH {
    mrh_filter_desc desc;
};

class SceneMRH final : public GPUSceneI {
public:
    SceneMRH(mrh_context* ctx, std::u8string path) : ctx(ctx), path(std::move(path)) {}
    ~SceneMRH() override
    {
        if (scene) mrh_scene_destroy(scene);
    }

    size_t UsedGPUMemory() const override
    {
        uint64_t b = 0;
        if (scene) mrh_scene_bytes(scene, &b);
        return (size_t)b;
    }
    size_t UsedCPUMemory() const override { return sizeof(*this) + cameras.size() * sizeof(mrh_camera_desc); }

    // GPUSceneJson::LoadScene: reads the file and builds the device scene
    SceneError LoadScene(double) override
    {
        const auto t0 = std::chrono::steady_clock::now();
        if (!ctx) return SceneError(SceneError::PRIMITIVE_TYPE_INTERNAL_ERROR, "the tracer system is not initialized");
        if (scene) {
            mrh_scene_destroy(scene);
            scene = nullptr;
        }
        const std::string p(path.begin(), path.end());
        char msg[1024] = {};
        mrh_scene_file* f = nullptr;
        int32_t rc = mrh_scene_file_load(p.c_str(), MRH_SCENE_FILE_STRICT, &f, msg, sizeof(msg));
        if (rc == MRH_ERR_UNSUPPORTED) {
            // the group the reference has no logic for, named in msg
            const std::string m = msg;
            auto has = [&](const char* a, const char* b) {
                return m.find(a) != std::string::npos || m.find(b) != std::string::npos;
            };
            SceneError::Type t = SceneError::NO_LOGIC_FOR_MATERIAL;
            if (has("primitive", "Primitive")) t = SceneError::NO_LOGIC_FOR_PRIMITIVE;
            else if (has("light", "Light")) t = SceneError::NO_LOGIC_FOR_LIGHT;
            else if (has("camera", "Camera")) t = SceneError::NO_LOGIC_FOR_CAMERA;
            else if (has("transform", "Transform")) t = SceneError::NO_LOGIC_FOR_TRANSFORM;
            else if (has("medium", "Medium")) t = SceneError::NO_LOGIC_FOR_MEDIUM;
            return SceneError(t, m);
        }
        if (rc != MRH_OK) {
            const std::string m = msg;
            const bool missing = m.find("open") != std::string::npos || m.find("not found") != std::string::npos;
            return SceneError(missing ? SceneError::FILE_NOT_FOUND : SceneError::JSON_FILE_PARSE_ERROR, m);
        }
        mrh_scene_desc sd = {};
        mrh_scene_file_desc(f, &sd);
        uint32_t nc = 0;
        mrh_scene_file_cameras(f, nullptr, 0, &nc);
        cameras.resize(nc);
        mrh_scene_file_cameras(f, cameras.data(), nc, &nc);
        n_tris = sd.n;
        n_materials = sd.n_lambert + sd.n_metal + sd.n_dielectric + sd.n_reflect;
        n_lights = sd.n_emissive;
        rc = mrh_scene_create(ctx, &sd, &scene);
        mrh_scene_file_destroy(f);
        if (rc != MRH_OK) {
            scene = nullptr;
            return SceneError(SceneError::PRIMITIVE_TYPE_INTERNAL_ERROR, mrh_last_error(ctx));
        }
        load_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return SceneError::OK;
    }
    // scenes here have no time-varying transforms (the reference's
    // ChangeTime re-evaluates animated transforms only)
    SceneError ChangeTime(double) override
    {
        return scene ? SceneError(SceneError::OK) : SceneError(SceneError::PRIMITIVE_TYPE_INTERNAL_ERROR, "scene not loaded");
    }

    Vector2i MaxMatIds() const override { return Vector2i(MRH_BATCH_REFLECT, (int)n_materials); }
    Vector2i MaxAccelIds() const override { return Vector2i(1, 1); }
    HitKey BaseBoundaryMaterial() const override { return HitKey::InvalidKey; }
    uint32_t HitStructUnionSize() const override { return sizeof(float) * 2; }   // triangle barycentrics
    double MaxSceneTime() const override { return 0.0; }
    uint32_t CameraCount() const override { return (uint32_t)cameras.size(); }

    const NamedList<CPULightGPtr>& Lights() const override { return lights; }
    const NamedList<CPUCameraGPtr>& Cameras() const override { return cams; }
    const NamedList<CPUTransformGPtr>& Transforms() const override { return transforms; }
    const NamedList<CPUMediumGPtr>& Mediums() const override { return mediums; }
    uint16_t BaseMediumIndex() const override { return 0; }
    uint32_t IdentityTransformIndex() const override { return 0; }
    uint32_t BoundaryTransformIndex() const override { return 0; }
    const WorkBatchCreationInfo& WorkBatchInfo() const override { return work; }
    const BoundaryWorkBatchCreationInfo& BoundarWorkBatchInfo() const override { return bwork; }
    const AcceleratorBatchMap& AcceleratorBatchMappings() const override { return accmap; }
    const GPUBaseAccelPtr& BaseAccelerator() const override { return base; }
    const std::map<NameGPUPair, GPUMatGPtr>& MaterialGroups() const override { return matgroups; }
    const NamedList<GPUAccelGPtr>& AcceleratorGroups() const override { return accgroups; }
    const NamedList<GPUPrimGPtr>& PrimitiveGroups() const override { return primgroups; }

    SceneAnalyticData AnalyticData() const override
    {
        SceneAnalyticData a = {};
        a.sceneName = std::string(path.begin(), path.end());
        a.sceneLoadTime = load_s;
        a.sceneUpdateTime = 0.0;
        a.groupCounts.fill(0);
        a.groupCounts[SceneAnalyticData::MATERIAL] = n_materials;
        a.groupCounts[SceneAnalyticData::PRIMITIVE] = n_tris ? 1u : 0u;
        a.groupCounts[SceneAnalyticData::LIGHT] = n_lights;
        a.groupCounts[SceneAnalyticData::CAMERA] = (uint32_t)cameras.size();
        a.groupCounts[SceneAnalyticData::ACCELERATOR] = 1;
        a.accKeyMax = MaxAccelIds();
        a.workKeyMax = MaxMatIds();
        return a;
    }

    mrh_scene* Handle() const { return scene; }
    const mrh_camera_desc* Camera(uint32_t i) const { return i < cameras.size() ? &cameras[i] : nullptr; }

private:
    mrh_context* ctx;
    std::u8string path;
    mrh_scene* scene = nullptr;
    std::vector<mrh_camera_desc> cameras;
    uint32_t n_tris = 0, n_materials = 0, n_lights = 0;
    double load_s = 0.0;
    NamedList<CPULightGPtr> lights;
    NamedList<CPUCameraGPtr> cams;
    NamedList<CPUTransformGPtr> transforms;
    NamedList<CPUMediumGPtr> mediums;
    WorkBatchCreationInfo work;
    BoundaryWorkBatchCreationInfo bwork;
    AcceleratorBatchMap accmap;
    GPUBaseAccelPtr base{nullptr, nullptr};
    std::map<NameGPUPair, GPUMatGPtr> matgroups;
    NamedList<GPUAccelGPtr> accgroups;
    NamedList<GPUPrimGPtr> primgroups;
};

// PathTracer (PathTracer.cu) and WFPGTracer (WFPGTracer.cu) over mrh_tracer
class TracerMRH final : public GPUTracerI {
public:
    TracerMRH(mrh_context* ctx, SceneMRH* scene, const TracerParameters& p, bool guided)
        : ctx(ctx), scene(scene), params(p), guided(guided) {}
    ~TracerMRH() override { Release(); }

    void AttachTracerCallbacks(TracerCallbacksI& cb) override { callbacks = &cb; }

    TracerError Initialize() override
    {
        if (!ctx || !scene || !scene->Handle())
            return TracerError(TracerError::UNABLE_TO_INITIALIZE_TRACER, "no loaded scene");
        return TracerError::OK;
    }

    // PathTracer::SetOptions (PathTracer.cu:120-142) and WFPGTracer::SetOptions
    // (WFPGTracer.cu:1355-1435): every option is required, with its type.
    TracerError SetOptions(const OptionsI& o) override
    {
        TracerError e = TracerError::OK;
        int32_t samples = 1;
        uint32_t depth = 10, rr = 3;
        bool nee = true, mis = false;
        std::string sampler;
        if ((e = o.GetInt(samples, "Samples")) != TracerError::OK) return e;
        if ((e = o.GetUInt(depth, "MaxDepth")) != TracerError::OK) return e;
        if ((e = o.GetBool(nee, "NextEventEstimation")) != TracerError::OK) return e;
        if ((e = o.GetUInt(rr, "RussianRouletteStart")) != TracerError::OK) return e;
        if ((e = o.GetBool(mis, "DirectLightMIS")) != TracerError::OK) return e;
        if (!guided) {
            bool mega = false;
            if ((e = o.GetBool(mega, "RunAsMegaKernel")) != TracerError::OK) return e;
        }
        if ((e = o.GetString(sampler, "NEESampler")) != TracerError::OK) return e;
        // LightSamplerCommon::StringToLightSamplerType (Tracers.cpp:5-24)
        if (sampler != "Uniform") return TracerError::UNABLE_TO_INITIALIZE_TRACER;
        if (samples < 1) return TracerError(TracerError::UNABLE_TO_INITIALIZE_TRACER, "Samples must be >= 1");
        if (guided) {
            uint32_t level = 10, minLevel = 0, binRays = 512, u = 0;
            float ratio = 0.5f, f = 0.0f;
            bool skipPG = false, purePG = false, b = false;
            std::string mode, s;
            if ((e = o.GetUInt(level, "OctreeLevel")) != TracerError::OK) return e;
            if ((e = o.GetUInt(minLevel, "MinRayBinLevel")) != TracerError::OK) return e;
            if ((e = o.GetUInt(binRays, "BinRayCount")) != TracerError::OK) return e;
            if ((e = o.GetString(mode, "RenderMode")) != TracerError::OK) return e;
            // StringToWFPGRenderMode (WFPGCommon.h:44-71): only "Render" draws the
            // scene; the SVO debug views have no counterpart here
            static const char* modes[] = {"Render", "SVOFalseColor", "SVONormal", "SVOInitialHit",
                                          "SVORadiance", "Binning", "GuidingFactor"};
            bool known = false;
            for (const char* m : modes) known |= (mode == m);
            if (!known) return TracerError::TRACER_INTERNAL_ERROR;
            if (mode != "Render")
                return TracerError(TracerError::UNABLE_TO_INITIALIZE_TRACER, "RenderMode " + mode + " is a debug view");
            if ((e = o.GetUInt(u, "SVODebugRenderIter")) != TracerError::OK) return e;
            if ((e = o.GetUInt(u, "SVORenderLevel")) != TracerError::OK) return e;
            if ((e = o.GetString(s, "InitialSVO")) != TracerError::OK) return e;
            float alpha = 1.0f;
            bool product = true;
            if ((e = o.GetFloat(alpha, "RFieldFilterAlpha")) != TracerError::OK) return e;
            if ((e = o.GetBool(skipPG, "SkipPG")) != TracerError::OK) return e;
            if ((e = o.GetBool(purePG, "PurePG")) != TracerError::OK) return e;
            if ((e = o.GetFloat(ratio, "BXDF-GuideMISRatio")) != TracerError::OK) return e;
            if ((e = o.GetBool(product, "DoProductPathGuiding")) != TracerError::OK) return e;
            if ((e = o.GetBool(b, "OptiXTraceSVO")) != TracerError::OK) return e;
            if ((e = o.GetBool(b, "UseSceneAccelerator")) != TracerError::OK) return e;
            if ((e = o.GetUInt(u, "OptixTraceBufferSize")) != TracerError::OK) return e;
            if ((e = o.GetBool(b, "OmitDirectRadiance")) != TracerError::OK) return e;
            if ((e = o.GetBool(b, "AdjustParallax")) != TracerError::OK) return e;
            if ((e = o.GetUInt(u, "GuideBounceAmount")) != TracerError::OK) return e;
            if ((e = o.GetBool(b, "WeightedCombineSamples")) != TracerError::OK) return e;
            if ((e = o.GetBool(b, "PGDumpDataStruct")) != TracerError::OK) return e;
            if ((e = o.GetUInt(u, "PGDataStructDumpIntervalExp")) != TracerError::OK) return e;
            if ((e = o.GetString(s, "PGDataStructDumpName")) != TracerError::OK) return e;
            if ((e = o.GetFloat(f, "GuidingDisableRatio")) != TracerError::OK) return e;
            if (level < 1 || level > 10)
                return TracerError(TracerError::UNABLE_TO_INITIALIZE_TRACER, "OctreeLevel must be in [1, 10]");
            if (minLevel > 10)
                return TracerError(TracerError::UNABLE_TO_INITIALIZE_TRACER, "MinRayBinLevel must be in [0, 10]");
            if (!(alpha > 0.0f))
                return TracerError(TracerError::UNABLE_TO_INITIALIZE_TRACER, "RFieldFilterAlpha must be > 0");
            octreeLevel = level;
            binRayCount = binRays;
            minRayBinLevel = minLevel;
            filterAlpha = alpha;
            productPG = product;
            // SkipPG: BxDF sampling only; PurePG: guide sampling only
            misRatio = skipPG ? 0.0f : (purePG ? 1.0f : ratio);
        }
        opt.samples = (uint32_t)samples;
        opt.depth = depth;
        opt.rr = rr;
        opt.nee = nee;
        opt.mis = mis;
        dirty = true;
        return TracerError::OK;
    }

    void SetParameters(const TracerParameters& p) override
    {
        if (p.seed != params.seed) dirty = true;
        params = p;
    }

    // PathTracer::AskOptions (PathTracer.cu:356-371)
    void AskOptions() override
    {
        if (!callbacks) return;
        VariableList list;
        list.emplace("Samples", OptionVariable((int64_t)opt.samples));
        list.emplace("MaxDepth", OptionVariable((int64_t)opt.depth));
        list.emplace("NextEventEstimation", OptionVariable(opt.nee));
        list.emplace("DirectLightMIS", OptionVariable(opt.mis));
        list.emplace("RussianRouletteStart", OptionVariable((int64_t)opt.rr));
        list.emplace("NEESampler", OptionVariable(std::string("Uniform")));
        if (guided) {
            list.emplace("OctreeLevel", OptionVariable((int64_t)octreeLevel));
            list.emplace("BinRayCount", OptionVariable((int64_t)binRayCount));
            list.emplace("BXDF-GuideMISRatio", OptionVariable(misRatio));
            list.emplace("MinRayBinLevel", OptionVariable((int64_t)minRayBinLevel));
            list.emplace("RFieldFilterAlpha", OptionVariable(filterAlpha));
            list.emplace("DoProductPathGuiding", OptionVariable(productPG));
        } else {
            list.emplace("RunAsMegaKernel", OptionVariable(false));
        }
        callbacks->SendCurrentOptions(::Options(std::move(list)));
    }
    void AskParameters() override
    {
        if (callbacks) callbacks->SendCurrentParameters(params);
    }

    void GenerateWork(uint32_t cameraId) override
    {
        const mrh_camera_desc* c = scene ? scene->Camera(cameraId) : nullptr;
        if (!c) {
            Report(TracerError(TracerError::UNABLE_TO_GENERATE_WORK, "no scene camera " + std::to_string(cameraId)));
            return;
        }
        StartFrame(*c);
    }
    // GPUTracer::GenerateCameraWithTransform (GPUTracer.cu:389-414): the
    // scene camera with the visor's position, gaze point and up vector
    void GenerateWork(const VisorTransform& t, uint32_t cameraId) override
    {
        const mrh_camera_desc* c = scene ? scene->Camera(cameraId) : nullptr;
        if (!c) {
            Report(TracerError(TracerError::UNABLE_TO_GENERATE_WORK, "no scene camera " + std::to_string(cameraId)));
            return;
        }
        mrh_camera_desc cam = *c;
        for (int a = 0; a < 3; a++) {
            cam.position[a] = t.position[a];
            cam.gaze[a] = t.gazePoint[a];
            cam.up[a] = t.up[a];
        }
        StartFrame(cam);
    }
    void GenerateWork(const GPUCameraI&) override
    {
        Report(TracerError(TracerError::UNABLE_TO_GENERATE_WORK,
                           "device-side GPUCameraI objects belong to TracerCUDA; use a scene camera or a VisorTransform"));
    }

    bool Render() override
    {
        if (!tracer || crashed) return false;
        int32_t more = 0;
        int32_t rc = mrh_tracer_render(tracer, &more);
        if (rc != MRH_OK) {
            Report(FromStatus(rc, ctx, TracerError::TRACER_INTERNAL_ERROR));
            crashed = true;
            return false;
        }
        return more != 0;
    }

    // GPUTracer::Finalize (GPUTracer.cu:466-500): the image portion and its
    // per-pixel sample counts to the visor.  The library keeps the running
    // mean of every frame since the last ResetImage, so the portion is reset
    // on the visor and sent whole: value = mean x samples, count = samples.
    void Finalize() override
    {
        if (!tracer || crashed) return;
        int32_t rc = mrh_tracer_finalize(tracer, dImage);
        if (rc != MRH_OK) return (void)Report(FromStatus(rc, ctx, TracerError::TRACER_INTERNAL_ERROR));
        const double ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - frameStart).count();
        if (!callbacks) return;
        std::vector<float> rgb((size_t)width * height * 3);
        rc = mrh_memcpy(ctx, rgb.data(), dImage, rgb.size() * sizeof(float));
        if (rc != MRH_OK) return (void)Report(FromStatus(rc, ctx, TracerError::TRACER_INTERNAL_ERROR));
        uint32_t frames = 0;
        mrh_tracer_stats(tracer, nullptr, &frames);
        const uint32_t spp = opt.samples * opt.samples;
        const float count = (float)((uint64_t)frames * spp);
        const int ch = (format == PixelFormat::RGBA_FLOAT) ? 4 : 3;
        const Vector2i s = PortionStart(), e = PortionEnd();
        const size_t pw = (size_t)(e[0] - s[0]), ph = (size_t)(e[1] - s[1]);
        const size_t n = pw * ph;
        const size_t offset = n * ch * sizeof(float);
        std::vector<Byte> data(offset + n * sizeof(float));
        float* px = reinterpret_cast<float*>(data.data());
        float* cnt = reinterpret_cast<float*>(data.data() + offset);
        for (size_t y = 0; y < ph; y++)
            for (size_t x = 0; x < pw; x++) {
                const size_t src = ((size_t)(s[1] + y) * width + (size_t)(s[0] + x)) * 3;
                const size_t dst = y * pw + x;
                for (int c = 0; c < 3; c++) px[dst * ch + c] = rgb[src + c] * count;
                if (ch == 4) px[dst * ch + 3] = count;
                cnt[dst] = count;
            }
        callbacks->SendImageSectionReset(s, e);
        callbacks->SendImage(std::move(data), format, offset, spp, s, e);
        mrh_render_stats st = {};
        mrh_tracer_stats(tracer, &st, nullptr);
        TracerAnalyticData a = {};
        a.iterationTime = (float)ms;
        a.throughput = ms > 0.0 ? (double)width * height * spp / (ms * 1e3) : 0.0;
        a.throughputSuffix = "M paths/s";
        a.workPerPixel = frames ? (double)st.segments / ((double)width * height * frames) : 0.0;
        a.workPerPixelSuffix = "rays/pixel";
        a.usedGPUMemoryMiB = (double)TotalGPUMemoryUsed() / (1024.0 * 1024.0);
        a.totalGPUMemoryMiB = a.usedGPUMemoryMiB;
        callbacks->SendAnalyticData(a);
    }

    // ImageMemory::SetPixelFormat: the float formats the tracer writes
    void SetImagePixelFormat(PixelFormat f) override
    {
        if (f != PixelFormat::RGBA_FLOAT && f != PixelFormat::RGB_FLOAT) {
            Report(TracerError::IMEM_UNKNOWN_PIXEL_FORMAT);
            return;
        }
        format = f;
    }
    void ReportionImage(Vector2i start, Vector2i end) override
    {
        start_ = start;
        end_ = end;
    }
    void ResizeImage(Vector2i resolution) override
    {
        if (resolution[0] <= 0 || resolution[1] <= 0) {
            Report(TracerError(TracerError::UNABLE_TO_GENERATE_WORK, "empty resolution"));
            return;
        }
        if ((uint32_t)resolution[0] != width || (uint32_t)resolution[1] != height) dirty = true;
        width = (uint32_t)resolution[0];
        height = (uint32_t)resolution[1];
    }
    // GPUTracer::ResetImage (GPUTracer.cu:437-449)
    void ResetImage() override
    {
        if (tracer) mrh_tracer_reset_image(tracer);
        if (callbacks) callbacks->SendImageSectionReset(PortionStart(), PortionEnd());
    }
    void AttachReconFilter(GPUReconFilterI* f) override
    {
        filter = reinterpret_cast<FilterMRH*>(f);
        dirty = true;
    }
    size_t TotalGPUMemoryUsed() const override
    {
        mrh_render_params rp = RenderParams();
        uint64_t b = 0;
        mrh_render_bytes(&rp, &b);
        return (size_t)b + (scene ? scene->UsedGPUMemory() : 0);
    }

private:
    mrh_render_params RenderParams() const
    {
        mrh_render_params rp = {};
        rp.width = width;
        rp.height = height;
        rp.samples = opt.samples;
        rp.max_depth = opt.depth;
        rp.rr_start = opt.rr;
        rp.nee = opt.nee ? 1u : 0u;
        rp.direct_light_mis = opt.mis ? 1u : 0u;
        rp.seed = params.seed;
        return rp;
    }

    Vector2i PortionStart() const
    {
        return Vector2i(std::max(0, std::min(start_[0], (int)width)), std::max(0, std::min(start_[1], (int)height)));
    }
    Vector2i PortionEnd() const
    {
        const Vector2i s = PortionStart();
        return Vector2i(std::max(s[0], std::min(end_[0], (int)width)), std::max(s[1], std::min(end_[1], (int)height)));
    }

    // (re)creates the device tracer when the image size, options, seed or
    // filter changed, then generates the frame's camera rays
    void StartFrame(const mrh_camera_desc& cam)
    {
        if (crashed) return;
        if (!scene || !scene->Handle()) {
            Report(TracerError(TracerError::UNABLE_TO_GENERATE_WORK, "no loaded scene"));
            return;
        }
        if (dirty || !tracer) {
            Release();
            mrh_render_params rp = RenderParams();
            int32_t rc = mrh_tracer_create(ctx, scene->Handle(), &rp, &tracer);
            if (rc == MRH_OK && filter) rc = mrh_tracer_set_filter(tracer, &filter->desc);
            if (rc == MRH_OK && guided) {
                // the reference's own guide (voxel octree, per-bin cone-traced
                // fields): its payload learns over the first 16 frames (the
                // vertices of frame k guide frame k + 1) and then stays fixed
                mrh_wfpg_params wp = {};
                wp.iterations = 16;
                wp.svo_depth = octreeLevel;
                wp.guide_ratio = misRatio;
                wp.min_bin_level = minRayBinLevel;
                wp.bin_ray_count = binRayCount;
                wp.filter_alpha = filterAlpha;
                wp.boundary_radiance = 0.0f;
                wp.product = productPG ? 1u : 0u;
                rc = mrh_tracer_set_wfpg(tracer, &wp);
            }
            void* img = nullptr;
            if (rc == MRH_OK) rc = mrh_device_alloc(ctx, &img, (size_t)width * height * 3 * sizeof(float));
            dImage = (float*)img;
            if (rc != MRH_OK) {
                Report(FromStatus(rc, ctx, TracerError::UNABLE_TO_INITIALIZE_TRACER));
                Release();
                return;
            }
            dirty = false;
        }
        frameStart = std::chrono::steady_clock::now();
        int32_t rc = mrh_tracer_generate_work(tracer, &cam);
        if (rc != MRH_OK) Report(FromStatus(rc, ctx, TracerError::UNABLE_TO_GENERATE_WORK));
    }

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

    struct Opt {
        uint32_t samples = 1, depth = 10, rr = 3;
        bool nee = true, mis = false;
    };

    mrh_context* ctx;
    SceneMRH* scene;
    TracerParameters params;
    bool guided;
    Opt opt;
    uint32_t octreeLevel = 10, binRayCount = 512, minRayBinLevel = 5;
    float misRatio = 0.5f, filterAlpha = 1.0f;
    bool productPG = true;
    TracerCallbacksI* callbacks = nullptr;
    FilterMRH* filter = nullptr;
    mrh_tracer* tracer = nullptr;
    float* dImage = nullptr;
    uint32_t width = 256, height = 256;
    Vector2i start_ = Zero2i, end_ = BaseConstants::IMAGE_MAX_SIZE;
    PixelFormat format = PixelFormat::RGBA_FLOAT;
    bool dirty = true, crashed = false;
    std::chrono::steady_clock::time_point frameStart;
};

void DestroyTracer(GPUTracerI* t) { delete t; }

// TracerSystemCUDA (TracerSystemCUDA.cpp:19-141)
class TracerSystemMRH final : public TracerSystemI {
public:
    ~TracerSystemMRH() override
    {
        scene.reset();
        if (ctx) mrh_context_destroy(ctx);
    }

    TracerError Initialize(const std::vector<SurfaceLoaderSharedLib>&, ScenePartitionerType partitioner) override
    {
        if (!ctx && mrh_context_create(0, &ctx) != MRH_OK) {
            ctx = nullptr;
            return TracerError::UNABLE_TO_INITIALIZE_TRACER;
        }
        if (partitioner != ScenePartitionerType::SINGLE_GPU) return TracerError::UNKNOWN_SCENE_PARTITIONER_TYPE;
        return TracerError::OK;
    }

    void ClearScene() override { scene.reset(); }

    void GenerateScene(GPUSceneI*& newScene, const std::u8string& scenePath, SceneLoadFlags) override
    {
        scene = std::make_unique<SceneMRH>(ctx, scenePath);
        newScene = scene.get();
    }

    TracerError GenerateTracer(GPUTracerPtr& tracer, const TracerParameters& params, const Options& opts,
                               const std::string& tracerType) override
    {
        if (!scene || (tracerType != "PathTracer" && tracerType != "WFPGTracer"))
            return TracerError::NO_LOGIC_FOR_TRACER;
        tracer = GPUTracerPtr(new TracerMRH(ctx, scene.get(), params, tracerType == "WFPGTracer"), &DestroyTracer);
        TracerError e = tracer->SetOptions(opts);
        if (e != TracerError::OK) {
            tracer = nullptr;
            return e;
        }
        return TracerError::OK;
    }

    // TracerSystemCUDA::GenerateReconFilter (TracerSystemCUDA.cpp:109-141)
    // with GPUReconFilter{Box,Tent,Gaussian,Mitchell}'s options
    TracerError GenerateReconFilter(GPUReconFilterI*& filterPtr, const Options& o) override
    {
        TracerError err = TracerError::OK;
        std::string type;
        if ((err = o.GetString(type, "type")) != TracerError::OK)
            return TracerError(err, "Filter Options has to have \"type\" field");
        float radius = 0.0f;
        if ((err = o.GetFloat(radius, "radius")) != TracerError::OK)
            return TracerError(err, "Filter Options has to have \"radius\" field");
        filter.reset();
        static const char* names[] = {"Box", "Tent", "Gaussian", "Mitchell"};
        uint32_t k = 0;
        while (k < 4 && type != names[k]) k++;
        if (k == 4) return TracerError(TracerError::TRACER_INTERNAL_ERROR, "Unable to find generator for filter " + type);
        mrh_filter_desc d = {};
        d.type = k;
        d.radius = radius;
        if (k == MRH_FILTER_GAUSSIAN && (err = o.GetFloat(d.alpha, "alpha")) != TracerError::OK)
            return TracerError(TracerError::TRACER_INTERNAL_ERROR, "Gaussian Filter parameter \"alpha\" not found");
        if (k == MRH_FILTER_MITCHELL) {
            if ((err = o.GetFloat(d.b, "b")) != TracerError::OK || (err = o.GetFloat(d.c, "c")) != TracerError::OK)
                return TracerError(TracerError::TRACER_INTERNAL_ERROR, "Mitchell Filter parameters \"b\", \"c\" not found");
        }
        if (!(d.radius >= 0.0f) || d.radius > 16.0f || (k != MRH_FILTER_BOX && d.radius == 0.0f))
            return TracerError(TracerError::TRACER_INTERNAL_ERROR, "bad filter radius");
        filter = std::make_unique<FilterMRH>(FilterMRH{d});
        filterPtr = reinterpret_cast<GPUReconFilterI*>(filter.get());
        return TracerError::OK;
    }

private:
    mrh_context* ctx = nullptr;
    std::unique_ptr<SceneMRH> scene;
    std::unique_ptr<FilterMRH> filter;
};

}  // namespace

extern "C" MRAY_DLL_EXPORT TracerSystemI* GenerateTracerSystem() { return new TracerSystemMRH(); }

extern "C" MRAY_DLL_EXPORT void DeleteTracerSystem(TracerSystemI* system) { delete system; }
