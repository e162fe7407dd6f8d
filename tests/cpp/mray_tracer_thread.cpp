// mray_tracer_thread.cpp — drives libTracerMRH.so through MRay's own plugin
// interfaces, compiled from the reference's headers, the way MRay does:
// the factory pair is resolved from the DLL by name (SharedLib::GenerateObject,
// RayLib/SharedLib.h:52-84) and the calls follow TracerThread::LoopWork and
// RecreateTracer (RayLib/TracerThread.cpp:44-230).  Images arrive through
// TracerCallbacksI::SendImage and are accumulated as the visor does
// (SendImageSectionReset, then pixel sums and per-pixel sample counts).
//
//   mray_tracer_thread cpu <libTracerMRH.so> <scene.json>
//       checks that need no GPU: the factory, Initialize's error on a machine
//       without a device, tracer option parsing and its errors, filter
//       options and their errors, work without a loaded scene
//   mray_tracer_thread gpu <libTracerMRH.so> <scene.json> <tracer> <w> <h>
//                          <frames> <samples> <seed> <box0|tent1> <out.bin>
//       the full loop: Initialize, GenerateScene + LoadScene, GenerateTracer,
//       callbacks, Initialize, GenerateReconFilter (Box radius 0 or Tent
//       radius 1) +
//       AttachReconFilter, RGBA_FLOAT, ResizeImage, ReportionImage, then
//       `frames` x (GenerateWork(0), while (Render()), Finalize()); writes the
//       visor's image (w*h*3 floats, bottom row first) to out.bin
//
// Prints one line per check; exit status 0 when every check passed.
#include "RayLib/AnalyticData.h"
#include "RayLib/GPUTracerI.h"
#include "RayLib/Options.h"
#include "RayLib/SceneError.h"
#include "RayLib/TracerCallbacksI.h"
#include "RayLib/TracerError.h"
#include "RayLib/TracerSystemI.h"
#include "RayLib/VisorTransform.h"

#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

int fails = 0;
void check(bool ok, const std::string& what)
{
    std::printf("%s %s\n", ok ? "ok" : "FAIL", what.c_str());
    std::fflush(stdout);
    if (!ok) fails++;
}

// The visor side of TracerCallbacksI (VisorGL::AccumulatePortion: running
// pixel sums and sample counts of the image portions sent)
struct Visor : TracerCallbacksI {
    int w = 0, h = 0;
    std::vector<double> sum, cnt;
    std::vector<TracerError::Type> errors;
    uint32_t cameraCount = 0, images = 0;
    Options lastOptions;
    bool gotOptions = false, gotAnalytics = false;
    TracerAnalyticData analytics = {};

    void Resize(int W, int H)
    {
        w = W;
        h = H;
        sum.assign((size_t)w * h * 3, 0.0);
        cnt.assign((size_t)w * h, 0.0);
    }
    void SendCrashSignal() override {}
    void SendLog(const std::string) override {}
    void SendError(TracerError e) override
    {
        std::printf("  tracer error: %s\n", static_cast<std::string>(e).c_str());
        errors.push_back(static_cast<TracerError::Type>(e));
    }
    void SendAnalyticData(TracerAnalyticData a) override
    {
        analytics = a;
        gotAnalytics = true;
    }
    void SendSceneAnalyticData(SceneAnalyticData) override {}
    void SendImageSectionReset(Vector2i start, Vector2i end) override
    {
        for (int y = std::max(0, start[1]); y < std::min(h, end[1]); y++)
            for (int x = std::max(0, start[0]); x < std::min(w, end[0]); x++) {
                const size_t i = (size_t)y * w + x;
                sum[i * 3] = sum[i * 3 + 1] = sum[i * 3 + 2] = 0.0;
                cnt[i] = 0.0;
            }
    }
    void SendImage(const std::vector<Byte> data, PixelFormat f, size_t offset, uint32_t, Vector2i start,
                   Vector2i end) override
    {
        images++;
        const int ch = (f == PixelFormat::RGBA_FLOAT) ? 4 : 3;
        const int pw = end[0] - start[0], ph = end[1] - start[1];
        const float* px = reinterpret_cast<const float*>(data.data());
        const float* c = reinterpret_cast<const float*>(data.data() + offset);
        for (int y = 0; y < ph; y++)
            for (int x = 0; x < pw; x++) {
                const size_t s = (size_t)y * pw + x;
                const size_t d = (size_t)(start[1] + y) * w + (start[0] + x);
                for (int k = 0; k < 3; k++) sum[d * 3 + k] += px[s * ch + k];
                cnt[d] += c[s];
            }
    }
    void SendCurrentOptions(Options o) override
    {
        lastOptions = std::move(o);
        gotOptions = true;
    }
    void SendCurrentParameters(TracerParameters) override {}
    void SendCurrentTransform(VisorTransform) override {}
    void SendCurrentSceneCameraCount(uint32_t n) override { cameraCount = n; }

    std::vector<float> Image() const
    {
        std::vector<float> out((size_t)w * h * 3, 0.0f);
        for (size_t i = 0; i < cnt.size(); i++)
            for (int k = 0; k < 3; k++) out[i * 3 + k] = cnt[i] > 0.0 ? (float)(sum[i * 3 + k] / cnt[i]) : 0.0f;
        return out;
    }
};

VariableList PathTracerOptions(int64_t samples)
{
    VariableList v;
    v.emplace("Samples", OptionVariable(samples));
    v.emplace("MaxDepth", OptionVariable((int64_t)8));
    v.emplace("NextEventEstimation", OptionVariable(false));
    v.emplace("RussianRouletteStart", OptionVariable((int64_t)3));
    v.emplace("DirectLightMIS", OptionVariable(false));
    v.emplace("RunAsMegaKernel", OptionVariable(false));
    v.emplace("NEESampler", OptionVariable(std::string("Uniform")));
    return v;
}

// WFPGTracer's option set (WFPGTracer.cu:1355-1435)
VariableList WFPGOptions(int64_t samples)
{
    VariableList v = PathTracerOptions(samples);
    v.erase("RunAsMegaKernel");
    v.emplace("OctreeLevel", OptionVariable((int64_t)8));
    v.emplace("MinRayBinLevel", OptionVariable((int64_t)4));
    v.emplace("BinRayCount", OptionVariable((int64_t)64));
    v.emplace("RenderMode", OptionVariable(std::string("Render")));
    v.emplace("SVODebugRenderIter", OptionVariable((int64_t)0));
    v.emplace("SVORenderLevel", OptionVariable((int64_t)0));
    v.emplace("InitialSVO", OptionVariable(std::string("")));
    v.emplace("RFieldFilterAlpha", OptionVariable(1.0f));
    v.emplace("SkipPG", OptionVariable(false));
    v.emplace("PurePG", OptionVariable(false));
    v.emplace("BXDF-GuideMISRatio", OptionVariable(0.5f));
    v.emplace("DoProductPathGuiding", OptionVariable(false));   // reproducible path for path
    v.emplace("OptiXTraceSVO", OptionVariable(false));
    v.emplace("UseSceneAccelerator", OptionVariable(false));
    v.emplace("OptixTraceBufferSize", OptionVariable((int64_t)0));
    v.emplace("OmitDirectRadiance", OptionVariable(false));
    v.emplace("AdjustParallax", OptionVariable(false));
    v.emplace("GuideBounceAmount", OptionVariable((int64_t)0));
    v.emplace("WeightedCombineSamples", OptionVariable(false));
    v.emplace("PGDumpDataStruct", OptionVariable(false));
    v.emplace("PGDataStructDumpIntervalExp", OptionVariable((int64_t)0));
    v.emplace("PGDataStructDumpName", OptionVariable(std::string("")));
    v.emplace("GuidingDisableRatio", OptionVariable(0.0f));
    return v;
}

Options FilterOptions(const char* type, float radius)
{
    VariableList v;
    v.emplace("type", OptionVariable(std::string(type)));
    v.emplace("radius", OptionVariable(radius));
    return Options(std::move(v));
}

struct Lib {
    void* h = nullptr;
    TracerSystemI* (*gen)() = nullptr;
    void (*del)(TracerSystemI*) = nullptr;
};

bool Open(const char* path, Lib& L)
{
    L.h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!L.h) {
        std::printf("FAIL dlopen %s: %s\n", path, dlerror());
        return false;
    }
    L.gen = reinterpret_cast<TracerSystemI* (*)()>(dlsym(L.h, "GenerateTracerSystem"));
    L.del = reinterpret_cast<void (*)(TracerSystemI*)>(dlsym(L.h, "DeleteTracerSystem"));
    return L.gen && L.del;
}

int CpuChecks(const Lib& L, const char* scenePath)
{
    TracerSystemI* sys = L.gen();
    check(sys != nullptr, "GenerateTracerSystem returns a system");
    TracerError e = sys->Initialize({}, ScenePartitionerType::SINGLE_GPU);
    check(e == TracerError::UNABLE_TO_INITIALIZE_TRACER, "Initialize without a device: UNABLE_TO_INITIALIZE_TRACER");
    GPUSceneI* scene = nullptr;
    const std::string sp = scenePath;
    sys->GenerateScene(scene, std::u8string(sp.begin(), sp.end()));
    check(scene != nullptr, "GenerateScene returns a scene object");
    check(scene->LoadScene(0.0) != SceneError::OK, "LoadScene without an initialized system fails");

    GPUTracerPtr tracer(nullptr, nullptr);
    TracerParameters params = {false, false, 0u};
    Options good(PathTracerOptions(2));
    check(sys->GenerateTracer(tracer, params, good, "AOTracer") == TracerError::NO_LOGIC_FOR_TRACER,
          "unknown tracer type: NO_LOGIC_FOR_TRACER");
    {
        VariableList v = PathTracerOptions(2);
        v.erase("MaxDepth");
        e = sys->GenerateTracer(tracer, params, Options(std::move(v)), "PathTracer");
        check(e == TracerError::OPTION_NOT_FOUND && !tracer, "missing MaxDepth: OPTION_NOT_FOUND, no tracer");
    }
    {
        VariableList v = PathTracerOptions(2);
        v["NextEventEstimation"] = OptionVariable((int64_t)1);
        e = sys->GenerateTracer(tracer, params, Options(std::move(v)), "PathTracer");
        check(e == TracerError::OPTION_TYPE_MISMATCH, "NextEventEstimation as an integer: OPTION_TYPE_MISMATCH");
    }
    {
        VariableList v = PathTracerOptions(2);
        v["MaxDepth"] = OptionVariable((int64_t)-1);
        e = sys->GenerateTracer(tracer, params, Options(std::move(v)), "PathTracer");
        check(e == TracerError::OPTION_TYPE_MISMATCH, "negative MaxDepth (GetUInt range): OPTION_TYPE_MISMATCH");
    }
    {
        VariableList v = PathTracerOptions(2);
        v["NEESampler"] = OptionVariable(std::string("Power"));
        e = sys->GenerateTracer(tracer, params, Options(std::move(v)), "PathTracer");
        check(e == TracerError::UNABLE_TO_INITIALIZE_TRACER, "unknown NEESampler: UNABLE_TO_INITIALIZE_TRACER");
    }
    {
        VariableList v = WFPGOptions(2);
        v.erase("GuidingDisableRatio");
        e = sys->GenerateTracer(tracer, params, Options(std::move(v)), "WFPGTracer");
        check(e == TracerError::OPTION_NOT_FOUND, "WFPGTracer without GuidingDisableRatio: OPTION_NOT_FOUND");
    }
    {
        VariableList v = WFPGOptions(2);
        v["RenderMode"] = OptionVariable(std::string("Wireframe"));
        e = sys->GenerateTracer(tracer, params, Options(std::move(v)), "WFPGTracer");
        check(e == TracerError::TRACER_INTERNAL_ERROR, "unknown RenderMode: TRACER_INTERNAL_ERROR");
    }
    e = sys->GenerateTracer(tracer, params, Options(WFPGOptions(2)), "WFPGTracer");
    check(e == TracerError::OK && tracer, "WFPGTracer with its full option set");
    e = sys->GenerateTracer(tracer, params, good, "PathTracer");
    check(e == TracerError::OK && tracer, "PathTracer with its full option set");

    Visor visor;
    tracer->AttachTracerCallbacks(visor);
    tracer->AskOptions();
    uint32_t maxDepth = 0;
    int32_t samples = 0;
    check(visor.gotOptions && visor.lastOptions.GetUInt(maxDepth, "MaxDepth") == TracerError::OK && maxDepth == 8 &&
              visor.lastOptions.GetInt(samples, "Samples") == TracerError::OK && samples == 2,
          "AskOptions sends the options back (MaxDepth 8, Samples 2)");
    check(tracer->Initialize() == TracerError::UNABLE_TO_INITIALIZE_TRACER, "tracer Initialize without a loaded scene");
    tracer->GenerateWork(0u);
    check(!visor.errors.empty() && visor.errors.back() == TracerError::UNABLE_TO_GENERATE_WORK,
          "GenerateWork without a loaded scene reports UNABLE_TO_GENERATE_WORK");
    check(!tracer->Render(), "Render without work returns false");
    tracer->SetImagePixelFormat(PixelFormat::RGBA8_UNORM);
    check(visor.errors.back() == TracerError::IMEM_UNKNOWN_PIXEL_FORMAT, "8-bit pixel format: IMEM_UNKNOWN_PIXEL_FORMAT");

    GPUReconFilterI* f = nullptr;
    {
        VariableList v;
        v.emplace("radius", OptionVariable(1.0f));
        e = sys->GenerateReconFilter(f, Options(std::move(v)));
        check(e == TracerError::OPTION_NOT_FOUND && static_cast<std::string>(e).find("type") != std::string::npos,
              "filter without \"type\": OPTION_NOT_FOUND naming the field");
    }
    {
        VariableList v;
        v.emplace("type", OptionVariable(std::string("Tent")));
        v.emplace("radius", OptionVariable((int64_t)1));
        e = sys->GenerateReconFilter(f, Options(std::move(v)));
        check(e == TracerError::OPTION_TYPE_MISMATCH, "integer radius: OPTION_TYPE_MISMATCH (GetFloat)");
    }
    check(sys->GenerateReconFilter(f, FilterOptions("Lanczos", 1.0f)) == TracerError::TRACER_INTERNAL_ERROR,
          "unknown filter type: TRACER_INTERNAL_ERROR");
    check(sys->GenerateReconFilter(f, FilterOptions("Gaussian", 1.0f)) == TracerError::TRACER_INTERNAL_ERROR,
          "Gaussian without alpha: TRACER_INTERNAL_ERROR");
    f = nullptr;
    check(sys->GenerateReconFilter(f, FilterOptions("Tent", 1.0f)) == TracerError::OK && f != nullptr,
          "Tent radius 1: a filter object");
    tracer->AttachReconFilter(f);
    tracer = nullptr;
    L.del(sys);
    check(true, "DeleteTracerSystem");
    return fails;
}

int GpuRun(const Lib& L, int argc, char** argv)
{
    if (argc != 12) return 2;
    const char* scenePath = argv[3];
    const std::string type = argv[4];
    const int W = std::atoi(argv[5]), H = std::atoi(argv[6]), frames = std::atoi(argv[7]);
    const int samples = std::atoi(argv[8]);
    const uint32_t seed = (uint32_t)std::strtoul(argv[9], nullptr, 10);
    const std::string filterName = argv[10];
    const char* outPath = argv[11];

    TracerSystemI* sys = L.gen();
    check(sys->Initialize({}, ScenePartitionerType::MULTI_GPU_MATERIAL) == TracerError::UNKNOWN_SCENE_PARTITIONER_TYPE,
          "MULTI_GPU_MATERIAL partitioner: UNKNOWN_SCENE_PARTITIONER_TYPE");
    check(sys->Initialize({}, ScenePartitionerType::SINGLE_GPU) == TracerError::OK, "Initialize (single GPU)");
    Visor visor;
    visor.Resize(W, H);
    // TracerThread::LoopWork: scene, then tracer, then work
    GPUSceneI* scene = nullptr;
    {
        const std::string missing = std::string(scenePath) + ".missing.json";
        sys->GenerateScene(scene, std::u8string(missing.begin(), missing.end()));
        check(scene->LoadScene(0.0) == SceneError::FILE_NOT_FOUND, "LoadScene of a missing file: FILE_NOT_FOUND");
    }
    const std::string sp = scenePath;
    sys->GenerateScene(scene, std::u8string(sp.begin(), sp.end()));
    SceneError se = scene->LoadScene(0.0);
    check(se == SceneError::OK, "LoadScene: " + static_cast<std::string>(se));
    if (se != SceneError::OK) return 1;
    visor.SendCurrentSceneCameraCount(scene->CameraCount());
    check(visor.cameraCount >= 1, "scene cameras");
    check(scene->UsedGPUMemory() > 0, "scene holds device memory");

    GPUTracerPtr tracer(nullptr, nullptr);
    TracerParameters params = {false, false, seed};
    Options opts(type == "WFPGTracer" ? WFPGOptions(samples) : PathTracerOptions(samples));
    TracerError e = sys->GenerateTracer(tracer, params, opts, type);
    check(e == TracerError::OK, "GenerateTracer " + type);
    if (e != TracerError::OK) return 1;
    tracer->AttachTracerCallbacks(visor);
    check(tracer->Initialize() == TracerError::OK, "tracer Initialize");
    GPUReconFilterI* filter = nullptr;
    const bool tent = (filterName == "tent1");
    check(sys->GenerateReconFilter(filter, tent ? FilterOptions("Tent", 1.0f) : FilterOptions("Box", 0.0f)) ==
              TracerError::OK,
          tent ? "Tent filter, radius 1" : "Box filter, radius 0");
    tracer->AttachReconFilter(filter);
    tracer->SetImagePixelFormat(PixelFormat::RGBA_FLOAT);
    tracer->ResizeImage(Vector2i(W, H));
    tracer->ReportionImage(Zero2i, BaseConstants::IMAGE_MAX_SIZE);
    for (int k = 0; k < frames; k++) {
        tracer->GenerateWork(0u);
        int bounces = 0;
        while (tracer->Render()) bounces++;
        tracer->Finalize();
        check(bounces > 0, "frame " + std::to_string(k) + ": " + std::to_string(bounces) + " Render() calls");
    }
    check(visor.errors.empty(), "no tracer errors");
    check(visor.images == (uint32_t)frames && visor.gotAnalytics, "one image and analytics per Finalize");
    check(tracer->TotalGPUMemoryUsed() >= scene->UsedGPUMemory(), "TotalGPUMemoryUsed covers the scene");
    const std::vector<float> img = visor.Image();
    double mean = 0.0;
    for (float v : img) mean += v;
    mean /= (double)img.size();
    check(mean > 0.0, "the image is not black");
    if (FILE* fo = std::fopen(outPath, "wb")) {
        std::fwrite(img.data(), sizeof(float), img.size(), fo);
        std::fclose(fo);
    } else {
        check(false, "write the image");
    }
    // a VisorTransform frame and a quarter portion after ResetImage
    VisorTransform vt = {};
    vt.position = Vector3f(0.0f, 1.0f, 3.0f);
    vt.gazePoint = Vector3f(0.0f, 1.0f, 0.0f);
    vt.up = Vector3f(0.0f, 1.0f, 0.0f);
    tracer->ReportionImage(Zero2i, Vector2i(W / 2, H / 2));
    tracer->ResetImage();
    tracer->GenerateWork(vt, 0u);
    while (tracer->Render()) {}
    tracer->Finalize();
    check(visor.errors.empty() && visor.images == (uint32_t)frames + 1, "VisorTransform frame of a quarter portion");
    tracer = nullptr;
    L.del(sys);
    return fails;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s cpu|gpu <libTracerMRH.so> <scene.json> ...\n", argv[0]);
        return 2;
    }
    Lib L;
    if (!Open(argv[2], L)) return 1;
    check(true, "dlopen + dlsym GenerateTracerSystem / DeleteTracerSystem");
    const std::string mode = argv[1];
    int r = (mode == "cpu") ? CpuChecks(L, argv[3]) : GpuRun(L, argc, argv);
    std::printf("%s\n", r == 0 ? "ALL OK" : "FAILURES");
    return r == 0 ? 0 : 1;
}
