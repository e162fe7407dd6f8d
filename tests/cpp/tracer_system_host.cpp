// tracer_system_host.cpp — a host driving the library the way MRay's
// TracerThread drives TracerCUDA: load the DLL by name, resolve the factory
// pair (SharedLib::GenerateObject, RayLib/SharedLib.h:52-84), initialise the
// system, hand it a scene, create a "PathTracer", and loop
// GenerateWork -> while (Render()) -> Finalize (RayLib/TracerThread.cpp:144-160).
//
//   tracer_system_host <libmrh.so> <out.bin> <width> <height> <frames>
//                      [<filtered.bin> [<scene.json> <scene.bin>]]
// (filtered.bin: one more frame through a Tent filter of radius 1;
//  scene.bin: one frame of scene.json loaded by GenerateScene(path))
//
// Writes the final mean image (float RGB) to out.bin and prints one line per
// check.  Exit status 0 when every check passed.
#include "mrh_tracer.hpp"

#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

namespace {

struct Callbacks : mrh::TracerCallbacksI {
    std::vector<float> image;
    uint32_t frames = 0;
    int errors = 0;
    void SendError(mrh::TracerError e) override
    {
        std::printf("error: %s\n", std::string(e).c_str());
        errors++;
    }
    void SendImage(const float* rgb, uint32_t w, uint32_t h, uint32_t f) override
    {
        image.assign(rgb, rgb + (size_t)w * h * 3);
        frames = f;
    }
};

int fails = 0;
void check(bool ok, const char* what)
{
    std::printf("%s %s\n", ok ? "ok" : "FAIL", what);
    if (!ok) fails++;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 6 && argc != 7 && argc != 9) {
        std::fprintf(stderr,
                     "usage: %s <libmrh.so> <out.bin> <width> <height> <frames> [<filtered.bin> [<scene.json> "
                     "<scene.bin>]]\n",
                     argv[0]);
        return 2;
    }
    const uint32_t W = (uint32_t)std::atoi(argv[3]), H = (uint32_t)std::atoi(argv[4]);
    const int frames = std::atoi(argv[5]);
    void* lib = dlopen(argv[1], RTLD_NOW | RTLD_LOCAL);
    if (!lib) {
        std::fprintf(stderr, "dlopen: %s\n", dlerror());
        return 2;
    }
    using GenFn = mrh::TracerSystemI* (*)();
    using DelFn = void (*)(mrh::TracerSystemI*);
    using CornellFn = int32_t (*)(uint32_t, mrh_scene_buffers*);
    auto gen = (GenFn)dlsym(lib, "GenerateTracerSystem");
    auto del = (DelFn)dlsym(lib, "DeleteTracerSystem");
    auto cornell = (CornellFn)dlsym(lib, "mrh_gen_cornell");
    check(gen && del && cornell, "factory symbols resolved");
    if (!gen || !del || !cornell) return 1;

    // scene description from the library's synthetic Cornell generator
    mrh_scene_buffers b = {};
    cornell(4, &b);
    mrh::SceneDescription desc;
    desc.triangles.resize((size_t)b.n * 9);
    desc.hitKeys.resize(b.n);
    desc.lambert.resize((size_t)b.n_lambert * 3);
    desc.emissive.resize((size_t)b.n_emissive * 3);
    b.tris = desc.triangles.data();
    b.prim_key = desc.hitKeys.data();
    b.lambert = desc.lambert.data();
    b.emissive = desc.emissive.data();
    check(cornell(4, &b) == MRH_OK, "scene generated");
    desc.cameras.push_back(b.camera);
    desc.accelerator = "BasicBVH";   // the reference's type name maps to the LBVH

    {
        std::unique_ptr<mrh::TracerSystemI, DelFn> sys(gen(), del);
        check(sys->Initialize(0) == mrh::TracerError::OK, "TracerSystemI::Initialize");
        mrh::GPUSceneI* scene = nullptr;
        check(sys->GenerateScene(scene, std::u8string(u8"scene.gltf")) == mrh::SceneError::NO_LOADER,
              "formats other than JSON scenes report NO_LOADER");
        check(sys->GenerateScene(scene, std::u8string(u8"/nonexistent/scene.json")) ==
                  mrh::SceneError::SCENE_FILE_ERROR,
              "a missing scene file reports SCENE_FILE_ERROR");
        check(sys->GenerateScene(scene, desc) == mrh::SceneError::OK && scene, "GenerateScene");
        mrh::TracerParameters params;
        params.seed = 0;
        mrh::Options opts;
        opts.Set("Samples", int64_t(1));
        opts.Set("MaxDepth", int64_t(8));
        opts.Set("RussianRouletteStart", int64_t(3));
        opts.Set("NextEventEstimation", false);
        mrh::GPUTracerPtr tracer;
        check(sys->GenerateTracer(tracer, params, opts, "WFPGTracerX") == mrh::TracerError::NO_LOGIC_FOR_TRACER,
              "unknown tracer type reports NO_LOGIC_FOR_TRACER");
        mrh::Options bad;
        bad.Set("MaxDepth", 8.5);
        check(sys->GenerateTracer(tracer, params, bad, "PathTracer") == mrh::TracerError::OPTION_TYPE_MISMATCH,
              "option type mismatch reported");
        check(sys->GenerateTracer(tracer, params, opts, "PathTracer") == mrh::TracerError::OK && tracer,
              "GenerateTracer(PathTracer)");
        Callbacks cb;
        tracer->AttachTracerCallbacks(cb);
        tracer->ResizeImage(W, H);
        check(tracer->Initialize() == mrh::TracerError::OK, "GPUTracerI::Initialize");
        for (int f = 0; f < frames; f++) {
            tracer->GenerateWork(0u);
            int bounces = 0;
            while (tracer->Render()) bounces++;
            check(bounces == 7, "Render() returns true for MaxDepth - 1 bounces");
            tracer->Finalize();
        }
        check(cb.errors == 0 && cb.frames == (uint32_t)frames && cb.image.size() == (size_t)W * H * 3,
              "Finalize sent the image");
        check(tracer->TotalGPUMemoryUsed() > 0, "TotalGPUMemoryUsed");
        FILE* out = std::fopen(argv[2], "wb");
        if (out) {
            std::fwrite(cb.image.data(), sizeof(float), cb.image.size(), out);
            std::fclose(out);
        }
        tracer.reset();

        // reconstruction filter (TracerSystemI::GenerateReconFilter +
        // GPUTracerI::AttachReconFilter)
        mrh::GPUReconFilterI* filter = nullptr;
        mrh::Options fo;
        fo.Set("radius", 1.0);
        check(sys->GenerateReconFilter(filter, fo) != mrh::TracerError::OK && !filter,
              "a filter without \"type\" is refused");
        fo.Set("type", std::string("Gaussian"));
        check(sys->GenerateReconFilter(filter, fo) != mrh::TracerError::OK, "Gaussian without alpha is refused");
        fo.Set("type", std::string("Lanczos"));
        check(sys->GenerateReconFilter(filter, fo) != mrh::TracerError::OK, "unknown filter type is refused");
        fo.Set("type", std::string("Tent"));
        check(sys->GenerateReconFilter(filter, fo) == mrh::TracerError::OK && filter &&
                  std::string(filter->Type()) == "Tent",
              "GenerateReconFilter(Tent, radius 1)");
        if (argc >= 7 && filter) {
            check(sys->GenerateTracer(tracer, params, opts, "PathTracer") == mrh::TracerError::OK && tracer,
                  "GenerateTracer for the filtered frame");
            Callbacks fcb;
            tracer->AttachTracerCallbacks(fcb);
            tracer->AttachReconFilter(filter);
            tracer->ResizeImage(W, H);
            check(tracer->Initialize() == mrh::TracerError::OK, "Initialize with a filter attached");
            tracer->GenerateWork(0u);
            while (tracer->Render()) {}
            tracer->Finalize();
            check(fcb.errors == 0 && fcb.image.size() == (size_t)W * H * 3, "filtered frame sent");
            FILE* fout = std::fopen(argv[6], "wb");
            if (fout) {
                std::fwrite(fcb.image.data(), sizeof(float), fcb.image.size(), fout);
                std::fclose(fout);
            }
            tracer.reset();
        }

        // a scene file through GenerateScene(path)
        if (argc == 9) {
            mrh::GPUSceneI* js = nullptr;
            const std::string sp = argv[7];
            const mrh::SceneError se = sys->GenerateScene(js, std::u8string(sp.begin(), sp.end()));
            check(se == mrh::SceneError::OK && js && js->CameraCount() > 0, "GenerateScene(scene.json)");
            if (js) {
                check(sys->GenerateTracer(tracer, params, opts, "PathTracer") == mrh::TracerError::OK && tracer,
                      "GenerateTracer for the scene file");
                Callbacks scb;
                tracer->AttachTracerCallbacks(scb);
                tracer->ResizeImage(W, H);
                check(tracer->Initialize() == mrh::TracerError::OK, "Initialize on the scene file");
                tracer->GenerateWork(0u);
                while (tracer->Render()) {}
                tracer->Finalize();
                check(scb.errors == 0 && scb.image.size() == (size_t)W * H * 3, "scene-file frame sent");
                FILE* sout = std::fopen(argv[8], "wb");
                if (sout) {
                    std::fwrite(scb.image.data(), sizeof(float), scb.image.size(), sout);
                    std::fclose(sout);
                }
                tracer.reset();
            }
        }
    }
    dlclose(lib);
    std::printf("%s\n", fails ? "FAILED" : "ALL OK");
    return fails ? 1 : 0;
}
