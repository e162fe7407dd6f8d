// mrh_tracer.hpp — C++ host interface of the MI355X tracer library, shaped
// like MRay's plugin interfaces so a host that drives TracerCUDA through
// them can drive this library the same way.
//
// Mirrors (names, argument meaning, error behaviour):
//   TracerSystemI   RayLib/TracerSystemI.h:34-51
//   GPUTracerI      RayLib/GPUTracerI.h:30-72
//   GPUSceneI       RayLib/GPUSceneI.h:22
//   TracerError     RayLib/TracerError.h (enum subset, same names)
//   Options         RayLib/Options.h:24-132 (Get -> OPTION_NOT_FOUND /
//                   OPTION_TYPE_MISMATCH)
//   TracerParameters RayLib/TracerStructs.h:64-69
//   factory         TracerCUDA/EntryPoint.h:12-16 (GenerateTracerSystem /
//                   DeleteTracerSystem, exported with C linkage)
//
// Scenes are handed over as in-memory descriptions (triangles, HitKeys,
// material tables, cameras); the reference's JSON scene loader is not part
// of this library (GenerateScene from a path reports SceneError::NO_LOADER).
#pragma once

#include "mrh.h"

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <variant>
#include <vector>

namespace mrh {

struct TracerError {
    enum Type {
        OK,
        NO_LOGIC_SET,
        CPU_OUT_OF_MEMORY,
        GPU_OUT_OF_MEMORY,
        OPTION_NOT_FOUND,
        OPTION_TYPE_MISMATCH,
        UNABLE_TO_CONSTRUCT_ACCELERATOR,
        UNABLE_TO_GENERATE_WORK,
        UNABLE_TO_INITIALIZE_TRACER,
        NO_LOGIC_FOR_TRACER,
        TRACER_INTERNAL_ERROR,
        END
    };
    TracerError(Type t = OK, std::string extra = {}) : type(t), extra(std::move(extra)) {}
    operator Type() const { return type; }
    operator std::string() const
    {
        static const char* const names[] = {
            "OK", "No logic set", "CPU out of memory", "GPU out of memory", "Option not found",
            "Option type mismatch", "Unable to construct accelerator", "Unable to generate work",
            "Unable to initialize tracer", "No logic found for tracer", "Tracer internal error"};
        std::string s = (type < END) ? names[type] : "Unknown error";
        if (!extra.empty()) s += ": " + extra;
        return s;
    }
    // mrh_status of a C-ABI call -> TracerError (the C-ABI's mrh_last_error in extra)
    static TracerError FromStatus(int32_t status, const char* message);

    Type type;
    std::string extra;
};

struct SceneError {
    enum Type {
        OK,
        NO_LOADER,                // a scene file format this library does not read
        EMPTY_SCENE,
        INVALID_HIT_KEY,          // a primitive names a missing material / batch
        ACCELERATOR_TYPE_UNKNOWN,
        GPU_OUT_OF_MEMORY,
        INTERNAL_ERROR,
        SCENE_FILE_ERROR,         // unreadable or malformed scene / mesh file
        UNSUPPORTED_FEATURE,      // a scene-file feature this library does not render
        END
    };
    SceneError(Type t = OK, std::string extra = {}) : type(t), extra(std::move(extra)) {}
    operator Type() const { return type; }
    operator std::string() const
    {
        static const char* const names[] = {"OK", "No scene loader for this format", "Empty scene",
                                            "Invalid hit key", "Unknown accelerator type",
                                            "GPU out of memory", "Internal error", "Scene file error",
                                            "Unsupported scene feature"};
        std::string s = (type < END) ? names[type] : "Unknown error";
        if (!extra.empty()) s += ": " + extra;
        return s;
    }
    Type type;
    std::string extra;
};

using OptionVariant = std::variant<bool, int64_t, double, std::string>;

class Options {
public:
    void Set(const std::string& name, OptionVariant v) { values[name] = std::move(v); }
    TracerError GetBool(bool& out, const std::string& name) const { return GetT(out, name); }
    TracerError GetInt(int64_t& out, const std::string& name) const { return GetT(out, name); }
    TracerError GetDouble(double& out, const std::string& name) const { return GetT(out, name); }
    TracerError GetString(std::string& out, const std::string& name) const { return GetT(out, name); }
    // Options::GetFloat: a number given as an integer reads as well
    TracerError GetNumber(double& out, const std::string& name) const
    {
        int64_t i = 0;
        if (GetT(i, name) == TracerError::OK) {
            out = (double)i;
            return TracerError::OK;
        }
        return GetT(out, name);
    }
    bool Has(const std::string& name) const { return values.count(name) != 0; }

private:
    template <class T>
    TracerError GetT(T& out, const std::string& name) const
    {
        auto it = values.find(name);
        if (it == values.end()) return TracerError(TracerError::OPTION_NOT_FOUND, name);
        if (!std::holds_alternative<T>(it->second))
            return TracerError(TracerError::OPTION_TYPE_MISMATCH, name);
        out = std::get<T>(it->second);
        return TracerError::OK;
    }
    std::map<std::string, OptionVariant> values;
};

struct TracerParameters {
    bool forceOptiX = false;   // accepted and ignored: there is no OptiX path
    bool verbose = false;
    uint32_t seed = 0;
};

// Responses from the tracer (RayLib/TracerCallbacksI.h); defaults do nothing.
class TracerCallbacksI {
public:
    virtual ~TracerCallbacksI() = default;
    virtual void SendLog(const std::string&) {}
    virtual void SendError(TracerError) {}
    // mean radiance, float RGB, rows bottom-up, after `frames` finished frames
    virtual void SendImage(const float* rgb, uint32_t width, uint32_t height, uint32_t frames) {}
    // M paths/s and frame time, as RayTracer::UpdateFrameAnalytics (RayTracer.cu:38-62)
    virtual void SendAnalyticData(double mpaths_per_s, double frame_ms) {}
};

// In-memory scene: one triangle primitive group, the four material batches
// and the cameras.  accelerator: "LBVH" (or its alias "BasicBVH").
struct SceneDescription {
    std::vector<float> triangles;     // [n][9]
    std::vector<uint32_t> hitKeys;    // [n], batch:8 | id:24
    std::vector<float> lambert;       // [..][3]
    std::vector<float> metal;         // [..][10]
    std::vector<float> dielectric;    // [..][4]
    std::vector<float> emissive;      // [..][3]
    std::vector<float> reflect;       // [..][4]
    std::vector<mrh_camera_desc> cameras;
    std::string accelerator = "LBVH";
    uint32_t mortonBits = 30;
};

class GPUSceneI {
public:
    virtual ~GPUSceneI() = default;
    virtual size_t PrimitiveCount() const = 0;
    virtual uint32_t CameraCount() const = 0;
    virtual const mrh_camera_desc& Camera(uint32_t id) const = 0;
    virtual size_t UsedGPUMemory() const = 0;
    virtual mrh_scene* Handle() const = 0;
};

// GPUReconFilterI (TracerCUDA/GPUReconFilterI.h): a reconstruction filter
// made by TracerSystemI::GenerateReconFilter and attached to tracers
class GPUReconFilterI {
public:
    static constexpr const char* RADIUS_OPTION_NAME = "radius";
    static constexpr const char* TYPE_OPTION_NAME = "type";
    virtual ~GPUReconFilterI() = default;
    virtual const char* Type() const = 0;   // "Box", "Tent", "Gaussian", "Mitchell"
    virtual const mrh_filter_desc& Desc() const = 0;
};

class GPUTracerI {
public:
    virtual ~GPUTracerI() = default;
    virtual void AttachTracerCallbacks(TracerCallbacksI&) = 0;
    // RayLib/GPUTracerI.h:69; frames are filtered from the next GenerateWork
    // on (without a filter each pixel is the mean of its own samples)
    virtual void AttachReconFilter(GPUReconFilterI* filter) = 0;
    virtual TracerError Initialize() = 0;
    virtual TracerError SetOptions(const Options&) = 0;
    virtual void SetParameters(const TracerParameters&) = 0;
    virtual void AskOptions() = 0;
    virtual void AskParameters() = 0;
    virtual void GenerateWork(uint32_t cameraId) = 0;
    virtual void GenerateWork(const mrh_camera_desc& camera) = 0;
    virtual bool Render() = 0;     // one bounce; false when the frame is done
    virtual void Finalize() = 0;   // image -> TracerCallbacksI::SendImage
    virtual void ResizeImage(uint32_t width, uint32_t height) = 0;
    virtual void ResetImage() = 0;
    virtual size_t TotalGPUMemoryUsed() const = 0;
    // host copy of the current mean image (width*height*3 floats)
    virtual TracerError GetImage(std::vector<float>& rgb) const = 0;
};

using GPUTracerPtr = std::unique_ptr<GPUTracerI>;

class TracerSystemI {
public:
    virtual ~TracerSystemI() = default;
    // device index instead of surface-loader libraries and a partitioner
    // type; the single-GPU partition is the only one
    virtual TracerError Initialize(int32_t device) = 0;
    virtual void ClearScene() = 0;
    virtual SceneError GenerateScene(GPUSceneI*& scene, const SceneDescription& desc) = 0;
    // the reference's JSON scene files (mrh_scene_file_load, strict: any
    // feature it cannot render is an UNSUPPORTED_FEATURE error)
    virtual SceneError GenerateScene(GPUSceneI*& scene, const std::u8string& scenePath) = 0;
    // tracerType: "PathTracer"
    virtual TracerError GenerateTracer(GPUTracerPtr& tracer, const TracerParameters& params,
                                       const Options& opts, const std::string& tracerType) = 0;
    // TracerSystemCUDA::GenerateReconFilter (TracerSystemCUDA.cpp:109-140):
    // "type" and "radius" are required, plus "alpha" (Gaussian) or "b" and
    // "c" (Mitchell); the system owns the filter (one at a time)
    virtual TracerError GenerateReconFilter(GPUReconFilterI*& filter, const Options& filterOptions) = 0;
};

}  // namespace mrh

extern "C" {
mrh::TracerSystemI* GenerateTracerSystem();
void DeleteTracerSystem(mrh::TracerSystemI* system);
}
