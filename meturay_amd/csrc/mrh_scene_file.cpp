// mrh_scene_file.cpp — scene files: the reference's JSON scene format
// (GPUSceneJson.cpp:690-885 and SceneNodeJson; sections Cameras, Lights,
// Mediums, Transforms, Materials, Primitives, Surfaces, LightSurfaces,
// CameraSurfaces) and Wavefront OBJ meshes (read by the reference through
// its assimp loader, AssimpSurfaceLoaders/MetaSurfaceLoader.cpp), flattened
// into what this library renders: world-space triangles, HitKeys and the
// material tables of mrh_scene_desc, plus the cameras.
//
// Supported: Pinhole cameras; Null and Primitive lights (a primitive light's
// triangles become emissive with its radiance); Vacuum and Homogeneous
// media (the ior of Refract); Identity and Single transforms (matrix4x4 in
// row or column major, or trs = T * Rz * Ry * Rx * S with angles in
// degrees, GPUTransformSingle.cu:44-100); LambertC / Lambert with a constant
// albedo, Reflect, Refract, Metal; Triangle primitives inline
// (nodeTriangle, nodeTriangleIndexed) or from OBJ files (innerIndex = mesh
// in the file; meshes split at o / g / usemtl, polygons fan-triangulated,
// degenerate triangles dropped).  Vertex normals (OBJ vn, the JSON "normal"
// lists) become per-corner shading normals, transformed by the inverse
// transpose and renormalised; triangles of a scene that has normals
// elsewhere but none of their own get their face normal (aiProcess_GenNormals
// of MetaSurfaceLoader.cpp:27), substituted spheres their radial normals.
// Everything else (textures, Unreal, analytic
// spheres, other lights and cameras, boundary lights) is an error, or, with
// MRH_SCENE_FILE_SUBSTITUTE, replaced (grey or constant-albedo Lambert for
// materials, a 64 x 32 UV tessellation for spheres) or skipped, each noted
// in the warnings.  Triangles are two-sided (cullFace is not honoured).
#include "mrh.h"
#include "mrh_json.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

struct mrh_scene_file {
    std::vector<float> tris;
    std::vector<float> normals;   // [n][9] per-corner shading normals, or empty
    std::vector<uint32_t> keys;
    std::vector<float> lambert, metal, dielectric, emissive, reflect;
    std::vector<mrh_camera_desc> cameras;
    std::string warnings;
    // instanced mode: meshes (one per primitive in first-use order) and one
    // instance per surface primitive / light surface
    bool instanced = false;
    std::vector<uint32_t> mesh_first, inst_mesh, inst_key;
    std::vector<float> xform;
    mrh_instancing_desc inst_desc = {};
};

namespace {

using mrh::json::Value;

struct Unsupported : std::runtime_error {
    explicit Unsupported(const std::string& m) : std::runtime_error(m) {}
};

struct Mat4 {
    double m[16];   // row major, p' = M p
};

Mat4 identity()
{
    Mat4 r = {};
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0;
    return r;
}

Mat4 mul(const Mat4& a, const Mat4& b)
{
    Mat4 r = {};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double s = 0.0;
            for (int k = 0; k < 4; k++) s += a.m[4 * i + k] * b.m[4 * k + j];
            r.m[4 * i + j] = s;
        }
    return r;
}

// rotation by `deg` degrees about a unit axis (TransformGen::Rotate,
// RayLib/Matrix.hpp:924-961: the right-handed rotation)
Mat4 rotate(double deg, double x, double y, double z)
{
    const double a = deg * M_PI / 180.0, c = std::cos(a), s = std::sin(a), t = 1.0 - c;
    Mat4 r = identity();
    r.m[0] = c + x * x * t;     r.m[1] = x * y * t - z * s; r.m[2] = x * z * t + y * s;
    r.m[4] = x * y * t + z * s; r.m[5] = c + y * y * t;     r.m[6] = y * z * t - x * s;
    r.m[8] = x * z * t - y * s; r.m[9] = y * z * t + x * s; r.m[10] = c + z * z * t;
    return r;
}

// unit face normal of a triangle [9] (zero for a degenerate one)
void face_normal(const float* t, float n[3])
{
    const double ax = t[3] - t[0], ay = t[4] - t[1], az = t[5] - t[2];
    const double bx = t[6] - t[0], by = t[7] - t[1], bz = t[8] - t[2];
    const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
    const double l = std::sqrt(cx * cx + cy * cy + cz * cz);
    n[0] = l > 0 ? (float)(cx / l) : 0.0f;
    n[1] = l > 0 ? (float)(cy / l) : 0.0f;
    n[2] = l > 0 ? (float)(cz / l) : 0.0f;
}

// Wavefront OBJ: v, vn and f lines (v, v/t, v//n, v/t/n; negative indices
// count back from the vertices / normals read so far, positive ones may
// refer to entries further down the file); a new mesh starts at o, g and
// usemtl once the current one has faces; polygons are fans around their
// first vertex.  Meshes without faces are dropped.  `normals`, when given,
// receives each mesh's per-corner normals [..][9]: empty for a mesh none of
// whose corners names one, else the named normal or, for corners naming
// none, the face normal.
int32_t obj_read(const char* file, std::vector<std::vector<float>>* meshes,
                 std::vector<std::vector<float>>* normals = nullptr)
{
    std::ifstream f(file);
    if (!f) return MRH_ERR_INVALID;
    std::vector<float> pos, nrm;
    std::vector<std::vector<long>> faces(1), fnrm(1);   // vertex / normal indices of each mesh's corners
    std::vector<bool> has_nrm(1, false);
    std::string line;
    auto next_mesh = [&]() {
        if (!faces.back().empty()) {
            faces.emplace_back();
            fnrm.emplace_back();
            has_nrm.push_back(false);
        }
    };
    while (std::getline(f, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        const size_t a = line.find_first_not_of(" \t");
        if (a == std::string::npos || line[a] == '#') continue;
        std::istringstream ls(line.substr(a));
        std::string tag;
        ls >> tag;
        if (tag == "v") {
            float x = 0, y = 0, z = 0;
            ls >> x >> y >> z;
            pos.push_back(x);
            pos.push_back(y);
            pos.push_back(z);
        } else if (tag == "vn") {
            float x = 0, y = 0, z = 0;
            ls >> x >> y >> z;
            nrm.push_back(x);
            nrm.push_back(y);
            nrm.push_back(z);
        } else if (tag == "o" || tag == "g" || tag == "usemtl") {
            next_mesh();
        } else if (tag == "f") {
            std::vector<long> idx, nidx;
            std::string tok;
            while (ls >> tok) {
                const size_t s1 = tok.find('/');
                const long v = std::stol(tok.substr(0, s1));
                if (v == 0) return MRH_ERR_INVALID;
                idx.push_back(v > 0 ? v - 1 : (long)(pos.size() / 3) + v);
                long n = -1;
                const size_t s2 = (s1 == std::string::npos) ? s1 : tok.find('/', s1 + 1);
                if (s2 != std::string::npos && s2 + 1 < tok.size()) {
                    const long vn = std::stol(tok.substr(s2 + 1));
                    if (vn == 0) return MRH_ERR_INVALID;
                    n = vn > 0 ? vn - 1 : (long)(nrm.size() / 3) + vn;
                    if (n < 0) return MRH_ERR_INVALID;
                    has_nrm.back() = true;
                }
                nidx.push_back(n);
            }
            for (size_t t = 1; t + 1 < idx.size(); t++)
                for (size_t k : {(size_t)0, t, t + 1}) {
                    faces.back().push_back(idx[k]);
                    fnrm.back().push_back(nidx[k]);
                }
        }
    }
    const long nv = (long)(pos.size() / 3), nn = (long)(nrm.size() / 3);
    for (size_t m = 0; m < faces.size(); m++) {
        const auto& mesh = faces[m];
        if (mesh.empty()) continue;
        std::vector<float> tris;
        tris.reserve(mesh.size() * 3);
        for (long k : mesh) {
            if (k < 0 || k >= nv) return MRH_ERR_INVALID;
            for (int c = 0; c < 3; c++) tris.push_back(pos[3 * (size_t)k + c]);
        }
        if (normals) {
            std::vector<float> cn;
            if (has_nrm[m]) {
                cn.resize(tris.size());
                for (size_t t = 0; t < mesh.size() / 3; t++) {
                    float fn[3];
                    face_normal(&tris[9 * t], fn);
                    for (int c = 0; c < 3; c++) {
                        const long k = fnrm[m][3 * t + c];
                        if (k >= nn) return MRH_ERR_INVALID;
                        for (int a = 0; a < 3; a++) cn[9 * t + 3 * c + a] = k < 0 ? fn[a] : nrm[3 * (size_t)k + a];
                    }
                }
            }
            normals->push_back(std::move(cn));
        }
        meshes->push_back(std::move(tris));
    }
    return MRH_OK;
}

struct Light {
    bool null = true;
    uint32_t primitive = 0;
    float radiance[3] = {0, 0, 0};
};

struct MatRef {
    uint32_t batch, index;
};

class Loader {
public:
    Loader(const std::string& path, bool relaxed, bool instanced, mrh_scene_file& out)
        : path(path), relaxed(relaxed), instanced(instanced), out(out)
    {
        const size_t slash = path.find_last_of('/');
        dir = (slash == std::string::npos) ? std::string() : path.substr(0, slash + 1);
    }

    void run()
    {
        std::ifstream f(path, std::ios::binary);
        if (!f) throw std::runtime_error("cannot open " + path);
        std::stringstream ss;
        ss << f.rdbuf();
        const std::string text = ss.str();
        const Value root = mrh::json::parse(text);
        if (!root.is_object()) throw std::runtime_error("scene file is not a JSON object");
        transforms[~0u] = identity();
        read_transforms(root);
        read_media(root);
        read_materials(root);
        read_primitives(root);
        read_lights(root);
        read_surfaces(root);
        read_light_surfaces(root);
        read_cameras(root);
        if (out.keys.empty()) throw std::runtime_error("scene has no triangles");
        if (instanced) {
            out.instanced = true;
            out.mesh_first.push_back((uint32_t)out.keys.size());
        }
    }

private:
    std::string path, dir;
    bool relaxed, instanced;
    std::map<uint32_t, uint32_t> mesh_of;   // instanced mode: primitive id -> mesh
    mrh_scene_file& out;
    std::map<uint32_t, Mat4> transforms;
    std::map<uint32_t, float> media_ior;
    std::map<uint32_t, MatRef> materials;
    std::map<uint32_t, std::vector<float>> prims;     // local-space triangles [..][9]
    std::map<uint32_t, std::vector<float>> prim_nrm;  // their corner normals [..][9], if any
    std::map<uint32_t, Light> lights;
    struct ObjFile {
        std::vector<std::vector<float>> meshes, normals;
    };
    std::map<std::string, ObjFile> obj_cache;
    bool nrm_dropped = false;

    // strict: fail with "what is not supported"; relaxed: note what was done
    void unsupported(const std::string& what, const std::string& action = "skipped")
    {
        if (!relaxed) throw Unsupported(what + " is not supported");
        out.warnings += what + ": " + action + "\n";
    }

    static const Value& need(const Value& node, const char* key)
    {
        const Value* v = node.get(key);
        if (!v) throw std::runtime_error(std::string("missing \"") + key + "\"");
        return *v;
    }

    static std::vector<uint32_t> ids(const Value& node)
    {
        const Value& v = need(node, "id");
        std::vector<uint32_t> r;
        if (v.is_array())
            for (const Value& e : v.array()) r.push_back((uint32_t)e.number());
        else
            r.push_back((uint32_t)v.number());
        return r;
    }

    // the value of `key` for the i-th of n ids: the value itself for one id,
    // else its i-th element (SceneNodeJson's per-id access)
    static const Value& per_id(const Value& node, const char* key, size_t i, size_t n)
    {
        const Value& v = need(node, key);
        if (n == 1) return v;
        if (!v.is_array() || v.array().size() != n)
            throw std::runtime_error(std::string("\"") + key + "\" needs one entry per id");
        return v.array()[i];
    }

    static void vec3(const Value& v, double out3[3])
    {
        const auto& a = v.array();
        if (a.size() != 3) throw std::runtime_error("expected 3 numbers");
        for (int k = 0; k < 3; k++) out3[k] = a[k].number();
    }

    static std::string type_of(const Value& node) { return need(node, "type").string(); }

    void read_transforms(const Value& root)
    {
        const Value* sec = root.get("Transforms");
        if (!sec) return;
        for (const Value& node : sec->array()) {
            const std::vector<uint32_t> id = ids(node);
            const std::string t = type_of(node);
            for (size_t i = 0; i < id.size(); i++) {
                if (t == "Identity") {
                    transforms[id[i]] = identity();
                } else if (t == "Single") {
                    const std::string layout = need(node, "layout").string();
                    Mat4 m = identity();
                    if (layout == "matrix4x4") {
                        const Value* major = node.get("major");
                        const bool row = !major || major->string() == "row";
                        const auto& a = per_id(node, "matrix", i, id.size()).array();
                        if (a.size() != 16) throw std::runtime_error("matrix4x4 needs 16 numbers");
                        for (int k = 0; k < 16; k++) {
                            const int r = k / 4, c = k % 4;
                            m.m[row ? k : (4 * c + r)] = a[k].number();
                        }
                    } else if (layout == "trs") {
                        double t3[3] = {0, 0, 0}, r3[3] = {0, 0, 0}, s3[3] = {1, 1, 1};
                        if (node.get("translate")) vec3(per_id(node, "translate", i, id.size()), t3);
                        if (node.get("rotate")) vec3(per_id(node, "rotate", i, id.size()), r3);
                        if (node.get("scale")) vec3(per_id(node, "scale", i, id.size()), s3);
                        Mat4 s = identity();
                        s.m[0] = s3[0];
                        s.m[5] = s3[1];
                        s.m[10] = s3[2];
                        Mat4 tr = identity();
                        tr.m[3] = t3[0];
                        tr.m[7] = t3[1];
                        tr.m[11] = t3[2];
                        m = mul(tr, mul(rotate(r3[2], 0, 0, 1), mul(rotate(r3[1], 0, 1, 0), mul(rotate(r3[0], 1, 0, 0), s))));
                    } else {
                        throw std::runtime_error("unknown transform layout " + layout);
                    }
                    transforms[id[i]] = m;
                } else {
                    unsupported("transform type " + t, "identity used");
                    transforms[id[i]] = identity();
                }
            }
        }
    }

    void read_media(const Value& root)
    {
        const Value* sec = root.get("Mediums");
        if (!sec) return;
        for (const Value& node : sec->array()) {
            const std::vector<uint32_t> id = ids(node);
            const std::string t = type_of(node);
            for (size_t i = 0; i < id.size(); i++) {
                if (t == "Vacuum") media_ior[id[i]] = 1.0f;
                else if (t == "Homogeneous")
                    media_ior[id[i]] = node.get("ior") ? (float)per_id(node, "ior", i, id.size()).number() : 1.0f;
                else {
                    unsupported("medium type " + t, "vacuum used");
                    media_ior[id[i]] = 1.0f;
                }
            }
        }
    }

    // constant RGB of a material attribute, or false for a texture reference
    static bool constant_rgb(const Value& v, float rgb[3])
    {
        if (!v.is_array()) return false;
        double d[3];
        vec3(v, d);
        for (int k = 0; k < 3; k++) rgb[k] = (float)d[k];
        return true;
    }

    uint32_t add(std::vector<float>& table, std::initializer_list<float> vals, size_t width)
    {
        const uint32_t index = (uint32_t)(table.size() / width);
        table.insert(table.end(), vals.begin(), vals.end());
        return index;
    }

    void read_materials(const Value& root)
    {
        const Value* sec = root.get("Materials");
        if (!sec) return;
        for (const Value& node : sec->array()) {
            const std::vector<uint32_t> id = ids(node);
            const std::string t = type_of(node);
            for (size_t i = 0; i < id.size(); i++) {
                const size_t n = id.size();
                float a[3] = {0.5f, 0.5f, 0.5f};
                const bool has_albedo = node.get("albedo") != nullptr;
                const bool const_albedo = has_albedo && constant_rgb(per_id(node, "albedo", i, n), a);
                if ((t == "LambertC" || t == "Lambert") && const_albedo) {
                    materials[id[i]] = {MRH_BATCH_LAMBERT, add(out.lambert, {a[0], a[1], a[2]}, 3)};
                } else if (t == "Reflect" && const_albedo) {
                    const float r = node.get("roughness") ? (float)per_id(node, "roughness", i, n).number() : 0.0f;
                    materials[id[i]] = {MRH_BATCH_REFLECT, add(out.reflect, {a[0], a[1], a[2], r}, 4)};
                } else if (t == "Refract" && const_albedo) {
                    const uint32_t med = (uint32_t)per_id(node, "medium", i, n).number();
                    auto it = media_ior.find(med);
                    if (it == media_ior.end()) throw std::runtime_error("Refract names unknown medium " + std::to_string(med));
                    materials[id[i]] = {MRH_BATCH_DIELECTRIC, add(out.dielectric, {a[0], a[1], a[2], it->second}, 4)};
                } else if (t == "Metal") {
                    double eta[3], k[3], spec[3] = {1, 1, 1};
                    vec3(per_id(node, "eta", i, n), eta);
                    vec3(per_id(node, "k", i, n), k);
                    if (node.get("specular")) vec3(per_id(node, "specular", i, n), spec);
                    const float r = node.get("roughness") ? (float)per_id(node, "roughness", i, n).number() : 0.0f;
                    materials[id[i]] = {MRH_BATCH_METAL,
                                        add(out.metal, {r, (float)eta[0], (float)eta[1], (float)eta[2], (float)k[0],
                                                        (float)k[1], (float)k[2], (float)spec[0], (float)spec[1],
                                                        (float)spec[2]}, 10)};
                } else {
                    unsupported("material " + std::to_string(id[i]) + " of type " + t +
                                    (has_albedo && !const_albedo ? " with a textured albedo" : ""),
                                "Lambert of albedo " + std::to_string(a[0]) + ", " + std::to_string(a[1]) + ", " +
                                    std::to_string(a[2]));
                    materials[id[i]] = {MRH_BATCH_LAMBERT, add(out.lambert, {a[0], a[1], a[2]}, 3)};
                }
            }
        }
    }

    // OBJ meshes of a file (positions and corner normals)
    const ObjFile& obj(const std::string& file)
    {
        auto it = obj_cache.find(file);
        if (it != obj_cache.end()) return it->second;
        ObjFile o;
        if (obj_read(file.c_str(), &o.meshes, &o.normals) != MRH_OK) throw std::runtime_error("cannot read OBJ " + file);
        return obj_cache[file] = std::move(o);
    }

    // a "normal" list of vec3 into corner normals
    static void normals_of(const std::vector<Value>& list, std::vector<float>& out3)
    {
        for (const Value& p : list) {
            double v[3];
            vec3(p, v);
            for (double c : v) out3.push_back((float)c);
        }
    }

    void read_primitives(const Value& root)
    {
        const Value* sec = root.get("Primitives");
        if (!sec) return;
        for (const Value& node : sec->array()) {
            const std::vector<uint32_t> id = ids(node);
            const std::string t = type_of(node);
            const std::string name = node.get("name") ? node.get("name")->string() : std::string();
            const size_t n = id.size();
            for (size_t i = 0; i < n; i++) {
                std::vector<float>& tris = prims[id[i]];
                if (t == "Triangle" && name == ".nodeTriangle") {
                    const auto& pos = per_id(node, "position", i, n).array();
                    if (pos.size() % 3) throw std::runtime_error("nodeTriangle positions are not triangles");
                    for (const Value& p : pos) {
                        double v[3];
                        vec3(p, v);
                        for (double c : v) tris.push_back((float)c);
                    }
                    if (node.get("normal")) {
                        const auto& nl = per_id(node, "normal", i, n).array();
                        if (nl.size() != pos.size())
                            throw std::runtime_error("nodeTriangle needs one normal per position");
                        normals_of(nl, prim_nrm[id[i]]);
                    }
                } else if (t == "Triangle" && name == ".nodeTriangleIndexed") {
                    const auto& pos = need(node, "position").array();
                    const auto& idx = per_id(node, "index", i, n).array();
                    if (idx.size() % 3) throw std::runtime_error("nodeTriangleIndexed indices are not triangles");
                    const Value* nv = node.get("normal");
                    if (nv && nv->array().size() != pos.size())
                        throw std::runtime_error("nodeTriangleIndexed needs one normal per position");
                    for (const Value& e : idx) {
                        const size_t k = (size_t)e.number();
                        if (k >= pos.size()) throw std::runtime_error("nodeTriangleIndexed index out of range");
                        double v[3];
                        vec3(pos[k], v);
                        for (double c : v) tris.push_back((float)c);
                        if (nv) {
                            vec3(nv->array()[k], v);
                            for (double c : v) prim_nrm[id[i]].push_back((float)c);
                        }
                    }
                } else if (t == "Triangle" && !name.empty() && name.compare(0, 5, ".node") != 0) {
                    const Value* tag = node.get("tag");
                    if (tag && tag->string() != "assimp_") {
                        unsupported("primitive " + std::to_string(id[i]) + " with surface loader " + tag->string());
                        continue;
                    }
                    const std::string file = (name[0] == '/') ? name : dir + name;
                    const ObjFile& o = obj(file);
                    const size_t inner = node.get("innerIndex") ? (size_t)per_id(node, "innerIndex", i, n).number() : 0;
                    if (inner >= o.meshes.size())
                        throw std::runtime_error(file + " has no mesh " + std::to_string(inner));
                    tris = o.meshes[inner];
                    if (!o.normals[inner].empty()) prim_nrm[id[i]] = o.normals[inner];
                } else if (t == "Sphere" && name == ".nodeSphere") {
                    double c[3];
                    vec3(per_id(node, "position", i, n), c);
                    const double r = per_id(node, "radius", i, n).number();
                    unsupported("primitive " + std::to_string(id[i]) + ", an analytic sphere,", "tessellated (64 x 32)");
                    tessellate_sphere(c, r, tris, prim_nrm[id[i]]);
                } else {
                    unsupported("primitive " + std::to_string(id[i]) + " of type " + t + " " + name);
                }
            }
        }
    }

    static void tessellate_sphere(const double c[3], double r, std::vector<float>& tris, std::vector<float>& nrm)
    {
        const int slices = 64, stacks = 32;
        auto pt = [&](int st, int sl, double* p) {
            const double th = M_PI * st / stacks, ph = 2.0 * M_PI * sl / slices;
            p[0] = c[0] + r * std::sin(th) * std::cos(ph);
            p[1] = c[1] + r * std::cos(th);
            p[2] = c[2] + r * std::sin(th) * std::sin(ph);
        };
        for (int st = 0; st < stacks; st++)
            for (int sl = 0; sl < slices; sl++) {
                double a[3], b[3], d[3], e[3];
                pt(st, sl, a);
                pt(st + 1, sl, b);
                pt(st + 1, sl + 1, d);
                pt(st, sl + 1, e);
                auto tri = [&](const double* p, const double* q, const double* s) {
                    for (const double* v : {p, q, s})
                        for (int k = 0; k < 3; k++) {
                            tris.push_back((float)v[k]);
                            nrm.push_back((float)((v[k] - c[k]) / r));   // radial
                        }
                };
                // wound counter-clockwise seen from outside: the face
                // normals agree with the radial shading normals
                if (st != 0) tri(a, e, b);
                if (st != stacks - 1) tri(b, e, d);
            }
    }

    void read_lights(const Value& root)
    {
        const Value* sec = root.get("Lights");
        if (!sec) return;
        for (const Value& node : sec->array()) {
            const std::vector<uint32_t> id = ids(node);
            const std::string t = type_of(node);
            for (size_t i = 0; i < id.size(); i++) {
                Light l;
                if (t == "Primitive") {
                    l.null = false;
                    l.primitive = (uint32_t)per_id(node, "primitive", i, id.size()).number();
                    if (!constant_rgb(per_id(node, "radiance", i, id.size()), l.radiance)) {
                        unsupported("light " + std::to_string(id[i]) + " with a textured radiance");
                        l.null = true;
                    }
                } else if (t != "Null") {
                    unsupported("light " + std::to_string(id[i]) + " of type " + t);
                }
                lights[id[i]] = l;
            }
        }
    }

    // corner normals of the next triangle (world triangle w, normals wn or
    // NULL); the first triangle with normals gives the flat ones before it
    // their face normals
    void add_normals(const float* w, const float* wn)
    {
        if (!wn && out.normals.empty()) return;
        float fn[3];
        while (out.normals.size() < 9 * out.keys.size()) {
            face_normal(&out.tris[out.normals.size()], fn);
            for (int c = 0; c < 3; c++) out.normals.insert(out.normals.end(), fn, fn + 3);
        }
        if (wn) {
            out.normals.insert(out.normals.end(), wn, wn + 9);
        } else {
            face_normal(w, fn);
            for (int c = 0; c < 3; c++) out.normals.insert(out.normals.end(), fn, fn + 3);
        }
    }

    void emit(const std::vector<float>& local, const std::vector<float>& lnrm, const Mat4& m, uint32_t key)
    {
        // normals transform by the inverse transpose: the cofactor matrix,
        // negated for a mirroring transform, before renormalising
        double cof[9];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) {
                const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
                cof[3 * r + c] = m.m[4 * r1 + c1] * m.m[4 * r2 + c2] - m.m[4 * r1 + c2] * m.m[4 * r2 + c1];
            }
        const double det = m.m[0] * cof[0] + m.m[1] * cof[1] + m.m[2] * cof[2];
        const double sgn = det < 0 ? -1.0 : 1.0;
        const bool smooth = lnrm.size() == local.size();
        for (size_t t = 0; t + 9 <= local.size(); t += 9) {
            float w[9], wn[9];
            for (int v = 0; v < 3; v++) {
                const double x = local[t + 3 * v], y = local[t + 3 * v + 1], z = local[t + 3 * v + 2];
                for (int r = 0; r < 3; r++)
                    w[3 * v + r] = (float)(m.m[4 * r] * x + m.m[4 * r + 1] * y + m.m[4 * r + 2] * z + m.m[4 * r + 3]);
            }
            // degenerate triangles are dropped (FindDegenerates of the mesh loader)
            const double ax = w[3] - w[0], ay = w[4] - w[1], az = w[5] - w[2];
            const double bx = w[6] - w[0], by = w[7] - w[1], bz = w[8] - w[2];
            const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
            if (cx * cx + cy * cy + cz * cz == 0.0) continue;
            if (smooth) {
                float fn[3];
                face_normal(w, fn);
                for (int v = 0; v < 3; v++) {
                    const double x = lnrm[t + 3 * v], y = lnrm[t + 3 * v + 1], z = lnrm[t + 3 * v + 2];
                    double n[3], l = 0.0;
                    for (int r = 0; r < 3; r++) {
                        n[r] = sgn * (cof[3 * r] * x + cof[3 * r + 1] * y + cof[3 * r + 2] * z);
                        l += n[r] * n[r];
                    }
                    l = std::sqrt(l);
                    // a zero or non-finite normal falls back to the face normal
                    for (int r = 0; r < 3; r++) wn[3 * v + r] = (l > 0 && std::isfinite(l)) ? (float)(n[r] / l) : fn[r];
                }
            }
            out.tris.insert(out.tris.end(), w, w + 9);
            add_normals(w, smooth ? wn : nullptr);
            out.keys.push_back(key);
        }
    }

    const Mat4& transform_of(const Value& node)
    {
        const Value* t = node.get("transform");
        const uint32_t id = t ? (uint32_t)t->number() : ~0u;
        auto it = transforms.find(id);
        if (it == transforms.end()) throw std::runtime_error("unknown transform " + std::to_string(id));
        return it->second;
    }

    static std::vector<uint32_t> id_list(const Value& v)
    {
        std::vector<uint32_t> r;
        if (v.is_array())
            for (const Value& e : v.array()) r.push_back((uint32_t)e.number());
        else
            r.push_back((uint32_t)v.number());
        return r;
    }

    // Instanced mode: the primitive's triangles become a mesh (once, in
    // local space, degenerate ones dropped; keys of its first use, which the
    // instances override anyway) and the surface an instance of it.
    void instance(uint32_t prim, const std::vector<float>& local, const Mat4& m, uint32_t key)
    {
        auto it = mesh_of.find(prim);
        if (it == mesh_of.end()) {
            const uint32_t before = (uint32_t)out.keys.size();
            emit(local, {}, identity(), key);
            if (out.keys.size() == before) return;   // nothing but degenerate triangles
            it = mesh_of.emplace(prim, (uint32_t)out.mesh_first.size()).first;
            out.mesh_first.push_back(before);
        }
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 4; c++) out.xform.push_back((float)m.m[4 * r + c]);
        out.inst_mesh.push_back(it->second);
        out.inst_key.push_back(key);
    }

    void place(uint32_t prim, const std::vector<float>& local, const Mat4& m, uint32_t key)
    {
        auto n = prim_nrm.find(prim);
        if (instanced) {
            if (n != prim_nrm.end() && !n->second.empty() && !nrm_dropped) {
                // the two-level tracer shades flat
                out.warnings += "vertex normals in instanced mode: flat shading used\n";
                nrm_dropped = true;
            }
            instance(prim, local, m, key);
        } else {
            static const std::vector<float> none;
            emit(local, n != prim_nrm.end() ? n->second : none, m, key);
        }
    }

    void read_surfaces(const Value& root)
    {
        const Value* sec = root.get("Surfaces");
        if (!sec) return;
        for (const Value& node : sec->array()) {
            const Mat4& m = transform_of(node);
            const std::vector<uint32_t> mats = id_list(need(node, "material"));
            const std::vector<uint32_t> prs = id_list(need(node, "primitive"));
            if (mats.size() != prs.size()) throw std::runtime_error("surface material / primitive counts differ");
            for (size_t k = 0; k < prs.size(); k++) {
                auto mi = materials.find(mats[k]);
                if (mi == materials.end()) throw std::runtime_error("unknown material " + std::to_string(mats[k]));
                auto pi = prims.find(prs[k]);
                if (pi == prims.end()) throw std::runtime_error("unknown primitive " + std::to_string(prs[k]));
                place(prs[k], pi->second, m, MRH_KEY(mi->second.batch, mi->second.index));
            }
        }
    }

    void read_light_surfaces(const Value& root)
    {
        const Value* sec = root.get("LightSurfaces");
        if (!sec) return;
        for (const Value& node : sec->array()) {
            const uint32_t lid = (uint32_t)need(node, "light").number();
            auto li = lights.find(lid);
            if (li == lights.end()) throw std::runtime_error("unknown light " + std::to_string(lid));
            if (li->second.null) continue;
            auto pi = prims.find(li->second.primitive);
            if (pi == prims.end()) throw std::runtime_error("unknown light primitive " + std::to_string(li->second.primitive));
            const uint32_t e = add(out.emissive, {li->second.radiance[0], li->second.radiance[1], li->second.radiance[2]}, 3);
            place(li->second.primitive, pi->second, transform_of(node), MRH_KEY(MRH_BATCH_EMISSIVE, e));
        }
    }

    void add_camera(const Value& node)
    {
        const std::string t = type_of(node);
        if (t != "Pinhole") {
            unsupported("camera type " + t);
            return;
        }
        mrh_camera_desc c = {};
        double v[3];
        vec3(need(node, "position"), v);
        for (int k = 0; k < 3; k++) c.position[k] = (float)v[k];
        vec3(need(node, "gaze"), v);
        for (int k = 0; k < 3; k++) c.gaze[k] = (float)v[k];
        vec3(need(node, "up"), v);
        for (int k = 0; k < 3; k++) c.up[k] = (float)v[k];
        const auto& fov = need(node, "fov").array();
        c.fov_x = (float)(fov.at(0).number() * M_PI / 180.0);
        const auto& planes = need(node, "planes").array();
        c.near_plane = (float)planes.at(0).number();
        c.far_plane = (float)planes.at(1).number();
        out.cameras.push_back(c);
    }

    void read_cameras(const Value& root)
    {
        const Value* sec = root.get("Cameras");
        if (!sec) return;
        std::map<uint32_t, const Value*> byid;
        for (const Value& node : sec->array())
            for (uint32_t id : ids(node)) byid[id] = &node;
        const Value* cs = root.get("CameraSurfaces");
        if (cs && cs->is_array() && !cs->array().empty()) {
            for (const Value& node : cs->array()) {
                const uint32_t id = (uint32_t)need(node, "camera").number();
                auto it = byid.find(id);
                if (it == byid.end()) throw std::runtime_error("unknown camera " + std::to_string(id));
                add_camera(*it->second);
            }
        } else {
            for (const Value& node : sec->array()) add_camera(node);
        }
    }
};

void copy_msg(char* msg, uint32_t len, const std::string& s)
{
    if (!msg || !len) return;
    const size_t n = std::min<size_t>(len - 1, s.size());
    std::memcpy(msg, s.data(), n);
    msg[n] = 0;
}

}  // namespace

extern "C" {

int32_t mrh_scene_file_load(const char* path, uint32_t flags, mrh_scene_file** out, char* msg, uint32_t msg_len)
{
    if (!path || !out) return MRH_ERR_INVALID;
    *out = nullptr;
    copy_msg(msg, msg_len, "");
    mrh_scene_file* f = new mrh_scene_file();
    try {
        Loader(path, (flags & MRH_SCENE_FILE_SUBSTITUTE) != 0, (flags & MRH_SCENE_FILE_INSTANCED) != 0, *f).run();
    } catch (const Unsupported& e) {
        copy_msg(msg, msg_len, std::string(path) + ": " + e.what());
        delete f;
        return MRH_ERR_UNSUPPORTED;
    } catch (const std::exception& e) {
        copy_msg(msg, msg_len, std::string(path) + ": " + e.what());
        delete f;
        return MRH_ERR_INVALID;
    }
    if (f->instanced) {
        mrh_instancing_desc& I = f->inst_desc;
        I.mesh_first = f->mesh_first.data();
        I.n_meshes = (uint32_t)f->mesh_first.size() - 1;
        I.xform = f->xform.data();
        I.inst_mesh = f->inst_mesh.data();
        I.inst_key = f->inst_key.data();
        I.n_instances = (uint32_t)f->inst_mesh.size();
    }
    *out = f;
    return MRH_OK;
}

int32_t mrh_scene_file_desc(const mrh_scene_file* f, mrh_scene_desc* d)
{
    if (!f || !d) return MRH_ERR_INVALID;
    std::memset(d, 0, sizeof(*d));
    d->tris = f->tris.data();
    d->prim_key = f->keys.data();
    d->n = (uint32_t)f->keys.size();
    d->lambert = f->lambert.data();
    d->n_lambert = (uint32_t)(f->lambert.size() / 3);
    d->metal = f->metal.data();
    d->n_metal = (uint32_t)(f->metal.size() / 10);
    d->dielectric = f->dielectric.data();
    d->n_dielectric = (uint32_t)(f->dielectric.size() / 4);
    d->emissive = f->emissive.data();
    d->n_emissive = (uint32_t)(f->emissive.size() / 3);
    d->reflect = f->reflect.data();
    d->n_reflect = (uint32_t)(f->reflect.size() / 4);
    d->morton_bits = 30;
    if (f->normals.size() == f->tris.size() && !f->normals.empty()) d->normals = f->normals.data();
    if (f->instanced) d->instancing = &f->inst_desc;
    return MRH_OK;
}

int32_t mrh_scene_file_cameras(const mrh_scene_file* f, mrh_camera_desc* out, uint32_t max, uint32_t* count)
{
    if (!f || !count || (max && !out)) return MRH_ERR_INVALID;
    *count = (uint32_t)f->cameras.size();
    for (uint32_t i = 0; i < max && i < *count; i++) out[i] = f->cameras[i];
    return MRH_OK;
}

const char* mrh_scene_file_warnings(const mrh_scene_file* f) { return f ? f->warnings.c_str() : ""; }

int32_t mrh_scene_file_destroy(mrh_scene_file* f)
{
    if (!f) return MRH_ERR_INVALID;
    delete f;
    return MRH_OK;
}

int32_t mrh_obj_mesh_counts(const char* path, uint32_t* meshes, uint32_t* triangles, uint32_t max)
{
    if (!path || !meshes) return MRH_ERR_INVALID;
    std::vector<std::vector<float>> m;
    try {
        if (obj_read(path, &m) != MRH_OK) return MRH_ERR_INVALID;
    } catch (const std::exception&) {
        return MRH_ERR_INVALID;
    }
    *meshes = (uint32_t)m.size();
    for (uint32_t i = 0; i < max && i < m.size() && triangles; i++) triangles[i] = (uint32_t)(m[i].size() / 9);
    return MRH_OK;
}

}  // extern "C"
