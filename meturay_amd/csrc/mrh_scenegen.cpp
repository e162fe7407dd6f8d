// mrh_scenegen.cpp — deterministic synthetic workloads (host only).
//
// Random streams are PCG32 (O'Neill 2014) with fixed stream ids per
// quantity, so every generator is reproducible from its seed alone.
#include "mrh.h"

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

namespace {

struct Pcg {
    uint64_t state = 0, inc = 1;
    Pcg(uint64_t seed, uint64_t seq)
    {
        inc = (seq << 1u) | 1u;
        next();
        state += seed;
        next();
    }
    uint32_t next()
    {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    float uniform() { return (float)(next() >> 8) * 0x1.0p-24f; }
};

struct Mesh {
    std::vector<float> tris;
    std::vector<uint32_t> keys;
    void add(const float* a, const float* b, const float* c, uint32_t key)
    {
        tris.insert(tris.end(), a, a + 3);
        tris.insert(tris.end(), b, b + 3);
        tris.insert(tris.end(), c, c + 3);
        keys.push_back(key);
    }
};

struct Tables {
    std::vector<float> lambert, metal, dielectric, emissive;
};

int32_t emit(const Mesh& m, const Tables& t, const mrh_camera_desc& cam, mrh_scene_buffers* out)
{
    if (!out) return MRH_ERR_INVALID;
    const uint32_t n = (uint32_t)m.keys.size();
    const uint32_t nl = (uint32_t)(t.lambert.size() / 3), nm = (uint32_t)(t.metal.size() / 10);
    const uint32_t nd = (uint32_t)(t.dielectric.size() / 4), ne = (uint32_t)(t.emissive.size() / 3);
    const bool fill = out->tris != nullptr;
    if (fill && (out->n < n || out->n_lambert < nl || out->n_metal < nm || out->n_dielectric < nd ||
                 out->n_emissive < ne))
        return MRH_ERR_INVALID;
    out->n = n;
    out->n_lambert = nl;
    out->n_metal = nm;
    out->n_dielectric = nd;
    out->n_emissive = ne;
    out->camera = cam;
    if (!fill) return MRH_OK;
    std::memcpy(out->tris, m.tris.data(), m.tris.size() * 4);
    if (out->prim_key) std::memcpy(out->prim_key, m.keys.data(), m.keys.size() * 4);
    if (out->lambert && nl) std::memcpy(out->lambert, t.lambert.data(), t.lambert.size() * 4);
    if (out->metal && nm) std::memcpy(out->metal, t.metal.data(), t.metal.size() * 4);
    if (out->dielectric && nd) std::memcpy(out->dielectric, t.dielectric.data(), t.dielectric.size() * 4);
    if (out->emissive && ne) std::memcpy(out->emissive, t.emissive.data(), t.emissive.size() * 4);
    return MRH_OK;
}

void quad(Mesh& m, const float* a, const float* b, const float* c, const float* d, uint32_t key)
{
    m.add(a, b, c, key);
    m.add(a, c, d, key);
}

}  // namespace

extern "C" {

int32_t mrh_gen_random_tris(uint32_t n, uint32_t seed, float* tris)
{
    if (!tris && n) return MRH_ERR_INVALID;
    Pcg c(seed, 1), o(seed, 2);
    for (uint32_t i = 0; i < n; i++) {
        float cx = c.uniform(), cy = c.uniform(), cz = c.uniform();
        float* t = tris + (size_t)i * 9;
        for (int k = 0; k < 3; k++) {
            t[3 * k + 0] = cx + (o.uniform() * 2.0f - 1.0f) * 0.01f;
            t[3 * k + 1] = cy + (o.uniform() * 2.0f - 1.0f) * 0.01f;
            t[3 * k + 2] = cz + (o.uniform() * 2.0f - 1.0f) * 0.01f;
        }
    }
    return MRH_OK;
}

int32_t mrh_gen_random_rays(uint32_t n, uint32_t seed, float* rays)
{
    if (!rays && n) return MRH_ERR_INVALID;
    Pcg po(seed, 3), pd(seed, 4);
    const size_t P = n;
    for (uint32_t i = 0; i < n; i++) {
        rays[i] = po.uniform();
        rays[P + i] = po.uniform();
        rays[2 * P + i] = po.uniform();
        // uniform on S^2: z = 1 - 2u, phi = 2 pi v
        double z = 1.0 - 2.0 * (double)pd.uniform();
        double phi = 2.0 * M_PI * (double)pd.uniform();
        double r = std::sqrt(std::fmax(0.0, 1.0 - z * z));
        rays[3 * P + i] = (float)(r * std::cos(phi));
        rays[4 * P + i] = (float)(r * std::sin(phi));
        rays[5 * P + i] = (float)z;
        rays[6 * P + i] = FLT_MAX;
    }
    return MRH_OK;
}

int32_t mrh_gen_cornell(uint32_t subdiv, mrh_scene_buffers* out)
{
    if (subdiv > 7) return MRH_ERR_INVALID;
    Mesh m;
    Tables t;
    // materials: Lambert 0 white, 1 red, 2 blue; emissive 0
    t.lambert = {0.75f, 0.75f, 0.75f, 0.75f, 0.25f, 0.25f, 0.25f, 0.25f, 0.75f};
    t.emissive = {15.0f, 15.0f, 15.0f};
    const uint32_t white = MRH_KEY(MRH_BATCH_LAMBERT, 0), red = MRH_KEY(MRH_BATCH_LAMBERT, 1),
                   blue = MRH_KEY(MRH_BATCH_LAMBERT, 2), light = MRH_KEY(MRH_BATCH_EMISSIVE, 0);
    // closed box [-1,1]^3: 6 walls x 2 triangles
    const float v[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                           {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    quad(m, v[0], v[1], v[5], v[4], white);   // floor  y = -1
    quad(m, v[3], v[7], v[6], v[2], white);   // ceiling y = +1
    quad(m, v[0], v[3], v[2], v[1], white);   // back   z = -1
    quad(m, v[4], v[5], v[6], v[7], white);   // front  z = +1
    quad(m, v[0], v[4], v[7], v[3], red);     // left   x = -1
    quad(m, v[1], v[2], v[6], v[5], blue);    // right  x = +1
    // 2-triangle ceiling light
    const float l0[3] = {-0.25f, 0.99f, -0.25f}, l1[3] = {0.25f, 0.99f, -0.25f},
                l2[3] = {0.25f, 0.99f, 0.25f}, l3[3] = {-0.25f, 0.99f, 0.25f};
    quad(m, l0, l1, l2, l3, light);
    // icosphere of radius 0.3 at the origin
    const float g = 1.6180339887498949f;
    std::vector<std::array<float, 3>> P = {{-1, g, 0}, {1, g, 0}, {-1, -g, 0}, {1, -g, 0},
                                           {0, -1, g}, {0, 1, g}, {0, -1, -g}, {0, 1, -g},
                                           {g, 0, -1}, {g, 0, 1}, {-g, 0, -1}, {-g, 0, 1}};
    for (auto& p : P) {
        float l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        for (auto& c : p) c /= l;
    }
    std::vector<std::array<uint32_t, 3>> F = {
        {0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
        {11, 10, 2}, {10, 7, 6}, {7, 1, 8},  {3, 9, 4},  {3, 4, 2},   {3, 2, 6}, {3, 6, 8},
        {3, 8, 9},  {4, 9, 5},  {2, 4, 11},  {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};
    for (uint32_t s = 0; s < subdiv; s++) {
        std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            std::array<float, 3> c = {(P[a][0] + P[b][0]) * 0.5f, (P[a][1] + P[b][1]) * 0.5f,
                                      (P[a][2] + P[b][2]) * 0.5f};
            float l = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            for (auto& x : c) x /= l;
            P.push_back(c);
            uint32_t id = (uint32_t)P.size() - 1;
            mid[key] = id;
            return id;
        };
        std::vector<std::array<uint32_t, 3>> G;
        G.reserve(F.size() * 4);
        for (auto& f : F) {
            uint32_t a = midpoint(f[0], f[1]), b = midpoint(f[1], f[2]), c = midpoint(f[2], f[0]);
            G.push_back({f[0], a, c});
            G.push_back({f[1], b, a});
            G.push_back({f[2], c, b});
            G.push_back({a, b, c});
        }
        F.swap(G);
    }
    for (auto& f : F) {
        float a[3], b[3], c[3];
        for (int k = 0; k < 3; k++) {
            a[k] = P[f[0]][k] * 0.3f;
            b[k] = P[f[1]][k] * 0.3f;
            c[k] = P[f[2]][k] * 0.3f;
        }
        m.add(a, b, c, white);
    }
    mrh_camera_desc cam = {};
    cam.position[0] = 0.0f; cam.position[1] = 0.0f; cam.position[2] = 0.95f;
    cam.gaze[0] = 0.0f; cam.gaze[1] = 0.0f; cam.gaze[2] = 0.0f;
    cam.up[0] = 0.0f; cam.up[1] = 1.0f; cam.up[2] = 0.0f;
    cam.fov_x = 1.3962634f;   // 80 degrees
    cam.near_plane = 1.0e-3f;
    cam.far_plane = 100.0f;
    return emit(m, t, cam, out);
}

int32_t mrh_gen_sphere_grid(uint32_t grid, uint32_t stacks, uint32_t slices, uint32_t seed,
                            mrh_scene_buffers* out)
{
    if (grid == 0 || stacks < 3 || slices < 3) return MRH_ERR_INVALID;
    Mesh m;
    Tables t;
    Pcg rm(seed, 5), rd(seed, 6);
    const float radius = 0.4f, spacing = 1.0f, amp = 0.05f;
    const double pi = M_PI;
    m.tris.reserve((size_t)grid * grid * 2 * slices * (stacks - 1) * 9);
    m.keys.reserve((size_t)grid * grid * 2 * slices * (stacks - 1));
    std::vector<float> vx((size_t)(stacks + 1) * slices * 3);
    for (uint32_t gi = 0; gi < grid; gi++) {
        for (uint32_t gj = 0; gj < grid; gj++) {
            // material: 5% emissive, else Lambert / GGX conductor / dielectric
            uint32_t key;
            float u = rm.uniform();
            if (u < 0.05f) {
                key = MRH_KEY(MRH_BATCH_EMISSIVE, t.emissive.size() / 3);
                t.emissive.insert(t.emissive.end(), {10.0f, 10.0f, 10.0f});
            } else {
                uint32_t kind = rm.next() % 3u;
                if (kind == 0) {
                    key = MRH_KEY(MRH_BATCH_LAMBERT, t.lambert.size() / 3);
                    for (int c = 0; c < 3; c++) t.lambert.push_back(0.1f + 0.8f * rm.uniform());
                } else if (kind == 1) {
                    key = MRH_KEY(MRH_BATCH_METAL, t.metal.size() / 10);
                    float r = 0.05f + 0.45f * rm.uniform();
                    // gold-like conductor: eta, k per RGB; white specular tint
                    t.metal.insert(t.metal.end(), {r, 0.143f, 0.374f, 1.442f, 3.983f, 2.385f, 1.603f,
                                                   1.0f, 1.0f, 1.0f});
                } else {
                    key = MRH_KEY(MRH_BATCH_DIELECTRIC, t.dielectric.size() / 4);
                    t.dielectric.insert(t.dielectric.end(), {1.0f, 1.0f, 1.0f, 1.5f});
                }
            }
            // displaced UV sphere: r(theta, phi) = R + A sin(f1 theta + p1) sin(f2 phi + p2)
            const double f1 = 2.0 + (double)(rd.next() % 4u), f2 = 2.0 + (double)(rd.next() % 4u);
            const double p1 = 2.0 * pi * (double)rd.uniform(), p2 = 2.0 * pi * (double)rd.uniform();
            const double cx = ((double)gi - (grid - 1) * 0.5) * spacing;
            const double cz = ((double)gj - (grid - 1) * 0.5) * spacing;
            const double cy = 0.0;
            auto vertex = [&](uint32_t k, uint32_t l, float* o) {
                double th = pi * (double)k / (double)stacks;
                double ph = 2.0 * pi * (double)(l % slices) / (double)slices;
                if (k == 0 || k == stacks) ph = 0.0;   // single pole vertex
                double r = radius + amp * std::sin(f1 * th + p1) * std::sin(f2 * ph + p2);
                o[0] = (float)(cx + r * std::sin(th) * std::cos(ph));
                o[1] = (float)(cy + r * std::cos(th));
                o[2] = (float)(cz + r * std::sin(th) * std::sin(ph));
            };
            for (uint32_t k = 0; k <= stacks; k++)
                for (uint32_t l = 0; l < slices; l++) vertex(k, l, &vx[((size_t)k * slices + l) * 3]);
            auto V = [&](uint32_t k, uint32_t l) { return &vx[((size_t)k * slices + (l % slices)) * 3]; };
            for (uint32_t l = 0; l < slices; l++) m.add(V(0, 0), V(1, l), V(1, l + 1), key);
            for (uint32_t k = 1; k + 1 < stacks; k++)
                for (uint32_t l = 0; l < slices; l++) {
                    m.add(V(k, l), V(k + 1, l), V(k + 1, l + 1), key);
                    m.add(V(k, l), V(k + 1, l + 1), V(k, l + 1), key);
                }
            for (uint32_t l = 0; l < slices; l++) m.add(V(stacks - 1, l), V(stacks, 0), V(stacks - 1, l + 1), key);
        }
    }
    const float half = (grid - 1) * 0.5f * spacing;
    mrh_camera_desc cam = {};
    cam.position[0] = 0.0f; cam.position[1] = 0.45f * half + 2.0f; cam.position[2] = 1.1f * half + 3.0f;
    cam.gaze[0] = 0.0f; cam.gaze[1] = 0.0f; cam.gaze[2] = -0.1f * half;
    cam.up[0] = 0.0f; cam.up[1] = 1.0f; cam.up[2] = 0.0f;
    cam.fov_x = 1.0471976f;   // 60 degrees
    cam.near_plane = 1.0e-3f;
    cam.far_plane = 1.0e4f;
    return emit(m, t, cam, out);
}

}  // extern "C"
