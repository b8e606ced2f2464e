// Flattened-scene builder shared by the procedural generators and the .p2 loader: triangle
// soup with per-triangle material / emitter ids, material and emitter tables, pinhole camera,
// exposed through phg_scene_host (include/phg.h).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

#include "../../include/phg.h"

namespace phg {
namespace sb {

struct F3 { float x, y, z; };
inline F3 f3(float x, float y, float z) { return F3{x, y, z}; }
inline F3 add(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline F3 sub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline F3 mul(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
inline F3 crs(F3 a, F3 b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline F3 nrm(F3 a) { const float l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); return mul(a, 1.0f / l); }
inline float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

struct Icosphere {
    std::vector<F3> v;
    std::vector<uint32_t> f;  // 3 per face
};

inline Icosphere make_icosphere(uint32_t subdiv) {
    Icosphere s;
    const float t = (1.0f + std::sqrt(5.0f)) / 2.0f;
    const F3 base[12] = {f3(-1, t, 0), f3(1, t, 0), f3(-1, -t, 0), f3(1, -t, 0), f3(0, -1, t), f3(0, 1, t),
                         f3(0, -1, -t), f3(0, 1, -t), f3(t, 0, -1), f3(t, 0, 1), f3(-t, 0, -1), f3(-t, 0, 1)};
    for (const F3& p : base) s.v.push_back(nrm(p));
    const uint32_t faces[60] = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2,
                                10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5,
                                2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    s.f.assign(faces, faces + 60);
    for (uint32_t it = 0; it < subdiv; ++it) {
        std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            const auto key = a < b ? std::make_pair(a, b) : std::make_pair(b, a);
            auto found = mid.find(key);
            if (found != mid.end()) return found->second;
            const uint32_t id = (uint32_t)s.v.size();
            s.v.push_back(nrm(mul(add(s.v[a], s.v[b]), 0.5f)));
            mid[key] = id;
            return id;
        };
        std::vector<uint32_t> nf;
        nf.reserve(s.f.size() * 4);
        for (size_t i = 0; i < s.f.size(); i += 3) {
            const uint32_t a = s.f[i], b = s.f[i + 1], c = s.f[i + 2];
            const uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            const uint32_t q[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            nf.insert(nf.end(), q, q + 12);
        }
        s.f.swap(nf);
    }
    return s;
}

struct SceneBuf {
    phg_scene_host host;
    std::vector<float> verts;
    std::vector<uint32_t> tri_mat;
    std::vector<int32_t> tri_emit;
    std::vector<phg_material> mats;
    std::vector<phg_emitter> emits;
    std::vector<uint32_t> emit_tris;

    uint32_t add_material(uint32_t type, F3 c, float rough, float ior_out, float ior_in) {
        phg_material m{};
        m.type = type; m.rgb[0] = c.x; m.rgb[1] = c.y; m.rgb[2] = c.z;
        m.roughness = rough; m.ior_outer = ior_out; m.ior_inner = ior_in;
        mats.push_back(m);
        return (uint32_t)mats.size() - 1;
    }
    // conductor with the exact Fresnel of a complex inner index n + ik (outer index ior_out)
    uint32_t add_exact_conductor(F3 n, F3 k, float rough, float ior_out) {
        phg_material m{};
        m.type = PHG_MAT_CONDUCTOR;
        m.rgb[0] = m.rgb[1] = m.rgb[2] = 1.0f;
        m.roughness = rough; m.ior_outer = ior_out; m.ior_inner = 1.0f;
        m.fresnel = PHG_FRESNEL_EXACT;
        m.n[0] = n.x; m.n[1] = n.y; m.n[2] = n.z;
        m.k[0] = k.x; m.k[1] = k.y; m.k[2] = k.z;
        mats.push_back(m);
        return (uint32_t)mats.size() - 1;
    }
    void tri(F3 a, F3 b, F3 c, uint32_t mat, int32_t emitter = -1) {
        const float v[9] = {a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z};
        verts.insert(verts.end(), v, v + 9);
        tri_mat.push_back(mat);
        tri_emit.push_back(emitter);
        if (emitter >= 0) emit_tris.push_back((uint32_t)tri_mat.size() - 1);
    }
    // quad p0 p1 p2 p3 (counter-clockwise seen from the side its normal points to)
    void quad(F3 p0, F3 p1, F3 p2, F3 p3, uint32_t mat, int32_t emitter = -1) {
        tri(p0, p1, p2, mat, emitter);
        tri(p0, p2, p3, mat, emitter);
    }
    // tessellated quad with nu x nv cells
    void grid(F3 p0, F3 eu, F3 ev, uint32_t nu, uint32_t nv, uint32_t mat) {
        for (uint32_t j = 0; j < nv; ++j)
            for (uint32_t i = 0; i < nu; ++i) {
                const F3 a = add(p0, add(mul(eu, (float)i / nu), mul(ev, (float)j / nv)));
                const F3 b = add(p0, add(mul(eu, (float)(i + 1) / nu), mul(ev, (float)j / nv)));
                const F3 c = add(p0, add(mul(eu, (float)(i + 1) / nu), mul(ev, (float)(j + 1) / nv)));
                const F3 d = add(p0, add(mul(eu, (float)i / nu), mul(ev, (float)(j + 1) / nv)));
                quad(a, b, c, d, mat);
            }
    }
    // axis-aligned light quad facing -y at height y
    void light_down(float x0, float x1, float z0, float z1, float y, F3 radiance, uint32_t mat) {
        const int32_t e = (int32_t)emits.size();
        phg_emitter em{};
        em.radiance[0] = radiance.x; em.radiance[1] = radiance.y; em.radiance[2] = radiance.z;
        em.tri_begin = (uint32_t)emit_tris.size();
        em.tri_count = 2;
        emits.push_back(em);
        // (x1-x0, 0, 0) x (0, 0, z1-z0) points to -y
        tri(f3(x0, y, z0), f3(x1, y, z0), f3(x0, y, z1), mat, e);
        tri(f3(x1, y, z0), f3(x1, y, z1), f3(x0, y, z1), mat, e);
    }
    // box of size (sx, sy, sz) standing on y = 0, rotated by deg around y, centred at (cx, cz)
    void box(float cx, float cz, float sx, float sy, float sz, float deg, uint32_t mat) {
        const float r = deg * 3.14159265358979f / 180.0f, c = std::cos(r), s = std::sin(r);
        auto P = [&](float x, float y, float z) {
            const float px = x - sx * 0.5f, pz = z - sz * 0.5f;
            return f3(cx + c * px + s * pz, y, cz - s * px + c * pz);
        };
        const F3 v000 = P(0, 0, 0), v100 = P(sx, 0, 0), v010 = P(0, sy, 0), v110 = P(sx, sy, 0);
        const F3 v001 = P(0, 0, sz), v101 = P(sx, 0, sz), v011 = P(0, sy, sz), v111 = P(sx, sy, sz);
        quad(v010, v011, v111, v110, mat);  // top (+y)
        quad(v000, v100, v101, v001, mat);  // bottom (-y)
        quad(v000, v010, v110, v100, mat);  // -z
        quad(v001, v101, v111, v011, mat);  // +z
        quad(v000, v001, v011, v010, mat);  // -x
        quad(v100, v110, v111, v101, mat);  // +x
    }
    void camera(F3 pos, F3 target, F3 up_hint, float fov_deg) {
        const F3 fwd = nrm(sub(target, pos));
        const F3 right = nrm(crs(fwd, up_hint));
        const F3 up = crs(right, fwd);
        phg_camera& c = host.desc.camera;
        std::memset(&c, 0, sizeof(c));
        c.pos[0] = pos.x; c.pos[1] = pos.y; c.pos[2] = pos.z;
        c.right[0] = right.x; c.right[1] = right.y; c.right[2] = right.z;
        c.up[0] = up.x; c.up[1] = up.y; c.up[2] = up.z;
        c.forward[0] = fwd.x; c.forward[1] = fwd.y; c.forward[2] = fwd.z;
        c.tan_half_fov = (float)std::tan((double)fov_deg * 3.14159265358979323846 / 360.0);
    }
    void finish() {
        phg_scene_desc& d = host.desc;
        d.num_tris = (uint32_t)tri_mat.size();
        d.num_materials = (uint32_t)mats.size();
        d.num_emitters = (uint32_t)emits.size();
        d.num_emitter_tris = (uint32_t)emit_tris.size();
        d.verts = verts.data();
        d.tri_material = tri_mat.data();
        d.tri_emitter = tri_emit.data();
        d.materials = mats.data();
        d.emitters = emits.data();
        d.emitter_tris = emit_tris.data();
    }
};

}  // namespace sb
}  // namespace phg
