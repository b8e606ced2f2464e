// TEST INFRASTRUCTURE ONLY: the CPU oracle built with AddressSanitizer and
// UndefinedBehaviorSanitizer (host code, g++), driven through its C entry points on small
// scenes covering every estimator, material, filter and query: a closed Lambertian box with a
// ceiling light, a rough gold-like conductor (Schlick and exact Fresnel), a glass cube and a
// random triangle soup.  Exits non-zero on a sanitizer report (halt_on_error) or a bad result.
#include "../../oracle/bneept_oracle.cpp"

#include <cstdio>
#include <random>

namespace {

struct SceneBuild {
    std::vector<float> verts;
    std::vector<uint32_t> mat;
    std::vector<int32_t> emit;
    std::vector<orc_material> mats;
    std::vector<orc_emitter> emitters;
    std::vector<uint32_t> etris;
    void tri(const float* a, const float* b, const float* c, uint32_t m, int32_t e) {
        verts.insert(verts.end(), a, a + 3); verts.insert(verts.end(), b, b + 3); verts.insert(verts.end(), c, c + 3);
        mat.push_back(m); emit.push_back(e);
        if (e >= 0) etris.push_back((uint32_t)mat.size() - 1);
    }
    void quad(const float* p0, const float* p1, const float* p2, const float* p3, uint32_t m, int32_t e = -1) {
        tri(p0, p1, p2, m, e); tri(p0, p2, p3, m, e);
    }
    void box(float x0, float y0, float z0, float x1, float y1, float z1, uint32_t m, bool inward) {
        const float c[8][3] = {{x0, y0, z0}, {x1, y0, z0}, {x1, y1, z0}, {x0, y1, z0},
                               {x0, y0, z1}, {x1, y0, z1}, {x1, y1, z1}, {x0, y1, z1}};
        const int f[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4}, {3, 7, 6, 2}, {0, 4, 7, 3}, {1, 2, 6, 5}};
        for (const auto& q : f) {
            if (inward) quad(c[q[3]], c[q[2]], c[q[1]], c[q[0]], m);
            else quad(c[q[0]], c[q[1]], c[q[2]], c[q[3]], m);
        }
    }
    orc_scene scene(const orc_camera& cam) const {
        orc_scene s{};
        s.num_tris = (uint32_t)mat.size(); s.num_materials = (uint32_t)mats.size();
        s.num_emitters = (uint32_t)emitters.size(); s.num_emitter_tris = (uint32_t)etris.size();
        s.verts = verts.data(); s.tri_material = mat.data(); s.tri_emitter = emit.data();
        s.materials = mats.data(); s.emitters = emitters.data(); s.emitter_tris = etris.data();
        s.camera = cam;
        return s;
    }
};

orc_material material(uint32_t type, float r, float g, float b, float rough, float ior) {
    orc_material m{};
    m.type = type; m.rgb[0] = r; m.rgb[1] = g; m.rgb[2] = b; m.roughness = rough; m.ior_outer = 1.0f; m.ior_inner = ior;
    return m;
}

int fails = 0;
void check(bool ok, const char* what) {
    if (!ok) { std::printf("FAIL %s\n", what); ++fails; }
}

}  // namespace

int main() {
    SceneBuild b;
    b.mats.push_back(material(ORC_MAT_ABSORBER, 0, 0, 0, 0, 1));              // 0 light surface
    b.mats.push_back(material(ORC_MAT_LAMBERT, 0.7f, 0.6f, 0.5f, 0, 1));      // 1 walls
    b.mats.push_back(material(ORC_MAT_CONDUCTOR, 1.0f, 0.77f, 0.34f, 0.3f, 1)); // 2 Schlick conductor
    orc_material ex = material(ORC_MAT_CONDUCTOR, 1, 1, 1, 0.4f, 1);
    ex.fresnel = 1; ex.n[0] = 0.143f; ex.n[1] = 0.374f; ex.n[2] = 1.442f; ex.k[0] = 3.983f; ex.k[1] = 2.385f; ex.k[2] = 1.603f;
    b.mats.push_back(ex);                                                      // 3 exact conductor
    b.mats.push_back(material(ORC_MAT_DIELECTRIC, 1, 1, 1, 0.1f, 1.5f));       // 4 glass
    b.box(0, 0, 0, 1, 1, 1, 1, true);
    b.box(0.15f, 0, 0.2f, 0.4f, 0.3f, 0.45f, 2, false);
    b.box(0.55f, 0, 0.5f, 0.8f, 0.5f, 0.75f, 3, false);
    b.box(0.3f, 0.5f, 0.55f, 0.5f, 0.7f, 0.75f, 4, false);
    const float l0[3] = {0.4f, 0.999f, 0.4f}, l1[3] = {0.6f, 0.999f, 0.4f}, l2[3] = {0.6f, 0.999f, 0.6f},
                l3[3] = {0.4f, 0.999f, 0.6f};
    orc_emitter e{};
    e.radiance[0] = 17; e.radiance[1] = 12; e.radiance[2] = 4; e.tri_begin = 0; e.tri_count = 2;
    b.emitters.push_back(e);
    b.quad(l0, l1, l2, l3, 0, 0);
    orc_camera cam{};
    cam.pos[0] = 0.5f; cam.pos[1] = 0.5f; cam.pos[2] = -1.3f;
    cam.right[0] = 1; cam.up[1] = 1; cam.forward[2] = 1; cam.tan_half_fov = 0.36f;
    const orc_scene sc = b.scene(cam);
    const uint32_t W = 12, H = 10;
    std::vector<double> film(8 * W * H);
    for (uint32_t est = 0; est < 3; ++est)
        for (uint32_t cap : {5u, 10000u}) {
            orc_render_opts o{W, H, 0, 4, cap, 2, 7, 1, est};
            orc_render_stats st{};
            std::fill(film.begin(), film.end(), 0.0);
            check(orc_render(&sc, &o, film.data(), &st) == 0, "render");
            check(st.paths == (uint64_t)W * H * 4, "path count");
            for (double v : film) check(std::isfinite(v), "finite film");
        }
    for (int filter : {0, 1, 2, 3, 4}) {
        orc_render_opts o{W, H, 0, 2, 8, 2, 9, 1, 0};
        std::fill(film.begin(), film.end(), 0.0);
        check(orc_render_filtered(&sc, &o, filter, film.data()) == 0, "filtered render");
    }
    {
        std::vector<uint32_t> pix = {0, 5, 17, W * H - 1}, smp = {0, 1, 2, 3};
        std::vector<float> out(3 * pix.size());
        orc_render_opts o{W, H, 0, 0, 10000, 1, 3, 0, 0};
        orc_render_paths(&sc, &o, pix.data(), smp.data(), (uint32_t)pix.size(), out.data());
    }
    // ray queries on a random soup
    std::mt19937 rng(5);
    std::uniform_real_distribution<float> U(-1.0f, 1.0f);
    std::vector<float> soup(9 * 300), rays(8 * 500);
    for (float& v : soup) v = 3.0f * U(rng);
    for (size_t i = 0; i < 500; ++i) {
        float* r = &rays[8 * i];
        r[0] = 4 * U(rng); r[1] = 4 * U(rng); r[2] = 4 * U(rng); r[3] = 0;
        float d[3] = {U(rng), U(rng), U(rng)};
        const float l = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) + 1e-6f;
        r[4] = d[0] / l; r[5] = d[1] / l; r[6] = d[2] / l; r[7] = 3.4e38f;
    }
    std::vector<int32_t> tri(500), tri2(500);
    std::vector<float> t(500), t2(500), bary(1500);
    std::vector<uint32_t> flags(500);
    orc_intersect_brute(soup.data(), 300, rays.data(), 500, tri.data(), t.data(), bary.data(), flags.data(), 2);
    orc_intersect_refbvh(soup.data(), 300, rays.data(), 500, tri2.data(), t2.data(), 2);
    for (size_t i = 0; i < 500; ++i) check(flags[i] & 1 || tri[i] == tri2[i], "reference BVH equals brute force");
    // surface optics of every material
    const float N[3] = {0, 1, 0}, V[3] = {0.6f, 0.8f, 0.0f};
    std::vector<float> u(3 * 256), L(3 * 256), w(3 * 256), rel(256), f(3 * 256), pdf(256);
    std::vector<int32_t> ok(256), ok2(256);
    for (float& x : u) x = 0.5f * (U(rng) + 1.0f);
    for (const orc_material& m : b.mats) {
        orc_bsdf_sample_batch(&m, N, V, 256, u.data(), L.data(), w.data(), rel.data(), ok.data());
        orc_bsdf_eval_batch(&m, N, V, 256, L.data(), f.data(), pdf.data(), ok2.data());
    }
    std::printf("sanitized oracle: %s\n", fails ? "FAILED" : "ok");
    return fails ? 1 : 0;
}
