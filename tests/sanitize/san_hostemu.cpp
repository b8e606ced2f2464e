// TEST INFRASTRUCTURE ONLY: the product's BVH builder, traversal (closest and any hit, with the
// stack spilling past its on-chip ring), guiding-model math and slot order compiled for the
// host with AddressSanitizer and UndefinedBehaviorSanitizer, driven on random and deeply
// nested triangle sets.  Exits non-zero on a sanitizer report or a traversal overflow.
#include "../hostemu/hostemu.cpp"

#include <cmath>
#include <cstdio>
#include <random>

int main() {
    int fails = 0;
    std::mt19937 rng(11);
    std::uniform_real_distribution<float> U(-1.0f, 1.0f);
    for (int kind = 0; kind < 2; ++kind) {
        const uint32_t ntri = kind == 0 ? 2000 : 600;
        std::vector<float> v(9 * ntri);
        for (uint32_t i = 0; i < ntri; ++i) {
            // kind 0: random soup; kind 1: triangles shrinking geometrically towards a point (deep BVH)
            const float s = kind == 0 ? 1.0f : std::pow(0.97f, (float)i);
            const float c[3] = {kind == 0 ? 10 * U(rng) : s, 10 * U(rng) * (kind == 0 ? 1 : 0.02f * s),
                                10 * U(rng) * (kind == 0 ? 1 : 0.02f * s)};
            for (int k = 0; k < 9; ++k) v[9 * i + k] = c[k % 3] + (kind == 0 ? 1.0f : 0.3f * s) * U(rng);
        }
        const uint32_t n = 4000;
        std::vector<float> rays(8 * n);
        for (uint32_t i = 0; i < n; ++i) {
            float* r = &rays[8 * i];
            const float sc = kind == 0 ? 12.0f : 0.05f;
            r[0] = sc * U(rng); r[1] = sc * U(rng); r[2] = sc * U(rng); r[3] = 0;
            float d[3] = {U(rng), U(rng), U(rng)};
            const float l = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) + 1e-6f;
            r[4] = d[0] / l; r[5] = d[1] / l; r[6] = d[2] / l; r[7] = 3.4e38f;
        }
        std::vector<int32_t> prim(n);
        std::vector<float> t(n), bb(n), cc(n);
        std::vector<uint32_t> nv(n), nt(n);
        for (uint32_t leaf : {1u, 4u, 8u})
            for (uint32_t width : {2u, 4u})
                for (int any = 0; any < 2; ++any) {
                    const int need = emu_intersect(v.data(), ntri, rays.data(), n, any, leaf, 16, width, prim.data(), t.data(),
                                                   bb.data(), cc.data(), nv.data(), nt.data());
                    if (need < 0) { std::printf("FAIL traversal overflow (kind %d)\n", kind); ++fails; }
                }
    }
    // guiding-model math on random network outputs
    std::vector<float> o(41 * 64), w3(3 * 64), weight(64), bp(64), sel(64), lrn(64), grad(41 * 64), pdf(64), out3(3 * 64),
        x9(9 * 64), enc(64 * 64), sp(64);
    for (float& x : o) x = 3.0f * U(rng);
    for (uint32_t i = 0; i < 64; ++i) {
        float d[3] = {U(rng), U(rng), U(rng)};
        const float l = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) + 1e-6f;
        for (int k = 0; k < 3; ++k) w3[3 * i + k] = d[k] / l;
        weight[i] = 0.5f * (U(rng) + 1.0f); bp[i] = 0.5f * (U(rng) + 1.0f); sel[i] = 0.5f; lrn[i] = (i & 1u) ? bp[i] : 0.0f;
    }
    for (float& x : x9) x = 0.5f * (U(rng) + 1.0f);
    emu_guide_output_grad(o.data(), 64, w3.data(), weight.data(), bp.data(), sel.data(), lrn.data(), grad.data());
    emu_guide_sel(o.data(), 64, sp.data());
    emu_guide_pdf(o.data(), 64, w3.data(), pdf.data());
    std::vector<float> u3(3 * 64);
    for (float& x : u3) x = 0.5f * (U(rng) + 1.0f);
    emu_guide_sample(o.data(), 64, u3.data(), out3.data());
    emu_guide_encode(x9.data(), 64, enc.data());
    for (float x : pdf) if (!(x >= 0.0f) || !std::isfinite(x)) { std::printf("FAIL guide pdf\n"); ++fails; break; }
    // slot order of ragged image sizes
    for (uint32_t W : {1u, 7u, 9u, 130u})
        for (uint32_t H : {1u, 5u, 17u}) {
            std::vector<uint32_t> pos(W * H), sop(W * H);
            emu_pixel_order(W, H, pos.data(), sop.data());
            for (uint32_t q = 0; q < W * H; ++q)
                if (sop[pos[q]] != q) { std::printf("FAIL slot order %ux%u\n", W, H); ++fails; break; }
        }
    std::printf("sanitized host traversal: %s\n", fails ? "FAILED" : "ok");
    return fails ? 1 : 0;
}
