// Procedural workloads as flattened scenes (phg_scene_generate).
//
// Cornell box: 5 walls + 2 boxes (34 triangles) + a 0.25 x 0.25 ceiling light (2 triangles).
// Procedural: displaced icosphere instances + a tessellated ground + area-light quads.
// Occluded room: a room lit only through a 0.1-wide slit in an interior wall.
// Terrain: displaced icospheres over an FBM height-field grid, 20M triangles by default.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <utility>
#include <vector>

#include "../../include/phg.h"
#include "scene_buf.h"

namespace {

using namespace phg::sb;


// splitmix64-based deterministic RNG for scene parameters
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ULL + 0x1234567ULL) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double u() { return (double)(next() >> 11) * 0x1p-53; }
    float uf(float a, float b) { return (float)(a + (b - a) * u()); }
};

// 3-D gradient noise on an integer lattice with hashed gradients, and its fractal sum.
static uint32_t hash3(int x, int y, int z, uint32_t seed) {
    uint32_t h = seed * 0x27d4eb2dU;
    h ^= (uint32_t)x * 0x8da6b343U; h = (h << 13) | (h >> 19);
    h ^= (uint32_t)y * 0xd8163841U; h = (h << 13) | (h >> 19);
    h ^= (uint32_t)z * 0xcb1ab31fU;
    h ^= h >> 16; h *= 0x85ebca6bU; h ^= h >> 13; h *= 0xc2b2ae35U; h ^= h >> 16;
    return h;
}
static float grad(uint32_t h, float x, float y, float z) {
    switch (h % 12u) {
        case 0: return x + y; case 1: return -x + y; case 2: return x - y; case 3: return -x - y;
        case 4: return x + z; case 5: return -x + z; case 6: return x - z; case 7: return -x - z;
        case 8: return y + z; case 9: return -y + z; case 10: return y - z; default: return -y - z;
    }
}
static float fade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }
static float lerpf(float a, float b, float t) { return a + (b - a) * t; }
static float gnoise(F3 p, uint32_t seed) {
    const int xi = (int)std::floor(p.x), yi = (int)std::floor(p.y), zi = (int)std::floor(p.z);
    const float xf = p.x - xi, yf = p.y - yi, zf = p.z - zi;
    const float u = fade(xf), v = fade(yf), w = fade(zf);
    float c[8];
    for (int k = 0; k < 8; ++k) {
        const int dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
        c[k] = grad(hash3(xi + dx, yi + dy, zi + dz, seed), xf - dx, yf - dy, zf - dz);
    }
    const float x00 = lerpf(c[0], c[1], u), x10 = lerpf(c[2], c[3], u);
    const float x01 = lerpf(c[4], c[5], u), x11 = lerpf(c[6], c[7], u);
    return lerpf(lerpf(x00, x10, v), lerpf(x01, x11, v), w);
}
static float fbm(F3 p, uint32_t seed, int octaves) {
    float sum = 0.0f, amp = 0.5f;
    for (int o = 0; o < octaves; ++o) {
        sum += amp * gnoise(p, seed + (uint32_t)o * 101u);
        p = mul(p, 2.0f);
        amp *= 0.5f;
    }
    return sum;
}

static void gen_cornell(SceneBuf& s, const phg_gen_opts& o) {
    const uint32_t absorber = s.add_material(PHG_MAT_ABSORBER, f3(0, 0, 0), 0, 1, 1);
    const uint32_t white = s.add_material(PHG_MAT_LAMBERT, f3(0.73f, 0.73f, 0.73f), 0, 1, 1);
    const uint32_t red = s.add_material(PHG_MAT_LAMBERT, f3(0.65f, 0.05f, 0.05f), 0, 1, 1);
    const uint32_t green = s.add_material(PHG_MAT_LAMBERT, f3(0.12f, 0.45f, 0.15f), 0, 1, 1);
    // walls of the unit cube, normals facing inwards
    s.quad(f3(0, 0, 0), f3(0, 0, 1), f3(1, 0, 1), f3(1, 0, 0), white);  // floor y=0 (+y)
    s.quad(f3(0, 1, 0), f3(1, 1, 0), f3(1, 1, 1), f3(0, 1, 1), white);  // ceiling y=1 (-y)
    s.quad(f3(0, 0, 1), f3(0, 1, 1), f3(1, 1, 1), f3(1, 0, 1), white);  // back z=1 (-z)
    s.quad(f3(1, 0, 0), f3(1, 0, 1), f3(1, 1, 1), f3(1, 1, 0), red);    // x=1 (-x), left in view
    s.quad(f3(0, 0, 0), f3(0, 1, 0), f3(0, 1, 1), f3(0, 0, 1), green);  // x=0 (+x), right in view
    s.box(0.35f, 0.35f, 0.3f, 0.3f, 0.3f, -17.0f, white);
    s.box(0.65f, 0.65f, 0.3f, 0.6f, 0.3f, 17.0f, white);
    s.light_down(0.375f, 0.625f, 0.375f, 0.625f, 1.0f - 1e-4f, f3(17, 12, 4), absorber);
    const float fov = o.fov_deg > 0 ? o.fov_deg : 40.0f;
    s.camera(f3(0.5f, 0.5f, -1.374f), f3(0.5f, 0.5f, 0.0f), f3(0, 1, 0), fov);
}

static void add_displaced_sphere(SceneBuf& s, const Icosphere& ico, F3 c, float radius, float amp,
                                 uint32_t noise_seed, uint32_t mat) {
    std::vector<F3> p(ico.v.size());
    for (size_t i = 0; i < ico.v.size(); ++i) {
        const F3 n = ico.v[i];
        const float d = fbm(mul(n, 2.0f), noise_seed, 4);
        p[i] = add(c, mul(n, radius + amp * d));
    }
    for (size_t i = 0; i < ico.f.size(); i += 3) s.tri(p[ico.f[i]], p[ico.f[i + 1]], p[ico.f[i + 2]], mat);
}

static void gen_procedural(SceneBuf& s, const phg_gen_opts& o) {
    Rng rng(o.seed ? o.seed : 2025);
    const uint32_t ninst = o.num_instances ? o.num_instances : 64;
    const uint32_t nlights = o.num_lights ? o.num_lights : 16;
    const uint32_t subdiv = o.subdiv ? o.subdiv : 4;
    const uint32_t target = o.target_tris ? o.target_tris : (1u << 20);
    const Icosphere ico = make_icosphere(subdiv);
    const uint32_t absorber = s.add_material(PHG_MAT_ABSORBER, f3(0, 0, 0), 0, 1, 1);
    const uint32_t ground = s.add_material(PHG_MAT_LAMBERT, f3(0.6f, 0.6f, 0.6f), 0, 1, 1);
    for (uint32_t i = 0; i < ninst; ++i) {
        const F3 c = f3(rng.uf(-10, 10), rng.uf(-10, 10), rng.uf(-10, 10));
        const float radius = rng.uf(0.5f, 1.5f);
        const float rough = rng.uf(0.05f, 0.6f);
        uint32_t mat;
        if (i % 2 == 0) {
            mat = s.add_material(PHG_MAT_CONDUCTOR, f3(rng.uf(0.5f, 1), rng.uf(0.5f, 1), rng.uf(0.5f, 1)), rough, 1, 1);
        } else {
            mat = s.add_material(PHG_MAT_DIELECTRIC, f3(1, 1, 1), rough, 1.0f, rng.uf(1.3f, 1.7f));
        }
        add_displaced_sphere(s, ico, c, radius, 0.1f, (uint32_t)(o.seed * 977u + i), mat);
    }
    // ground plane at y = -12 tessellated so that spheres + ground reach the triangle target
    // (2 per cell); the light quads come on top
    const uint32_t have = (uint32_t)s.tri_mat.size();
    uint32_t cells = target > have ? (target - have) / 2 : 1;
    uint32_t nz = (uint32_t)std::sqrt((double)cells * 512.0 / 704.0);
    if (nz == 0) nz = 1;
    uint32_t nx = cells / nz;
    if (nx == 0) nx = 1;
    // (ex, 0, 0) x (0, 0, ez) -> -y; build with eu = +z, ev = +x so the normal is +y
    s.grid(f3(-30, -12, -30), f3(0, 0, 60), f3(60, 0, 0), nz, nx, ground);
    for (uint32_t l = 0; l < nlights; ++l) {
        const float x = rng.uf(-12, 10), z = rng.uf(-12, 10);
        const float r = rng.uf(1, 20);
        s.light_down(x, x + 2.0f, z, z + 2.0f, 13.0f, f3(r, r, r), absorber);
    }
    s.camera(f3(0, 12, -24), f3(0, -6, 0), f3(0, 1, 0), o.fov_deg > 0 ? o.fov_deg : 60.0f);
}

// Large procedural world: displaced icosphere instances over an FBM terrain grid, mixed GGX
// conductors and dielectrics, many area lights.
static void gen_terrain(SceneBuf& s, const phg_gen_opts& o) {
    Rng rng(o.seed ? o.seed : 4242);
    const uint32_t ninst = o.num_instances ? o.num_instances : 256;
    const uint32_t nlights = o.num_lights ? o.num_lights : 64;
    const uint32_t subdiv = o.subdiv ? o.subdiv : 4;
    const uint32_t target = o.target_tris ? o.target_tris : 20000000u;
    const Icosphere ico = make_icosphere(subdiv);
    const uint32_t absorber = s.add_material(PHG_MAT_ABSORBER, f3(0, 0, 0), 0, 1, 1);
    const uint32_t ground = s.add_material(PHG_MAT_CONDUCTOR, f3(0.55f, 0.5f, 0.45f), 0.7f, 1, 1);
    const uint32_t have_est = ninst * (uint32_t)(ico.f.size() / 3) + 2 * nlights;
    const uint32_t cells = target > have_est ? (target - have_est) / 2 : 1;
    const uint32_t n = std::max<uint32_t>(1, (uint32_t)std::sqrt((double)cells));
    const float x0 = -50.0f, size = 100.0f, base = -12.0f, amp = 4.0f;
    s.verts.reserve(9ull * ((size_t)2 * n * n + have_est));
    s.tri_mat.reserve((size_t)2 * n * n + have_est);
    s.tri_emit.reserve((size_t)2 * n * n + have_est);
    auto height = [&](float x, float z) { return base + amp * fbm(f3(x * 0.04f, 0.0f, z * 0.04f), o.seed * 31u + 7u, 5); };
    std::vector<float> row0(n + 1), row1(n + 1);
    for (uint32_t j = 0; j <= n; ++j) row0[j] = height(x0 + size * (float)j / n, x0);
    for (uint32_t i = 0; i < n; ++i) {
        const float z0 = x0 + size * (float)i / n, z1 = x0 + size * (float)(i + 1) / n;
        for (uint32_t j = 0; j <= n; ++j) row1[j] = height(x0 + size * (float)j / n, z1);
        for (uint32_t j = 0; j < n; ++j) {
            const float xa = x0 + size * (float)j / n, xb = x0 + size * (float)(j + 1) / n;
            // counter-clockwise seen from +y
            s.quad(f3(xa, row0[j], z0), f3(xa, row1[j], z1), f3(xb, row1[j + 1], z1), f3(xb, row0[j + 1], z0), ground);
        }
        std::swap(row0, row1);
    }
    for (uint32_t i = 0; i < ninst; ++i) {
        const float x = rng.uf(-40, 40), z = rng.uf(-40, 40);
        const float radius = rng.uf(0.6f, 2.0f);
        const F3 c = f3(x, height(x, z) + radius + rng.uf(0.0f, 6.0f), z);
        const float rough = rng.uf(0.05f, 0.6f);
        uint32_t mat;
        if (i % 2 == 0) {
            mat = s.add_material(PHG_MAT_CONDUCTOR, f3(rng.uf(0.5f, 1), rng.uf(0.5f, 1), rng.uf(0.5f, 1)), rough, 1, 1);
        } else {
            mat = s.add_material(PHG_MAT_DIELECTRIC, f3(1, 1, 1), rough, 1.0f, rng.uf(1.3f, 1.7f));
        }
        add_displaced_sphere(s, ico, c, radius, 0.15f, (uint32_t)(o.seed * 977u + i), mat);
    }
    for (uint32_t l = 0; l < nlights; ++l) {
        const float x = rng.uf(-45, 42), z = rng.uf(-45, 42);
        const float r = rng.uf(1, 20);
        s.light_down(x, x + 3.0f, z, z + 3.0f, 20.0f, f3(r, r, r), absorber);
    }
    s.camera(f3(0, 10, -60), f3(0, -8, 0), f3(0, 1, 0), o.fov_deg > 0 ? o.fov_deg : 60.0f);
}

static void gen_room(SceneBuf& s, const phg_gen_opts& o) {
    Rng rng(o.seed ? o.seed : 7);
    const uint32_t target = o.target_tris ? o.target_tris : 200000;
    const uint32_t absorber = s.add_material(PHG_MAT_ABSORBER, f3(0, 0, 0), 0, 1, 1);
    const uint32_t wall = s.add_material(PHG_MAT_CONDUCTOR, f3(0.75f, 0.75f, 0.75f), rng.uf(0.3f, 0.5f), 1, 1);
    const uint32_t floor_m = s.add_material(PHG_MAT_CONDUCTOR, f3(0.6f, 0.55f, 0.5f), rng.uf(0.2f, 0.5f), 1, 1);
    const uint32_t inner = s.add_material(PHG_MAT_CONDUCTOR, f3(0.7f, 0.7f, 0.7f), rng.uf(0.3f, 0.5f), 1, 1);
    const uint32_t tess = 40;
    // room [0,4] x [0,3] x [0,4]; normals facing inwards
    s.grid(f3(0, 0, 0), f3(0, 0, 4), f3(4, 0, 0), tess, tess, floor_m);   // floor (+y)
    s.grid(f3(0, 3, 0), f3(4, 0, 0), f3(0, 0, 4), tess, tess, wall);      // ceiling (-y)
    s.grid(f3(0, 0, 4), f3(0, 3, 0), f3(4, 0, 0), tess, tess, wall);      // z = 4 (-z)
    s.grid(f3(0, 0, 0), f3(4, 0, 0), f3(0, 3, 0), tess, tess, wall);      // z = 0 (+z)
    s.grid(f3(0, 0, 0), f3(0, 3, 0), f3(0, 0, 4), tess, tess, wall);      // x = 0 (+x)
    s.grid(f3(4, 0, 0), f3(0, 0, 4), f3(0, 3, 0), tess, tess, wall);      // x = 4 (-x)
    // interior wall at x in [2.95, 3.05] with a full-height slit z in [1.95, 2.05]
    const float wx0 = 2.95f, wx1 = 3.05f, g0 = 1.95f, g1 = 2.05f;
    s.grid(f3(wx0, 0, 0), f3(0, 3, 0), f3(0, 0, g0), 8, 8, inner);       // face towards the room, part 1
    s.grid(f3(wx0, 0, g1), f3(0, 3, 0), f3(0, 0, 4 - g1), 8, 8, inner);  // part 2
    s.grid(f3(wx1, 0, 0), f3(0, 0, g0), f3(0, 3, 0), 8, 8, inner);       // chamber side
    s.grid(f3(wx1, 0, g1), f3(0, 0, 4 - g1), f3(0, 3, 0), 8, 8, inner);
    s.grid(f3(wx0, 0, g0), f3(wx1 - wx0, 0, 0), f3(0, 3, 0), 1, 8, inner);  // slit jambs
    s.grid(f3(wx0, 0, g1), f3(0, 3, 0), f3(wx1 - wx0, 0, 0), 8, 1, inner);
    // the hidden light: 0.5 x 0.5, radiance 50, chamber ceiling, facing down
    s.light_down(3.275f, 3.775f, 0.75f, 1.25f, 2.999f, f3(50, 50, 50), absorber);
    // furniture: displaced spheres to reach the triangle target
    const Icosphere ico = make_icosphere(4);
    uint32_t i = 0;
    while (s.tri_mat.size() + ico.f.size() / 3 <= target) {
        const F3 c = f3(rng.uf(0.5f, 2.5f), rng.uf(0.3f, 1.2f), rng.uf(0.5f, 3.5f));
        const float rough = rng.uf(0.1f, 0.5f);
        const uint32_t mat = s.add_material(PHG_MAT_CONDUCTOR, f3(rng.uf(0.4f, 0.95f), rng.uf(0.4f, 0.95f), rng.uf(0.4f, 0.95f)), rough, 1, 1);
        add_displaced_sphere(s, ico, c, rng.uf(0.1f, 0.3f), 0.02f, 31u * i + o.seed, mat);
        ++i;
    }
    s.camera(f3(0.3f, 1.6f, 0.3f), f3(3.0f, 1.2f, 2.4f), f3(0, 1, 0), o.fov_deg > 0 ? o.fov_deg : 70.0f);
}

}  // namespace

extern "C" phg_status phg_scene_generate(const phg_gen_opts* opts, phg_scene_host** out) {
    if (!opts || !out) return PHG_ERR_INVALID_ARG;
    SceneBuf* s = new (std::nothrow) SceneBuf();
    if (!s) return PHG_ERR_OUT_OF_MEMORY;
    std::memset(&s->host, 0, sizeof(s->host));
    try {
        switch (opts->kind) {
            case PHG_GEN_CORNELL: gen_cornell(*s, *opts); break;
            case PHG_GEN_PROCEDURAL: gen_procedural(*s, *opts); break;
            case PHG_GEN_OCCLUDED_ROOM: gen_room(*s, *opts); break;
            case PHG_GEN_TERRAIN: gen_terrain(*s, *opts); break;
            default: delete s; return PHG_ERR_INVALID_ARG;
        }
    } catch (const std::bad_alloc&) {
        delete s;
        return PHG_ERR_OUT_OF_MEMORY;
    }
    s->finish();
    *out = &s->host;  // host is the first member: the pointer converts back in phg_scene_free
    return PHG_OK;
}

extern "C" void phg_scene_free(phg_scene_host* scene) {
    if (!scene) return;
    delete reinterpret_cast<SceneBuf*>(scene);
}
