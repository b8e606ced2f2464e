// C ABI of the BNEEPT hot path (include/phg.h): context, scene upload, batched ray
// queries, wavefront render loop, FP64 film, statistics.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "../../include/phg.h"
#include "bvh_build.h"
#include "guide.h"
#include "kernels.h"

namespace phg {
namespace {

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;  // elements
    void free_() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
    hipError_t ensure(size_t count) {
        if (count <= n && p) return hipSuccess;
        free_();
        void* q = nullptr;
        const hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
        if (e != hipSuccess) return e;
        p = static_cast<T*>(q); n = std::max<size_t>(count, 1);
        return hipSuccess;
    }
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool have_scene = false;
    // scene
    DevBuf<WideNode> nodes;
    DevBuf<TriRec> tris;
    DevBuf<phg_material> mats;
    DevBuf<EmitterRec> emitters;
    DevBuf<float> emit_cdf;
    DevBuf<EmitTri> emit_tris;
    SceneDev sc{};
    uint32_t stack_need = 0, bvh_depth = 0, bvh_nodes = 0;
    double bvh_build_ms = 0;
    // film
    DevBuf<double> film;
    uint32_t film_w = 0, film_h = 0;
    // wavefront buffers
    DevBuf<uint64_t> rng;
    DevBuf<float4> T, Wem, prevP, L, hits, shadow_c;
    DevBuf<phg_ray> ext0, ext1, shadow;
    DevBuf<uint32_t> ext_pid0, ext_pid1, shadow_pid;
    DevBuf<uint32_t> counts;
    DevBuf<int32_t> spill;       // traversal-stack spill of the persistent traversal grid
    DevBuf<unsigned long long> tcount;
    DevBuf<uint32_t> errflags;
    // intersect staging
    DevBuf<phg_ray> q_rays;
    DevBuf<phg_hit> q_hits;
    DevBuf<uint32_t> q_nvis, q_ntri;
    // guiding network
    DevBuf<float> g_params, g_m, g_v, g_grads, g_bias;
    DevBuf<uint16_t> g_wt, g_wn;
    DevBuf<float> g_lobes;       // per queued extension ray
    DevBuf<float4> g_rec;        // per path, per recorded bounce
    DevBuf<float> g_train;       // collected training records
    DevBuf<float> g_q;           // staging for the guide entry points
    DevBuf<uint32_t> g_counters;
    bool g_init = false, g_ready = false;
    uint32_t g_steps = 0, g_pass_salt = 0;
    float g_keep = 1.0f;
    GuideScene g_scene{};
    // stats
    phg_stats stats{};
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    uint32_t cu_count = 256;
    uint32_t trace_blocks_per_cu = 8;   // resident blocks per compute unit, from the occupancy query
    uint32_t shade_blocks_per_cu = 4;
};

#define PHG_HIP(ctx, expr)                                                                        \
    do {                                                                                          \
        const hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                                   \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                       \
            return e_ == hipErrorOutOfMemory ? PHG_ERR_OUT_OF_MEMORY : PHG_ERR_HIP;               \
        }                                                                                         \
    } while (0)

// ---- host-side float math identical to the device (built with -ffp-contract=off) --------
struct H3 { float x, y, z; };
static inline H3 hsub(H3 a, H3 b) { return H3{a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline H3 hcross(H3 a, H3 b) { return H3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
static inline float hdot(H3 a, H3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float htri_area(H3 a, H3 b, H3 c) { return std::sqrt(hdot(hcross(hsub(b, a), hsub(c, a)), hcross(hsub(b, a), hsub(c, a)))) * 0.5f; }
static inline float hluminance(const float* c) {
    return c[0] * 0.21263900587151027f + c[1] * 0.715168678767756f + c[2] * 0.07219231536073371f;
}

static phg_status set_err(Ctx* c, phg_status s, const std::string& msg) { c->err = msg; return s; }

static GuideNetDev guide_net(Ctx* c) {
    GuideNetDev n{};
    n.params = c->g_params.p; n.wt = c->g_wt.p; n.wn = c->g_wn.p; n.bias = c->g_bias.p;
    n.params_mut = c->g_params.p; n.wt_mut = c->g_wt.p; n.wn_mut = c->g_wn.p; n.bias_mut = c->g_bias.p;
    return n;
}

// He-uniform initialisation (fan-in), zero biases, from a splitmix64 stream.
static std::vector<float> guide_default_params(uint64_t seed) {
    std::vector<float> p(kGuideParamCount, 0.0f);
    const int offs[11] = {0, 4096, 4160, 8256, 8320, 12416, 12480, 16576, 16640, 19200, 19240};
    uint64_t st = seed * 0x9E3779B97F4A7C15ULL + 0xD1B54A32D192ED03ULL;
    auto next = [&]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    };
    for (int l = 0; l < 5; ++l) {
        const float lim = std::sqrt(6.0f / 64.0f);
        for (int i = offs[2 * l]; i < offs[2 * l + 1]; ++i)
            p[i] = (float)(((double)(next() >> 11) * 0x1p-53) * 2.0 - 1.0) * lim;
    }
    return p;
}

static phg_status guide_alloc(Ctx* c) {
    if (c->g_init) return PHG_OK;
    PHG_HIP(c, c->g_params.ensure(kGuideParamCount));
    PHG_HIP(c, c->g_m.ensure(kGuideParamCount));
    PHG_HIP(c, c->g_v.ensure(kGuideParamCount));
    PHG_HIP(c, c->g_grads.ensure(kGuideParamCount));
    PHG_HIP(c, c->g_bias.ensure(5 * 64));
    PHG_HIP(c, c->g_wt.ensure(5 * 4096));
    PHG_HIP(c, c->g_wn.ensure(5 * 4096));
    PHG_HIP(c, c->g_counters.ensure(4));
    c->g_init = true;
    return PHG_OK;
}

static phg_status guide_load(Ctx* c, const float* host_params) {
    phg_status st = guide_alloc(c);
    if (st != PHG_OK) return st;
    PHG_HIP(c, hipMemcpyAsync(c->g_params.p, host_params, kGuideParamCount * sizeof(float), hipMemcpyHostToDevice, c->stream));
    PHG_HIP(c, guide_fill(c->g_m.p, kGuideParamCount, 0.0f, c->stream));
    PHG_HIP(c, guide_fill(c->g_v.p, kGuideParamCount, 0.0f, c->stream));
    PHG_HIP(c, guide_fill(c->g_grads.p, kGuideParamCount, 0.0f, c->stream));
    PHG_HIP(c, guide_prepare(guide_net(c), c->stream));
    PHG_HIP(c, hipStreamSynchronize(c->stream));
    c->g_steps = 0;
    return PHG_OK;
}

// `steps` Adam steps over device records, minibatches taken round robin
static phg_status guide_train_device(Ctx* c, const float* d_rec, uint32_t nrec, uint32_t batch, float lr, uint32_t steps) {
    if (nrec == 0 || steps == 0) return PHG_OK;
    const uint32_t bs = std::min(batch ? batch : nrec, nrec);
    const uint32_t grid = std::min<uint32_t>(c->cu_count, (bs + 31) / 32);
    uint32_t off = 0;
    for (uint32_t k = 0; k < steps; ++k) {
        const uint32_t n = std::min(bs, nrec - off);
        PHG_HIP(c, guide_gradients(guide_net(c), d_rec + (size_t)off * kGuideRecordFloats, n, c->g_grads.p, grid, c->stream));
        PHG_HIP(c, guide_adam(guide_net(c), c->g_grads.p, c->g_m.p, c->g_v.p, lr, (int)(++c->g_steps), c->stream));
        PHG_HIP(c, guide_prepare(guide_net(c), c->stream));
        off += n;
        if (off >= nrec) off = 0;
    }
    return PHG_OK;
}

}  // namespace
}  // namespace phg

using namespace phg;

extern "C" {

phg_status phg_create(void** out_ctx, const phg_device_opts* opts) {
    if (!out_ctx) return PHG_ERR_INVALID_ARG;
    *out_ctx = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PHG_ERR_NO_DEVICE;
    const int dev = opts ? opts->device : 0;
    if (dev < 0 || dev >= ndev) return PHG_ERR_INVALID_ARG;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return PHG_ERR_NO_DEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return PHG_ERR_NO_DEVICE;
    Ctx* c = new Ctx();
    c->device = dev;
    c->cu_count = (uint32_t)prop.multiProcessorCount;
    if (hipSetDevice(dev) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return PHG_ERR_HIP;
    }
    if (resident_blocks(&c->trace_blocks_per_cu, &c->shade_blocks_per_cu) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return PHG_ERR_HIP;
    }
    for (auto& e : c->ev) (void)hipEventCreate(&e);
    *out_ctx = c;
    return PHG_OK;
}

phg_status phg_destroy(void* ctx) {
    Ctx* c = static_cast<Ctx*>(ctx);
    if (!c) return PHG_ERR_INVALID_ARG;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->nodes.free_(); c->tris.free_(); c->mats.free_(); c->emitters.free_(); c->emit_cdf.free_(); c->emit_tris.free_();
    c->film.free_(); c->rng.free_(); c->T.free_(); c->Wem.free_(); c->prevP.free_(); c->L.free_(); c->hits.free_();
    c->shadow_c.free_(); c->ext0.free_(); c->ext1.free_(); c->shadow.free_(); c->ext_pid0.free_(); c->ext_pid1.free_();
    c->shadow_pid.free_(); c->spill.free_(); c->counts.free_(); c->tcount.free_(); c->errflags.free_(); c->q_rays.free_(); c->q_hits.free_();
    c->q_nvis.free_(); c->q_ntri.free_();
    c->g_params.free_(); c->g_m.free_(); c->g_v.free_(); c->g_grads.free_(); c->g_bias.free_(); c->g_wt.free_();
    c->g_wn.free_(); c->g_lobes.free_(); c->g_rec.free_(); c->g_train.free_(); c->g_q.free_(); c->g_counters.free_();
    for (auto& e : c->ev) if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return PHG_OK;
}

const char* phg_last_error(void* ctx) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return c ? c->err.c_str() : "null context";
}

phg_status phg_upload_scene(void* ctx, const phg_scene_desc* s, const phg_bvh_opts* bo) {
    Ctx* c = static_cast<Ctx*>(ctx);
    if (!c || !s) return PHG_ERR_INVALID_ARG;
    if (s->num_tris == 0 || !s->verts || !s->tri_material || !s->tri_emitter || !s->materials || s->num_materials == 0)
        return set_err(c, PHG_ERR_INVALID_ARG, "scene needs triangles, per-triangle material/emitter ids and materials");
    if (s->num_emitters > 0 && (!s->emitters || !s->emitter_tris))
        return set_err(c, PHG_ERR_INVALID_ARG, "emitters declared without emitter arrays");
    for (uint32_t t = 0; t < s->num_tris; ++t) {
        if (s->tri_material[t] >= s->num_materials) return set_err(c, PHG_ERR_INVALID_ARG, "triangle material id out of range");
        if (s->tri_emitter[t] >= (int32_t)s->num_emitters) return set_err(c, PHG_ERR_INVALID_ARG, "triangle emitter id out of range");
    }
    for (uint32_t m = 0; m < s->num_materials; ++m)
        if (s->materials[m].type > PHG_MAT_DIELECTRIC) return set_err(c, PHG_ERR_INVALID_ARG, "unknown material type");
    for (uint32_t e = 0; e < s->num_emitters; ++e) {
        const phg_emitter& em = s->emitters[e];
        if (em.tri_count == 0 || (uint64_t)em.tri_begin + em.tri_count > s->num_emitter_tris)
            return set_err(c, PHG_ERR_INVALID_ARG, "emitter triangle range out of bounds");
        for (uint32_t k = 0; k < em.tri_count; ++k)
            if (s->emitter_tris[em.tri_begin + k] >= s->num_tris) return set_err(c, PHG_ERR_INVALID_ARG, "emitter triangle id out of range");
    }
    PHG_HIP(c, hipSetDevice(c->device));
    BvhBuildParams bp;
    if (bo) {
        if (bo->max_leaf_tris) bp.max_leaf_tris = bo->max_leaf_tris;
        if (bo->sah_bins) bp.sah_bins = bo->sah_bins;
        bp.nthreads = bo->nthreads;
    }
    BvhBuildResult bvh;
    const auto t0 = std::chrono::steady_clock::now();
    try {
        bvh = build_wide_bvh(s->verts, s->num_tris, bp);
    } catch (const std::exception& ex) {
        return set_err(c, PHG_ERR_INVALID_ARG, ex.what());
    }
    c->bvh_build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (bvh.stack_need > 4096)
        return set_err(c, PHG_ERR_UNSUPPORTED, "BVH needs a traversal stack deeper than 4096 entries");
    // triangles in leaf order
    std::vector<TriRec> tr(s->num_tris);
    std::vector<uint32_t> pos_of(s->num_tris);
    for (uint32_t k = 0; k < s->num_tris; ++k) {
        const uint32_t t = bvh.tri_order[k];
        pos_of[t] = k;
        TriRec& r = tr[k];
        std::memcpy(r.a, s->verts + 9 * (size_t)t, 9 * sizeof(float));
        r.prim = (int32_t)t;
        r.mat = s->tri_material[t];
        r.emitter = s->tri_emitter[t];
    }
    // emitters: power-favouring CDF (ESPowerFavoring::update, TPwcDistribution1D ctor)
    const uint32_t ne = s->num_emitters;
    std::vector<EmitterRec> er(std::max(ne, 1u));
    std::vector<float> cdf(ne + 1, 0.0f);
    std::vector<EmitTri> etris(std::max(s->num_emitter_tris, 1u));
    for (uint32_t i = 0; i < s->num_emitter_tris; ++i) {
        const uint32_t t = s->emitter_tris[i];
        const float* v = s->verts + 9 * (size_t)t;
        etris[i].tri = pos_of[t];
        etris[i].area = htri_area(H3{v[0], v[1], v[2]}, H3{v[3], v[4], v[5]}, H3{v[6], v[7], v[8]});
    }
    if (ne > 0) {
        std::vector<float> w(ne, 0.0f);
        for (uint32_t e = 0; e < ne; ++e) {
            const phg_emitter& em = s->emitters[e];
            float total = 0.0f;
            for (uint32_t k = 0; k < em.tri_count; ++k) {
                const float flux = hluminance(em.radiance) * etris[em.tri_begin + k].area * 3.14159265358979323846f;
                total += flux;
            }
            w[e] = total;
        }
        const float delta = 1.0f / (float)ne;
        for (uint32_t i = 1; i <= ne; ++i) cdf[i] = cdf[i - 1] + w[i - 1] * delta;
        const float rcpSum = 1.0f / cdf[ne];
        cdf[0] = 0.0f; cdf[ne] = 1.0f;
        if (std::isfinite(rcpSum)) { for (uint32_t i = 1; i < ne; ++i) cdf[i] *= rcpSum; }
        else { for (uint32_t i = 1; i < ne; ++i) cdf[i] = (float)i / (float)ne; }
        for (uint32_t e = 0; e < ne; ++e) {
            const phg_emitter& em = s->emitters[e];
            EmitterRec& r = er[e];
            std::memset(&r, 0, sizeof(r));
            r.radiance[0] = em.radiance[0]; r.radiance[1] = em.radiance[1]; r.radiance[2] = em.radiance[2];
            r.pick_pdf = cdf[e + 1] - cdf[e];
            r.tri_begin = em.tri_begin; r.tri_count = em.tri_count;
        }
    }
    PHG_HIP(c, c->nodes.ensure(bvh.nodes.size()));
    PHG_HIP(c, c->tris.ensure(tr.size()));
    PHG_HIP(c, c->mats.ensure(s->num_materials));
    PHG_HIP(c, c->emitters.ensure(er.size()));
    PHG_HIP(c, c->emit_cdf.ensure(cdf.size()));
    PHG_HIP(c, c->emit_tris.ensure(etris.size()));
    PHG_HIP(c, hipMemcpy(c->nodes.p, bvh.nodes.data(), bvh.nodes.size() * sizeof(WideNode), hipMemcpyHostToDevice));
    PHG_HIP(c, hipMemcpy(c->tris.p, tr.data(), tr.size() * sizeof(TriRec), hipMemcpyHostToDevice));
    PHG_HIP(c, hipMemcpy(c->mats.p, s->materials, s->num_materials * sizeof(phg_material), hipMemcpyHostToDevice));
    PHG_HIP(c, hipMemcpy(c->emitters.p, er.data(), er.size() * sizeof(EmitterRec), hipMemcpyHostToDevice));
    PHG_HIP(c, hipMemcpy(c->emit_cdf.p, cdf.data(), cdf.size() * sizeof(float), hipMemcpyHostToDevice));
    PHG_HIP(c, hipMemcpy(c->emit_tris.p, etris.data(), etris.size() * sizeof(EmitTri), hipMemcpyHostToDevice));
    c->sc.nodes = c->nodes.p;
    c->sc.tris = c->tris.p;
    c->sc.mats = c->mats.p;
    c->sc.emitters = c->emitters.p;
    c->sc.emit_cdf = c->emit_cdf.p;
    c->sc.emit_tris = c->emit_tris.p;
    c->sc.num_emitters = ne;
    c->sc.cam = s->camera;
    {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (size_t i = 0; i < 3ull * s->num_tris; ++i) {
            const float* v = s->verts + 3 * i;
            for (int k = 0; k < 3; ++k) {
                if (std::isfinite(v[k])) { lo[k] = std::min(lo[k], v[k]); hi[k] = std::max(hi[k], v[k]); }
            }
        }
        for (int k = 0; k < 3; ++k) {
            const float e = hi[k] - lo[k];
            c->g_scene.lo[k] = std::isfinite(lo[k]) ? lo[k] : 0.0f;
            c->g_scene.inv_ext[k] = (e > 0.0f && std::isfinite(e)) ? 1.0f / e : 1.0f;
        }
    }
    c->stack_need = bvh.stack_need;
    c->bvh_depth = bvh.depth;
    c->bvh_nodes = (uint32_t)bvh.nodes.size();
    c->have_scene = true;
    c->stats.bvh_nodes = c->bvh_nodes;
    c->stats.bvh_depth = c->bvh_depth;
    c->stats.bvh_stack_need = c->stack_need;
    c->stats.bvh_build_ms = c->bvh_build_ms;
    return PHG_OK;
}

static uint32_t trace_grid(const Ctx* c) { return c->cu_count * c->trace_blocks_per_cu; }
static uint32_t spill_cap(const Ctx* c) {
    return c->stack_need > (uint32_t)kTraceRing ? c->stack_need - (uint32_t)kTraceRing : 0u;
}
static phg_status ensure_spill(Ctx* c) {
    const size_t n = (size_t)trace_grid(c) * kTraceBlock * std::max<uint32_t>(spill_cap(c), 1u);
    PHG_HIP(c, c->spill.ensure(n));
    return PHG_OK;
}

static phg_status do_intersect(Ctx* c, const phg_ray* rays, uint64_t n, phg_hit* hits, int query, int ptr_kind,
                               uint32_t* nvis, uint32_t* ntri) {
    if (!c->have_scene) return set_err(c, PHG_ERR_NO_SCENE, "no scene uploaded");
    if (n == 0) return PHG_OK;
    if (!rays || !hits || (query != PHG_QUERY_CLOSEST && query != PHG_QUERY_ANY) ||
        (ptr_kind != PHG_PTR_HOST && ptr_kind != PHG_PTR_DEVICE))
        return set_err(c, PHG_ERR_INVALID_ARG, "bad intersect arguments");
    PHG_HIP(c, hipSetDevice(c->device));
    PHG_HIP(c, c->errflags.ensure(1));
    PHG_HIP(c, hipMemsetAsync(c->errflags.p, 0, sizeof(uint32_t), c->stream));
    const phg_ray* d_rays = rays;
    phg_hit* d_hits = hits;
    uint32_t* d_nvis = nullptr;
    uint32_t* d_ntri = nullptr;
    if (ptr_kind == PHG_PTR_HOST) {
        PHG_HIP(c, c->q_rays.ensure(n));
        PHG_HIP(c, c->q_hits.ensure(n));
        PHG_HIP(c, hipMemcpyAsync(c->q_rays.p, rays, n * sizeof(phg_ray), hipMemcpyHostToDevice, c->stream));
        d_rays = c->q_rays.p;
        d_hits = c->q_hits.p;
        if (nvis) {
            PHG_HIP(c, c->q_nvis.ensure(n));
            PHG_HIP(c, c->q_ntri.ensure(n));
            d_nvis = c->q_nvis.p; d_ntri = c->q_ntri.p;
        }
    } else {
        d_nvis = nvis; d_ntri = ntri;
    }
    PHG_HIP(c, hipEventRecord(c->ev[0], c->stream));
    {
        const phg_status st = ensure_spill(c);
        if (st != PHG_OK) return st;
    }
    PHG_HIP(c, launch_intersect(c->sc, d_rays, n, d_hits, query == PHG_QUERY_ANY, d_nvis, d_ntri, c->errflags.p,
                                c->spill.p, spill_cap(c), trace_grid(c), c->stream));
    PHG_HIP(c, hipEventRecord(c->ev[1], c->stream));
    if (ptr_kind == PHG_PTR_HOST) {
        PHG_HIP(c, hipMemcpyAsync(hits, d_hits, n * sizeof(phg_hit), hipMemcpyDeviceToHost, c->stream));
        if (nvis) {
            PHG_HIP(c, hipMemcpyAsync(nvis, d_nvis, n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
            PHG_HIP(c, hipMemcpyAsync(ntri, d_ntri, n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        }
    }
    uint32_t flags = 0;
    PHG_HIP(c, hipMemcpyAsync(&flags, c->errflags.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    PHG_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    PHG_HIP(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    if (query == PHG_QUERY_ANY) c->stats.shadow_ms = ms; else c->stats.trace_ms = ms;
    if (flags) return set_err(c, PHG_ERR_UNSUPPORTED, "traversal stack overflow");
    return PHG_OK;
}

phg_status phg_intersect(void* ctx, const phg_ray* rays, uint64_t n, phg_hit* hits, int query, int ptr_kind) {
    Ctx* c = static_cast<Ctx*>(ctx);
    if (!c) return PHG_ERR_INVALID_ARG;
    return do_intersect(c, rays, n, hits, query, ptr_kind, nullptr, nullptr);
}

phg_status phg_intersect_counted(void* ctx, const phg_ray* rays, uint64_t n, phg_hit* hits, uint32_t* node_visits,
                                 uint32_t* tri_tests, int ptr_kind) {
    Ctx* c = static_cast<Ctx*>(ctx);
    if (!c) return PHG_ERR_INVALID_ARG;
    if (!node_visits || !tri_tests) return set_err(c, PHG_ERR_INVALID_ARG, "counter outputs required");
    return do_intersect(c, rays, n, hits, PHG_QUERY_CLOSEST, ptr_kind, node_visits, tri_tests);
}

phg_status phg_film_clear(void* ctx, uint32_t width, uint32_t height) {
    Ctx* c = static_cast<Ctx*>(ctx);
    if (!c || width == 0 || height == 0) return PHG_ERR_INVALID_ARG;
    PHG_HIP(c, hipSetDevice(c->device));
    const uint64_t n = 8ull * width * height;
    PHG_HIP(c, c->film.ensure(n));
    PHG_HIP(c, launch_fill_zero(c->film.p, n, c->stream));
    PHG_HIP(c, hipStreamSynchronize(c->stream));
    c->film_w = width; c->film_h = height;
    return PHG_OK;
}

phg_status phg_film_read(void* ctx, double* out, uint64_t out_len) {
    Ctx* c = static_cast<Ctx*>(ctx);
    if (!c || !out) return PHG_ERR_INVALID_ARG;
    const uint64_t n = 8ull * c->film_w * c->film_h;
    if (n == 0) return set_err(c, PHG_ERR_INVALID_ARG, "film not initialised");
    if (out_len < n) return set_err(c, PHG_ERR_INVALID_ARG, "output buffer too small");
    PHG_HIP(c, hipSetDevice(c->device));
    PHG_HIP(c, hipMemcpyAsync(out, c->film.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PHG_HIP(c, hipStreamSynchronize(c->stream));
    return PHG_OK;
}

phg_status phg_render(void* ctx, const phg_render_opts* o) {
    Ctx* c = static_cast<Ctx*>(ctx);
    if (!c || !o) return PHG_ERR_INVALID_ARG;
    if (!c->have_scene) return set_err(c, PHG_ERR_NO_SCENE, "no scene uploaded");
    if (o->width != c->film_w || o->height != c->film_h)
        return set_err(c, PHG_ERR_INVALID_ARG, "film size differs from render size; call phg_film_clear first");
    const uint64_t npix = (uint64_t)o->width * o->height;
    if (npix == 0 || npix > (1ull << 26)) return set_err(c, PHG_ERR_INVALID_ARG, "image size out of range");
    if (o->seed >= (1ull << 24)) return set_err(c, PHG_ERR_INVALID_ARG, "seed must be below 2^24");
    if ((uint64_t)o->spp_begin + o->spp > (1ull << 40) / npix) return set_err(c, PHG_ERR_INVALID_ARG, "sample index range too large");
    if (o->spp == 0) return PHG_OK;
    PHG_HIP(c, hipSetDevice(c->device));
    // samples per pass: fill ~4M paths in flight
    uint32_t k = o->spp_per_pass;
    if (k == 0) k = (uint32_t)std::max<uint64_t>(1, (4ull << 20) / npix);
    k = std::min(k, o->spp);
    uint64_t cap = npix * k;
    if (cap > 0xffffffffull) return set_err(c, PHG_ERR_INVALID_ARG, "too many paths per pass");
    PHG_HIP(c, c->rng.ensure(cap));
    PHG_HIP(c, c->T.ensure(cap)); PHG_HIP(c, c->Wem.ensure(cap)); PHG_HIP(c, c->prevP.ensure(cap)); PHG_HIP(c, c->L.ensure(cap));
    PHG_HIP(c, c->hits.ensure(cap)); PHG_HIP(c, c->shadow_c.ensure(cap));
    PHG_HIP(c, c->ext0.ensure(cap)); PHG_HIP(c, c->ext1.ensure(cap)); PHG_HIP(c, c->shadow.ensure(cap));
    PHG_HIP(c, c->ext_pid0.ensure(cap)); PHG_HIP(c, c->ext_pid1.ensure(cap)); PHG_HIP(c, c->shadow_pid.ensure(cap));
    PHG_HIP(c, c->counts.ensure(2 * kCountSlots));
    PHG_HIP(c, c->tcount.ensure(8));
    PHG_HIP(c, c->errflags.ensure(1));
    PHG_HIP(c, hipMemsetAsync(c->tcount.p, 0, 8 * sizeof(unsigned long long), c->stream));
    PHG_HIP(c, hipMemsetAsync(c->errflags.p, 0, sizeof(uint32_t), c->stream));
    PHG_HIP(c, launch_clear_counts(c->counts.p, 2 * kCountSlots, c->stream));
    PathBuffers pb{};
    pb.rng = c->rng.p; pb.T = c->T.p; pb.Wem = c->Wem.p; pb.prevP = c->prevP.p; pb.L = c->L.p;
    pb.ext[0] = c->ext0.p; pb.ext[1] = c->ext1.p; pb.ext_pid[0] = c->ext_pid0.p; pb.ext_pid[1] = c->ext_pid1.p;
    pb.hits = c->hits.p; pb.shadow = c->shadow.p; pb.shadow_pid = c->shadow_pid.p; pb.shadow_c = c->shadow_c.p;
    pb.counts = c->counts.p; pb.traversal_counters = c->tcount.p; pb.error_flags = c->errflags.p; pb.film = c->film.p;
    {
        const phg_status st = ensure_spill(c);
        if (st != PHG_OK) return st;
    }
    pb.spill = c->spill.p;
    pb.spill_cap = spill_cap(c);
    const uint32_t count_trav = (o->flags & PHG_RENDER_COUNT_TRAVERSAL) ? 1u : 0u;
    const bool guided = (o->flags & PHG_RENDER_GUIDED) != 0;
    const bool train = (o->flags & PHG_RENDER_TRAIN) != 0;
    uint32_t rec_bounces = 0;
    uint32_t train_cap = 0;
    const uint32_t batch = o->guide_batch ? o->guide_batch : (1u << 18);
    if (guided || train) {
        if (!c->g_init) {
            const std::vector<float> p0 = guide_default_params(o->seed + 1);
            phg_status st = guide_load(c, p0.data());
            if (st != PHG_OK) return st;
        }
        PHG_HIP(c, c->g_lobes.ensure(cap * (size_t)kGuideOutputs));
    }
    if (train) {
        rec_bounces = std::min<uint32_t>(kGuideRecBounces, o->max_bounces + 1);
        PHG_HIP(c, c->g_rec.ensure((size_t)rec_bounces * cap * kRecFloat4));
        // fresh records for every step of a pass (no record is reused within the pass)
        train_cap = std::max<uint32_t>(4u, o->guide_steps ? o->guide_steps : 4u) * batch;
        PHG_HIP(c, c->g_train.ensure((size_t)train_cap * kGuideRecordFloats));
    }
    pb.lobes = c->g_lobes.p;
    pb.rec = c->g_rec.p;
    double t_ginfer = 0, t_gtrain = 0;
    uint64_t g_queries = 0, g_records = 0;
    const uint32_t grid_trace = trace_grid(c);     // persistent-style grid-stride launches
    const uint32_t grid_shade = c->cu_count * c->shade_blocks_per_cu;
    double t_trace = 0, t_shadow = 0, t_shade = 0;
    uint64_t rays_primary = 0;
    uint32_t n_trace = 0, n_shadow = 0;
    const auto wall0 = std::chrono::steady_clock::now();
    for (uint32_t s0 = 0; s0 < o->spp; s0 += k) {
        const uint32_t kk = std::min(k, o->spp - s0);
        PassParams pp{};
        pp.width = o->width; pp.height = o->height; pp.npix = (uint32_t)npix;
        pp.paths = (uint32_t)(npix * kk); pp.sample_begin = o->spp_begin + s0; pp.max_bounces = o->max_bounces;
        pp.seed = o->seed;
        pp.guide_prob = (guided && c->g_ready) ? o->guide_prob : 0.0f;
        pp.record = train ? 1u : 0u;
        pp.rec_bounces = rec_bounces;
        rays_primary += pp.paths;
        PHG_HIP(c, hipEventRecord(c->ev[0], c->stream));
        PHG_HIP(c, launch_generate(c->sc, pp, pb, c->stream));
        PHG_HIP(c, hipEventRecord(c->ev[1], c->stream));
        PHG_HIP(c, hipStreamSynchronize(c->stream));
        float ms = 0; PHG_HIP(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1])); t_shade += ms;
        for (uint32_t b = 0;; ++b) {
            const bool use_guide = guided && pp.guide_prob > 0.0f && b < o->max_bounces;
            PHG_HIP(c, hipEventRecord(c->ev[0], c->stream));
            PHG_HIP(c, launch_extend(c->sc, pb, b, grid_trace, count_trav, c->stream));
            PHG_HIP(c, hipEventRecord(c->ev[1], c->stream));
            if (use_guide) {
                PHG_HIP(c, guide_infer_hits(guide_net(c), pb.hits, pb.ext[b & 1], c->sc.tris, c->g_scene,
                                            c->counts.p + b % kCountSlots, c->g_lobes.p, c->cu_count * 2, c->stream));
            }
            PHG_HIP(c, hipEventRecord(c->ev[4], c->stream));
            PHG_HIP(c, launch_shade(c->sc, pp, pb, b, grid_shade, guided || train, c->stream));
            PHG_HIP(c, hipEventRecord(c->ev[2], c->stream));
            PHG_HIP(c, launch_shadow(c->sc, pb, b, grid_trace, count_trav, rec_bounces, pp.paths, c->stream));
            PHG_HIP(c, hipEventRecord(c->ev[3], c->stream));
            uint32_t next[2] = {0, 0};
            PHG_HIP(c, hipMemcpyAsync(&next[0], c->counts.p + (b + 1) % kCountSlots, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
            PHG_HIP(c, hipMemcpyAsync(&next[1], c->counts.p + b % kCountSlots, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
            PHG_HIP(c, hipStreamSynchronize(c->stream));
            float a = 0, bb = 0, cc = 0, gi = 0;
            PHG_HIP(c, hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
            PHG_HIP(c, hipEventElapsedTime(&gi, c->ev[1], c->ev[4]));
            PHG_HIP(c, hipEventElapsedTime(&bb, c->ev[4], c->ev[2]));
            PHG_HIP(c, hipEventElapsedTime(&cc, c->ev[2], c->ev[3]));
            t_trace += a; t_shade += bb; t_shadow += cc; t_ginfer += gi;
            if (use_guide) g_queries += next[1];
            const uint32_t next0 = next[0];
            ++n_trace; ++n_shadow;
            if (next0 == 0 || b >= o->max_bounces) break;
        }
        PHG_HIP(c, hipEventRecord(c->ev[0], c->stream));
        PHG_HIP(c, launch_film(pp, pb, kk, c->stream));
        PHG_HIP(c, hipEventRecord(c->ev[1], c->stream));
        PHG_HIP(c, hipStreamSynchronize(c->stream));
        PHG_HIP(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1])); t_shade += ms;
        if (train) {
            // online training on this pass's radiance samples
            PHG_HIP(c, hipEventRecord(c->ev[0], c->stream));
            PHG_HIP(c, hipMemsetAsync(c->g_counters.p, 0, 4 * sizeof(uint32_t), c->stream));
            PHG_HIP(c, guide_collect(pb, pp, c->g_scene, c->g_train.p, train_cap, c->g_counters.p, c->g_keep,
                                     0.0f, ++c->g_pass_salt, c->stream));
            uint32_t cnt[2] = {0, 0};
            PHG_HIP(c, hipMemcpyAsync(cnt, c->g_counters.p, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
            PHG_HIP(c, hipStreamSynchronize(c->stream));
            const uint32_t nrec = std::min(cnt[0], train_cap);
            const uint32_t steps = o->guide_steps ? o->guide_steps : 4;
            phg_status gs = guide_train_device(c, c->g_train.p, nrec, batch, o->guide_lr > 0 ? o->guide_lr : 1e-3f, steps);
            if (gs != PHG_OK) return gs;
            PHG_HIP(c, hipEventRecord(c->ev[1], c->stream));
            PHG_HIP(c, hipStreamSynchronize(c->stream));
            PHG_HIP(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1])); t_gtrain += ms;
            g_records += nrec;
            // keep about as many records as the next pass's steps consume
            const double want = (double)steps * batch;
            c->g_keep = cnt[1] > 0 ? (float)std::min(1.0, 1.25 * want / (double)cnt[1]) : 1.0f;
            if (nrec > 0) c->g_ready = true;
            pp.guide_prob = 0.0f;  // (next pass re-reads g_ready)
        }
    }
    uint32_t flags = 0;
    unsigned long long tc[8];
    PHG_HIP(c, hipMemcpyAsync(&flags, c->errflags.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    PHG_HIP(c, hipMemcpyAsync(tc, c->tcount.p, sizeof(tc), hipMemcpyDeviceToHost, c->stream));
    PHG_HIP(c, hipStreamSynchronize(c->stream));
    c->stats.render_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    c->stats.paths = npix * (uint64_t)o->spp;
    c->stats.rays_primary = rays_primary;
    c->stats.rays_extension = tc[4];
    c->stats.rays_shadow = tc[5];
    c->stats.node_visits = tc[0] + tc[2];
    c->stats.tri_tests = tc[1] + tc[3];
    c->stats.trace_node_visits = tc[0];
    c->stats.trace_tri_tests = tc[1];
    c->stats.shadow_node_visits = tc[2];
    c->stats.shadow_tri_tests = tc[3];
    c->stats.trace_launches = n_trace;
    c->stats.shadow_launches = n_shadow;
    c->stats.guide_infer_ms = t_ginfer;
    c->stats.guide_train_ms = t_gtrain;
    c->stats.guide_queries = g_queries;
    c->stats.guide_records = g_records;
    c->stats.guide_steps = c->g_steps;
    c->stats.trace_ms = t_trace;
    c->stats.shadow_ms = t_shadow;
    c->stats.shade_ms = t_shade;
    if (flags) return set_err(c, PHG_ERR_UNSUPPORTED, "traversal stack overflow");
    return PHG_OK;
}

phg_status phg_guide_set_params(void* ctx, const float* params, uint32_t n) {
    Ctx* c = static_cast<Ctx*>(ctx);
    if (!c || !params) return PHG_ERR_INVALID_ARG;
    if (n != kGuideParamCount) return set_err(c, PHG_ERR_INVALID_ARG, "expected PHG_GUIDE_PARAMS values");
    PHG_HIP(c, hipSetDevice(c->device));
    phg_status st = guide_load(c, params);
    if (st == PHG_OK) c->g_ready = true;
    return st;
}

phg_status phg_guide_get_params(void* ctx, float* params, uint32_t n) {
    Ctx* c = static_cast<Ctx*>(ctx);
    if (!c || !params) return PHG_ERR_INVALID_ARG;
    if (n != kGuideParamCount) return set_err(c, PHG_ERR_INVALID_ARG, "expected PHG_GUIDE_PARAMS values");
    if (!c->g_init) return set_err(c, PHG_ERR_INVALID_ARG, "guiding network not initialised");
    PHG_HIP(c, hipSetDevice(c->device));
    PHG_HIP(c, hipMemcpyAsync(params, c->g_params.p, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    PHG_HIP(c, hipStreamSynchronize(c->stream));
    return PHG_OK;
}

phg_status phg_guide_forward(void* ctx, const float* features, uint64_t n, float* out) {
    Ctx* c = static_cast<Ctx*>(ctx);
    if (!c || !features || !out || n >= (1ull << 31)) return PHG_ERR_INVALID_ARG;
    if (!c->g_init) return set_err(c, PHG_ERR_INVALID_ARG, "guiding network not initialised");
    if (n == 0) return PHG_OK;
    PHG_HIP(c, hipSetDevice(c->device));
    PHG_HIP(c, c->g_q.ensure(n * (9 + kGuideOutputs)));
    float* d_in = c->g_q.p;
    float* d_out = c->g_q.p + n * 9;
    PHG_HIP(c, hipMemcpyAsync(d_in, features, n * 9 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    PHG_HIP(c, guide_infer_features(guide_net(c), d_in, (uint32_t)n, d_out, 1, c->cu_count * 2, c->stream));
    PHG_HIP(c, hipMemcpyAsync(out, d_out, n * kGuideOutputs * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    PHG_HIP(c, hipStreamSynchronize(c->stream));
    return PHG_OK;
}

phg_status phg_guide_gradients(void* ctx, const float* records, uint64_t n, float* grads) {
    Ctx* c = static_cast<Ctx*>(ctx);
    if (!c || !records || !grads || n == 0 || n >= (1ull << 31)) return PHG_ERR_INVALID_ARG;
    if (!c->g_init) return set_err(c, PHG_ERR_INVALID_ARG, "guiding network not initialised");
    PHG_HIP(c, hipSetDevice(c->device));
    PHG_HIP(c, c->g_q.ensure(n * kGuideRecordFloats));
    PHG_HIP(c, hipMemcpyAsync(c->g_q.p, records, n * kGuideRecordFloats * sizeof(float), hipMemcpyHostToDevice, c->stream));
    PHG_HIP(c, guide_fill(c->g_grads.p, kGuideParamCount, 0.0f, c->stream));
    const uint32_t grid = std::min<uint32_t>(c->cu_count, (uint32_t)((n + 31) / 32));
    PHG_HIP(c, guide_gradients(guide_net(c), c->g_q.p, (uint32_t)n, c->g_grads.p, grid, c->stream));
    PHG_HIP(c, hipMemcpyAsync(grads, c->g_grads.p, kGuideParamCount * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    PHG_HIP(c, guide_fill(c->g_grads.p, kGuideParamCount, 0.0f, c->stream));
    PHG_HIP(c, hipStreamSynchronize(c->stream));
    return PHG_OK;
}

phg_status phg_guide_train(void* ctx, const float* records, uint64_t n, float lr, uint32_t steps) {
    Ctx* c = static_cast<Ctx*>(ctx);
    if (!c || !records || n == 0 || n >= (1ull << 31)) return PHG_ERR_INVALID_ARG;
    if (!c->g_init) return set_err(c, PHG_ERR_INVALID_ARG, "guiding network not initialised");
    PHG_HIP(c, hipSetDevice(c->device));
    PHG_HIP(c, c->g_q.ensure(n * kGuideRecordFloats));
    PHG_HIP(c, hipMemcpyAsync(c->g_q.p, records, n * kGuideRecordFloats * sizeof(float), hipMemcpyHostToDevice, c->stream));
    phg_status st = guide_train_device(c, c->g_q.p, (uint32_t)n, (uint32_t)n, lr, steps);
    if (st != PHG_OK) return st;
    PHG_HIP(c, hipStreamSynchronize(c->stream));
    c->g_ready = true;
    return PHG_OK;
}

phg_status phg_get_stats(void* ctx, phg_stats* out) {
    Ctx* c = static_cast<Ctx*>(ctx);
    if (!c || !out) return PHG_ERR_INVALID_ARG;
    *out = c->stats;
    return PHG_OK;
}

}  // extern "C"
