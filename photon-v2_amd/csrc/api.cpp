// C ABI of the BNEEPT hot path (include/phg.h): context, scene upload, batched ray
// queries, wavefront render loop, FP64 film, statistics.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <thread>
#include <vector>

#include "../../include/phg.h"
#include "bvh_build.h"
#include "color_energy.h"
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

// Where an uploaded scene lives in the shared device arrays.
struct SceneInfo {
    uint32_t root = 0;            // first node of its BVH
    uint32_t emitter_begin = 0, num_emitters = 0, cdf_begin = 0;
    uint32_t num_tris = 0, num_nodes = 0;
    phg_camera cam{};
    GuideScene box{};             // bounds, for the guiding network's position features
};

// Paths per wavefront pass unless tuned: large passes keep every bounce launch full (the
// path state is 216 B per path, 14 GB at 64M paths).  Renders that train the guiding network
// use passes of at most kTrainPassPaths unless told otherwise: the network is updated after
// every pass, and each pass keeps 768 B of training records per path (12 GB at 16M).
constexpr uint32_t kDefaultPassPaths = 32u << 20;
// Bounces per captured chunk of an unbounded or guided render: even (the ping-pong queue of
// bounce m0 + k is that of k) and a divisor of the counter ring (the slot of m0 + k follows
// from the device bounce base).
constexpr uint32_t kChunkBounces = 8;
static_assert(kChunkBounces % 2 == 0 && kCountSlots % kChunkBounces == 0, "bounce chunk must be even and divide the counter ring");
constexpr uint32_t kTrainPassPaths = 16u << 20;

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
    std::vector<SceneInfo> scenes;
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
    DevBuf<uint32_t> fetch;       // region counters of the traversal kernels' work distribution
    DevBuf<uint32_t> path_seg;   // per path: its pass segment
    DevBuf<JobDev> jobs;         // job table of the current render
    DevBuf<PassParams> d_pass;   // pass parameters and segments, rewritten before every pass
    DevBuf<PassSeg> d_segs;
    PassParams* h_pass = nullptr;  // pinned staging of the above
    PassSeg* h_segs = nullptr;
    DevBuf<double> bfilm;        // films of the last batch, one after another
    std::vector<JobDev> bfilm_jobs;
    DevBuf<int32_t> spill;       // traversal-stack spill of the persistent traversal grid
    DevBuf<unsigned long long> tcount;
    DevBuf<uint32_t> errflags;
    // intersect staging
    DevBuf<phg_ray> q_rays;
    DevBuf<phg_hit> q_hits;
    DevBuf<uint32_t> q_nvis, q_ntri;
    // guiding network
    DevBuf<float> g_params, g_m, g_v, g_grads, g_bias;
    DevBuf<uint16_t> g_wt, g_wn, g_wp;
    DevBuf<float> g_bp;
    DevBuf<float> g_lobes;       // per queued extension ray
    DevBuf<float4> g_rec;        // per path, per recorded bounce
    DevBuf<float> g_train;       // collected training records
    DevBuf<float> g_q;           // staging for the guide entry points
    DevBuf<uint32_t> g_counters;
    DevBuf<uint32_t> g_blockcnt;     // per-block record counts of the collection scan
    DevBuf<uint16_t> g_ws_x, g_ws_dz0, g_ws_dz1;   // training-step scratch (GuideWorkspace)
    DevBuf<float> g_ws_part;
    bool g_init = false, g_ready = false;
    uint32_t g_steps = 0, g_pass_salt = 0;
    float g_keep = 1.0f;
    GuideScene g_scene{};
    // stats
    phg_stats stats{};
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    uint32_t cu_count = 256;
    uint32_t trace_block = 128;         // threads per block of the path traversal kernels
    uint32_t trace_blocks_per_cu = 8;   // resident blocks per compute unit, from the occupancy query
    uint32_t shade_blocks_per_cu = 4;
    uint32_t paths_per_pass = kDefaultPassPaths; // paths in flight per wavefront pass
    void* noalloc_scratch = nullptr;    // this device's copy of the allocation-free query's scratch
};


#define PHG_HIP(ctx, expr)                                                                        \
    do {                                                                                          \
        const hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                                   \
            (void)hipGetLastError(); /* do not leave the error to the next launch check */        \
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
static inline float hcolor_energy(const float* c) { return (float)color_energy(c[0], c[1], c[2]); }

static phg_status set_err(Ctx* c, phg_status s, const std::string& msg) { c->err = msg; return s; }

// Runs the body of an exported call: nothing throws across the C ABI.  Host allocations that
// fail (std::bad_alloc) become PHG_ERR_OUT_OF_MEMORY, any other exception PHG_ERR_INVALID_ARG;
// the message goes to phg_last_error.
template <typename F>
static phg_status guarded(void* ctx, F&& body) {
    Ctx* c = static_cast<Ctx*>(ctx);
    try {
        return body();
    } catch (const std::bad_alloc&) {
        if (c) c->err = "host memory allocation failed";
        return PHG_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        if (c) c->err = e.what();
        return PHG_ERR_INVALID_ARG;
    }
}

static GuideNetDev guide_net(Ctx* c) {
    GuideNetDev n{};
    n.params = c->g_params.p; n.wt = c->g_wt.p; n.wn = c->g_wn.p; n.bias = c->g_bias.p;
    n.params_mut = c->g_params.p; n.wt_mut = c->g_wt.p; n.wn_mut = c->g_wn.p; n.bias_mut = c->g_bias.p;
    n.wp = c->g_wp.p; n.bp = c->g_bp.p; n.wp_mut = c->g_wp.p; n.bp_mut = c->g_bp.p;
    return n;
}

// He-uniform initialisation (fan-in), zero biases, from a splitmix64 stream.
static std::vector<float> guide_default_params(uint64_t seed) {
    std::vector<float> p(kGuideParamCount, 0.0f);
    const int offs[11] = {0, 4096, 4160, 8256, 8320, 12416, 12480, 16576, 16640, 19264, 19305};
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
    // the selection logit starts at 0 for every query (selection probability halfway)
    for (int i = 0; i < 64; ++i) p[offs[8] + i * 41 + 40] = 0.0f;
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
    PHG_HIP(c, c->g_wp.ensure(5 * 4096));
    PHG_HIP(c, c->g_bp.ensure(5 * 64));
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

// Scratch of training steps of up to nrec records.
static phg_status guide_workspace(Ctx* c, uint32_t nrec, GuideWorkspace* ws) {
    const uint32_t cap = (nrec + 31u) & ~31u;
    const uint32_t parts = c->cu_count * 4, slices = 32;
    PHG_HIP(c, c->g_ws_x.ensure((size_t)kGuideNetLayers * cap * 64));
    PHG_HIP(c, c->g_ws_dz0.ensure((size_t)cap * 64));
    PHG_HIP(c, c->g_ws_dz1.ensure((size_t)cap * 64));
    const size_t part_floats = (size_t)kGuideNetLayers * (parts + 4) * (64 * 64 + 64);
    PHG_HIP(c, c->g_ws_part.ensure(part_floats + (size_t)kGuideNetLayers * slices * (64 * 64 + 64)));
    ws->X = c->g_ws_x.p; ws->dz0 = c->g_ws_dz0.p; ws->dz1 = c->g_ws_dz1.p; ws->partial = c->g_ws_part.p;
    ws->partial2 = c->g_ws_part.p + part_floats;
    ws->slices = slices;
    ws->cap_records = cap;
    ws->max_parts = parts;
    ws->max_blocks = c->cu_count * 4;
    return PHG_OK;
}

// `steps` Adam steps over device records, minibatches taken round robin
static phg_status guide_train_device(Ctx* c, const float* d_rec, uint32_t nrec, uint32_t batch, float lr, uint32_t steps,
                                     uint64_t* processed = nullptr) {
    if (nrec == 0 || steps == 0) return PHG_OK;
    const uint32_t bs = std::min(batch ? batch : nrec, nrec);
    GuideWorkspace ws{};
    {
        const phg_status st = guide_workspace(c, bs, &ws);
        if (st != PHG_OK) return st;
    }
    uint32_t off = 0;
    for (uint32_t k = 0; k < steps; ++k) {
        const uint32_t n = std::min(bs, nrec - off);
        PHG_HIP(c, guide_gradients(guide_net(c), d_rec + (size_t)off * kGuideRecordFloats, n, c->g_grads.p, ws, c->stream));
        PHG_HIP(c, guide_adam(guide_net(c), c->g_grads.p, c->g_m.p, c->g_v.p, lr, (int)(++c->g_steps), c->stream));
        PHG_HIP(c, guide_prepare(guide_net(c), c->stream));
        off += n;
        if (processed) *processed += n;
        if (off >= nrec) off = 0;
    }
    return PHG_OK;
}

}  // namespace
}  // namespace phg

using namespace phg;

extern "C" {

phg_status phg_create(void** out_ctx, const phg_device_opts* opts) {
    return guarded(nullptr, [&]() -> phg_status {
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
        if (resident_blocks(c->trace_block, &c->trace_blocks_per_cu, &c->shade_blocks_per_cu) != hipSuccess ||
            noalloc_scratch_address(dev, &c->noalloc_scratch) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipStreamDestroy(c->stream);
            delete c;
            return PHG_ERR_HIP;
        }
        for (auto& e : c->ev) (void)hipEventCreate(&e);
        *out_ctx = c;
        return PHG_OK;
    });
}

phg_status phg_destroy(void* ctx) {
    return guarded(ctx, [&]() -> phg_status {
        Ctx* c = static_cast<Ctx*>(ctx);
        if (!c) return PHG_ERR_INVALID_ARG;
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        c->nodes.free_(); c->tris.free_(); c->mats.free_(); c->emitters.free_(); c->emit_cdf.free_(); c->emit_tris.free_();
        c->film.free_(); c->rng.free_(); c->T.free_(); c->Wem.free_(); c->prevP.free_(); c->L.free_(); c->hits.free_();
        c->shadow_c.free_(); c->ext0.free_(); c->ext1.free_(); c->shadow.free_(); c->ext_pid0.free_(); c->ext_pid1.free_();
        c->shadow_pid.free_(); c->path_seg.free_(); c->jobs.free_(); c->d_pass.free_(); c->d_segs.free_(); c->bfilm.free_();
        if (c->h_pass) (void)hipHostFree(c->h_pass);
        c->spill.free_(); c->counts.free_(); c->fetch.free_(); c->tcount.free_(); c->errflags.free_(); c->q_rays.free_(); c->q_hits.free_();
        c->q_nvis.free_(); c->q_ntri.free_();
        c->g_params.free_(); c->g_m.free_(); c->g_v.free_(); c->g_grads.free_(); c->g_bias.free_(); c->g_wt.free_();
        c->g_ws_x.free_(); c->g_ws_dz0.free_(); c->g_ws_dz1.free_(); c->g_ws_part.free_();
        c->g_wn.free_(); c->g_wp.free_(); c->g_bp.free_(); c->g_lobes.free_(); c->g_rec.free_(); c->g_train.free_(); c->g_q.free_(); c->g_counters.free_(); c->g_blockcnt.free_();
        for (auto& e : c->ev) if (e) (void)hipEventDestroy(e);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return PHG_OK;
    });
}

const char* phg_last_error(void* ctx) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return c ? c->err.c_str() : "null context";
}

}  // extern "C"

namespace phg {
namespace {

// One scene prepared on the host: BVH, triangles in leaf order and emitter tables, all with
// scene-local indices (rebased when the scenes are packed together).
struct HostScene {
    std::vector<WideNode> nodes;
    std::vector<TriRec> tris;
    std::vector<phg_material> mats;
    std::vector<EmitterRec> emitters;
    std::vector<float> cdf;         // num_emitters + 1
    std::vector<EmitTri> etris;
    phg_camera cam{};
    GuideScene box{};
    uint32_t stack_need = 0, depth = 0;
    double build_ms = 0.0;
    std::string err;
    phg_status status = PHG_OK;
};

static phg_status scene_fail(HostScene& h, phg_status st, const char* msg) { h.status = st; h.err = msg; return st; }

static phg_status prepare_scene(const phg_scene_desc* s, const BvhBuildParams& bp, HostScene& h) {
    if (s->num_tris == 0 || !s->verts || !s->tri_material || !s->tri_emitter || !s->materials || s->num_materials == 0)
        return scene_fail(h, PHG_ERR_INVALID_ARG, "scene needs triangles, per-triangle material/emitter ids and materials");
    if (s->num_emitters > 0 && (!s->emitters || !s->emitter_tris))
        return scene_fail(h, PHG_ERR_INVALID_ARG, "emitters declared without emitter arrays");
    for (uint32_t t = 0; t < s->num_tris; ++t) {
        if (s->tri_material[t] >= s->num_materials) return scene_fail(h, PHG_ERR_INVALID_ARG, "triangle material id out of range");
        if (s->tri_emitter[t] >= (int32_t)s->num_emitters) return scene_fail(h, PHG_ERR_INVALID_ARG, "triangle emitter id out of range");
    }
    for (uint32_t m = 0; m < s->num_materials; ++m) {
        const phg_material& mt = s->materials[m];
        if (mt.type > PHG_MAT_DIELECTRIC) return scene_fail(h, PHG_ERR_INVALID_ARG, "unknown material type");
        if (mt.type == PHG_MAT_CONDUCTOR && mt.fresnel > PHG_FRESNEL_EXACT)
            return scene_fail(h, PHG_ERR_INVALID_ARG, "unknown conductor Fresnel");
        if (mt.type == PHG_MAT_CONDUCTOR && mt.fresnel == PHG_FRESNEL_EXACT && !(mt.ior_outer > 0.0f))
            return scene_fail(h, PHG_ERR_INVALID_ARG, "exact conductor Fresnel needs a positive outer index");
    }
    for (uint32_t e = 0; e < s->num_emitters; ++e) {
        const phg_emitter& em = s->emitters[e];
        if (em.tri_count == 0 || (uint64_t)em.tri_begin + em.tri_count > s->num_emitter_tris)
            return scene_fail(h, PHG_ERR_INVALID_ARG, "emitter triangle range out of bounds");
        for (uint32_t k = 0; k < em.tri_count; ++k)
            if (s->emitter_tris[em.tri_begin + k] >= s->num_tris)
                return scene_fail(h, PHG_ERR_INVALID_ARG, "emitter triangle id out of range");
    }
    BvhBuildResult bvh;
    const auto t0 = std::chrono::steady_clock::now();
    try {
        bvh = build_wide_bvh(s->verts, s->num_tris, bp);
    } catch (const std::exception& ex) {
        h.err = ex.what();
        h.status = PHG_ERR_INVALID_ARG;
        return h.status;
    }
    h.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (bvh.stack_need > 4096)
        return scene_fail(h, PHG_ERR_UNSUPPORTED, "BVH needs a traversal stack deeper than 4096 entries");
    // triangles in leaf order
    h.tris.resize(s->num_tris);
    std::vector<uint32_t> pos_of(s->num_tris);
    for (uint32_t k = 0; k < s->num_tris; ++k) {
        const uint32_t t = bvh.tri_order[k];
        pos_of[t] = k;
        TriRec& r = h.tris[k];
        std::memcpy(r.a, s->verts + 9 * (size_t)t, 9 * sizeof(float));
        r.prim = (int32_t)t;
        r.mat = s->tri_material[t];
        r.emitter = s->tri_emitter[t];
    }
    // emitters: power-favouring CDF (ESPowerFavoring::update, TPwcDistribution1D ctor); a
    // triangle's weight is its approximate radiant flux, colour energy x area x pi
    // (DiffuseSurfaceEmitter::calcRadiantFluxApprox, DiffuseSurfaceEmitter.cpp:55-86)
    const uint32_t ne = s->num_emitters;
    h.emitters.assign(ne, EmitterRec{});
    h.cdf.assign(ne + 1, 0.0f);
    h.etris.resize(s->num_emitter_tris);
    for (uint32_t i = 0; i < s->num_emitter_tris; ++i) {
        const uint32_t t = s->emitter_tris[i];
        const float* v = s->verts + 9 * (size_t)t;
        h.etris[i].tri = pos_of[t];
        h.etris[i].area = htri_area(H3{v[0], v[1], v[2]}, H3{v[3], v[4], v[5]}, H3{v[6], v[7], v[8]});
    }
    if (ne > 0) {
        std::vector<float> w(ne, 0.0f);
        for (uint32_t e = 0; e < ne; ++e) {
            const phg_emitter& em = s->emitters[e];
            float total = 0.0f;
            for (uint32_t k = 0; k < em.tri_count; ++k) {
                const float flux = hcolor_energy(em.radiance) * h.etris[em.tri_begin + k].area * 3.14159265358979323846f;
                total += flux;
            }
            w[e] = total;
        }
        std::vector<float>& cdf = h.cdf;
        const float delta = 1.0f / (float)ne;
        for (uint32_t i = 1; i <= ne; ++i) cdf[i] = cdf[i - 1] + w[i - 1] * delta;
        const float rcpSum = 1.0f / cdf[ne];
        cdf[0] = 0.0f; cdf[ne] = 1.0f;
        if (std::isfinite(rcpSum)) { for (uint32_t i = 1; i < ne; ++i) cdf[i] *= rcpSum; }
        else { for (uint32_t i = 1; i < ne; ++i) cdf[i] = (float)i / (float)ne; }
        for (uint32_t e = 0; e < ne; ++e) {
            const phg_emitter& em = s->emitters[e];
            EmitterRec& r = h.emitters[e];
            std::memset(&r, 0, sizeof(r));
            r.radiance[0] = em.radiance[0]; r.radiance[1] = em.radiance[1]; r.radiance[2] = em.radiance[2];
            r.pick_pdf = cdf[e + 1] - cdf[e];
            r.tri_begin = em.tri_begin; r.tri_count = em.tri_count;
        }
    }
    h.mats.assign(s->materials, s->materials + s->num_materials);
    h.cam = s->camera;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < 3ull * s->num_tris; ++i) {
        const float* v = s->verts + 3 * i;
        for (int k = 0; k < 3; ++k) {
            if (std::isfinite(v[k])) { lo[k] = std::min(lo[k], v[k]); hi[k] = std::max(hi[k], v[k]); }
        }
    }
    for (int k = 0; k < 3; ++k) {
        const float e = hi[k] - lo[k];
        h.box.lo[k] = std::isfinite(lo[k]) ? lo[k] : 0.0f;
        h.box.inv_ext[k] = (e > 0.0f && std::isfinite(e)) ? 1.0f / e : 1.0f;
    }
    h.nodes = std::move(bvh.nodes);
    h.stack_need = bvh.stack_need;
    h.depth = bvh.depth;
    return PHG_OK;
}

// Pack prepared scenes into the shared device arrays, rebasing node, triangle, material,
// emitter and CDF indices by each scene's offsets.
static phg_status upload_arena(Ctx* c, std::vector<HostScene>& hs) {
    size_t nn = 0, nt = 0, nm = 0, ne = 0, nc = 0, net = 0;
    for (const HostScene& h : hs) {
        nn += h.nodes.size(); nt += h.tris.size(); nm += h.mats.size(); ne += h.emitters.size();
        nc += h.cdf.size(); net += h.etris.size();
    }
    if (nt >= (1ull << kLeafFirstBits) || nn >= 0x7fffffffull)
        return set_err(c, PHG_ERR_UNSUPPORTED, "scenes exceed 2^28 triangles in total");
    std::vector<WideNode> nodes; nodes.reserve(nn);
    std::vector<TriRec> tris; tris.reserve(nt);
    std::vector<phg_material> mats; mats.reserve(nm);
    std::vector<EmitterRec> emitters; emitters.reserve(std::max<size_t>(ne, 1));
    std::vector<float> cdf; cdf.reserve(nc);
    std::vector<EmitTri> etris; etris.reserve(std::max<size_t>(net, 1));
    std::vector<SceneInfo> info(hs.size());
    uint32_t stack_need = 0, depth = 0;
    for (size_t si = 0; si < hs.size(); ++si) {
        HostScene& h = hs[si];
        const uint32_t n0 = (uint32_t)nodes.size(), t0 = (uint32_t)tris.size(), m0 = (uint32_t)mats.size();
        const uint32_t e0 = (uint32_t)emitters.size(), c0 = (uint32_t)cdf.size(), et0 = (uint32_t)etris.size();
        for (WideNode w : h.nodes) {
            for (int k = 0; k < 4; ++k) {
                const int32_t ch = w.child[k];
                if (ch == kEmptyChild) continue;
                if (ch < 0) {
                    const uint32_t u = (uint32_t)ch;
                    const uint32_t first = u & ((1u << kLeafFirstBits) - 1u);
                    w.child[k] = (int32_t)((u & ~((1u << kLeafFirstBits) - 1u)) | (first + t0));
                } else {
                    w.child[k] = ch + (int32_t)n0;
                }
            }
            nodes.push_back(w);
        }
        for (TriRec r : h.tris) {
            r.mat += m0;
            if (r.emitter >= 0) r.emitter += (int32_t)e0;
            tris.push_back(r);
        }
        mats.insert(mats.end(), h.mats.begin(), h.mats.end());
        for (EmitterRec r : h.emitters) { r.tri_begin += et0; emitters.push_back(r); }
        cdf.insert(cdf.end(), h.cdf.begin(), h.cdf.end());
        for (EmitTri e : h.etris) { e.tri += t0; etris.push_back(e); }
        SceneInfo& in = info[si];
        in.root = n0; in.emitter_begin = e0; in.num_emitters = (uint32_t)h.emitters.size(); in.cdf_begin = c0;
        in.cam = h.cam; in.box = h.box; in.num_tris = (uint32_t)h.tris.size(); in.num_nodes = (uint32_t)h.nodes.size();
        stack_need = std::max(stack_need, h.stack_need);
        depth = std::max(depth, h.depth);
        // host copies are no longer needed
        std::vector<WideNode>().swap(h.nodes);
        std::vector<TriRec>().swap(h.tris);
    }
    if (emitters.empty()) emitters.push_back(EmitterRec{});
    if (etris.empty()) etris.push_back(EmitTri{});
    // traversal addresses nodes by 32-bit byte offsets
    if ((uint64_t)nodes.size() * sizeof(WideNode) > 0xFFFFFFFFull) {
        c->err = "BVH of " + std::to_string(nodes.size()) + " nodes exceeds the 4 GiB node arena";
        return PHG_ERR_OUT_OF_MEMORY;
    }
    PHG_HIP(c, hipSetDevice(c->device));
    PHG_HIP(c, hipStreamSynchronize(c->stream));
    PHG_HIP(c, c->nodes.ensure(nodes.size()));
    PHG_HIP(c, c->tris.ensure(tris.size()));
    PHG_HIP(c, c->mats.ensure(mats.size()));
    PHG_HIP(c, c->emitters.ensure(emitters.size()));
    PHG_HIP(c, c->emit_cdf.ensure(cdf.size()));
    PHG_HIP(c, c->emit_tris.ensure(etris.size()));
    PHG_HIP(c, hipMemcpy(c->nodes.p, nodes.data(), nodes.size() * sizeof(WideNode), hipMemcpyHostToDevice));
    PHG_HIP(c, hipMemcpy(c->tris.p, tris.data(), tris.size() * sizeof(TriRec), hipMemcpyHostToDevice));
    PHG_HIP(c, hipMemcpy(c->mats.p, mats.data(), mats.size() * sizeof(phg_material), hipMemcpyHostToDevice));
    PHG_HIP(c, hipMemcpy(c->emitters.p, emitters.data(), emitters.size() * sizeof(EmitterRec), hipMemcpyHostToDevice));
    PHG_HIP(c, hipMemcpy(c->emit_cdf.p, cdf.data(), cdf.size() * sizeof(float), hipMemcpyHostToDevice));
    PHG_HIP(c, hipMemcpy(c->emit_tris.p, etris.data(), etris.size() * sizeof(EmitTri), hipMemcpyHostToDevice));
    c->sc.nodes = c->nodes.p;
    c->sc.tris = c->tris.p;
    c->sc.mats = c->mats.p;
    c->sc.emitters = c->emitters.p;
    c->sc.emit_cdf = c->emit_cdf.p;
    c->sc.emit_tris = c->emit_tris.p;
    c->scenes = std::move(info);
    c->g_scene = c->scenes[0].box;
    c->stack_need = stack_need;
    c->bvh_depth = depth;
    c->bvh_nodes = (uint32_t)nodes.size();
    c->have_scene = true;
    c->stats.bvh_nodes = c->bvh_nodes;
    c->stats.bvh_depth = c->bvh_depth;
    c->stats.bvh_stack_need = c->stack_need;
    c->stats.bvh_build_ms = c->bvh_build_ms;
    return PHG_OK;
}

static phg_status upload_scenes(Ctx* c, const phg_scene_desc* scenes, uint32_t n, const phg_bvh_opts* bo) {
    BvhBuildParams bp;
    if (bo) {
        if (bo->max_leaf_tris) bp.max_leaf_tris = bo->max_leaf_tris;
        if (bo->sah_bins) bp.sah_bins = bo->sah_bins;
        if (bo->width) bp.width = bo->width;
        bp.nthreads = bo->nthreads;
    }
    std::vector<HostScene> hs(n);
    const auto t0 = std::chrono::steady_clock::now();
    if (n == 1) {
        prepare_scene(&scenes[0], bp, hs[0]);
    } else {
        // one scene per worker, each build single-threaded
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const unsigned workers = std::min<unsigned>(n, bo && bo->nthreads ? bo->nthreads : hw);
        BvhBuildParams bp1 = bp;
        bp1.nthreads = 1;
        std::atomic<uint32_t> next{0};
        std::vector<std::thread> pool;
        for (unsigned w = 0; w < workers; ++w)
            pool.emplace_back([&] {
                for (uint32_t i = next.fetch_add(1); i < n; i = next.fetch_add(1)) prepare_scene(&scenes[i], bp1, hs[i]);
            });
        for (auto& t : pool) t.join();
    }
    c->bvh_build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (uint32_t i = 0; i < n; ++i)
        if (hs[i].status != PHG_OK) return set_err(c, hs[i].status, "scene " + std::to_string(i) + ": " + hs[i].err);
    return upload_arena(c, hs);
}

}  // namespace
}  // namespace phg

extern "C" {

phg_status phg_upload_scene(void* ctx, const phg_scene_desc* s, const phg_bvh_opts* bo) {
    return guarded(ctx, [&]() -> phg_status {
        Ctx* c = static_cast<Ctx*>(ctx);
        if (!c || !s) return PHG_ERR_INVALID_ARG;
        return upload_scenes(c, s, 1, bo);
    });
}

phg_status phg_upload_scenes(void* ctx, const phg_scene_desc* scenes, uint32_t n, const phg_bvh_opts* bo) {
    return guarded(ctx, [&]() -> phg_status {
        Ctx* c = static_cast<Ctx*>(ctx);
        if (!c || !scenes || n == 0) return PHG_ERR_INVALID_ARG;
        return upload_scenes(c, scenes, n, bo);
    });
}

static uint32_t trace_grid(const Ctx* c) { return c->cu_count * c->trace_blocks_per_cu; }
static uint32_t trace_threads(const Ctx* c) { return trace_grid(c) * c->trace_block; }
static uint32_t spill_cap(const Ctx* c) {
    return c->stack_need > (uint32_t)kTraceRing ? c->stack_need - (uint32_t)kTraceRing : 0u;
}
static phg_status ensure_spill(Ctx* c) {
    const size_t n = (size_t)trace_threads(c) * std::max<uint32_t>(spill_cap(c), 1u);
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
                                c->spill.p, spill_cap(c), trace_threads(c) / kTraceBlock, c->stream));
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
    return guarded(ctx, [&]() -> phg_status {
        Ctx* c = static_cast<Ctx*>(ctx);
        if (!c) return PHG_ERR_INVALID_ARG;
        return do_intersect(c, rays, n, hits, query, ptr_kind, nullptr, nullptr);
    });
}

phg_status phg_intersect_noalloc(void* ctx, const phg_ray* rays, uint64_t n, phg_hit* hits, int query) {
    return guarded(ctx, [&]() -> phg_status {
        Ctx* c = static_cast<Ctx*>(ctx);
        if (!c) return PHG_ERR_INVALID_ARG;
        if (!c->have_scene) return set_err(c, PHG_ERR_NO_SCENE, "no scene uploaded");
        if (n == 0) return PHG_OK;
        if (!rays || !hits || (query != PHG_QUERY_CLOSEST && query != PHG_QUERY_ANY))
            return set_err(c, PHG_ERR_INVALID_ARG, "bad intersect arguments");
        // the grid is sized to the scratch: blocks of kTraceBlock lanes, each lane with its
        // share of the traversal-stack spill (the kernel streams any number of rays per lane)
        const size_t lane_entries = std::max<uint32_t>(spill_cap(c), 1u);
        const size_t block_bytes = (size_t)kTraceBlock * lane_entries * sizeof(int32_t);
        const size_t fit_blocks = (kNoAllocScratchBytes - kNoAllocFlagBytes) / block_bytes;
        if (fit_blocks == 0)
            return set_err(c, PHG_ERR_UNSUPPORTED, "BVH too deep: one block's traversal-stack spill exceeds the "
                                                   "allocation-free query's scratch");
        const uint32_t grid = (uint32_t)std::min<size_t>(trace_threads(c) / kTraceBlock, fit_blocks);
        uint32_t* d_flags = static_cast<uint32_t*>(c->noalloc_scratch);
        int32_t* d_spill = reinterpret_cast<int32_t*>(static_cast<unsigned char*>(c->noalloc_scratch) + kNoAllocFlagBytes);
        PHG_HIP(c, hipSetDevice(c->device));
        PHG_HIP(c, launch_clear_counts(d_flags, 1, c->stream));   // a fill kernel, no memset node
        PHG_HIP(c, hipEventRecord(c->ev[0], c->stream));
        PHG_HIP(c, launch_intersect(c->sc, rays, n, hits, query == PHG_QUERY_ANY, nullptr, nullptr, d_flags, d_spill,
                                    spill_cap(c), grid, c->stream));
        PHG_HIP(c, hipEventRecord(c->ev[1], c->stream));
        uint32_t flags = 0;
        PHG_HIP(c, hipMemcpyAsync(&flags, d_flags, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        PHG_HIP(c, hipStreamSynchronize(c->stream));
        float ms = 0.0f;
        PHG_HIP(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        if (query == PHG_QUERY_ANY) c->stats.shadow_ms = ms; else c->stats.trace_ms = ms;
        if (flags) return set_err(c, PHG_ERR_UNSUPPORTED, "traversal stack overflow");
        return PHG_OK;
    });
}

phg_status phg_intersect_counted(void* ctx, const phg_ray* rays, uint64_t n, phg_hit* hits, uint32_t* node_visits,
                                 uint32_t* tri_tests, int ptr_kind) {
    return guarded(ctx, [&]() -> phg_status {
        Ctx* c = static_cast<Ctx*>(ctx);
        if (!c) return PHG_ERR_INVALID_ARG;
        if (!node_visits || !tri_tests) return set_err(c, PHG_ERR_INVALID_ARG, "counter outputs required");
        return do_intersect(c, rays, n, hits, PHG_QUERY_CLOSEST, ptr_kind, node_visits, tri_tests);
    });
}

phg_status phg_film_clear(void* ctx, uint32_t width, uint32_t height) {
    return guarded(ctx, [&]() -> phg_status {
        Ctx* c = static_cast<Ctx*>(ctx);
        if (!c || width == 0 || height == 0) return PHG_ERR_INVALID_ARG;
        PHG_HIP(c, hipSetDevice(c->device));
        const uint64_t n = 8ull * width * height;
        PHG_HIP(c, c->film.ensure(n));
        PHG_HIP(c, launch_fill_zero(c->film.p, n, c->stream));
        PHG_HIP(c, hipStreamSynchronize(c->stream));
        c->film_w = width; c->film_h = height;
        return PHG_OK;
    });
}

phg_status phg_film_read(void* ctx, double* out, uint64_t out_len) {
    return guarded(ctx, [&]() -> phg_status {
        Ctx* c = static_cast<Ctx*>(ctx);
        if (!c || !out) return PHG_ERR_INVALID_ARG;
        const uint64_t n = 8ull * c->film_w * c->film_h;
        if (n == 0) return set_err(c, PHG_ERR_INVALID_ARG, "film not initialised");
        if (out_len < n) return set_err(c, PHG_ERR_INVALID_ARG, "output buffer too small");
        PHG_HIP(c, hipSetDevice(c->device));
        PHG_HIP(c, hipMemcpyAsync(out, c->film.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        PHG_HIP(c, hipStreamSynchronize(c->stream));
        return PHG_OK;
    });
}

}  // extern "C"

namespace phg {
namespace {

struct RenderMode {
    bool count_traversal = false, guided = false, train = false;
    bool graph = false;           // replay a captured bounce loop per pass (bounded depth, unguided)
    float guide_prob = 0.5f, guide_lr = 1e-3f;
    uint32_t guide_batch = 1u << 18, guide_steps = 4, guide_depth = 0;
};

struct PassPlan {
    std::vector<PassSeg> segs;
    uint32_t paths = 0, film_threads = 0, max_bounces = 0;
};

// Render the planned passes of the given jobs into c->film / c->bfilm (per job film_base).
// job_done_ms (optional, one per job) receives the time from the start to the end of the
// pass that completes each job.
static phg_status render_passes(Ctx* c, const std::vector<JobDev>& jobs, const std::vector<PassPlan>& passes,
                                double* film, const RenderMode& m, std::vector<double>* job_done_ms) {
    uint32_t cap = 0;
    for (const PassPlan& pl : passes) {
        cap = std::max(cap, pl.paths);
        if (pl.segs.empty() || pl.segs.size() > kMaxSegs) return set_err(c, PHG_ERR_INVALID_ARG, "bad pass plan");
    }
    if (cap == 0) return PHG_OK;
    // a fixed selection probability lies in [0, 1]; a negative one selects the learned one
    if (m.guided && !(m.guide_prob <= 1.0f))
        return set_err(c, PHG_ERR_INVALID_ARG, "guide_prob must be at most 1 (negative: learned per query)");
    PHG_HIP(c, hipSetDevice(c->device));
    PHG_HIP(c, c->rng.ensure(cap));
    PHG_HIP(c, c->T.ensure(cap)); PHG_HIP(c, c->Wem.ensure(cap)); PHG_HIP(c, c->prevP.ensure(cap)); PHG_HIP(c, c->L.ensure(cap));
    // the ray queues: 8 regions of queue_region_cap(paths) slots (a pass of fewer paths uses
    // smaller regions of the same buffers)
    const size_t qcap = queue_capacity(cap);
    PHG_HIP(c, c->hits.ensure(qcap)); PHG_HIP(c, c->shadow_c.ensure(qcap));
    PHG_HIP(c, c->ext0.ensure(qcap)); PHG_HIP(c, c->ext1.ensure(qcap)); PHG_HIP(c, c->shadow.ensure(qcap));
    PHG_HIP(c, c->ext_pid0.ensure(qcap)); PHG_HIP(c, c->ext_pid1.ensure(qcap)); PHG_HIP(c, c->shadow_pid.ensure(qcap));
    PHG_HIP(c, c->path_seg.ensure(cap));
    // + the bounce base of captured chunks and the pass's queue region size
    PHG_HIP(c, c->counts.ensure(kCountSlots + 2));
    PHG_HIP(c, c->fetch.ensure(kFetchWords));
    PHG_HIP(c, c->tcount.ensure(8));
    PHG_HIP(c, c->errflags.ensure(1));
    PHG_HIP(c, c->jobs.ensure(jobs.size()));
    PHG_HIP(c, c->d_pass.ensure(1));
    PHG_HIP(c, c->d_segs.ensure(kMaxSegs));
    if (!c->h_pass) {
        void* hp = nullptr;
        PHG_HIP(c, hipHostMalloc(&hp, sizeof(PassParams) + kMaxSegs * sizeof(PassSeg), hipHostMallocDefault));
        c->h_pass = static_cast<PassParams*>(hp);
        c->h_segs = reinterpret_cast<PassSeg*>(static_cast<char*>(hp) + sizeof(PassParams));
    }
    PHG_HIP(c, hipMemcpyAsync(c->jobs.p, jobs.data(), jobs.size() * sizeof(JobDev), hipMemcpyHostToDevice, c->stream));
    PHG_HIP(c, hipMemsetAsync(c->tcount.p, 0, 8 * sizeof(unsigned long long), c->stream));
    PHG_HIP(c, hipMemsetAsync(c->errflags.p, 0, sizeof(uint32_t), c->stream));
    PHG_HIP(c, launch_clear_counts(c->counts.p, kCountSlots, c->stream));
    PHG_HIP(c, launch_clear_counts(c->fetch.p, kFetchWords, c->stream));
    SceneDev sc = c->sc;
    sc.jobs = c->jobs.p;
    PathBuffers pb{};
    pb.rng = c->rng.p; pb.T = c->T.p; pb.Wem = c->Wem.p; pb.prevP = c->prevP.p; pb.L = c->L.p;
    pb.ext[0] = c->ext0.p; pb.ext[1] = c->ext1.p; pb.ext_pid[0] = c->ext_pid0.p; pb.ext_pid[1] = c->ext_pid1.p;
    pb.hits = c->hits.p; pb.shadow = c->shadow.p; pb.shadow_pid = c->shadow_pid.p; pb.shadow_c = c->shadow_c.p;
    pb.counts = c->counts.p; pb.traversal_counters = c->tcount.p; pb.error_flags = c->errflags.p; pb.film = film;
    pb.path_seg = c->path_seg.p;
    pb.bounce_base = c->counts.p + kCountSlots;
    pb.fetch = c->fetch.p;
    {
        const phg_status st = ensure_spill(c);
        if (st != PHG_OK) return st;
    }
    pb.spill = c->spill.p;
    pb.spill_cap = spill_cap(c);
    const bool guided = m.guided, train = m.train;
    uint32_t rec_bounces = 0;
    uint32_t train_cap = 0;
    const uint32_t batch = m.guide_batch ? m.guide_batch : (1u << 18);
    if (guided || train) {
        if (!c->g_init) {
            const std::vector<float> p0 = guide_default_params(jobs[0].seed + 1);
            phg_status st = guide_load(c, p0.data());
            if (st != PHG_OK) return st;
        }
        PHG_HIP(c, c->g_lobes.ensure(queue_capacity(cap) * kGuideOutputs));
    }
    uint32_t max_b = 0;
    for (const PassPlan& pl : passes) max_b = std::max(max_b, pl.max_bounces);
    if (train) {
        rec_bounces = std::min<uint32_t>(kGuideRecBounces, max_b + 1);
        // the network is queried at the first guide_depth vertices only: fit it to those
        if (m.guide_depth > 0) rec_bounces = std::min(rec_bounces, m.guide_depth);
        PHG_HIP(c, c->g_rec.ensure((size_t)rec_bounces * cap * kRecFloat4));
        // fresh records for every step of a pass (no record is reused within the pass)
        train_cap = std::max<uint32_t>(4u, m.guide_steps ? m.guide_steps : 4u) * batch;
        PHG_HIP(c, c->g_train.ensure((size_t)train_cap * kGuideRecordFloats));
    }
    pb.lobes = c->g_lobes.p;
    pb.rec = c->g_rec.p;
    double t_ginfer = 0, t_gtrain = 0;
    uint64_t g_records = 0;
    const uint32_t grid_trace = trace_grid(c);     // persistent-style grid-stride launches
    const uint32_t grid_shade = c->cu_count * c->shade_blocks_per_cu;
    double t_trace = 0, t_shadow = 0, t_shade = 0;
    uint64_t rays_primary = 0, paths_total = 0;
    uint32_t n_trace = 0, n_shadow = 0;
    const uint32_t count_trav = m.count_traversal ? 1u : 0u;
    const PassParams* ppd = c->d_pass.p;
    // Captured pass: generate, max_b + 1 bounce iterations (extend, shade, shadow; a bounce
    // whose queue is empty costs three near-empty launches), film.  Replayed once per pass
    // after the pass parameters are copied; timed as a whole (no per-kernel times).
    hipGraphExec_t gexec = nullptr;
    const bool use_graph = m.graph && !guided && !train && max_b + 2 < kCountSlots;
    // Otherwise (unbounded depth, guided or training renders) a graph mode replays captured
    // chunks of kChunkBounces bounces -- extend, inference where guided, shade, shadow, then
    // the bounce base advanced -- until the extension queue is empty: one host round trip per
    // chunk instead of per bounce.  Chunks differ only in which bounces sample the lobes, so
    // a graph is captured per pattern (bit k: inference + guided shading at bounce m0 + k).
    const bool use_chunks = m.graph && !use_graph;
    struct ChunkGraphs {
        std::vector<std::pair<uint32_t, hipGraphExec_t>> g;
        ~ChunkGraphs() { for (auto& e : g) (void)hipGraphExecDestroy(e.second); }
    } chunk_graphs;
    auto chunk_graph = [&](uint32_t infer_mask, uint32_t guided_mask, hipGraphExec_t* out) -> phg_status {
        const uint32_t key = infer_mask | (guided_mask << 16);
        for (auto& e : chunk_graphs.g)
            if (e.first == key) { *out = e.second; return PHG_OK; }
        hipGraph_t graph = nullptr;
        PHG_HIP(c, hipStreamSynchronize(c->stream));   // nothing of the pass in flight while capturing
        PHG_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        hipError_t ce = hipSuccess;
        for (uint32_t k = 0; k < kChunkBounces && ce == hipSuccess; ++k) {
            ce = launch_extend(sc, pb, k, grid_trace, c->trace_block, count_trav, c->stream);
            if (ce == hipSuccess && (infer_mask >> k & 1u))
                ce = guide_infer_hits(guide_net(c), pb.hits, pb.ext[k & 1u], sc.tris, c->g_scene, c->fetch.p,
                                      pb.bounce_base, k, c->g_lobes.p, c->tcount.p + 6, c->cu_count * 2, c->stream);
            if (ce == hipSuccess) ce = launch_shade(sc, ppd, pb, k, grid_shade, (guided_mask >> k & 1u) != 0, c->stream);
            if (ce == hipSuccess) ce = launch_shadow(sc, ppd, pb, k, grid_trace, c->trace_block, count_trav, c->stream);
        }
        if (ce == hipSuccess) ce = launch_advance_bounce(pb.bounce_base, kChunkBounces, c->stream);
        const hipError_t ee = hipStreamEndCapture(c->stream, &graph);
        if (ce != hipSuccess || ee != hipSuccess) {
            (void)hipGetLastError();
            if (graph) (void)hipGraphDestroy(graph);
            c->err = std::string("bounce-chunk capture: ") + hipGetErrorString(ce != hipSuccess ? ce : ee);
            return PHG_ERR_HIP;
        }
        hipGraphExec_t ex = nullptr;
        const hipError_t ie = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) {
            (void)hipGetLastError();
            c->err = std::string("bounce-chunk graph: ") + hipGetErrorString(ie);
            return PHG_ERR_HIP;
        }
        chunk_graphs.g.emplace_back(key, ex);
        *out = ex;
        return PHG_OK;
    };
    if (use_graph) {
        hipGraph_t graph = nullptr;
        PHG_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        hipError_t ce = launch_generate(sc, ppd, pb, grid_shade, c->stream);
        for (uint32_t b = 0; b <= max_b && ce == hipSuccess; ++b) {
            ce = launch_extend(sc, pb, b, grid_trace, c->trace_block, count_trav, c->stream);
            if (ce == hipSuccess) ce = launch_shade(sc, ppd, pb, b, grid_shade, false, c->stream);
            if (ce == hipSuccess) ce = launch_shadow(sc, ppd, pb, b, grid_trace, c->trace_block, count_trav, c->stream);
        }
        if (ce == hipSuccess) ce = launch_film(sc, ppd, pb, grid_shade, c->stream);
        const hipError_t ee = hipStreamEndCapture(c->stream, &graph);
        if (ce != hipSuccess || ee != hipSuccess) {
            (void)hipGetLastError();
            if (graph) (void)hipGraphDestroy(graph);
            c->err = std::string("bounce-loop capture: ") + hipGetErrorString(ce != hipSuccess ? ce : ee);
            return PHG_ERR_HIP;
        }
        const hipError_t ie = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) {
            (void)hipGetLastError();
            c->err = std::string("bounce-loop graph: ") + hipGetErrorString(ie);
            return PHG_ERR_HIP;
        }
    }
    double t_graph = 0;
    struct GraphGuard {
        hipGraphExec_t g;
        ~GraphGuard() { if (g) (void)hipGraphExecDestroy(g); }
    } graph_guard{gexec};
    const auto wall0 = std::chrono::steady_clock::now();
    for (const PassPlan& pl : passes) {
        PassParams pp{};
        pp.paths = pl.paths;
        pp.num_segs = (uint32_t)pl.segs.size();
        pp.film_threads = pl.film_threads;
        pp.max_bounces = pl.max_bounces;
        pp.guide_prob = (guided && c->g_ready) ? m.guide_prob : 0.0f;
        pp.guide_depth = m.guide_depth;
        pp.record = train ? 1u : 0u;
        pp.rec_bounces = rec_bounces;
        pp.segs = c->d_segs.p;
        *c->h_pass = pp;
        std::memcpy(c->h_segs, pl.segs.data(), pl.segs.size() * sizeof(PassSeg));
        PHG_HIP(c, hipMemcpyAsync(c->d_pass.p, c->h_pass, sizeof(PassParams), hipMemcpyHostToDevice, c->stream));
        PHG_HIP(c, hipMemcpyAsync(c->d_segs.p, c->h_segs, pl.segs.size() * sizeof(PassSeg), hipMemcpyHostToDevice, c->stream));
        rays_primary += pp.paths;
        paths_total += pp.paths;
        if (gexec) {
            PHG_HIP(c, hipEventRecord(c->ev[0], c->stream));
            PHG_HIP(c, hipGraphLaunch(gexec, c->stream));
            PHG_HIP(c, hipEventRecord(c->ev[1], c->stream));
            PHG_HIP(c, hipStreamSynchronize(c->stream));
            float ms = 0;
            PHG_HIP(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
            t_graph += ms;
            n_trace += max_b + 1; n_shadow += max_b + 1;
            if (job_done_ms) {
                const double now = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
                for (const PassSeg& sg : pl.segs) (*job_done_ms)[sg.job] = now;
            }
            continue;
        }
        float ms = 0;
        if (use_chunks) {
            PHG_HIP(c, hipEventRecord(c->ev[0], c->stream));
            PHG_HIP(c, launch_generate(sc, ppd, pb, grid_shade, c->stream));
            for (uint32_t m0 = 0;; m0 += kChunkBounces) {
                uint32_t infer_mask = 0, guided_mask = 0;
                for (uint32_t k = 0; k < kChunkBounces; ++k) {
                    const uint32_t b = m0 + k;
                    const bool in_depth = m.guide_depth == 0 || b < m.guide_depth;
                    if (guided && pp.guide_prob != 0.0f && b < pl.max_bounces && in_depth) infer_mask |= 1u << k;
                    if ((train || guided) && in_depth) guided_mask |= 1u << k;
                }
                hipGraphExec_t ex = nullptr;
                const phg_status gs = chunk_graph(infer_mask, guided_mask, &ex);
                if (gs != PHG_OK) return gs;
                PHG_HIP(c, hipGraphLaunch(ex, c->stream));
                uint32_t next = 0;
                PHG_HIP(c, hipMemcpyAsync(&next, c->counts.p + (m0 + kChunkBounces) % kCountSlots, sizeof(uint32_t),
                                          hipMemcpyDeviceToHost, c->stream));
                PHG_HIP(c, hipStreamSynchronize(c->stream));
                n_trace += kChunkBounces; n_shadow += kChunkBounces;
                if (next == 0 || m0 + kChunkBounces > pl.max_bounces) break;
            }
            PHG_HIP(c, launch_film(sc, ppd, pb, grid_shade, c->stream));
            PHG_HIP(c, hipEventRecord(c->ev[1], c->stream));
            PHG_HIP(c, hipStreamSynchronize(c->stream));
            PHG_HIP(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
            t_graph += ms;
        } else {
        PHG_HIP(c, hipEventRecord(c->ev[0], c->stream));
        PHG_HIP(c, launch_generate(sc, ppd, pb, grid_shade, c->stream));
        PHG_HIP(c, hipEventRecord(c->ev[1], c->stream));
        PHG_HIP(c, hipStreamSynchronize(c->stream));
        PHG_HIP(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1])); t_shade += ms;
        for (uint32_t b = 0;; ++b) {
            const bool use_guide = guided && pp.guide_prob != 0.0f && b < pl.max_bounces &&
                                   (m.guide_depth == 0 || b < m.guide_depth);
            PHG_HIP(c, hipEventRecord(c->ev[0], c->stream));
            PHG_HIP(c, launch_extend(sc, pb, b, grid_trace, c->trace_block, count_trav, c->stream));
            PHG_HIP(c, hipEventRecord(c->ev[1], c->stream));
            if (use_guide) {
                PHG_HIP(c, guide_infer_hits(guide_net(c), pb.hits, pb.ext[b & 1], sc.tris, c->g_scene, c->fetch.p,
                                            pb.bounce_base, b, c->g_lobes.p, c->tcount.p + 6, c->cu_count * 2,
                                            c->stream));
            }
            PHG_HIP(c, hipEventRecord(c->ev[4], c->stream));
            // the guided shading variant where lobes are sampled or training records written
            // (records are kept for bounces < guide_depth); deeper bounces shade as plain
            // BNEEPT (the same estimator with no lobe term, fewer random numbers per vertex)
            const bool guided_bounce = (train || guided) && (m.guide_depth == 0 || b < m.guide_depth);
            PHG_HIP(c, launch_shade(sc, ppd, pb, b, grid_shade, guided_bounce, c->stream));
            PHG_HIP(c, hipEventRecord(c->ev[2], c->stream));
            PHG_HIP(c, launch_shadow(sc, ppd, pb, b, grid_trace, c->trace_block, count_trav, c->stream));
            PHG_HIP(c, hipEventRecord(c->ev[3], c->stream));
            uint32_t next = 0;
            PHG_HIP(c, hipMemcpyAsync(&next, c->counts.p + (b + 1) % kCountSlots, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
            PHG_HIP(c, hipStreamSynchronize(c->stream));
            float a = 0, bb = 0, cc = 0, gi = 0;
            PHG_HIP(c, hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
            PHG_HIP(c, hipEventElapsedTime(&gi, c->ev[1], c->ev[4]));
            PHG_HIP(c, hipEventElapsedTime(&bb, c->ev[4], c->ev[2]));
            PHG_HIP(c, hipEventElapsedTime(&cc, c->ev[2], c->ev[3]));
            t_trace += a; t_shade += bb; t_shadow += cc; t_ginfer += gi;
            ++n_trace; ++n_shadow;
            if (next == 0 || b >= pl.max_bounces) break;
        }
        PHG_HIP(c, hipEventRecord(c->ev[0], c->stream));
        PHG_HIP(c, launch_film(sc, ppd, pb, grid_shade, c->stream));
        PHG_HIP(c, hipEventRecord(c->ev[1], c->stream));
        PHG_HIP(c, hipStreamSynchronize(c->stream));
        PHG_HIP(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1])); t_shade += ms;
        }
        if (job_done_ms) {
            const double now = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
            for (const PassSeg& sg : pl.segs) (*job_done_ms)[sg.job] = now;  // the last pass of a job wins
        }
        if (train) {
            // online training on this pass's radiance samples
            PHG_HIP(c, hipEventRecord(c->ev[0], c->stream));
            PHG_HIP(c, c->g_blockcnt.ensure(2 * (size_t)guide_collect_blocks(pp.paths)));
            PHG_HIP(c, guide_collect(pb, pp, c->g_scene, c->g_train.p, train_cap, c->g_counters.p, c->g_blockcnt.p,
                                     c->g_keep, ++c->g_pass_salt, c->stream));
            uint32_t cnt[2] = {0, 0};
            PHG_HIP(c, hipMemcpyAsync(cnt, c->g_counters.p, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
            PHG_HIP(c, hipStreamSynchronize(c->stream));
            const uint32_t nrec = std::min(cnt[0], train_cap);
            const uint32_t steps = m.guide_steps ? m.guide_steps : 4;
            phg_status gs = guide_train_device(c, c->g_train.p, nrec, batch, m.guide_lr > 0 ? m.guide_lr : 1e-3f, steps,
                                               &g_records);
            if (gs != PHG_OK) return gs;
            PHG_HIP(c, hipEventRecord(c->ev[1], c->stream));
            PHG_HIP(c, hipStreamSynchronize(c->stream));
            PHG_HIP(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1])); t_gtrain += ms;
            // keep about as many records as the next pass's steps consume
            const double want = (double)steps * batch;
            c->g_keep = cnt[1] > 0 ? (float)std::min(1.0, 1.25 * want / (double)cnt[1]) : 1.0f;
            if (nrec > 0) c->g_ready = true;
        }
    }
    uint32_t flags = 0;
    unsigned long long tc[8];
    PHG_HIP(c, hipMemcpyAsync(&flags, c->errflags.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    PHG_HIP(c, hipMemcpyAsync(tc, c->tcount.p, sizeof(tc), hipMemcpyDeviceToHost, c->stream));
    PHG_HIP(c, hipStreamSynchronize(c->stream));
    c->stats.render_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    c->stats.paths = paths_total;
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
    c->stats.guide_queries = tc[6];
    c->stats.guide_records = g_records;
    c->stats.guide_steps = c->g_steps;
    c->stats.trace_ms = t_trace;
    c->stats.shadow_ms = t_shadow;
    c->stats.shade_ms = t_shade;
    c->stats.graph_ms = t_graph;
    if (flags) return set_err(c, PHG_ERR_UNSUPPORTED, "traversal stack overflow");
    return PHG_OK;
}

// A sample-source value the renderer takes: a kind, and for Halton a permutation and a sequence.
static bool sampler_ok(uint32_t sampler) {
    const uint32_t kind = sampler & 0xFFu;
    if (kind == PHG_SAMPLER_HALTON) return (sampler & ~0x33FFu) == 0u;
    return kind <= PHG_SAMPLER_STRATIFIED && (sampler >> 8) == 0u;
}

// The n-th good prime (0-based): a prime p_k with p_k^2 > p_{k-i} p_{k+i} for every
// 1 <= i <= k (math_table.h GOOD_PRIME lists the first 512, for the Halton leaps).
static uint32_t good_prime(uint32_t n) {
    static const std::vector<uint32_t> good = [] {
        std::vector<uint32_t> primes;
        for (uint32_t v = 2; primes.size() < 16384; ++v) {
            bool p = true;
            for (uint32_t q : primes) {
                if (q * q > v) break;
                if (v % q == 0) { p = false; break; }
            }
            if (p) primes.push_back(v);
        }
        std::vector<uint32_t> out;
        for (size_t k = 1; k < primes.size() / 2 && out.size() < 512; ++k) {
            bool ok = true;
            for (size_t i = 1; i <= k && ok; ++i)
                ok = (uint64_t)primes[k] * primes[k] > (uint64_t)primes[k - i] * primes[k + i];
            if (ok) out.push_back(primes[k]);
        }
        return out;
    }();
    return good[n];
}

// moremur bit mixer (Math/hash.ipp:104-115), for host-side keyed choices
static uint64_t mix64(uint64_t v) {
    v ^= v >> 27; v *= 0x3C79AC492BA7B653ULL;
    v ^= v >> 33; v *= 0x1C69B3F74AC4AE35ULL;
    v ^= v >> 27;
    return v;
}

// Index step of a Halton job's sequence (SGHalton.cpp:73-98): 1; the 257th good prime for
// "leap"; for "random-leap" one of the good primes 152..512, keyed by the seed here.
static uint32_t halton_leap(uint32_t sampler, uint64_t seed) {
    if ((sampler & 0xFFu) != PHG_SAMPLER_HALTON) return 1u;
    const uint32_t seq = sampler & 0x3000u;
    if (seq == PHG_HALTON_SEQ_LEAP) return good_prime(256);
    if (seq == PHG_HALTON_SEQ_RANDOM_LEAP) return good_prime(151u + (uint32_t)(mix64(seed ^ 0x51ED270B27ACE8CDULL) % 361u));
    return 1u;
}

static JobDev make_job(const SceneInfo& si, uint32_t width, uint32_t height, uint32_t max_bounces, uint64_t seed,
                       uint64_t film_base, uint32_t estimator, uint32_t filter, uint32_t sampler) {
    JobDev j{};
    j.filter = filter;
    j.sampler = sampler;
    j.halton_leap = halton_leap(sampler, seed);
    // BVPTDL is BVPT capped at one bounce
    j.estimator = estimator == PHG_ESTIMATOR_BNEEPT ? PHG_ESTIMATOR_BNEEPT : PHG_ESTIMATOR_BVPT;
    if (estimator == PHG_ESTIMATOR_BVPTDL) max_bounces = std::min(max_bounces, 1u);
    j.cam = si.cam;
    j.root = si.root;
    j.width = width; j.height = height; j.npix = width * height;
    j.emitter_begin = si.emitter_begin; j.num_emitters = si.num_emitters; j.cdf_begin = si.cdf_begin;
    j.max_bounces = max_bounces;
    j.seed = seed;
    j.film_base = film_base;
    return j;
}

}  // namespace
}  // namespace phg

extern "C" {

phg_status phg_render(void* ctx, const phg_render_opts* o) {
    return guarded(ctx, [&]() -> phg_status {
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
        // samples per pass: fill the tuned number of paths in flight
        uint32_t k = o->spp_per_pass;
        if (k == 0) {
            uint64_t ppp = c->paths_per_pass;
            if (o->flags & PHG_RENDER_TRAIN) ppp = std::min<uint64_t>(ppp, kTrainPassPaths);
            k = (uint32_t)std::max<uint64_t>(1, ppp / npix);
        }
        k = std::min(k, o->spp);
        if (npix * k > 0xffffffffull) return set_err(c, PHG_ERR_INVALID_ARG, "too many paths per pass");
        if (o->estimator > PHG_ESTIMATOR_BVPTDL) return set_err(c, PHG_ERR_INVALID_ARG, "unknown estimator");
        if (o->estimator != PHG_ESTIMATOR_BNEEPT && (o->flags & (PHG_RENDER_GUIDED | PHG_RENDER_TRAIN)))
            return set_err(c, PHG_ERR_UNSUPPORTED, "guiding needs the BNEEPT estimator");
        if (o->filter > PHG_FILTER_BH) return set_err(c, PHG_ERR_INVALID_ARG, "unknown filter");
        if (!sampler_ok(o->sampler)) return set_err(c, PHG_ERR_INVALID_ARG, "unknown sampler");
        std::vector<JobDev> jobs{
            make_job(c->scenes[0], o->width, o->height, o->max_bounces, o->seed, 0, o->estimator, o->filter, o->sampler)};
        std::vector<PassPlan> passes;
        for (uint32_t s0 = 0; s0 < o->spp; s0 += k) {
            const uint32_t kk = std::min(k, o->spp - s0);
            PassPlan pl;
            PassSeg sg{};
            sg.job = 0; sg.sample_begin = o->spp_begin + s0; sg.samples = kk; sg.path_offset = 0;
            sg.pix_offset = 0; sg.npix = (uint32_t)npix;
            pl.segs.push_back(sg);
            pl.paths = (uint32_t)(npix * kk);
            pl.film_threads = (uint32_t)npix;
            pl.max_bounces = o->max_bounces;
            passes.push_back(std::move(pl));
        }
        RenderMode m;
        m.count_traversal = (o->flags & PHG_RENDER_COUNT_TRAVERSAL) != 0;
        m.guided = (o->flags & PHG_RENDER_GUIDED) != 0;
        m.train = (o->flags & PHG_RENDER_TRAIN) != 0;
        m.graph = (o->flags & PHG_RENDER_GRAPH) != 0;
        m.guide_prob = o->guide_prob; m.guide_lr = o->guide_lr; m.guide_batch = o->guide_batch; m.guide_steps = o->guide_steps;
        m.guide_depth = o->guide_depth;
        c->g_scene = c->scenes[0].box;
        return render_passes(c, jobs, passes, c->film.p, m, nullptr);
    });
}

phg_status phg_render_batch(void* ctx, const phg_batch_job* in, uint32_t n, const phg_batch_opts* bo,
                            phg_batch_result* results) {
    return guarded(ctx, [&]() -> phg_status {
        Ctx* c = static_cast<Ctx*>(ctx);
        if (!c || !in || n == 0) return PHG_ERR_INVALID_ARG;
        if (!c->have_scene) return set_err(c, PHG_ERR_NO_SCENE, "no scene uploaded");
        uint64_t cap = bo && bo->paths_per_pass ? bo->paths_per_pass : (uint64_t)c->paths_per_pass;
        if (!(bo && bo->paths_per_pass) && bo && (bo->flags & PHG_BATCH_TRAIN)) cap = std::min<uint64_t>(cap, kTrainPassPaths);
        if (cap > 0xffffffffull) return set_err(c, PHG_ERR_INVALID_ARG, "paths_per_pass too large");
        std::vector<JobDev> jobs(n);
        uint64_t film_doubles = 0;
        for (uint32_t j = 0; j < n; ++j) {
            const phg_batch_job& b = in[j];
            const uint64_t npix = (uint64_t)b.width * b.height;
            if (b.scene >= c->scenes.size()) return set_err(c, PHG_ERR_INVALID_ARG, "job scene index out of range");
            if (npix == 0 || npix > (1ull << 26)) return set_err(c, PHG_ERR_INVALID_ARG, "job image size out of range");
            if (b.seed >= (1u << 24)) return set_err(c, PHG_ERR_INVALID_ARG, "job seed must be below 2^24");
            if (b.spp == 0 || (uint64_t)b.spp > (1ull << 40) / npix) return set_err(c, PHG_ERR_INVALID_ARG, "job spp out of range");
            if (b.estimator > PHG_ESTIMATOR_BVPTDL) return set_err(c, PHG_ERR_INVALID_ARG, "unknown job estimator");
            if (b.filter > PHG_FILTER_BH) return set_err(c, PHG_ERR_INVALID_ARG, "unknown job filter");
            if (!sampler_ok(b.sampler)) return set_err(c, PHG_ERR_INVALID_ARG, "unknown job sampler");
            jobs[j] = make_job(c->scenes[b.scene], b.width, b.height, b.max_bounces, b.seed, film_doubles, b.estimator,
                               b.filter, b.sampler);
            if (b.flags & PHG_JOB_CAMERA) jobs[j].cam = b.camera;
            film_doubles += 8 * npix;
        }
        // pass plan: jobs in the requested order, each pass filled first-fit with whole samples
        std::vector<uint32_t> order(n);
        for (uint32_t j = 0; j < n; ++j) order[j] = j;
        if (bo && bo->order == PHG_BATCH_ORDER_SHORTEST_FIRST)
            std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
                return (uint64_t)jobs[a].npix * in[a].spp < (uint64_t)jobs[b].npix * in[b].spp;
            });
        std::vector<uint32_t> left(n), next(n, 0);
        for (uint32_t j = 0; j < n; ++j) left[j] = in[j].spp;
        std::vector<PassPlan> passes;
        size_t first = 0;
        while (first < n) {
            PassPlan pl;
            uint64_t used = 0;
            for (size_t oi = first; oi < n && pl.segs.size() < kMaxSegs; ++oi) {
                const uint32_t j = order[oi];
                if (left[j] == 0) continue;
                const uint64_t npix = jobs[j].npix;
                uint64_t fit = (cap - std::min(cap, used)) / npix;
                if (fit == 0 && used == 0) fit = 1;  // a job larger than a pass: one sample per pass
                if (fit == 0) continue;
                const uint32_t take = (uint32_t)std::min<uint64_t>(fit, left[j]);
                PassSeg sg{};
                sg.job = j; sg.sample_begin = next[j]; sg.samples = take; sg.path_offset = (uint32_t)used;
                sg.pix_offset = pl.film_threads; sg.npix = (uint32_t)npix;
                pl.segs.push_back(sg);
                used += (uint64_t)take * npix;
                pl.film_threads += (uint32_t)npix;
                pl.max_bounces = std::max(pl.max_bounces, in[j].max_bounces);
                next[j] += take; left[j] -= take;
            }
            pl.paths = (uint32_t)used;
            passes.push_back(std::move(pl));
            while (first < n && left[order[first]] == 0) ++first;
        }
        PHG_HIP(c, hipSetDevice(c->device));
        PHG_HIP(c, c->bfilm.ensure(film_doubles));
        PHG_HIP(c, launch_fill_zero(c->bfilm.p, film_doubles, c->stream));
        c->bfilm_jobs.assign(jobs.begin(), jobs.end());
        RenderMode m;
        m.count_traversal = bo && (bo->flags & PHG_BATCH_COUNT_TRAVERSAL);
        m.graph = bo && (bo->flags & PHG_BATCH_GRAPH);
        m.guided = bo && (bo->flags & PHG_BATCH_GUIDED);
        m.train = bo && (bo->flags & PHG_BATCH_TRAIN);
        if (m.guided || m.train) {
            for (uint32_t j = 0; j < n; ++j) {
                if (in[j].scene != in[0].scene)
                    return set_err(c, PHG_ERR_UNSUPPORTED, "guided batches render one scene (the network is shared)");
                if (in[j].estimator != PHG_ESTIMATOR_BNEEPT)
                    return set_err(c, PHG_ERR_UNSUPPORTED, "guiding needs the BNEEPT estimator");
            }
            m.guide_prob = bo->guide_prob; m.guide_lr = bo->guide_lr; m.guide_batch = bo->guide_batch;
            m.guide_steps = bo->guide_steps;
            m.guide_depth = bo->guide_depth;
            c->g_scene = c->scenes[in[0].scene].box;
        }
        std::vector<double> done(n, 0.0);
        const phg_status st = render_passes(c, jobs, passes, c->bfilm.p, m, &done);
        if (st != PHG_OK) return st;
        c->stats.batch_passes = (uint32_t)passes.size();
        if (results)
            for (uint32_t j = 0; j < n; ++j) {
                results[j].done_ms = done[j];
                results[j].paths = (uint64_t)jobs[j].npix * in[j].spp;
            }
        return PHG_OK;
    });
}

phg_status phg_batch_film_read(void* ctx, uint32_t job, double* out, uint64_t out_len) {
    return guarded(ctx, [&]() -> phg_status {
        Ctx* c = static_cast<Ctx*>(ctx);
        if (!c || !out) return PHG_ERR_INVALID_ARG;
        if (job >= c->bfilm_jobs.size()) return set_err(c, PHG_ERR_INVALID_ARG, "no such job in the last batch");
        const JobDev& j = c->bfilm_jobs[job];
        const uint64_t n = 8ull * j.npix;
        if (out_len < n) return set_err(c, PHG_ERR_INVALID_ARG, "output buffer too small");
        PHG_HIP(c, hipSetDevice(c->device));
        PHG_HIP(c, hipMemcpyAsync(out, c->bfilm.p + j.film_base, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        PHG_HIP(c, hipStreamSynchronize(c->stream));
        return PHG_OK;
    });
}

phg_status phg_guide_set_params(void* ctx, const float* params, uint32_t n) {
    return guarded(ctx, [&]() -> phg_status {
        Ctx* c = static_cast<Ctx*>(ctx);
        if (!c || !params) return PHG_ERR_INVALID_ARG;
        if (n != kGuideParamCount) return set_err(c, PHG_ERR_INVALID_ARG, "expected PHG_GUIDE_PARAMS values");
        PHG_HIP(c, hipSetDevice(c->device));
        phg_status st = guide_load(c, params);
        if (st == PHG_OK) c->g_ready = true;
        return st;
    });
}

phg_status phg_guide_get_params(void* ctx, float* params, uint32_t n) {
    return guarded(ctx, [&]() -> phg_status {
        Ctx* c = static_cast<Ctx*>(ctx);
        if (!c || !params) return PHG_ERR_INVALID_ARG;
        if (n != kGuideParamCount) return set_err(c, PHG_ERR_INVALID_ARG, "expected PHG_GUIDE_PARAMS values");
        if (!c->g_init) return set_err(c, PHG_ERR_INVALID_ARG, "guiding network not initialised");
        PHG_HIP(c, hipSetDevice(c->device));
        PHG_HIP(c, hipMemcpyAsync(params, c->g_params.p, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        PHG_HIP(c, hipStreamSynchronize(c->stream));
        return PHG_OK;
    });
}

static phg_status guide_infer_host(void* ctx, const float* features, uint64_t n, float* out, int raw) {
    return guarded(ctx, [&]() -> phg_status {
        Ctx* c = static_cast<Ctx*>(ctx);
        if (!c || !features || !out || n >= (1ull << 31)) return PHG_ERR_INVALID_ARG;
        if (!c->g_init) return set_err(c, PHG_ERR_INVALID_ARG, "guiding network not initialised");
        if (n == 0) return PHG_OK;
        PHG_HIP(c, hipSetDevice(c->device));
        PHG_HIP(c, c->g_q.ensure(n * (9 + kGuideOutputs)));
        float* d_in = c->g_q.p;
        float* d_out = c->g_q.p + n * 9;
        PHG_HIP(c, hipMemcpyAsync(d_in, features, n * 9 * sizeof(float), hipMemcpyHostToDevice, c->stream));
        PHG_HIP(c, guide_infer_features(guide_net(c), d_in, (uint32_t)n, d_out, raw, c->cu_count * 2, c->stream));
        PHG_HIP(c, hipMemcpyAsync(out, d_out, n * kGuideOutputs * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        PHG_HIP(c, hipStreamSynchronize(c->stream));
        return PHG_OK;
    });
}

phg_status phg_guide_forward(void* ctx, const float* features, uint64_t n, float* out) {
    return guide_infer_host(ctx, features, n, out, 1);
}

phg_status phg_guide_lobes(void* ctx, const float* features, uint64_t n, float* out) {
    return guide_infer_host(ctx, features, n, out, 0);
}

phg_status phg_guide_gradients(void* ctx, const float* records, uint64_t n, float* grads) {
    return guarded(ctx, [&]() -> phg_status {
        Ctx* c = static_cast<Ctx*>(ctx);
        if (!c || !records || !grads || n == 0 || n >= (1ull << 31)) return PHG_ERR_INVALID_ARG;
        if (!c->g_init) return set_err(c, PHG_ERR_INVALID_ARG, "guiding network not initialised");
        PHG_HIP(c, hipSetDevice(c->device));
        PHG_HIP(c, c->g_q.ensure(n * kGuideRecordFloats));
        PHG_HIP(c, hipMemcpyAsync(c->g_q.p, records, n * kGuideRecordFloats * sizeof(float), hipMemcpyHostToDevice, c->stream));
        PHG_HIP(c, guide_fill(c->g_grads.p, kGuideParamCount, 0.0f, c->stream));
        GuideWorkspace ws{};
        {
            const phg_status st = guide_workspace(c, (uint32_t)n, &ws);
            if (st != PHG_OK) return st;
        }
        PHG_HIP(c, guide_gradients(guide_net(c), c->g_q.p, (uint32_t)n, c->g_grads.p, ws, c->stream));
        PHG_HIP(c, hipMemcpyAsync(grads, c->g_grads.p, kGuideParamCount * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        PHG_HIP(c, guide_fill(c->g_grads.p, kGuideParamCount, 0.0f, c->stream));
        PHG_HIP(c, hipStreamSynchronize(c->stream));
        return PHG_OK;
    });
}

phg_status phg_guide_train(void* ctx, const float* records, uint64_t n, float lr, uint32_t steps) {
    return guarded(ctx, [&]() -> phg_status {
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
    });
}

phg_status phg_set_tuning(void* ctx, const phg_tuning* t) {
    return guarded(ctx, [&]() -> phg_status {
        Ctx* c = static_cast<Ctx*>(ctx);
        if (!c || !t) return PHG_ERR_INVALID_ARG;
        const uint32_t block = t->trace_block ? t->trace_block : 128;
        if (block != 64 && block != 128 && block != 256) return set_err(c, PHG_ERR_INVALID_ARG, "trace_block must be 64, 128 or 256");
        if (t->paths_per_pass != 0 && (t->paths_per_pass < (1u << 16) || t->paths_per_pass > (1u << 26)))
            return set_err(c, PHG_ERR_INVALID_ARG, "paths_per_pass must be in [2^16, 2^26]");
        PHG_HIP(c, hipSetDevice(c->device));
        PHG_HIP(c, hipStreamSynchronize(c->stream));
        uint32_t tb = 0, sb = 0;
        PHG_HIP(c, resident_blocks(block, &tb, &sb));
        c->trace_block = block;
        c->trace_blocks_per_cu = tb;
        c->shade_blocks_per_cu = sb;
        c->paths_per_pass = t->paths_per_pass ? t->paths_per_pass : kDefaultPassPaths;
        return PHG_OK;
    });
}

phg_status phg_get_tuning(void* ctx, phg_tuning* t) {
    return guarded(ctx, [&]() -> phg_status {
        Ctx* c = static_cast<Ctx*>(ctx);
        if (!c || !t) return PHG_ERR_INVALID_ARG;
        std::memset(t, 0, sizeof(*t));
        t->trace_block = c->trace_block;
        t->paths_per_pass = c->paths_per_pass;
        t->trace_blocks_per_cu = c->trace_blocks_per_cu;
        return PHG_OK;
    });
}

phg_status phg_autotune_batch(void* ctx, const phg_scene_desc* scenes, uint32_t nscenes, const phg_batch_job* jobs,
                              uint32_t njobs, const phg_autotune_opts* ao, phg_tune_trial* best, phg_tune_trial* trials,
                              uint32_t max_trials, uint32_t* ntrials) {
    return guarded(ctx, [&]() -> phg_status {
        Ctx* c = static_cast<Ctx*>(ctx);
        if (!c || !scenes || nscenes == 0 || !jobs || njobs == 0 || !best) return PHG_ERR_INVALID_ARG;
        for (uint32_t j = 0; j < njobs; ++j)
            if (jobs[j].scene >= nscenes) return set_err(c, PHG_ERR_INVALID_ARG, "job scene index out of range");
        const uint32_t want = ao && ao->sample_jobs ? ao->sample_jobs : 8;
        const uint32_t sspp = ao && ao->sample_spp ? ao->sample_spp : 4;
        const uint32_t reps = ao && ao->repeats ? ao->repeats : 2;
        // sample: jobs spread evenly over the work-size order, with their scenes renumbered
        std::vector<uint32_t> order(njobs);
        for (uint32_t j = 0; j < njobs; ++j) order[j] = j;
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
            return (uint64_t)jobs[a].width * jobs[a].height * jobs[a].spp < (uint64_t)jobs[b].width * jobs[b].height * jobs[b].spp;
        });
        const uint32_t ns = std::min(want, njobs);
        std::vector<phg_batch_job> sjobs;
        std::vector<phg_scene_desc> sscenes;
        std::vector<int64_t> remap(nscenes, -1);
        for (uint32_t k = 0; k < ns; ++k) {
            const uint32_t j = order[ns > 1 ? (uint64_t)k * (njobs - 1) / (ns - 1) : 0];
            phg_batch_job b = jobs[j];
            if (remap[b.scene] < 0) { remap[b.scene] = (int64_t)sscenes.size(); sscenes.push_back(scenes[b.scene]); }
            b.scene = (uint32_t)remap[b.scene];
            b.spp = std::min(b.spp, sspp);
            sjobs.push_back(b);
        }
        uint64_t paths = 0;
        for (const phg_batch_job& b : sjobs) paths += (uint64_t)b.width * b.height * b.spp;
        std::vector<phg_tune_trial> all;
        auto measure = [&](phg_tune_trial& t) -> phg_status {
            phg_tuning tu{};
            tu.trace_block = t.trace_block;
            tu.paths_per_pass = t.paths_per_pass;
            phg_status st = phg_set_tuning(c, &tu);
            if (st != PHG_OK) return st;
            phg_batch_opts bo{};
            bo.flags = PHG_BATCH_GRAPH;
            double best_ms = INFINITY;
            for (uint32_t r = 0; r <= reps; ++r) {  // the first render warms up
                const auto t0 = std::chrono::steady_clock::now();
                st = phg_render_batch(c, sjobs.data(), (uint32_t)sjobs.size(), &bo, nullptr);
                if (st != PHG_OK) return st;
                const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                if (r > 0) best_ms = std::min(best_ms, ms);
            }
            t.msamples_per_s = (double)paths / (best_ms * 1e3);
            all.push_back(t);
            return PHG_OK;
        };
        auto upload = [&](uint32_t leaf, uint32_t width) -> phg_status {
            phg_bvh_opts bo{};
            bo.max_leaf_tris = leaf;
            bo.width = width;
            return phg_upload_scenes(c, sscenes.data(), (uint32_t)sscenes.size(), &bo);
        };
        phg_tune_trial cur{4, 4, 128, kDefaultPassPaths, 0.0};
        phg_status st = upload(cur.max_leaf_tris, cur.width);
        if (st != PHG_OK) return st;
        // stage 1: launch parameters on the default BVH
        phg_tune_trial top{};
        top.msamples_per_s = -1.0;
        for (uint32_t block : {64u, 128u, 256u})
            for (uint32_t ppp : {2u << 20, 4u << 20, 8u << 20, 16u << 20, 32u << 20, 64u << 20}) {
                phg_tune_trial t = cur;
                t.trace_block = block;
                t.paths_per_pass = ppp;
                st = measure(t);
                if (st != PHG_OK) return st;
                if (t.msamples_per_s > top.msamples_per_s) top = t;
            }
        // stage 2: BVH shape with the best launch parameters
        for (uint32_t leaf : {2u, 4u, 8u})
            for (uint32_t width : {2u, 4u}) {
                if (leaf == 4 && width == 4) continue;  // measured in stage 1
                st = upload(leaf, width);
                if (st != PHG_OK) return st;
                phg_tune_trial t = top;
                t.max_leaf_tris = leaf;
                t.width = width;
                st = measure(t);
                if (st != PHG_OK) return st;
                if (t.msamples_per_s > top.msamples_per_s) top = t;
            }
        *best = top;
        phg_tuning tu{};
        tu.trace_block = top.trace_block;
        tu.paths_per_pass = top.paths_per_pass;
        st = phg_set_tuning(c, &tu);
        if (st != PHG_OK) return st;
        if (trials)
            for (uint32_t i = 0; i < std::min<uint32_t>(max_trials, (uint32_t)all.size()); ++i) trials[i] = all[i];
        if (ntrials) *ntrials = (uint32_t)all.size();
        return PHG_OK;
    });
}

phg_status phg_get_stats(void* ctx, phg_stats* out) {
    return guarded(ctx, [&]() -> phg_status {
        Ctx* c = static_cast<Ctx*>(ctx);
        if (!c || !out) return PHG_ERR_INVALID_ARG;
        *out = c->stats;
        return PHG_OK;
    });
}

}  // extern "C"
