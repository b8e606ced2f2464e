// Host-visible declarations of the device scene, wavefront buffers and kernel launchers.
#pragma once

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>
#include <stdint.h>

#include "../../include/phg.h"
#include "bvh_build.h"

namespace phg {

constexpr uint32_t kTraceBlock = 128;   // ray-query kernel block (2 waves); path kernels take 64, 128 or 256
constexpr int kTraceRing = 16;          // LDS ring entries per lane; older entries spill to global memory
constexpr uint32_t kShadeBlock = 256;
constexpr uint32_t kCountSlots = 64;    // queue counters, used modulo this many bounces

struct alignas(16) TriRec {       // 48 bytes: one triangle in BVH leaf order
    float a[3], b[3], c[3];
    int32_t prim;                 // input triangle id
    uint32_t mat;
    int32_t emitter;              // -1 if not emissive
};
static_assert(sizeof(TriRec) == 48, "triangle record is 48 bytes");

struct alignas(16) EmitterRec {   // 32 bytes
    float radiance[3];
    float pick_pdf;               // ESPowerFavoring pdfDiscrete
    uint32_t tri_begin, tri_count;
    uint32_t pad[2];
};

struct alignas(8) EmitTri {       // one emitter triangle: its BVH-order index and area
    uint32_t tri;
    float area;
};

// One render job: a camera over one of the uploaded scenes, its film and its sample stream.
// The scenes share the node / triangle / material / emitter arrays (indices rebased at
// upload); a job addresses its own part through these offsets.
struct alignas(16) JobDev {
    phg_camera cam;
    uint32_t root;                // first BVH node of the job's scene
    uint32_t width, height, npix;
    uint32_t emitter_begin;       // the scene's emitters: [emitter_begin, emitter_begin + num_emitters)
    uint32_t num_emitters;
    uint32_t cdf_begin;           // the scene's emitter-pick CDF: num_emitters + 1 floats
    uint32_t max_bounces;
    uint64_t seed;                // sample-stream seed (< 2^24)
    uint64_t film_base;           // first double of the job's film (8 per pixel)
    uint32_t estimator;           // PHG_ESTIMATOR_BNEEPT, or BVPT (BVPTDL runs as BVPT capped at one bounce)
    uint32_t filter;              // PHG_FILTER_*: reconstruction filter of the film
    uint32_t sampler;             // PHG_SAMPLER_* (Halton: with its permutation and sequence bits)
    uint32_t halton_leap;         // Halton: index step of the sequence (1 but for the leap sequences)
};
static_assert(sizeof(JobDev) == 128, "job record is 128 bytes");

// A pass renders whole samples of one or more jobs: segment s holds `samples` consecutive
// samples of job `job`, starting at sample index sample_begin, as pass paths
// [path_offset, path_offset + samples * npix) in sample-major order; its pixels are film
// threads [pix_offset, pix_offset + npix).
struct alignas(16) PassSeg {
    uint32_t job, sample_begin, samples, path_offset;
    uint32_t pix_offset, npix, pad[2];
};
constexpr uint32_t kMaxSegs = 4096;

struct SceneDev {
    const WideNode* nodes;
    const TriRec* tris;
    const phg_material* mats;
    const EmitterRec* emitters;
    const float* emit_cdf;
    const EmitTri* emit_tris;
    const JobDev* jobs;
};

struct PassParams {
    uint32_t paths;          // paths in this pass (sum over the segments)
    uint32_t num_segs;
    uint32_t film_threads;   // sum of the segments' pixels
    uint32_t max_bounces;    // largest over the pass's jobs (bounce-loop length)
    float guide_prob;        // guided shading: probability of sampling the learned lobes
    uint32_t record;         // guided shading: write training records
    uint32_t rec_bounces;    // bounces with a record slot
    uint32_t guide_depth;    // guided shading at bounces < guide_depth only (0 = every bounce)
    const PassSeg* segs;     // num_segs entries, ascending path_offset and pix_offset
};

constexpr uint32_t kGuideRecBounces = 8;   // training records are kept for the first bounces
constexpr uint32_t kRecFloat4 = 6;         // P+pdf, N, V, omega, Wem, L before

// Structure-of-arrays path state and queues; all device pointers.
struct PathBuffers {
    uint64_t* rng;           // PCG32 state per path
    float4* T;               // throughput rgb, rrScale
    float4* Wem;             // previous BSDF weight x throughput rgb, previous BSDF pdf
    float4* prevP;           // previous vertex position, flags (bits)
    float4* L;               // accumulated radiance
    phg_ray* ext[2];         // extension-ray queues (ping-pong)
    uint32_t* ext_pid[2];
    float4* hits;            // per queued extension ray: tri (bits), bary a, b, c
    phg_ray* shadow;         // shadow-ray queue
    uint32_t* shadow_pid;
    float4* shadow_c;        // unoccluded contribution
    uint32_t* counts;        // [0, kCountSlots): entries of each bounce's extension queue (k_generate,
                             // then the shadow kernel of the bounce before sums its regions)
    unsigned long long* traversal_counters;  // node visits / tri tests: ext, ext, shadow, shadow
    uint32_t* error_flags;
    double* film;            // 8 doubles per pixel
    // guided shading
    const float* lobes;      // per queued extension ray: 8 x (mu xyz, kappa, weight)
    float4* rec;             // training records [bounce][path][kRecFloat4]
    int32_t* spill;          // traversal-stack spill, spill_cap entries per traversal thread
    uint32_t spill_cap;
    uint32_t* path_seg;      // per path: its pass segment
    uint32_t* bounce_base;   // [0]: first bounce of the running chunk of bounces (the stage
                             // kernels take their bounce as bounce_base[0] + their chunk
                             // offset); [1]: the queue region size of the running pass
    uint32_t* fetch;         // work counters of the stage kernels (fetch_slot), kFetchWords
};

// Region counters, 8 per queue pass and bounce, in a ring of kCountSlots bounces, each on a
// 128-byte line of its own (kFetchStride words apart): the work counters of the dynamic work
// distribution of the stage kernels (extension-ray traversal, shadow-ray traversal, shading)
// and the fill counts of the regions of the extension- and shadow-ray queues (kPushExt,
// kPushShadow).
//
// A queue of a pass is 8 regions of queue_region_cap(paths) slots; region r is
// [r cap, r cap + count_r).  The shading kernel pushes the rays of a vertex taken from region r
// of its input queue into region r of the output queues, so no region receives more entries
// than its input region held (at most one ray per queue and vertex), and the camera rays
// fill the regions in path order.  Device-scope atomics on one address serialize (about 10 ns
// each on MI355X), so one counter per queue limited the shading kernel to one 64-vertex push
// per 10 ns; 8 counters take 8 at once, and each XCD's waves drain and fill mostly their own.
enum : uint32_t { kFetchExtend = 0, kFetchShadow = 1, kFetchShade = 2, kPushExt = 3, kPushShadow = 4, kFetchKinds = 5 };
constexpr uint32_t kFetchStride = 32;
constexpr uint32_t kFetchWords = kFetchKinds * kCountSlots * 8u * kFetchStride;
__host__ __device__ inline uint32_t fetch_slot(uint32_t kind, uint32_t bounce) {
    return (kind * kCountSlots + bounce % kCountSlots) * 8u * kFetchStride;
}
__host__ __device__ inline uint32_t queue_region_cap(uint32_t paths) { return ((paths + 7u) / 8u + 63u) & ~63u; }
// slots a queue of a pass of `paths` paths spans (its 8 regions)
__host__ __device__ inline uint64_t queue_capacity(uint32_t paths) { return 8ull * queue_region_cap(paths); }

// Dense position d in [0, n) of a queue's entries -> its slot (regions in order).
struct QueueMap {
    uint32_t cap = 0, n = 0;
    uint32_t pre[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // entries before region r
    bool identity = true;                        // a plain array (no regions)
    // the fill counts of the 8 regions (kPushExt / kPushShadow counters of one bounce)
    __host__ __device__ void load(const uint32_t* region_counts, uint32_t region_cap) {
        cap = region_cap;
        identity = false;
        uint32_t run = 0;
        for (int r = 0; r < 8; ++r) {
            pre[r] = run;
            run += region_counts[r * kFetchStride];
        }
        n = run;
    }
    __host__ __device__ uint32_t slot(uint32_t d) const {
        if (identity) return d;
        uint32_t r = 0, p = 0;
        for (int k = 1; k < 8; ++k)
            if (d >= pre[k]) { r = (uint32_t)k; p = pre[k]; }
        return r * cap + (d - p);
    }
};

// Size of the device scratch array of the allocation-free ray query; its address on `device`
// (looked up on the first call per device, which makes that device current).
constexpr size_t kNoAllocScratchBytes = 16u << 20;
constexpr size_t kNoAllocFlagBytes = 256;
hipError_t noalloc_scratch_address(int device, void** out);

hipError_t launch_intersect(const SceneDev& sc, const phg_ray* rays, uint64_t n, phg_hit* hits, int any,
                            uint32_t* nvis, uint32_t* ntri, uint32_t* flags, int32_t* spill, uint32_t spill_cap,
                            uint32_t grid, hipStream_t st);
// Wavefront stages.  Pass parameters are read from device memory (ppd), so that a captured
// bounce loop can be replayed for every pass after the host rewrites them.
hipError_t launch_generate(const SceneDev& sc, const PassParams* ppd, const PathBuffers& pb, uint32_t grid, hipStream_t st);
// Blocks of the path traversal kernels (trace_block threads: 64, 128 or 256) and of the
// shading kernel (kShadeBlock threads) that one compute unit holds at once; sizes the
// persistent grids.
hipError_t resident_blocks(uint32_t trace_block, uint32_t* trace_blocks_per_cu, uint32_t* shade_blocks_per_cu);
// The stage kernels of bounce bounce_base[0] + bk (bk: offset within the running chunk; eager
// loops keep bounce_base[0] = 0 and pass the bounce itself).
hipError_t launch_extend(const SceneDev& sc, const PathBuffers& pb, uint32_t bk, uint32_t grid, uint32_t block,
                         uint32_t count, hipStream_t st);
hipError_t launch_shadow(const SceneDev& sc, const PassParams* ppd, const PathBuffers& pb, uint32_t bk, uint32_t grid,
                         uint32_t block, uint32_t count, hipStream_t st);
hipError_t launch_shade(const SceneDev& sc, const PassParams* ppd, const PathBuffers& pb, uint32_t bk, uint32_t grid,
                        bool guided, hipStream_t st);
hipError_t launch_film(const SceneDev& sc, const PassParams* ppd, const PathBuffers& pb, uint32_t grid, hipStream_t st);
hipError_t launch_clear_counts(uint32_t* counts, uint32_t n, hipStream_t st);
// bounce_base[0] += chunk: ends a captured chunk of bounces
hipError_t launch_advance_bounce(uint32_t* bounce_base, uint32_t chunk, hipStream_t st);
hipError_t launch_fill_zero(double* p, uint64_t n, hipStream_t st);

}  // namespace phg
