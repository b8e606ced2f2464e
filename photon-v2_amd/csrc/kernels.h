// Host-visible declarations of the device scene, wavefront buffers and kernel launchers.
#pragma once

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>
#include <stdint.h>

#include "../../include/phg.h"
#include "bvh_build.h"

namespace phg {

constexpr uint32_t kTraceBlock = 128;   // traversal kernels: 2 waves, LDS stack ring [slot][128]
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

struct SceneDev {
    const WideNode* nodes;
    const TriRec* tris;
    const phg_material* mats;
    const EmitterRec* emitters;
    const float* emit_cdf;        // num_emitters + 1
    const EmitTri* emit_tris;
    uint32_t num_emitters;
    uint32_t pad;
    phg_camera cam;
};

struct PassParams {
    uint32_t width, height;
    uint32_t npix;
    uint32_t paths;          // paths in this pass = npix * samples in the pass
    uint32_t sample_begin;   // sample index of path 0
    uint32_t max_bounces;
    uint64_t seed;
    float guide_prob;        // guided shading: probability of sampling the learned lobes
    uint32_t record;         // guided shading: write training records
    uint32_t rec_bounces;    // bounces with a record slot
    uint32_t pad;
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
    uint32_t* counts;        // [0, kCountSlots) extension counts, [kCountSlots, 2*kCountSlots) shadow counts
    unsigned long long* traversal_counters;  // node visits / tri tests: ext, ext, shadow, shadow
    uint32_t* error_flags;
    double* film;            // 8 doubles per pixel
    // guided shading
    const float* lobes;      // per queued extension ray: 8 x (mu xyz, kappa, weight)
    float4* rec;             // training records [bounce][path][kRecFloat4]
    int32_t* spill;          // traversal-stack spill, spill_cap entries per traversal thread
    uint32_t spill_cap;
};

hipError_t launch_intersect(const SceneDev& sc, const phg_ray* rays, uint64_t n, phg_hit* hits, int any,
                            uint32_t* nvis, uint32_t* ntri, uint32_t* flags, int32_t* spill, uint32_t spill_cap,
                            uint32_t grid, hipStream_t st);
hipError_t launch_generate(const SceneDev& sc, const PassParams& pp, const PathBuffers& pb, hipStream_t st);
// Blocks of the traversal kernels (kTraceBlock threads) and of the shading kernel
// (kShadeBlock threads) that one compute unit holds at once; sizes the persistent grids.
hipError_t resident_blocks(uint32_t* trace_blocks_per_cu, uint32_t* shade_blocks_per_cu);
hipError_t launch_extend(const SceneDev& sc, const PathBuffers& pb, uint32_t bounce, uint32_t grid, uint32_t count,
                         hipStream_t st);
hipError_t launch_shadow(const SceneDev& sc, const PathBuffers& pb, uint32_t bounce, uint32_t grid, uint32_t count,
                         uint32_t rec_bounces, uint32_t npaths, hipStream_t st);
hipError_t launch_shade(const SceneDev& sc, const PassParams& pp, const PathBuffers& pb, uint32_t bounce, uint32_t grid,
                        bool guided, hipStream_t st);
hipError_t launch_film(const PassParams& pp, const PathBuffers& pb, uint32_t spp_in_pass, hipStream_t st);
hipError_t launch_clear_counts(uint32_t* counts, uint32_t n, hipStream_t st);
hipError_t launch_fill_zero(double* p, uint64_t n, hipStream_t st);

}  // namespace phg
