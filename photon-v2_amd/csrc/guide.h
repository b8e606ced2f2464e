// Host-visible declarations of the guiding-MLP kernels.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "kernels.h"

namespace phg {

struct GuideScene {
    float lo[3];
    float inv_ext[3];
};

// bf16 arrays are carried as uint16_t bit patterns on the host side
struct GuideNetDev {
    const float* params;     // fp32 master parameters (kGuideParamCount)
    const uint16_t* wt;      // 5 x [64 out][64 in] bf16
    const uint16_t* wn;      // 5 x [64 in][64 out] bf16
    const float* bias;       // 5 x 64 fp32
    float* params_mut;
    uint16_t* wt_mut;
    uint16_t* wn_mut;
    float* bias_mut;
    const uint16_t* wp;      // 5 x 4096 bf16 in the matrix-core operand order of inference
    const float* bp;         // 5 x 64 fp32 biases in the same output order
    uint16_t* wp_mut;
    float* bp_mut;
};

constexpr uint32_t kGuideParamCount = 19305;
// record: 9 features, sampled direction (3), weight, BSDF density, selection probability used,
// and the sampling density of the direction if that probability was the network's own (else 0)
constexpr uint32_t kGuideRecordFloats = 16;
constexpr uint32_t kGuideOutputs = 41;      // raw outputs: 8 lobes x 5, selection logit
constexpr uint32_t kGuideNetLayers = 5;     // dense layers of the network (scratch sizing)

hipError_t guide_prepare(const GuideNetDev& net, hipStream_t st);
hipError_t guide_infer_features(const GuideNetDev& net, const float* features, uint32_t n, float* out, int raw,
                                uint32_t grid, hipStream_t st);
// Lobes of the hits queued at bounce bounce_base[0] + bk (the entries of that bounce's
// extension-queue regions, whose fill counters are the kPushExt counters in `regions`; region
// size bounce_base[1]), each written at its queue slot; adds the number of queries to *queries.
hipError_t guide_infer_hits(const GuideNetDev& net, const float4* hits, const phg_ray* rays, const TriRec* tris,
                            const GuideScene& gs, const uint32_t* regions, const uint32_t* bounce_base, uint32_t bk,
                            float* out, unsigned long long* queries, uint32_t grid, hipStream_t st);
// Scratch of a training step (device memory): layer inputs X (5 x cap x 64 bf16), two dZ
// buffers (cap x 64 bf16 each), per-block dW / db partials (5 x (max_parts + 4) x 4160 floats)
// and their first-level sums (5 x slices x 4160 floats).
struct GuideWorkspace {
    uint16_t* X;
    uint16_t* dz0;
    uint16_t* dz1;
    float* partial;
    float* partial2;
    uint32_t cap_records;    // records per step it holds (multiple of 32)
    uint32_t max_parts;      // waves of the backward kernels
    uint32_t max_blocks;     // blocks of the forward kernel
    uint32_t slices;         // first-level reduction slices
};
// Gradient of the mean record loss over nrec records, written to grads (kGuideParamCount floats).
hipError_t guide_gradients(const GuideNetDev& net, const float* rec, uint32_t nrec, float* grads,
                           const GuideWorkspace& ws, hipStream_t st);
hipError_t guide_adam(const GuideNetDev& net, float* grads, float* m, float* v, float lr, int step, hipStream_t st);
hipError_t guide_fill(float* p, uint32_t n, float v, hipStream_t st);
// Training records of a finished pass (incident radiance along each recorded sampled
// direction), in path order: counters[0] = records kept (at most `capacity` are written),
// counters[1] = records valid before the keep test; block_counts holds 2 x
// guide_collect_blocks(paths) entries of scratch.
uint32_t guide_collect_blocks(uint32_t paths);
hipError_t guide_collect(const PathBuffers& pb, const PassParams& pp, const GuideScene& gs, float* out_rec,
                         uint32_t capacity, uint32_t* counters, uint32_t* block_counts, float keep_prob,
                         uint32_t salt, hipStream_t st);

}  // namespace phg
