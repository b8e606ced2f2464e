// TEST INFRASTRUCTURE ONLY: runs the product's wide-BVH build and its traversal code
// (photon-v2_amd/csrc/traverse.h, compiled for the host) on the CPU, so BVH conservativeness
// and traversal semantics can be checked against the oracle without a GPU.  Never loaded by
// the product.
#include <cstring>
#include <vector>

#include "../../photon-v2_amd/csrc/traverse.h"

using namespace phg;

extern "C" int emu_intersect(const float* verts, uint32_t ntri, const float* rays, uint32_t n, int any,
                             uint32_t max_leaf, uint32_t bins, uint32_t width, int32_t* out_prim, float* out_t, float* out_b,
                             float* out_c, uint32_t* out_nvis, uint32_t* out_ntri) {
    BvhBuildParams p;
    p.max_leaf_tris = max_leaf ? max_leaf : 4;
    p.width = width ? width : 4;
    p.sah_bins = bins ? bins : 16;
    BvhBuildResult bvh = build_wide_bvh(verts, ntri, p);
    const uint32_t spill_cap = bvh.stack_need > (uint32_t)kTraceRing ? bvh.stack_need - (uint32_t)kTraceRing : 0u;
    std::vector<int32_t> spill(spill_cap + 1);
    std::vector<TriRec> tr(ntri);
    for (uint32_t k = 0; k < ntri; ++k) {
        const uint32_t t = bvh.tri_order[k];
        std::memcpy(tr[k].a, verts + 9 * (size_t)t, 9 * sizeof(float));
        tr[k].prim = (int32_t)t; tr[k].mat = 0; tr[k].emitter = -1;
    }
    int32_t stack[kTraceRing];
    uint32_t overflow = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * (size_t)i;
        uint32_t nv = 0, nt = 0;
        TraceResult res = any ? traverse<kTraceRing, true, true>(bvh.nodes.data(), tr.data(), 0, mk(r[0], r[1], r[2]),
                                                                mk(r[4], r[5], r[6]), r[3], r[7], stack, 1, spill.data(),
                                                                spill_cap, &nv, &nt, &overflow)
                              : traverse<kTraceRing, false, true>(bvh.nodes.data(), tr.data(), 0, mk(r[0], r[1], r[2]),
                                                                 mk(r[4], r[5], r[6]), r[3], r[7], stack, 1, spill.data(),
                                                                 spill_cap, &nv, &nt, &overflow);
        out_prim[i] = res.tri >= 0 ? tr[res.tri].prim : -1;
        out_t[i] = res.t; out_b[i] = res.b; out_c[i] = res.c;
        out_nvis[i] = nv; out_ntri[i] = nt;
    }
    return overflow ? -2 : (int)bvh.stack_need;
}

// ---- guiding-model device math, compiled for the host
#include "../../photon-v2_amd/csrc/guide_device.h"

extern "C" void emu_guide_output_grad(const float* o, uint32_t n, const float* w3, const float* weight,
                                      const float* bsdf_pdf, const float* sel, const float* samp_pdf, float* out) {
    for (uint32_t i = 0; i < n; ++i)
        guide_output_grad(o + kGuideOut * (size_t)i, mk(w3[3 * i], w3[3 * i + 1], w3[3 * i + 2]), weight[i], bsdf_pdf[i],
                          sel[i], samp_pdf[i], out + kGuideOut * (size_t)i);
}

extern "C" void emu_guide_sel(const float* o, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = guide_sel(o[kGuideOut * (size_t)i + kGuideSel]);
}

extern "C" void emu_guide_pdf(const float* o, uint32_t n, const float* w3, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        float lobes[kGuideLobeVals];
        guide_decode(o + kGuideOut * (size_t)i, lobes);
        out[i] = guide_pdf(lobes, mk(w3[3 * i], w3[3 * i + 1], w3[3 * i + 2]));
    }
}

extern "C" void emu_guide_sample(const float* o40, uint32_t n, const float* u3, float* out3) {
    float lobes[kGuideLobeVals];
    guide_decode(o40, lobes);
    for (uint32_t i = 0; i < n; ++i) {
        const V3 d = guide_sample(lobes, u3[3 * i], u3[3 * i + 1], u3[3 * i + 2]);
        out3[3 * i] = d.x; out3[3 * i + 1] = d.y; out3[3 * i + 2] = d.z;
    }
}

extern "C" void emu_guide_encode(const float* x9, uint32_t n, float* out64) {
    for (uint32_t i = 0; i < n; ++i)
        for (int j = 0; j < 64; ++j) out64[64 * (size_t)i + j] = guide_encode(x9 + 9 * (size_t)i, j);
}

#include "../../photon-v2_amd/csrc/pixel_order.h"

// Pixel of every path slot of a W x H job (the product's slot order), and the slot
// tile_slot gives back for each pixel.
extern "C" void emu_pixel_order(uint32_t W, uint32_t H, uint32_t* pixel_of_slot, uint32_t* slot_of_pixel) {
    for (uint32_t q = 0; q < W * H; ++q) pixel_of_slot[q] = phg::tile_pixel(q, W, H);
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x) slot_of_pixel[y * W + x] = phg::tile_slot(x, y, W, H);
}

// The product's Halton source for the host: dimension dims[i] of point index[i] of a job with
// the given sampler value, leap and seed (pt_device.h halton_value).
extern "C" void emu_halton_values(const uint32_t* dims, const uint64_t* index, uint32_t n, uint32_t sampler,
                                  uint32_t leap, uint64_t seed, float* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = halton_value(dims[i], index[i], sampler, leap, seed);
}

// The product's Halton raster positions of batch `sample` of a W x H job (filter: PHG_FILTER_*)
extern "C" void emu_halton_raster(uint32_t W, uint32_t H, uint32_t sample, uint32_t filter, uint32_t sampler,
                                  uint32_t leap, uint64_t seed, double* xy) {
    JobDev job{};
    job.width = W; job.height = H; job.npix = W * H;
    job.filter = filter; job.sampler = sampler; job.halton_leap = leap; job.seed = seed;
    for (uint32_t p = 0; p < W * H; ++p) halton_raster(job, p, sample, xy + 2 * (size_t)p, xy + 2 * (size_t)p + 1);
}
