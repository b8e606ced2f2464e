// gfx950 kernels of the BNEEPT path: 4-wide BVH traversal (closest and any hit), and the
// wavefront stages generate -> extend -> shade -> shadow -> film.
//
// Traversal: one ray per lane, 128-byte node fetches (8 x 16-byte loads), child boxes
// tested against the shrinking segment with the reference's robust slab test, a per-lane
// traversal stack in LDS laid out [depth][lane] (bank-conflict free), near-first order.
// Queues between stages are compacted with a wave64 ballot + popcount prefix and one
// atomic per wave.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "kernels.h"
#include "pt_device.h"
#include "traverse.h"
#include "pixel_order.h"
#include "guide_device.h"

namespace phg {

// ---------------------------------------------------------------------------------------
// launch geometry helpers
// ---------------------------------------------------------------------------------------
// Blocks are handed to the 8 XCDs round robin.  Remap so that consecutive logical blocks
// (coherent rays) run on the same XCD and share its L2.
__device__ __forceinline__ uint32_t xcd_block_id() {
    const uint32_t b = blockIdx.x, n = gridDim.x;
    if ((n & 7u) != 0u) return b;
    return (b & 7u) * (n >> 3) + (b >> 3);
}

// Queue pushes of a wave: each lane with pred pushes one entry into region reg (its own) of a
// queue whose region fill counters are counters[r kFetchStride], and gets its slot
// r cap + position.  The lanes of a region take consecutive positions (ballot + popcount) and
// one atomic per region the wave pushes to (nearly always one) reserves them.
__device__ __forceinline__ uint32_t wave_push_region(uint32_t* counters, uint32_t reg, uint32_t cap, bool pred) {
    const uint32_t lane = __lane_id();
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64u - lane));
    uint64_t left = __ballot(pred);
    uint32_t slot = 0u;
    while (left != 0ull) {
        const uint32_t leader = (uint32_t)__ffsll((long long)left) - 1u;
        const uint32_t r = __builtin_amdgcn_readlane(reg, (int)leader);
        const bool mine = pred && reg == r;
        const uint64_t grp = __ballot(mine);
        uint32_t base = 0u;
        if (lane == leader) base = atomicAdd(&counters[r * kFetchStride], (uint32_t)__popcll(grp));
        base = __builtin_amdgcn_readlane(base, (int)leader);
        if (mine) slot = r * cap + base + (uint32_t)__popcll(grp & below);
        left &= ~grp;
    }
    return slot;
}

// ---------------------------------------------------------------------------------------
// 4-wide BVH traversal
// ---------------------------------------------------------------------------------------
// ---------------------------------------------------------------------------------------
// Batched ray queries (phg_intersect): rays as phg_ray, hits as phg_hit with input ids.
// ---------------------------------------------------------------------------------------
template <int RING, bool ANY, bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void k_intersect(SceneDev sc, const phg_ray* __restrict__ rays, uint64_t n,
                                                          phg_hit* __restrict__ hits, uint32_t* __restrict__ nvis_out,
                                                          uint32_t* __restrict__ ntri_out, uint32_t* __restrict__ flags,
                                                          int32_t* __restrict__ spill, uint32_t spill_cap) {
    __shared__ int32_t stack_mem[RING * kTraceBlock];
    const uint32_t gtid = xcd_block_id() * kTraceBlock + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * kTraceBlock;
    uint32_t ovf = 0;
    Traversal<RING, ANY, COUNT> tv;
    tv.attach(stack_mem + threadIdx.x, kTraceBlock, spill + (size_t)gtid * spill_cap, spill_cap);
    uint64_t i = gtid;
    uint32_t nvis = 0, ntri = 0;
    if (i < n) {
        const float4* rp = reinterpret_cast<const float4*>(rays + i);
        const float4 r0 = rp[0], r1 = rp[1];
        tv.begin(mk(r0.x, r0.y, r0.z), mk(r1.x, r1.y, r1.z), r0.w, r1.w, 0);
    }
    while (i < n) {
        if (!tv.step(sc.nodes, sc.tris, &nvis, &ntri, &ovf)) continue;
        const TraceResult tr = tv.r;
        phg_hit h;
        if (tr.tri >= 0) {
            h.t = ANY ? 0.0f : tr.t;
            h.prim = ANY ? 1 : sc.tris[tr.tri].prim;
            h.b = ANY ? 0.0f : tr.b;
            h.c = ANY ? 0.0f : tr.c;
        } else {
            h.t = 0.0f; h.prim = -1; h.b = 0.0f; h.c = 0.0f;
        }
        hits[i] = h;
        if (COUNT) { nvis_out[i] = nvis; ntri_out[i] = ntri; }
        // this lane's next ray
        nvis = 0; ntri = 0;
        i += stride;
        if (i < n) {
            const float4* rp = reinterpret_cast<const float4*>(rays + i);
            const float4 r0 = rp[0], r1 = rp[1];
            tv.begin(mk(r0.x, r0.y, r0.z), mk(r1.x, r1.y, r1.z), r0.w, r1.w, 0);
        }
    }
    if (ovf) atomicOr(flags, 1u);
}

// ---------------------------------------------------------------------------------------
// Wavefront BNEEPT
// ---------------------------------------------------------------------------------------
enum : uint32_t { kFlagDieAfterEmission = 1u };



// Passes of at most this many segments find a path's segment by searching the (cached) segment
// table instead of storing and re-reading a per-path index (three search steps at most).
constexpr uint32_t kSegSearchMax = 8;

// Pass segment holding pass path p: the last segment whose path_offset is <= p.
__device__ __forceinline__ uint32_t find_seg(const PassSeg* __restrict__ segs, uint32_t nseg, uint32_t p) {
    uint32_t lo = 0, hi = nseg;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (segs[mid].path_offset <= p) lo = mid; else hi = mid;
    }
    return lo;
}

// pixel and sample index of pass path p within its segment's job
__device__ __forceinline__ void path_ids(const PassSeg& sg, const JobDev& job, uint32_t p, uint32_t* pixel,
                                         uint32_t* sample) {
    const uint32_t local = p - sg.path_offset;
    const uint32_t s = local / sg.npix;
    *pixel = tile_pixel(local - s * sg.npix, job.width, job.height);
    *sample = sg.sample_begin + s;
}

// Surfaces where guided shading samples the lobes: Lambertian ones and GGX conductors of
// roughness >= kGuideMinRoughness.  Glossier conductors and dielectrics sample their BSDF alone
// (selection probability 0, so the one-sample MIS density is the BSDF's): lobes fitted to
// f cos L_i cannot follow such sharp lobes, and mixing them in only takes samples from the
// BSDF (measured on 8 views of the 20M-triangle terrain world: per-sample variance 1.38x
// plain with lobes at every surface, 1.04x with this rule; no change on the occluded room).
constexpr float kGuideMinRoughness = 0.25f;
__device__ __forceinline__ bool guide_samples_material(const phg_material& m) {
    return m.type == PHG_MAT_LAMBERT || (m.type == PHG_MAT_CONDUCTOR && m.roughness >= kGuideMinRoughness);
}

// Extension and shadow rays in the path queues: (origin, root node of the job's BVH as int
// bits), (direction, tmax); the segment starts at t = 0.
__device__ __forceinline__ void queue_ray(phg_ray* q, uint32_t slot, V3 o, uint32_t root, V3 d, float tmax) {
    float4* rq = reinterpret_cast<float4*>(q + slot);
    rq[0] = make_float4(o.x, o.y, o.z, __uint_as_float(root));
    rq[1] = make_float4(d.x, d.y, d.z, tmax);
}

__global__ __launch_bounds__(kShadeBlock) void k_generate(SceneDev sc, const PassParams* __restrict__ ppd, PathBuffers pb) {
    const PassParams pp = *ppd;
    const uint32_t stride = gridDim.x * kShadeBlock;
    const uint32_t cap = queue_region_cap(pp.paths);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        pb.counts[0] = pp.paths;
        pb.bounce_base[0] = 0u;
        pb.bounce_base[1] = cap;
    }
    if (blockIdx.x == 0 && threadIdx.x < 8u) {
        // path i's camera ray is slot i: the regions fill in order
        const uint32_t r = threadIdx.x, rbeg = r * cap;
        pb.fetch[fetch_slot(kFetchExtend, 0u) + r * kFetchStride] = 0u;
        pb.fetch[fetch_slot(kPushExt, 0u) + r * kFetchStride] = pp.paths <= rbeg ? 0u : (pp.paths - rbeg < cap ? pp.paths - rbeg : cap);
    }
    for (uint32_t i = blockIdx.x * kShadeBlock + threadIdx.x; i < pp.paths; i += stride) {
        const uint32_t si = find_seg(pp.segs, pp.num_segs, i);
        const PassSeg sg = pp.segs[si];
        const JobDev& job = sc.jobs[sg.job];
        uint32_t pixel, sample;
        path_ids(sg, job, i, &pixel, &sample);
        Stream s = stream_init(path_key(job.seed, pixel, sample, job.npix));
        const uint32_t px = pixel % job.width, py = pixel / job.width;
        const float ux = s.u(), uy = s.u();
        float rx = (float)px + ux, ry = (float)py + uy;
        if ((job.sampler & 0xFFu) == PHG_SAMPLER_HALTON) {
            // Halton point sample N + pixel over the whole sample window (the stream's two
            // numbers are drawn all the same, so later dimensions keep their stream positions)
            double hx, hy;
            halton_raster(job, pixel, sample, &hx, &hy);
            rx = (float)hx;
            ry = (float)hy;
        } else if (job.filter != PHG_FILTER_BOX) {
            rx = (float)window_raster(px, ux, job.width, kFilterHalf);
            ry = (float)window_raster(py, uy, job.height, kFilterHalf);
        }
        // PinholeCamera::genReceiveRayDir with the horizontal-FOV sensor of SingleLensObserver
        const phg_camera& c = job.cam;
        const float cx = (2.0f * (rx / (float)job.width) - 1.0f) * c.tan_half_fov;
        const float cy = (2.0f * (ry / (float)job.height) - 1.0f) * c.tan_half_fov * ((float)job.height / (float)job.width);
        const V3 right = mk(c.right[0], c.right[1], c.right[2]);
        const V3 up = mk(c.up[0], c.up[1], c.up[2]);
        const V3 fwd = mk(c.forward[0], c.forward[1], c.forward[2]);
        const V3 d = normalize(right * cx + up * cy + fwd);
        queue_ray(pb.ext[0], i, mk(c.pos[0], c.pos[1], c.pos[2]), job.root, d, kFltMax);
        pb.ext_pid[0][i] = i;
        if (pp.num_segs > kSegSearchMax) pb.path_seg[i] = si;
        pb.rng[i] = s.state;
        // The throughput (one) is not stored: k_shade takes it as one at the camera vertex.
        // The MIS state and previous vertex are not read there either (no previous vertex),
        // so they are stored only for a stratified or Halton job, whose first surface's
        // dimensions u0, u2, u3, u4 and u5 travel in them to k_shade, which then holds none of
        // this code (it would push the shading kernel past its registers).
        float4 w0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), p0 = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0u));
        if (job.sampler == PHG_SAMPLER_STRATIFIED) {
            // the stream values k_shade draws at the first vertex (u0 .. u5, after the raster
            // pair) are the strata's jitters: stages (u0, u2), (u3, u4), u5
            Stream t = s;
            const float j0 = t.u();
            (void)t.u();
            const float j2 = t.u(), j3 = t.u(), j4 = t.u(), j5 = t.u();
            w0 = make_float4(strat_2d(pixel, job.npix, job.seed, sample, 0u, 0u, j0),
                             strat_2d(pixel, job.npix, job.seed, sample, 0u, 1u, j2),
                             strat_2d(pixel, job.npix, job.seed, sample, 1u, 0u, j3),
                             strat_2d(pixel, job.npix, job.seed, sample, 1u, 1u, j4));
            p0.x = strat_1d(pixel, job.npix, job.seed, sample, j5);
        } else if ((job.sampler & 0xFFu) == PHG_SAMPLER_HALTON) {
            const uint64_t hidx = (uint64_t)sample * job.npix + pixel;
            w0 = make_float4(halton_value(2u, hidx, job.sampler, job.halton_leap, job.seed),
                             halton_value(3u, hidx, job.sampler, job.halton_leap, job.seed),
                             halton_value(4u, hidx, job.sampler, job.halton_leap, job.seed),
                             halton_value(5u, hidx, job.sampler, job.halton_leap, job.seed));
            p0.x = halton_value(6u, hidx, job.sampler, job.halton_leap, job.seed);
        }
        if (job.sampler != PHG_SAMPLER_INDEPENDENT) {
            pb.Wem[i] = w0;
            pb.prevP[i] = p0;
        }
        pb.L[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (pp.record)
            for (uint32_t b = 0; b < pp.rec_bounces; ++b)
                pb.rec[((size_t)b * pp.paths + i) * kRecFloat4] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// Every ray of a path queue, traced by the lanes of the persistent grid; a lane starts its
// next ray as soon as its current one finishes.  Queue rays are (origin, root node bits),
// (direction, tmax) with tmin = 0; done(i, result) consumes each finished ray.
//
// Two schedules.  A plain grid-stride split (lane l of wave w takes rays k S + 64 w + l) left
// most of every launch's tail to a few waves: the sequencer favours older waves, so the
// younger ones of each CU finished up to 2x later, and with S a multiple of the image's tile
// rows each wave of the camera bounce saw the same tile columns every round.
//  - coherent (the camera rays, queued in 8 x 8 pixel tiles): static rounds, wave w taking the
//    64-ray group (w + 7 k) mod W of round k, so that a wave's rounds walk across the image;
//    a wave's 64 lanes trace one tile together, which keeps their node fetches shared.
//  - otherwise dynamic: the queue is split into 8 regions, one per XCD (blocks are dispatched
//    to the XCDs round robin); a wave takes chunks of kFetchChunk consecutive rays from its
//    XCD's region counter (one atomic per chunk) and gives them to its idle lanes in lane
//    order; once its region is drained it takes chunks from the next kFetchSteal regions, then
//    stops: every region is drained by its own XCD's waves, so stealing only evens out the
//    end.  fetch: 8 region counters kFetchStride words apart (one 128-byte line each: atomics
//    on one line serialize), zero at launch.
// Which lane traces a ray does not change its result.
constexpr uint32_t kFetchChunk = 256;
constexpr uint32_t kFetchSteal = 2;
constexpr uint32_t kRoundRotate = 7;
static_assert(kFetchChunk % 64 == 0, "chunks keep 64-ray groups aligned");

template <int RING, uint32_t BLOCK, bool ANY, bool COUNT, typename Done>
__device__ __forceinline__ void trace_queue(const SceneDev& sc, const phg_ray* __restrict__ queue, uint32_t n,
                                            const uint32_t* __restrict__ rcount, uint32_t cap,
                                            bool coherent, uint32_t* __restrict__ fetch, int32_t* stk,
                                            int32_t* spill, uint32_t spill_cap, uint32_t* nvis, uint32_t* ntri,
                                            uint32_t* ovf, Done done) {
    Traversal<RING, ANY, COUNT> tv;
    tv.attach(stk, BLOCK, spill, spill_cap);
    const uint32_t lane = __lane_id();
    uint32_t i = 0u;
    auto start = [&](uint32_t k) {
        i = k;
        const float4* rp = reinterpret_cast<const float4*>(queue + k);
        const float4 r0 = rp[0], r1 = rp[1];
        tv.begin(mk(r0.x, r0.y, r0.z), mk(r1.x, r1.y, r1.z), 0.0f, r1.w, (int32_t)__float_as_uint(r0.w));
    };
    if (coherent) {
        const uint32_t stride = gridDim.x * BLOCK, nw = stride / 64u, rot = kRoundRotate % nw;
        uint32_t group = (xcd_block_id() * BLOCK + threadIdx.x) / 64u;
        uint64_t base = 0u;
        if ((uint64_t)group * 64u + lane >= n) return;
        start(group * 64u + lane);
        while (true) {
            if (!tv.step(sc.nodes, sc.tris, nvis, ntri, ovf)) continue;
            done(i, tv.r);
            base += stride;
            group += rot;
            if (group >= nw) group -= nw;
            const uint64_t k = base + (uint64_t)group * 64u + lane;
            if (k >= n) break;
            start((uint32_t)k);
        }
        return;
    }
    const uint64_t below = lane == 0 ? 0ull : ((~0ull) >> (64u - lane));
    // the queue's regions: region r is [r cap, r cap + rcount[r kFetchStride]) (whole 64-ray
    // groups but the last)
    // wave-uniform: the region being drained, the regions still to try, the wave's chunk
    // (a grid of fewer than 8 blocks leaves regions without home waves: its waves try them all)
    uint32_t reg = blockIdx.x & 7u, left = gridDim.x >= 8u ? 1u + kFetchSteal : 8u, cnext = 0u, cend = 0u;
    bool has = false;
    while (true) {
        const uint64_t need = __ballot(!has);
        if (need != 0ull) {
            uint32_t m = (uint32_t)__popcll(need);
            uint32_t rank = (uint32_t)__popcll(need & below);
            while (m != 0u && left != 0u) {
                if (cnext == cend) {
                    const uint32_t first = (uint32_t)__ffsll((long long)__ballot(1)) - 1u;
                    uint32_t got = 0u;
                    if (lane == first) got = atomicAdd(&fetch[reg * kFetchStride], kFetchChunk);
                    got = __builtin_amdgcn_readfirstlane(got);
                    const uint32_t rbeg = reg * cap;
                    const uint32_t rend = rbeg + rcount[reg * kFetchStride];
                    if (rbeg >= rend || got >= rend - rbeg) {
                        reg = (reg + 1u) & 7u;
                        --left;
                        continue;
                    }
                    cnext = rbeg + got;
                    cend = rend - cnext < kFetchChunk ? rend : cnext + kFetchChunk;
                }
                const uint32_t take = cend - cnext < m ? cend - cnext : m;
                if (!has) {
                    if (rank < take) {
                        start(cnext + rank);
                        has = true;
                    } else {
                        rank -= take;
                    }
                }
                cnext += take;
                m -= take;
            }
        }
        if (!has) break;
        if (!tv.step(sc.nodes, sc.tris, nvis, ntri, ovf)) continue;
        done(i, tv.r);
        has = false;
    }
}

// The traversal kernels keep 7 waves per SIMD (at most 72 VGPRs): their persistent grids hold
// 7 waves per SIMD, and a wave fewer measured 2-4% slower.
template <int RING, uint32_t BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(7))) void k_extend(SceneDev sc, PathBuffers pb, uint32_t bk, uint32_t count_traversal) {
    const uint32_t bounce = pb.bounce_base[0] + bk;
    __shared__ int32_t stack_mem[RING * BLOCK];
    int32_t* spill = pb.spill + (size_t)(xcd_block_id() * BLOCK + threadIdx.x) * pb.spill_cap;
    // entries queued for this bounce (k_generate's count, or the one the previous bounce's
    // shadow kernel summed from the regions), and the queue's regions
    const uint32_t n = pb.counts[bounce % kCountSlots];
    const uint32_t cap = pb.bounce_base[1];
    const uint32_t* rcount = pb.fetch + fetch_slot(kPushExt, bounce);
    // the region counters the shade stage of this bounce pushes to (this bounce's shadow rays,
    // the next bounce's extension rays) start from zero, as do the work counters
    if (blockIdx.x == 0 && threadIdx.x == 0 && bounce > 0) atomicAdd(&pb.traversal_counters[4], (unsigned long long)n);
    if (blockIdx.x == 0 && threadIdx.x < 8u) {
        pb.fetch[fetch_slot(kFetchExtend, bounce + 1u) + threadIdx.x * kFetchStride] = 0u;
        pb.fetch[fetch_slot(kFetchShadow, bounce) + threadIdx.x * kFetchStride] = 0u;
        pb.fetch[fetch_slot(kFetchShade, bounce) + threadIdx.x * kFetchStride] = 0u;
        pb.fetch[fetch_slot(kPushExt, bounce + 1u) + threadIdx.x * kFetchStride] = 0u;
        pb.fetch[fetch_slot(kPushShadow, bounce) + threadIdx.x * kFetchStride] = 0u;
    }
    const uint32_t q = bounce & 1u;
    uint32_t nvis = 0, ntri = 0, ovf = 0;
    if (count_traversal)
        trace_queue<RING, BLOCK, false, true>(sc, pb.ext[q], n, rcount, cap, bounce == 0u, pb.fetch + fetch_slot(kFetchExtend, bounce), stack_mem + threadIdx.x, spill, pb.spill_cap, &nvis, &ntri, &ovf,
                                       [&](uint32_t i, const TraceResult& tr) {
                                           pb.hits[i] = make_float4(__int_as_float(tr.tri), tr.a, tr.b, tr.c);
                                       });
    else
        trace_queue<RING, BLOCK, false, false>(sc, pb.ext[q], n, rcount, cap, bounce == 0u, pb.fetch + fetch_slot(kFetchExtend, bounce), stack_mem + threadIdx.x, spill, pb.spill_cap, &nvis, &ntri, &ovf,
                                        [&](uint32_t i, const TraceResult& tr) {
                                            pb.hits[i] = make_float4(__int_as_float(tr.tri), tr.a, tr.b, tr.c);
                                        });
    if (count_traversal) {
        atomicAdd(&pb.traversal_counters[0], (unsigned long long)nvis);
        atomicAdd(&pb.traversal_counters[1], (unsigned long long)ntri);
    }
    if (ovf) atomicOr(pb.error_flags, 1u);
}

template <int RING, uint32_t BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(7))) void k_shadow(SceneDev sc, const PassParams* __restrict__ ppd, PathBuffers pb,
                                                  uint32_t bk, uint32_t count_traversal) {
    const uint32_t bounce = pb.bounce_base[0] + bk;
    const uint32_t rec_bounces = ppd->record ? ppd->rec_bounces : 0u, npaths = ppd->paths;
    __shared__ int32_t stack_mem[RING * BLOCK];
    int32_t* spill = pb.spill + (size_t)(xcd_block_id() * BLOCK + threadIdx.x) * pb.spill_cap;
    const uint32_t cap = pb.bounce_base[1];
    const uint32_t* rcount = pb.fetch + fetch_slot(kPushShadow, bounce);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // the shading of this bounce is complete: the totals of its shadow rays (statistics)
        // and of the next bounce's extension rays (the next bounce's traversal, inference and
        // the host's test for an empty queue)
        uint32_t ns = 0u, ne = 0u;
        const uint32_t* ecount = pb.fetch + fetch_slot(kPushExt, bounce + 1u);
        for (uint32_t r = 0; r < 8u; ++r) {
            ns += rcount[r * kFetchStride];
            ne += ecount[r * kFetchStride];
        }
        pb.counts[(bounce + 1) % kCountSlots] = ne;
        atomicAdd(&pb.traversal_counters[5], (unsigned long long)ns);
    }
    uint32_t nvis = 0, ntri = 0, ovf = 0;
    auto unoccluded = [&](uint32_t i, const TraceResult& tr) {
        if (tr.tri < 0) {
            const uint32_t pid = pb.shadow_pid[i];
            const float4 c = pb.shadow_c[i];
            float4 L = pb.L[pid];
            L.x += c.x; L.y += c.y; L.z += c.z;
            pb.L[pid] = L;
            if (bounce < rec_bounces) {
                // NEE at this vertex is not radiance arriving along the sampled direction
                float4* lb = pb.rec + ((size_t)bounce * npaths + pid) * kRecFloat4 + 5;
                float4 v = *lb;
                v.x += c.x; v.y += c.y; v.z += c.z;
                *lb = v;
            }
        }
    };
    if (count_traversal)
        trace_queue<RING, BLOCK, true, true>(sc, pb.shadow, 0u, rcount, cap, false, pb.fetch + fetch_slot(kFetchShadow, bounce), stack_mem + threadIdx.x, spill, pb.spill_cap, &nvis, &ntri, &ovf,
                                      unoccluded);
    else
        trace_queue<RING, BLOCK, true, false>(sc, pb.shadow, 0u, rcount, cap, false, pb.fetch + fetch_slot(kFetchShadow, bounce), stack_mem + threadIdx.x, spill, pb.spill_cap, &nvis, &ntri, &ovf,
                                       unoccluded);
    if (count_traversal) {
        atomicAdd(&pb.traversal_counters[2], (unsigned long long)nvis);
        atomicAdd(&pb.traversal_counters[3], (unsigned long long)ntri);
    }
    if (ovf) atomicOr(pb.error_flags, 1u);
}


// One bounce-loop iteration of BNEEPTEstimator::estimate (BNEEPTEstimator.cpp:84-255) for
// every path whose extension ray of this bounce is queued: emission found by the previous
// BSDF sample (MIS against NEE), then NEE at this vertex (shadow ray queued with its
// contribution), then the BSDF sample that extends the path, throughput update and
// Russian roulette.  The queue pushes ballot over the lanes that reach them (one atomic per
// wave and queue).
// GUIDED: the extension direction comes from one-sample MIS between the BSDF sampler and
// the guiding network's lobes (probability guide_prob), and the combined density replaces
// the BSDF pdf in both MIS weights against NEE, which keeps the estimator unbiased.
// Register budget: left alone the compiler takes ~180 VGPRs (2 waves per SIMD); capping it at
// 128 (4 waves, with some spills to scratch) hides more of the kernel's scattered loads and
// measured 15% faster on the 1M-triangle, 1024^2 workload (5 or 6 waves spill more and lose).
template <bool GUIDED>
__global__ __launch_bounds__(kShadeBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_shade(SceneDev sc, const PassParams* __restrict__ ppd, PathBuffers pb,
                                                       uint32_t bk) {
    const uint32_t bounce = pb.bounce_base[0] + bk;
    const PassParams& pp = *ppd;   // fields read where used (uniform loads; no registers held)
    const uint32_t cap = pb.bounce_base[1];
    const uint32_t* rcount = pb.fetch + fetch_slot(kPushExt, bounce);
    uint32_t* push_shadow = pb.fetch + fetch_slot(kPushShadow, bounce);
    uint32_t* push_ext = pb.fetch + fetch_slot(kPushExt, bounce + 1u);
    const uint32_t qin = bounce & 1u, qout = qin ^ 1u;
    // 64-vertex groups handed out dynamically, as the traversal kernels' incoherent queues:
    // chunks of kFetchChunk vertices from the XCD's region of the queue, then from the next
    // kFetchSteal regions (wave-uniform state; one atomic per chunk)
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t* fetch = pb.fetch + fetch_slot(kFetchShade, bounce);
    uint32_t reg = blockIdx.x & 7u, left = gridDim.x >= 8u ? 1u + kFetchSteal : 8u, cnext = 0u, cend = 0u;
    // Queue entries whose ray left the scene need no shading, and after the first bounce they
    // are most of the queue (rays from the ground that go to the sky).  Each wave therefore
    // gathers the queue indices of the entries that hit a surface into its own LDS list and
    // shades them 64 at a time, so its lanes are not left idle beside the misses.  hl holds
    // nl (wave-uniform) indices, and hr the queue region of each (whose region of the output
    // queues the vertex pushes to); a 64-entry group adds at most 64 to fewer than 64.
    __shared__ uint32_t s_hitlist[kShadeBlock / 64u][128];
    __shared__ uint8_t s_hitreg[kShadeBlock / 64u][128];
    uint32_t* hl = s_hitlist[threadIdx.x >> 6];
    uint8_t* hr = s_hitreg[threadIdx.x >> 6];
    const uint64_t lanes_below = lane == 0u ? 0ull : (~0ull >> (64u - lane));
    uint32_t nl = 0u;
    for (;;) {
        while (nl < 64u && left != 0u) {
            if (cnext == cend) {
                uint32_t got = 0u;
                if (lane == 0u) got = atomicAdd(&fetch[reg * kFetchStride], kFetchChunk);
                got = __builtin_amdgcn_readfirstlane(got);
                const uint32_t rbeg = reg * cap;
                const uint32_t rend = rbeg + rcount[reg * kFetchStride];
                if (rbeg >= rend || got >= rend - rbeg) {
                    reg = (reg + 1u) & 7u;
                    --left;
                    continue;
                }
                cnext = rbeg + got;
                cend = rend - cnext < kFetchChunk ? rend : cnext + kFetchChunk;
            }
            const uint32_t j = cnext + lane;
            const bool hit = j < cend && __float_as_int(pb.hits[j].x) >= 0;
            cnext = cend - cnext < 64u ? cend : cnext + 64u;
            const uint64_t hm = __ballot(hit);
            if (hit) {
                const uint32_t pos = nl + (uint32_t)__popcll(hm & lanes_below);
                hl[pos] = j;
                hr[pos] = (uint8_t)reg;
            }
            nl += (uint32_t)__popcll(hm);
        }
        if (nl == 0u) break;
        // the list's stores and loads cross lanes of this wave only: LDS serves one wave's
        // accesses in order, and the wave barrier keeps the compiler from reordering them
        __builtin_amdgcn_wave_barrier();
        const bool have = lane < nl;
        const uint32_t i = have ? hl[lane] : 0u;
        const uint32_t ireg = have ? hr[lane] : 0u;
        __builtin_amdgcn_wave_barrier();
        if (nl > 64u) {
            const uint32_t rest = nl - 64u;
            const uint32_t carry = lane < rest ? hl[64u + lane] : 0u;
            const uint8_t carry_r = lane < rest ? hr[64u + lane] : (uint8_t)0;
            __builtin_amdgcn_wave_barrier();
            if (lane < rest) {
                hl[lane] = carry;
                hr[lane] = carry_r;
            }
            nl = rest;
        } else {
            nl = 0u;
        }
        __builtin_amdgcn_wave_barrier();
        if (have) {
            const uint32_t pid = pb.ext_pid[qin][i];
            const float4 h = pb.hits[i];
            const int32_t tri = __float_as_int(h.x);
            const float4 rd = reinterpret_cast<const float4*>(pb.ext[qin] + i)[1];
            const V3 dir = mk(rd.x, rd.y, rd.z);
            bool stop = tri < 0;
            SurfaceHit X;
            uint32_t mat = 0;
            int32_t emitter = -1;
            if (!stop) {
                // the triangle's vertices are read again where the escape tests need them,
                // instead of being held in registers through the NEE and BSDF stages
                const TriRec& tr = sc.tris[tri];
                mat = tr.mat;
                emitter = tr.emitter;
                TriVerts tv;
                load_tri(sc.tris, (uint32_t)tri, &tv.a, &tv.b, &tv.c);
                X = make_hit(tv.a, tv.b, tv.c, (uint32_t)tri, h.y, h.z, h.w);
                if (!side_agreed(X, dir)) stop = true;  // SurfaceTracer.h:126-143
            }
            const uint32_t nseg = pp.num_segs;
            const PassSeg sg = pp.segs[nseg <= kSegSearchMax ? find_seg(pp.segs, nseg, pid) : pb.path_seg[pid]];
            const JobDev& job = sc.jobs[sg.job];
            const bool bvpt = job.estimator != PHG_ESTIMATOR_BNEEPT;
            // a camera vertex's throughput is one (k_generate does not store it)
            float4 T4 = bounce == 0u ? make_float4(1.0f, 1.0f, 1.0f, 1.0f) : pb.T[pid];
            if (!stop && emitter >= 0) {
                // the radiance so far is read and written here only (not held through the
                // NEE and BSDF stages)
                float4 Lacc = pb.L[pid];
                bool l_dirty = false;
                // DiffuseSurfaceEmitterBase::evalEnergyFromSurface: front face only
                const EmitterRec em = sc.emitters[emitter];
                const bool front = dot(neg(dir), X.Ns) > 0.0f;
                const RGB Le = front ? rgb3(em.radiance[0], em.radiance[1], em.radiance[2]) : rgb1(0.0f);
                if (bvpt) {
                    // BVPTEstimator.cpp:68-78: emission times the throughput, the throughput
                    // clamped to [0, 1e9] in place
                    T4.x = safe_clamp(T4.x, 0.0f, 1e9f); T4.y = safe_clamp(T4.y, 0.0f, 1e9f);
                    T4.z = safe_clamp(T4.z, 0.0f, 1e9f);
                    Lacc.x += Le.r * T4.x; Lacc.y += Le.g * T4.y; Lacc.z += Le.b * T4.z;
                    l_dirty = true;
                } else if (bounce == 0) {
                    // zero-bounce emission (SurfaceTracer.h:239-259)
                    Lacc.x += Le.r; Lacc.y += Le.g; Lacc.z += Le.b; l_dirty = true;
                } else if (!rgb_zero(Le)) {
                    // BSDF-sampled emitter hit, MIS against NEE (BNEEPTEstimator.cpp:178-217)
                    const float4 prevP = pb.prevP[pid];
                    const float4 wem = pb.Wem[pid];
                    const V3 Xp = mk(prevP.x, prevP.y, prevP.z);
                    // neeSamplePdfWUnoccluded -> ESPowerFavoring::calcDirectPdf
                    float neePdf = 0.0f;
                    const V3 e2t = Xp - X.P;
                    if (dot(e2t, X.Ns) > 0.0f) {
                        TriVerts tv;
                        load_tri(sc.tris, (uint32_t)tri, &tv.a, &tv.b, &tv.c);
                        const float pdfA = 1.0f / tri_area(tv.a, tv.b, tv.c);
                        float pdfW = pdfA_to_pdfW(pdfA, e2t, X.Ns);
                        const float gp = em.tri_count >= 2 ? 1.0f / (float)em.tri_count : 1.0f;
                        pdfW = pdfW * gp;
                        if (pdf_ok(pdfW)) {
                            const float pdf = pdfW * em.pick_pdf;
                            neePdf = pdf_ok(pdf) ? pdf : 0.0f;
                        }
                    }
                    const float bsdfPdf = wem.w;
                    if (pdf_ok(bsdfPdf)) {
                        const float w = mis_power(bsdfPdf, neePdf);
                        RGB weight = muls(rgb3(wem.x, wem.y, wem.z), w);
                        weight = rational_clamp(weight);
                        Lacc.x += Le.r * weight.r; Lacc.y += Le.g * weight.g; Lacc.z += Le.b * weight.b;
                        l_dirty = true;
                    }
                }
                if (l_dirty) pb.L[pid] = Lacc;
            }
            if (!stop && bounce > 0) {
                const uint32_t flags = __float_as_uint(pb.prevP[pid].w);
                if (flags & kFlagDieAfterEmission) stop = true;
            }
            if (!stop && bounce >= job.max_bounces) stop = true;
            if (!stop) {
                uint32_t pixel, sample;
                path_ids(sg, job, pid, &pixel, &sample);
                const uint64_t key = path_key(job.seed, pixel, sample, job.npix);
                Stream s = stream_resume(key, pb.rng[pid]);
                // the vertex's random numbers, drawn in stream order where they are first
                // needed (NEE: 4, then BSDF: 3 + Russian roulette: 1, + 4 when guided), so that
                // no more of them are live at once than the stage uses
                float u0 = s.u();
                const float u1 = s.u();
                float u2 = s.u(), u3 = s.u();
                // a stratified or Halton job's first-surface dimensions (u0, u2, u3, u4, u5),
                // computed by k_generate into the first vertex's (still unused) MIS state
                const bool src0 = bounce == 0u && job.sampler != PHG_SAMPLER_INDEPENDENT;
                if (src0) {
                    const float4 hv = pb.Wem[pid];
                    u0 = hv.x; u2 = hv.y; u3 = hv.z;
                }
                const float* lob = GUIDED ? pb.lobes + (size_t)i * kGuideDec : nullptr;
                const bool guide_here = pp.guide_depth == 0 || bounce < pp.guide_depth;
                const phg_material& m = sc.mats[mat];
                // the lobes' selection probability: guide_prob, or the network's own for this
                // query where guide_prob < 0
                const float alpha = (GUIDED && guide_here && pp.guide_prob != 0.0f && guide_samples_material(m) &&
                                     guide_lobes_ok(lob)) ? (pp.guide_prob > 0.0f ? pp.guide_prob : lob[kGuideSel]) : 0.0f;
                if (m.type == PHG_MAT_ABSORBER) stop = true;
                if (!stop) {
                    const RGB T = rgb3(T4.x, T4.y, T4.z);
                    float rrScale = T4.w;
                    const V3 V = dir * -1.0f;
                    // ---- NEE (TDirectLightEstimator.ipp:70-102, BNEEPTEstimator.cpp:103-149)
                    if (!bvpt && job.num_emitters > 0) {
                        const uint32_t ei = cdf_sample(sc.emit_cdf + job.cdf_begin, job.num_emitters, u0);
                        const EmitterRec em = sc.emitters[job.emitter_begin + ei];
                        uint32_t k = (uint32_t)(u1 * (float)em.tri_count);
                        if (k >= em.tri_count) k = em.tri_count - 1u;
                        const EmitTri et = sc.emit_tris[em.tri_begin + k];
                        TriVerts te;
                        load_tri(sc.tris, et.tri, &te.a, &te.b, &te.c);
                        const float A = sqrtf(u2);
                        const float B = u3;
                        const float e0 = 1.0f - A, e1 = A * (1.0f - B), e2 = B * A;
                        const float pdfA = 1.0f / et.area;
                        const SurfaceHit Xe = make_hit(te.a, te.b, te.c, et.tri, e0, e1, e2);
                        const V3 e2t = X.P - Xe.P;
                        if (dot(e2t, Xe.Ns) > 0.0f) {
                            float pdf = pdfA_to_pdfW(pdfA, e2t, Xe.Ns);
                            if (em.tri_count >= 2) pdf = pdf * (1.0f / (float)em.tri_count);
                            pdf = pdf * em.pick_pdf;
                            V3 so, sd; float st;
                            TriVerts tv;
                            load_tri(sc.tris, (uint32_t)tri, &tv.a, &tv.b, &tv.c);
                            if (pdf_ok(pdf) && side_agreed(X, Xe.P - X.P) && mutual_escape(X, tv, Xe, te, &so, &sd, &st)) {
                                const V3 Ld = normalize(Xe.P - X.P);
                                const BsdfEval be = bsdf_eval(m, X, Ld, V);
                                if (be.ok) {
                                    float bp = bsdf_pdf(m, X, Ld, V);
                                    if (!pdf_ok(bp)) bp = 0.0f;
                                    if (GUIDED && alpha > 0.0f) bp = alpha * guide_pdf(lob, Ld) + (1.0f - alpha) * bp;
                                    const float w = mis_power(pdf, bp);
                                    RGB weight = muls(be.f, absdot(X.Ns, Ld));
                                    weight = mulc(weight, T);
                                    weight = muls(weight, w / pdf);
                                    weight = rational_clamp(weight);
                                    const RGB sh_c = rgb3(em.radiance[0] * weight.r, em.radiance[1] * weight.g,
                                                          em.radiance[2] * weight.b);
                                    if (!rgb_zero(sh_c)) {
                                        // queued at once (its lanes ballot together), so the
                                        // shadow ray is not carried through the BSDF stage
                                        const uint32_t slot_s = wave_push_region(push_shadow, ireg, cap, true);
                                        queue_ray(pb.shadow, slot_s, so, job.root, sd, st);
                                        pb.shadow_pid[slot_s] = pid;
                                        pb.shadow_c[slot_s] = make_float4(sh_c.r, sh_c.g, sh_c.b, 0.0f);
                                    }
                                }
                            }
                        }
                    }
                    // ---- BSDF sample + extension (BNEEPTEstimator.cpp:151-255)
                    float u4 = s.u(), u5 = s.u();
                    const float u6 = s.u(), u7 = s.u();
                    if (src0) {
                        u4 = pb.Wem[pid].w;
                        u5 = pb.prevP[pid].x;
                    }
                    float u8 = 0.0f, u9 = 0.0f, u10 = 0.0f, u11 = 0.0f;
                    if (GUIDED) { u8 = s.u(); u9 = s.u(); u10 = s.u(); u11 = s.u(); }
                    bool ok = false;
                    V3 Ls = mk(0.0f, 0.0f, 0.0f);
                    RGB ws = rgb1(0.0f);
                    float relIor = 1.0f, pdfc = 0.0f;
                    float bsdf_pdf_ls = 0.0f;  // BSDF density of the chosen direction (training records)
                    if (GUIDED && u8 < alpha) {
                        // direction from the learned lobes, weighted by the combined density
                        Ls = guide_sample(lob, u9, u10, u11);
                        const BsdfEval be = bsdf_eval(m, X, Ls, V);
                        float bp = bsdf_pdf(m, X, Ls, V);
                        if (!pdf_ok(bp)) bp = 0.0f;
                        bsdf_pdf_ls = bp;
                        pdfc = alpha * guide_pdf(lob, Ls) + (1.0f - alpha) * bp;
                        if (be.ok && pdf_ok(pdfc)) {
                            const float cosl = absdot(X.Ns, Ls);
                            ws = muls(be.f, cosl / pdfc);
                            if (m.type == PHG_MAT_DIELECTRIC && opposite_hemi(X, V, Ls)) {
                                float etaI = m.ior_outer, etaT = m.ior_inner;
                                if (dot(X.Ns, Ls) < 0.0f) { const float t = etaI; etaI = etaT; etaT = t; }
                                relIor = etaT / etaI;
                            }
                            ok = measurable(ws, cosl) && side_agreed(X, Ls);
                        }
                    } else {
                        const BsdfSample bs = bsdf_sample(m, X, V, u4, u5, u6);
                        if (bs.ok) {
                            Ls = bs.L; ws = bs.w; relIor = bs.relIor;
                            const float bp = bsdf_pdf(m, X, bs.L, V);
                            bsdf_pdf_ls = pdf_ok(bp) ? bp : 0.0f;
                            if (GUIDED && alpha > 0.0f) {
                                pdfc = alpha * guide_pdf(lob, Ls) + (1.0f - alpha) * bp;
                                ok = pdf_ok(bp) && pdf_ok(pdfc);
                                if (ok) ws = muls(ws, bp / pdfc);
                            } else {
                                pdfc = bp;
                                ok = true;
                            }
                        }
                    }
                    if (ok) {
                        V3 off, ud;
                        TriVerts tv;
                        load_tri(sc.tris, (uint32_t)tri, &tv.a, &tv.b, &tv.c);
                        iterative_offset(X, tv, Ls, &off, &ud);
                        const RGB wem = mulc(ws, T);
                        RGB Tn = mulc(T, ws);
                        rrScale /= relIor * relIor;
                        bool die = false;
                        if (bounce >= 3) {  // RussianRoulette.h:30-55
                            const float surv = safe_clamp(luminance(muls(Tn, rrScale)), 0.0f, 0.95f);
                            if (u7 < surv) Tn = muls(Tn, 1.0f / surv);
                            else die = true;
                        }
                        if (!die && rgb_zero(Tn)) die = true;
                        // BVPT ends a path that Russian roulette (or a zero throughput) stops;
                        // BNEEPT still follows the sampled direction to weigh the emission it
                        // finds against NEE, then stops
                        if (!(bvpt && die)) {
                            pb.T[pid] = make_float4(Tn.r, Tn.g, Tn.b, rrScale);
                            pb.Wem[pid] = make_float4(wem.r, wem.g, wem.b, pdfc);
                            pb.prevP[pid] = make_float4(X.P.x, X.P.y, X.P.z, __uint_as_float(die ? kFlagDieAfterEmission : 0u));
                            pb.rng[pid] = s.state;
                            const uint32_t slot_e = wave_push_region(push_ext, ireg, cap, true);
                            queue_ray(pb.ext[qout], slot_e, X.P + off, job.root, ud, kFltMax);
                            pb.ext_pid[qout][slot_e] = pid;
                            if (GUIDED && pp.record && bounce < pp.rec_bounces && !guide_samples_material(m)) {
                                // no record where the lobes are never sampled (their fit there
                                // would only take capacity from the surfaces they serve); the
                                // marker lets the collection skip to the path's next vertex
                                pb.rec[((size_t)bounce * pp.paths + pid) * kRecFloat4] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
                            } else if (GUIDED && pp.record && bounce < pp.rec_bounces) {
                                float4* rc = pb.rec + ((size_t)bounce * pp.paths + pid) * kRecFloat4;
                                rc[0] = make_float4(X.P.x, X.P.y, X.P.z, pdfc);
                                rc[1] = make_float4(X.Ns.x, X.Ns.y, X.Ns.z, bsdf_pdf_ls);
                                rc[2] = make_float4(V.x, V.y, V.z, alpha);
                                // the sampling density where the selection probability is learned
                                rc[3] = make_float4(ud.x, ud.y, ud.z, pp.guide_prob < 0.0f ? pdfc : 0.0f);
                                rc[4] = make_float4(T.r, T.g, T.b, 0.0f);  // throughput before this vertex
                                const float4 Lnow = pb.L[pid];   // radiance so far, this vertex's emission included
                                rc[5] = make_float4(Lnow.x, Lnow.y, Lnow.z, 0.0f);
                            }
                        }
                    }
                }
            }
        }
    }
}

// Reconstruction filters of SampleFilter (Core/Filmic/SampleFilter.cpp:8-56), evaluated at the
// offset from the pixel centre to the sample, in pixels; all three span 4 x 4 pixels.
//   Gaussian: sigma 0.5, minus its value at the corner (2, 2) ("submerged"), floored at 0
//   Mitchell-Netravali: cubic with b = c = 1/3 (Mitchell & Netravali 1988)
//   Blackman-Harris: 4-term window of radius 2 (Math/Function/TBlackmanHarris2D.h:42-60)
__device__ double film_filter(uint32_t f, double x, double y) {
    if (f == PHG_FILTER_GAUSSIAN) {
        const double edge = exp(-(4.0 / 0.5) - (4.0 / 0.5));
        const double g = exp(-(x * x) / 0.5 - (y * y) / 0.5);
        return fmax(g - edge, 0.0);
    }
    if (f == PHG_FILTER_MN) {
        auto mn = [](double v) {
            const double b = 1.0 / 3.0, c = 1.0 / 3.0, a = fabs(v);
            if (a < 1.0) return ((12.0 - 9.0 * b - 6.0 * c) * a * a * a + (-18.0 + 12.0 * b + 6.0 * c) * a * a + (6.0 - 2.0 * b)) / 6.0;
            if (a < 2.0) return ((-b - 6.0 * c) * a * a * a + (6.0 * b + 30.0 * c) * a * a + (-12.0 * b - 48.0 * c) * a +
                                 (8.0 * b + 24.0 * c)) / 6.0;
            return 0.0;
        };
        return mn(x) * mn(y);
    }
    auto bh = [](double v) {
        const double n = (v * 0.5 + 1.0) * 0.5, pi = 3.14159265358979323846;
        return 0.35875 - 0.48829 * cos(2.0 * pi * n) + 0.14128 * cos(4.0 * pi * n) - 0.01168 * cos(6.0 * pi * n);
    };
    return bh(x) * bh(y);
}

// Filtered film, gathered per pixel: every sample of this pass whose filter footprint covers
// the pixel centre (the range HdrRgbFilm::addRgbSample splats to, HdrRgbFilm.cpp:82-122, with
// half size 2) adds its radiance times the filter weight.  A sample's raster position comes
// from its stratum (the pixel that keys its stream) and the first two numbers of its stream,
// placed in the soft-edged sample window (window_raster); only strata whose part of the
// window reaches within the footprint of this pixel are visited.
__device__ void film_gather(const JobDev& job, const PassSeg& sg, const PathBuffers& pb, uint32_t x) {
    const int W = (int)job.width, H = (int)job.height;
    const int px = (int)(x % job.width), py = (int)(x / job.width);
    double2* f = reinterpret_cast<double2*>(pb.film + job.film_base + 8 * (size_t)x);
    double2 f0 = f[0], f1 = f[1], f2 = f[2], f3 = f[3];
    // samples that splat to px lie in [px - 1.5, px + 2.5]; stratum q covers
    // [lo + q s, lo + (q + 1) s] with lo = 0.5 - 2 and s = (n + 3) / n
    const double lo = 0.5 - kFilterHalf;
    const double sx = ((double)W + 2.0 * kFilterHalf - 1.0) / (double)W;
    const double sy = ((double)H + 2.0 * kFilterHalf - 1.0) / (double)H;
    const int qx0 = max((int)floor(((double)px - 1.5 - lo) / sx) - 1, 0);
    const int qx1 = min((int)floor(((double)px + 2.5 - lo) / sx) + 1, W - 1);
    const int qy0 = max((int)floor(((double)py - 1.5 - lo) / sy) - 1, 0);
    const int qy1 = min((int)floor(((double)py + 2.5 - lo) / sy) + 1, H - 1);
    for (int qy = qy0; qy <= qy1; ++qy)
        for (int qx = qx0; qx <= qx1; ++qx) {
            const uint32_t q = (uint32_t)(qy * W + qx);
            const uint32_t slot = tile_slot((uint32_t)qx, (uint32_t)qy, job.width, job.height);
            for (uint32_t s = 0; s < sg.samples; ++s) {
                Stream st = stream_init(path_key(job.seed, q, sg.sample_begin + s, job.npix));
                const float ux = st.u(), uy = st.u();
                const double xp = window_raster((uint32_t)qx, ux, job.width, kFilterHalf);
                const double yp = window_raster((uint32_t)qy, uy, job.height, kFilterHalf);
                // pixels the sample splats to: ceil(p - 2.5) .. floor(p + 1.5), within the image
                if (px < (int)ceil(xp - 2.5) || px > (int)floor(xp + 1.5)) continue;
                if (py < (int)ceil(yp - 2.5) || py > (int)floor(yp + 1.5)) continue;
                const double w = film_filter(job.filter, (double)px - (xp - 0.5), (double)py - (yp - 0.5));
                const float4 L = pb.L[sg.path_offset + (size_t)s * sg.npix + slot];
                f0.x += w * L.x; f0.y += w * L.y; f1.x += w * L.z; f1.y += w;
                f2.x += w * (double)L.x * L.x; f2.y += w * (double)L.y * L.y; f3.x += w * (double)L.z * L.z;
            }
        }
    f[0] = f0; f[1] = f1; f[2] = f2; f[3] = f3;
}

// Film of a Halton job: the samples land anywhere in the sample window, so each path splats
// its radiance into the pixels whose centres its filter covers, with FP64 atomics (the
// reference's ranges, HdrRgbFilm.cpp:82-122: within the image, bounds inclusive, so a box
// sample on a pixel edge counts in both pixels; the order of the additions, and so the last
// bits of the sums, vary from run to run).
__device__ void film_scatter_halton(const JobDev& job, const PassSeg& sg, const PathBuffers& pb, uint32_t x) {
    const double hs = job.filter == PHG_FILTER_BOX ? 0.5 : kFilterHalf;
    const uint32_t py = x / job.width;
    const uint32_t slot = tile_slot(x - py * job.width, py, job.width, job.height);
    for (uint32_t s = 0; s < sg.samples; ++s) {
        double xp, yp;
        halton_raster(job, x, sg.sample_begin + s, &xp, &yp);
        const float4 L = pb.L[sg.path_offset + (size_t)s * sg.npix + slot];
        const double fminx = fmax(xp - hs, 0.0), fminy = fmax(yp - hs, 0.0);
        const double fmaxx = fmin(xp + hs, (double)job.width), fmaxy = fmin(yp + hs, (double)job.height);
        // within [0, W) x [0, H): fmin >= 0 and fmax <= W, H
        const int x0 = (int)ceil(fminx - 0.5), y0 = (int)ceil(fminy - 0.5);
        const int x1 = (int)floor(fmaxx - 0.5) + 1, y1 = (int)floor(fmaxy - 0.5) + 1;
        for (int yy = y0; yy < y1; ++yy)
            for (int xx = x0; xx < x1; ++xx) {
                const double w = job.filter == PHG_FILTER_BOX
                                     ? 1.0 : film_filter(job.filter, (double)xx - (xp - 0.5), (double)yy - (yp - 0.5));
                double* f = pb.film + job.film_base + 8 * ((size_t)yy * job.width + (size_t)xx);
                atomicAdd(f + 0, w * L.x); atomicAdd(f + 1, w * L.y); atomicAdd(f + 2, w * L.z);
                atomicAdd(f + 3, w);
                atomicAdd(f + 4, w * (double)L.x * L.x); atomicAdd(f + 5, w * (double)L.y * L.y);
                atomicAdd(f + 6, w * (double)L.z * L.z);
            }
    }
}

// FP64 film accumulation, owner computes: each pixel sums its own samples of this pass in
// sample order (box filter, HdrRgbFilm::addRgbSample, HdrRgbFilm.cpp:82-122), or gathers the
// filtered contributions of the samples around it.  Halton jobs scatter (film_scatter_halton).
__global__ __launch_bounds__(kShadeBlock) void k_film(SceneDev sc, const PassParams* __restrict__ ppd, PathBuffers pb) {
    const uint32_t nseg = ppd->num_segs, total = ppd->film_threads;
    const PassSeg* __restrict__ segs = ppd->segs;
    for (uint32_t t = blockIdx.x * kShadeBlock + threadIdx.x; t < total; t += gridDim.x * kShadeBlock) {
        uint32_t lo = 0, hi = nseg;  // segment of film thread t: last pix_offset <= t
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (segs[mid].pix_offset <= t) lo = mid; else hi = mid;
        }
        const PassSeg sg = segs[lo];
        const uint32_t x = t - sg.pix_offset;
        const JobDev& job = sc.jobs[sg.job];
        if ((job.sampler & 0xFFu) == PHG_SAMPLER_HALTON) {
            film_scatter_halton(job, sg, pb, x);
            continue;
        }
        if (job.filter != PHG_FILTER_BOX) {
            film_gather(job, sg, pb, x);
            continue;
        }
        double2* f = reinterpret_cast<double2*>(pb.film + job.film_base + 8 * (size_t)x);
        double2 f0 = f[0], f1 = f[1], f2 = f[2], f3 = f[3];
        const uint32_t py = x / job.width;
        const uint32_t slot = tile_slot(x - py * job.width, py, job.width, job.height);
        for (uint32_t s = 0; s < sg.samples; ++s) {
            const float4 L = pb.L[sg.path_offset + (size_t)s * sg.npix + slot];
            f0.x += L.x; f0.y += L.y; f1.x += L.z; f1.y += 1.0;
            f2.x += (double)L.x * L.x; f2.y += (double)L.y * L.y; f3.x += (double)L.z * L.z;
        }
        f[0] = f0; f[1] = f1; f[2] = f2; f[3] = f3;
    }
}

__global__ void k_clear_counts(uint32_t* counts, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) counts[i] = 0u;
}

__global__ void k_advance_bounce(uint32_t* bounce_base, uint32_t chunk) {
    if (threadIdx.x == 0) bounce_base[0] += chunk;
}

__global__ void k_fill_double(double* p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// ---------------------------------------------------------------------------------------
// Scratch of the allocation-free ray query (phg_intersect_noalloc): its error flag in the
// first 256 bytes, then the traversal-stack spill of its grid.  Each GPU holds its own copy at
// its own address, looked up once per device with that device current; host code uses only
// that address, never the array's name (which is a host-side stand-in).
// ---------------------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

hipError_t noalloc_scratch_address(int device, void** out) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void* addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices || !out) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != device) e = hipSetDevice(device);
        if (e != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[device] = p;
    }
    *out = addr[device];
    return hipSuccess;
}

// ---------------------------------------------------------------------------------------
// host launch wrappers
// ---------------------------------------------------------------------------------------
static inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

template <int RING>
static hipError_t launch_intersect_t(const SceneDev& sc, const phg_ray* rays, uint64_t n, phg_hit* hits, int any,
                                     uint32_t* nvis, uint32_t* ntri, uint32_t* flags, int32_t* spill, uint32_t spill_cap,
                                     uint32_t grid, hipStream_t st) {
    const uint32_t nb = (uint32_t)std::min<uint64_t>(grid, blocks_for(n, kTraceBlock));
    if (nvis) {
        hipLaunchKernelGGL((k_intersect<RING, false, true>), dim3(nb), dim3(kTraceBlock), 0, st, sc, rays, n, hits, nvis, ntri, flags, spill, spill_cap);
    } else if (any) {
        hipLaunchKernelGGL((k_intersect<RING, true, false>), dim3(nb), dim3(kTraceBlock), 0, st, sc, rays, n, hits, nvis, ntri, flags, spill, spill_cap);
    } else {
        hipLaunchKernelGGL((k_intersect<RING, false, false>), dim3(nb), dim3(kTraceBlock), 0, st, sc, rays, n, hits, nvis, ntri, flags, spill, spill_cap);
    }
    return hipGetLastError();
}

hipError_t launch_intersect(const SceneDev& sc, const phg_ray* rays, uint64_t n, phg_hit* hits, int any,
                            uint32_t* nvis, uint32_t* ntri, uint32_t* flags, int32_t* spill, uint32_t spill_cap,
                            uint32_t grid, hipStream_t st) {
    if (n == 0) return hipSuccess;
    return launch_intersect_t<kTraceRing>(sc, rays, n, hits, any, nvis, ntri, flags, spill, spill_cap, grid, st);
}

hipError_t launch_generate(const SceneDev& sc, const PassParams* ppd, const PathBuffers& pb, uint32_t grid, hipStream_t st) {
    hipLaunchKernelGGL(k_generate, dim3(grid), dim3(kShadeBlock), 0, st, sc, ppd, pb);
    return hipGetLastError();
}

hipError_t launch_extend(const SceneDev& sc, const PathBuffers& pb, uint32_t bk, uint32_t grid, uint32_t block,
                         uint32_t count, hipStream_t st) {
    switch (block) {
        case 64: hipLaunchKernelGGL((k_extend<kTraceRing, 64>), dim3(grid), dim3(64), 0, st, sc, pb, bk, count); break;
        case 128: hipLaunchKernelGGL((k_extend<kTraceRing, 128>), dim3(grid), dim3(128), 0, st, sc, pb, bk, count); break;
        case 256: hipLaunchKernelGGL((k_extend<kTraceRing, 256>), dim3(grid), dim3(256), 0, st, sc, pb, bk, count); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_shadow(const SceneDev& sc, const PassParams* ppd, const PathBuffers& pb, uint32_t bk, uint32_t grid,
                         uint32_t block, uint32_t count, hipStream_t st) {
    switch (block) {
        case 64: hipLaunchKernelGGL((k_shadow<kTraceRing, 64>), dim3(grid), dim3(64), 0, st, sc, ppd, pb, bk, count); break;
        case 128: hipLaunchKernelGGL((k_shadow<kTraceRing, 128>), dim3(grid), dim3(128), 0, st, sc, ppd, pb, bk, count); break;
        case 256: hipLaunchKernelGGL((k_shadow<kTraceRing, 256>), dim3(grid), dim3(256), 0, st, sc, ppd, pb, bk, count); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_shade(const SceneDev& sc, const PassParams* ppd, const PathBuffers& pb, uint32_t bk, uint32_t grid,
                        bool guided, hipStream_t st) {
    if (guided) hipLaunchKernelGGL(k_shade<true>, dim3(grid), dim3(kShadeBlock), 0, st, sc, ppd, pb, bk);
    else hipLaunchKernelGGL(k_shade<false>, dim3(grid), dim3(kShadeBlock), 0, st, sc, ppd, pb, bk);
    return hipGetLastError();
}

template <uint32_t BLOCK>
static hipError_t trace_occupancy(int* a, int* b) {
    hipError_t err = hipOccupancyMaxActiveBlocksPerMultiprocessor(a, k_extend<kTraceRing, BLOCK>, BLOCK, 0);
    if (err == hipSuccess) err = hipOccupancyMaxActiveBlocksPerMultiprocessor(b, k_shadow<kTraceRing, BLOCK>, BLOCK, 0);
    return err;
}

hipError_t resident_blocks(uint32_t trace_block, uint32_t* trace_blocks_per_cu, uint32_t* shade_blocks_per_cu) {
    int a = 0, b = 0, c = 0, d = 0, e = 0;
    hipError_t err = trace_block == 64 ? trace_occupancy<64>(&a, &b)
                   : trace_block == 128 ? trace_occupancy<128>(&a, &b)
                   : trace_block == 256 ? trace_occupancy<256>(&a, &b) : hipErrorInvalidValue;
    if (err == hipSuccess)
        err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, k_intersect<kTraceRing, false, false>, kTraceBlock, 0);
    if (err == hipSuccess) err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&d, k_shade<false>, kShadeBlock, 0);
    if (err == hipSuccess) err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&e, k_shade<true>, kShadeBlock, 0);
    if (err != hipSuccess) return err;
    // the intersect kernel runs 128-thread blocks: express its residency in trace blocks
    const int c_tb = (int)((uint64_t)c * kTraceBlock / trace_block);
    *trace_blocks_per_cu = (uint32_t)std::max(1, std::min(a, std::min(b, c_tb)));
    *shade_blocks_per_cu = (uint32_t)std::max(1, std::min(d, e));
    return hipSuccess;
}

hipError_t launch_film(const SceneDev& sc, const PassParams* ppd, const PathBuffers& pb, uint32_t grid, hipStream_t st) {
    hipLaunchKernelGGL(k_film, dim3(grid), dim3(kShadeBlock), 0, st, sc, ppd, pb);
    return hipGetLastError();
}

hipError_t launch_clear_counts(uint32_t* counts, uint32_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_clear_counts, dim3(blocks_for(n, 256)), dim3(256), 0, st, counts, n);
    return hipGetLastError();
}

hipError_t launch_advance_bounce(uint32_t* bounce_base, uint32_t chunk, hipStream_t st) {
    hipLaunchKernelGGL(k_advance_bounce, dim3(1), dim3(64), 0, st, bounce_base, chunk);
    return hipGetLastError();
}

hipError_t launch_fill_zero(double* p, uint64_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_double, dim3(blocks_for(n, 256)), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

}  // namespace phg
