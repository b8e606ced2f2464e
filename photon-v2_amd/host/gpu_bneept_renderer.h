// GpuBneeptRenderer: host-side mirror of the reference's renderer seam for the GPU BNEEPT
// path.  Same method names and meaning as ph::Renderer (Core/Renderer/Renderer.h:38-93);
// where the reference takes cooked engine objects (CoreCookedUnit, VisualWorld) this class
// takes the flattened scene they cook into (phg_scene_desc).  All device work goes through
// the C ABI of include/phg.h.
//
// Rendering proceeds in progressive chunks of samples so that the async queries observe
// progress the way EqualSamplingRenderer's work reports do
// (Core/Renderer/Sampling/EqualSamplingRenderer.cpp:205-322).
#pragma once

#include <atomic>
#include <cstdint>
#include <deque>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/phg.h"

namespace phg {

// RenderStats of the reference: integer stat 0 = finished paths per pixel,
// real stat 0 = paths per second (EqualSamplingRenderer.cpp:258-280).
struct RenderStats {
    uint64_t paths_per_pixel = 0;
    double paths_per_second = 0.0;
};

struct RenderProgress {
    uint64_t total_work = 0;   // samples per pixel requested
    uint64_t work_done = 0;    // samples per pixel finished
    uint64_t elapsed_ms = 0;
};

// Linear RGB frame, row 0 = bottom row (HdrRgbFrame as developed by HdrRgbFilm::developRegion).
struct HdrRgbFrame {
    uint32_t width = 0, height = 0;
    std::vector<float> rgb;  // 3 floats per pixel
};

struct GpuBneeptSettings {
    uint32_t width = 0, height = 0;
    uint32_t spp = 16;
    uint32_t max_bounces = 10000;   // reference semantics: MAX_RAY_BOUNCES
    uint64_t seed = 0;
    uint32_t chunk_spp = 4;         // samples per pixel per progressive chunk
    int32_t device = 0;
    uint32_t estimator = PHG_ESTIMATOR_BNEEPT;
    uint32_t filter = PHG_FILTER_BOX;
    uint32_t sampler = PHG_SAMPLER_INDEPENDENT;

    // Settings of a .p2 scene (frame size, samples, estimator, filter, sample source, bounce cap).
    static GpuBneeptSettings fromP2(const phg_p2_settings& p2);
};

// A rectangle of pixels [x0, x1) x [y0, y1), raster y up (Region = TAABB2D<int64>,
// Core/Scheduler/Region.h:8).
struct Region {
    int64_t x0 = 0, y0 = 0, x1 = 0, y1 = 0;
    int64_t area() const { return x1 > x0 && y1 > y0 ? (x1 - x0) * (y1 - y0) : 0; }
};

// Core/Renderer/ERegionStatus.h: later entries take precedence when regions merge.
enum class ERegionStatus { Invalid = 0, Finished = 1, Updating = 2 };

// Core/Renderer/RenderRegionStatus.h: a region and what happened to it.
struct RenderRegionStatus {
    Region region;
    ERegionStatus status = ERegionStatus::Invalid;
};

class GpuBneeptRenderer {
public:
    explicit GpuBneeptRenderer(const GpuBneeptSettings& settings);
    ~GpuBneeptRenderer();
    GpuBneeptRenderer(const GpuBneeptRenderer&) = delete;
    GpuBneeptRenderer& operator=(const GpuBneeptRenderer&) = delete;

    // Renderer::doUpdate: upload the cooked (flattened) scene, build the BVH, clear the film.
    void doUpdate(const phg_scene_desc& scene);
    // Renderer::doRender: render all samples (blocking), chunk by chunk.
    void doRender();
    // Renderer::retrieveFrame: developed frame (sum / weight per pixel).  Layer 0 only.
    void retrieveFrame(std::size_t layerIndex, HdrRgbFrame& out_frame);
    // Renderer::asyncQueryRenderStats / asyncQueryRenderProgress: safe during doRender.
    RenderStats asyncQueryRenderStats() const;
    RenderProgress asyncQueryRenderProgress() const;
    // Renderer::asyncPeekFrame: the pixels of `region` as of the last finished chunk, developed
    // into out_frame (sized to the whole frame if it is not already; pixels outside the region
    // are left as they are), as EqualSamplingRenderer::asyncPeekFrame develops the main film
    // over a region (EqualSamplingRenderer.cpp:217-234).
    void asyncPeekFrame(std::size_t layerIndex, const Region& region, HdrRgbFrame& out_frame);
    // Renderer::getObservationInfo: names of the stats above.
    std::vector<std::string> getObservationInfo() const;
    // Renderer::asyncPollUpdatedRegions: regions whose work finished since the last poll, oldest
    // first, up to `capacity` of them; returns how many were written.  The frame is covered by
    // 128 x 128 tiles (EqualSamplingRenderer's block size, EqualSamplingRenderer.cpp:61); when
    // a chunk of samples finishes, every tile's work unit of that chunk has finished and each
    // tile is queued as ERegionStatus::Finished (EqualSamplingRenderer::asyncAddUpdatedRegion,
    // EqualSamplingRenderer.cpp:243-258).
    std::size_t asyncPollUpdatedRegions(RenderRegionStatus* out_regions, std::size_t capacity);
    // phSaveFrame: the developed frame as a .pfm file (PfmFile::save layout).
    void saveFrame(const std::string& path);

private:
    void check(phg_status st, const char* what);
    void develop_locked(const Region& region, HdrRgbFrame& out);

    static constexpr int64_t kTile = 128;

    GpuBneeptSettings m_settings;
    void* m_ctx = nullptr;
    std::vector<double> m_film;          // host copy of the last finished chunk's film
    mutable std::mutex m_filmMutex;
    std::atomic<uint64_t> m_sppDone{0};
    std::deque<RenderRegionStatus> m_regionQueue;   // guarded by m_regionMutex
    std::mutex m_regionMutex;
    std::atomic<uint64_t> m_elapsedMs{0};
    std::atomic<double> m_pathsPerSecond{0.0};
};

}  // namespace phg
