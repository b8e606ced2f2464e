#include "gpu_bneept_renderer.h"

#include <algorithm>
#include <chrono>

namespace phg {

GpuBneeptRenderer::GpuBneeptRenderer(const GpuBneeptSettings& settings) : m_settings(settings) {
    if (settings.width == 0 || settings.height == 0) throw std::invalid_argument("GpuBneeptRenderer: empty frame");
    phg_device_opts o{settings.device, 0};
    check(phg_create(&m_ctx, &o), "phg_create");
}

GpuBneeptRenderer::~GpuBneeptRenderer() {
    if (m_ctx) phg_destroy(m_ctx);
}

void GpuBneeptRenderer::check(phg_status st, const char* what) {
    if (st != PHG_OK) {
        std::string msg = std::string(what) + " failed (status " + std::to_string((int)st) + ")";
        if (m_ctx) msg += ": " + std::string(phg_last_error(m_ctx));
        throw std::runtime_error(msg);
    }
}

void GpuBneeptRenderer::doUpdate(const phg_scene_desc& scene) {
    check(phg_upload_scene(m_ctx, &scene, nullptr), "phg_upload_scene");
    check(phg_film_clear(m_ctx, m_settings.width, m_settings.height), "phg_film_clear");
    std::lock_guard<std::mutex> lock(m_filmMutex);
    m_film.assign(8ull * m_settings.width * m_settings.height, 0.0);
    m_sppDone = 0;
    {
        std::lock_guard<std::mutex> rl(m_regionMutex);
        m_regionQueue.clear();
    }
    m_elapsedMs = 0;
    m_pathsPerSecond = 0.0;
}

void GpuBneeptRenderer::doRender() {
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t chunk = std::max(1u, m_settings.chunk_spp);
    std::vector<double> staging(8ull * m_settings.width * m_settings.height);
    for (uint32_t s = 0; s < m_settings.spp; s += chunk) {
        phg_render_opts o{};
        o.width = m_settings.width;
        o.height = m_settings.height;
        o.spp = std::min(chunk, m_settings.spp - s);
        o.spp_begin = s;
        o.max_bounces = m_settings.max_bounces;
        o.seed = m_settings.seed;
        o.estimator = m_settings.estimator;
        o.filter = m_settings.filter;
        o.sampler = m_settings.sampler;
        check(phg_render(m_ctx, &o), "phg_render");
        check(phg_film_read(m_ctx, staging.data(), staging.size()), "phg_film_read");
        {
            std::lock_guard<std::mutex> lock(m_filmMutex);
            m_film.swap(staging);
        }
        m_sppDone = s + o.spp;
        {
            // every tile's work unit of this chunk is done
            std::lock_guard<std::mutex> rl(m_regionMutex);
            for (int64_t y = 0; y < (int64_t)m_settings.height; y += kTile)
                for (int64_t x = 0; x < (int64_t)m_settings.width; x += kTile) {
                    RenderRegionStatus r;
                    r.region = Region{x, y, std::min<int64_t>(x + kTile, m_settings.width),
                                      std::min<int64_t>(y + kTile, m_settings.height)};
                    r.status = ERegionStatus::Finished;
                    m_regionQueue.push_back(r);
                }
        }
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        m_elapsedMs = (uint64_t)ms;
        m_pathsPerSecond = ms > 0 ? (double)m_sppDone * m_settings.width * m_settings.height / (ms / 1e3) : 0.0;
    }
}

void GpuBneeptRenderer::develop_locked(const Region& region, HdrRgbFrame& out) {
    const int64_t W = m_settings.width, H = m_settings.height;
    if (out.width != m_settings.width || out.height != m_settings.height || out.rgb.size() != 3ull * W * H) {
        out.width = m_settings.width;
        out.height = m_settings.height;
        out.rgb.assign(3ull * W * H, 0.0f);
    }
    const int64_t x0 = std::max<int64_t>(region.x0, 0), x1 = std::min<int64_t>(region.x1, W);
    const int64_t y0 = std::max<int64_t>(region.y0, 0), y1 = std::min<int64_t>(region.y1, H);
    for (int64_t y = y0; y < y1; ++y)
        for (int64_t x = x0; x < x1; ++x) {
            const size_t i = (size_t)(y * W + x);
            const double w = m_film[8 * i + 3];
            const double rw = w == 0.0 ? 0.0 : 1.0 / w;  // HdrRgbFilm.cpp:186-193
            for (int k = 0; k < 3; ++k) out.rgb[3 * i + k] = (float)(m_film[8 * i + k] * rw);
        }
}

void GpuBneeptRenderer::retrieveFrame(std::size_t layerIndex, HdrRgbFrame& out_frame) {
    asyncPeekFrame(layerIndex, Region{0, 0, m_settings.width, m_settings.height}, out_frame);
}

void GpuBneeptRenderer::asyncPeekFrame(std::size_t layerIndex, const Region& region, HdrRgbFrame& out_frame) {
    if (layerIndex != 0) throw std::out_of_range("GpuBneeptRenderer has one layer");
    std::lock_guard<std::mutex> lock(m_filmMutex);
    develop_locked(region, out_frame);
}

RenderStats GpuBneeptRenderer::asyncQueryRenderStats() const {
    RenderStats s;
    s.paths_per_pixel = m_sppDone.load();
    s.paths_per_second = m_pathsPerSecond.load();
    return s;
}

RenderProgress GpuBneeptRenderer::asyncQueryRenderProgress() const {
    RenderProgress p;
    p.total_work = m_settings.spp;
    p.work_done = m_sppDone.load();
    p.elapsed_ms = m_elapsedMs.load();
    return p;
}

std::vector<std::string> GpuBneeptRenderer::getObservationInfo() const {
    return {"paths/pixel", "paths/second"};
}

std::size_t GpuBneeptRenderer::asyncPollUpdatedRegions(RenderRegionStatus* out_regions, std::size_t capacity) {
    if (capacity == 0 || !out_regions) return 0;
    std::lock_guard<std::mutex> rl(m_regionMutex);
    std::size_t n = 0;
    while (n < capacity && !m_regionQueue.empty()) {
        out_regions[n++] = m_regionQueue.front();
        m_regionQueue.pop_front();
    }
    return n;
}

void GpuBneeptRenderer::saveFrame(const std::string& path) {
    std::lock_guard<std::mutex> lock(m_filmMutex);
    check(phg_film_write_pfm(m_film.data(), m_settings.width, m_settings.height, path.c_str()), "phg_film_write_pfm");
}

GpuBneeptSettings GpuBneeptSettings::fromP2(const phg_p2_settings& p2) {
    GpuBneeptSettings s;
    s.width = p2.width;
    s.height = p2.height;
    s.spp = p2.spp;
    s.max_bounces = p2.max_bounces;
    s.estimator = p2.estimator;
    s.filter = p2.filter;
    s.sampler = p2.sampler;
    return s;
}

}  // namespace phg
