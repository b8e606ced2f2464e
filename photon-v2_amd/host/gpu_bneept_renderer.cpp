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
    m_sppPolled = 0;
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
        check(phg_render(m_ctx, &o), "phg_render");
        check(phg_film_read(m_ctx, staging.data(), staging.size()), "phg_film_read");
        {
            std::lock_guard<std::mutex> lock(m_filmMutex);
            m_film.swap(staging);
        }
        m_sppDone = s + o.spp;
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        m_elapsedMs = (uint64_t)ms;
        m_pathsPerSecond = ms > 0 ? (double)m_sppDone * m_settings.width * m_settings.height / (ms / 1e3) : 0.0;
    }
}

void GpuBneeptRenderer::develop_locked(HdrRgbFrame& out) {
    out.width = m_settings.width;
    out.height = m_settings.height;
    out.rgb.assign(3ull * out.width * out.height, 0.0f);
    for (size_t i = 0; i < (size_t)out.width * out.height; ++i) {
        const double w = m_film[8 * i + 3];
        const double rw = w == 0.0 ? 0.0 : 1.0 / w;  // HdrRgbFilm.cpp:186-193
        for (int k = 0; k < 3; ++k) out.rgb[3 * i + k] = (float)(m_film[8 * i + k] * rw);
    }
}

void GpuBneeptRenderer::retrieveFrame(std::size_t layerIndex, HdrRgbFrame& out_frame) {
    if (layerIndex != 0) throw std::out_of_range("GpuBneeptRenderer has one layer");
    std::lock_guard<std::mutex> lock(m_filmMutex);
    develop_locked(out_frame);
}

void GpuBneeptRenderer::asyncPeekFrame(std::size_t layerIndex, HdrRgbFrame& out_frame) {
    retrieveFrame(layerIndex, out_frame);
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
    const uint64_t done = m_sppDone.load();
    if (capacity == 0 || !out_regions || done == m_sppPolled.load()) return 0;
    m_sppPolled = done;
    RenderRegionStatus r;
    r.x1 = m_settings.width;
    r.y1 = m_settings.height;
    r.finished = done >= m_settings.spp;
    out_regions[0] = r;
    return 1;
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
    return s;
}

}  // namespace phg
