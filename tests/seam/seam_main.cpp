// Drives the renderer seam (photon-v2_amd/host/gpu_bneept_renderer.h) the way the engine's
// front ends drive ph::Renderer: load a .p2 scene, update, render on a worker thread while the
// caller polls progress, statistics and updated regions, then save the frame.
//   seam_main <scene.p2> <out.pfm>
// Prints one line per poll and a final summary line.
#include <chrono>
#include <cstdio>
#include <thread>

#include "../../include/phg.h"
#include "../../photon-v2_amd/host/gpu_bneept_renderer.h"

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <scene.p2> <out.pfm>\n", argv[0]);
        return 2;
    }
    phg_p2_settings p2{};
    phg_scene_host* scene = nullptr;
    char err[512] = {0};
    if (phg_scene_load_p2(argv[1], &p2, &scene, err, sizeof(err)) != PHG_OK) {
        std::fprintf(stderr, "load failed: %s\n", err);
        return 1;
    }
    try {
        phg::GpuBneeptSettings s = phg::GpuBneeptSettings::fromP2(p2);
        s.chunk_spp = 4;
        s.seed = 17;
        phg::GpuBneeptRenderer r(s);
        r.doUpdate(scene->desc);
        phg_scene_free(scene);
        std::thread worker([&] { r.doRender(); });
        uint64_t polls = 0, regions = 0;
        bool finished = false;
        while (!finished) {
            phg::RenderRegionStatus reg[4];
            const std::size_t n = r.asyncPollUpdatedRegions(reg, 4);
            regions += n;
            if (n && reg[0].finished) finished = true;
            const phg::RenderProgress p = r.asyncQueryRenderProgress();
            const phg::RenderStats st = r.asyncQueryRenderStats();
            if (n) std::printf("poll %llu: %llu/%llu spp, %.0f paths/s\n", (unsigned long long)polls,
                               (unsigned long long)p.work_done, (unsigned long long)p.total_work, st.paths_per_second);
            ++polls;
            if (!finished) std::this_thread::sleep_for(std::chrono::milliseconds(1));
        }
        worker.join();
        r.saveFrame(argv[2]);
        const phg::RenderProgress p = r.asyncQueryRenderProgress();
        std::printf("done %u x %u, %llu spp, %llu region updates, estimator %u, filter %u\n", s.width, s.height,
                    (unsigned long long)p.work_done, (unsigned long long)regions, s.estimator, s.filter);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
