// Film output: develop the FP64 film (sum / weight per pixel, HdrRgbFilm::developRegion,
// Core/Filmic/HdrRgbFilm.cpp:167-207) and write it as a Portable Float Map the way
// PfmFile::save does (DataIO/PfmFile.cpp:109-145): "PF", "<width> <height>", the endianness
// indicator -1 (little endian), each on its own line, then RGB float32 triplets left to right,
// bottom row first.
#include <cstdint>
#include <cstdio>
#include <new>
#include <string>
#include <vector>

#include "../../include/phg.h"

extern "C" phg_status phg_film_develop(const double* film, uint32_t width, uint32_t height, float* out_rgb) {
    if (!film || !out_rgb || width == 0 || height == 0) return PHG_ERR_INVALID_ARG;
    const uint64_t n = (uint64_t)width * height;
    for (uint64_t i = 0; i < n; ++i) {
        const double* f = film + 8 * i;
        const double w = f[3];
        const double rw = w == 0.0 ? 0.0 : 1.0 / w;
        for (int k = 0; k < 3; ++k) out_rgb[3 * i + k] = (float)(f[k] * rw);
    }
    return PHG_OK;
}

extern "C" phg_status phg_film_write_pfm(const double* film, uint32_t width, uint32_t height, const char* path) {
    if (!film || !path || width == 0 || height == 0) return PHG_ERR_INVALID_ARG;
    std::vector<float> rgb;
    try {
        rgb.resize(3ull * width * height);
    } catch (const std::bad_alloc&) {  // nothing throws across the C ABI
        return PHG_ERR_OUT_OF_MEMORY;
    }
    phg_status st = phg_film_develop(film, width, height, rgb.data());
    if (st != PHG_OK) return st;
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return PHG_ERR_INVALID_ARG;
    const std::string header = "PF\n" + std::to_string(width) + " " + std::to_string(height) + "\n-1\n";
    bool ok = std::fwrite(header.data(), 1, header.size(), fp) == header.size();
    ok = ok && std::fwrite(rgb.data(), sizeof(float), rgb.size(), fp) == rgb.size();  // little-endian host
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? PHG_OK : PHG_ERR_INVALID_ARG;
}
