// Order of a job's path slots over its pixels (device and host).
#pragma once

#include <stdint.h>

namespace phg {

// Path slots of one sample of a job cover its pixels in 8 x 8 tiles -- bands of 8 rows, the
// tiles of a band left to right, rows within a tile -- so that the 64 camera rays of a wave
// leave one tile of the image (coherent traversal).  Edge bands and tiles are narrower.
// Results do not depend on the order: every sample stream is keyed by its pixel.
__host__ __device__ inline uint32_t tile_pixel(uint32_t q, uint32_t W, uint32_t H) {
    const uint32_t band = q / (8u * W);
    const uint32_t r = q - band * 8u * W;
    const uint32_t bh = (H - 8u * band < 8u ? H - 8u * band : 8u);
    uint32_t t, yi, xi;
    if (bh == 8u) {
        t = r >> 6;
        const uint32_t rr = r & 63u;
        if (8u * t + 8u <= W) { yi = rr >> 3; xi = rr & 7u; }
        else { const uint32_t tw = W - 8u * t; yi = rr / tw; xi = rr - yi * tw; }
    } else {
        t = r / (8u * bh);
        const uint32_t rr = r - t * 8u * bh, tw = (W - 8u * t < 8u ? W - 8u * t : 8u);
        yi = rr / tw; xi = rr - yi * tw;
    }
    return (8u * band + yi) * W + 8u * t + xi;
}

// Inverse of tile_pixel for the pixel at (x, y).
__host__ __device__ inline uint32_t tile_slot(uint32_t x, uint32_t y, uint32_t W, uint32_t H) {
    const uint32_t band = y >> 3, t = x >> 3;
    const uint32_t bh = (H - 8u * band < 8u ? H - 8u * band : 8u), tw = (W - 8u * t < 8u ? W - 8u * t : 8u);
    return band * 8u * W + t * 8u * bh + (y & 7u) * tw + (x & 7u);
}

}  // namespace phg
