// Per-lane device functions of the BNEEPT path: vector math, watertight intersection,
// slab test, hit records, GGX/Lambert/Fresnel surface optics, surface escaping, emitter
// sampling.  Built with -ffp-contract=off so every float operation rounds where written,
// in the same order as the reference C++ (each function cites the file it follows).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

#define PHG_HD __host__ __device__ inline

namespace phg {

// ---- vector math in the reference's operation order (Math/TVector3.ipp) ----------------
struct V3 { float x, y, z; };
PHG_HD V3 mk(float x, float y, float z) { V3 v; v.x = x; v.y = y; v.z = z; return v; }
PHG_HD V3 operator+(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
PHG_HD V3 operator-(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
PHG_HD V3 operator*(V3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
PHG_HD V3 neg(V3 a) { return mk(-a.x, -a.y, -a.z); }
PHG_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PHG_HD float absdot(V3 a, V3 b) { return fabsf(dot(a, b)); }
PHG_HD V3 cross(V3 a, V3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
PHG_HD float length(V3 a) { return sqrtf(dot(a, a)); }
PHG_HD V3 normalize(V3 a) { const float r = 1.0f / length(a); return a * r; }
PHG_HD bool is_zero(V3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
PHG_HD bool is_finite(V3 a) { return __builtin_isfinite(a.x) && __builtin_isfinite(a.y) && __builtin_isfinite(a.z); }
PHG_HD V3 safe_normalize(V3 a, V3 fb) {
    const V3 n = normalize(a);
    return (!is_zero(n) && is_finite(n)) ? n : fb;
}
PHG_HD V3 reflect(V3 v, V3 n) { return v - n * (2.0f * dot(n, v)); }
PHG_HD V3 wsum(V3 a, float wa, V3 b, float wb, V3 c, float wc) { return a * wa + b * wb + c * wc; }
PHG_HD float clampf(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }
PHG_HD float safe_clamp(float v, float lo, float hi) { return __builtin_isfinite(v) ? clampf(v, lo, hi) : lo; }

struct RGB { float r, g, b; };
PHG_HD RGB rgb1(float v) { RGB c; c.r = v; c.g = v; c.b = v; return c; }
PHG_HD RGB rgb3(float r, float g, float b) { RGB c; c.r = r; c.g = g; c.b = b; return c; }
PHG_HD bool rgb_zero(RGB c) { return c.r == 0.0f && c.g == 0.0f && c.b == 0.0f; }
PHG_HD bool rgb_finite(RGB c) { return __builtin_isfinite(c.r) && __builtin_isfinite(c.g) && __builtin_isfinite(c.b); }
PHG_HD RGB mulc(RGB a, RGB b) { return rgb3(a.r * b.r, a.g * b.g, a.b * b.b); }
PHG_HD RGB muls(RGB a, float s) { return rgb3(a.r * s, a.g * s, a.b * s); }
PHG_HD RGB rational_clamp(RGB c) {  // BNEEPTEstimator.cpp:264-268
    return rgb3(safe_clamp(c.r, 0.0f, 1e9f), safe_clamp(c.g, 0.0f, 1e9f), safe_clamp(c.b, 0.0f, 1e9f));
}
PHG_HD float luminance(RGB c) {  // linear-sRGB Y (color_spaces.ipp:129)
    return c.r * 0.21263900587151027f + c.g * 0.715168678767756f + c.b * 0.07219231536073371f;
}

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = 6.28318530717958647692f;
constexpr float kRcpPi = 0.31830988618379067154f;
constexpr float kRcpSqrt2 = 0.70710678118654752440f;
constexpr float kFltMax = 3.40282346638528859812e+38f;
constexpr float kFltMin = 1.17549435082228750797e-38f;

// ---- sample stream: PCG32 (Math/Random/Pcg32.h:72-124), moremur (Math/hash.ipp:104-115) --
PHG_HD uint64_t moremur(uint64_t v) {
    v ^= v >> 27; v *= 0x3C79AC492BA7B653ULL;
    v ^= v >> 33; v *= 0x1C69B3F74AC4AE35ULL;
    v ^= v >> 27;
    return v;
}
constexpr uint64_t kPcgMul = 0x5851F42D4C957F2DULL;

struct Stream {
    uint64_t state, inc;
    PHG_HD uint32_t next() {
        const uint64_t old = state;
        state = old * kPcgMul + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31));
    }
    // odd multiples of 2^-24: strictly inside (0, 1)
    PHG_HD float u() { const uint32_t b = next(); return (float)(((b >> 9) << 1) | 1u) * 0x1p-24f; }
};
PHG_HD uint64_t path_key(uint64_t seed, uint64_t pixel, uint64_t sample, uint64_t npix) {
    return (seed << 40) ^ (sample * npix + pixel);
}
PHG_HD Stream stream_init(uint64_t key) {
    Stream s; s.state = 0u; s.inc = (key << 1u) | 1u;
    s.next(); s.state += moremur(key); s.next();
    return s;
}
PHG_HD Stream stream_resume(uint64_t key, uint64_t state) {
    Stream s; s.state = state; s.inc = (key << 1u) | 1u; return s;
}

// ---- stratified sample source (Core/SampleGenerator/SGStratified.cpp:16-154) -------------
// A batch of the reference's stratified source is one sample of every pixel: its raster stage
// has one stratum per pixel (a uniform offset in the pixel, as the independent source draws),
// and its 5-D ray stage -- the first five SampleFlow dimensions the estimator reads at the
// first surface (ReceiverSamplingWork.cpp:76-84, SampleFlow.h:105-169) -- is generated as a
// 2-D, a 2-D and a 1-D stage (SampleGenerator.cpp:81-135: SGStratified has no >= 3-D
// generator), each a set of jittered strata handed to the batch's N samples in shuffled order
// (SGStratified.cpp:24-90): k x k strata with k = floor(sqrt N), x fastest, in whole passes and
// the rest uniform (reviseDimSizeHints, SGStratified.cpp:112-152), or N strata in 1-D.
// Here the shuffle is strat_permute, a keyed bijection of the pixels (the reference shuffles
// with its thread-local random source), and a jitter is the path's own stream value.  In this
// renderer's stream order at the first vertex the stages are (u0 emitter pick, u2 light point
// x), (u3 light point y, u4 BSDF direction x) and u5 (BSDF direction y); the triangle pick u1
// is drawn from the reference's global random source (GroupedDiffuseSurfaceEmitter.cpp:68-69)
// and stays independent, as do all later dimensions (SampleFlow.h:157-169).

// Bijection of [0, n) keyed by key: an 8-round Feistel network on the smallest even number of
// bits that holds n (round function: moremur of key, round and half), walked until the image
// falls inside [0, n) (at most 4 steps on average), then rotated by a key-derived offset so
// that every index's image is uniform over keys (fewer rounds leave pairs of images visibly
// non-uniform on small n).
PHG_HD uint32_t strat_permute(uint32_t i, uint32_t n, uint64_t key) {
    if (n <= 1u) return 0u;
    uint32_t bits = 0;
    while (bits < 32u && ((n - 1u) >> bits) != 0u) ++bits;
    bits += bits & 1u;
    const uint32_t half = bits / 2u, mask = (1u << half) - 1u;
    uint32_t x = i;
    do {
        uint32_t L = x >> half, R = x & mask;
        for (uint32_t r = 0; r < 8u; ++r) {
            const uint32_t F = (uint32_t)moremur(key ^ ((uint64_t)r << 32) ^ R) & mask;
            const uint32_t nL = R;
            R = L ^ F;
            L = nL;
        }
        x = (L << half) | R;
    } while (x >= n);
    const uint32_t off = (uint32_t)(((moremur(key ^ 0xD1B54A32D192ED03ULL) >> 32) * (uint64_t)n) >> 32);
    x += off;
    return x >= n ? x - n : x;
}

// key of stratified stage `stage` (0, 1: the 2-D stages, 2: the 1-D stage) of sample batch
// `sample` of a job with stream seed `seed` (< 2^24)
PHG_HD uint64_t strat_key(uint64_t seed, uint32_t sample, uint32_t stage) {
    return moremur((seed << 40) ^ ((uint64_t)sample << 2) ^ (uint64_t)stage ^ 0x9E3779B97F4A7C15ULL);
}

// (c + u) / k rounded to float, kept below 1 (u in (0, 1) is the stratum's jitter)
PHG_HD float strat_value(uint32_t c, uint32_t k, float u) {
    const float v = (float)(((double)c + (double)u) / (double)k);
    return v < 1.0f ? v : 0x1.fffffep-1f;
}

PHG_HD uint32_t strat_isqrt(uint32_t n) {
    uint32_t k = (uint32_t)sqrt((double)n);
    while ((uint64_t)k * k > n) --k;
    while ((uint64_t)(k + 1u) * (k + 1u) <= n) ++k;
    return k;
}

// dimension `axis` (0: x, 1: y) of 2-D stage `stage` for the sample of `pixel` among npix
PHG_HD float strat_2d(uint32_t pixel, uint32_t npix, uint64_t seed, uint32_t sample, uint32_t stage, uint32_t axis,
                      float u) {
    const uint32_t k = strat_isqrt(npix), per = k * k, passes = npix / per;
    const uint32_t s = strat_permute(pixel, npix, strat_key(seed, sample, stage));
    if (s >= passes * per) return u;   // beyond the whole passes: uniform
    const uint32_t c = s % per;
    return strat_value(axis == 0u ? c % k : c / k, k, u);
}

PHG_HD float strat_1d(uint32_t pixel, uint32_t npix, uint64_t seed, uint32_t sample, float u) {
    return strat_value(strat_permute(pixel, npix, strat_key(seed, sample, 2u)), npix, u);
}

// ---- Halton sample source (Core/SampleGenerator/SGHalton.cpp:22-206, Halton/halton.h,
// Math/Random/Quasi/radical_inverse.h) -------------------------------------------------------
// Sample si of batch b (a batch is one raster sample per pixel, si the pixel index) is Halton
// point b N + si; its dimension d is the radical inverse in the d-th prime base of
// start_d + (b N + si) leap.  Dimensions 0, 1 (bases 2, 3) place the raster sample over the
// whole sample window; dimensions 2..6 (bases 5, 7, 11, 13, 17) are the ray stage's five, read
// at the first surface as u0, u2, u3, u4, u5 (the stratified source's stages), each through the
// chosen digit permutation.  Values are FP32 as the reference computes them (a float scale
// that shrinks by 1/base per digit times the reversed digits converted to float, clamped to
// [0, 1]); a value that rounds to 1 is used as the largest float below 1.
constexpr uint32_t kHaltonBase[7] = {2u, 3u, 5u, 7u, 11u, 13u, 17u};
enum : uint32_t { kHaltonPermFixed = 0, kHaltonPermNone = 1, kHaltonPermPerDigit = 2, kHaltonPermOwen = 3 };
enum : uint32_t { kHaltonSeqOriginal = 0, kHaltonSeqLeap = 1, kHaltonSeqRandomLeap = 2, kHaltonSeqRandomStart = 3 };
PHG_HD uint32_t halton_perm(uint32_t sampler) { return (sampler >> 8) & 3u; }
PHG_HD uint32_t halton_seq(uint32_t sampler) { return (sampler >> 12) & 3u; }

// Kensler's hashed permutation ("Correlated Multi-Jittered Sampling", 2013; the reference's
// math::permuted_index, Math/hash.ipp:206-243): element i of a permutation of [0, l) keyed by p
PHG_HD uint32_t halton_permute(uint32_t i, uint32_t l, uint32_t p) {
    uint32_t w = l - 1u;
    w |= w >> 1; w |= w >> 2; w |= w >> 4; w |= w >> 8; w |= w >> 16;
    do {
        i ^= p; i *= 0xe170893du; i ^= p >> 16; i ^= (i & w) >> 4;
        i ^= p >> 8; i *= 0x0929eb3fu; i ^= p >> 23; i ^= (i & w) >> 1;
        i *= 1u | p >> 27; i *= 0x6935fa69u; i ^= (i & w) >> 11; i *= 0x74dcb303u;
        i ^= (i & w) >> 2; i *= 0x9e501cc3u; i ^= (i & w) >> 2; i *= 0xc860a3dfu;
        i &= w; i ^= i >> 5;
    } while (i >= l);
    return (i + p) % l;
}

// digits of the base of each dimension that can still change a float in [0, 1): the count n
// of the loop "while 1 - (b - 1) s < 1: n += 1, s *= 1 / b" in FP32 with every operation
// rounded (radical_inverse.h num_meaningful_digits); a table, since a fused multiply-add
// would change the loop's last comparison
constexpr uint32_t kHaltonDigits[7] = {25u, 17u, 12u, 10u, 9u, 8u, 8u};

PHG_HD float halton_below_one(float v) { return v < 1.0f ? v : 0x1.fffffep-1f; }

// unpermuted radical inverse of v in base b (the raster dimensions, and every dimension with
// permutation "none")
PHG_HD float halton_ri(uint32_t b, uint64_t v) {
    if (b == 2u) return (float)__builtin_bitreverse64(v) * 0x1p-64f;
    const float rcp = 1.0f / (float)b;
    const uint64_t safe = (~0ull - (b - 1u)) / b;
    uint64_t rev = 0;
    float scale = 1.0f;
    while (v > 0u && rev <= safe) {
        const uint64_t q = v / b;
        rev = rev * b + (v - q * b);
        scale *= rcp;
        v = q;
    }
    return fminf(fmaxf((float)rev * scale, 0.0f), 1.0f);
}

// key of the keyed digit permutations: dimension d (and digit position k for per-digit) of a
// job with stream seed `seed`
PHG_HD uint32_t halton_key(uint64_t seed, uint32_t d, uint32_t k) {
    return (uint32_t)(moremur((seed << 24) ^ ((uint64_t)d << 12) ^ (uint64_t)k ^ 0xA0761D6478BD642FULL) >> 32);
}

// permuted radical inverse of v in the base of dimension d (d >= 2): exactly kHaltonDigits[d]
// digits, each passed through the permutation -- fixed: one keyed permutation of the digit
// values per dimension; per-digit: one per dimension and digit position; owen: keyed by the
// dimension and every digit already emitted (halton.h OwenScrambledPermuter)
PHG_HD float halton_ri_perm(uint32_t d, uint64_t v, uint32_t perm, uint64_t seed) {
    const uint32_t b = kHaltonBase[d];
    const float rcp = 1.0f / (float)b;
    const uint64_t safe = (~0ull - (b - 1u)) / b;
    const uint32_t ndig = kHaltonDigits[d];
    const uint32_t fixed_key = halton_key(seed, d, 0xFFFu);
    uint64_t rev = 0, owen_hash = moremur(((uint64_t)d << 4) + 1u), owen_digits = 0;
    float scale = 1.0f;
    for (uint32_t k = 0; k < ndig && rev <= safe; ++k) {
        const uint64_t q = v / b;
        const uint32_t digit = (uint32_t)(v - q * b);
        uint32_t pd;
        if (perm == kHaltonPermOwen) {
            owen_hash = moremur(owen_hash ^ owen_digits);
            pd = halton_permute(digit, b, (uint32_t)owen_hash);
            owen_digits = owen_digits * b + pd;
        } else if (perm == kHaltonPermPerDigit) {
            pd = halton_permute(digit, b, halton_key(seed, d, k));
        } else {
            pd = halton_permute(digit, b, fixed_key);
        }
        rev = rev * b + pd;
        scale *= rcp;
        v = q;
    }
    return fminf(fmaxf((float)rev * scale, 0.0f), 1.0f);
}

// dimension d (0..6) of Halton point `index` (b N + si) of a job: sampler carries the
// permutation and sequence, leap the index step (1 but for the leap sequences)
PHG_HD float halton_value(uint32_t d, uint64_t index, uint32_t sampler, uint32_t leap, uint64_t seed) {
    uint64_t v = index * (uint64_t)leap;
    if (halton_seq(sampler) == kHaltonSeqRandomStart)
        v += moremur((seed << 24) ^ ((uint64_t)d << 8) ^ 0x2545F4914F6CDD1DULL);
    const uint32_t perm = halton_perm(sampler);
    const float x = (d < 2u || perm == kHaltonPermNone) ? halton_ri(kHaltonBase[d], v) : halton_ri_perm(d, v, perm, seed);
    return halton_below_one(x);
}

// ---- watertight ray/triangle (Math/Geometry/TWatertightTriangle.ipp:16-149) -------------
PHG_HD bool wt_intersect(V3 va, V3 vb, V3 vc, V3 org, V3 dir, float minT, float maxT,
                                             float* out_t, float* out_a, float* out_b, float* out_c) {
    V3 d = dir;
    V3 a = va - org, b = vb - org, c = vc - org;
    if (fabsf(d.x) > fabsf(d.y)) {
        if (fabsf(d.x) > fabsf(d.z)) {
            d = mk(d.y, d.z, d.x); a = mk(a.y, a.z, a.x); b = mk(b.y, b.z, b.x); c = mk(c.y, c.z, c.x);
        }
    } else {
        if (fabsf(d.y) > fabsf(d.z)) {
            d = mk(d.z, d.x, d.y); a = mk(a.z, a.x, a.y); b = mk(b.z, b.x, b.y); c = mk(c.z, c.x, c.y);
        }
    }
    const float rz = 1.0f / d.z;
    const float sx = -d.x * rz;
    const float sy = -d.y * rz;
    const float sz = rz;
    a.x += sx * a.z; a.y += sy * a.z;
    b.x += sx * b.z; b.y += sy * b.z;
    c.x += sx * c.z; c.y += sy * c.z;
    float ea = b.x * c.y - b.y * c.x;
    float eb = c.x * a.y - c.y * a.x;
    float ec = a.x * b.y - a.y * b.x;
    if (ea == 0.0f || eb == 0.0f || ec == 0.0f) {
        const double ea64 = (double)b.x * (double)c.y - (double)b.y * (double)c.x;
        const double eb64 = (double)c.x * (double)a.y - (double)c.y * (double)a.x;
        const double ec64 = (double)a.x * (double)b.y - (double)a.y * (double)b.x;
        ea = (float)ea64; eb = (float)eb64; ec = (float)ec64;
    }
    if ((ea < 0.0f || eb < 0.0f || ec < 0.0f) && (ea > 0.0f || eb > 0.0f || ec > 0.0f)) return false;
    const float det = ea + eb + ec;
    if (det == 0.0f || !__builtin_isfinite(det)) return false;
    a.z *= sz; b.z *= sz; c.z *= sz;
    const float ts = ea * a.z + eb * b.z + ec * c.z;
    if (det > 0.0f) {
        if (ts < minT * det || ts > maxT * det) return false;
    } else {
        if (ts > minT * det || ts < maxT * det) return false;
    }
    const float rd = 1.0f / det;
    *out_a = ea * rd; *out_b = eb * rd; *out_c = ec * rd;
    *out_t = ts * rd;
    return true;
}

PHG_HD void load_tri(const TriRec* tris, uint32_t i, V3* a, V3* b, V3* c) {
    const float4* p = reinterpret_cast<const float4*>(tris + i);
    const float4 q0 = p[0], q1 = p[1], q2 = p[2];
    *a = mk(q0.x, q0.y, q0.z); *b = mk(q0.w, q1.x, q1.y); *c = mk(q1.z, q1.w, q2.x);
}

// ---- hit records (PTriangle::calcHitDetail PTriangle.cpp:66-135, HitInfo.h:140-170) -----
struct SurfaceHit {
    uint32_t tri;
    V3 P, Ng, Ns, bx, bz;  // shading basis: x = bx, y = Ns, z = bz
};

PHG_HD V3 face_normal(V3 a, V3 b, V3 c) {  // TBasicTriangle.ipp:116-121, 268-275
    const V3 cr = cross(b - a, c - a);
    const float cf = dot(cr, cr);
    const bool degenerate = !(cf > 0.0f) || __builtin_isinf(cf);
    return !degenerate ? safe_normalize(cr, mk(0.0f, 1.0f, 0.0f)) : mk(0.0f, 1.0f, 0.0f);
}
PHG_HD float tri_area(V3 a, V3 b, V3 c) { return length(cross(b - a, c - a)) * 0.5f; }

// TOrthonormalBasis3.ipp:16-39 makeFromUnitY
PHG_HD void basis_from_unit_y(V3 y, V3* bx, V3* bz) {
    V3 x;
    if (fabsf(y.y) < kRcpSqrt2) {
        x = mk(-y.z, 0.0f, y.x);
        x = x * (1.0f / sqrtf(x.x * x.x + x.z * x.z));
    } else {
        x = mk(y.y, -y.x, 0.0f);
        x = x * (1.0f / sqrtf(x.x * x.x + x.y * x.y));
    }
    *bx = x;
    *bz = cross(x, y);
}

PHG_HD SurfaceHit make_hit(V3 a, V3 b, V3 c, uint32_t tri, float ba, float bb, float bc) {
    SurfaceHit h;
    h.tri = tri;
    h.P = wsum(a, ba, b, bb, c, bc);
    h.Ng = face_normal(a, b, c);
    h.Ns = normalize(wsum(h.Ng, ba, h.Ng, bb, h.Ng, bc));
    basis_from_unit_y(h.Ns, &h.bx, &h.bz);
    return h;
}
PHG_HD V3 to_local(const SurfaceHit& h, V3 w) { return mk(dot(h.bx, w), dot(h.Ns, w), dot(h.bz, w)); }
PHG_HD V3 to_world(const SurfaceHit& h, V3 l) { return h.bx * l.x + h.Ns * l.y + h.bz * l.z; }

// Strict sidedness (Core/LTA/SidednessAgreement.h:104-283)
PHG_HD bool side_agreed(const SurfaceHit& X, V3 v) { return dot(X.Ng, v) * dot(X.Ns, v) > 0.0f; }
PHG_HD bool same_hemi(const SurfaceHit& X, V3 a, V3 b) {
    return side_agreed(X, a) && side_agreed(X, b) && dot(a, X.Ng) * dot(b, X.Ng) > 0.0f;
}
PHG_HD bool opposite_hemi(const SurfaceHit& X, V3 a, V3 b) {
    return side_agreed(X, a) && side_agreed(X, b) && dot(a, X.Ng) * dot(b, X.Ng) < 0.0f;
}

// ---- GGX (Core/SurfaceBehavior/Property/{IsoTrowbridgeReitz,TrowbridgeReitz,ShapeInvariantMicrofacet}) --
PHG_HD float mat_alpha(const phg_material& m) {
    const float a = m.roughness * m.roughness;
    return a > 0.001f ? a : 0.001f;
}
PHG_HD float ggx_lambda(V3 N, V3 w, float alpha) {
    const float a2 = alpha * alpha;
    const float nod = dot(N, w);
    const float nod2 = fminf(nod * nod, 1.0f);
    return 0.5f * (-1.0f + sqrtf(1.0f + a2 * (1.0f / nod2 - 1.0f)));
}
PHG_HD float ggx_D(V3 N, V3 H, float alpha) {
    const float noh = dot(N, H);
    const float a2 = alpha * alpha;
    const float noh2 = noh * noh;
    const float inner = noh2 * (a2 - 1.0f) + 1.0f;
    const float den = kPi * inner * inner;
    return a2 / den;
}
PHG_HD float ggx_G(V3 N, V3 H, V3 L, V3 V, float alpha) {
    const float nol = dot(N, L), nov = dot(N, V), hol = dot(H, L), hov = dot(H, V);
    if (!(hol * nol > 0.0f && hov * nov > 0.0f)) return 0.0f;
    const float ll = ggx_lambda(N, L, alpha), lv = ggx_lambda(N, V, alpha);
    return 1.0f / (1.0f + ll + lv);
}
PHG_HD float ggx_vndf_pdf(V3 N, V3 H, V3 V, float alpha) {
    const float lv = ggx_lambda(N, V, alpha);
    const float g1v = 1.0f / (1.0f + lv);
    const float ndf = ggx_D(N, H, alpha);
    return g1v * ndf * fabsf(dot(H, V) / dot(N, V));
}
PHG_HD V3 ggx_sample_vh(const SurfaceHit& X, V3 N, V3 V, float alpha, float u0, float u1) {
    const V3 upperV = dot(V, N) >= 0.0f ? V : neg(V);
    const V3 lv = to_local(X, upperV);
    const V3 sv = normalize(mk(lv.x * alpha, lv.y, lv.z * alpha));
    const float phi = kTwoPi * u0;
    const float y = fmaf(1.0f - u1, 1.0f + sv.y, -sv.y);
    const float st = sqrtf(clampf(1.0f - y * y, 0.0f, 1.0f));
    const float z = st * cosf(phi);
    const float x = st * sinf(phi);
    const V3 h = mk(x, y, z) + sv;
    const V3 lh = mk(h.x * alpha, h.y, h.z * alpha);
    return safe_normalize(to_world(X, lh), N);
}

// Fresnel (SchlickApproxConductorFresnel.cpp:35-43, ExactDielectricFresnel.cpp:12-41,
// DielectricFresnel.cpp:18-42)
PHG_HD RGB fresnel_schlick(const phg_material& m, float cosI) {
    const float c = fabsf(cosI);
    // the reference's std::pow(float, int) evaluates in double; (1 - c)^2 is exact in double
    // and two more roundings leave the float result equal to the rounded pow except within a
    // few double ulps of a float rounding boundary -- without a double-precision pow (its
    // 64-bit temporaries set the shading kernel's register peak)
    const double x = (double)(1.0f - c), x2 = x * x;
    const float p = (float)(x2 * x2 * x);
    return rgb3((1.0f - m.rgb[0]) * p + m.rgb[0], (1.0f - m.rgb[1]) * p + m.rgb[1], (1.0f - m.rgb[2]) * p + m.rgb[2]);
}
// ExactConductorFresnel.cpp:66-101 per channel, with setIors' precomputation (:103-116):
// en2 = (n / eta_outer)^2, ek2 = (k / eta_outer)^2 (squares as products).
PHG_HD float fresnel_conductor_exact1(float etaOuter, float n, float k, float cosThetaI) {
    const float en = n / etaOuter, ek = k / etaOuter;
    const float en2 = en * en, ek2 = ek * ek;
    const float e2_sub_k2 = en2 - ek2, four_e2k2 = en2 * ek2 * 4.0f;
    const float cosI = fabsf(cosThetaI);
    const float cosI2 = cosI * cosI;
    const float sinI2 = 1.0f - cosI * cosI;
    const float t0 = e2_sub_k2 - sinI2;
    const float a2plusb2 = sqrtf(t0 * t0 + four_e2k2);
    const float a = sqrtf((a2plusb2 + t0) * 0.5f);
    const float t1 = a2plusb2 + cosI2;
    const float t2 = a * (2.0f * cosI);
    const float t3 = a2plusb2 * cosI2 + sinI2 * sinI2;
    const float t4 = t2 * sinI2;
    const float Rs = (t1 - t2) / (t1 + t2);
    const float Rp = Rs * ((t3 - t4) / (t3 + t4));
    return (Rs + Rp) * 0.5f;
}
// the conductor's Fresnel term as its interface info selects it (ConductiveInterfaceInfo.cpp)
PHG_HD RGB fresnel_conductor(const phg_material& m, float cosI) {
    if (m.fresnel == PHG_FRESNEL_EXACT)
        return rgb3(fresnel_conductor_exact1(m.ior_outer, m.n[0], m.k[0], cosI),
                    fresnel_conductor_exact1(m.ior_outer, m.n[1], m.k[1], cosI),
                    fresnel_conductor_exact1(m.ior_outer, m.n[2], m.k[2], cosI));
    return fresnel_schlick(m, cosI);
}
PHG_HD float fresnel_dielectric(float etaOuter, float etaInner, float cosThetaI) {
    float etaI = etaOuter, etaT = etaInner, cosI = cosThetaI;
    if (cosThetaI < 0.0f) { const float t = etaI; etaI = etaT; etaT = t; cosI = fabsf(cosI); }
    const float ratio = etaI / etaT;
    const float sinT2 = (1.0f - cosI * cosI) * ratio * ratio;
    if (sinT2 >= 1.0f) return 1.0f;
    const float cosT = sqrtf(1.0f - sinT2);
    const float rpar = (etaT * cosI - etaI * cosT) / (etaT * cosI + etaI * cosT);
    const float rper = (etaI * cosI - etaT * cosT) / (etaI * cosI + etaT * cosT);
    return 0.5f * (rpar * rpar + rper * rper);
}
PHG_HD bool refract_dir(float etaOuter, float etaInner, V3 I, V3 N, V3* out) {
    const float ion = dot(I, N);
    const float sgn = ion < 0.0f ? -1.0f : 1.0f;
    const float ratio = sgn < 0.0f ? etaInner / etaOuter : etaOuter / etaInner;
    const float sq = 1.0f - ratio * ratio * (1.0f - ion * ion);
    if (sq <= 0.0f) return false;
    const float nf = ratio * ion - sgn * sqrtf(sq);
    const float iff = -ratio;
    *out = normalize(N * nf + I * iff);
    return true;
}
PHG_HD float reflect_prob(float F) { return clampf(F, 0.0001f, 1.0f - 0.0001f); }

// BsdfHelper.h:14-45
PHG_HD bool half_vec_same_hemi(V3 L, V3 V, V3 N, V3* H) {
    V3 h = L + V;
    if (is_zero(h)) return false;
    h = normalize(h);
    const float nh = dot(N, h);
    const float s = (float)((0.0f < nh) - (0.0f > nh));
    h = h * s;
    *H = h;
    return !is_zero(h);
}

struct BsdfEval { bool ok; RGB f; };
struct BsdfSample { bool ok; V3 L; RGB w; float relIor; };

// SurfaceOptics::calcBsdf + LambertianReflector/OpaqueMicrofacet/TranslucentMicrofacet::calcBsdf
PHG_HD BsdfEval bsdf_eval(const phg_material& m, const SurfaceHit& X, V3 L, V3 V) {
    BsdfEval e; e.ok = false; e.f = rgb1(0.0f);
    if (!side_agreed(X, L) || !side_agreed(X, V)) return e;
    const V3 N = X.Ns;
    if (m.type == PHG_MAT_LAMBERT) {
        if (!same_hemi(X, L, V)) return e;
        e.f = muls(rgb3(m.rgb[0], m.rgb[1], m.rgb[2]), kRcpPi);
        e.ok = rgb_finite(e.f);
    } else if (m.type == PHG_MAT_CONDUCTOR) {
        const float nol = dot(N, L), nov = dot(N, V);
        if (!same_hemi(X, L, V)) return e;
        V3 H;
        if (!half_vec_same_hemi(L, V, N, &H)) return e;
        const float hol = dot(H, L);
        const float a = mat_alpha(m);
        const RGB F = fresnel_conductor(m, hol);
        const float D = ggx_D(N, H, a);
        const float G = ggx_G(N, H, L, V, a);
        e.f = muls(F, D * G / (4.0f * fabsf(nov * nol)));
        e.ok = rgb_finite(e.f);
    } else if (m.type == PHG_MAT_DIELECTRIC) {
        const float nol = dot(N, L), nov = dot(N, V);
        const float nolnov = nol * nov;
        if (nolnov == 0.0f) return e;
        const float a = mat_alpha(m);
        if (same_hemi(X, L, V)) {
            V3 H;
            if (!half_vec_same_hemi(L, V, N, &H)) return e;
            const float hol = dot(H, L);
            const float F = fresnel_dielectric(m.ior_outer, m.ior_inner, hol);
            const float D = ggx_D(N, H, a);
            const float G = ggx_G(N, H, L, V, a);
            e.f = rgb1(F * (D * G / (4.0f * fabsf(nolnov))));
            e.ok = rgb_finite(e.f);
        } else if (opposite_hemi(X, L, V)) {
            float etaI = m.ior_outer, etaT = m.ior_inner;
            if (nol < 0.0f) { const float t = etaI; etaI = etaT; etaT = t; }
            V3 H = L * (-etaI) + V * (-etaT);
            if (is_zero(H)) return e;
            H = normalize(H);
            if (dot(N, H) < 0.0f) H = H * -1.0f;
            const float hov = dot(H, V);
            const float hol = dot(H, L);
            const float F = 1.0f - fresnel_dielectric(m.ior_outer, m.ior_inner, hol);
            const float D = ggx_D(N, H, a);
            const float G = ggx_G(N, H, L, V, a);
            const float transport = etaT / etaI;
            const float iorTerm = transport * etaI / (etaI * hol + etaT * hov);
            if (!__builtin_isfinite(iorTerm)) return e;
            const float dotTerm = fabsf(hol * hov / nolnov);
            e.f = rgb1(F * (D * G * dotTerm * (iorTerm * iorTerm)));
            e.ok = rgb_finite(e.f);
        }
    }
    return e;
}

// SurfaceOptics::calcBsdfPdf + material calcBsdfPdf; a result that is not finite and > 0
// means "no pdf" (lta::PDF::operator bool, Core/LTA/PDF.h).
PHG_HD float bsdf_pdf(const phg_material& m, const SurfaceHit& X, V3 L, V3 V) {
    if (!side_agreed(X, L) || !side_agreed(X, V)) return 0.0f;
    const V3 N = X.Ns;
    if (m.type == PHG_MAT_LAMBERT) {
        if (!same_hemi(X, L, V)) return 0.0f;
        return absdot(L, N) * kRcpPi;
    } else if (m.type == PHG_MAT_CONDUCTOR) {
        if (!same_hemi(X, L, V)) return 0.0f;
        V3 H;
        if (!half_vec_same_hemi(L, V, N, &H)) return 0.0f;
        const float pdf = ggx_vndf_pdf(N, H, V, mat_alpha(m));
        return pdf / (4.0f * absdot(H, L));
    } else if (m.type == PHG_MAT_DIELECTRIC) {
        const float a = mat_alpha(m);
        if (same_hemi(X, L, V)) {
            V3 H;
            if (!half_vec_same_hemi(L, V, N, &H)) return 0.0f;
            const float F = fresnel_dielectric(m.ior_outer, m.ior_inner, dot(H, L));
            const float rp = reflect_prob(F);
            const float pdf = ggx_vndf_pdf(N, H, V, a);
            return pdf / (4.0f * absdot(H, L)) * rp;
        } else if (opposite_hemi(X, L, V)) {
            const float nol = dot(N, L);
            float etaI = m.ior_outer, etaT = m.ior_inner;
            if (nol < 0.0f) { const float t = etaI; etaI = etaT; etaT = t; }
            V3 H = L * (-etaI) + V * (-etaT);
            if (is_zero(H)) return 0.0f;
            H = normalize(H);
            if (dot(N, H) < 0.0f) H = H * -1.0f;
            const float hov = dot(H, V), hol = dot(H, L);
            const float iorTerm = etaI * hol + etaT * hov;
            const float mult = fabsf(etaI * etaI * hol) / (iorTerm * iorTerm);
            const float F = fresnel_dielectric(m.ior_outer, m.ior_inner, hol);
            const float tp = 1.0f - reflect_prob(F);
            const float pdf = ggx_vndf_pdf(N, H, V, a);
            return pdf * mult * tp;
        }
        return 0.0f;
    }
    return 0.0f;
}
PHG_HD bool pdf_ok(float p) { return __builtin_isfinite(p) && p > 0.0f; }
PHG_HD bool measurable(RGB w, float cosv) { return rgb_finite(w) && 0.0f < cosv && cosv < 1.1f; }

// SurfaceOptics::genBsdfSample + material genBsdfSample.  u0, u1: direction; u2: lobe pick.
PHG_HD BsdfSample bsdf_sample(const phg_material& m, const SurfaceHit& X, V3 V, float u0, float u1, float u2) {
    BsdfSample s; s.ok = false; s.L = mk(0.0f, 0.0f, 0.0f); s.w = rgb1(0.0f); s.relIor = 1.0f;
    if (!side_agreed(X, V)) return s;
    const V3 N = X.Ns;
    if (m.type == PHG_MAT_LAMBERT) {  // LambertianReflector.cpp:54-84, THemisphere.h:104-119
        const float phi = kTwoPi * u0;
        const float yv = sqrtf(u1);
        const float yr = sqrtf(1.0f - yv * yv);
        V3 L = mk(sinf(phi) * yr, yv, cosf(phi) * yr);
        L = normalize(to_world(X, L));
        if (dot(V, N) < 0.0f) L = L * -1.0f;
        s.w = rgb3(m.rgb[0], m.rgb[1], m.rgb[2]);
        s.ok = measurable(s.w, absdot(N, L));
        s.L = L;
    } else if (m.type == PHG_MAT_CONDUCTOR) {  // OpaqueMicrofacet.cpp:93-123
        const float a = mat_alpha(m);
        const V3 H = ggx_sample_vh(X, N, V, a, u0, u1);
        const V3 L = normalize(reflect(neg(V), H));
        const float nov = dot(N, V);
        const float hol = dot(H, L);
        const float dotTerms = fabsf(hol / nov);
        const RGB F = fresnel_conductor(m, hol);
        const float G = ggx_G(N, H, L, V, a);
        const float D = ggx_D(N, H, a);
        const float pdf = ggx_vndf_pdf(N, H, V, a);
        s.w = muls(F, G * D * dotTerms / pdf);
        s.ok = measurable(s.w, absdot(N, L));
        s.L = L;
    } else if (m.type == PHG_MAT_DIELECTRIC) {  // TranslucentMicrofacet.cpp:154-273
        const float a = mat_alpha(m);
        const V3 H = ggx_sample_vh(X, N, V, a, u0, u1);
        float F = fresnel_dielectric(m.ior_outer, m.ior_inner, dot(H, V));
        const float rp = reflect_prob(F);
        V3 L;
        if (u2 < rp) {
            L = normalize(reflect(neg(V), H));
            if (!same_hemi(X, V, L)) return s;
            F = F / rp;
        } else {
            V3 R;
            if (!refract_dir(m.ior_outer, m.ior_inner, V, H, &R) || !opposite_hemi(X, V, R)) return s;
            L = R;
            F = 1.0f - fresnel_dielectric(m.ior_outer, m.ior_inner, dot(H, L));
            float etaI = m.ior_outer, etaT = m.ior_inner;
            if (dot(N, L) < 0.0f) { const float t = etaI; etaI = etaT; etaT = t; }
            const float rel = etaT / etaI;
            s.relIor = rel;
            F = F * (rel * rel);
            F = F / (1.0f - rp);
        }
        const float pdf = ggx_vndf_pdf(N, H, V, a);
        const float nov = dot(N, V), hov = dot(H, V);
        const float dotTerms = fabsf(hov / nov);
        const float G = ggx_G(N, H, L, V, a);
        const float D = ggx_D(N, H, a);
        s.w = rgb1(F * (G * D * dotTerms / pdf));
        s.ok = measurable(s.w, absdot(N, L));
        s.L = L;
    } else {
        return s;
    }
    if (s.ok && !side_agreed(X, s.L)) s.ok = false;
    return s;
}

// ---- escaping a surface (Core/LTA/SurfaceHitRefinery.cpp:62-263, iterative mode) --------
constexpr float kMeanErr = 5e-8f;
constexpr float kMaxErr = 5e-8f * 5e2f;
constexpr int kMaxEscapeIters = 256;

PHG_HD bool reintersect(V3 a, V3 b, V3 c, V3 org, V3 dir, float minT, float maxT, float* t) {
    float x, y, z;
    return wt_intersect(a, b, c, org, dir, minT, maxT, t, &x, &y, &z);
}

struct TriVerts { V3 a, b, c; };

PHG_HD void iterative_offset(const SurfaceHit& X, const TriVerts& tv, V3 dir, V3* out_offset, V3* out_dir) {
    const V3 N = X.Ng;
    const V3 offsetDir = dot(N, dir) > 0.0f ? N : neg(N);
    const V3 escDir = normalize(dir);
    const float escCos = clampf(dot(escDir, offsetDir), 0.0f, 1.0f);
    float init = length(X.P) * kMeanErr;
    if (!__builtin_isfinite(init)) init = fmaxf(length(X.P) * 1e-6f, 2e-4f);
    init = init + kFltMin;
    float maxD = init, minD = 0.0f;
    for (int it = 0; it < kMaxEscapeIters; ++it) {
        float t;
        if (!reintersect(tv.a, tv.b, tv.c, X.P + offsetDir * maxD, escDir, 0.0f, kFltMax, &t)) break;
        maxD += t * escCos;
        minD = maxD;
        maxD *= 2.0f;
    }
    int nref = 0;
    for (int it = 0; it < kMaxEscapeIters && nref < 1; ++it) {
        const float mid = (minD + maxD) * 0.5f;
        if (!(minD < mid && mid < maxD)) break;
        float t;
        if (!reintersect(tv.a, tv.b, tv.c, X.P + offsetDir * mid, escDir, 0.0f, kFltMax, &t)) { maxD = mid; ++nref; }
        else minD = clampf(mid + t * escCos, mid, maxD);
    }
    *out_offset = offsetDir * maxD;
    *out_dir = escDir;
}

PHG_HD bool mutual_escape(const SurfaceHit& X, const TriVerts& tx, const SurfaceHit& X2, const TriVerts& t2,
                                     V3* org, V3* dir, float* maxT) {
    const V3 x2x = X2.P - X.P;
    V3 off, ud;
    iterative_offset(X, tx, x2x, &off, &ud);
    float init = length(X2.P) * kMaxErr;
    if (!__builtin_isfinite(init)) init = fmaxf(length(X2.P) * 1e-3f, 2e-4f);
    init = init + kFltMin;
    const V3 o = X.P + off;
    const float dist = length(o - X2.P);
    float maxD = init, minD = 0.0f;
    for (int it = 0; it < kMaxEscapeIters && dist > maxD; ++it) {
        float t;
        if (!reintersect(t2.a, t2.b, t2.c, o, ud, 0.0f, dist - maxD, &t)) break;
        maxD = fmaxf(dist - t, maxD);
        minD = maxD;
        maxD *= 2.0f;
    }
    int nref = 0;
    for (int it = 0; it < kMaxEscapeIters && dist > maxD && nref < 1; ++it) {
        const float mid = (minD + maxD) * 0.5f;
        if (!(minD < mid && mid < maxD)) break;
        float t;
        if (!reintersect(t2.a, t2.b, t2.c, o, ud, 0.0f, dist - mid, &t)) { maxD = mid; ++nref; }
        else minD = clampf(dist - t, mid, maxD);
    }
    const float len = fmaxf(dist - maxD, 0.0f);
    if (!(len > 0.0f && __builtin_isfinite(len))) return false;
    *org = o; *dir = ud; *maxT = len;
    return true;
}

// ---- emitters (ESPowerFavoring.cpp:60-97, GroupedDiffuseSurfaceEmitter.cpp:58-84,
//      DiffuseSurfaceEmitterBase.cpp:34-89, SurfaceEmitter.cpp:20-61, lta.h:17-25) ----------
PHG_HD float pdfA_to_pdfW(float pdfA, V3 posToTarget, V3 N) {
    const float d2 = dot(posToTarget, posToTarget);
    const float s = pdfA / dot(normalize(posToTarget), N) * d2;
    return __builtin_isfinite(s) ? fabsf(s) : 0.0f;
}
// TPwcDistribution1D::sampleDiscrete: lower_bound over cdf[0..n]
PHG_HD uint32_t cdf_sample(const float* cdf, uint32_t n, float u) {
    uint32_t lo = 0, len = n + 1;
    while (len > 0) {
        const uint32_t half = len >> 1;
        if (cdf[lo + half] < u) { lo = lo + half + 1; len = len - half - 1; }
        else len = half;
    }
    if (lo == 0) {
        for (uint32_t i = 0; i < n; ++i) if (cdf[i + 1] - cdf[i] > 0.0f) return i;
        return 0;
    }
    return lo - 1;
}
PHG_HD float mis_power(float p0, float p1) { return (p0 * p0) / (p0 * p0 + p1 * p1); }

// ---- film sample window (TSamplingFilm::updateSampleDimensions, TSamplingFilm.ipp:103-134) ---
// A film with a reconstruction filter of half size `half` (pixels) draws its samples over the
// image's pixel centres expanded by `half` (soft edge, softness 1), a batch's n strata per
// axis spread evenly over that window (ReceiverSamplingWork.cpp:76-101): stratum x with offset
// u lands at lo + ((x + u) / n) * extent, in FP64.  A unit-size (box) filter's window is the
// image itself and its samples sit at x + u in FP32.
constexpr double kFilterHalf = 2.0;   // Gaussian, Mitchell-Netravali and Blackman-Harris: 4 x 4 pixels
PHG_HD double window_raster(uint32_t x, float u, uint32_t n, double half) {
    const double lo = 0.5 - half;
    const double ext = ((double)n - 0.5 + half) - lo;
    return lo + (((double)x + (double)u) / (double)n) * ext;
}

// Raster position (FP64, as the film splats it) of the Halton source's sample of `pixel` in
// batch `sample`: Halton dimensions 0 and 1 of point sample N + pixel over the sample window
// (the image, or the soft-edged window of a filtered film: lo + v * extent).
PHG_HD void halton_raster(const JobDev& job, uint32_t pixel, uint32_t sample, double* x, double* y) {
    const uint64_t index = (uint64_t)sample * job.npix + pixel;
    const double half = job.filter == PHG_FILTER_BOX ? 0.5 : kFilterHalf;
    const float vx = halton_value(0u, index, job.sampler, job.halton_leap, job.seed);
    const float vy = halton_value(1u, index, job.sampler, job.halton_leap, job.seed);
    const double lo = 0.5 - half;
    *x = lo + (double)vx * (((double)job.width - 0.5 + half) - lo);
    *y = lo + (double)vy * (((double)job.height - 0.5 + half) - lo);
}

}  // namespace phg
