// Path-guiding model shared by the MLP kernels and the guided shade stage: feature
// encoding, decoding of the network output into a mixture of 8 von Mises-Fisher lobes,
// mixture pdf, lobe sampling, and the gradient of the radiance-weighted log-likelihood.
// (The reference renderer has no guiding; the specification is the oracle in
// oracle/guiding_oracle.py, which these functions follow operation for operation where
// it matters: encoding order, clamps and the kappa -> 0 limits.)
#pragma once

#include "guide.h"
#include "pt_device.h"

namespace phg {

constexpr int kGuideIn = 64;
constexpr int kGuideHid = 64;
constexpr int kGuideLayers = 5;
constexpr int kGuideLobes = 8;
constexpr int kGuideLobeVals = 5 * kGuideLobes;  // 40: per lobe mean direction (3), log kappa, weight logit
constexpr int kGuideSel = kGuideLobeVals;        // output 40: logit of the lobes' selection probability
constexpr int kGuideOut = kGuideLobeVals + 1;    // 41 raw outputs per query
constexpr int kGuideDec = kGuideLobeVals + 1;    // decoded: 8 x (mu, kappa, weight), then the selection probability
constexpr int kGuideOutPad = 48;                 // output rows padded to a multiple of 16
constexpr float kKappaLogMin = -8.0f, kKappaLogMax = 7.0f;
// learned selection probability of the lobes in the one-sample MIS: kSelMin + (kSelMax -
// kSelMin) sigmoid(output 40), so the BSDF always keeps a share of the samples
constexpr float kSelMin = 0.05f, kSelMax = 0.95f;

// flat FP32 parameter blob: W0 b0 W1 b1 W2 b2 W3 b3 W4 b4 with W stored [in][out]
constexpr int kGuideParamOffsets[11] = {0, 4096, 4160, 8256, 8320, 12416, 12480, 16576, 16640, 19264, 19305};
constexpr int kGuideParams = 19305;

// 9 feature scalars in [0, 1]: position in the scene box, normal, outgoing direction
PHG_HD void guide_features(const GuideScene& gs, V3 P, V3 N, V3 V, float* x) {
    x[0] = (P.x - gs.lo[0]) * gs.inv_ext[0];
    x[1] = (P.y - gs.lo[1]) * gs.inv_ext[1];
    x[2] = (P.z - gs.lo[2]) * gs.inv_ext[2];
    x[3] = 0.5f * (N.x + 1.0f); x[4] = 0.5f * (N.y + 1.0f); x[5] = 0.5f * (N.z + 1.0f);
    x[6] = 0.5f * (V.x + 1.0f); x[7] = 0.5f * (V.y + 1.0f); x[8] = 0.5f * (V.z + 1.0f);
}

// encoding value j (0..63) of the 9 features: raw, then sin/cos at frequencies pi, 2pi, 4pi.
// On the device the hardware sine/cosine (v_sin_f32 / v_cos_f32, argument in revolutions)
// serve: the arguments stay within 2 revolutions and every value is rounded to bf16 next.
PHG_HD float guide_encode(const float* x, int j) {
    if (j < 9) return x[j];
    if (j == 63) return 1.0f;
    const int r = j - 9, band = r / 18, q = r % 18, i = q >> 1;
#if defined(__HIP_DEVICE_COMPILE__)
    const float rev = (float)(1 << band) * 0.5f * x[i];
    return (q & 1) ? __builtin_amdgcn_cosf(rev) : __builtin_amdgcn_sinf(rev);
#else
    const float a = (float)(1 << band) * kPi * x[i];
    return (q & 1) ? cosf(a) : sinf(a);
#endif
}

// kappa / (2 pi (1 - exp(-2 kappa))), limit 1 / (4 pi) as kappa -> 0
PHG_HD float vmf_norm(float k) {
    if (k < 1e-6f) return 1.0f / (4.0f * kPi);
    return k / (2.0f * kPi * (-expm1f(-2.0f * k)));
}

// the lobes' selection probability from raw output 40
PHG_HD float guide_sel(float z) {
    return kSelMin + (kSelMax - kSelMin) / (1.0f + expf(-z));
}

// decode the 40 lobe outputs into 8 x (mu xyz, kappa, weight)
PHG_HD void guide_decode(const float* o, float* lobes) {
    float mx = -INFINITY;
    for (int j = 0; j < kGuideLobes; ++j) mx = fmaxf(mx, o[5 * j + 4]);
    float sum = 0.0f;
    for (int j = 0; j < kGuideLobes; ++j) sum += expf(o[5 * j + 4] - mx);
    for (int j = 0; j < kGuideLobes; ++j) {
        const float rx = o[5 * j], ry = o[5 * j + 1], rz = o[5 * j + 2];
        const float rl = sqrtf(rx * rx + ry * ry + rz * rz);
        float* L = lobes + 5 * j;
        if (rl > 0.0f) { L[0] = rx / rl; L[1] = ry / rl; L[2] = rz / rl; }
        else { L[0] = 0.0f; L[1] = 0.0f; L[2] = 1.0f; }
        L[3] = expf(clampf(o[5 * j + 3], kKappaLogMin, kKappaLogMax));
        L[4] = expf(o[5 * j + 4] - mx) / sum;
    }
}

// false if any decoded lobe value is not finite (a diverged network): the caller then
// samples the BSDF alone at this vertex, which keeps the estimator unbiased
PHG_HD bool guide_lobes_ok(const float* lobes) {
    float acc = 0.0f;
    for (int k = 0; k < kGuideDec; ++k) acc += lobes[k];
    return __builtin_isfinite(acc);
}

PHG_HD float guide_pdf(const float* lobes, V3 w) {
    float q = 0.0f;
    for (int j = 0; j < kGuideLobes; ++j) {
        const float* L = lobes + 5 * j;
        const float c = L[0] * w.x + L[1] * w.y + L[2] * w.z;
        q += L[4] * vmf_norm(L[3]) * expf(L[3] * (c - 1.0f));
    }
    return q;
}

// pick a lobe with u0, sample it with (u1, u2) (Wood 1994 / Jakob 2012 inversion)
PHG_HD V3 guide_sample(const float* lobes, float u0, float u1, float u2) {
    int j = 0;
    float acc = 0.0f;
    for (; j < kGuideLobes - 1; ++j) {
        acc += lobes[5 * j + 4];
        if (u0 < acc) break;
    }
    const float* L = lobes + 5 * j;
    const float k = L[3];
    float w;
    if (k < 1e-6f) w = 1.0f - 2.0f * u1;
    else w = 1.0f + logf(u1 + (1.0f - u1) * expf(-2.0f * k)) / k;
    w = clampf(w, -1.0f, 1.0f);
    const float s = sqrtf(fmaxf(0.0f, 1.0f - w * w));
    const float phi = kTwoPi * u2;
    const V3 mu = mk(L[0], L[1], L[2]);
    V3 bx, bz;
    basis_from_unit_y(mu, &bx, &bz);
    return normalize(bx * (s * cosf(phi)) + mu * w + bz * (s * sinf(phi)));
}

// Output gradients for one record (matches guiding_oracle.output_grad): direction w drawn
// with density samp_pdf, weight = integrand / that density.  q = the lobes' density at w,
// b = the BSDF density there, a = the selection probability of the lobes: the network's own
// (output 40) where samp_pdf > 0 (the record's selection probability was learned), else sel.
//   outputs 0..39: d(-weight * log(a q + (1 - a) b)), the radiance-weighted log-likelihood of
//     the one-sample MIS density (b = 0 or a = 0 gives d(-weight * log q));
//   output 40 (learned, a > 0, b > 0; else 0): -weight^2 samp_pdf (q - b) / (a q + (1 - a) b)^2
//     da/dz, the one-sample estimate of d/dz of the second moment of integrand / (a q + (1 - a)
//     b), which is the one-sample MIS estimator's variance plus a constant.
PHG_HD void guide_output_grad(const float* o, V3 w, float weight, float bsdf_pdf, float sel, float samp_pdf, float* g) {
    float lobes[kGuideLobeVals];
    guide_decode(o, lobes);
    float p[kGuideLobes], c[kGuideLobes];
    float q = 0.0f;
    for (int j = 0; j < kGuideLobes; ++j) {
        const float* L = lobes + 5 * j;
        c[j] = L[0] * w.x + L[1] * w.y + L[2] * w.z;
        p[j] = vmf_norm(L[3]) * expf(L[3] * (c[j] - 1.0f));
        q += L[4] * p[j];
    }
    // d log(a q + (1 - a) b) = (a q / (a q + (1 - a) b)) d log q
    const bool learned = samp_pdf > 0.0f && sel > 0.0f;
    const float a = learned ? guide_sel(o[kGuideSel]) : sel;
    const bool mix = a > 0.0f && bsdf_pdf > 0.0f;
    const float dens = mix ? a * q + (1.0f - a) * bsdf_pdf : q;
    const float share = mix ? a * q / dens : 1.0f;
    const float scale = q > 0.0f ? -weight * share / q : 0.0f;
    // selection logit: da/dz = (a - kSelMin) (kSelMax - a) / (kSelMax - kSelMin)
    g[kGuideSel] = (learned && mix && dens > 0.0f)
                       ? -weight * weight * samp_pdf * (q - bsdf_pdf) / (dens * dens) *
                             ((a - kSelMin) * (kSelMax - a) / (kSelMax - kSelMin))
                       : 0.0f;
    for (int j = 0; j < kGuideLobes; ++j) {
        const float* L = lobes + 5 * j;
        const float k = L[3];
        const float wp = L[4] * p[j];
        const float rx = o[5 * j], ry = o[5 * j + 1], rz = o[5 * j + 2];
        const float rl = sqrtf(rx * rx + ry * ry + rz * rz);
        const float inv = rl > 0.0f ? 1.0f / rl : 1.0f;
        const float f = scale * wp * k * inv;
        g[5 * j + 0] = f * (w.x - c[j] * L[0]);
        g[5 * j + 1] = f * (w.y - c[j] * L[1]);
        g[5 * j + 2] = f * (w.z - c[j] * L[2]);
        const float lk = o[5 * j + 3];
        float dl;  // 1/k - 2/expm1(2k) - 1
        if (k < 1e-2f) dl = -k / 3.0f + k * k * k / 45.0f;
        else dl = 1.0f / k - 2.0f / expm1f(2.0f * k) - 1.0f;
        g[5 * j + 3] = (lk > kKappaLogMin && lk < kKappaLogMax) ? scale * wp * (dl + c[j]) * k : 0.0f;
        g[5 * j + 4] = scale * (wp - L[4] * q);
    }
}

}  // namespace phg
