// Energy of a linear-sRGB colour as the reference estimates it (math::estimate_color_energy,
// Math/Color/color_spaces.ipp:743-764): Smits' up-sampling to a spectrum
// (color_spaces.ipp:341-476) times the D65 spectrum normalised to unit sum (emitted radiance),
// summed over the 100 samples of [350, 850] nm.  The estimate is linear in Smits' basis
// weights, so it is the colour's branch of Smits' algorithm over seven basis energies (white,
// cyan, magenta, yellow, red, green, blue), derived from the reference's tabulated spectra by
// tests/golden/make_color_energy.py.  Emitter-pick weights (DiffuseSurfaceEmitter.cpp:55-86)
// and area-light power (AAreaLight.cpp:55-58, Spectrum::putEnergy) use it.
#pragma once

namespace phg {

inline double color_energy(double r, double g, double b) {
    static const double E[7] = {0.9996779720137005, 0.5398497088854213, 0.7349389630228806, 0.7044563722431154, 0.4544199699880077, 0.2582578706735675, 0.3028703986116803};
    // the smallest channel weighs the white basis, the middle one's excess the complementary
    // (cyan / magenta / yellow) basis of the smallest, the largest one's excess its primary
    double lo, mid, hi;
    int second, third;
    if (r <= g && r <= b) { lo = r; mid = g < b ? g : b; hi = g < b ? b : g; second = 1; third = g <= b ? 6 : 5; }
    else if (g <= r && g <= b) { lo = g; mid = r < b ? r : b; hi = r < b ? b : r; second = 2; third = r <= b ? 6 : 4; }
    else { lo = b; mid = r < g ? r : g; hi = r < g ? g : r; second = 3; third = r <= g ? 5 : 4; }
    const double s = E[0] * lo + E[second] * (mid - lo) + E[third] * (hi - mid);
    return s > 0.0 ? s : 0.0;
}

}  // namespace phg
