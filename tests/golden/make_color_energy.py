"""Writes tests/golden/color_energy.json: the reference's colour-energy estimate of linear-sRGB
emitted radiance, which sets the emitter-pick weights (DiffuseSurfaceEmitter.cpp:55-86,
math::estimate_color_energy, Math/Color/color_spaces.ipp:743-764).

The estimate up-samples the colour to a spectrum with Smits' method over the reference's
tabulated illuminant-E bases (Math/Color/color_spaces.ipp:341-476, data in
Math/Color/spectral_data.cpp), multiplies it by the D65 spectrum normalised to unit sum (EMR
usage), and sums the 100 samples of [350, 850] nm (Common/config.h:149-151,
Math/Color/spectral_samples.ipp:38-43).  Every tabulated curve is resampled by averaging its
piecewise-linear interpolation (end values held beyond the data) over each 5 nm interval
(spectral_samples.ipp:67-109, TPiecewiseLinear1D.h, TAnalyticalIntegrator1D.ipp).

The estimate is linear in Smits' basis weights, so it reduces to seven constants -- the energy
of each basis spectrum -- and the colour's branch of Smits' algorithm.  This script reads the
data tables as text from the reference checkout (present only where the fixture is made),
restates the resampling, and writes the seven constants and energies of test colours.
Run from the repository root: python tests/golden/make_color_energy.py
"""
import json
import os
import re

import numpy as np

REF = "/root/reference/Engine/Engine/Source/Math/Color/spectral_data.cpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "color_energy.json")
NUM, LO, HI = 100, 350.0, 850.0
BASES = ["white", "cyan", "magenta", "yellow", "red", "green", "blue"]


def table(text, name):
    m = re.search(r"constexpr \w+ " + re.escape(name) + r"_data\s*=\s*\{\{(.*?)\}\};", text, re.S)
    if not m:
        raise SystemExit("table not found: " + name)
    return np.array([float(v) for v in m.group(1).replace("\n", " ").split(",") if v.strip()], np.float64)


def interval_average(xs, ys, a, b):
    """Mean over [a, b] of the piecewise-linear curve through (xs, ys), end values held."""
    if b < xs[0]:
        return ys[0]
    if a > xs[-1]:
        return ys[-1]
    area = 0.0
    if a <= xs[0]:
        area += (xs[0] - a) * ys[0]
    if b >= xs[-1]:
        area += (b - xs[-1]) * ys[-1]
    for i in range(len(xs) - 1):
        x0, x1 = max(a, xs[i]), min(b, xs[i + 1])
        if x0 < x1:
            f = lambda x: (x - xs[i]) / (xs[i + 1] - xs[i]) * (ys[i + 1] - ys[i]) + ys[i]
            area += (f(x0) + f(x1)) * (x1 - x0) / 2.0
    return area / (b - a)


def resample(xs, ys):
    step = (HI - LO) / NUM
    return np.array([interval_average(xs, ys, LO + i * step, LO + (i + 1) * step) for i in range(NUM)])


def energy(c, e):
    r, g, b = c
    w = dict.fromkeys(BASES, 0.0)
    if r <= g and r <= b:
        w["white"] = r
        if g <= b: w["cyan"], w["blue"] = g - r, b - g
        else: w["cyan"], w["green"] = b - r, g - b
    elif g <= r and g <= b:
        w["white"] = g
        if r <= b: w["magenta"], w["blue"] = r - g, b - r
        else: w["magenta"], w["red"] = b - g, r - b
    else:
        w["white"] = b
        if r <= g: w["yellow"], w["green"] = r - b, g - r
        else: w["yellow"], w["red"] = g - b, r - g
    s = sum(w[k] * e[k] for k in BASES)
    return s if s > 0 else 0.0


def main():
    text = open(REF).read()
    d65 = resample(table(text, "CIE_D65_wavelengths_nm"), table(text, "CIE_D65_values"))
    d65 = d65 / d65.sum()
    wl = table(text, "smits_linear_sRGB_to_spectrum_E_wavelengths_nm")
    e = {k: float((resample(wl, table(text, "smits_linear_sRGB_to_spectrum_E_" + k)) * d65).sum()) for k in BASES}
    rng = np.random.default_rng(2024)
    colours = [[1, 1, 1], [0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [0, 1, 1], [1, 0, 1],
               [0.2, 0.5, 0.9], [0.9, 0.5, 0.2], [5.0, 2.5, 1.25], [0.3, 0.3, 0.7], [-1.0, 0.0, 0.0]]
    colours += rng.uniform(0.0, 10.0, (24, 3)).round(4).tolist()
    vec = [[*c, energy(c, e)] for c in colours]
    json.dump({"constants": e, "vectors": vec}, open(OUT, "w"), indent=1)
    print(json.dumps(e, indent=1))


if __name__ == "__main__":
    main()
