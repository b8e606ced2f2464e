"""Gray furnace with glass objects on the CPU oracle: the reference's energy check of the
GGX dielectric (Main/RenderTest/tests/test_gray_furnace_box.py:31-37,48-49,79-80 — BNEEPT
glass sphere and glass cube inside walls of radiance 0.5, image mean 0.5 within 1e-3).

The dielectric at the roughness floor (alpha 0.001) neither absorbs nor emits, so any error
in the reflect / transmit split (F vs 1 - F, the pick probability it is divided by), in the
refracted direction or in the sample weight's G D |H.V| / (|N.V| pdf) term moves the mean
away from 0.5.  An outside camera cannot see the (eta_t / eta_i)^2 radiance factor of
TranslucentMicrofacet.cpp:236-241 (it enters and leaves), so a second closed form puts the
camera inside the glass: there every pixel is 0.5 * 1.5^2 = 1.125, which fails if the factor
is dropped, inverted or applied to the reflected lobe.

The bar is the image mean, as in the reference: per-pixel sample distributions here are a
spike at the expected value plus rare Russian-roulette outcomes, so per-pixel normal
intervals do not apply."""
import numpy as np
import pytest

from oracle import oracle as orc
from helpers import film_mean_se, glass_furnace_scene


@pytest.mark.parametrize("shape", ["sphere", "cube"])
def test_oracle_glass_furnace_mean_is_half(shape):
    sc = glass_furnace_scene(shape)
    film, _ = orc.render(sc, 48, 48, 256, max_bounces=10000, seed=41)
    mean, se = film_mean_se(film)
    print(f"[furnace] oracle glass {shape}: mean={mean:.6f} se={se:.2e}")
    assert abs(mean - 0.5) < 1e-3                       # the reference's bar


def test_oracle_camera_inside_glass_sees_eta_squared():
    sc = glass_furnace_scene("sphere", camera="inside")
    film, _ = orc.render(sc, 32, 32, 64, max_bounces=10000, seed=43)
    mean, se = film_mean_se(film)
    print(f"[furnace] oracle inside glass: mean={mean:.6f} se={se:.2e} expect=1.125")
    assert abs(mean - 1.125) < 5 * se + 1e-5
    assert abs(mean - 1.125) < 2.5e-3


@pytest.mark.parametrize("fault,camera", [(1, "inside"), (2, "outside"), (3, "outside")])
def test_glass_furnace_detects_planted_mistakes(fault, camera):
    """Each mistake planted in the oracle's transmission lobe (1: no (eta_t / eta_i)^2 factor,
    2: F in place of 1 - F, 3: no division by the transmission pick probability) moves the
    furnace far outside its bar; without it the same render passes."""
    L = orc.lib()
    sc = glass_furnace_scene("sphere", camera=camera)
    expect = 1.125 if camera == "inside" else 0.5
    try:
        L.orc_set_fault(fault)
        film, _ = orc.render(sc, 16, 16, 16, max_bounces=10000, seed=45)
    finally:
        L.orc_set_fault(0)
    mean, se = film_mean_se(film)
    print(f"[furnace] planted mistake {fault} ({camera}): mean={mean:.5f} expect={expect}")
    assert abs(mean - expect) > max(0.02, 8 * se)
