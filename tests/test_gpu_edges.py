"""Edge cases of the GPU path: empty and one-element inputs, a scene without lights, images
whose sides are not multiples of the 8x8 pixel tiles, the smallest bounce cap, and the errors
the C ABI reports for calls it cannot serve (phg_status codes, phg_last_error text)."""
import numpy as np
import pytest

import phg
from oracle import oracle as orc
from helpers import FLT_MAX, pixel_z, random_soup, scene_from_tris

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = phg.Context(0)
    yield c
    c.close()


def test_empty_ray_batch(ctx):
    ctx.upload(scene_from_tris(random_soup(10, 1)))
    for any_hit in (False, True):
        hits = ctx.intersect(np.zeros((0, 8), np.float32), any_hit=any_hit)
        assert hits.shape == (0,)


def test_one_triangle_scene(ctx):
    """A BVH of a single leaf: rays through a grid of points around the triangle, parallel to
    its normal; hits and t equal the oracle's brute force bit for bit away from the edges."""
    v = np.array([[0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0]], np.float32)
    ctx.upload(scene_from_tris(v))
    g = np.linspace(-0.45, 1.45, 39, dtype=np.float32)
    x, y = np.meshgrid(g, g)
    rays = np.zeros((x.size, 8), np.float32)
    rays[:, 0], rays[:, 1], rays[:, 2] = x.ravel(), y.ravel(), -2.0
    rays[:, 6] = 1.0
    rays[:, 7] = FLT_MAX
    tri, t, _, flags = orc.intersect_brute(v, rays)
    ok = (flags & 1) == 0
    hits = ctx.intersect(rays)
    assert np.array_equal(hits["prim"][ok], tri[ok])
    h = ok & (tri >= 0)
    assert np.array_equal(hits["t"][h].view(np.uint32), t[h].view(np.uint32))
    occl = ctx.intersect(rays, any_hit=True)   # prim 1 if occluded, -1 otherwise
    assert np.array_equal(occl["prim"][ok] >= 0, tri[ok] >= 0)
    inside = (x.ravel() > 0.01) & (y.ravel() > 0.01) & (x.ravel() + y.ravel() < 0.99)
    assert np.all(tri[inside] == 0) and np.count_nonzero(tri < 0) > 0
    assert np.allclose(t[inside], 2.0)


@pytest.mark.parametrize("estimator", [phg.PHG_ESTIMATOR_BNEEPT, phg.PHG_ESTIMATOR_BVPT, phg.PHG_ESTIMATOR_BVPTDL])
def test_scene_without_lights_renders_black(ctx, estimator):
    sc = scene_from_tris(random_soup(200, 3, spread=2.0))   # no emitters at all
    W, H, spp = 24, 16, 4
    ctx.upload(sc)
    ctx.film_clear(W, H)
    ctx.render(spp, max_bounces=8, seed=1, estimator=estimator)
    f = ctx.film()
    assert np.all(f[..., [0, 1, 2, 4, 5, 6]] == 0.0)
    assert np.all(f[..., 3] == spp)


@pytest.mark.parametrize("W,H", [(1, 1), (7, 5), (9, 17), (130, 3)])
def test_ragged_image_sizes_match_oracle(ctx, W, H):
    """Image sides that are not multiples of the 8x8 slot tiles: every pixel receives exactly
    its own samples (weights equal the oracle's), and values agree per pixel."""
    sc = phg.Scene.generate(phg.GEN_CORNELL)
    spp = 16
    ctx.upload(sc)
    ctx.film_clear(W, H)
    ctx.render(spp, max_bounces=5, seed=31)
    gpu = ctx.film()
    ref, _ = orc.render(sc, W, H, spp, max_bounces=5, seed=31)
    assert np.array_equal(gpu[..., 3], ref[..., 3])
    ident = np.all(np.isclose(gpu[..., :3], ref[..., :3], rtol=1e-6, atol=1e-9), axis=-1).mean()
    z = pixel_z(gpu, ref)
    print(f"[edges] case=ragged {W}x{H} identical_pixels={ident:.3f} max|z|={np.max(np.abs(z)):.2f}", flush=True)
    assert np.all(np.abs(z) < 4.5)
    if W * H >= 30:
        assert ident >= 0.75


def test_zero_bounce_cap_is_emission_seen_by_the_camera(ctx):
    """max_bounces 0: a camera path adds only the emission of its primary hit (no NEE, no
    extension), so the image is the light seen directly, equal to the oracle's."""
    sc = phg.Scene.generate(phg.GEN_CORNELL)
    W = H = 32
    ctx.upload(sc)
    ctx.film_clear(W, H)
    ctx.render(4, max_bounces=0, seed=2)
    gpu = ctx.film()
    st = ctx.stats()
    ref, _ = orc.render(sc, W, H, 4, max_bounces=0, seed=2)
    assert st["rays_shadow"] == 0
    assert np.count_nonzero(gpu[..., 0]) > 0           # the light is in view
    assert np.array_equal(gpu[..., 3], ref[..., 3])
    assert np.allclose(gpu, ref, rtol=1e-6, atol=1e-9)


def test_abi_reports_errors():
    cornell = phg.Scene.generate(phg.GEN_CORNELL)
    with phg.Context(0) as c:
        with pytest.raises(phg.PhgError, match="NO_SCENE"):
            c.intersect(np.zeros((4, 8), np.float32))
        c.film_clear(8, 8)
        with pytest.raises(phg.PhgError, match="NO_SCENE"):
            c.render(1)
        c.upload(cornell)
        with pytest.raises(phg.PhgError, match="INVALID_ARG"):
            c.render(1, seed=1 << 24)                    # sample-stream seeds are below 2^24
        with pytest.raises(phg.PhgError, match="UNSUPPORTED"):
            c.render(1, guided=True, estimator=phg.PHG_ESTIMATOR_BVPT)
        with pytest.raises(phg.PhgError, match="INVALID_ARG"):
            c.set_tuning(trace_block=96)
        c.width = 16                                     # a render size other than the film's
        with pytest.raises(phg.PhgError, match="film size differs"):
            c.render(1)
        c.width = 8
        c.render(2, max_bounces=3, seed=4)               # the context still works
        assert np.all(c.film()[..., 3] == 2)
