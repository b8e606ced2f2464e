"""Ray-batch parity of the HIP traversal + watertight test against the CPU oracle.

Bar: hit/miss and primitive id identical, hit distance and barycentrics bit-identical,
for every ray except exact-distance ties (counted; the reference resolves them by test
order, TLinearDepthFirstBinaryBvh.ipp:79-92).  Rays whose closest hit lies within 1e-6 of a
triangle edge are counted and reported; they are compared like all others."""
import numpy as np
import pytest

import phg
from oracle import oracle as orc
from helpers import FLT_MAX, nested_soup, random_rays, random_soup, rays_towards, scene_from_tris

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = phg.Context(0)
    yield c
    c.close()


def compare_closest(ctx, verts, rays, name):
    print(f"[intersect] case={name} tris={len(verts)} rays={len(rays)}", flush=True)
    ctx.upload(scene_from_tris(verts))
    hits = ctx.intersect(rays)
    tri, t, bary, flags = orc.intersect_brute(verts, rays)
    tie = (flags & 1) != 0
    edge = (flags & 2) != 0
    ok = ~tie
    assert np.array_equal(hits["prim"][ok], tri[ok]), \
        f"{name}: {np.count_nonzero(hits['prim'][ok] != tri[ok])} primitive-id mismatches"
    h = ok & (tri >= 0)
    assert np.array_equal(hits["t"][h].view(np.uint32), t[h].view(np.uint32)), f"{name}: hit distance differs"
    assert np.array_equal(hits["b"][h].view(np.uint32), bary[h, 1].view(np.uint32)), f"{name}: barycentric b differs"
    assert np.array_equal(hits["c"][h].view(np.uint32), bary[h, 2].view(np.uint32)), f"{name}: barycentric c differs"
    # ties: both sides found a hit at the same distance (within 4 ulp)
    th = tie & (tri >= 0)
    if th.any():
        assert np.all(hits["prim"][th] >= 0)
        ulp = np.abs(hits["t"][th].view(np.int32).astype(np.int64) - t[th].view(np.int32).astype(np.int64))
        assert ulp.max() <= 4
    print(f"[intersect] case={name} hits={np.count_nonzero(tri >= 0)} ties={np.count_nonzero(tie)} "
          f"edge_near={np.count_nonzero(edge)}", flush=True)
    return hits


def test_cornell_camera_and_random_rays(ctx):
    sc = phg.Scene.generate(phg.GEN_CORNELL)
    rays = np.concatenate([rays_towards(sc.verts, 100_000, 1), random_rays(sc.verts, 100_000, 2)])
    compare_closest(ctx, sc.verts, rays, "cornell")


@pytest.mark.parametrize("n,seed", [(1, 3), (7, 4), (2000, 5)])
def test_random_soups(ctx, n, seed):
    v = random_soup(n, seed)
    rays = np.concatenate([rays_towards(v, 100_000, seed + 10, jitter=0.01), random_rays(v, 50_000, seed + 20)])
    compare_closest(ctx, v, rays, f"soup{n}")


def test_segments_and_misses(ctx):
    v = random_soup(500, 9)
    rays = rays_towards(v, 50_000, 11)
    rng = np.random.default_rng(12)
    # finite segments: some end before the surface, some start after it
    rays[:, 3] = rng.uniform(0, 5, len(rays)).astype(np.float32)
    rays[:, 7] = (rays[:, 3] + rng.uniform(0, 30, len(rays))).astype(np.float32)
    away = rays.copy()
    away[:, 4:7] *= -1
    compare_closest(ctx, v, np.concatenate([rays, away]), "segments")


def test_degenerate_and_nonfinite_triangles(ctx):
    # TWatertightTriangle_test.cpp:62-80: a vertex at +inf must not produce a hit
    inf, nan = np.inf, np.nan
    v = np.array([[0, 0, 0, 1, 1, 0, 1, -1, inf],
                  [0, 0, 2, 1, 0, 2, 2, 0, 2],            # zero area (collinear)
                  [0, 0, 3, 1, nan, 3, 0, 1, 3],
                  [-5, -5, 4, 5, -5, 4, 0, 5, 4]], np.float32)
    rays = np.array([[0.1, 0, -1, 0, 0.1, 0, 1, FLT_MAX],   # reference segment of the test
                     [0.5, 0, -1, 0, 0, 0, 1, FLT_MAX],
                     [0.2, 0.2, -1, 0, 0, 0, 1, FLT_MAX]], np.float32)
    rays[:, 4:7] /= np.linalg.norm(rays[:, 4:7], axis=1, keepdims=True)
    ctx.upload(scene_from_tris(v))
    hits = ctx.intersect(rays)
    tri, t, bary, flags = orc.intersect_brute(v, rays)
    assert np.array_equal(hits["prim"], tri)
    assert np.all(tri == 3), "only the finite, non-degenerate triangle may be hit"


def test_shared_edge_is_watertight(ctx):
    # two triangles sharing the edge x=0; rays exactly through the edge must not leak
    v = np.array([[0, -1, 0, 0, 1, 0, -1, 0, 0],
                  [0, -1, 0, 1, 0, 0, 0, 1, 0]], np.float32)
    ys = np.linspace(-0.9, 0.9, 1001, dtype=np.float32)
    rays = np.zeros((len(ys), 8), np.float32)
    rays[:, 1] = ys
    rays[:, 2] = -1
    rays[:, 6] = 1
    rays[:, 7] = FLT_MAX
    ctx.upload(scene_from_tris(v))
    hits = ctx.intersect(rays)
    assert np.all(hits["prim"] >= 0), "a ray through the shared edge missed both triangles"
    tri, t, bary, flags = orc.intersect_brute(v, rays)
    assert np.array_equal(hits["prim"], tri)


def test_deep_bvh_spills_the_stack_ring(ctx):
    # traversal stack deeper than the on-chip ring: entries spill to the overflow area
    v = nested_soup(1000, 21, ratio=0.98)
    rays = np.concatenate([rays_towards(v, 20_000, 22), random_rays(v, 10_000, 23)])
    rays[:, 0:3] *= 0.05
    compare_closest(ctx, v, rays, "deep_bvh")
    assert ctx.stats()["bvh_stack_need"] > 32
    tri = orc.intersect_brute(v, rays)[0]
    assert np.array_equal(ctx.intersect(rays, any_hit=True)["prim"] >= 0, tri >= 0)


def test_procedural_against_reference_bvh(ctx):
    sc = phg.Scene.generate(phg.GEN_PROCEDURAL, seed=3, target_tris=200_000, num_instances=16, num_lights=4)
    rays = np.concatenate([rays_towards(sc.verts, 60_000, 21), random_rays(sc.verts, 60_000, 22)])
    print(f"[intersect] case=procedural tris={sc.num_tris} rays={len(rays)}", flush=True)
    ctx.upload(sc)
    hits = ctx.intersect(rays)
    tri, t = orc.intersect_refbvh(sc.verts, rays)
    same = hits["prim"] == tri
    both = same & (tri >= 0)
    assert np.array_equal(hits["t"][both].view(np.uint32), t[both].view(np.uint32))
    diff = ~same
    if diff.any():
        # only exact-distance ties may differ; confirm with brute force
        _, tb, _, fl = orc.intersect_brute(sc.verts, rays[diff])
        assert np.all((fl & 1) != 0), f"{np.count_nonzero((fl & 1) == 0)} non-tie mismatches"
    print(f"[intersect] case=procedural mismatching_ties={np.count_nonzero(diff)}", flush=True)


def test_any_hit_matches_closest(ctx):
    sc = phg.Scene.generate(phg.GEN_PROCEDURAL, seed=5, target_tris=100_000, num_instances=8, num_lights=2)
    rays = rays_towards(sc.verts, 100_000, 31)
    rng = np.random.default_rng(32)
    rays[:, 7] = rng.uniform(0, 40, len(rays)).astype(np.float32)
    ctx.upload(sc)
    closest = ctx.intersect(rays)
    occl = ctx.intersect(rays, any_hit=True)
    assert np.array_equal(occl["prim"] >= 0, closest["prim"] >= 0)


def test_counted_variant(ctx):
    sc = phg.Scene.generate(phg.GEN_CORNELL)
    rays = rays_towards(sc.verts, 10_000, 41)
    ctx.upload(sc)
    plain = ctx.intersect(rays)
    counted, nv, nt = ctx.intersect_counted(rays)
    assert np.array_equal(plain, counted)
    assert nv.min() >= 1 and nt.sum() > 0
