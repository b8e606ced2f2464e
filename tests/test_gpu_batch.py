"""Batched rendering (phg_render_batch): several scenes and many jobs share wavefront passes.

Bar: every job's FP64 film is bit-identical to rendering that job alone with phg_render —
the sample streams depend only on (seed, pixel, sample) and each pixel accumulates its
samples in order, so how jobs are packed into passes must not change a single bit."""
import numpy as np
import pytest

import phg

pytestmark = pytest.mark.gpu


def scenes():
    return [phg.Scene.generate(phg.GEN_CORNELL),
            phg.Scene.generate(phg.GEN_PROCEDURAL, seed=5, target_tris=30_000, num_instances=4, num_lights=3),
            phg.Scene.generate(phg.GEN_OCCLUDED_ROOM, seed=3, target_tris=20_000)]


JOBS = [dict(scene=0, width=64, height=48, spp=9, max_bounces=5, seed=11),
        dict(scene=1, width=96, height=64, spp=5, max_bounces=8, seed=12, sampler=phg.SAMPLER_STRATIFIED),
        dict(scene=2, width=40, height=40, spp=17, max_bounces=10000, seed=13),
        dict(scene=0, width=33, height=17, spp=3, max_bounces=2, seed=14,
             sampler=phg.SAMPLER_HALTON | phg.HALTON_PERM_OWEN),
        dict(scene=1, width=200, height=150, spp=2, max_bounces=4, seed=15)]


def same_film(a, b, job):
    """Equal films: bit for bit, or for a Halton job (splatted with FP64 atomics, so the order
    of the additions varies) equal counts and sums equal to rounding."""
    if job.get("sampler", 0) & 0xFF == phg.SAMPLER_HALTON:
        return np.array_equal(a[..., 3], b[..., 3]) and np.allclose(a, b, rtol=1e-12, atol=1e-12)
    return np.array_equal(a, b)


def single(scene, job):
    with phg.Context(0) as c:
        c.upload(scene)
        c.film_clear(job["width"], job["height"])
        c.render(job["spp"], max_bounces=job["max_bounces"], seed=job["seed"], sampler=job.get("sampler", 0))
        return c.film()


@pytest.mark.parametrize("paths_per_pass,order", [(0, phg.BATCH_ORDER_GIVEN), (20_000, phg.BATCH_ORDER_GIVEN),
                                                  (20_000, phg.BATCH_ORDER_SHORTEST_FIRST)])
def test_batch_films_equal_single_renders(paths_per_pass, order):
    sc = scenes()
    print(f"[batch] case paths_per_pass={paths_per_pass} order={order} jobs={len(JOBS)}", flush=True)
    with phg.Context(0) as c:
        c.upload_many(sc)
        res = c.render_batch(JOBS, paths_per_pass=paths_per_pass, order=order)
        films = [c.batch_film(i) for i in range(len(JOBS))]
        st = c.stats()
    print(f"[batch] passes={st['batch_passes']} done_ms={[round(r[0], 2) for r in res]}", flush=True)
    for i, job in enumerate(JOBS):
        assert res[i][1] == job["width"] * job["height"] * job["spp"]
        ref = single(sc[job["scene"]], job)
        assert same_film(films[i], ref, job), f"job {i} differs from its single render"
    if paths_per_pass:
        # the 200x150 job exceeds a pass: one sample per pass
        assert st["batch_passes"] >= 2


@pytest.mark.parametrize("njobs", [8, 9])
def test_segment_lookup_either_side_of_the_search_limit(njobs):
    """A pass of at most 8 job segments finds a path's segment by searching the segment table;
    a pass of more stores a per-path index.  One pass of 8 and one of 9 small jobs (stratified
    and Halton ones among them, whose first-vertex dimensions travel through the path state):
    every job's film equals its single render."""
    sc = scenes()
    samplers = [0, phg.SAMPLER_STRATIFIED, phg.SAMPLER_HALTON | phg.HALTON_PERM_FIXED]
    jobs = [dict(scene=k % 3, width=24 + 4 * k, height=16 + 2 * k, spp=2 + k % 4, max_bounces=3 + k % 5,
                 seed=40 + k, sampler=samplers[k % 3]) for k in range(njobs)]
    print(f"[batch] case segments={njobs}", flush=True)
    with phg.Context(0) as c:
        c.upload_many(sc)
        res = c.render_batch(jobs)
        films = [c.batch_film(i) for i in range(njobs)]
        st = c.stats()
    assert st["batch_passes"] == 1, "all jobs must share one pass"
    for i, job in enumerate(jobs):
        assert res[i][1] == job["width"] * job["height"] * job["spp"]
        assert same_film(films[i], single(sc[job["scene"]], job), job), f"job {i} differs from its single render"


def test_batch_reports_completion_order():
    sc = scenes()[:2]
    jobs = [dict(scene=1, width=256, height=256, spp=8, max_bounces=8, seed=1),
            dict(scene=0, width=32, height=32, spp=4, max_bounces=5, seed=2)]
    with phg.Context(0) as c:
        c.upload_many(sc)
        given = c.render_batch(jobs, paths_per_pass=1 << 17)
        sjf = c.render_batch(jobs, paths_per_pass=1 << 17, order=phg.BATCH_ORDER_SHORTEST_FIRST)
    print(f"[batch] given={given} shortest_first={sjf}", flush=True)
    # shortest first finishes the small job in the first pass
    assert sjf[1][0] <= sjf[0][0]
    assert all(r[0] > 0 for r in given + sjf)


def test_captured_bounce_loop_equals_eager():
    """PHG_RENDER_GRAPH replays one captured pass (generate, every bounce, film) per pass;
    its films must equal the eager loop's bit for bit, replay after replay, with eager
    renders and synchronisations in between."""
    sc = phg.Scene.generate(phg.GEN_PROCEDURAL, seed=9, target_tris=50_000, num_instances=6, num_lights=4)
    W, H = 80, 60
    with phg.Context(0) as c:
        c.upload(sc)
        films = []
        for graph in (False, True, False, True):
            c.film_clear(W, H)
            c.render(24, max_bounces=8, seed=21, spp_per_pass=3, graph=graph)   # 8 passes = 8 replays
            films.append(c.film())
            st = c.stats()
            print(f"[batch] graph={graph} trace_ms={st['trace_ms']:.3f} graph_ms={st['graph_ms']:.3f} "
                  f"launches={st['trace_launches']}", flush=True)
            assert (st["graph_ms"] if graph else st["trace_ms"]) > 0 and st["rays_extension"] > 0
        for f in films[1:]:
            assert np.array_equal(f, films[0])
        # counters agree too
        c.film_clear(W, H)
        c.render(4, max_bounces=8, seed=21, count_traversal=True)
        eager = c.stats()
        c.film_clear(W, H)
        c.render(4, max_bounces=8, seed=21, count_traversal=True, graph=True)
        graph = c.stats()
        for k in ("rays_extension", "rays_shadow", "trace_node_visits", "trace_tri_tests", "shadow_node_visits"):
            assert eager[k] == graph[k], k


def test_batch_graph_equals_eager():
    sc = scenes()
    jobs = [dict(j, max_bounces=min(j["max_bounces"], 12)) for j in JOBS]
    with phg.Context(0) as c:
        c.upload_many(sc)
        c.render_batch(jobs, paths_per_pass=20_000)
        eager = [c.batch_film(i) for i in range(len(jobs))]
        c.render_batch(jobs, paths_per_pass=20_000, graph=True)
        graph = [c.batch_film(i) for i in range(len(jobs))]
        print(f"[batch] graph passes={c.stats()['batch_passes']}", flush=True)
    for i in range(len(jobs)):
        assert same_film(eager[i], graph[i], jobs[i]), i


def test_launch_tuning_does_not_change_films():
    sc = phg.Scene.generate(phg.GEN_PROCEDURAL, seed=4, target_tris=40_000, num_instances=5, num_lights=3)
    W, H = 72, 40
    films = []
    with phg.Context(0) as c:
        c.upload(sc)
        for block, ppp in ((128, 0), (64, 1 << 16), (256, 3 << 16)):
            c.set_tuning(trace_block=block, paths_per_pass=ppp)
            t = c.tuning()
            print(f"[batch] tuning {t}", flush=True)
            assert t["trace_block"] == block and t["trace_blocks_per_cu"] >= 1
            c.film_clear(W, H)
            c.render(12, max_bounces=8, seed=5, graph=(block == 256))
            films.append(c.film())
    for f in films[1:]:
        assert np.array_equal(f, films[0])


@pytest.mark.parametrize("leaf,width", [(1, 2), (8, 3), (2, 4)])
def test_bvh_shapes_keep_intersection_parity(leaf, width):
    from oracle import oracle as orc
    from helpers import random_rays, random_soup, rays_towards, scene_from_tris
    v = random_soup(3000, 31)
    rays = np.concatenate([rays_towards(v, 30_000, 32), random_rays(v, 20_000, 33)])
    with phg.Context(0) as c:
        c.upload(scene_from_tris(v), max_leaf_tris=leaf, width=width)
        hits = c.intersect(rays)
    tri, t, bary, flags = orc.intersect_brute(v, rays)
    ok = (flags & 1) == 0
    assert np.array_equal(hits["prim"][ok], tri[ok])
    h = ok & (tri >= 0)
    assert np.array_equal(hits["t"][h].view(np.uint32), t[h].view(np.uint32))


def test_autotuner_picks_a_measured_configuration():
    sc = scenes()
    with phg.Context(0) as c:
        best, trials = c.autotune_batch(sc, JOBS, sample_jobs=4, sample_spp=2, repeats=1)
        print(f"[batch] autotune best={best} trials={len(trials)}", flush=True)
        assert len(trials) == 3 * 6 + 5   # blocks x pass sizes, then BVH leaf x width
        assert best["msamples_per_s"] == max(t["msamples_per_s"] for t in trials)
        assert c.tuning()["trace_block"] == best["trace_block"]
        # the tuned context renders the real jobs, and the films are still the single-job films
        c.upload_many(sc, max_leaf_tris=best["max_leaf_tris"], width=best["width"])
        c.render_batch(JOBS[:2], graph=True)
        film0 = c.batch_film(0)
    with phg.Context(0) as c:
        c.upload(sc[JOBS[0]["scene"]], max_leaf_tris=best["max_leaf_tris"], width=best["width"])
        c.film_clear(JOBS[0]["width"], JOBS[0]["height"])
        c.render(JOBS[0]["spp"], max_bounces=JOBS[0]["max_bounces"], seed=JOBS[0]["seed"])
        assert np.array_equal(c.film(), film0)


def test_job_cameras_render_their_own_views():
    sc = phg.Scene.generate(phg.GEN_CORNELL)
    cam = phg.camera_look_at([0.2, 0.7, -0.8], [0.55, 0.35, 0.6], fov_deg=55)
    jobs = [dict(scene=0, width=48, height=40, spp=6, max_bounces=5, seed=3),
            dict(scene=0, width=48, height=40, spp=6, max_bounces=5, seed=3, camera=sc.camera),
            dict(scene=0, width=40, height=30, spp=6, max_bounces=5, seed=4, camera=cam)]
    with phg.Context(0) as c:
        c.upload_many([sc])
        c.render_batch(jobs, paths_per_pass=1 << 16)
        films = [c.batch_film(i) for i in range(3)]
    assert np.array_equal(films[0], films[1])       # the scene's own camera, given explicitly
    moved = phg.Scene(sc.verts, sc.tri_material, sc.tri_emitter, sc.materials, sc.emitters, sc.emitter_tris, cam)
    ref = single(moved, jobs[2])
    assert np.array_equal(films[2], ref)


def test_guided_batch_of_views_matches_plain():
    """Guided rendering of several views sharing one network, trained online on all of
    them, agrees with plain rendering per pixel (99% CIs)."""
    from helpers import pixel_z
    sc = phg.Scene.generate(phg.GEN_CORNELL)
    cams = [sc.camera, phg.camera_look_at([0.15, 0.8, -0.9], [0.5, 0.3, 0.7], fov_deg=50)]
    jobs = [dict(scene=0, width=32, height=32, spp=256, max_bounces=10000, seed=40 + k, camera=cams[k])
            for k in range(2)]
    with phg.Context(0) as c:
        c.upload_many([sc])
        train = [dict(j, spp=16, seed=90 + i) for i, j in enumerate(jobs)]
        c.render_batch(train, paths_per_pass=1 << 14, train=True, guided=True, guide_batch=8192)
        st = c.stats()
        assert st["guide_records"] > 0
        c.render_batch(jobs, paths_per_pass=1 << 16, guided=True)
        g = [c.batch_film(i) for i in range(2)]
        st = c.stats()
        assert st["guide_queries"] > 0
        c.render_batch([dict(j, seed=j["seed"] + 100) for j in jobs], paths_per_pass=1 << 16)
        p = [c.batch_film(i) for i in range(2)]
    for k in range(2):
        z = pixel_z(g[k], p[k])
        frac = np.mean(np.abs(z) > 2.5758)
        print(f"[batch] guided view {k}: outside_99CI={frac:.4f}", flush=True)
        assert frac < 0.02


def test_guided_batch_needs_one_scene():
    sc = scenes()[:2]
    with phg.Context(0) as c:
        c.upload_many(sc)
        with pytest.raises(phg.PhgError):
            c.render_batch(JOBS[:2], guided=True)


def test_captured_chunks_equal_eager_unbounded():
    """Unbounded renders (the reference's 10000-bounce cap) replay captured chunks of 8
    bounces until the extension queue drains: films and ray counts equal the eager loop's,
    replay after replay, and a bounce cap that falls inside a chunk stops the paths there."""
    sc = phg.Scene.generate(phg.GEN_OCCLUDED_ROOM, seed=3, target_tris=20_000)
    W, H = 64, 48
    with phg.Context(0) as c:
        c.upload(sc)
        for cap in (10000, 13):
            films, stats = [], []
            for graph in (False, True, False, True):
                c.film_clear(W, H)
                c.render(12, max_bounces=cap, seed=8, spp_per_pass=4, graph=graph)
                films.append(c.film())
                stats.append(c.stats())
            print(f"[batch] chunks cap={cap} launches eager={stats[0]['trace_launches']} "
                  f"graph={stats[1]['trace_launches']} graph_ms={stats[1]['graph_ms']:.2f}", flush=True)
            assert stats[1]["graph_ms"] > 0
            for f in films[1:]:
                assert np.array_equal(f, films[0])
            for k in ("rays_extension", "rays_shadow"):
                assert stats[1][k] == stats[0][k], k


@pytest.mark.parametrize("guide_depth", [0, 1, 11])
def test_captured_chunks_equal_eager_guided(guide_depth):
    """Guided renders in captured chunks (inference and guided shading inside the graph, at
    the bounces guide_depth selects): with the same network the films and query counts equal
    the eager loop's; training renders (records, collection and Adam steps after every pass)
    reach the same network and films from the same start."""
    sc = phg.Scene.generate(phg.GEN_OCCLUDED_ROOM, seed=3, target_tris=20_000)
    W, H = 48, 40
    out = {}
    for graph in (False, True):
        with phg.Context(0) as c:
            c.upload(sc)
            c.film_clear(W, H)
            c.render(8, max_bounces=10000, seed=3, spp_per_pass=2, guided=True, train=True, guide_batch=4096,
                     guide_steps=2, guide_depth=guide_depth, graph=graph)
            trained = c.film()
            params = c.guide_get_params()
            c.film_clear(W, H)
            c.render(16, max_bounces=10000, seed=4, spp_per_pass=4, guided=True, guide_depth=guide_depth, graph=graph)
            out[graph] = (trained, params, c.film(), c.stats())
    st_e, st_g = out[False][3], out[True][3]
    print(f"[batch] guided chunks depth={guide_depth} queries eager={st_e['guide_queries']} "
          f"graph={st_g['guide_queries']}", flush=True)
    assert st_g["guide_queries"] == st_e["guide_queries"] > 0
    assert np.array_equal(out[True][1], out[False][1])
    assert np.array_equal(out[True][0], out[False][0])
    assert np.array_equal(out[True][2], out[False][2])
