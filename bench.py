"""Benchmark of the GPU BNEEPT hot path (one JSON line on stdout).

Workload (one step): plain BNEEPT over the 1M-triangle procedural scene — 64 displaced
icospheres + tessellated ground (2^20 triangles) + 16 area-light quads, GGX conductors and
dielectrics — at 1024x1024, 64 samples per pixel, 8 bounces, FP32 shading, FP64 film.
Inputs are resident on the GPU before timing starts (scene and BVH uploaded once).
The headline value times the production mode: the autotuner's pass size (64M paths on this
workload, so the 64-spp render is one pass) and each pass replays one captured HIP graph of
the whole bounce loop.  The same steps also run eagerly (one host round trip
per bounce) with HIP events around every kernel launch; those give the kernel times.

Reported beside the throughput:
  roofline      the closest-hit traversal kernel: algorithmic bytes
                (48 per ray + 128 per node visited + 48 per triangle tested, counted by the
                same kernel in a counted render of the same paths) over its summed HIP-event
                launch durations, against the 8 TB/s HBM peak.
  cpu_baseline  the CPU oracle (scalar C++ restatement of the reference path) on this
                host's cores, over a bounded sample of the same image (every k-th row).
  batch         48 independent jobs (procedural scenes of 1e4..2e6 triangles, 128^2..1024^2,
                4..256 spp, log-uniform, depth 8) rendered by phg_render_batch in shared passes:
                aggregate Msamples/s and per-job completion p50 / p99.
  scale         guided + batched at scale: 8 views of uneven resolution (2048^2 down to 640x360)
                of a 20M-triangle terrain world (256 displaced icospheres, FBM terrain grid, 64
                area lights), depth 12, one network shared by all views and trained online, in
                shared passes; plain captured-graph passes of the same views beside it.
  convergence   time to relMSE 0.01 on the occluded-light room (~199k triangles, light behind
                a wall with a 0.1-wide slit, GGX conductors), 1024^2, guided vs plain BNEEPT: both
                render progressively against a plain reference of the same scene; the guided
                run trains its network online (included in its time).
"""
import argparse
import json
import os
import platform
import sys
import time
import traceback

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "photon-v2_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBPS = 8000.0
COUNTER_PROFILE = os.path.join(ROOT, "profiles", "bneept_procedural1M_counters.json")
# VALU issue peak: one wave64 vector instruction per compute unit per clock (4 SIMDs, each
# taking 4 clocks per wave64 instruction), 256 compute units at the 2.4 GHz peak engine clock
VALU_PEAK_GINST = 256 * 2.4


def profiled_kernel(kernel_prefix, config):
    """(per-launch counter figures, profiled configuration) of a kernel of this workload from
    the committed rocprofv3 profile (profiles/bneept_procedural1M_counters.json: trace time,
    VALU instructions, VALU busy fraction, lane utilization, bytes past L2, L2 hit rate) under
    launch configuration `config` ("leaf,width,block,paths_per_pass").  When that exact
    configuration was not profiled, one with the same pass size and block size stands in, then
    one with the same pass size (the BVH leaf size and the block size move these figures by
    under 2%); (None, None) otherwise."""
    try:
        with open(COUNTER_PROFILE) as f:
            prof = json.load(f)
    except (OSError, ValueError):
        return None, None
    if not config:
        return None, None
    configs = prof.get("configs", {})
    leaf, width, block, paths = config.split(",")
    same_pb = sorted(k for k in configs if k.split(",")[2:] == [block, paths])
    same_p = sorted(k for k in configs if k.split(",")[3] == paths)
    for key in [config] + same_pb + same_p:
        for name, v in configs.get(key, {}).get("kernels", {}).items():
            if name.removeprefix("void ").startswith(kernel_prefix):
                return v, key
    return None, None


# Algorithmic bytes of one shading-kernel vertex (k_shade), the most it moves: the queue entry
# (path id 4, hit 16, ray direction 16), path state read (throughput, previous vertex 16 each,
# generator state 8, segment index 4, radiance 16), triangle record 48 and material 64, next
# event (emitter-pick CDF 16, emitter 32, emitter triangle 8, light triangle 48), path state
# written (throughput, weight, previous vertex 16 each, generator state 8), and the two queue
# pushes (extension ray 32 + id 4, shadow ray 32 + id 4 + contribution 16).
SHADE_BYTES_PER_VERTEX = (4 + 16 + 16) + (16 + 16 + 8 + 4 + 16) + (48 + 64) + (16 + 32 + 8 + 48) + \
    (16 + 16 + 16 + 8) + (32 + 4 + 32 + 4 + 16)


METRIC = "Msamples/s at 1 GPU; time to relMSE 0.01, guided vs plain GPU BNEEPT"


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def lscpu():
    """Model name and core counts as lscpu reports them (empty when lscpu is missing)."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return {}
    keep = {"Model name": "model", "CPU(s)": "cpus", "Thread(s) per core": "threads_per_core",
            "Core(s) per socket": "cores_per_socket", "Socket(s)": "sockets"}
    info = {}
    for line in out.splitlines():
        k, _, v = line.partition(":")
        if k.strip() in keep:
            info[keep[k.strip()]] = v.strip()
    return info


def batch_leg(args):
    """Many independent render jobs in shared wavefront passes."""
    import phg
    from phg import workloads as wl

    specs, jobs = wl.batch_workload(args.batch_jobs, seed=0)
    t0 = time.time()
    scenes = wl.generate_scenes(specs)
    gen_s = time.time() - t0
    with phg.Context(0) as ctx:
        t0 = time.time()
        best, trials = ctx.autotune_batch(scenes, jobs, sample_spp=16)
        tune_s = time.time() - t0
        untuned = next(t for t in trials if t["max_leaf_tris"] == 4 and t["width"] == 4 and t["trace_block"] == 128
                       and t["paths_per_pass"] == 32 << 20)
        print(f"[bench] batch autotune: {len(trials)} trials in {tune_s:.1f}s, best {best} "
              f"(default {untuned['msamples_per_s']:.0f} Msamples/s on the sample)", file=sys.stderr, flush=True)
        t0 = time.time()
        ctx.upload_many(scenes, max_leaf_tris=best["max_leaf_tris"], width=best["width"])
        up_s = time.time() - t0
        ntris = sum(s.num_tris for s in scenes)
        del scenes
        print(f"[bench] batch: {len(jobs)} jobs, {ntris} triangles, scenes {gen_s:.1f}s, upload+BVH {up_s:.1f}s",
              file=sys.stderr, flush=True)
        ctx.render_batch(jobs, graph=True)          # warm-up (same work)
        t0 = time.perf_counter()
        res = ctx.render_batch(jobs, graph=True)
        wall = time.perf_counter() - t0
        st = ctx.stats()
        t0 = time.perf_counter()
        eres = ctx.render_batch(jobs)
        ewall = time.perf_counter() - t0
        # the same jobs packed shortest first: the small jobs finish in the first passes
        t0 = time.perf_counter()
        sres = ctx.render_batch(jobs, graph=True, order=phg.BATCH_ORDER_SHORTEST_FIRST)
        swall = time.perf_counter() - t0
    paths = sum(p for _, p in res)
    done = [d for d, _ in res]
    print(f"[bench] batch: {paths / 1e6:.0f}M paths in {wall:.2f}s, {st['batch_passes']} passes", file=sys.stderr,
          flush=True)
    return {"jobs": len(jobs), "triangles": ntris, "paths": paths, "seconds": round(wall, 4),
            "msamples_per_s": round(paths / wall / 1e6, 3), "mrays_per_s": round(
                (st["rays_primary"] + st["rays_extension"] + st["rays_shadow"]) / wall / 1e6, 3),
            "p50_done_ms": round(wl.percentile(done, 50), 2), "p99_done_ms": round(wl.percentile(done, 99), 2),
            "passes": st["batch_passes"], "order": "given", "max_bounces": 8,
            "mode": "captured-graph passes", "tuning": best, "autotune_trials": len(trials),
            "autotune_seconds": round(tune_s, 2), "untuned_sample_msamples_per_s": round(untuned["msamples_per_s"], 3),
            "eager": {"seconds": round(ewall, 4), "msamples_per_s": round(paths / ewall / 1e6, 3),
                      "p50_done_ms": round(wl.percentile([d for d, _ in eres], 50), 2),
                      "p99_done_ms": round(wl.percentile([d for d, _ in eres], 99), 2)},
            "shortest_first": {"seconds": round(swall, 4), "msamples_per_s": round(paths / swall / 1e6, 3),
                               "p50_done_ms": round(wl.percentile([d for d, _ in sres], 50), 2),
                               "p99_done_ms": round(wl.percentile([d for d, _ in sres], 99), 2),
                               "mode": "captured-graph passes, jobs packed shortest first"}}


SCALE_VIEWS = [(2048, 2048), (1920, 1080), (1600, 900), (1280, 720), (1024, 1024), (960, 540), (800, 800), (640, 360)]


# Algorithmic multiply-add work of the guiding MLP (64-64-64-64-64-41, 2 flops per multiply-add):
# inference = forward; a training record = forward + dW for every layer + dZ propagation
# through layers 4..1.
MLP_FWD_FLOPS = 2 * (4 * 64 * 64 + 64 * 41)
MLP_TRAIN_FLOPS = 2 * MLP_FWD_FLOPS + 2 * (64 * 41 + 3 * 64 * 64)
MFMA_BF16_PEAK_TFLOPS = 2500.0


def guide_prob_label(p: float):
    """The lobes' selection probability as reported: the fixed value, or "learned" (negative)."""
    return "learned per query (network output 40, in [0.05, 0.95])" if p < 0 else p


def ordinal(n: int) -> str:
    return f"{n}{'th' if 10 <= n % 100 <= 20 else {1: 'st', 2: 'nd', 3: 'rd'}.get(n % 10, 'th')}"


def mlp_utilisation(st):
    """Matrix-core utilisation of the guiding MLP over a render's inference and training time."""
    out = {"peak_tflops": MFMA_BF16_PEAK_TFLOPS, "dtype": "bf16 (fp32 accumulation)"}
    if st["guide_infer_ms"] > 0:
        t = st["guide_queries"] * MLP_FWD_FLOPS / (st["guide_infer_ms"] / 1e3) / 1e12
        out["infer"] = {"queries": st["guide_queries"], "ms": round(st["guide_infer_ms"], 2), "tflops": round(t, 2),
                        "frac": round(t / MFMA_BF16_PEAK_TFLOPS, 4)}
    if st["guide_train_ms"] > 0:
        t = st["guide_records"] * MLP_TRAIN_FLOPS / (st["guide_train_ms"] / 1e3) / 1e12
        out["train"] = {"records": st["guide_records"], "ms": round(st["guide_train_ms"], 2), "tflops": round(t, 2),
                        "frac": round(t / MFMA_BF16_PEAK_TFLOPS, 4),
                        "timed": "record collection + gradient kernels + Adam (HIP events)"}
    return out


def adam_step_utilisation(ctx, np, n=1 << 18):
    """Matrix-core utilisation of one training step (forward, backward, reductions, Adam) on n
    random records, timed as the difference of a 33-step and a 1-step phg_guide_train call."""
    rng = np.random.default_rng(0)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rec = np.concatenate([rng.random((n, 9)), d, rng.exponential(1.0, (n, 1)), rng.exponential(0.5, (n, 1)),
                          np.full((n, 1), 0.5), rng.exponential(0.5, (n, 1))], axis=1).astype(np.float32)
    t0 = time.perf_counter()
    ctx.guide_train(rec, lr=1e-3, steps=1)
    t1 = time.perf_counter()
    ctx.guide_train(rec, lr=1e-3, steps=33)
    t2 = time.perf_counter()
    ms = max(((t2 - t1) - (t1 - t0)) / 32 * 1e3, 1e-6)
    t = n * MLP_TRAIN_FLOPS / (ms / 1e3) / 1e12
    return {"records": n, "ms": round(ms, 4), "tflops": round(t, 2), "frac": round(t / MFMA_BF16_PEAK_TFLOPS, 4),
            "timed": "one Adam step alone: host clock, 33 steps minus 1 step over the same uploaded records"}


def scale_leg(args):
    """20M-triangle terrain world, 8 views of uneven resolution, guided with one shared network."""
    import numpy as np
    import phg

    t0 = time.time()
    scene = phg.Scene.generate(phg.GEN_TERRAIN)
    gen_s = time.time() - t0
    jobs = []
    for k, (w, h) in enumerate(SCALE_VIEWS):
        a = 2 * np.pi * k / len(SCALE_VIEWS)
        cam = phg.camera_look_at([55 * np.sin(a), 6 + 2 * k, -55 * np.cos(a)], [0, -8, 0], fov_deg=60)
        jobs.append(dict(scene=0, width=w, height=h, spp=args.scale_spp, max_bounces=12, seed=100 + k, camera=cam))
    npix = sum(w * h for w, h in SCALE_VIEWS)
    with phg.Context(0) as ctx:
        t0 = time.time()
        ctx.upload_many([scene])
        up_s = time.time() - t0
        # the plain passes run 64M paths at a time (the headline's tuned pass size; tuning this
        # scene would rebuild its 20M-triangle BVH per trial); training passes stay smaller
        ctx.set_tuning(trace_block=128, paths_per_pass=64 << 20)
        st0 = ctx.stats()
        ntris = scene.num_tris
        del scene
        print(f"[bench] scale: {ntris} triangles (generated {gen_s:.1f}s), BVH {st0['bvh_build_ms'] / 1e3:.1f}s, "
              f"upload {up_s:.1f}s, {len(jobs)} views, {npix} pixels", file=sys.stderr, flush=True)
        ctx.render_batch([dict(j, spp=1) for j in jobs], count_traversal=True)
        cs = ctx.stats()
        rays1 = cs["rays_primary"] + cs["rays_extension"]
        nodes_per_ray, tris_per_ray = cs["trace_node_visits"] / rays1, cs["trace_tri_tests"] / rays1
        conv = scale_convergence(args, ctx, jobs) if args.scale_ref_spp > 0 else None
        t0 = time.perf_counter()
        gres = ctx.render_batch(jobs, guided=True, train=True, guide_steps=32)
        g_s = time.perf_counter() - t0
        gs = ctx.stats()
        paths = sum(p for _, p in gres)
        print(f"[bench] scale guided+train: {paths / 1e6:.0f}M paths in {g_s:.2f}s, guide infer "
              f"{gs['guide_infer_ms']:.0f} ms, train {gs['guide_train_ms']:.0f} ms", file=sys.stderr, flush=True)
        ctx.render_batch(jobs, graph=True)            # warm-up
        t0 = time.perf_counter()
        ctx.render_batch(jobs, graph=True)
        p_s = time.perf_counter() - t0
        ps = ctx.stats()
        ctx.render_batch(jobs)
        es = ctx.stats()
    rays_c = es["rays_primary"] + es["rays_extension"]
    bytes_c = rays_c * (48 + 128 * nodes_per_ray + 48 * tris_per_ray)
    ach = bytes_c / (es["trace_ms"] / 1e3) / 1e9
    print(f"[bench] scale plain: {p_s:.2f}s, k_extend {ach:.0f} GB/s", file=sys.stderr, flush=True)
    return {"scene": f"terrain world, {ntris} triangles, 64 area lights", "views": [list(v) for v in SCALE_VIEWS],
            "spp": args.scale_spp, "max_bounces": 12, "paths": paths,
            "guided_trained": {"seconds": round(g_s, 4), "msamples_per_s": round(paths / g_s / 1e6, 3),
                               "guide_infer_ms": round(gs["guide_infer_ms"], 2),
                               "guide_train_ms": round(gs["guide_train_ms"], 2), "guide_steps": gs["guide_steps"],
                               "mlp": mlp_utilisation(gs)},
            "plain_graph": {"seconds": round(p_s, 4), "msamples_per_s": round(paths / p_s / 1e6, 3),
                            "mrays_per_s": round((ps["rays_primary"] + ps["rays_extension"] + ps["rays_shadow"]) / p_s / 1e6, 3)},
            "k_extend_roofline": {"achieved_gbps": round(ach, 1), "frac": round(ach / HBM_PEAK_GBPS, 4),
                                  "nodes_per_ray": round(nodes_per_ray, 3), "tris_per_ray": round(tris_per_ray, 3),
                                  "timed_in": "eager plain render of the same views (HIP events)"},
            "bvh_build_s": round(st0["bvh_build_ms"] / 1e3, 2), "convergence": conv}


def scale_convergence(args, ctx, jobs):
    """Time to relMSE 0.01 over all 8 views, plain (captured bounce loops) and guided (one
    network for all views, trained on the first samples, lobes at the first path vertex),
    against a plain reference per view with its own variance subtracted.  Runs before any
    other guided render on this context, so the network starts untrained."""
    from phg import convergence as cv
    t0 = time.perf_counter()
    for s0 in range(0, args.scale_ref_spp, 256):
        ctx.render_batch([dict(j, spp=min(256, args.scale_ref_spp - s0), seed=5000 + 16 * s0 + k)
                          for k, j in enumerate(jobs)], graph=True)
        films = [ctx.batch_film(k) for k in range(len(jobs))]
        acc = films if s0 == 0 else [a + b for a, b in zip(acc, films)]
        print(f"[bench] scale reference: {s0 + 256} spp, {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    refs = [cv.mean_and_var_of_mean(f) for f in acc]
    del acc, films
    ref_s = time.perf_counter() - t0
    plain = cv.progressive_batch(ctx, jobs, args.scale_conv_max_spp, refs, seed_base=100000, graph=True, target=0.005)
    gtot = {}
    guided = cv.progressive_batch(ctx, jobs, args.scale_conv_max_spp, refs, seed_base=200000, guided=True,
                                  train_spp=16, guide_steps=32, guide_depth=1, guide_prob=args.guide_prob,
                                  target=0.005, totals=gtot, graph=True)
    tp, tg = cv.time_to(plain, 0.01), cv.time_to(guided, 0.01)
    for name, cur in (("plain", plain), ("guided", guided)):
        print(f"[bench] scale {name}: " + " ".join(f"{s}spp/{t:.2f}s/{e:.2e}" for s, t, e in cur), file=sys.stderr,
              flush=True)
    return {"reference_spp": args.scale_ref_spp, "reference_s": round(ref_s, 2), "relmse_target": 0.01,
            "plain_s": None if tp is None else round(tp, 4), "guided_s": None if tg is None else round(tg, 4),
            "guided_speedup": None if (tp is None or tg is None) else round(tp / tg, 3),
            "guided": {"guide_prob": guide_prob_label(args.guide_prob), "guide_depth": 1, "train_spp": 16,
                       "steps_per_pass": 32, "mlp": mlp_utilisation(gtot)},
            "plain_curve": [[s, round(t, 4), float(f"{e:.4e}")] for s, t, e in plain],
            "guided_curve": [[s, round(t, 4), float(f"{e:.4e}")] for s, t, e in guided]}


def convergence_leg(args):
    """Time to relMSE 0.01, plain vs guided, on the occluded-light room."""
    import numpy as np
    import phg
    from phg import convergence as cv

    res = args.conv_res
    scene = phg.Scene.generate(phg.GEN_OCCLUDED_ROOM)
    with phg.Context(0) as ctx:
        ctx.upload(scene)
        ctx.film_clear(res, res)
        t0 = time.perf_counter()
        chunk = 1024
        for s0 in range(0, args.conv_ref_spp, chunk):
            ctx.render(min(chunk, args.conv_ref_spp - s0), seed=7, spp_begin=s0)
            if args.conv_ref_spp > 8 * chunk and (s0 // chunk) % 8 == 7:
                print(f"[bench] convergence reference: {s0 + chunk} spp, {time.perf_counter() - t0:.1f}s",
                      file=sys.stderr, flush=True)
        ref_s = time.perf_counter() - t0
        R, vR = cv.mean_and_var_of_mean(ctx.film())
        print(f"[bench] convergence reference: {args.conv_ref_spp} spp at {res}x{res} in {ref_s:.1f}s "
              f"(its own relMSE {float(np.mean(vR / (R * R + cv.EPS))):.2e})", file=sys.stderr, flush=True)
        plain = cv.progressive(ctx, res, res, args.conv_max_spp, R, vR, seed=1, target=0.005, graph=True)
    with phg.Context(0) as ctx:
        ctx.upload(scene)
        gtot = {}
        guided = cv.progressive(ctx, res, res, args.conv_max_spp, R, vR, seed=1, guided=True,
                                train_spp=args.conv_train_spp, guide_prob=args.guide_prob, guide_steps=32, target=0.005,
                                guide_depth=1, totals=gtot, graph=True)
        # the captured chunks carry no per-kernel times: the trained network's inference rate
        # comes from one eager guided render of the same image
        ctx.film_clear(res, res)
        ctx.render(16, seed=99, guided=True, guide_depth=1, guide_prob=args.guide_prob)
        ist = ctx.stats()
        gtot["guide_infer_ms"], gtot["guide_queries"] = ist["guide_infer_ms"], ist["guide_queries"]
        # one Adam step on 2^18 records alone (after the measurements above; random records of
        # the same layout): host clock over 33 steps minus 1 step, each call uploading the same
        # records and synchronising at its end
        train_step = adam_step_utilisation(ctx, np)
    tp, tg = cv.time_to(plain, 0.01), cv.time_to(guided, 0.01)
    # per-sample error at the last checkpoint both curves reach: relMSE x spp (the variance of
    # one sample per pixel, relative), and seconds per sample
    common = max(s for s, _, _ in guided if any(s == q for q, _, _ in plain))
    pe = next((t, e) for s, t, e in plain if s == common)
    ge = next((t, e) for s, t, e in guided if s == common)
    per_sample = {"spp": common, "plain_relmse_x_spp": round(pe[1] * common, 4),
                  "guided_relmse_x_spp": round(ge[1] * common, 4),
                  "variance_gain": round(pe[1] / ge[1], 4), "plain_s": round(pe[0], 4), "guided_s": round(ge[0], 4),
                  "note": "guided time includes its training passes"}
    for name, cur in (("plain", plain), ("guided", guided)):
        print(f"[bench] {name}: " + " ".join(f"{s}spp/{t:.2f}s/{e:.2e}" for s, t, e in cur), file=sys.stderr,
              flush=True)
    return {"scene": f"occluded-light room, {scene.num_tris} triangles, GGX conductors, light behind a slit",
            "width": res, "height": res, "reference_spp": args.conv_ref_spp, "max_bounces": 10000,
            "relmse_target": 0.01, "plain_s": None if tp is None else round(tp, 4),
            "guided_s": None if tg is None else round(tg, 4),
            "guided_speedup": None if (tp is None or tg is None) else round(tp / tg, 3),
            "guided": {"guide_prob": guide_prob_label(args.guide_prob), "guide_depth": 1,
                       "train_spp": args.conv_train_spp, "steps_per_pass": 32,
                       "training_pass_paths": "<= 16M (16 spp at 1024^2)", "records": "first path vertex", "lr": 1e-3,
                       "batch": 1 << 18, "mlp": dict(mlp_utilisation(gtot), train_step=train_step)},
            "per_sample": per_sample,
            "mode": "both legs replay captured bounce chunks (8 bounces per graph) until the path queue drains",
            "plain_curve": [[s, round(t, 4), float(f"{e:.4e}")] for s, t, e in plain],
            "guided_curve": [[s, round(t, 4), float(f"{e:.4e}")] for s, t, e in guided]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=1024)
    ap.add_argument("--height", type=int, default=1024)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--bounces", type=int, default=8)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline sample time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-conv", action="store_true", help="skip the guided-vs-plain convergence leg")
    ap.add_argument("--guide-prob", type=float, default=-1.0,
                    help="guided legs: probability of sampling the learned lobes (negative: the network's own "
                         "per-query probability)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-jobs leg")
    ap.add_argument("--batch-jobs", type=int, default=48)
    ap.add_argument("--no-scale", action="store_true", help="skip the 20M-triangle multi-view leg")
    ap.add_argument("--scale-spp", type=int, default=16)
    ap.add_argument("--scale-ref-spp", type=int, default=1024,
                    help="spp per view of the 8-view reference for time to relMSE 0.01 (0 skips it)")
    ap.add_argument("--scale-conv-max-spp", type=int, default=1024)
    ap.add_argument("--no-tune", action="store_true", help="default launch and BVH parameters (no autotuner)")
    ap.add_argument("--tuning", default=None, metavar="LEAF,WIDTH,BLOCK,PATHS",
                    help="fixed BVH leaf size, BVH width, traversal block and paths per pass instead of the autotuner")
    ap.add_argument("--conv-res", type=int, default=1024)
    ap.add_argument("--conv-ref-spp", type=int, default=4096,
                    help="spp of the plain reference; its per-pixel variance of the mean is subtracted from every "
                         "relMSE, so 4096 spp estimates the error against the converged image without bias "
                         "(65536 spp at 1024^2 takes about 4 minutes)")
    ap.add_argument("--conv-max-spp", type=int, default=8192)
    ap.add_argument("--conv-train-spp", type=int, default=48)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import phg

    W, H, spp, B = args.width, args.height, args.spp, args.bounces
    t0 = time.time()
    scene = phg.Scene.generate(phg.GEN_PROCEDURAL, seed=2025)
    gen_s = time.time() - t0
    print(f"[bench] scene: {scene.num_tris} triangles, {len(scene.emitters)} lights ({gen_s:.1f}s)",
          file=sys.stderr, flush=True)
    ctx = phg.Context(0)
    tuned = None
    if args.tuning:
        leaf, width, block, ppp = (int(x) for x in args.tuning.split(","))
        ctx.set_tuning(trace_block=block, paths_per_pass=ppp)
        tuned = dict(max_leaf_tris=leaf, width=width, trace_block=block, paths_per_pass=ppp, trials=0, seconds=0.0)
    elif not args.no_tune:
        t0 = time.time()
        job = dict(scene=0, width=W, height=H, spp=spp, max_bounces=B, seed=1)
        best, trials = ctx.autotune_batch([scene], [job], sample_jobs=1, sample_spp=spp)
        tuned = dict(best, trials=len(trials), seconds=round(time.time() - t0, 2))
        print(f"[bench] autotune: {tuned}", file=sys.stderr, flush=True)
    t0 = time.time()
    if tuned:
        ctx.upload(scene, max_leaf_tris=tuned["max_leaf_tris"], width=tuned["width"])
    else:
        ctx.upload(scene)
    upload_s = time.time() - t0
    launch_tuning = tuned if tuned else ctx.tuning()
    traffic_key = (f"{tuned['max_leaf_tris']},{tuned['width']},{tuned['trace_block']},{tuned['paths_per_pass']}"
                   if tuned else None)
    prof_extend, profiled_key = profiled_kernel("phg::k_extend", traffic_key)
    prof_shadow, _ = profiled_kernel("phg::k_shadow", traffic_key)
    prof_shade, _ = profiled_kernel("phg::k_shade<false>", traffic_key)
    traffic_bytes = round(prof_extend["traffic_bytes"]) if prof_extend else None
    st0 = ctx.stats()
    print(f"[bench] BVH: {st0['bvh_nodes']} nodes, depth {st0['bvh_depth']}, build {st0['bvh_build_ms']:.0f} ms, "
          f"upload {upload_s:.1f}s", file=sys.stderr, flush=True)

    # counted render of one sample per pixel: node visits / triangle tests per traversal
    ctx.film_clear(W, H)
    ctx.render(1, max_bounces=B, seed=1, count_traversal=True)
    cs = ctx.stats()
    rays_closest_1 = cs["rays_primary"] + cs["rays_extension"]
    nodes_per_ray = cs["trace_node_visits"] / rays_closest_1
    tris_per_ray = cs["trace_tri_tests"] / rays_closest_1

    for _ in range(args.warmup):
        ctx.film_clear(W, H)
        ctx.render(spp, max_bounces=B, seed=1)
    times, trace_ms, trace_launch, rays = [], 0.0, 0, [0, 0, 0]
    for s in range(args.steps):
        ctx.film_clear(W, H)
        t = time.perf_counter()
        ctx.render(spp, max_bounces=B, seed=1)
        times.append(time.perf_counter() - t)
        st = ctx.stats()
        trace_ms += st["trace_ms"]
        trace_launch += st["trace_launches"]
        rays[0] += st["rays_primary"]
        rays[1] += st["rays_extension"]
        rays[2] += st["rays_shadow"]
        print(f"[bench] step {s}: {times[-1] * 1e3:.1f} ms, trace {st['trace_ms']:.1f} ms, "
              f"shadow {st['shadow_ms']:.1f} ms, shade {st['shade_ms']:.1f} ms", file=sys.stderr, flush=True)
    gtimes = []
    for s in range(args.steps):
        ctx.film_clear(W, H)
        t = time.perf_counter()
        ctx.render(spp, max_bounces=B, seed=1, graph=True)
        gtimes.append(time.perf_counter() - t)
        print(f"[bench] graph step {s}: {gtimes[-1] * 1e3:.1f} ms", file=sys.stderr, flush=True)
    total_s = sum(times)
    samples = W * H * spp * args.steps
    msps_eager = samples / total_s / 1e6
    graph_s = sum(gtimes)
    msps = samples / graph_s / 1e6
    rays_closest = rays[0] + rays[1]
    # algorithmic bytes of the closest-hit traversal kernel over all its launches
    bytes_closest = rays_closest * (32 + 16 + nodes_per_ray * 128 + tris_per_ray * 48)
    achieved = bytes_closest / (trace_ms / 1e3) / 1e9
    img = phg.develop(ctx.film())
    ctx.close()

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        # the timed CPU build: -O3 -march=native for this host, built here (outside any timing)
        orc.build_native(force=True)
        threads = os.cpu_count() or 1
        # probe, then size the row sample to about args.cpu_seconds of CPU work
        step = 64
        _, pst = orc.render(scene, W, H, spp, max_bounces=B, seed=1, nthreads=threads, row_step=256, native=True)
        per_path_s = pst["render_ms"] / 1e3 / max(pst["paths"], 1)
        want_paths = args.cpu_seconds / max(per_path_s, 1e-9)
        step = int(max(1, min(256, round(W * H * spp / max(want_paths, 1.0)))))
        _, cst = orc.render(scene, W, H, spp, max_bounces=B, seed=1, nthreads=threads, row_step=step, native=True)
        cpu = {"value": cst["paths"] / (cst["render_ms"] / 1e3) / 1e6, "unit": "Msamples/s", "cores": threads,
               "kind": "port",
               "sample": f"every {ordinal(step)} row of the same {W}x{H} image at {spp} spp ({cst['paths']} paths, "
                         f"{cst['render_ms'] / 1e3:.1f} s), oracle BNEEPT, {threads} threads on {cpu_model()}",
               "build": f"g++ {orc.NATIVE_FLAGS}", "lscpu": lscpu()}
        # one core, on a sample of about a third of the time
        per_path_1 = per_path_s * threads
        step1 = int(max(1, min(1024, round(W * H * spp / max(args.cpu_seconds / 3 / per_path_1, 1.0)))))
        _, sst = orc.render(scene, W, H, spp, max_bounces=B, seed=1, nthreads=1, row_step=step1, native=True)
        cpu["single_core"] = {"value": sst["paths"] / (sst["render_ms"] / 1e3) / 1e6, "unit": "Msamples/s",
                              "sample": f"every {ordinal(step1)} row ({sst['paths']} paths, {sst['render_ms'] / 1e3:.1f} s), "
                                        "1 thread"}

    def leg(fn):
        """An auxiliary leg that fails is reported in its JSON field (traceback on stderr); the
        headline line is still printed."""
        try:
            return fn(args)
        except Exception as e:  # noqa: BLE001
            traceback.print_exc()
            return {"error": f"{type(e).__name__}: {e}"}

    batch = None if args.no_batch else leg(batch_leg)
    scale = None if args.no_scale else leg(scale_leg)
    conv = None if args.no_conv else leg(convergence_leg)

    avg_launch_s = trace_ms / max(trace_launch, 1) / 1e3
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic_bytes, "cache_served": True,
        "kernel": "k_extend (closest-hit traversal)",
        "note": "achieved = algorithmic bytes (48 + 128 per node visit + 48 per triangle test, per ray) per "
                "launch / launch time: this scene's nodes and triangles (~85 MB) are served by L2 and the 256 MB "
                "Infinity Cache, so frac is not a fraction of what HBM delivers; hbm_measured is (traffic past L2 "
                "per launch) / launch time, and valu (the bound that binds) the vector-instruction issue rate "
                "against its peak",
        "timed_in": "the eager steps: HIP events on the library stream around each launch",
        "launches": trace_launch, "avg_launch_ms": round(avg_launch_s * 1e3, 4),
        "bytes_per_launch": round(bytes_closest / max(trace_launch, 1)),
        "nodes_per_ray": round(nodes_per_ray, 3), "tris_per_ray": round(tris_per_ray, 3),
        "binding": "valu",
    }
    if prof_extend:
        roofline["hbm_measured"] = {
            "achieved": round(prof_extend["traffic_bytes"] / avg_launch_s / 1e9, 1), "peak": HBM_PEAK_GBPS,
            "unit": "GB/s", "frac": round(prof_extend["traffic_bytes"] / avg_launch_s / 1e9 / HBM_PEAK_GBPS, 4),
            "l2_hit_rate": round(prof_extend["l2_hit_rate"], 4)}
        ach_v = prof_extend["insts_valu"] / avg_launch_s / 1e9
        roofline["valu"] = {
            "achieved": round(ach_v, 1), "peak": VALU_PEAK_GINST, "unit": "G wave64 VALU instructions/s",
            "frac": round(ach_v / VALU_PEAK_GINST, 4), "busy": round(prof_extend["valu_busy"], 4),
            "lane_utilization": round(prof_extend["lane_utilization"], 4),
            "insts_per_launch": round(prof_extend["insts_valu"])}
        roofline["profile"] = (f"profiles/bneept_procedural1M_counters.json, launch configuration {profiled_key} "
                               f"(this run: {traffic_key}); per-launch averages of rocprofv3 SQ_INSTS_VALU, "
                               "SQ_ACTIVE_INST_VALU, SQ_THREAD_CYCLES_VALU, GRBM_GUI_ACTIVE, FETCH_SIZE x2.0 + "
                               "WRITE_SIZE x1.0, TCC_HIT/MISS over the 64-spp render's launches")
    kern = {}
    for name, pr in (("k_shadow", prof_shadow), ("k_shade", prof_shade)):
        if not pr:
            continue
        t = pr["trace_avg_ms"] / 1e3
        e = {"avg_launch_ms": round(pr["trace_avg_ms"], 4),
             "valu": {"frac": round(pr["insts_valu"] / t / 1e9 / VALU_PEAK_GINST, 4), "busy": round(pr["valu_busy"], 4),
                      "lane_utilization": round(pr["lane_utilization"], 4)},
             "hbm_measured": {"frac": round(pr["traffic_bytes"] / t / 1e9 / HBM_PEAK_GBPS, 4),
                              "l2_hit_rate": round(pr["l2_hit_rate"], 4)}}
        if name == "k_shade":
            # vertices shaded per launch: every queued extension ray (camera rays included)
            verts = (rays[0] + rays[1]) / max(trace_launch, 1)
            alg = verts * SHADE_BYTES_PER_VERTEX / t / 1e9
            e["algorithmic"] = {"bytes_per_vertex": SHADE_BYTES_PER_VERTEX, "vertices_per_launch": round(verts),
                                "achieved": round(alg, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                                "frac": round(alg / HBM_PEAK_GBPS, 4)}
        kern[name] = e
    if kern:
        roofline["other_kernels"] = kern
        roofline["other_kernels_timed_in"] = "launch times from the committed kernel trace of the same configuration"

    out = {
        "metric": METRIC,
        "value": round(msps, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(graph_s / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"plain BNEEPT, procedural {scene.num_tris}-triangle scene, {W}x{H}, {spp} spp, "
                               f"{B} bounces", "width": W, "height": H, "spp": spp, "max_bounces": B,
                   "triangles": scene.num_tris, "lights": int(len(scene.emitters))},
        "mode": "each pass replays one captured HIP graph of generate, 9 bounces and film",
        "tuning": launch_tuning,
        "mrays_per_s": round(sum(rays) / graph_s / 1e6, 3),
        "mrays_per_s_by_class": {"primary": round(rays[0] / graph_s / 1e6, 3),
                                 "extension": round(rays[1] / graph_s / 1e6, 3),
                                 "shadow": round(rays[2] / graph_s / 1e6, 3)},
        "eager": {"ms_per_step": round(total_s / args.steps * 1e3, 3), "msamples_per_s": round(msps_eager, 3),
                  "note": "same passes with a host round trip per bounce and HIP events around every launch "
                          "(the roofline's kernel times come from these steps)"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batch": batch,
        "scale": scale,
        "convergence": conv,
        "image_mean": [round(float(x), 6) for x in img.mean(axis=(0, 1))],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
