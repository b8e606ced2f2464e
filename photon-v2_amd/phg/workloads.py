"""Benchmark workloads built from the procedural scene generator (phg_scene_generate).

batch_workload: the batched-rendering workload — n independent jobs, job j a procedural
scene generated with seed j whose triangle count is log-uniform in [1e4, 2e6] (icosphere
instances scaled with it, the tessellated ground filling the rest), a square image whose side
is log-uniform in [128, 1024], spp log-uniform in [4, 256], depth 8.
"""
from __future__ import annotations

import numpy as np


def batch_workload(n: int = 48, seed: int = 0, max_bounces: int = 8):
    """Returns (scene specs, jobs): spec j = generator arguments of scene j, job j = the
    phg_render_batch job over scene j."""
    rng = np.random.default_rng(seed)
    specs, jobs = [], []
    for j in range(n):
        tris = int(round(np.exp(rng.uniform(np.log(1e4), np.log(2e6)))))
        side = int(round(np.exp(rng.uniform(np.log(128), np.log(1024)))))
        spp = int(round(np.exp(rng.uniform(np.log(4), np.log(256)))))
        instances = max(1, int(round(64 * tris / float(1 << 20))))
        specs.append(dict(seed=j, target_tris=tris, num_instances=instances, num_lights=16))
        jobs.append(dict(scene=j, width=side, height=side, spp=spp, max_bounces=max_bounces, seed=j))
    return specs, jobs


def generate_scenes(specs):
    import phg
    return [phg.Scene.generate(phg.GEN_PROCEDURAL, **s) for s in specs]


def percentile(values, q):
    """Nearest-rank percentile (q in [0, 100])."""
    v = np.sort(np.asarray(values, np.float64))
    k = max(0, int(np.ceil(q / 100.0 * len(v))) - 1)
    return float(v[k])
