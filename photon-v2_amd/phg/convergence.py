"""Time to a relative-MSE target: progressive renders against a converged reference.

relMSE = mean over pixels and channels of (I - R)^2 / (R^2 + 0.01), with R the reference
render.  The reference is itself a Monte Carlo mean; its FP64 film holds per-pixel sums of
squares, so the expected contribution of its own variance (the variance of its per-pixel
mean) is subtracted, which leaves an unbiased estimate of the relMSE against the converged
image.  Renders are timed on the host around the render calls only (film reads and the
error computation sit outside the clock), with checkpoints at doubling sample counts.
"""
from __future__ import annotations

import time

import numpy as np

EPS = 1e-2


def mean_and_var_of_mean(film: np.ndarray):
    """Per-pixel mean radiance and the variance of that mean from an FP64 film."""
    n = film[..., 3:4]
    safe = np.where(n > 0, n, 1.0)
    m = film[..., 0:3] / safe
    var = np.maximum(film[..., 4:7] / safe - m * m, 0.0) / np.maximum(safe - 1.0, 1.0)
    return m, var


def relmse(image: np.ndarray, ref_mean: np.ndarray, ref_var: np.ndarray | None = None) -> float:
    d2 = (image - ref_mean) ** 2
    if ref_var is not None:
        d2 = d2 - ref_var
    return float(np.mean(d2 / (ref_mean * ref_mean + EPS)))


def time_to(curve, target: float):
    """Seconds at which a (spp, seconds, relMSE) curve first reaches `target`, interpolated
    log-log between the checkpoints that bracket it; None if it never does."""
    prev = None
    for spp, t, e in curve:
        if e <= target:
            if prev is None or prev[2] <= 0 or e <= 0:
                return float(t)
            _, t0, e0 = prev
            a = (np.log(target) - np.log(e0)) / (np.log(e) - np.log(e0))
            return float(np.exp(np.log(t0) + a * (np.log(t) - np.log(t0))))
        prev = (spp, t, e)
    return None


def progressive(ctx, width: int, height: int, max_spp: int, ref_mean, ref_var, seed: int = 1,
                max_bounces: int = 10000, guided: bool = False, train_spp: int = 0, guide_prob: float = 0.5,
                guide_steps: int = 16, guide_lr: float = 1e-3, target: float | None = None,
                guide_depth: int = 0, totals: dict | None = None, train_spp_per_pass: int = 0, graph: bool = False):
    """Render 1, 2, 4, ... max_spp samples per pixel progressively on one film; returns the
    (spp, seconds, relMSE) checkpoints.  Guided renders train the network on every pass of
    the first `train_spp` samples and keep sampling with it afterwards.  Stops early once
    `target` is reached.  `totals`, if given, accumulates the renders' guiding statistics
    (guide_infer_ms, guide_queries, guide_train_ms, guide_records).  `train_spp_per_pass` sets
    the samples per pass of the training renders (0 = the library's choice); `graph` replays
    captured bounce loops / chunks (PHG_RENDER_GRAPH)."""
    ctx.film_clear(width, height)
    total, spp, nxt, out = 0.0, 0, 1, []
    while spp < max_spp:
        end = min(nxt, max_spp)
        while spp < end:
            train = guided and spp < train_spp
            stop = min(end, train_spp) if train else end
            t0 = time.perf_counter()
            ctx.render(stop - spp, max_bounces=max_bounces, seed=seed, spp_begin=spp, guided=guided, train=train,
                       guide_prob=guide_prob, guide_steps=guide_steps, guide_lr=guide_lr, guide_depth=guide_depth,
                       spp_per_pass=train_spp_per_pass if train else 0, graph=graph)
            total += time.perf_counter() - t0
            if totals is not None:
                st = ctx.stats()
                for k in ("guide_infer_ms", "guide_queries", "guide_train_ms", "guide_records"):
                    totals[k] = totals.get(k, 0) + st[k]
            spp = stop
        m, _ = mean_and_var_of_mean(ctx.film())
        e = relmse(m, ref_mean, ref_var)
        out.append((spp, total, e))
        if target is not None and e <= target:
            break
        nxt *= 2
    return out


def progressive_batch(ctx, jobs, max_spp: int, refs, seed_base: int = 1000, guided: bool = False,
                      train_spp: int = 0, guide_prob: float = 0.5, guide_steps: int = 16, guide_lr: float = 1e-3,
                      guide_depth: int = 0, graph: bool = False, target: float | None = None,
                      totals: dict | None = None):
    """progressive() for several jobs rendered together (phg_render_batch), each against its
    own reference (mean, variance of the mean).  Checkpoints at 1, 2, 4, ... max_spp samples
    per pixel of every job; each increment renders with seeds of its own (independent samples)
    and its films are added on the host (films hold sums, so the sum is the film of all the
    samples).  relMSE is the mean over every pixel and channel of all jobs.  Guided renders
    train on the increments below `train_spp`.  Returns the (spp, seconds, relMSE) checkpoints."""
    films = None
    total, spp, nxt, inc, out = 0.0, 0, 1, 0, []
    while spp < max_spp:
        end = min(nxt, max_spp)
        while spp < end:
            train = guided and spp < train_spp
            stop = min(end, train_spp) if train else end
            batch = [dict(j, spp=stop - spp, seed=(seed_base + inc * len(jobs) + k) % (1 << 24))
                     for k, j in enumerate(jobs)]
            t0 = time.perf_counter()
            ctx.render_batch(batch, graph=graph, guided=guided, train=train, guide_prob=guide_prob,
                             guide_steps=guide_steps, guide_lr=guide_lr, guide_depth=guide_depth)
            total += time.perf_counter() - t0
            if totals is not None:
                st = ctx.stats()
                for k in ("guide_infer_ms", "guide_queries", "guide_train_ms", "guide_records"):
                    totals[k] = totals.get(k, 0) + st[k]
            f = [ctx.batch_film(k) for k in range(len(jobs))]
            films = f if films is None else [a + b for a, b in zip(films, f)]
            spp, inc = stop, inc + 1
        num, cnt = 0.0, 0
        for film, (R, vR) in zip(films, refs):
            m, _ = mean_and_var_of_mean(film)
            d2 = (m - R) ** 2 - vR
            num += float(np.sum(d2 / (R * R + EPS)))
            cnt += d2.size
        e = num / cnt
        out.append((spp, total, e))
        if target is not None and e <= target:
            break
        nxt *= 2
    return out
