"""Restatement of the reference's stratified sample source, per sample batch — test
infrastructure only (tests/ use it; the product never does).

One batch of a W x H image (ReceiverSamplingWork.cpp:76-84) draws N = W * H samples from two
stages of SGStratified:
  raster stage, 2-D, size hints (W, H): one jittered point per pixel stratum, then the
      samples are shuffled (SGStratified.cpp:42-90, SamplesND.h:101-113);
  ray stage, 5-D, identity size hints: SGStratified has no >= 3-D generator, so the stage is
      split into 2-D, 2-D and 1-D stages (SampleGenerator.cpp:81-135, popFirstND), each
      revised by reviseDimSizeHints (SGStratified.cpp:112-152) -- a 2-D stage gets
      floor(sqrt(N)) x floor(sqrt(N)) strata, filled in whole passes with the remainder
      uniformly random -- and each shuffled on its own; a 1-D stage has N strata
      (SGStratified.cpp:24-40).
Sample i of the batch pairs raster point i with ray-stage row i, whose 5 values are the first
dimensions its SampleFlow hands out (SampleFlow.h:180-190); later dimensions are uniform
random.  The shuffles make the pixel a sample lands in independent of its ray-stage row.

The random source is numpy's generator in place of math::Random: the restatement fixes the
joint distribution of one batch, not the reference's bit stream.
"""
from __future__ import annotations

import math

import numpy as np


def revise_dim_size_hints(num_samples: int, hints):
    """SGStratified::reviseDimSizeHints: scale the hints towards num_samples (never above)."""
    hints = [max(int(h), 1) for h in hints]
    size = 1
    for h in hints:
        size *= h
    shrink = (num_samples / size) ** (1.0 / len(hints))
    out = [max(int(shrink * h), 1) for h in hints]
    while math.prod(out) > num_samples:
        k = int(np.argmax(out))
        out[k] = out[k] - 1 if out[k] > 1 else 1
    return out


def stratified_2d(n: int, sx: int, sy: int, rng) -> np.ndarray:
    """SGStratified::genSamples2D: whole passes over sx x sy jittered strata (x fastest), the
    remainder uniform random, then a shuffle of the n samples."""
    out = np.empty((n, 2))
    per = sx * sy
    passes = n // per
    if passes:
        ys, xs = np.meshgrid(np.arange(sy), np.arange(sx), indexing="ij")
        base = np.stack([xs.ravel(), ys.ravel()], 1).astype(np.float64)
        grid = np.tile(base, (passes, 1))
        jitter = rng.random((passes * per, 2))
        out[:passes * per, 0] = (grid[:, 0] + jitter[:, 0]) / sx
        out[:passes * per, 1] = (grid[:, 1] + jitter[:, 1]) / sy
    out[passes * per:] = rng.random((n - passes * per, 2))
    return out[rng.permutation(n)]


def stratified_1d(n: int, rng) -> np.ndarray:
    """SGStratified::genSamples1D: n jittered strata, shuffled."""
    return ((np.arange(n) + rng.random(n)) / n)[rng.permutation(n)]


def stratified_batch(width: int, height: int, rng):
    """One batch: (pixel index of each sample, raster offset in the pixel (n, 2), first five
    SampleFlow dimensions (n, 5))."""
    n = width * height
    sx, sy = revise_dim_size_hints(n, [width, height])
    raster = stratified_2d(n, sx, sy, rng)
    px = np.minimum((raster[:, 0] * width).astype(np.int64), width - 1)
    py = np.minimum((raster[:, 1] * height).astype(np.int64), height - 1)
    offset = np.stack([raster[:, 0] * width - px, raster[:, 1] * height - py], 1)
    a, b = revise_dim_size_hints(n, [1, 1])
    flow = np.empty((n, 5))
    flow[:, 0:2] = stratified_2d(n, a, b, rng)
    flow[:, 2:4] = stratified_2d(n, a, b, rng)
    flow[:, 4] = stratified_1d(n, rng)
    return py * width + px, offset, flow
