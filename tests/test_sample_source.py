"""The reference's stratified sample source (oracle/sample_source.py) against what the
mathematics fixes, and the property that makes it interchangeable with the renderer's
independent per-pixel samples for every per-pixel statistic: within a batch each pixel gets
one sample whose raster offset and first five dimensions are uniformly distributed, and
batches are independent.  Stratification only correlates the pixels of one batch, which the
image mean shows and no per-pixel mean or variance does."""
import math

import numpy as np
import pytest

from oracle import sample_source as ss


@pytest.mark.parametrize("n,hints,want", [
    (65536, [256, 256], [256, 256]),   # raster stage of a 256 x 256 image: one stratum per pixel
    (65536, [1, 1], [256, 256]),       # 2-D ray stage: sqrt(N) x sqrt(N)
    (1000, [1, 1], [31, 31]),          # floor(sqrt(1000)) = 31, 961 <= 1000
    (12, [4, 3], [4, 3]),
    (10, [4, 3], [3, 2]),              # shrink sqrt(10 / 12) = 0.913 -> floor(3.65), floor(2.74)
    (1, [8, 8], [1, 1]),
])
def test_revised_size_hints(n, hints, want):
    got = ss.revise_dim_size_hints(n, hints)
    assert got == want
    assert 1 <= math.prod(got) <= n


def test_one_raster_sample_per_pixel_and_full_strata():
    rng = np.random.default_rng(3)
    for w, h in [(16, 9), (7, 7), (1, 5)]:
        pix, off, flow = ss.stratified_batch(w, h, rng)
        n = w * h
        assert np.array_equal(np.bincount(pix, minlength=n), np.ones(n))
        assert off.min() >= 0.0 and off.max() < 1.0
        assert flow.min() >= 0.0 and flow.max() < 1.0
        # 2-D ray stages: every one of the k x k strata holds at least N // k^2 samples
        k = math.isqrt(n)
        passes = n // (k * k)
        for d in (0, 2):
            cell = np.minimum((flow[:, d] * k).astype(int), k - 1) * k + np.minimum((flow[:, d + 1] * k).astype(int), k - 1)
            assert np.bincount(cell, minlength=k * k).min() >= passes
        # 1-D stage: exactly one sample per stratum
        assert np.array_equal(np.sort(np.minimum((flow[:, 4] * n).astype(int), n - 1)), np.arange(n))


def test_per_pixel_statistics_match_independent_sampling():
    """Per pixel, over batches: mean and variance of an integrand of the raster offset and the
    five stratified dimensions equal those of independent uniform samples (closed forms), while
    the variance of the image mean drops well below the independent value Var[f] / N."""
    rng = np.random.default_rng(11)
    w, h, batches = 6, 6, 4000
    n = w * h

    def f(off, flow):  # E[f] = 0.5 + 0.5 + 0.25 + 0.3 = 1.55 under independent uniforms
        return off[:, 0] + flow[:, 0] + flow[:, 2] * flow[:, 4] + (flow[:, 1] < 0.3)

    mean_f = 1.55
    # Var[f] = 1/12 + 1/12 + Var[u2 u4] + 0.3 * 0.7, Var[u2 u4] = 1/9 - 1/16
    var_f = 1 / 12 + 1 / 12 + (1 / 9 - 1 / 16) + 0.21
    vals = np.empty((batches, n))
    for b in range(batches):
        pix, off, flow = ss.stratified_batch(w, h, rng)
        vals[b, pix] = f(off, flow)
    m = vals.mean(0)
    v = vals.var(0, ddof=1)
    # per-pixel means: z-scores against the closed form
    z = (m - mean_f) / math.sqrt(var_f / batches)
    assert np.mean(np.abs(z) > 2.5758) < 0.1 and np.abs(z).max() < 4.5
    # per-pixel variances: the pooled variance within 3% of Var[f] (4000 x 36 samples)
    assert abs(v.mean() / var_f - 1.0) < 0.03
    # the image mean: stratification across pixels removes most of its variance
    img = vals.mean(1)
    assert img.var(ddof=1) < 0.5 * var_f / n
