"""The reference's stratified sample source (oracle/sample_source.py) against what the
mathematics fixes, and the property that makes it interchangeable with the renderer's
independent per-pixel samples for every per-pixel statistic: within a batch each pixel gets
one sample whose raster offset and first five dimensions are uniformly distributed, and
batches are independent.  Stratification only correlates the pixels of one batch, which the
image mean shows and no per-pixel mean or variance does.

The renderer's stratified source (the oracle's strat_* functions, which the GPU reproduces bit
for bit in tests/test_gpu_render.py) is pinned against the same properties: its keyed shuffle
is a bijection of the pixels, each batch fills the SGStratified strata exactly as the
restatement does, the jitter survives inside each stratum, and per-pixel statistics equal
independent sampling while the image mean's variance drops."""
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


# ---- the renderer's stratified source (oracle strat_*)
from oracle import oracle as orc


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 16, 17, 63, 64, 65, 255, 1000, 4096, 4097])
def test_strat_shuffle_is_a_bijection(n):
    for key in (0, 1, 0x9E3779B97F4A7C15, 2**64 - 1):
        v = [orc.strat_perm(i, n, key) for i in range(n)]
        assert sorted(v) == list(range(n))


def test_strat_shuffle_is_uniform_over_keys():
    """Over keys, the image of one index is uniform and of two indices uniform over ordered
    pairs (chi-square at the 1e-4 level): a stand-in for a uniformly random shuffle."""
    from scipy import stats
    n, keys = 6, 30000
    first = np.zeros(n)
    pairs = np.zeros((n, n))
    for key in range(keys):
        k = orc.strat_key(key, 0, 0)
        a, b = orc.strat_perm(0, n, k), orc.strat_perm(1, n, k)
        first[a] += 1
        pairs[a, b] += 1
    assert stats.chisquare(first).pvalue > 1e-4
    off = pairs[~np.eye(n, dtype=bool)]
    assert np.all(np.diag(pairs) == 0)
    assert stats.chisquare(off).pvalue > 1e-4


def _strat_value(c, k, u):
    v = ((c.astype(np.float64) + u.astype(np.float64)) / k).astype(np.float32)
    return np.where(v < 1.0, v, np.nextafter(np.float32(1.0), np.float32(0.0)))


@pytest.mark.parametrize("n", [36, 1000, 37 * 29, 4096, 5])
def test_strat_batch_fills_the_sgstratified_strata(n):
    """One batch: the shuffled stratum of every pixel (per stage) fills SGStratified's strata --
    k x k cells in whole passes, the rest uniform, N cells in 1-D -- and each value is its
    stratum's cell plus the pixel's own jitter, (cell + u) / k rounded to float."""
    rng = np.random.default_rng(n)
    jit = (rng.integers(0, 1 << 23, (n, 5)) * 2 + 1).astype(np.float32) * np.float32(2.0 ** -24)
    seed, sample = 12345, 7
    out = orc.strat_batch(n, seed, sample, jit)
    assert out.min() > 0.0 and out.max() < 1.0
    k = math.isqrt(n)
    assert ss.revise_dim_size_hints(n, [1, 1]) == [k, k]
    passes = n // (k * k)
    whole = passes * k * k
    for stage, (a, b) in enumerate(((0, 1), (2, 3))):
        key = orc.strat_key(seed, sample, stage)
        r = np.array([orc.strat_perm(p, n, key) for p in range(n)])
        assert np.array_equal(np.sort(r), np.arange(n))
        cell = r % (k * k)
        in_pass = r < whole
        counts = np.bincount(cell[in_pass], minlength=k * k)
        assert np.all(counts == passes)               # every stratum of every whole pass once
        want_a = np.where(in_pass, _strat_value(cell % k, k, jit[:, a]), jit[:, a])
        want_b = np.where(in_pass, _strat_value(cell // k, k, jit[:, b]), jit[:, b])
        assert np.array_equal(out[:, a], want_a) and np.array_equal(out[:, b], want_b)
    r = np.array([orc.strat_perm(p, n, orc.strat_key(seed, sample, 2)) for p in range(n)])
    assert np.array_equal(np.sort(r), np.arange(n))   # 1-D: one sample per stratum
    assert np.array_equal(out[:, 4], _strat_value(r, n, jit[:, 4]))


def test_strat_per_pixel_statistics_match_independent_sampling():
    """As test_per_pixel_statistics_match_independent_sampling, for the renderer's source: per
    pixel over batches the five dimensions integrate like independent uniforms, the image mean's
    variance falls well below Var[f] / N."""
    rng = np.random.default_rng(5)
    n, batches = 36, 4000

    def f(d):  # E = 0.5 + 0.5 + 0.25 + 0.3 = 1.55
        return d[:, 0] + d[:, 1] + d[:, 2] * d[:, 4] + (d[:, 3] < 0.3)

    mean_f = 1.55
    var_f = 1 / 12 + 1 / 12 + (1 / 9 - 1 / 16) + 0.21
    vals = np.empty((batches, n))
    for b in range(batches):
        jit = rng.random((n, 5)).astype(np.float32)
        vals[b] = f(orc.strat_batch(n, 99, b, jit).astype(np.float64))
    m = vals.mean(0)
    z = (m - mean_f) / math.sqrt(var_f / batches)
    assert np.mean(np.abs(z) > 2.5758) < 0.1 and np.abs(z).max() < 4.5
    assert abs(vals.var(0, ddof=1).mean() / var_f - 1.0) < 0.03
    assert vals.mean(1).var(ddof=1) < 0.5 * var_f / n
