"""The Halton sample source (SGHalton.cpp:22-206 with Halton/halton.h and
Math/Random/Quasi/radical_inverse.h) as the oracle restates it and the product computes it,
against what the mathematics fixes:

* the radical inverse against exact digit reversal (Fractions), bit-exact in base 2;
* the float digit count against an FP32 evaluation of its defining loop;
* Kensler's hashed permutation is a bijection of [0, l) for every key;
* every digit permutation (none, fixed, per-digit, Owen) keeps the Halton net property: over
  any b^k consecutive indices (or an arithmetic progression with a step prime to b, as the leap
  sequences use) the leading k digits of the radical inverse take every value once -- checked
  on the integer digits, so float rounding cannot blur it;
* the float values are those digits over b^count to FP32 accuracy;
* the good primes (the leap table) against their definition, brute force;
* the raster batch is a (2^a, 3^b) net over the image;
* the product's Halton code (compiled for the host) equals the oracle bit for bit;
* films: the splat counts follow the reference's inclusive pixel ranges, and a Halton render
  of the furnace keeps its closed form."""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as orc
from hostemu import emu
from helpers import furnace_scene

PRIMES = [2, 3, 5, 7, 11, 13, 17]
HALTON = 2
PERM = {"fixed": 0 << 8, "none": 1 << 8, "per-digit": 2 << 8, "owen": 3 << 8}
SEQ = {"original": 0, "leap": 1 << 12, "random-leap": 2 << 12, "random-start": 3 << 12}


def exact_radical_inverse(b, i):
    r, scale = Fraction(0), Fraction(1, b)
    while i > 0:
        r += (i % b) * scale
        i //= b
        scale /= b
    return r


def test_radical_inverse_closed_forms():
    for b in PRIMES:
        for i in list(range(300)) + [b ** 5 - 1, b ** 7 + 3, 123456789]:
            want = exact_radical_inverse(b, i)
            got = orc.radical_inverse(b, i)
            if b == 2 and i < (1 << 24):
                assert got == float(want), (b, i)
            else:
                # a float scale multiplied by 1/b (rounded) once per digit: about half an ulp
                # of error per digit
                nd = 0 if i == 0 else int(math.floor(math.log(i, b) + 1e-9)) + 1
                assert abs(got - float(want)) <= (nd + 2) * 2.0 ** -24, (b, i, got, float(want))
    assert orc.radical_inverse(3, 5) == pytest.approx(7 / 9, abs=1e-7)   # 5 = 12 base 3 -> 0.21
    assert orc.radical_inverse(5, 0) == 0.0


def test_float_digit_counts():
    """num_meaningful_digits<float>: digits n while 1 - (b - 1) b^-n < 1 with FP32 rounding."""
    f = np.float32
    for b in PRIMES:
        n, s, r = 0, f(1), f(1) / f(b)
        while f(1) - f(f(b - 1) * s) < f(1):
            n, s = n + 1, f(s * r)
        assert orc.halton_float_digits(b) == n
        # the last digit still moves a float below 1, the next would not
        assert (b - 1) * float(b) ** -(n - 1) > 2.0 ** -25


@pytest.mark.parametrize("l", [2, 3, 5, 7, 11, 13, 17, 64, 100, 1000, 4097])
def test_kensler_permutation_is_a_bijection(l):
    rng = np.random.default_rng(l)
    for key in rng.integers(0, 2 ** 32, 12):
        img = [orc.kensler_permute(i, l, int(key)) for i in range(l)]
        assert sorted(img) == list(range(l))
    # the keys reach many permutations (l = 2 maps every key to the identity; base 2 is never
    # permuted)
    if l > 2:
        perms = {tuple(orc.kensler_permute(i, l, k) for i in range(l)) for k in range(24)}
        assert len(perms) >= min(math.factorial(l), 4)


def leading_digits(dim, index, sampler, leap, seed, k):
    digits, count = orc_reversed(dim, index, sampler, leap, seed)
    b = PRIMES[dim]
    return (digits * b ** k) // b ** count


def orc_reversed(dim, index, sampler, leap, seed):
    import ctypes as C
    L = orc.lib()
    L.orc_halton_reversed.restype = C.c_uint64
    L.orc_halton_reversed.argtypes = [C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64,
                                      C.POINTER(C.c_uint32)]
    cnt = C.c_uint32(0)
    d = L.orc_halton_reversed(dim, index, sampler, leap, seed, C.byref(cnt))
    return int(d), int(cnt.value)


@pytest.mark.parametrize("perm", list(PERM))
@pytest.mark.parametrize("seq", ["original", "leap", "random-start"])
def test_digit_permutations_keep_the_net_property(perm, seq):
    sampler = HALTON | PERM[perm] | SEQ[seq]
    for seed in (0, 99):
        leap = orc.halton_leap(sampler, seed)
        for dim in range(2, 7):
            b = PRIMES[dim]
            for k in (1, 2):
                for start in (0, 3 * b ** k, 1000003):
                    cells = [leading_digits(dim, start + i, sampler, leap, seed, k) for i in range(b ** k)]
                    assert sorted(cells) == list(range(b ** k)), (perm, seq, dim, k, start)


@pytest.mark.parametrize("perm", list(PERM))
def test_float_values_follow_their_digits(perm):
    sampler = HALTON | PERM[perm]
    rng = np.random.default_rng(5)
    for dim in range(7):
        b = PRIMES[dim]
        for i in list(range(50)) + [int(x) for x in rng.integers(0, 2 ** 62, 50)]:
            digits, count = orc_reversed(dim, i, sampler, 1, 3)
            exact = digits / b ** count
            got = orc.halton_dim(dim, i, sampler, 1, 3)
            assert 0.0 <= got < 1.0
            assert abs(got - exact) <= (min(count, 25) + 2) * 2.0 ** -24, (perm, dim, i)


def test_owen_scrambling_is_keyed_by_the_dimension_only():
    owen, fixed = HALTON | PERM["owen"], HALTON | PERM["fixed"]
    a = [orc.halton_dim(d, i, owen, 1, 1) for d in range(2, 7) for i in range(40)]
    b = [orc.halton_dim(d, i, owen, 1, 2) for d in range(2, 7) for i in range(40)]
    assert a == b
    c = [orc.halton_dim(d, i, fixed, 1, 1) for d in range(2, 7) for i in range(40)]
    e = [orc.halton_dim(d, i, fixed, 1, 2) for d in range(2, 7) for i in range(40)]
    assert c != e
    # raster dimensions are never permuted
    assert [orc.halton_dim(0, i, owen, 1, 1) for i in range(8)] == [0.0, 0.5, 0.25, 0.75, 0.125, 0.625, 0.375, 0.875]


def test_good_primes_and_leaps():
    n = 200000    # the 512th good prime, 65099, needs primes up to 2 x its index (6502)
    sieve = np.ones(n, bool)
    sieve[:2] = False
    for k in range(2, int(n ** 0.5) + 1):
        if sieve[k]:
            sieve[k * k::k] = False
    p = np.nonzero(sieve)[0].astype(np.int64)
    good = []
    for k in range(1, len(p) // 2):
        if all(p[k] * p[k] > p[k - i] * p[k + i] for i in range(1, k + 1)):
            good.append(int(p[k]))
            if len(good) == 512:
                break
    assert good[:8] == [5, 11, 17, 29, 37, 41, 53, 59]
    assert [orc.good_prime(i) for i in range(512)] == good
    assert orc.halton_leap(HALTON | SEQ["leap"], 0) == good[256] == 19961
    leaps = {orc.halton_leap(HALTON | SEQ["random-leap"], s) for s in range(40)}
    assert leaps <= set(good[151:512]) and len(leaps) > 20
    assert orc.halton_leap(HALTON, 5) == 1 and orc.halton_leap(HALTON | SEQ["random-start"], 5) == 1


def test_raster_batch_is_a_2_3_net():
    """W x H = 32 x 27: batch s holds points s*864 .. s*864 + 863, one in every cell of the
    32 x 27 grid (dimensions 0 and 1 are never permuted); the raster positions lie in the
    window and equal v * W, v * H."""
    W, H = 32, 27
    for sampler in (HALTON, HALTON | PERM["owen"] | SEQ["random-start"]):
        for s in (0, 5):
            xy = orc.halton_raster(W, H, s, sampler, seed=4)
            if (sampler >> 12) & 3 == 0:
                cx = [leading_digits(0, s * W * H + p, sampler, 1, 4, 5) for p in range(W * H)]
                cy = [leading_digits(1, s * W * H + p, sampler, 1, 4, 3) for p in range(W * H)]
                assert len(set(zip(cx, cy))) == W * H
            assert np.all((xy[:, 0] >= 0) & (xy[:, 0] < W) & (xy[:, 1] >= 0) & (xy[:, 1] < H))
            vx = np.array([orc.halton_dim(0, s * W * H + p, sampler, 1, 4) for p in range(W * H)], np.float64)
            assert np.array_equal(xy[:, 0], vx * W)
    # a filtered film's window: half size 2 around the pixel centres
    xy = orc.halton_raster(W, H, 0, HALTON, half=2.0)
    assert xy[:, 0].min() >= -1.5 and xy[:, 0].max() < W + 1.5


@pytest.mark.parametrize("sampler", [HALTON | PERM[p] | SEQ[q] for p in PERM for q in SEQ])
def test_product_matches_oracle(sampler):
    rng = np.random.default_rng(sampler)
    for seed in (0, 7, 12345):
        leap = orc.halton_leap(sampler, seed)
        dims = rng.integers(0, 7, 1500).astype(np.uint32)
        idx = np.concatenate([np.arange(500), rng.integers(0, 2 ** 40, 500),
                              rng.integers(0, 2 ** 63, 500)]).astype(np.uint64)
        got = emu.halton_values(dims, idx, sampler, leap, seed)
        want = np.array([orc.halton_dim(int(d), int(i), sampler, leap, seed) for d, i in zip(dims, idx)], np.float32)
        assert np.array_equal(got, want)
    for filt, half in ((0, 0.5), (3, 2.0)):
        a = emu.halton_raster(24, 10, 3, filt, sampler, orc.halton_leap(sampler, 9), 9)
        b = orc.halton_raster(24, 10, 3, sampler, seed=9, half=half)
        assert np.array_equal(a, b)


def splat_counts(xy, W, H, half=0.5):
    """Per-pixel sample counts of a box-filtered film from the raster positions: pixel centres
    within [p - half, p + half], both ends included, inside the image."""
    cnt = np.zeros((H, W))
    for x, y in xy:
        x0 = math.ceil(max(x - half, 0.0) - 0.5)
        x1 = math.floor(min(x + half, W) - 0.5) + 1
        y0 = math.ceil(max(y - half, 0.0) - 0.5)
        y1 = math.floor(min(y + half, H) - 0.5) + 1
        cnt[y0:y1, x0:x1] += 1
    return cnt


def test_halton_film_counts_follow_the_splat_ranges():
    """A Halton render's sample counts are the inclusive box-splat counts of its raster
    positions: samples on a pixel edge (every x of the first points of a power-of-two-wide
    image) count in both pixels."""
    sc = furnace_scene(albedo=0.5, radiance=0.5)
    W, H, spp = 16, 12, 3
    film, _ = orc.render(sc, W, H, spp, max_bounces=2, seed=3, sampler=HALTON)
    want = sum(splat_counts(orc.halton_raster(W, H, s, HALTON, seed=3), W, H) for s in range(spp))
    assert np.array_equal(film[..., 3], want)
    assert want.sum() > W * H * spp      # edge samples are counted twice


@pytest.mark.parametrize("perm", ["fixed", "owen", "none"])
def test_halton_furnace_closed_form(perm):
    """White furnace (albedo 0.5, radiance 0.5): every pixel's expectation is 1 whatever the
    sample placement, so the Halton render's mean stays at the closed form."""
    sc = furnace_scene(albedo=0.5, radiance=0.5)
    W = H = 16
    film, _ = orc.render(sc, W, H, 64, max_bounces=10000, seed=1, sampler=HALTON | PERM[perm])
    n = film[..., 3].sum()
    mean = film[..., 0].sum() / n
    var = film[..., 4].sum() / n - mean * mean
    se = math.sqrt(max(var, 0.0) / n) * 2.0     # splats of one sample into two pixels correlate
    assert abs(mean - 1.0) < 4 * se + 1e-3
