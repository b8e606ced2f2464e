"""The relMSE / time-to-target arithmetic used to compare guided and plain rendering."""
import numpy as np

from phg import convergence as cv


def film_from_samples(x):
    """FP64 film (sum rgb, count, sum of squares) of per-pixel sample stacks x: (n, H, W, 3)."""
    n = x.shape[0]
    f = np.zeros(x.shape[1:3] + (8,))
    f[..., 0:3] = x.sum(0)
    f[..., 3] = n
    f[..., 4:7] = (x * x).sum(0)
    return f


def test_mean_and_variance_of_mean():
    rng = np.random.default_rng(0)
    x = rng.normal(2.0, 3.0, (64, 4, 5, 3))
    m, v = cv.mean_and_var_of_mean(film_from_samples(x))
    assert np.allclose(m, x.mean(0))
    assert np.allclose(v, x.var(0, ddof=1) / 64)


def test_reference_noise_correction_is_unbiased():
    # image and reference are independent noisy estimates of the same truth: subtracting the
    # reference's variance of the mean leaves E[relMSE] = the image's own relative variance
    rng = np.random.default_rng(1)
    truth = rng.uniform(0.0, 2.0, (32, 32, 3))
    sd, n_img, n_ref = 1.5, 16, 64
    est, raw = [], []
    for _ in range(200):
        ref = truth + rng.normal(0, sd, (n_ref, 32, 32, 3))
        R, vR = cv.mean_and_var_of_mean(film_from_samples(ref))
        img = truth + rng.normal(0, sd, (n_img, 32, 32, 3)).mean(0)
        est.append(cv.relmse(img, R, vR))
        raw.append(cv.relmse(img, R))
    expect = np.mean(sd * sd / n_img / (truth * truth + cv.EPS))
    # the denominator uses the noisy R, so compare through the same weights to first order
    assert abs(np.mean(est) - expect) < 0.1 * expect
    assert np.mean(raw) > np.mean(est) * 1.1


def test_time_to_interpolates_log_log():
    # relMSE = 1 / spp at 1 ms per sample: reaches 0.01 at 100 spp = 0.1 s
    curve = [(s, s * 1e-3, 1.0 / s) for s in (1, 2, 4, 8, 16, 32, 64, 128, 256)]
    assert abs(cv.time_to(curve, 0.01) - 0.1) < 1e-9
    assert cv.time_to(curve, 1e-4) is None
    assert cv.time_to(curve, 2.0) == curve[0][1]


class FakeBatchContext:
    """Stands in for phg.Context in progressive_batch: each job's film holds `spp` Gaussian
    samples per pixel around the job's truth, drawn from the job's seed."""

    def __init__(self, truths, sd):
        self.truths, self.sd, self.films, self.seeds = truths, sd, [], []

    def render_batch(self, jobs, **kw):
        self.films = []
        for j in jobs:
            self.seeds.append(j["seed"])
            rng = np.random.default_rng(j["seed"])
            t = self.truths[j["scene"]]
            self.films.append(film_from_samples(t + rng.normal(0, self.sd, (j["spp"],) + t.shape)))

    def batch_film(self, k):
        return self.films[k]


def test_progressive_batch_sums_independent_increments():
    rng = np.random.default_rng(2)
    truths = [rng.uniform(0.5, 2.0, (6, 9, 3)), rng.uniform(0.5, 2.0, (4, 5, 3))]
    jobs = [dict(scene=0, spp=1, seed=0), dict(scene=1, spp=1, seed=0)]
    ctx = FakeBatchContext(truths, 0.8)
    refs = [(t, np.zeros_like(t)) for t in truths]
    curve = cv.progressive_batch(ctx, jobs, 256, refs, seed_base=10)
    assert [c[0] for c in curve] == [1, 2, 4, 8, 16, 32, 64, 128, 256]
    assert len(set(ctx.seeds)) == len(ctx.seeds)           # every increment of every job: its own seed
    # against the exact truth relMSE falls as 1/spp: E[relMSE] = mean(sd^2 / spp / (t^2 + eps))
    w = np.concatenate([(0.8 ** 2 / (t * t + cv.EPS)).ravel() for t in truths]).mean()
    prod = np.array([e * s for s, _, e in curve])
    assert np.all(np.abs(prod / w - 1.0) < 0.35)
    slope = np.polyfit(np.log([c[0] for c in curve]), np.log([c[2] for c in curve]), 1)[0]
    assert -1.15 < slope < -0.85
