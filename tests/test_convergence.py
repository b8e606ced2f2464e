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
