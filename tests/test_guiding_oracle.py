"""Pins the guiding oracle: analytic gradients against central finite differences, the
mixture density against its normalisation, Adam against its closed-form first step."""
import numpy as np

from oracle import guiding_oracle as go


def data(n, seed):
    rng = np.random.default_rng(seed)
    e = go.encode(rng.random((n, 9)))
    om = rng.normal(size=(n, 3))
    om /= np.linalg.norm(om, axis=1, keepdims=True)
    return e, om, rng.random(n) + 0.1


def test_gradients_match_finite_differences():
    params = go.init_params(1)
    e, om, wt = data(16, 0)
    g = go.backward(params, e, om, wt)
    rng = np.random.default_rng(2)
    worst = 0.0
    for li, (W, b) in enumerate(params):
        for _ in range(6):
            i, j = rng.integers(0, W.shape[0]), rng.integers(0, W.shape[1])
            for sgn_h in (1e-6,):
                P1 = [(w.copy(), bb.copy()) for w, bb in params]
                P2 = [(w.copy(), bb.copy()) for w, bb in params]
                P1[li][0][i, j] += sgn_h
                P2[li][0][i, j] -= sgn_h
                fd = (go.loss(P1, e, om, wt) - go.loss(P2, e, om, wt)) / (2 * sgn_h)
                worst = max(worst, abs(fd - g[li][0][i, j]) / (abs(fd) + 1e-7))
        j = rng.integers(0, b.size)
        P1 = [(w.copy(), bb.copy()) for w, bb in params]
        P2 = [(w.copy(), bb.copy()) for w, bb in params]
        P1[li][1][j] += 1e-6
        P2[li][1][j] -= 1e-6
        fd = (go.loss(P1, e, om, wt) - go.loss(P2, e, om, wt)) / 2e-6
        worst = max(worst, abs(fd - g[li][1][j]) / (abs(fd) + 1e-7))
    assert worst < 1e-4


def test_output_gradient_small_and_large_kappa():
    rng = np.random.default_rng(3)
    o = rng.normal(size=(8, 40))
    o[:, 3::5] = rng.choice([-7.5, -3.0, 0.0, 2.0, 5.5], size=(8, 8))
    om = rng.normal(size=(8, 3))
    om /= np.linalg.norm(om, axis=1, keepdims=True)
    w = np.ones(8)
    g = go.output_grad(o, om, w)
    for k in range(40):
        d = np.zeros_like(o)
        d[:, k] = 1e-6
        fd = (-np.log(go.mixture_pdf(o + d, om)) + np.log(go.mixture_pdf(o - d, om))) / 2e-6
        assert np.allclose(g[:, k], fd, rtol=1e-4, atol=1e-6), k


def test_mixture_loss_gradient_matches_finite_differences():
    # -w log(a q + (1 - a) b): the lobes as one component of the sampling mixture
    rng = np.random.default_rng(5)
    o = rng.normal(size=(8, 40))
    o[:, 3::5] = rng.choice([-7.5, -3.0, 0.0, 2.0, 5.5], size=(8, 8))
    om = rng.normal(size=(8, 3))
    om /= np.linalg.norm(om, axis=1, keepdims=True)
    w = rng.random(8) + 0.5
    b = np.array([0.0, 0.1, 0.5, 2.0, 10.0, 0.3, 0.0, 1.0])
    a = np.array([0.5, 0.5, 0.3, 0.7, 0.5, 1.0, 0.0, 0.9])
    g = go.output_grad(o, om, w, b, a)

    def f(oo):
        return -w * np.log(go.record_density(go.mixture_pdf(oo, om), b, a))
    for k in range(40):
        d = np.zeros_like(o)
        d[:, k] = 1e-6
        fd = (f(o + d) - f(o - d)) / 2e-6
        assert np.allclose(g[:, k], fd, rtol=1e-4, atol=1e-6), k
    # b = 0 or a = 0 leaves the plain weighted log-likelihood gradient
    plain = go.output_grad(o, om, w)
    assert np.allclose(g[[0, 6]], plain[[0, 6]])
    # a = 1 also reduces to it (the BSDF component has no weight)
    assert np.allclose(g[5], plain[5])


def test_mixture_integrates_to_one():
    rng = np.random.default_rng(4)
    o = rng.normal(size=(1, 40))
    o[0, 3::5] = np.log([0.1, 1.0, 5.0, 20.0, 60.0, 0.5, 2.0, 8.0])
    # quadrature on a fine latitude-longitude grid
    nt, nphi = 2000, 1000
    th = (np.arange(nt) + 0.5) / nt * np.pi
    ph = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (np.pi / nt) * (2 * np.pi / nphi)).reshape(-1)
    q = go.mixture_pdf(np.repeat(o, len(d), 0), d)
    assert abs((q * dw).sum() - 1.0) < 1e-3


def test_adam_first_step_moves_by_learning_rate():
    params = go.init_params(5)
    e, om, wt = data(32, 6)
    g = go.backward(params, e, om, wt)
    m = [[np.zeros_like(W), np.zeros_like(b)] for W, b in params]
    v = [[np.zeros_like(W), np.zeros_like(b)] for W, b in params]
    new = go.adam_step(params, g, m, v, 1, lr=1e-3)
    for (W0, b0), (W1, b1), (gW, gb) in zip(params, new, g):
        # first Adam step: m_hat = g, v_hat = g^2 -> step = lr * g / (|g| + eps)
        assert np.allclose(W0 - W1, 1e-3 * gW / (np.abs(gW) + 1e-8), rtol=1e-9, atol=1e-15)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.005859375, -2.5, 1.01171875, 65504.0], np.float32)
    r = go.bf16_round(x)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-9 rounds up; 1 + 3*2^-8 is a tie -> even (1 + 2^-6)
    assert list(r[:5]) == [1.0, 1.0, 1.0078125, -2.5, 1.015625]
    assert r[5] == 65536.0
