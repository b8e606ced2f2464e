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
    o = rng.normal(size=(8, 41))
    o[:, 3:40:5] = rng.choice([-7.5, -3.0, 0.0, 2.0, 5.5], size=(8, 8))
    om = rng.normal(size=(8, 3))
    om /= np.linalg.norm(om, axis=1, keepdims=True)
    w = np.ones(8)
    g = go.output_grad(o, om, w)
    assert np.all(g[:, 40] == 0.0)   # no selection probability in this loss
    for k in range(40):
        d = np.zeros_like(o)
        d[:, k] = 1e-6
        fd = (-np.log(go.mixture_pdf(o + d, om)) + np.log(go.mixture_pdf(o - d, om))) / 2e-6
        assert np.allclose(g[:, k], fd, rtol=1e-4, atol=1e-6), k


def test_mixture_loss_gradient_matches_finite_differences():
    # -w log(a q + (1 - a) b): the lobes as one component of the sampling mixture
    rng = np.random.default_rng(5)
    o = rng.normal(size=(8, 41))
    o[:, 3:40:5] = rng.choice([-7.5, -3.0, 0.0, 2.0, 5.5], size=(8, 8))
    om = rng.normal(size=(8, 3))
    om /= np.linalg.norm(om, axis=1, keepdims=True)
    w = rng.random(8) + 0.5
    b = np.array([0.0, 0.1, 0.5, 2.0, 10.0, 0.3, 0.0, 1.0])
    a = np.array([0.5, 0.5, 0.3, 0.7, 0.5, 1.0, 0.0, 0.9])
    g = go.output_grad(o, om, w, b, a)
    assert np.all(g[:, 40] == 0.0)   # a fixed selection probability is not learned

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


def test_learned_selection_gradient_matches_finite_differences():
    """With the network's own selection probability a(z) = 0.05 + 0.9 sigmoid(output 40): the
    lobe outputs' gradients against central differences of -w log(a q + (1 - a) b), and the
    selection logit's against central differences of the one-sample second moment
    w^2 p_s / (a q + (1 - a) b) (p_s: the record's sampling density); records not to be learned
    (p_s = 0) or sampled by the BSDF alone (a = 0) keep the fixed-probability gradient."""
    rng = np.random.default_rng(9)
    o = rng.normal(size=(8, 41))
    o[:, 3:40:5] = rng.choice([-7.5, -3.0, 0.0, 2.0, 5.5], size=(8, 8))
    o[:, 40] = [-4.0, -1.0, 0.0, 0.5, 2.0, 6.0, 1.0, -2.0]
    om = rng.normal(size=(8, 3))
    om /= np.linalg.norm(om, axis=1, keepdims=True)
    w = rng.random(8) + 0.5
    b = np.array([0.2, 0.1, 0.5, 2.0, 10.0, 0.3, 0.0, 1.0])
    sel = np.array([0.5, 0.5, 0.3, 0.7, 0.5, 0.9, 0.5, 0.0])
    ps = np.array([0.3, 1.2, 0.05, 2.5, 7.0, 0.4, 0.6, 0.9])
    g = go.output_grad(o, om, w, b, sel, ps)

    def f(oo):
        q = go.mixture_pdf(oo, om)
        a = np.where((ps > 0) & (sel > 0), go.selection(oo), sel)
        return -w * np.log(go.record_density(q, b, a))
    for k in range(40):
        d = np.zeros_like(o)
        d[:, k] = 1e-6
        fd = (f(o + d) - f(o - d)) / 2e-6
        assert np.allclose(g[:, k], fd, rtol=1e-4, atol=1e-7), k
    d = np.zeros_like(o)
    d[:, 40] = 1e-6
    fd = (go.selection_objective(o + d, om, w, b, sel, ps) - go.selection_objective(o - d, om, w, b, sel, ps)) / 2e-6
    assert np.allclose(g[:, 40], fd, rtol=1e-4, atol=1e-7)
    assert np.all(fd[:6] != 0.0)
    # b = 0 (row 6) and a = 0 (row 7): no selection gradient, plain log-likelihood lobes
    assert g[6, 40] == 0.0 and g[7, 40] == 0.0
    plain = go.output_grad(o, om, w)
    assert np.allclose(g[[6, 7], :40], plain[[6, 7], :40])
    # the selection probability's range and midpoint
    assert np.allclose(go.selection(np.array([[0.0] * 41])), 0.5)
    zs = np.zeros((3, 41)); zs[:, 40] = [-50.0, 50.0, 0.0]
    s_ = go.selection(zs)
    assert abs(s_[0] - 0.05) < 1e-12 and abs(s_[1] - 0.95) < 1e-12


def sphere_grid(nt, nphi):
    th = (np.arange(nt) + 0.5) / nt * np.pi
    ph = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (np.pi / nt) * (2 * np.pi / nphi)).reshape(-1)
    return d, dw


def test_selection_gradient_is_unbiased_for_the_second_moment():
    """Directions drawn from the one-sample MIS density p = a q + (1 - a) b with weight
    g / p: the expected selection-logit gradient equals d/dz of the estimator's second
    moment M(z) = integral of g^2 / p, here by quadrature over the sphere (expectation =
    sum of p dw x per-direction gradient) against central differences of M.  Also at the
    minimiser of M the expected gradient vanishes, and M's minimiser is where the second
    moment is smallest over a scan of z."""
    rng = np.random.default_rng(21)
    o = rng.normal(size=(1, 41))
    o[0, 3:40:5] = np.log([30.0, 2.0, 8.0, 0.5, 15.0, 1.0, 4.0, 60.0])
    o[0, 0:5] = [0.6, 0.0, 0.8, np.log(12.0), 3.0]           # a dominant lobe on the integrand's peak
    d, dw = sphere_grid(200, 400)
    n = len(d)
    b = np.maximum(d[:, 2], 0.0) / np.pi                     # cosine density of the BSDF
    # a glossy peak the lobe covers plus a diffuse part only the BSDF covers: M has an interior
    # minimiser
    g_int = np.where(d[:, 2] > 0, np.maximum(d @ np.array([0.6, 0.0, 0.8]), 0.0) ** 12, 0.0) + 0.3 * b
    sel = np.full(n, 0.5)

    def moment(z):
        oo = o.copy(); oo[0, 40] = z
        oo = np.repeat(oo, n, 0)
        a = go.selection(oo)
        p = a * go.mixture_pdf(oo, d) + (1 - a) * b
        return np.sum(np.where(p > 0, g_int ** 2 / np.where(p > 0, p, 1.0), 0.0) * dw)

    def expected_grad(z):
        oo = o.copy(); oo[0, 40] = z
        oo = np.repeat(oo, n, 0)
        a = go.selection(oo)
        p = a * go.mixture_pdf(oo, d) + (1 - a) * b
        w = np.where(p > 0, g_int / np.where(p > 0, p, 1.0), 0.0)
        gr = go.output_grad(oo, d, w, b, sel, p)[:, 40]
        return np.sum(p * dw * gr)

    for z in (-2.0, 0.0, 1.5):
        fd = (moment(z + 1e-4) - moment(z - 1e-4)) / 2e-4
        assert abs(expected_grad(z) - fd) <= 1e-5 * abs(fd) + 1e-9, z
    zs = np.linspace(-6, 6, 49)
    ms = np.array([moment(z) for z in zs])
    zbest = zs[np.argmin(ms)]
    assert -6 < zbest < 6          # an interior minimiser for this integrand
    assert abs(expected_grad(zbest)) < abs(expected_grad(zbest - 1.0))
    assert abs(expected_grad(zbest)) < abs(expected_grad(zbest + 1.0))


def test_selection_head_gradient_stops_at_the_last_hidden_layer():
    """Parameter gradients with learned selection: the hidden layers and the lobe columns of
    the output layer follow the log-likelihood with a held at its current value (central
    differences), the selection column follows the second-moment objective (central
    differences), and no part of the selection objective reaches the hidden layers."""
    params = go.init_params(3)
    rng = np.random.default_rng(5)
    e, om, wt = data(16, 4)
    b = rng.exponential(0.4, 16) + 0.05
    sel, ps = np.full(16, 0.5), rng.exponential(0.5, 16) + 0.05
    g = go.backward(params, e, om, wt, bsdf_pdf=b, sel=sel, samp_pdf=ps)
    _, o = go.forward(params, e)
    a_now = go.selection(o)

    def kl(P):
        return go.loss(P, e, om, wt, bsdf_pdf=b, sel=a_now)

    def second(P):
        _, oo = go.forward(P, e)
        return go.selection_objective(oo, om, wt, b, sel, ps).mean()

    def fd(fn, li, kind, idx):
        P1 = [(w.copy(), bb.copy()) for w, bb in params]
        P2 = [(w.copy(), bb.copy()) for w, bb in params]
        P1[li][kind][idx] += 1e-6
        P2[li][kind][idx] -= 1e-6
        return (fn(P1) - fn(P2)) / 2e-6

    worst = 0.0
    for li in range(go.N_LAYERS):
        W = params[li][0]
        for _ in range(5):
            i, j = rng.integers(0, W.shape[0]), rng.integers(0, W.shape[1] if li < go.N_LAYERS - 1 else 40)
            v = fd(kl, li, 0, (i, j))
            worst = max(worst, abs(v - g[li][0][i, j]) / (abs(v) + 1e-7))
    for i in (0, 7, 33, 63):
        v = fd(second, go.N_LAYERS - 1, 0, (i, 40))
        worst = max(worst, abs(v - g[-1][0][i, 40]) / (abs(v) + 1e-7))
    v = fd(second, go.N_LAYERS - 1, 1, 40)
    worst = max(worst, abs(v - g[-1][1][40]) / (abs(v) + 1e-7))
    assert worst < 1e-4
    # the hidden layers' gradients do not change when only the selection objective changes
    g2 = go.backward(params, e, om, wt * 3.0, bsdf_pdf=b, sel=sel, samp_pdf=ps)
    for li in range(go.N_LAYERS - 1):
        assert np.allclose(g2[li][0], 3.0 * g[li][0], rtol=1e-9, atol=1e-15)


def test_mixture_integrates_to_one():
    rng = np.random.default_rng(4)
    o = rng.normal(size=(1, 41))
    o[0, 3:40:5] = np.log([0.1, 1.0, 5.0, 20.0, 60.0, 0.5, 2.0, 8.0])
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
