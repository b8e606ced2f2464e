"""Path-guiding network on the GPU against its CPU oracle (oracle/guiding_oracle.py), and
the guided estimator against closed forms and plain BNEEPT.

Tolerances: the GPU feeds BF16 operands to the matrix cores and accumulates in FP32; the
oracle rounds the same operands to BF16 and accumulates in FP64, so outputs agree to a few
BF16 ulps (2% of the output scale) and gradients to 5% in norm (cosine > 0.995)."""
import numpy as np
import pytest

import phg
from oracle import guiding_oracle as go
from helpers import furnace_scene, image_mean_z, pixel_z

pytestmark = pytest.mark.gpu


def flat(params):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in params]).astype(np.float32)


def unflat(v):
    out, off = [], 0
    for fi, fo in go.LAYER_SHAPES:
        W = v[off:off + fi * fo].reshape(fi, fo).astype(np.float64)
        off += fi * fo
        b = v[off:off + fo].astype(np.float64)
        off += fo
        out.append((W, b))
    return out


@pytest.fixture(scope="module")
def ctx():
    c = phg.Context(0)
    c.upload(phg.Scene.generate(phg.GEN_CORNELL))
    yield c
    c.close()


def random_records(n, seed, mixture=True, learned=False):
    """Records of 16 floats: features, direction, weight, BSDF density, selection probability,
    sampling density where the selection probability is learned (0: fixed)."""
    rng = np.random.default_rng(seed)
    x = rng.random((n, 9))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.exponential(1.0, n)
    b = rng.exponential(0.5, n) * (rng.random(n) < 0.8) if mixture else np.zeros(n)
    a = rng.choice([0.3, 0.5, 0.7], n) if mixture else np.zeros(n)
    ps = rng.exponential(0.5, n) * (rng.random(n) < 0.7) if learned else np.zeros(n)
    return np.concatenate([x, d, w[:, None], b[:, None], a[:, None], ps[:, None]], axis=1).astype(np.float32)


def oracle_grads(params, rec):
    e = go.encode(rec[:, :9].astype(np.float64))
    return go.backward(params, e, rec[:, 9:12], rec[:, 12], bf16=True, bsdf_pdf=rec[:, 13], sel=rec[:, 14],
                       samp_pdf=rec[:, 15])


def test_forward_matches_oracle(ctx):
    params = go.init_params(3)
    ctx.guide_set_params(flat(params))
    x = np.random.default_rng(4).random((1000, 9)).astype(np.float32)
    gpu = ctx.guide_forward(x)
    _, ref = go.forward(unflat(flat(params)), go.encode(x.astype(np.float64)), bf16=True)
    scale = np.abs(ref).max()
    err = np.abs(gpu - ref).max()
    print(f"[guide] forward max_abs_err={err:.3e} scale={scale:.3e}", flush=True)
    assert err <= 0.02 * max(1.0, scale)


def test_decoded_lobes_match_oracle(ctx):
    """The decoded mixture the guided shade stage reads: unit means, kappa = exp(clamped log
    kappa), softmax weights and the selection probability.  Decoding the GPU's own raw outputs
    with the oracle must give the GPU's lobes to float precision (hardware rcp / rsqrt / exp,
    about 1 ulp each); against the oracle's forward they agree to the BF16 tolerance."""
    params = go.init_params(11)
    # larger output-layer weights spread kappa over its clamp range and the weights over
    # several lobes
    W4, b4 = params[-1]
    params[-1] = (W4 * 8.0, b4)
    ctx.guide_set_params(flat(params))
    x = np.random.default_rng(12).random((1337, 9)).astype(np.float32)   # not a multiple of 64
    raw = ctx.guide_forward(x)
    lob = ctx.guide_lobes(x).astype(np.float64)
    mu, kappa, w = go.decode(raw)
    L = lob[:, :40].reshape(-1, 8, 5)
    assert np.allclose(L[:, :, 0:3], mu, rtol=0, atol=2e-6)
    assert np.allclose(L[:, :, 3], kappa, rtol=2e-6, atol=0)
    assert np.allclose(L[:, :, 4], w, rtol=4e-6, atol=1e-7)
    assert np.allclose(lob[:, 40], go.selection(raw), rtol=1e-6, atol=0)
    assert np.allclose(np.linalg.norm(L[:, :, 0:3], axis=2), 1.0, atol=2e-6)
    assert np.allclose(L[:, :, 4].sum(axis=1), 1.0, atol=2e-6)
    # against the oracle's own forward pass (BF16 operands)
    _, ref = go.forward(unflat(flat(params)), go.encode(x.astype(np.float64)), bf16=True)
    rmu, rk, rw = go.decode(ref)
    spread = np.ptp(np.log(kappa))
    print(f"[guide] lobes: log-kappa spread {spread:.2f}, max weight {w.max():.3f}", flush=True)
    assert spread > 4.0 and w.max() > 0.5
    cos = np.sum(L[:, :, 0:3] * rmu, axis=2)
    assert np.quantile(cos, 0.01) > 0.99
    assert np.abs(np.log(L[:, :, 3]) - np.log(rk)).max() < 0.5
    assert np.abs(L[:, :, 4] - rw).max() < 0.1


def test_inference_over_many_iterations_per_wave(ctx):
    """2^20 + 777 queries: each wave of the inference kernel runs several iterations of its
    software pipeline (next query fetched, next features made between the layers and the
    stores).  Decoding the raw outputs gives the lobes everywhere, and sampled rows (first,
    last, across the iteration boundaries) match the oracle's forward pass."""
    params = go.init_params(13)
    ctx.guide_set_params(flat(params))
    n = (1 << 20) + 777
    x = np.random.default_rng(14).random((n, 9)).astype(np.float32)
    raw = ctx.guide_forward(x)
    lob = ctx.guide_lobes(x).astype(np.float64)
    mu, kappa, w = go.decode(raw)
    L = lob[:, :40].reshape(-1, 8, 5)
    assert np.allclose(L[:, :, 0:3], mu, rtol=0, atol=2e-6)
    assert np.allclose(L[:, :, 3], kappa, rtol=2e-6, atol=0)
    assert np.allclose(L[:, :, 4], w, rtol=4e-6, atol=1e-7)
    rng = np.random.default_rng(15)
    rows = np.unique(np.concatenate([np.arange(64), np.arange(n - 64, n), (1 << 18) + np.arange(-32, 32),
                                     rng.integers(0, n, 4000)]))
    _, ref = go.forward(unflat(flat(params)), go.encode(x[rows].astype(np.float64)), bf16=True)
    scale = np.abs(ref).max()
    err = np.abs(raw[rows] - ref).max()
    print(f"[guide] {n} queries: max_abs_err={err:.3e} on {rows.size} rows, scale={scale:.3e}", flush=True)
    assert err <= 0.02 * max(1.0, scale)


@pytest.mark.parametrize("learned", [False, True])
def test_gradients_match_oracle(ctx, learned):
    params = go.init_params(5)
    ctx.guide_set_params(flat(params))
    rec = random_records(4096, 6, learned=learned)
    gpu = ctx.guide_gradients(rec).astype(np.float64)
    grads = oracle_grads(unflat(flat(params)), rec)
    ref = np.concatenate([np.concatenate([gW.ravel(), gb.ravel()]) for gW, gb in grads])
    rel = np.linalg.norm(gpu - ref) / np.linalg.norm(ref)
    cos = gpu @ ref / (np.linalg.norm(gpu) * np.linalg.norm(ref))
    print(f"[guide] gradient rel_err={rel:.3e} cosine={cos:.6f}", flush=True)
    assert rel < 0.05 and cos > 0.995
    # per layer, too
    off = 0
    for li, (gW, gb) in enumerate(grads):
        n = gW.size + gb.size
        g_l, r_l = gpu[off:off + n], ref[off:off + n]
        off += n
        assert np.linalg.norm(g_l - r_l) <= 0.08 * np.linalg.norm(r_l) + 1e-9, li


def test_adam_steps_follow_oracle(ctx):
    params = go.init_params(7)
    ctx.guide_set_params(flat(params))
    rec = random_records(2048, 8, learned=True)
    ctx.guide_train(rec, lr=1e-3, steps=3)
    gpu = ctx.guide_get_params().astype(np.float64)
    p = unflat(flat(params))
    m = [[np.zeros_like(W), np.zeros_like(b)] for W, b in p]
    v = [[np.zeros_like(W), np.zeros_like(b)] for W, b in p]
    for t in range(1, 4):
        g = oracle_grads(p, rec)
        p = go.adam_step(p, g, m, v, t)
    ref = flat(p).astype(np.float64)
    moved_ref = ref - flat(params)
    moved_gpu = gpu - flat(params)
    # each Adam step moves a parameter by about lr; compare the displacement
    cos = moved_gpu @ moved_ref / (np.linalg.norm(moved_gpu) * np.linalg.norm(moved_ref))
    print(f"[guide] adam displacement cosine={cos:.4f}", flush=True)
    assert cos > 0.97


def test_training_fits_a_known_distribution(ctx):
    """Directions drawn from a vMF whose mean depends on the position feature: training
    must raise the mean log-likelihood of held-out samples well above the uniform 1/(4 pi)."""
    rng = np.random.default_rng(9)
    n = 1 << 16

    def draw(n):
        x = rng.random((n, 9))
        mu = np.stack([np.cos(2 * np.pi * x[:, 0]), np.sin(2 * np.pi * x[:, 0]), 0.3 * np.ones(n)], 1)
        mu /= np.linalg.norm(mu, axis=1, keepdims=True)
        k = 20.0
        u = rng.random(n)
        w = 1 + np.log(u + (1 - u) * np.exp(-2 * k)) / k
        phi = 2 * np.pi * rng.random(n)
        s = np.sqrt(np.maximum(0, 1 - w * w))
        t1 = np.cross(mu, np.array([0, 0, 1.0]))
        t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
        t2 = np.cross(mu, t1)
        d = mu * w[:, None] + t1 * (s * np.cos(phi))[:, None] + t2 * (s * np.sin(phi))[:, None]
        return np.concatenate([x, d, np.ones((n, 1)), np.zeros((n, 3))], 1).astype(np.float32)

    train = draw(n)
    test = draw(4096)
    ctx.guide_set_params(flat(go.init_params(10)))
    for _ in range(20):
        ctx.guide_train(train, lr=3e-3, steps=10)
    o = ctx.guide_forward(test[:, :9])
    ll = np.log(go.mixture_pdf(o.astype(np.float64), test[:, 9:12].astype(np.float64))).mean()
    print(f"[guide] held-out mean log-likelihood={ll:.3f} (uniform={np.log(1 / (4 * np.pi)):.3f})", flush=True)
    assert ll > np.log(1 / (4 * np.pi)) + 1.5


def slit_records(n, rng, learned=False):
    """Floor points of a room (y = 0, x < 2.9) lit only through a slit at x = 3, z in
    [1.95, 2.05]: cosine-sampled directions, record weight 5 through the slit and 0.01
    elsewhere (the ratio f cos L_i / pdf for a white Lambertian floor)."""
    P = np.stack([rng.uniform(0.1, 2.9, n), np.zeros(n), rng.uniform(0.1, 3.9, n)], 1)
    N = np.tile([0.0, 1.0, 0.0], (n, 1))
    V = rng.normal(size=(n, 3))
    V[:, 1] = np.abs(V[:, 1])
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    u1, u2 = rng.random(n), rng.random(n)
    r, ph = np.sqrt(u1), 2 * np.pi * u2
    D = np.stack([r * np.cos(ph), np.sqrt(1 - u1), r * np.sin(ph)], 1)
    ok = D[:, 0] > 1e-6
    t = np.where(ok, (3.0 - P[:, 0]) / np.where(ok, D[:, 0], 1.0), np.inf)
    through = ok & (np.abs(P[:, 2] + t * D[:, 2] - 2.0) < 0.05) & (t * D[:, 1] < 3.0)
    w = np.where(through, 5.0, 0.01)
    x = go.features(P, N, V, [0, 0, 0], [4, 3, 4])
    if learned:
        # records whose selection probability is learned: the BSDF (cosine) density, a nominal
        # selection 0.5 and the density the directions were drawn with (cosine)
        tail = np.stack([D[:, 1] / np.pi, np.full(n, 0.5), D[:, 1] / np.pi], 1)
    else:
        tail = np.zeros((n, 3))
    rec = np.concatenate([x, D, w[:, None], tail], 1).astype(np.float32)
    return rec, through


def test_training_gain_on_a_slit_lit_floor(ctx):
    """What the trained network gains over BSDF (cosine) sampling alone: after 512 Adam steps
    (lr 1e-3, 2^16-record batches) on slit-lit floor records, the lobes put about 20x the
    cosine density's mass through the slit and the one-sample MIS mixture (a = 0.5) estimates
    the floor's irradiance integral with less variance (1.5x measured; the lobes do not yet
    cover the whole slit, and directions they miss keep only half the cosine density).
    Second moments under the mixture are computed by importance reweighting cosine samples:
    E_mix[(g / p)^2] = E_cos[w^2 c / p] with g = w c."""
    rng = np.random.default_rng(11)
    ctx.guide_set_params(flat(go.init_params(12)))
    test, through = slit_records(200_000, rng)
    for _ in range(8):
        ctx.guide_train(slit_records(1 << 16, rng)[0], lr=1e-3, steps=64)
    o = ctx.guide_forward(test[:, :9]).astype(np.float64)
    q = go.mixture_pdf(o, test[:, 9:12].astype(np.float64))
    c = test[:, 10].astype(np.float64) / np.pi
    w = test[:, 12].astype(np.float64)
    mean = w.mean()
    var_bsdf = (w * w).mean() - mean ** 2
    a = 0.5
    var_mix = (w * w * c / (a * q + (1 - a) * c)).mean() - mean ** 2
    mass_q = np.mean(np.where(through, q / c, 0.0))
    print(f"[guide] slit mass cosine={through.mean():.4f} trained={mass_q:.4f} "
          f"variance cosine={var_bsdf:.4f} mixture={var_mix:.4f} gain={var_bsdf / var_mix:.2f}x", flush=True)
    assert mass_q > 5 * through.mean()
    assert var_bsdf / var_mix > 1.25


def test_learned_selection_on_a_slit_lit_floor(ctx):
    """Training on records of the one-sample MIS mixture with the learned selection
    probability (its head trained on the estimator's second moment): the lobes still gather
    mass through the slit, and the mixture at the learned probability estimates the floor's
    irradiance integral with less variance than cosine sampling alone and no more than the
    mixture at the fixed 0.5 (same reweighting as the fixed-probability test)."""
    rng = np.random.default_rng(13)
    ctx.guide_set_params(flat(go.init_params(12)))
    test, through = slit_records(200_000, rng, learned=True)
    for _ in range(8):
        ctx.guide_train(slit_records(1 << 16, rng, learned=True)[0], lr=1e-3, steps=64)
    o = ctx.guide_forward(test[:, :9]).astype(np.float64)
    q = go.mixture_pdf(o, test[:, 9:12].astype(np.float64))
    a = go.selection(o)
    c = test[:, 10].astype(np.float64) / np.pi
    w = test[:, 12].astype(np.float64)
    mean = w.mean()
    var_bsdf = (w * w).mean() - mean ** 2
    var_half = (w * w * c / (0.5 * q + 0.5 * c)).mean() - mean ** 2
    var_learned = (w * w * c / (a * q + (1 - a) * c)).mean() - mean ** 2
    mass_q = np.mean(np.where(through, q / c, 0.0))
    print(f"[guide] learned selection: mean a={a.mean():.3f} range=[{a.min():.3f}, {a.max():.3f}] slit mass "
          f"cosine={through.mean():.4f} trained={mass_q:.4f} variance cosine={var_bsdf:.4f} a=0.5 {var_half:.4f} "
          f"learned {var_learned:.4f} gain={var_bsdf / var_learned:.2f}x", flush=True)
    assert np.all((a >= go.SEL_MIN - 1e-6) & (a <= go.SEL_MAX + 1e-6))
    assert mass_q > 5 * through.mean()
    assert var_learned < var_bsdf
    assert var_learned <= 1.05 * var_half


@pytest.mark.parametrize("albedo,radiance,guide_prob", [(0.5, 0.5, 0.5), (0.5, 0.5, -1.0)])
def test_guided_furnace_stays_unbiased(albedo, radiance, guide_prob):
    """Guided BNEEPT with online training, at a fixed and at the learned selection
    probability: the furnace's closed form Le / (1 - rho) holds."""
    sc = furnace_scene(albedo=albedo, radiance=radiance)
    with phg.Context(0) as c:
        c.upload(sc)
        c.film_clear(32, 32)
        c.render(8, max_bounces=10000, seed=1, spp_per_pass=1, guided=True, train=True, guide_batch=4096,
                 guide_steps=4, guide_prob=guide_prob)
        c.film_clear(32, 32)
        c.render(64, max_bounces=10000, seed=2, spp_per_pass=1, guided=True, train=False, guide_prob=guide_prob)
        film = c.film()
        st = c.stats()
    expect = radiance / (1 - albedo)
    n = film[..., 3].sum()
    mean = film[..., 0].sum() / n
    se = np.sqrt(max(film[..., 4].sum() / n - mean * mean, 0) / n)
    print(f"[guide] furnace guided mean={mean:.6f} expect={expect:.6f} se={se:.2e} queries={st['guide_queries']}",
          flush=True)
    assert st["guide_queries"] > 0
    assert abs(mean - expect) < 4 * se + 1e-4


@pytest.mark.parametrize("kind,guide_depth,guide_prob", [("cornell", 0, 0.5), ("cornell", 2, 0.5), ("ggx", 0, 0.5),
                                                         ("cornell", 0, -1.0), ("ggx", 0, -1.0)])
def test_guided_matches_plain_in_expectation(kind, guide_depth, guide_prob):
    """Guided renders (lobes at every vertex, or only at the first two; a fixed or the
    learned per-query selection probability) agree with plain ones per pixel within 99%
    confidence intervals: the Lambertian Cornell box, and a scene of GGX conductors and
    dielectrics of roughness 0.05..0.6, where the lobes are sampled at the rougher conductors
    and the other surfaces sample their BSDF alone."""
    if kind == "cornell":
        sc = phg.Scene.generate(phg.GEN_CORNELL)
    else:
        sc = phg.Scene.generate(phg.GEN_PROCEDURAL, seed=6, target_tris=30_000, num_instances=12, num_lights=6,
                                subdiv=3)
        rough = sc.materials["roughness"][sc.materials["type"] == phg.MAT_CONDUCTOR]
        assert np.any(rough >= 0.25) and np.any(rough < 0.25)
    W = H = 32
    with phg.Context(0) as c:
        c.upload(sc)
        c.film_clear(W, H)
        c.render(16, max_bounces=10000, seed=3, spp_per_pass=1, guided=True, train=True, guide_batch=8192,
                 guide_prob=guide_prob)
        c.film_clear(W, H)
        c.render(256, max_bounces=10000, seed=4, spp_per_pass=4, guided=True, guide_depth=guide_depth,
                 guide_prob=guide_prob)
        guided = c.film()
        st = c.stats()
        c.film_clear(W, H)
        c.render(256, max_bounces=10000, seed=5, spp_per_pass=4)
        plain = c.film()
    z = pixel_z(guided, plain)
    frac = np.mean(np.abs(z) > 2.5758)
    mg = guided[..., :3].sum() / guided[..., 3].sum()
    mp = plain[..., :3].sum() / plain[..., 3].sum()
    print(f"[guide] {kind} prob={guide_prob} guided vs plain: outside_99CI={frac:.4f} mean_guided={mg:.5f} "
          f"mean_plain={mp:.5f}", flush=True)
    zm = image_mean_z(guided, plain)
    print(f"[guide] {kind} image-mean z per channel {np.round(zm, 2)}", flush=True)
    assert frac < 0.02
    assert np.all(np.abs(zm) < 4.0)
    if kind == "cornell":
        assert abs(mg - mp) < 0.01 * mp
    if guide_depth:
        # the network is queried at the first guide_depth bounces only
        assert 0 < st["guide_queries"] <= guide_depth * st["paths"]


def test_training_renders_keep_small_passes():
    """A render that trains the network updates it after every pass; without an explicit pass
    size its passes hold at most 16M paths (64 samples of a 1024^2 image = 4 passes of 16
    samples), whatever the context's tuned pass size."""
    sc = phg.Scene.generate(phg.GEN_CORNELL)
    with phg.Context(0) as c:
        c.set_tuning(trace_block=128, paths_per_pass=64 << 20)
        c.upload(sc)
        c.film_clear(1024, 1024)
        c.render(64, max_bounces=5, seed=3, guided=True, train=True, guide_steps=2, guide_batch=4096)
        assert c.stats()["guide_steps"] == 4 * 2
        # a plain render on the same context uses the tuned size: one pass of at most 6 bounce
        # launches (depth 5) instead of four
        c.film_clear(1024, 1024)
        c.render(64, max_bounces=5, seed=3)
        st = c.stats()
        assert st["rays_primary"] == 64 * 1024 * 1024
        assert st["trace_launches"] <= 6
