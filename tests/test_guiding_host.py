"""The guiding model's per-lane math (photon-v2_amd/csrc/guide_device.h, compiled for the
host) against the guiding oracle: encoding, mixture pdf, loss gradient, lobe sampling."""
import numpy as np

from oracle import guiding_oracle as go
from hostemu import emu


def test_encoding_matches_oracle():
    x = np.random.default_rng(1).random((500, 9))
    assert np.allclose(emu.guide_encode(x), go.encode(x.astype(np.float32).astype(np.float64)), atol=2e-6)


def test_mixture_pdf_and_gradient_match_oracle():
    rng = np.random.default_rng(2)
    o = rng.normal(0, 1.5, (2000, 41))
    o[:, 3:40:5] = rng.uniform(-9, 8, (2000, 8))     # log kappa, including clamped values
    w = rng.normal(size=(2000, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    weight = rng.exponential(1.0, 2000)
    o32, w32 = o.astype(np.float32).astype(np.float64), w.astype(np.float32).astype(np.float64)
    p_ref = go.mixture_pdf(o32, w32)
    assert np.allclose(emu.guide_pdf(o, w), p_ref, rtol=2e-4, atol=1e-7)
    g_ref = go.output_grad(o32, w32, weight.astype(np.float32))
    g = emu.guide_output_grad(o, w, weight)
    scale = np.abs(g_ref).max(axis=1, keepdims=True) + 1e-12
    assert np.abs(g - g_ref).max() / 1.0 < 1e-3 * scale.max()
    assert np.all(np.abs(g - g_ref) <= 2e-3 * scale + 1e-6)
    # as a component of the one-sample MIS mixture with the BSDF density (some b = 0)
    b = rng.exponential(0.5, 2000) * (rng.random(2000) < 0.8)
    a = rng.choice([0.0, 0.3, 0.5, 1.0], 2000)
    b32, a32 = b.astype(np.float32).astype(np.float64), a.astype(np.float32).astype(np.float64)
    g_ref = go.output_grad(o32, w32, weight.astype(np.float32), b32, a32)
    g = emu.guide_output_grad(o, w, weight, b, a)
    scale = np.abs(g_ref).max(axis=1, keepdims=True) + 1e-12
    assert np.all(np.abs(g - g_ref) <= 2e-3 * scale + 1e-6)
    # with the network's own selection probability (a sampling density on most records)
    ps = rng.exponential(0.5, 2000) * (rng.random(2000) < 0.7)
    ps32 = ps.astype(np.float32).astype(np.float64)
    g_ref = go.output_grad(o32, w32, weight.astype(np.float32), b32, a32, ps32)
    g = emu.guide_output_grad(o, w, weight, b, a, ps)
    scale = np.abs(g_ref).max(axis=1, keepdims=True) + 1e-12
    assert np.all(np.abs(g - g_ref) <= 2e-3 * scale + 1e-6)
    assert np.allclose(emu.guide_sel(o), go.selection(o32), rtol=1e-6, atol=1e-7)


def test_lobe_sampling_follows_mixture_pdf():
    rng = np.random.default_rng(3)
    o = rng.normal(0, 1.0, 41)
    o[3:40:5] = np.log([0.5, 3.0, 10.0, 40.0, 0.01, 5.0, 100.0, 1.0])
    o[40] = 0.0
    u = rng.random((200_000, 3))
    d = emu.guide_sample(o, u)
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-5)
    # histogram over a coarse latitude-longitude grid vs the integrated pdf
    th = np.arccos(np.clip(d[:, 2], -1, 1))
    ph = np.arctan2(d[:, 1], d[:, 0]) % (2 * np.pi)
    nb_t, nb_p = 12, 24
    hist = np.histogram2d(th, ph, bins=[nb_t, nb_p], range=[[0, np.pi], [0, 2 * np.pi]])[0]
    tt = (np.arange(nb_t * 8) + 0.5) / (nb_t * 8) * np.pi
    pp = (np.arange(nb_p * 8) + 0.5) / (nb_p * 8) * 2 * np.pi
    T, P = np.meshgrid(tt, pp, indexing="ij")
    dirs = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (np.pi / (nb_t * 8)) * (2 * np.pi / (nb_p * 8))).reshape(-1)
    pdf = go.mixture_pdf(np.repeat(o[None].astype(np.float32).astype(np.float64), len(dirs), 0), dirs)
    exp = (pdf * dw).reshape(nb_t, 8, nb_p, 8).sum(axis=(1, 3)) * len(d)
    m = exp > 30
    chi2 = ((hist[m] - exp[m]) ** 2 / exp[m]).sum()
    assert chi2 < 1.5 * m.sum() + 30, (chi2, m.sum())
