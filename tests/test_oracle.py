"""Pins the CPU oracle before it is trusted as the checker.

Each check is chosen so that a plausible slip in the restatement (a dropped term, a wrong
sign, a swapped operand) fails: known answers from the reference's own unit tests
(restated), published known answers, closed forms, integrals and invariants."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as orc
from helpers import FLT_MAX, furnace_scene, random_rays, random_soup, rays_towards, scene_from_tris


def fp(a):
    return np.ascontiguousarray(np.asarray(a, np.float32))


# ---------------------------------------------------------------- PCG32 (Pcg32.h:72-124)
def pcg(seq=0, state=0, default=False, jump=0, n=8):
    out = np.zeros(n, np.uint32)
    orc.lib().orc_pcg32_generate(seq, state, 1 if default else 0, jump, n, out.ctypes.data)
    return out


def test_pcg32_published_known_answer():
    # pcg-c-basic demo: pcg32_srandom_r(&rng, 42u, 54u) -> first six outputs
    got = pcg(seq=54, state=42, n=6)
    assert [hex(x) for x in got] == ["0xa15c02b7", "0x7b47f409", "0xba1d3330", "0x83d2f293", "0xbfa4784b",
                                     "0xcbed606e"]


def test_pcg32_jump_ahead_consistency():
    # Pcg32_test.cpp:39-83: jumping d then drawing equals the d-th draw
    ref = pcg(default=True, n=1024)
    for d in (0, 1, 2, 3, 17, 100, 511, 1023):
        assert pcg(default=True, jump=d, n=1)[0] == ref[d]
    assert pcg(default=True, jump=12 + 34 + 56, n=1)[0] == ref[12 + 34 + 56]


def test_stream_samples_in_open_unit_interval_and_uniform():
    out = np.zeros(200_000, np.float32)
    orc.lib().orc_stream_samples(12345, len(out), out.ctypes.data)
    assert out.min() > 0.0 and out.max() < 1.0
    hist = np.histogram(out, bins=20, range=(0, 1))[0]
    exp = len(out) / 20
    chi2 = ((hist - exp) ** 2 / exp).sum()
    assert chi2 < 45  # 19 dof, p ~ 1e-3


def test_moremur_is_bijective_on_a_sample():
    L = orc.lib()
    xs = np.arange(1, 20001, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    ys = {L.orc_moremur(int(x)) for x in xs}
    assert len(ys) == len(xs)


# ---------------------------------------------- watertight triangle (TWatertightTriangle.ipp)
def tri_hit(tri9, ray8):
    t = np.zeros(1, np.float32)
    b = np.zeros(3, np.float32)
    h = orc.lib().orc_intersect_triangle(fp(tri9).ctypes.data, fp(ray8).ctypes.data, t.ctypes.data, b.ctypes.data)
    return bool(h), float(t[0]), b


def test_triangle_closed_form():
    # axis-aligned ray through (0.25, 0.25) of the right triangle at z = 5
    hit, t, b = tri_hit([0, 0, 5, 1, 0, 5, 0, 1, 5], [0.25, 0.25, 0, 0, 0, 0, 1, FLT_MAX])
    assert hit and t == 5.0
    assert list(b) == [0.5, 0.25, 0.25]
    # same along -x (permutation of the dominant axis) and -y
    hit, t, b = tri_hit([-3, 0, 0, -3, 1, 0, -3, 0, 1], [0, 0.5, 0.25, 0, -1, 0, 0, FLT_MAX])
    assert hit and t == 3.0 and list(b) == [0.25, 0.5, 0.25]
    hit, t, b = tri_hit([0, 2, 0, 0, 2, 1, 1, 2, 0], [0.5, 0, 0.25, 0, 0, 1, 0, FLT_MAX])
    assert hit and t == 2.0 and list(b) == [0.25, 0.25, 0.5]


def test_triangle_segment_range_is_inclusive():
    tri = [0, 0, 5, 1, 0, 5, 0, 1, 5]
    assert tri_hit(tri, [0.25, 0.25, 0, 0, 0, 0, 1, 5.0])[0]          # maxT == t
    assert not tri_hit(tri, [0.25, 0.25, 0, 0, 0, 0, 1, 4.999])[0]
    assert tri_hit(tri, [0.25, 0.25, 0, 5.0, 0, 0, 1, 9.0])[0]        # minT == t
    assert not tri_hit(tri, [0.25, 0.25, 0, 5.001, 0, 0, 1, 9.0])[0]
    assert not tri_hit(tri, [0.25, 0.25, 0, 0, 0, 0, -1, FLT_MAX])[0]  # behind


def test_triangle_nonfinite_vertices_never_hit():
    # TWatertightTriangle_test.cpp:62-80
    inf = np.inf
    assert not tri_hit([0, 0, 0, 1, 1, 0, 1, -1, inf], [0.1, 0, -1, 0, 0.1, 0, 1, FLT_MAX])[0]
    assert not tri_hit([0, 0, 0, 1, 0, np.nan, 0, 1, 0], [0.2, 0.2, -1, 0, 0, 0, 1, FLT_MAX])[0]
    assert not tri_hit([0, 0, 0, 1, 1, 0, 2, 2, 0], [0.5, 0.5, -1, 0, 0, 0, 1, FLT_MAX])[0]  # zero area


def test_triangle_t_against_fp64_plane_intersection():
    rng = np.random.default_rng(1)
    v = random_soup(1, 0)[0].astype(np.float64).reshape(3, 3)
    worst = 0.0
    for i in range(2000):
        w = rng.dirichlet((1, 1, 1))
        if w.min() < 1e-3:
            continue
        p = w @ v
        o = p + rng.normal(0, 3, 3)
        d = (p - o) / np.linalg.norm(p - o)
        hit, t, b = tri_hit(v.reshape(9), np.r_[o, 0, d, FLT_MAX].astype(np.float32))
        assert hit
        of, df = o.astype(np.float32).astype(np.float64), d.astype(np.float32).astype(np.float64)
        n = np.cross(v[1] - v[0], v[2] - v[0])
        t64 = np.dot(v[0] - of, n) / np.dot(df, n)
        worst = max(worst, abs(t - t64) / t64)
        bb = np.array(b, np.float64)
        assert abs(bb.sum() - 1) < 1e-5
    assert worst < 2e-5


def test_shared_edge_watertight_in_oracle():
    v = np.array([[0, -1, 0, 0, 1, 0, -1, 0, 0], [0, -1, 0, 1, 0, 0, 0, 1, 0]], np.float32)
    ys = np.linspace(-0.9, 0.9, 501, dtype=np.float32)
    rays = np.zeros((len(ys), 8), np.float32)
    rays[:, 1] = ys
    rays[:, 2] = -1
    rays[:, 6] = 1
    rays[:, 7] = FLT_MAX
    tri, t, _, flags = orc.intersect_brute(v, rays, nthreads=1)
    assert np.all(tri >= 0)


# -------------------------------------- reference BVH semantics vs brute force (conservative)
@pytest.mark.parametrize("n,seed", [(1, 1), (50, 2), (3000, 3)])
def test_reference_bvh_finds_every_closest_hit(n, seed):
    v = random_soup(n, seed)
    rays = np.concatenate([rays_towards(v, 20_000, seed), random_rays(v, 20_000, seed + 1)])
    tri_b, t_b, _, flags = orc.intersect_brute(v, rays)
    tri_r, t_r = orc.intersect_refbvh(v, rays)
    ok = (flags & 1) == 0
    assert np.array_equal(tri_b[ok], tri_r[ok])
    h = ok & (tri_b >= 0)
    assert np.array_equal(t_b[h].view(np.uint32), t_r[h].view(np.uint32))


# ---------------------------------------------------------- GGX (IsoTrowbridgeReitz.cpp)
def hemisphere_grid(nt=400, nphi=800):
    th = (np.arange(nt) + 0.5) / nt * (np.pi / 2)
    ph = (np.arange(nphi) + 0.5) / nphi * (2 * np.pi)
    T, P = np.meshgrid(th, ph, indexing="ij")
    dw = np.sin(T) * (np.pi / 2 / nt) * (2 * np.pi / nphi)
    # y-up local frame
    d = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1)
    return d.reshape(-1, 3), dw.reshape(-1), T.reshape(-1)


@pytest.mark.parametrize("alpha", [0.05, 0.2, 0.6])
def test_ggx_D_projected_area_integrates_to_one(alpha):
    H, dw, th = hemisphere_grid(1200 if alpha < 0.1 else 400, 400)
    N = [0, 1, 0]
    D, vndf, lam = orc.ggx_batch(N, N, alpha, H)
    assert abs((D * np.cos(th) * dw).sum() - 1.0) < 2e-3


@pytest.mark.parametrize("alpha,vtheta", [(0.1, 0.3), (0.4, 1.0), (0.7, 1.4)])
def test_ggx_vndf_integrates_to_one_and_lambda_closed_form(alpha, vtheta):
    V = [np.sin(vtheta), np.cos(vtheta), 0.0]
    H, dw, th = hemisphere_grid(1000, 400)
    D, vndf, lam = orc.ggx_batch([0, 1, 0], V, alpha, H)
    visible = (H @ np.asarray(V)) > 0  # the density of visible normals lives where H.V > 0
    assert abs((vndf * dw)[visible].sum() - 1.0) < 3e-3
    # Smith G1 of the GGX: 2 / (1 + sqrt(1 + alpha^2 tan^2 theta))
    _, _, lamv = orc.ggx_batch([0, 1, 0], V, alpha, np.array([V], np.float32))
    g1 = 1 / (1 + lamv[0])
    assert abs(g1 - 2 / (1 + np.sqrt(1 + alpha ** 2 * np.tan(vtheta) ** 2))) < 1e-5


def test_ggx_visible_normal_samples_follow_vndf():
    alpha, vtheta = 0.35, 0.8
    V = np.array([np.sin(vtheta), np.cos(vtheta), 0.0], np.float32)
    rng = np.random.default_rng(3)
    n = 100_000
    out = np.zeros(3, np.float32)
    L = orc.lib()
    Hs = np.zeros((n, 3), np.float32)
    u = rng.random((n, 2)).astype(np.float32)
    N = fp([0, 1, 0])
    for i in range(n):
        L.orc_ggx_sample_vh(N.ctypes.data, V.ctypes.data, C.c_float(alpha), C.c_float(u[i, 0]), C.c_float(u[i, 1]),
                            out.ctypes.data)
        Hs[i] = out
    cos_t = Hs[:, 1]
    bins = np.linspace(0, 1, 21)
    hist = np.histogram(cos_t, bins=bins)[0]
    # expected mass per cos(theta_h) bin from the vndf
    H, dw, th = hemisphere_grid(2000, 200)
    _, vndf, _ = orc.ggx_batch([0, 1, 0], V, alpha, H)
    visible = (H @ V.astype(np.float64)) > 0
    exp = np.histogram(np.cos(th), bins=bins, weights=vndf * dw * visible)[0] * n
    m = exp > 20
    chi2 = ((hist[m] - exp[m]) ** 2 / exp[m]).sum()
    assert chi2 < 2.5 * m.sum() + 20


# -------------------------------------------------- Fresnel (ExactDielectricFresnel.cpp etc.)
def test_fresnel_dielectric_closed_forms():
    L = orc.lib()
    for n1, n2 in [(1.0, 1.5), (1.0, 1.33), (1.2, 1.7)]:
        r0 = ((n1 - n2) / (n1 + n2)) ** 2
        assert abs(L.orc_fresnel_dielectric(n1, n2, 1.0) - r0) < 1e-6
        assert abs(L.orc_fresnel_dielectric(n1, n2, -1.0) - r0) < 1e-6
        # total internal reflection from the dense side beyond the critical angle
        crit = np.arcsin(n1 / n2)
        assert L.orc_fresnel_dielectric(n1, n2, -np.cos(crit + 0.05)) == 1.0
        # reciprocity: F(theta_i from outside) == F(theta_t from inside)
        ti = 0.7
        tt = np.arcsin(n1 / n2 * np.sin(ti))
        assert abs(L.orc_fresnel_dielectric(n1, n2, np.cos(ti)) - L.orc_fresnel_dielectric(n1, n2, -np.cos(tt))) < 1e-5


def test_fresnel_schlick_endpoints():
    L = orc.lib()
    out = np.zeros(3, np.float32)
    f0 = fp([0.9, 0.5, 0.1])
    L.orc_fresnel_schlick(f0.ctypes.data, C.c_float(1.0), out.ctypes.data)
    assert np.allclose(out, f0)
    L.orc_fresnel_schlick(f0.ctypes.data, C.c_float(0.0), out.ctypes.data)
    assert np.allclose(out, 1.0)
    L.orc_fresnel_schlick(f0.ctypes.data, C.c_float(-0.5), out.ctypes.data)
    exp = f0 + (1 - f0) * 0.5 ** 5
    assert np.allclose(out, exp, atol=1e-6)


def _exact_conductor(eo, n, k, cos):
    L = orc.lib()
    out = np.zeros(3, np.float32)
    n32, k32 = fp(n), fp(k)          # kept alive across the call
    L.orc_fresnel_conductor_exact(C.c_float(eo), n32.ctypes.data, k32.ctypes.data, C.c_float(cos), out.ctypes.data)
    return out


def _complex_fresnel(eo, n, k, cos):
    """Textbook reflectance of unpolarized light at a dielectric (index eo) / conductor
    (index n + ik) interface, in complex arithmetic (independent of the formula under test)."""
    eta = (np.asarray(n, np.float64) + 1j * np.asarray(k, np.float64)) / eo
    c = abs(cos)
    s2 = 1.0 - c * c
    root = np.sqrt(eta * eta - s2)
    rs = (c - root) / (c + root)
    rp = (eta * eta * c - root) / (eta * eta * c + root)
    return 0.5 * (np.abs(rs) ** 2 + np.abs(rp) ** 2)


def test_fresnel_exact_conductor_closed_forms():
    """ExactConductorFresnel (ExactConductorFresnel.cpp:66-101): normal incidence
    ((n - eo)^2 + k^2) / ((n + eo)^2 + k^2); 1 at grazing; k = 0 reduces to the exact
    dielectric Fresnel; every angle against the complex-arithmetic Fresnel equations; the sign
    of cos does not matter."""
    L = orc.lib()
    gold_n, gold_k = [0.143, 0.374, 1.442], [3.983, 2.385, 1.603]
    for eo in (1.0, 1.33):
        f = _exact_conductor(eo, gold_n, gold_k, 1.0)
        n, k = np.array(gold_n), np.array(gold_k)
        assert np.allclose(f, ((n - eo) ** 2 + k ** 2) / ((n + eo) ** 2 + k ** 2), rtol=2e-6)
        assert np.allclose(_exact_conductor(eo, gold_n, gold_k, 0.0), 1.0, atol=1e-6)
        for cos in np.linspace(0.02, 1.0, 50):
            got = _exact_conductor(eo, gold_n, gold_k, cos)
            assert np.allclose(got, _complex_fresnel(eo, gold_n, gold_k, cos), rtol=2e-5, atol=2e-6), cos
            assert np.array_equal(got, _exact_conductor(eo, gold_n, gold_k, -cos))
    for cos in np.linspace(0.05, 1.0, 20):
        d = L.orc_fresnel_dielectric(C.c_float(1.0), C.c_float(1.5), C.c_float(cos))
        c = _exact_conductor(1.0, [1.5] * 3, [0.0] * 3, cos)
        assert np.allclose(c, d, rtol=1e-5, atol=1e-7), cos


def test_exact_conductor_sample_weight_matches_integral():
    """A rough conductor with the exact Fresnel: the sampler's mean weight equals the
    quadrature of its f cos (same check as the Schlick conductors)."""
    m = orc.material(orc.ORC_MAT_CONDUCTOR, roughness=0.5, n=[0.2, 0.9, 1.4], k=[3.9, 2.4, 1.6])
    N = fp([0, 1, 0])
    V = fp([np.sin(0.7), np.cos(0.7), 0])
    rng = np.random.default_rng(3)
    u = rng.random((200_000, 3)).astype(np.float32)
    Ls, w, _, ok = orc.bsdf_sample_batch(m, N, V, u)
    mean_w = np.where(ok[:, None], w, 0).mean(0)
    # quadrature of f cos over the hemisphere
    th = (np.arange(400) + 0.5) / 400 * (np.pi / 2)
    ph = (np.arange(800) + 0.5) / 800 * (2 * np.pi)
    T, P = np.meshgrid(th, ph, indexing="ij")
    Lq = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1).reshape(-1, 3)
    fv, _, okq = orc.bsdf_eval_batch(m, N, V, Lq)
    dw = (np.pi / 2 / 400) * (2 * np.pi / 800) * np.sin(T).reshape(-1)
    integral = (np.where(okq[:, None], fv, 0) * Lq[:, 1:2] * dw[:, None]).sum(0)
    assert np.allclose(mean_w, integral, rtol=0.02), (mean_w, integral)


def test_half_vector_known_answers():
    # surface_optics_test.cpp:9-36
    L = orc.lib()
    h = np.zeros(3, np.float32)
    l1 = fp([0.5, 1, 0.5]) / np.linalg.norm([0.5, 1, 0.5])
    v1 = fp([-0.5, 1, -0.5]) / np.linalg.norm([-0.5, 1, -0.5])
    assert L.orc_half_vector(fp(l1).ctypes.data, fp(v1).ctypes.data, h.ctypes.data) == 1
    assert np.allclose(h, [0, 1, 0], atol=1e-7)
    assert L.orc_half_vector(fp([0, -1, 0]).ctypes.data, fp([0, -1, 0]).ctypes.data, h.ctypes.data) == 1
    assert np.allclose(h, [0, -1, 0])
    assert L.orc_half_vector(fp([1, 0, 0]).ctypes.data, fp([-1, 0, 0]).ctypes.data, h.ctypes.data) == 0


def test_basis_from_unit_y_is_orthonormal_and_right_handed():
    L = orc.lib()
    rng = np.random.default_rng(5)
    x, z = np.zeros(3, np.float32), np.zeros(3, np.float32)
    for _ in range(200):
        y = rng.normal(size=3)
        y = fp(y / np.linalg.norm(y))
        L.orc_basis_from_y(y.ctypes.data, x.ctypes.data, z.ctypes.data)
        M = np.stack([x, y, z]).astype(np.float64)
        assert np.allclose(M @ M.T, np.eye(3), atol=1e-5)
        assert np.allclose(np.cross(z, x), y, atol=1e-5)  # x cross y = z  <=>  z cross x = y


# ---------------------------------------------------- BSDF sampling vs evaluation consistency
MATS = {
    "lambert": orc.material(orc.ORC_MAT_LAMBERT, (0.8, 0.5, 0.2)),
    "conductor_rough": orc.material(orc.ORC_MAT_CONDUCTOR, (0.95, 0.64, 0.54), roughness=0.6),
    "conductor_glossy": orc.material(orc.ORC_MAT_CONDUCTOR, (0.9, 0.9, 0.9), roughness=0.25),
    "dielectric": orc.material(orc.ORC_MAT_DIELECTRIC, (1, 1, 1), roughness=0.5, ior_inner=1.5),
}


def sphere_grid(nt=600, nphi=600):
    th = (np.arange(nt) + 0.5) / nt * np.pi
    ph = (np.arange(nphi) + 0.5) / nphi * (2 * np.pi)
    T, P = np.meshgrid(th, ph, indexing="ij")
    dw = np.sin(T) * (np.pi / nt) * (2 * np.pi / nphi)
    d = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1)
    return d.reshape(-1, 3).astype(np.float32), dw.reshape(-1)


@pytest.mark.parametrize("name", list(MATS))
@pytest.mark.parametrize("vtheta", [0.2, 1.1])
def test_sample_weight_matches_integral_of_bsdf_cos(name, vtheta):
    """E[f cos / pdf] over the sampler must equal the quadrature of f cos over the sphere,
    and the pdf returned by calcBsdfPdf must integrate to at most one."""
    m = MATS[name]
    N = fp([0, 1, 0])
    V = fp([np.sin(vtheta), np.cos(vtheta), 0])
    rng = np.random.default_rng(7)
    u = rng.random((200_000, 3)).astype(np.float32)
    Ls, w, rel, ok = orc.bsdf_sample_batch(m, N, V, u)
    mc = np.where(ok[:, None], w, 0).astype(np.float64)
    mc_mean = mc.mean(0)
    mc_se = mc.std(0) / np.sqrt(len(mc))
    D, dw = sphere_grid()
    fv, pdf, eok = orc.bsdf_eval_batch(m, N, V, D)
    cosl = np.abs(D[:, 1]).astype(np.float64)
    fv = np.where(eok[:, None], fv, 0).astype(np.float64)
    quad = (fv * (cosl * dw)[:, None]).sum(0)
    assert np.all(np.abs(mc_mean - quad) < 5 * mc_se + 3e-3 * np.abs(quad) + 1e-4), (mc_mean, quad, mc_se)
    pdf = np.where(np.isfinite(pdf), pdf, 0).astype(np.float64)
    if name == "dielectric":
        # calcBsdfPdf (TranslucentMicrofacet.cpp:275-360) is also positive for half vectors that
        # no microfacet can realise (|H.V| in ShapeInvariantMicrofacet.h:124-138, no check that H
        # lies on the sides of L and V that Microfacet.h:131-145 requires); the BSDF is zero
        # there and the sampler never goes there.  The sampler's density is the rest.
        Vd = V.astype(np.float64)
        refl = D[:, 1] * Vd[1] > 0
        Hr = D + Vd
        etaI = np.where(D[:, 1] < 0, 1.5, 1.0)[:, None]
        etaT = np.where(D[:, 1] < 0, 1.0, 1.5)[:, None]
        Ht = -(etaI * D + etaT * Vd)
        Hh = np.where(refl[:, None], Hr, Ht)
        Hh *= np.sign(Hh[:, 1:2])
        nol, nov = D[:, 1], Vd[1]
        realisable = ((Hh * D).sum(1) * nol > 0) & ((Hh @ Vd) * nov > 0)
        pdf = np.where(realisable, pdf, 0)
    pint = (pdf * dw).sum()
    assert pint <= 1.0 + 5e-3 and pint > 0.5


def test_lambert_known_values():
    m = MATS["lambert"]
    N, V = fp([0, 1, 0]), fp([0, 1, 0])
    Lq = fp([[0.6, 0.8, 0.0]])
    fv, pdf, ok = orc.bsdf_eval_batch(m, N, V, Lq)
    assert ok[0] and np.allclose(fv[0], np.array([0.8, 0.5, 0.2]) / np.pi, rtol=1e-6)
    assert abs(pdf[0] - 0.8 / np.pi) < 1e-7
    Ls, w, rel, sok = orc.bsdf_sample_batch(m, N, V, np.full((4, 3), 0.3, np.float32))
    assert np.allclose(w, [0.8, 0.5, 0.2]) and np.all(rel == 1)
    # opposite hemispheres do not reflect
    fv, pdf, ok = orc.bsdf_eval_batch(m, N, V, fp([[0.6, -0.8, 0.0]]))
    assert not ok[0] and pdf[0] == 0


# ------------------------------------------- emitter pick distribution (distribution_test.cpp)
def pwc(weights, us=(0.5,)):
    w = fp(weights)
    n = len(w)
    cdf, pdf = np.zeros(n + 1, np.float32), np.zeros(n, np.float32)
    us = fp(us)
    picks = np.zeros(len(us), np.uint32)
    orc.lib().orc_pwc_distribution(w.ctypes.data, n, us.ctypes.data, len(us), cdf.ctypes.data, pdf.ctypes.data,
                                   picks.ctypes.data)
    return cdf, pdf, picks


def test_pwc_distribution_known_answers():
    assert np.allclose(pwc([1, 1])[1], [0.5, 0.5])
    assert np.allclose(pwc([2, 2, 2])[1], [1 / 3] * 3)
    assert np.allclose(pwc([99999])[1], [1.0])
    assert np.allclose(pwc([0])[1], [1.0])
    assert np.allclose(pwc([0, 0, 0])[1], [1 / 3] * 3)
    # zero-weight entries are never sampled (distribution_test.cpp:30-43)
    us = np.random.default_rng(0).random(20000)
    _, _, picks = pwc([0, 0, 3, 0, 4, 5, 0, 6], us)
    assert not np.isin(picks, [0, 1, 3, 6]).any()
    counts = np.bincount(picks, minlength=8)[[2, 4, 5, 7]] / len(us)
    assert np.allclose(counts, np.array([3, 4, 5, 6]) / 18, atol=0.01)


# ---------------------------------------------------------- film (HdrRgbFilm_test.cpp)
def splat(filter_id, W, H, xy, rgb):
    s = np.zeros((H, W, 4), np.float64)
    xy = np.ascontiguousarray(xy, np.float64)
    rgb = fp(rgb)
    orc.lib().orc_film_splat(filter_id, W, H, xy.ctypes.data, rgb.ctypes.data, len(xy), s.ctypes.data)
    out = np.zeros((H, W, 3), np.float32)
    orc.lib().orc_film_develop(W, H, s.ctypes.data, out.ctypes.data)
    return s, out


def test_film_asymmetric_filter_known_answer():
    # HdrRgbFilmTest.DevelopesToFrame, asymmetric Heaviside step: pixel value 3
    _, out = splat(3, 1, 2, [[0.0, 0.5 - 0.123456789], [0.75, 0.5 - 0.54321]], [[3, 3, 3], [1, 1, 1]])
    assert np.allclose(out[0, 0], 3.0)


def test_film_box_filter_is_per_pixel_mean():
    rng = np.random.default_rng(2)
    W, H = 4, 3
    xy = np.stack([rng.integers(0, W, 500) + rng.uniform(0.01, 0.99, 500),
                   rng.integers(0, H, 500) + rng.uniform(0.01, 0.99, 500)], 1)
    rgb = rng.random((500, 3))
    s, out = splat(0, W, H, xy, rgb)
    for y in range(H):
        for x in range(W):
            m = (np.floor(xy[:, 0]) == x) & (np.floor(xy[:, 1]) == y)
            assert s[y, x, 3] == m.sum()
            assert np.allclose(out[y, x], rgb[m].mean(0), rtol=1e-5)


def test_film_gaussian_weighted_mean():
    s, out = splat(1, 1, 2, [[0.2, 0.2], [0.4, 0.4]], [[0.7] * 3, [0.3] * 3])
    w = s[0, 0, 3]
    assert w > 0 and np.allclose(out[0, 0, 0], out[0, 0, 1]) and 0.3 < out[0, 0, 0] < 0.7


def test_film_filter_shapes_closed_forms():
    """One sample at the centre of pixel 2 of a 5 x 1 film: the weights at offsets -2..2 are
    f(dx) f(0).  Mitchell-Netravali (b = c = 1/3): f(0) = (6 - 2b)/6, f(1) = b/6, f(2) = 0;
    Blackman-Harris (radius 2): 1 at the centre, 0.35875 - 0.14128 at distance 1;
    submerged Gaussian (sigma 0.5): (exp(-2 dx^2) - exp(-16)) relative to the centre."""
    def w(fid):
        s, _ = splat(fid, 5, 1, [[2.5, 0.5]], [[1, 1, 1]])
        return s[0, :, 3]
    b = 1 / 3
    mn = w(4)
    assert np.allclose(mn, [0, b / 6 * (6 - 2 * b) / 6, ((6 - 2 * b) / 6) ** 2, b / 6 * (6 - 2 * b) / 6, 0], atol=1e-15)
    bh = w(2)
    assert np.isclose(bh[2], 1.0) and np.allclose(bh[[1, 3]], 0.35875 - 0.14128)
    assert np.allclose(bh[[0, 4]], 0.35875 - 0.48829 + 0.14128 - 0.01168, atol=1e-12)
    g = w(1)
    e = np.exp(-16.0)
    assert np.allclose(g[[1, 3]] / g[2], (np.exp(-2.0) - e) / (1 - e))


@pytest.mark.parametrize("fid", [1, 2, 3])
def test_filtered_film_soft_edge_weight_sums(fid):
    """fid: PHG_FILTER_GAUSSIAN, _MN, _BH.  Soft-edged sampling window (TSamplingFilm.ipp:103-134): the W x H samples of a batch
    are spread over the pixel centres expanded by the filter half size 2, i.e. a window of
    (W + 3) x (H + 3) pixels, so the whole 4 x 4 footprint of every pixel, border and corner
    pixels included, receives samples at the same density W H / ((W + 3)(H + 3)).  The expected
    weight sum per pixel and sample is then that density times the filter's integral: the
    closed forms (4 * 0.35875)^2 for Blackman-Harris and 1 for Mitchell-Netravali, a quadrature
    for the submerged Gaussian.  Without the soft edge a corner pixel would get about a third
    of an interior one's weight."""
    import phg
    W = H = 8
    spp = 16384
    sc = scene_from_tris(np.array([[0, 0, -50, 1, 0, -50, 0, 1, -50]], np.float32))   # behind the camera
    film = orc.render_filtered(sc, W, H, spp, fid, max_bounces=0, seed=9)
    wsum = film[..., 3] / spp
    density = W * H / ((W + 3) * (H + 3))
    if fid == 1:
        x = (np.arange(4000) + 0.5) / 4000 * 4 - 2
        X, Y = np.meshgrid(x, x)
        g = np.maximum(np.exp(-(X * X + Y * Y) / 0.5) - np.exp(-16.0), 0).sum() * (4 / 4000) ** 2
        integral = g
    elif fid == 3:
        integral = (4 * 0.35875) ** 2
    else:
        integral = 1.0
    expect = density * integral
    dev = np.abs(wsum / expect - 1)
    print(f"[film] filter {fid}: expected {expect:.5f}, corner {wsum[0, 0]:.5f}, centre {wsum[4, 4]:.5f}, "
          f"max deviation {dev.max():.4f}")
    assert dev.max() < 0.02


# ---------------------------------------------------------------- camera (PinholeCamera.cpp)
def test_camera_rays_horizontal_fov():
    cam = orc.orc_camera()
    cam.pos[:] = [1, 2, 3]
    cam.right[:] = [1, 0, 0]
    cam.up[:] = [0, 1, 0]
    cam.forward[:] = [0, 0, 1]
    fov = np.radians(40)
    cam.tan_half_fov = np.tan(fov / 2)
    W, H = 256, 128
    raster = fp([[128, 64], [0, 64], [256, 64], [128, 0]])
    out = np.zeros((4, 8), np.float32)
    orc.lib().orc_camera_rays(C.byref(cam), W, H, raster.ctypes.data, 4, out.ctypes.data)
    d = out[:, 4:7]
    assert np.allclose(out[:, 0:3], [1, 2, 3])
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-6)
    assert np.allclose(d[0], [0, 0, 1], atol=1e-7)
    assert abs(np.degrees(np.arccos(d[1] @ [0, 0, 1])) - 20) < 1e-3 and d[1, 0] < 0
    assert d[2, 0] > 0  # raster x grows to the right
    vfov_half = np.arctan(np.tan(fov / 2) * H / W)
    assert abs(np.arccos(d[3] @ [0, 0, 1]) - vfov_half) < 1e-5 and d[3, 1] < 0  # raster y = 0 is the bottom


# ------------------------------------------------------------------- NEE (ESPowerFavoring)
def test_nee_sample_pdf_round_trip():
    import phg
    sc = phg.Scene.generate(phg.GEN_PROCEDURAL, seed=3, target_tris=20_000, num_instances=4, num_lights=5, subdiv=2)
    sv = orc.SceneView(sc)
    L = orc.lib()
    rng = np.random.default_rng(4)
    ground = np.where(sc.tri_material == 1)[0]
    n_ok = 0
    for i in range(300):
        t = int(rng.choice(ground))
        bary = fp(rng.dirichlet((1, 1, 1)))
        u = fp(rng.random(4) * 0.998 + 0.001)
        pos, pdfw, le = np.zeros(3, np.float32), np.zeros(1, np.float32), np.zeros(3, np.float32)
        ok = L.orc_nee_sample(C.byref(sv.s), t, bary.ctypes.data, u.ctypes.data, pos.ctypes.data, pdfw.ctypes.data,
                              le.ctypes.data)
        if not ok:
            continue
        n_ok += 1
        # find the emitter triangle and barycentrics of the sampled point, then ask for the pdf
        ei = int(np.searchsorted(np.cumsum(np.ones(len(sc.emitters))), 0))
        best = None
        for et in sc.emitter_tris:
            v = sc.verts[et].reshape(3, 3).astype(np.float64)
            M = np.stack([v[0], v[1], v[2]], 1)
            try:
                b = np.linalg.lstsq(M, pos.astype(np.float64), rcond=None)[0]
            except np.linalg.LinAlgError:
                continue
            if np.all(b > -1e-4) and abs(b.sum() - 1) < 1e-4 and np.allclose(M @ b, pos, atol=1e-4):
                best = (int(et), fp(b))
                break
        assert best is not None
        p2 = L.orc_nee_pdf(C.byref(sv.s), t, bary.ctypes.data, best[0], best[1].ctypes.data)
        assert abs(p2 - pdfw[0]) <= 2e-3 * pdfw[0]
    assert n_ok > 50


# --------------------------------------------------------- end to end: furnace closed form
@pytest.mark.parametrize("albedo,radiance", [(0.5, 0.5), (0.8, 0.2)])
def test_oracle_furnace_closed_form(albedo, radiance):
    """Closed Lambertian enclosure emitting everywhere: E[L] = Le / (1 - rho).  Exercises
    NEE pdfs, both MIS weights, BSDF sample weights and Russian roulette together."""
    sc = furnace_scene(albedo=albedo, radiance=radiance)
    film, _ = orc.render(sc, 16, 16, 32, max_bounces=10000, seed=3)
    n = film[..., 3].sum()
    mean = film[..., 0].sum() / n
    var = film[..., 4].sum() / n - mean * mean
    se = np.sqrt(var / n)
    assert abs(mean - radiance / (1 - albedo)) < 4 * se + 1e-4


def test_oracle_furnace_bounce_cap_truncates_series():
    sc = furnace_scene(albedo=0.5, radiance=0.5)
    for B, exp in [(0, 0.5), (1, 0.75), (2, 0.875)]:
        film, _ = orc.render(sc, 16, 16, 64, max_bounces=B, seed=5)
        n = film[..., 3].sum()
        mean = film[..., 0].sum() / n
        se = np.sqrt(max(film[..., 4].sum() / n - mean * mean, 0) / n)
        assert abs(mean - exp) < 4 * se + 1e-4, (B, mean, exp)


def test_oracle_deterministic_per_path():
    import phg
    sc = phg.Scene.generate(phg.GEN_CORNELL)
    a, _ = orc.render(sc, 16, 16, 4, max_bounces=5, seed=1, nthreads=1)
    b, _ = orc.render(sc, 16, 16, 4, max_bounces=5, seed=1, nthreads=4)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("estimator", [1, 2])
def test_oracle_bvpt_furnace(estimator):
    """BVPT (unbounded) reaches the furnace's Le / (1 - rho); BVPTDL, capped at one bounce,
    reaches Le (1 + rho), the first two terms of the series."""
    albedo, radiance = 0.5, 0.5
    sc = furnace_scene(albedo=albedo, radiance=radiance)
    film, st = orc.render(sc, 16, 16, 64, max_bounces=10000, seed=7, estimator=estimator)
    n = film[..., 3].sum()
    mean = film[..., 0].sum() / n
    se = np.sqrt(max(film[..., 4].sum() / n - mean * mean, 0) / n)
    expect = radiance / (1 - albedo) if estimator == 1 else radiance * (1 + albedo)
    assert st["rays_shadow"] == 0                      # no next-event estimation
    assert abs(mean - expect) < 4 * se + 1e-4, (mean, expect)


def big_light_cornell():
    """The Cornell box with its ceiling light widened from 0.25 to 0.8: paths that must hit
    the light to count (BVPT) then have a moderate variance, so per-pixel intervals hold."""
    import phg
    sc = phg.Scene.generate(phg.GEN_CORNELL)
    v = sc.verts.reshape(-1, 3, 3).copy()
    for t in sc.emitter_tris:
        v[t, :, 0] = 0.5 + (v[t, :, 0] - 0.5) * 3.2
        v[t, :, 2] = 0.5 + (v[t, :, 2] - 0.5) * 3.2
    sc.verts = v.reshape(-1, 9)
    return sc


@pytest.mark.parametrize("estimator,cap", [(1, 5), (2, 10000)])
def test_oracle_bvpt_matches_bneept_in_expectation(estimator, cap):
    """Same integral, different sampling: BVPT at a bounce cap B and BNEEPT at B agree per
    pixel; BVPTDL agrees with BNEEPT capped at one bounce."""
    import phg
    from helpers import pixel_z
    sc = big_light_cornell()
    a, _ = orc.render(sc, 24, 24, 512, max_bounces=cap, seed=11, estimator=estimator)
    b, _ = orc.render(sc, 24, 24, 512, max_bounces=cap if estimator == 1 else 1, seed=12, estimator=0)
    z = pixel_z(a, b)
    frac = np.mean(np.abs(z) > 2.5758)
    ma, mb = a[..., :3].sum() / a[..., 3].sum(), b[..., :3].sum() / b[..., 3].sum()
    assert frac < 0.03, frac
    assert abs(ma - mb) < 0.03 * mb, (ma, mb)


def test_native_build_renders_the_same_film():
    """The timed CPU baseline is the same source built -O3 -march=native (contraction off): it
    renders the oracle's film bit for bit."""
    import phg
    orc.build_native()
    sc = phg.Scene.generate(phg.GEN_CORNELL)
    a, _ = orc.render(sc, 24, 24, 4, max_bounces=5, seed=3)
    b, _ = orc.render(sc, 24, 24, 4, max_bounces=5, seed=3, native=True)
    assert np.array_equal(a, b)


# ---- emitter-pick weights: the reference's colour-energy estimate
def test_color_energy_matches_the_reference_estimate():
    """math::estimate_color_energy of emitted linear-sRGB radiance (DiffuseSurfaceEmitter.cpp:
    55-86) against tests/golden/color_energy.json, which make_color_energy.py computes from the
    reference's Smits and D65 tables (resampled as the reference does): the oracle's seven basis
    energies and Smits branches reproduce every vector to float rounding."""
    import json
    import os
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "color_energy.json")))
    for r, g, b, e in gold["vectors"]:
        got = orc.color_energy([r, g, b])
        assert abs(got - e) <= 2e-7 * max(1.0, abs(e)), (r, g, b, got, e)


def test_color_energy_invariants():
    """What the construction fixes: white (1, 1, 1) has unit energy (the bases reproduce
    illuminant E and D65 is normalised), complementary bases add up to white, the estimate is
    positively homogeneous and continuous across Smits' branches, and never negative."""
    e = orc.color_energy
    assert abs(e([1, 1, 1]) - 1.0) < 1e-3
    for c, p in (([0, 1, 1], [1, 0, 0]), ([1, 0, 1], [0, 1, 0]), ([1, 1, 0], [0, 0, 1])):
        assert abs(e(c) + e(p) - e([1, 1, 1])) < 0.01
    rng = np.random.default_rng(8)
    for c in rng.uniform(0, 3, (50, 3)):
        assert abs(e(2.5 * c) - 2.5 * e(c)) <= 1e-6 * e(c) * 2.5 + 1e-7
        # continuity: a colour on a branch boundary and a nudge either side
        for i, j in ((0, 1), (1, 2), (0, 2)):
            d = c.copy(); d[j] = d[i]
            for eps in (1e-6, -1e-6):
                dd = d.copy(); dd[j] += eps
                assert abs(e(dd) - e(d)) < 1e-5
    assert e([-1, 0, 0]) == 0.0 and e([0, 0, 0]) == 0.0
