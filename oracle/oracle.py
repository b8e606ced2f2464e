"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so) — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this module, and
only as the checker / the timed CPU comparator.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

ORC_MAT_ABSORBER, ORC_MAT_LAMBERT, ORC_MAT_CONDUCTOR, ORC_MAT_DIELECTRIC = 0, 1, 2, 3
FLT_MAX = float(np.finfo(np.float32).max)


class orc_material(C.Structure):
    _fields_ = [("type", C.c_uint32), ("rgb", C.c_float * 3), ("roughness", C.c_float),
                ("ior_outer", C.c_float), ("ior_inner", C.c_float), ("fresnel", C.c_uint32),
                ("n", C.c_float * 3), ("k", C.c_float * 3), ("pad", C.c_float * 2)]


class orc_emitter(C.Structure):
    _fields_ = [("radiance", C.c_float * 3), ("tri_begin", C.c_uint32), ("tri_count", C.c_uint32),
                ("pad", C.c_uint32 * 3)]


class orc_camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("right", C.c_float * 3), ("up", C.c_float * 3),
                ("forward", C.c_float * 3), ("tan_half_fov", C.c_float), ("pad", C.c_float * 3)]


class orc_scene(C.Structure):
    _fields_ = [("num_tris", C.c_uint32), ("num_materials", C.c_uint32), ("num_emitters", C.c_uint32),
                ("num_emitter_tris", C.c_uint32), ("verts", C.c_void_p), ("tri_material", C.c_void_p),
                ("tri_emitter", C.c_void_p), ("materials", C.c_void_p), ("emitters", C.c_void_p),
                ("emitter_tris", C.c_void_p), ("camera", orc_camera)]


class orc_render_opts(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp_begin", C.c_uint32), ("spp", C.c_uint32),
                ("max_bounces", C.c_uint32), ("nthreads", C.c_uint32), ("seed", C.c_uint64),
                ("row_step", C.c_uint32), ("estimator", C.c_uint32), ("sampler", C.c_uint32), ("pad", C.c_uint32)]


class orc_render_stats(C.Structure):
    _fields_ = [("rays_primary", C.c_uint64), ("rays_extension", C.c_uint64), ("rays_shadow", C.c_uint64),
                ("paths", C.c_uint64), ("render_ms", C.c_double)]


NATIVE_LIB_PATH = os.path.join(_HERE, "_build", "liboracle_native.so")
NATIVE_FLAGS = "-O3 -march=native -ffp-contract=off -fno-fast-math"
_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def build_native(force: bool = False):
    """The timed CPU-baseline build (-O3 -march=native) for the host this runs on; force
    rebuilds it even when a build (perhaps for another host) is already there."""
    subprocess.run(["make", "-s"] + (["-B"] if force else []) + ["-C", _HERE, "native"], check=True)


def lib(native: bool = False) -> C.CDLL:
    """The oracle library; native=True loads the -O3 -march=native build (build_native())."""
    if native in _libs:
        return _libs[native]
    path = NATIVE_LIB_PATH if native else LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing; build it with `make -C oracle{' native' if native else ''}`")
    L = C.CDLL(path)
    vp, f, u32, u64, i32 = C.c_void_p, C.c_float, C.c_uint32, C.c_uint64, C.c_int32
    L.orc_pcg32_generate.argtypes = [u64, u64, C.c_int, u64, u32, vp]
    L.orc_moremur.argtypes = [u64]
    L.orc_moremur.restype = u64
    L.orc_set_fault.argtypes = [C.c_int]
    L.orc_set_fault.restype = None
    L.orc_stream_samples.argtypes = [u64, u32, vp]
    L.orc_intersect_triangle.argtypes = [vp, vp, vp, vp]
    L.orc_intersect_brute.argtypes = [vp, u32, vp, u32, vp, vp, vp, vp, u32]
    L.orc_intersect_refbvh.argtypes = [vp, u32, vp, u32, vp, vp, u32]
    L.orc_bsdf_eval.argtypes = [vp, vp, vp, vp, vp]
    L.orc_bsdf_pdf.argtypes = [vp, vp, vp, vp]
    L.orc_bsdf_pdf.restype = f
    L.orc_bsdf_sample.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.orc_bsdf_sample_batch.argtypes = [vp, vp, vp, u32, vp, vp, vp, vp, vp]
    L.orc_bsdf_eval_batch.argtypes = [vp, vp, vp, u32, vp, vp, vp, vp]
    L.orc_ggx_batch.argtypes = [vp, vp, f, u32, vp, vp, vp, vp]
    L.orc_ggx_D.argtypes = [vp, vp, f]
    L.orc_ggx_D.restype = f
    L.orc_ggx_G.argtypes = [vp, vp, vp, vp, f]
    L.orc_ggx_G.restype = f
    L.orc_ggx_vndf_pdf.argtypes = [vp, vp, vp, f]
    L.orc_ggx_vndf_pdf.restype = f
    L.orc_ggx_sample_vh.argtypes = [vp, vp, f, f, f, vp]
    L.orc_fresnel_dielectric.argtypes = [f, f, f]
    L.orc_fresnel_dielectric.restype = f
    L.orc_fresnel_schlick.argtypes = [vp, f, vp]
    L.orc_fresnel_conductor_exact.argtypes = [f, vp, vp, f, vp]
    L.orc_half_vector.argtypes = [vp, vp, vp]
    L.orc_basis_from_y.argtypes = [vp, vp, vp]
    L.orc_pwc_distribution.argtypes = [vp, u32, vp, u32, vp, vp, vp]
    L.orc_nee_sample.argtypes = [C.POINTER(orc_scene), i32, vp, vp, vp, vp, vp]
    L.orc_nee_pdf.argtypes = [C.POINTER(orc_scene), i32, vp, i32, vp]
    L.orc_nee_pdf.restype = f
    L.orc_camera_rays.argtypes = [C.POINTER(orc_camera), u32, u32, vp, u32, vp]
    L.orc_render.argtypes = [C.POINTER(orc_scene), C.POINTER(orc_render_opts), vp, C.POINTER(orc_render_stats)]
    L.orc_render_filtered.argtypes = [C.POINTER(orc_scene), C.POINTER(orc_render_opts), C.c_int, vp]
    L.orc_render_paths.argtypes = [C.POINTER(orc_scene), C.POINTER(orc_render_opts), vp, vp, u32, vp]
    L.orc_film_splat.argtypes = [C.c_int, C.c_int64, C.c_int64, vp, vp, u32, vp]
    L.orc_film_develop.argtypes = [C.c_int64, C.c_int64, vp, vp]
    _libs[native] = L
    return L


def f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, np.float32))


def material(mtype: int, rgb=(0.5, 0.5, 0.5), roughness=0.3, ior_outer=1.0, ior_inner=1.5, n=None,
             k=None) -> orc_material:
    """A material record; a conductor given n and k uses the exact conductor Fresnel."""
    m = orc_material()
    m.type = mtype
    m.rgb[:] = [float(x) for x in rgb]
    m.roughness = roughness
    m.ior_outer = ior_outer
    m.ior_inner = ior_inner
    if n is not None:
        m.fresnel = 1
        m.n[:] = [float(x) for x in n]
        m.k[:] = [float(x) for x in k]
    return m


class SceneView:
    """Holds an orc_scene pointing at the arrays of a scene object with the attributes of
    phg.Scene (verts, tri_material, tri_emitter, materials, emitters, emitter_tris, camera)."""

    def __init__(self, scene):
        self._keep = [np.ascontiguousarray(scene.verts, np.float32),
                      np.ascontiguousarray(scene.tri_material, np.uint32),
                      np.ascontiguousarray(scene.tri_emitter, np.int32),
                      np.ascontiguousarray(scene.materials),
                      np.ascontiguousarray(scene.emitters),
                      np.ascontiguousarray(scene.emitter_tris, np.uint32)]
        s = orc_scene()
        s.num_tris = self._keep[0].shape[0]
        s.num_materials = len(self._keep[3])
        s.num_emitters = len(self._keep[4])
        s.num_emitter_tris = len(self._keep[5])
        s.verts, s.tri_material, s.tri_emitter, s.materials, s.emitters, s.emitter_tris = \
            [k.ctypes.data if k.size else None for k in self._keep]
        cam = orc_camera()
        for k in ("pos", "right", "up", "forward"):
            getattr(cam, k)[:] = [float(x) for x in scene.camera[k]]
        cam.tan_half_fov = float(scene.camera["tan_half_fov"])
        s.camera = cam
        self.s = s


FILTER_IDS = {0: 0, 1: 1, 2: 4, 3: 2}   # PHG_FILTER_* (box, gaussian, mn, bh) -> oracle filter id


def render_filtered(scene, width, height, spp, filter, max_bounces=10000, seed=0, spp_begin=0, nthreads=None,
                    estimator=0, sampler=0):
    """Weighted FP64 film (H, W, 8) of an oracle render splatted with a PHG_FILTER_* filter."""
    L = lib()
    sv = SceneView(scene)
    film = np.zeros((height, width, 8), np.float64)
    o = orc_render_opts(width, height, spp_begin, spp, max_bounces, nthreads or os.cpu_count() or 1, seed, 1,
                        estimator, sampler)
    L.orc_render_filtered(C.byref(sv.s), C.byref(o), FILTER_IDS[filter], film.ctypes.data)
    return film


def render(scene, width, height, spp, max_bounces=10000, seed=0, spp_begin=0, nthreads=None, row_step=1,
           estimator=0, native=False, sampler=0):
    """FP64 film (H, W, 8) of the oracle render (estimator 0 BNEEPT, 1 BVPT, 2 BVPTDL; sampler 0
    independent, 1 stratified, 2 Halton with its permutation / sequence bits), plus ray counts
    (zero for Halton, whose samples are splatted).  native=True runs the -O3 -march=native
    build."""
    L = lib(native)
    sv = SceneView(scene)
    film = np.zeros((height, width, 8), np.float64)
    o = orc_render_opts(width, height, spp_begin, spp, max_bounces, nthreads or os.cpu_count() or 1, seed, row_step,
                        estimator, sampler)
    st = orc_render_stats()
    L.orc_render(C.byref(sv.s), C.byref(o), film.ctypes.data, C.byref(st))
    return film, {"rays_primary": st.rays_primary, "rays_extension": st.rays_extension, "rays_shadow": st.rays_shadow,
                  "paths": st.paths, "render_ms": st.render_ms}


def render_paths(scene, width, height, pixels, samples, max_bounces=10000, seed=0, sampler=0):
    L = lib()
    sv = SceneView(scene)
    pix = np.ascontiguousarray(pixels, np.uint32)
    smp = np.ascontiguousarray(samples, np.uint32)
    out = np.zeros((len(pix), 3), np.float32)
    o = orc_render_opts(width, height, 0, 0, max_bounces, 1, seed, 0, 0, sampler)
    L.orc_render_paths(C.byref(sv.s), C.byref(o), pix.ctypes.data, smp.ctypes.data, len(pix), out.ctypes.data)
    return out


def color_energy(rgb):
    """Colour energy of a linear-sRGB radiance (the emitter-pick weight per unit area)."""
    L = lib()
    L.orc_color_energy.restype = C.c_float
    L.orc_color_energy.argtypes = [C.c_void_p]
    a = np.ascontiguousarray(rgb, np.float32).reshape(3)
    return float(L.orc_color_energy(a.ctypes.data))


def halton_dim(dim, index, sampler, leap=1, seed=0):
    """Dimension dim (0..6) of Halton point `index` under a sampler value (kind 2 with its
    permutation and sequence bits) and leap."""
    L = lib()
    L.orc_halton_dim.restype = C.c_float
    L.orc_halton_dim.argtypes = [C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64]
    return L.orc_halton_dim(dim, index, sampler, leap, seed)


def kensler_permute(i, n, key):
    """Kensler's hashed permutation of [0, n) with key `key` (math::permuted_index)."""
    L = lib()
    L.orc_kensler_permute.restype = C.c_uint32
    L.orc_kensler_permute.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
    return L.orc_kensler_permute(i, n, key)


def halton_float_digits(base):
    L = lib()
    L.orc_halton_float_digits.restype = C.c_uint32
    L.orc_halton_float_digits.argtypes = [C.c_uint32]
    return L.orc_halton_float_digits(base)


def radical_inverse(base, value):
    L = lib()
    L.orc_radical_inverse.restype = C.c_float
    L.orc_radical_inverse.argtypes = [C.c_uint32, C.c_uint64]
    return L.orc_radical_inverse(base, value)


def good_prime(n):
    L = lib()
    L.orc_good_prime.restype = C.c_uint32
    L.orc_good_prime.argtypes = [C.c_uint32]
    return L.orc_good_prime(n)


def halton_leap(sampler, seed):
    L = lib()
    L.orc_halton_leap.restype = C.c_uint32
    L.orc_halton_leap.argtypes = [C.c_uint32, C.c_uint64]
    return L.orc_halton_leap(sampler, seed)


def halton_raster(width, height, sample, sampler, seed=0, half=0.5):
    """Raster positions (FP64) of the Halton source's batch `sample`: (W * H, 2), pixel order."""
    L = lib()
    L.orc_halton_raster.restype = None
    L.orc_halton_raster.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32, C.c_uint64,
                                    C.c_void_p]
    out = np.zeros((width * height, 2), np.float64)
    L.orc_halton_raster(width, height, sample, half, sampler, seed, out.ctypes.data)
    return out


def strat_perm(i, n, key):
    """The stratified source's keyed bijection of [0, n)."""
    L = lib()
    L.orc_strat_perm.restype = C.c_uint32
    L.orc_strat_perm.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64]
    return L.orc_strat_perm(i, n, key)


def strat_key(seed, sample, stage):
    L = lib()
    L.orc_strat_key.restype = C.c_uint64
    L.orc_strat_key.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32]
    return L.orc_strat_key(seed, sample, stage)


def strat_batch(npix, seed, sample, jitter5):
    """First-vertex dimensions (u0, u2, u3, u4, u5) of every pixel of one stratified batch, from
    the pixels' stream values jitter5 (npix, 5)."""
    L = lib()
    j = np.ascontiguousarray(jitter5, np.float32).reshape(npix, 5)
    out = np.zeros((npix, 5), np.float32)
    L.orc_strat_batch.argtypes = [C.c_uint32, C.c_uint64, C.c_uint32, C.c_void_p, C.c_void_p]
    L.orc_strat_batch(npix, seed, sample, j.ctypes.data, out.ctypes.data)
    return out


def intersect_brute(verts, rays, nthreads=None):
    L = lib()
    verts = f32(verts).reshape(-1, 9)
    rays = f32(rays).reshape(-1, 8)
    n = rays.shape[0]
    tri = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    bary = np.zeros((n, 3), np.float32)
    flags = np.zeros(n, np.uint32)
    L.orc_intersect_brute(verts.ctypes.data, verts.shape[0], rays.ctypes.data, n, tri.ctypes.data, t.ctypes.data,
                          bary.ctypes.data, flags.ctypes.data, nthreads or os.cpu_count() or 1)
    return tri, t, bary, flags


def intersect_refbvh(verts, rays, nthreads=None):
    L = lib()
    verts = f32(verts).reshape(-1, 9)
    rays = f32(rays).reshape(-1, 8)
    n = rays.shape[0]
    tri = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    L.orc_intersect_refbvh(verts.ctypes.data, verts.shape[0], rays.ctypes.data, n, tri.ctypes.data, t.ctypes.data,
                           nthreads or os.cpu_count() or 1)
    return tri, t


def bsdf_sample_batch(m, N, V, u):
    L = lib()
    u = f32(u).reshape(-1, 3)
    n = u.shape[0]
    Ls, w = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32)
    rel, ok = np.zeros(n, np.float32), np.zeros(n, np.int32)
    N, V = f32(N), f32(V)
    L.orc_bsdf_sample_batch(C.byref(m), N.ctypes.data, V.ctypes.data, n, u.ctypes.data, Ls.ctypes.data,
                            w.ctypes.data, rel.ctypes.data, ok.ctypes.data)
    return Ls, w, rel, ok.astype(bool)


def bsdf_eval_batch(m, N, V, Ls):
    L = lib()
    Ls = f32(Ls).reshape(-1, 3)
    n = Ls.shape[0]
    fv, pdf, ok = np.zeros((n, 3), np.float32), np.zeros(n, np.float32), np.zeros(n, np.int32)
    N, V = f32(N), f32(V)
    L.orc_bsdf_eval_batch(C.byref(m), N.ctypes.data, V.ctypes.data, n, Ls.ctypes.data, fv.ctypes.data,
                          pdf.ctypes.data, ok.ctypes.data)
    return fv, pdf, ok.astype(bool)


def ggx_batch(N, V, alpha, H):
    L = lib()
    H = f32(H).reshape(-1, 3)
    n = H.shape[0]
    D, vndf, lam = np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32)
    N, V = f32(N), f32(V)
    L.orc_ggx_batch(N.ctypes.data, V.ctypes.data, alpha, n, H.ctypes.data, D.ctypes.data, vndf.ctypes.data,
                    lam.ctypes.data)
    return D, vndf, lam
