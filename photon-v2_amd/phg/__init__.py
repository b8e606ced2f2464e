"""Python binding of the BNEEPT hot path's C ABI (include/phg.h).

The compute path is the HIP library ``photon-v2_amd/lib/libphg.so``; this module only
marshals arrays.  It never falls back to a CPU implementation: if the library or a GPU is
missing, the calls that need them raise.

Mirrors the reference's renderer seam: ``Context.upload`` takes the place of the renderer
update (Core/Renderer/Renderer.h:38-93, doUpdate), ``Context.render`` of doRender
(Core/Renderer/Sampling/EqualSamplingRenderer.cpp:119-203), ``Context.film`` of
retrieveFrame, ``Context.stats`` of asyncQueryRenderStats.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libphg.so")

PHG_OK = 0
STATUS_NAMES = {0: "OK", 1: "INVALID_ARG", 2: "HIP", 3: "OUT_OF_MEMORY", 4: "NO_SCENE", 5: "NO_DEVICE",
                6: "UNSUPPORTED"}

MAT_ABSORBER, MAT_LAMBERT, MAT_CONDUCTOR, MAT_DIELECTRIC = 0, 1, 2, 3
FRESNEL_SCHLICK, FRESNEL_EXACT = 0, 1
GEN_CORNELL, GEN_PROCEDURAL, GEN_OCCLUDED_ROOM, GEN_TERRAIN = 1, 2, 3, 4
QUERY_CLOSEST, QUERY_ANY = 0, 1
PTR_HOST, PTR_DEVICE = 0, 1
RENDER_COUNT_TRAVERSAL, RENDER_GUIDED, RENDER_TRAIN, RENDER_GRAPH = 1, 2, 4, 8
BATCH_ORDER_GIVEN, BATCH_ORDER_SHORTEST_FIRST = 0, 1
BATCH_COUNT_TRAVERSAL, BATCH_GRAPH, BATCH_GUIDED, BATCH_TRAIN = 1, 2, 4, 8
JOB_CAMERA = 1
PHG_ESTIMATOR_BNEEPT, PHG_ESTIMATOR_BVPT, PHG_ESTIMATOR_BVPTDL = 0, 1, 2
PHG_FILTER_BOX, PHG_FILTER_GAUSSIAN, PHG_FILTER_MN, PHG_FILTER_BH = 0, 1, 2, 3
SAMPLER_INDEPENDENT, SAMPLER_STRATIFIED, SAMPLER_HALTON = 0, 1, 2
# Halton permutation and sequence, OR-ed into SAMPLER_HALTON (phg.h PHG_HALTON_*)
HALTON_PERM_FIXED, HALTON_PERM_NONE, HALTON_PERM_PER_DIGIT, HALTON_PERM_OWEN = 0 << 8, 1 << 8, 2 << 8, 3 << 8
HALTON_SEQ_ORIGINAL, HALTON_SEQ_LEAP, HALTON_SEQ_RANDOM_LEAP, HALTON_SEQ_RANDOM_START = 0, 1 << 12, 2 << 12, 3 << 12
GUIDE_PARAMS, GUIDE_OUTPUTS, GUIDE_RECORD = 19305, 41, 16
FLT_MAX = float(np.finfo(np.float32).max)


class phg_material(C.Structure):
    _fields_ = [("type", C.c_uint32), ("rgb", C.c_float * 3), ("roughness", C.c_float),
                ("ior_outer", C.c_float), ("ior_inner", C.c_float), ("fresnel", C.c_uint32),
                ("n", C.c_float * 3), ("k", C.c_float * 3), ("pad", C.c_float * 2)]


class phg_emitter(C.Structure):
    _fields_ = [("radiance", C.c_float * 3), ("tri_begin", C.c_uint32), ("tri_count", C.c_uint32),
                ("pad", C.c_uint32 * 3)]


class phg_camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("right", C.c_float * 3), ("up", C.c_float * 3),
                ("forward", C.c_float * 3), ("tan_half_fov", C.c_float), ("pad", C.c_float * 3)]


class phg_batch_job(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("max_bounces", C.c_uint32), ("seed", C.c_uint32), ("flags", C.c_uint32), ("estimator", C.c_uint32),
                ("camera", phg_camera), ("filter", C.c_uint32), ("sampler", C.c_uint32), ("pad2", C.c_uint32 * 2)]


class phg_scene_desc(C.Structure):
    _fields_ = [("num_tris", C.c_uint32), ("num_materials", C.c_uint32), ("num_emitters", C.c_uint32),
                ("num_emitter_tris", C.c_uint32), ("verts", C.POINTER(C.c_float)),
                ("tri_material", C.POINTER(C.c_uint32)), ("tri_emitter", C.POINTER(C.c_int32)),
                ("materials", C.POINTER(phg_material)), ("emitters", C.POINTER(phg_emitter)),
                ("emitter_tris", C.POINTER(C.c_uint32)), ("camera", phg_camera)]


class phg_device_opts(C.Structure):
    _fields_ = [("device", C.c_int32), ("flags", C.c_uint32)]


class phg_bvh_opts(C.Structure):
    _fields_ = [("max_leaf_tris", C.c_uint32), ("sah_bins", C.c_uint32), ("nthreads", C.c_uint32),
                ("width", C.c_uint32)]


class phg_batch_opts(C.Structure):
    _fields_ = [("paths_per_pass", C.c_uint32), ("order", C.c_uint32), ("flags", C.c_uint32),
                ("guide_steps", C.c_uint32), ("guide_prob", C.c_float), ("guide_lr", C.c_float),
                ("guide_batch", C.c_uint32), ("guide_depth", C.c_uint32)]


class phg_batch_result(C.Structure):
    _fields_ = [("done_ms", C.c_double), ("paths", C.c_uint64)]


class phg_tuning(C.Structure):
    _fields_ = [("trace_block", C.c_uint32), ("paths_per_pass", C.c_uint32), ("trace_blocks_per_cu", C.c_uint32),
                ("pad", C.c_uint32)]


class phg_tune_trial(C.Structure):
    _fields_ = [("max_leaf_tris", C.c_uint32), ("width", C.c_uint32), ("trace_block", C.c_uint32),
                ("paths_per_pass", C.c_uint32), ("msamples_per_s", C.c_double)]


class phg_autotune_opts(C.Structure):
    _fields_ = [("sample_jobs", C.c_uint32), ("sample_spp", C.c_uint32), ("repeats", C.c_uint32), ("pad", C.c_uint32)]


class phg_p2_settings(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32), ("estimator", C.c_uint32),
                ("max_bounces", C.c_uint32), ("filter", C.c_uint32), ("warnings", C.c_uint32), ("sampler", C.c_uint32)]


class phg_render_opts(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32), ("spp_begin", C.c_uint32),
                ("max_bounces", C.c_uint32), ("spp_per_pass", C.c_uint32), ("seed", C.c_uint64),
                ("flags", C.c_uint32), ("guide_prob", C.c_float), ("guide_lr", C.c_float),
                ("guide_batch", C.c_uint32), ("guide_steps", C.c_uint32), ("estimator", C.c_uint32),
                ("filter", C.c_uint32), ("guide_depth", C.c_uint32), ("sampler", C.c_uint32)]


class phg_stats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("rays_primary", C.c_uint64), ("rays_extension", C.c_uint64),
                ("rays_shadow", C.c_uint64), ("node_visits", C.c_uint64), ("tri_tests", C.c_uint64),
                ("render_ms", C.c_double), ("trace_ms", C.c_double), ("shadow_ms", C.c_double),
                ("shade_ms", C.c_double), ("bvh_nodes", C.c_uint32), ("bvh_depth", C.c_uint32),
                ("bvh_stack_need", C.c_uint32), ("pad", C.c_uint32), ("bvh_build_ms", C.c_double),
                ("trace_node_visits", C.c_uint64), ("trace_tri_tests", C.c_uint64),
                ("shadow_node_visits", C.c_uint64), ("shadow_tri_tests", C.c_uint64),
                ("trace_launches", C.c_uint32), ("shadow_launches", C.c_uint32),
                ("guide_infer_ms", C.c_double), ("guide_train_ms", C.c_double), ("guide_queries", C.c_uint64),
                ("guide_records", C.c_uint64), ("guide_steps", C.c_uint32), ("batch_passes", C.c_uint32),
                ("graph_ms", C.c_double)]


class phg_gen_opts(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("seed", C.c_uint32), ("target_tris", C.c_uint32),
                ("num_instances", C.c_uint32), ("num_lights", C.c_uint32), ("subdiv", C.c_uint32),
                ("width", C.c_uint32), ("height", C.c_uint32), ("fov_deg", C.c_float), ("pad", C.c_uint32)]


class phg_scene_host(C.Structure):
    _fields_ = [("desc", phg_scene_desc)]


# every entry point declared in include/phg.h
EXPORTED = ["phg_create", "phg_destroy", "phg_last_error", "phg_upload_scene", "phg_intersect",
            "phg_intersect_counted", "phg_film_clear", "phg_film_read", "phg_render", "phg_get_stats",
            "phg_guide_set_params", "phg_guide_get_params", "phg_guide_forward", "phg_guide_lobes", "phg_guide_gradients",
            "phg_guide_train", "phg_scene_generate", "phg_scene_free", "phg_upload_scenes", "phg_render_batch",
            "phg_batch_film_read", "phg_set_tuning", "phg_get_tuning", "phg_autotune_batch", "phg_scene_parse_p2",
            "phg_scene_load_p2", "phg_film_develop", "phg_film_write_pfm", "phg_intersect_noalloc"]

_lib = None


def lib() -> C.CDLL:
    """Load libphg.so (built by __graft_entry__.build()); raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.phg_create.argtypes = [C.POINTER(vp), C.POINTER(phg_device_opts)]
    L.phg_destroy.argtypes = [vp]
    L.phg_last_error.argtypes = [vp]
    L.phg_last_error.restype = C.c_char_p
    L.phg_upload_scene.argtypes = [vp, C.POINTER(phg_scene_desc), C.POINTER(phg_bvh_opts)]
    L.phg_intersect.argtypes = [vp, vp, C.c_uint64, vp, C.c_int, C.c_int]
    L.phg_intersect_counted.argtypes = [vp, vp, C.c_uint64, vp, vp, vp, C.c_int]
    L.phg_intersect_noalloc.argtypes = [vp, vp, C.c_uint64, vp, C.c_int]
    L.phg_film_clear.argtypes = [vp, C.c_uint32, C.c_uint32]
    L.phg_film_read.argtypes = [vp, vp, C.c_uint64]
    L.phg_render.argtypes = [vp, C.POINTER(phg_render_opts)]
    L.phg_get_stats.argtypes = [vp, C.POINTER(phg_stats)]
    L.phg_guide_set_params.argtypes = [vp, vp, C.c_uint32]
    L.phg_guide_get_params.argtypes = [vp, vp, C.c_uint32]
    L.phg_guide_forward.argtypes = [vp, vp, C.c_uint64, vp]
    L.phg_guide_lobes.argtypes = [vp, vp, C.c_uint64, vp]
    L.phg_guide_gradients.argtypes = [vp, vp, C.c_uint64, vp]
    L.phg_guide_train.argtypes = [vp, vp, C.c_uint64, C.c_float, C.c_uint32]
    L.phg_scene_generate.argtypes = [C.POINTER(phg_gen_opts), C.POINTER(C.POINTER(phg_scene_host))]
    L.phg_scene_free.argtypes = [C.POINTER(phg_scene_host)]
    L.phg_scene_free.restype = None
    L.phg_film_develop.argtypes = [vp, C.c_uint32, C.c_uint32, vp]
    L.phg_film_write_pfm.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_char_p]
    L.phg_scene_parse_p2.argtypes = [C.c_char_p, C.POINTER(phg_p2_settings), C.POINTER(C.POINTER(phg_scene_host)),
                                     C.c_char_p, C.c_uint32]
    L.phg_scene_load_p2.argtypes = [C.c_char_p, C.POINTER(phg_p2_settings), C.POINTER(C.POINTER(phg_scene_host)),
                                    C.c_char_p, C.c_uint32]
    L.phg_upload_scenes.argtypes = [vp, C.POINTER(phg_scene_desc), C.c_uint32, C.POINTER(phg_bvh_opts)]
    L.phg_render_batch.argtypes = [vp, C.POINTER(phg_batch_job), C.c_uint32, C.POINTER(phg_batch_opts),
                                   C.POINTER(phg_batch_result)]
    L.phg_batch_film_read.argtypes = [vp, C.c_uint32, vp, C.c_uint64]
    L.phg_set_tuning.argtypes = [vp, C.POINTER(phg_tuning)]
    L.phg_get_tuning.argtypes = [vp, C.POINTER(phg_tuning)]
    L.phg_autotune_batch.argtypes = [vp, C.POINTER(phg_scene_desc), C.c_uint32, C.POINTER(phg_batch_job), C.c_uint32,
                                     C.POINTER(phg_autotune_opts), C.POINTER(phg_tune_trial), C.POINTER(phg_tune_trial),
                                     C.c_uint32, C.POINTER(C.c_uint32)]
    for name in EXPORTED:
        if name not in ("phg_last_error", "phg_scene_free"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


class PhgError(RuntimeError):
    pass


def camera_look_at(pos, target, up=(0.0, 1.0, 0.0), fov_deg: float = 60.0) -> dict:
    """Pinhole camera dict (Scene.camera layout) at pos looking at target; fov is horizontal."""
    pos, target, up = (np.asarray(v, np.float64) for v in (pos, target, up))
    fwd = target - pos
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, up)
    right /= np.linalg.norm(right)
    upv = np.cross(right, fwd)
    return {"pos": list(map(float, pos)), "right": list(map(float, right)), "up": list(map(float, upv)),
            "forward": list(map(float, fwd)), "tan_half_fov": float(np.tan(np.radians(fov_deg) / 2))}


def _camera_struct(cam: dict):
    return phg_camera((C.c_float * 3)(*cam["pos"]), (C.c_float * 3)(*cam["right"]), (C.c_float * 3)(*cam["up"]),
                      (C.c_float * 3)(*cam["forward"]), cam["tan_half_fov"], (C.c_float * 3)(0, 0, 0))


def _job_array(jobs):
    """jobs: dicts with scene, width, height, spp, max_bounces, seed and optionally camera."""
    arr = (phg_batch_job * len(jobs))()
    for i, j in enumerate(jobs):
        cam = j.get("camera")
        arr[i] = phg_batch_job(j["scene"], j["width"], j["height"], j["spp"], j.get("max_bounces", 10000),
                               j.get("seed", 0), JOB_CAMERA if cam else 0, j.get("estimator", 0),
                               _camera_struct(cam) if cam else phg_camera(), j.get("filter", 0), j.get("sampler", 0),
                               (C.c_uint32 * 2)(0, 0))
    return arr


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclass
class Scene:
    """A flattened world-space scene held in numpy arrays (the phg_scene_desc layout)."""
    verts: np.ndarray          # (T, 9) float32
    tri_material: np.ndarray   # (T,) uint32
    tri_emitter: np.ndarray    # (T,) int32
    materials: np.ndarray      # (M,) structured, phg_material layout
    emitters: np.ndarray       # (E,) structured, phg_emitter layout
    emitter_tris: np.ndarray   # (K,) uint32
    camera: dict

    MAT_DTYPE = np.dtype([("type", "<u4"), ("rgb", "<f4", 3), ("roughness", "<f4"), ("ior_outer", "<f4"),
                          ("ior_inner", "<f4"), ("fresnel", "<u4"), ("n", "<f4", 3), ("k", "<f4", 3),
                          ("pad", "<f4", 2)])
    EMIT_DTYPE = np.dtype([("radiance", "<f4", 3), ("tri_begin", "<u4"), ("tri_count", "<u4"), ("pad", "<u4", 3)])

    @property
    def num_tris(self) -> int:
        return int(self.verts.shape[0])

    @staticmethod
    def generate(kind: int, seed: int = 0, target_tris: int = 0, num_instances: int = 0, num_lights: int = 0,
                 subdiv: int = 0, width: int = 0, height: int = 0, fov_deg: float = 0.0) -> "Scene":
        L = lib()
        o = phg_gen_opts(kind, seed, target_tris, num_instances, num_lights, subdiv, width, height, fov_deg, 0)
        out = C.POINTER(phg_scene_host)()
        st = L.phg_scene_generate(C.byref(o), C.byref(out))
        if st != PHG_OK:
            raise PhgError(f"phg_scene_generate: {STATUS_NAMES.get(st, st)}")
        try:
            d = out.contents.desc
            T, M, E, K = d.num_tris, d.num_materials, d.num_emitters, d.num_emitter_tris
            verts = np.ctypeslib.as_array(d.verts, shape=(T * 9,)).reshape(T, 9).copy()
            tri_mat = np.ctypeslib.as_array(d.tri_material, shape=(T,)).copy()
            tri_emit = np.ctypeslib.as_array(d.tri_emitter, shape=(T,)).copy()
            mats = np.frombuffer(C.string_at(d.materials, M * Scene.MAT_DTYPE.itemsize), dtype=Scene.MAT_DTYPE).copy()
            emits = np.frombuffer(C.string_at(d.emitters, E * 32), dtype=Scene.EMIT_DTYPE).copy() if E else \
                np.zeros(0, Scene.EMIT_DTYPE)
            etris = np.ctypeslib.as_array(d.emitter_tris, shape=(K,)).copy() if K else np.zeros(0, np.uint32)
            cam = d.camera
            camera = {"pos": list(cam.pos), "right": list(cam.right), "up": list(cam.up),
                      "forward": list(cam.forward), "tan_half_fov": cam.tan_half_fov}
        finally:
            L.phg_scene_free(out)
        return Scene(verts, tri_mat, tri_emit, mats, emits, etris, camera)

    @staticmethod
    def _from_host(out) -> "Scene":
        L = lib()
        try:
            d = out.contents.desc
            T, M, E, K = d.num_tris, d.num_materials, d.num_emitters, d.num_emitter_tris
            verts = np.ctypeslib.as_array(d.verts, shape=(T * 9,)).reshape(T, 9).copy()
            tri_mat = np.ctypeslib.as_array(d.tri_material, shape=(T,)).copy()
            tri_emit = np.ctypeslib.as_array(d.tri_emitter, shape=(T,)).copy()
            mats = np.frombuffer(C.string_at(d.materials, M * Scene.MAT_DTYPE.itemsize), dtype=Scene.MAT_DTYPE).copy()
            emits = np.frombuffer(C.string_at(d.emitters, E * 32), dtype=Scene.EMIT_DTYPE).copy() if E else \
                np.zeros(0, Scene.EMIT_DTYPE)
            etris = np.ctypeslib.as_array(d.emitter_tris, shape=(K,)).copy() if K else np.zeros(0, np.uint32)
            cam = d.camera
            camera = {"pos": list(cam.pos), "right": list(cam.right), "up": list(cam.up),
                      "forward": list(cam.forward), "tan_half_fov": cam.tan_half_fov}
        finally:
            L.phg_scene_free(out)
        return Scene(verts, tri_mat, tri_emit, mats, emits, etris, camera)

    @staticmethod
    def parse_p2(text: str = None, path: str = None):
        """Scene and render settings (dict) from .p2 text or a .p2 file (phg_scene_parse_p2 /
        phg_scene_load_p2)."""
        L = lib()
        st = phg_p2_settings()
        out = C.POINTER(phg_scene_host)()
        err = C.create_string_buffer(512)
        if path is not None:
            rc = L.phg_scene_load_p2(path.encode(), C.byref(st), C.byref(out), err, 512)
        else:
            rc = L.phg_scene_parse_p2(text.encode(), C.byref(st), C.byref(out), err, 512)
        if rc != PHG_OK:
            raise PhgError(f"p2: {STATUS_NAMES.get(rc, rc)}: {err.value.decode()}")
        settings = {k: getattr(st, k) for k, _ in phg_p2_settings._fields_}
        return Scene._from_host(out), settings

    def desc(self) -> phg_scene_desc:
        """A phg_scene_desc pointing into this scene's arrays (keep the Scene alive)."""
        self.verts = np.ascontiguousarray(self.verts, np.float32)
        self.tri_material = np.ascontiguousarray(self.tri_material, np.uint32)
        self.tri_emitter = np.ascontiguousarray(self.tri_emitter, np.int32)
        self.materials = np.ascontiguousarray(self.materials, Scene.MAT_DTYPE)
        self.emitters = np.ascontiguousarray(self.emitters, Scene.EMIT_DTYPE)
        self.emitter_tris = np.ascontiguousarray(self.emitter_tris, np.uint32)
        d = phg_scene_desc()
        d.num_tris = self.num_tris
        d.num_materials = len(self.materials)
        d.num_emitters = len(self.emitters)
        d.num_emitter_tris = len(self.emitter_tris)
        d.verts = _ptr(self.verts, C.c_float)
        d.tri_material = _ptr(self.tri_material, C.c_uint32)
        d.tri_emitter = _ptr(self.tri_emitter, C.c_int32)
        d.materials = C.cast(self.materials.ctypes.data, C.POINTER(phg_material))
        d.emitters = C.cast(self.emitters.ctypes.data, C.POINTER(phg_emitter)) if len(self.emitters) else None
        d.emitter_tris = _ptr(self.emitter_tris, C.c_uint32) if len(self.emitter_tris) else None
        cam = phg_camera()
        for k in ("pos", "right", "up", "forward"):
            getattr(cam, k)[:] = [float(x) for x in self.camera[k]]
        cam.tan_half_fov = float(self.camera["tan_half_fov"])
        d.camera = cam
        return d


def rays_array(org: np.ndarray, dirs: np.ndarray, tmin=0.0, tmax=FLT_MAX) -> np.ndarray:
    """Pack rays into an (N, 8) float32 array in phg_ray layout."""
    n = org.shape[0]
    r = np.empty((n, 8), np.float32)
    r[:, 0:3] = org
    r[:, 3] = tmin
    r[:, 4:7] = dirs
    r[:, 7] = tmax
    return r


class Context:
    """One GPU context: scene, film and wavefront buffers live on the device."""

    def __init__(self, device: int = 0):
        self._L = lib()
        self._h = C.c_void_p()
        st = self._L.phg_create(C.byref(self._h), C.byref(phg_device_opts(device, 0)))
        if st != PHG_OK:
            raise PhgError(f"phg_create: {STATUS_NAMES.get(st, st)}")
        self._scene_ref = None
        self.width = self.height = 0

    def close(self):
        if self._h:
            self._L.phg_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int, what: str):
        if st != PHG_OK:
            msg = self._L.phg_last_error(self._h)
            raise PhgError(f"{what}: {STATUS_NAMES.get(st, st)}: {msg.decode() if msg else ''}")

    def upload(self, scene: Scene, max_leaf_tris: int = 4, sah_bins: int = 16, nthreads: int = 0, width: int = 4):
        d = scene.desc()
        self._check(self._L.phg_upload_scene(self._h, C.byref(d),
                                             C.byref(phg_bvh_opts(max_leaf_tris, sah_bins, nthreads, width))),
                    "phg_upload_scene")
        self._scene_ref = scene

    def upload_many(self, scenes, max_leaf_tris: int = 4, sah_bins: int = 16, nthreads: int = 0, width: int = 4):
        """Upload several scenes at once (phg_upload_scenes); jobs address them by index."""
        descs = (phg_scene_desc * len(scenes))(*[s.desc() for s in scenes])
        self._check(self._L.phg_upload_scenes(self._h, descs, len(scenes),
                                              C.byref(phg_bvh_opts(max_leaf_tris, sah_bins, nthreads, width))),
                    "phg_upload_scenes")
        self._scene_ref = list(scenes)
        self._batch_npix = []

    def render_batch(self, jobs, paths_per_pass: int = 0, order: int = BATCH_ORDER_GIVEN,
                     count_traversal: bool = False, graph: bool = False, guided: bool = False, train: bool = False,
                     guide_prob: float = 0.5, guide_lr: float = 1e-3, guide_batch: int = 0, guide_steps: int = 4,
                     guide_depth: int = 0):
        """jobs: sequence of dicts with scene, width, height, spp, max_bounces, seed.
        Returns per job (done_ms, paths)."""
        arr = _job_array(jobs)
        res = (phg_batch_result * len(jobs))()
        flags = (BATCH_COUNT_TRAVERSAL if count_traversal else 0) | (BATCH_GRAPH if graph else 0) | \
            (BATCH_GUIDED if guided else 0) | (BATCH_TRAIN if train else 0)
        opts = phg_batch_opts(paths_per_pass, order, flags, guide_steps, guide_prob, guide_lr, guide_batch, guide_depth)
        self._check(self._L.phg_render_batch(self._h, arr, len(jobs), C.byref(opts), res), "phg_render_batch")
        self._batch_shapes = [(j["height"], j["width"]) for j in jobs]
        return [(r.done_ms, r.paths) for r in res]

    def set_tuning(self, trace_block: int = 128, paths_per_pass: int = 0):
        self._check(self._L.phg_set_tuning(self._h, C.byref(phg_tuning(trace_block, paths_per_pass, 0, 0))),
                    "phg_set_tuning")

    def tuning(self) -> dict:
        t = phg_tuning()
        self._check(self._L.phg_get_tuning(self._h, C.byref(t)), "phg_get_tuning")
        return {"trace_block": t.trace_block, "paths_per_pass": t.paths_per_pass,
                "trace_blocks_per_cu": t.trace_blocks_per_cu}

    def autotune_batch(self, scenes, jobs, sample_jobs: int = 8, sample_spp: int = 4, repeats: int = 2):
        """Run the batch autotuner; returns (best, trials) as dicts.  Leaves the sample scenes
        uploaded and the best launch tuning set."""
        descs = (phg_scene_desc * len(scenes))(*[s.desc() for s in scenes])
        arr = _job_array(jobs)
        best = phg_tune_trial()
        trials = (phg_tune_trial * 64)()
        n = C.c_uint32()
        self._check(self._L.phg_autotune_batch(self._h, descs, len(scenes), arr, len(jobs),
                                               C.byref(phg_autotune_opts(sample_jobs, sample_spp, repeats, 0)),
                                               C.byref(best), trials, 64, C.byref(n)), "phg_autotune_batch")
        self._scene_ref = list(scenes)

        def d(t):
            return {"max_leaf_tris": t.max_leaf_tris, "width": t.width, "trace_block": t.trace_block,
                    "paths_per_pass": t.paths_per_pass, "msamples_per_s": t.msamples_per_s}
        return d(best), [d(trials[i]) for i in range(min(n.value, 64))]

    def batch_film(self, job: int) -> np.ndarray:
        h, w = self._batch_shapes[job]
        out = np.zeros((h, w, 8), np.float64)
        self._check(self._L.phg_batch_film_read(self._h, job, out.ctypes.data, out.size), "phg_batch_film_read")
        return out

    def intersect(self, rays: np.ndarray, any_hit: bool = False) -> np.ndarray:
        """rays: (N, 8) float32 -> hits (N, 4) viewed as t, prim, b, c (structured)."""
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        hits = np.zeros(n, dtype=[("t", "<f4"), ("prim", "<i4"), ("b", "<f4"), ("c", "<f4")])
        self._check(self._L.phg_intersect(self._h, rays.ctypes.data, n, hits.ctypes.data,
                                          QUERY_ANY if any_hit else QUERY_CLOSEST, PTR_HOST), "phg_intersect")
        return hits

    def intersect_counted(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        hits = np.zeros(n, dtype=[("t", "<f4"), ("prim", "<i4"), ("b", "<f4"), ("c", "<f4")])
        nv = np.zeros(n, np.uint32)
        nt = np.zeros(n, np.uint32)
        self._check(self._L.phg_intersect_counted(self._h, rays.ctypes.data, n, hits.ctypes.data, nv.ctypes.data,
                                                  nt.ctypes.data, PTR_HOST), "phg_intersect_counted")
        return hits, nv, nt

    def intersect_device(self, rays_ptr: int, n: int, hits_ptr: int, any_hit: bool = False):
        """Device-pointer variant (e.g. torch tensors' data_ptr())."""
        self._check(self._L.phg_intersect(self._h, C.c_void_p(rays_ptr), n, C.c_void_p(hits_ptr),
                                          QUERY_ANY if any_hit else QUERY_CLOSEST, PTR_DEVICE), "phg_intersect")

    def intersect_noalloc(self, rays_ptr: int, n: int, hits_ptr: int, any_hit: bool = False):
        """phg_intersect_noalloc on device pointers: no device allocation by the library."""
        self._check(self._L.phg_intersect_noalloc(self._h, C.c_void_p(rays_ptr), n, C.c_void_p(hits_ptr),
                                                  QUERY_ANY if any_hit else QUERY_CLOSEST), "phg_intersect_noalloc")

    def film_clear(self, width: int, height: int):
        self._check(self._L.phg_film_clear(self._h, width, height), "phg_film_clear")
        self.width, self.height = width, height

    def render(self, spp: int, max_bounces: int = 10000, seed: int = 0, spp_begin: int = 0,
               spp_per_pass: int = 0, count_traversal: bool = False, guided: bool = False, train: bool = False,
               guide_prob: float = 0.5, guide_lr: float = 1e-3, guide_batch: int = 0, guide_steps: int = 4,
               graph: bool = False, estimator: int = 0, filter: int = 0, guide_depth: int = 0, sampler: int = 0):
        flags = (RENDER_COUNT_TRAVERSAL if count_traversal else 0) | (RENDER_GUIDED if guided else 0) | \
            (RENDER_TRAIN if train else 0) | (RENDER_GRAPH if graph else 0)
        o = phg_render_opts(self.width, self.height, spp, spp_begin, max_bounces, spp_per_pass, seed, flags,
                            guide_prob, guide_lr, guide_batch, guide_steps, estimator, filter, guide_depth, sampler)
        self._check(self._L.phg_render(self._h, C.byref(o)), "phg_render")

    # ---- guiding network
    def guide_set_params(self, params: np.ndarray):
        p = np.ascontiguousarray(params, np.float32)
        self._check(self._L.phg_guide_set_params(self._h, p.ctypes.data, p.size), "phg_guide_set_params")

    def guide_get_params(self) -> np.ndarray:
        p = np.zeros(GUIDE_PARAMS, np.float32)
        self._check(self._L.phg_guide_get_params(self._h, p.ctypes.data, p.size), "phg_guide_get_params")
        return p

    def guide_forward(self, features: np.ndarray) -> np.ndarray:
        f = np.ascontiguousarray(features, np.float32).reshape(-1, 9)
        out = np.zeros((f.shape[0], GUIDE_OUTPUTS), np.float32)
        self._check(self._L.phg_guide_forward(self._h, f.ctypes.data, f.shape[0], out.ctypes.data), "phg_guide_forward")
        return out

    def guide_lobes(self, features: np.ndarray) -> np.ndarray:
        """Decoded mixture per query: 8 x (unit mean 3, kappa, weight), then the selection probability."""
        f = np.ascontiguousarray(features, np.float32).reshape(-1, 9)
        out = np.zeros((f.shape[0], GUIDE_OUTPUTS), np.float32)
        self._check(self._L.phg_guide_lobes(self._h, f.ctypes.data, f.shape[0], out.ctypes.data), "phg_guide_lobes")
        return out

    def guide_gradients(self, records: np.ndarray) -> np.ndarray:
        r = np.ascontiguousarray(records, np.float32).reshape(-1, GUIDE_RECORD)
        g = np.zeros(GUIDE_PARAMS, np.float32)
        self._check(self._L.phg_guide_gradients(self._h, r.ctypes.data, r.shape[0], g.ctypes.data),
                    "phg_guide_gradients")
        return g

    def guide_train(self, records: np.ndarray, lr: float = 1e-3, steps: int = 1):
        r = np.ascontiguousarray(records, np.float32).reshape(-1, GUIDE_RECORD)
        self._check(self._L.phg_guide_train(self._h, r.ctypes.data, r.shape[0], lr, steps), "phg_guide_train")

    def film(self) -> np.ndarray:
        """(H, W, 8) float64: sum r, g, b, weight, sum r^2, g^2, b^2, 0; row 0 = bottom row."""
        out = np.zeros((self.height, self.width, 8), np.float64)
        self._check(self._L.phg_film_read(self._h, out.ctypes.data, out.size), "phg_film_read")
        return out

    def stats(self) -> dict:
        s = phg_stats()
        self._check(self._L.phg_get_stats(self._h, C.byref(s)), "phg_get_stats")
        return {k: getattr(s, k) for k, _ in phg_stats._fields_ if k != "pad"}


def write_pfm(film: np.ndarray, path: str):
    """Develop an (H, W, 8) FP64 film and write it as a .pfm file (phg_film_write_pfm)."""
    f = np.ascontiguousarray(film, np.float64)
    h, w = f.shape[:2]
    st = lib().phg_film_write_pfm(f.ctypes.data, w, h, path.encode())
    if st != PHG_OK:
        raise PhgError(f"phg_film_write_pfm: {STATUS_NAMES.get(st, st)}")


def develop(film: np.ndarray) -> np.ndarray:
    """Mean radiance per pixel (HdrRgbFilm::developRegion, Core/Filmic/HdrRgbFilm.cpp:167-207)."""
    w = film[..., 3:4]
    return np.where(w > 0, film[..., 0:3] / np.where(w > 0, w, 1.0), 0.0)
