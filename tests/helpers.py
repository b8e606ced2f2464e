"""Shared inputs for the parity tests: random ray batches, triangle soups, per-pixel CIs."""
import numpy as np

FLT_MAX = float(np.finfo(np.float32).max)


def random_soup(n, seed, spread=10.0, size=1.0):
    """n random triangles in a cube of half-width `spread`."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, (n, 1, 3))
    v = c + rng.normal(0, size, (n, 3, 3))
    return v.reshape(n, 9).astype(np.float32)


def nested_soup(n, seed, ratio=0.97):
    """Triangles at geometrically shrinking sizes and distances from the origin: SAH peels
    one off at a time, so the BVH is deep and traversal needs a long stack."""
    rng = np.random.default_rng(seed)
    k = np.arange(n)
    s = ratio ** k
    c = np.stack([s, 0.2 * s * rng.standard_normal(n), 0.2 * s * rng.standard_normal(n)], 1)[:, None, :]
    v = c + 0.3 * s[:, None, None] * rng.uniform(-1, 1, (n, 3, 3))
    return v.reshape(n, 9).astype(np.float32)


def rays_towards(verts, n, seed, jitter=0.0):
    """Rays from random origins aimed at random points of random triangles (mostly hits)."""
    rng = np.random.default_rng(seed)
    v = verts.reshape(-1, 3, 3).astype(np.float64)
    tri = rng.integers(0, v.shape[0], n)
    w = rng.dirichlet((1, 1, 1), n)
    target = np.einsum("nk,nkd->nd", w, v[tri])
    lo, hi = v.reshape(-1, 3).min(0), v.reshape(-1, 3).max(0)
    ext = hi - lo
    org = rng.uniform(lo - 0.5 * ext, hi + 0.5 * ext, (n, 3))
    d = target - org + rng.normal(0, jitter, (n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.empty((n, 8), np.float32)
    r[:, 0:3] = org
    r[:, 3] = 0.0
    r[:, 4:7] = d
    r[:, 7] = FLT_MAX
    return r


def random_rays(verts, n, seed):
    """Rays with uniformly random directions from points around the scene."""
    rng = np.random.default_rng(seed)
    v = verts.reshape(-1, 3)
    lo, hi = v.min(0), v.max(0)
    org = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.empty((n, 8), np.float32)
    r[:, 0:3] = org
    r[:, 3] = 0.0
    r[:, 4:7] = d
    r[:, 7] = FLT_MAX
    return r


def pixel_z(film_a, film_b):
    """Per-pixel, per-channel z statistic of the difference of two independent-sample
    means, from FP64 films holding sum, count and sum of squares."""
    def mv(f):
        n = f[..., 3:4]
        m = f[..., 0:3] / n
        var = (f[..., 4:7] - n * m * m) / np.maximum(n - 1, 1)
        return m, np.maximum(var, 0) / n
    ma, va = mv(film_a)
    mb, vb = mv(film_b)
    se = np.sqrt(va + vb)
    d = ma - mb
    z = np.where(se > 0, d / np.where(se > 0, se, 1), np.where(np.abs(d) > 1e-6 * (np.abs(ma) + 1e-12), np.inf, 0.0))
    return z


def image_mean_z(film_a, film_b):
    """Per-channel z statistic of the difference of two independent renders' image means
    (mean over pixels of the per-pixel means), from the per-pixel variances of the mean."""
    def mv(f):
        n = f[..., 3:4]
        m = f[..., 0:3] / n
        var = np.maximum(f[..., 4:7] - n * m * m, 0) / np.maximum(n - 1, 1) / n
        P = m.shape[0] * m.shape[1]
        return m.reshape(-1, 3).mean(0), var.reshape(-1, 3).sum(0) / (P * P)
    ma, va = mv(film_a)
    mb, vb = mv(film_b)
    return (ma - mb) / np.sqrt(va + vb)


def scene_from_tris(verts, mat_type=1, rgb=(0.5, 0.5, 0.5), roughness=0.3, ior=1.5, camera=None):
    """phg.Scene with one material for all triangles and no emitters."""
    import phg
    n = verts.shape[0]
    mats = np.zeros(1, phg.Scene.MAT_DTYPE)
    mats["type"] = mat_type
    mats["rgb"] = rgb
    mats["roughness"] = roughness
    mats["ior_outer"] = 1.0
    mats["ior_inner"] = ior
    cam = camera or {"pos": [0, 0, -5], "right": [1, 0, 0], "up": [0, 1, 0], "forward": [0, 0, 1],
                     "tan_half_fov": 0.5}
    return phg.Scene(np.ascontiguousarray(verts, np.float32), np.zeros(n, np.uint32), np.full(n, -1, np.int32),
                     mats, np.zeros(0, phg.Scene.EMIT_DTYPE), np.zeros(0, np.uint32), cam)


def furnace_scene(albedo=0.5, radiance=0.5, mat_type=1, roughness=0.4):
    """Closed unit cube whose six inward-facing walls are all emitters with constant
    radiance and all share one material; with a Lambertian albedo rho every pixel's expected
    value is radiance / (1 - rho) (unbounded path length)."""
    import phg
    P = lambda x, y, z: np.array([x, y, z], np.float32)
    quads = [  # counter-clockwise seen from inside the cube
        (P(0, 0, 0), P(0, 0, 1), P(1, 0, 1), P(1, 0, 0)),  # floor, +y
        (P(0, 1, 0), P(1, 1, 0), P(1, 1, 1), P(0, 1, 1)),  # ceiling, -y
        (P(0, 0, 1), P(0, 1, 1), P(1, 1, 1), P(1, 0, 1)),  # z=1, -z
        (P(0, 0, 0), P(1, 0, 0), P(1, 1, 0), P(0, 1, 0)),  # z=0, +z
        (P(0, 0, 0), P(0, 1, 0), P(0, 1, 1), P(0, 0, 1)),  # x=0, +x
        (P(1, 0, 0), P(1, 0, 1), P(1, 1, 1), P(1, 1, 0)),  # x=1, -x
    ]
    verts, emit = [], []
    for qi, (a, b, c, d) in enumerate(quads):
        verts += [np.concatenate([a, b, c]), np.concatenate([a, c, d])]
        emit += [qi, qi]
    verts = np.array(verts, np.float32)
    n = len(verts)
    mats = np.zeros(1, phg.Scene.MAT_DTYPE)
    mats["type"] = mat_type
    mats["rgb"] = (albedo, albedo, albedo)
    mats["roughness"] = roughness
    mats["ior_outer"] = 1.0
    mats["ior_inner"] = 1.5
    emitters = np.zeros(6, phg.Scene.EMIT_DTYPE)
    emitters["radiance"] = radiance
    emitters["tri_begin"] = np.arange(6) * 2
    emitters["tri_count"] = 2
    cam = {"pos": [0.5, 0.5, 0.1], "right": [-1, 0, 0], "up": [0, 1, 0], "forward": [0, 0, 1], "tan_half_fov": 0.8}
    return phg.Scene(verts, np.zeros(n, np.uint32), np.array(emit, np.int32), mats, emitters,
                     np.arange(n, dtype=np.uint32), cam)


def icosphere(subdiv=3, radius=1.0, center=(0.0, 0.0, 0.0)):
    """(T, 9) float32 triangles of a subdivided icosahedron, counter-clockwise seen from
    outside (face normals point outwards)."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache, nf = {}, []

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[k] = len(verts) - 1
            return cache[k]
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    P = np.array(verts) * radius + np.asarray(center, np.float64)
    tri = P[np.array(faces)]                          # (T, 3, 3)
    n = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    outward = np.einsum("ij,ij->i", n, tri.mean(1) - np.asarray(center)) > 0
    tri[~outward] = tri[~outward][:, ::-1]
    return tri.reshape(-1, 9).astype(np.float32)


def box_tris(lo, hi, inward=False):
    """(12, 9) float32 triangles of an axis-aligned box; faces point outwards (or inwards)."""
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    c = lambda i, j, k: np.array([(lo, hi)[i][0], (lo, hi)[j][1], (lo, hi)[k][2]])
    quads = [(c(0, 0, 0), c(1, 0, 0), c(1, 0, 1), c(0, 0, 1)),   # y = lo, normal -y
             (c(0, 1, 0), c(0, 1, 1), c(1, 1, 1), c(1, 1, 0)),   # y = hi, +y
             (c(0, 0, 0), c(0, 1, 0), c(1, 1, 0), c(1, 0, 0)),   # z = lo, -z
             (c(0, 0, 1), c(1, 0, 1), c(1, 1, 1), c(0, 1, 1)),   # z = hi, +z
             (c(0, 0, 0), c(0, 0, 1), c(0, 1, 1), c(0, 1, 0)),   # x = lo, -x
             (c(1, 0, 0), c(1, 1, 0), c(1, 1, 1), c(1, 0, 1))]   # x = hi, +x
    tris = []
    for a, b, cc, d in quads:
        t1, t2 = [a, b, cc], [a, cc, d]
        if inward:
            t1, t2 = t1[::-1], t2[::-1]
        tris += [np.concatenate(t1), np.concatenate(t2)]
    return np.array(tris, np.float32)


def glass_furnace_scene(shape="sphere", camera="outside", radiance=0.5, ior=1.5, half=5.0, subdiv=5):
    """Gray furnace with a glass object: the six inward-facing walls of a box of half-width
    `half` are area lights of constant radiance (their material absorbs, as area-light
    surfaces do), and a smooth GGX dielectric (roughness 0, i.e. alpha at the 0.001 floor;
    IOR 1 outside, `ior` inside) sphere of radius 1 or cube of half-width 1 sits at the
    centre.  The dielectric neither absorbs nor emits, so with the camera outside the object
    every pixel's expected value is `radiance`; with the camera inside it, radiance carries
    the (ior_inner / ior_outer)^2 factor of transmission into the denser medium and every
    pixel's expected value is radiance * ior^2 (L / eta^2 is invariant along a path).
    The microfacet model scatters once, so masking at grazing angles loses a little energy
    (about 3e-4 of 0.5 for the sphere of 20480 facets); coarser spheres have more facets seen
    at grazing angles and heavier-tailed estimates."""
    import phg
    obj = icosphere(subdiv) if shape == "sphere" else box_tris((-1, -1, -1), (1, 1, 1))
    walls = box_tris((-half, -half, -half), (half, half, half), inward=True)
    verts = np.concatenate([walls, obj]).astype(np.float32)
    nw, n = len(walls), len(walls) + len(obj)
    mats = np.zeros(2, phg.Scene.MAT_DTYPE)
    mats[0]["type"] = phg.MAT_ABSORBER
    mats[1]["type"] = phg.MAT_DIELECTRIC
    mats[1]["roughness"] = 0.0
    mats[1]["ior_outer"] = 1.0
    mats[1]["ior_inner"] = ior
    tri_mat = np.concatenate([np.zeros(nw, np.uint32), np.ones(n - nw, np.uint32)])
    emitters = np.zeros(6, phg.Scene.EMIT_DTYPE)
    emitters["radiance"] = radiance
    emitters["tri_begin"] = np.arange(6) * 2
    emitters["tri_count"] = 2
    tri_emit = np.concatenate([np.repeat(np.arange(6), 2), np.full(n - nw, -1)]).astype(np.int32)
    if camera == "outside":
        cam = phg.camera_look_at([0.3, 0.4, -3.2], [0, 0, 0], fov_deg=55)
    else:
        cam = phg.camera_look_at([0.12, -0.07, -0.25], [0.3, 0.2, 1.0], fov_deg=70)
    return phg.Scene(verts, tri_mat, tri_emit, mats, emitters, np.arange(nw, dtype=np.uint32), cam)


def film_mean_se(film):
    """Mean over every sample of the film's first channel, and its standard error."""
    n = film[..., 3].sum()
    mean = film[..., 0].sum() / n
    var = max(film[..., 4].sum() / n - mean * mean, 0.0)
    return mean, np.sqrt(var / n)


class DeviceArray:
    """A caller-owned device buffer for tests of the device-pointer entry points, allocated with
    the HIP runtime directly (hipMalloc / hipMemcpy / hipFree through ctypes)."""
    _hip = None

    @classmethod
    def _lib(cls):
        import ctypes as C
        if cls._hip is None:
            import os
            path = "/opt/rocm/lib/libamdhip64.so"
            h = C.CDLL(path if os.path.exists(path) else "libamdhip64.so")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, host: np.ndarray):
        import ctypes as C
        self.host = np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        h = self._lib()
        assert h.hipMalloc(C.byref(self.ptr), max(self.host.nbytes, 1)) == 0
        assert h.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1) == 0   # host to device

    def read(self) -> np.ndarray:
        out = np.empty_like(self.host)
        h = self._lib()
        assert h.hipDeviceSynchronize() == 0
        assert h.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0               # device to host
        return out

    def free(self):
        if self.ptr:
            self._lib().hipFree(self.ptr)
            self.ptr = None
