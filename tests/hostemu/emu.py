"""ctypes binding of the test-only host build of the product's BVH build + traversal."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libhostemu.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        _lib = C.CDLL(LIB)
        vp = C.c_void_p
        _lib.emu_intersect.argtypes = [vp, C.c_uint32, vp, C.c_uint32, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32,
                                       vp, vp, vp, vp, vp, vp]
        _lib.emu_guide_output_grad.argtypes = [vp, C.c_uint32, vp, vp, vp, vp, vp, vp]
        _lib.emu_guide_sel.argtypes = [vp, C.c_uint32, vp]
        _lib.emu_guide_pdf.argtypes = [vp, C.c_uint32, vp, vp]
        _lib.emu_guide_sample.argtypes = [vp, C.c_uint32, vp, vp]
        _lib.emu_guide_encode.argtypes = [vp, C.c_uint32, vp]
        _lib.emu_pixel_order.argtypes = [C.c_uint32, C.c_uint32, vp, vp]
        _lib.emu_halton_values.argtypes = [vp, vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, vp]
        _lib.emu_halton_raster.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                           C.c_uint64, vp]
    return _lib


def intersect(verts, rays, any_hit=False, max_leaf=4, bins=16, width=4):
    verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 9)
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = rays.shape[0]
    prim = np.zeros(n, np.int32)
    t, b, c = (np.zeros(n, np.float32) for _ in range(3))
    nv, nt = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    st = lib().emu_intersect(verts.ctypes.data, verts.shape[0], rays.ctypes.data, n, int(any_hit), max_leaf, bins, width,
                             prim.ctypes.data, t.ctypes.data, b.ctypes.data, c.ctypes.data, nv.ctypes.data,
                             nt.ctypes.data)
    if st < 0:
        raise RuntimeError(f"emu_intersect failed ({st})")
    intersect.last_stack_need = st
    return prim, t, b, c, nv, nt


def _f(a, shape):
    return np.ascontiguousarray(np.asarray(a, np.float32).reshape(shape))


def guide_output_grad(o, w3, weight, bsdf_pdf=None, sel=None, samp_pdf=None):
    o = _f(o, (-1, 41)); n = o.shape[0]
    w3 = _f(w3, (n, 3)); weight = _f(weight, (n,))
    b = _f(np.zeros(n) if bsdf_pdf is None else bsdf_pdf, (n,))
    a = _f(np.zeros(n) if sel is None else sel, (n,))
    lrn = _f(np.zeros(n) if samp_pdf is None else samp_pdf, (n,))
    out = np.zeros((n, 41), np.float32)
    lib().emu_guide_output_grad(o.ctypes.data, n, w3.ctypes.data, weight.ctypes.data, b.ctypes.data, a.ctypes.data,
                                lrn.ctypes.data, out.ctypes.data)
    return out


def guide_sel(o):
    o = _f(o, (-1, 41)); n = o.shape[0]
    out = np.zeros(n, np.float32)
    lib().emu_guide_sel(o.ctypes.data, n, out.ctypes.data)
    return out


def guide_pdf(o, w3):
    o = _f(o, (-1, 41)); n = o.shape[0]
    w3 = _f(w3, (n, 3))
    out = np.zeros(n, np.float32)
    lib().emu_guide_pdf(o.ctypes.data, n, w3.ctypes.data, out.ctypes.data)
    return out


def guide_sample(o40, u3):
    o40 = _f(o40, (-1,))[:40].copy(); u3 = _f(u3, (-1, 3)); n = u3.shape[0]
    out = np.zeros((n, 3), np.float32)
    lib().emu_guide_sample(o40.ctypes.data, n, u3.ctypes.data, out.ctypes.data)
    return out


def guide_encode(x9):
    x9 = _f(x9, (-1, 9)); n = x9.shape[0]
    out = np.zeros((n, 64), np.float32)
    lib().emu_guide_encode(x9.ctypes.data, n, out.ctypes.data)
    return out


def pixel_order(W, H):
    """(pixel of each path slot, slot of each pixel) of a W x H job."""
    p = np.zeros(W * H, np.uint32)
    q = np.zeros(W * H, np.uint32)
    lib().emu_pixel_order(W, H, p.ctypes.data, q.ctypes.data)
    return p, q


def halton_values(dims, index, sampler, leap=1, seed=0):
    """The product's Halton dimension values (float32) for arrays of dimensions and indices."""
    d = np.ascontiguousarray(dims, np.uint32)
    i = np.ascontiguousarray(index, np.uint64)
    out = np.zeros(len(d), np.float32)
    lib().emu_halton_values(d.ctypes.data, i.ctypes.data, len(d), sampler, leap, seed, out.ctypes.data)
    return out


def halton_raster(W, H, sample, filter, sampler, leap=1, seed=0):
    out = np.zeros((W * H, 2), np.float64)
    lib().emu_halton_raster(W, H, sample, filter, sampler, leap, seed, out.ctypes.data)
    return out
