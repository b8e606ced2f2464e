"""The C-ABI library loads without a GPU and exports every entry point of include/phg.h."""
import os
import re
import subprocess

import numpy as np

import phg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "phg.h")).read()
    return sorted(set(re.findall(r"^\s*(?:phg_status|void|const char\*)\s+(phg_\w+)\s*\(", text, re.M)))


def test_header_declares_binding_list():
    assert declared_symbols() == sorted(phg.EXPORTED)


def test_library_exports_all_symbols():
    L = phg.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", phg.LIB_PATH], capture_output=True, text=True).stdout
    for name in declared_symbols():
        assert re.search(rf"\bT {name}\b", out), f"{name} not exported"


def test_scene_generation_on_host():
    c = phg.Scene.generate(phg.GEN_CORNELL)
    assert c.num_tris == 36 and len(c.emitters) == 1 and list(c.emitter_tris) == [34, 35]
    assert np.allclose(c.emitters["radiance"][0], [17, 12, 4])
    # the light faces down into the box
    v = c.verts[34].reshape(3, 3)
    n = np.cross(v[1] - v[0], v[2] - v[0])
    assert n[1] < 0 and abs(n[0]) < 1e-9 and abs(n[2]) < 1e-9
    p = phg.Scene.generate(phg.GEN_PROCEDURAL)
    assert p.num_tris == (1 << 20) + 32 and len(p.emitters) == 16
    kinds = np.unique(p.materials["type"])
    assert set(kinds) == {phg.MAT_ABSORBER, phg.MAT_LAMBERT, phg.MAT_CONDUCTOR, phg.MAT_DIELECTRIC}
    r = phg.Scene.generate(phg.GEN_OCCLUDED_ROOM)
    assert 150_000 <= r.num_tris <= 200_000 and len(r.emitters) == 1


def test_context_without_gpu_fails_loudly():
    import ctypes
    h = ctypes.c_void_p()
    st = phg.lib().phg_create(ctypes.byref(h), ctypes.byref(phg.phg_device_opts(99, 0)))
    assert st != phg.PHG_OK
