"""phg_intersect_noalloc (the allocation-free ray query, device pointers) against
phg_intersect on the same device buffers, bit for bit: a shallow BVH and a deep one whose
traversal stack spills into the scratch array, closest and any hit, two calls in a row with
the second smaller (the scratch keeps the previous call's contents).  The buffers are
allocated by the caller (hipMalloc through ctypes); the library allocates nothing for the query."""
import numpy as np
import pytest

import phg
from helpers import DeviceArray, nested_soup, random_soup, rays_towards, random_rays, scene_from_tris

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["soup", "deep"])
def test_noalloc_query_equals_phg_intersect(kind):
    verts = random_soup(3000, 71) if kind == "soup" else nested_soup(1000, 21, ratio=0.98)
    rays = np.concatenate([rays_towards(verts, 60_000, 73), random_rays(verts, 20_000, 74)])
    if kind == "deep":
        rays[:, 0:3] *= 0.05      # origins near the nest's centre: long traversal stacks
    print(f"[noalloc] case={kind} tris={len(verts)} rays={len(rays)}", flush=True)
    with phg.Context(0) as c:
        c.upload(scene_from_tris(verts))
        st = c.stats()
        print(f"[noalloc] bvh depth={st['bvh_depth']} stack_need={st['bvh_stack_need']}", flush=True)
        if kind == "deep":
            assert st["bvh_stack_need"] > 32          # the 16-entry ring spills
        for n, any_hit in ((len(rays), False), (len(rays) // 3, False), (len(rays), True), (len(rays) // 5, True)):
            d_rays = DeviceArray(rays[:n])
            ref = DeviceArray(np.zeros((n, 4), np.float32))
            out = DeviceArray(np.full((n, 4), -7.0, np.float32))
            try:
                c.intersect_device(d_rays.ptr.value, n, ref.ptr.value, any_hit=any_hit)
                c.intersect_noalloc(d_rays.ptr.value, n, out.ptr.value, any_hit=any_hit)
                a, b = ref.read(), out.read()
            finally:
                for d in (d_rays, ref, out):
                    d.free()
            same = np.array_equal(a.view(np.uint32), b.view(np.uint32))
            hits = int((b[:, 1].view(np.int32) >= 0).sum())
            print(f"[noalloc] n={n} any={any_hit} hits={hits} bit_equal={same}", flush=True)
            assert same
            assert hits > n // 4
