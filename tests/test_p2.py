"""The .p2 scene ingestion (phg_scene_parse_p2 / phg_scene_load_p2) on the host: syntax,
transforms (T * R * S, rotations composed new * old as TDecomposedTransform does), geometry
tessellation and orientation, material and light mapping, render settings."""
import os

import numpy as np
import pytest

import phg

HERE = os.path.dirname(os.path.abspath(__file__))
SCENE = os.path.join(HERE, "golden", "scenes", "glossy_box.p2")
REF_CORNELL = "/root/reference/scenes/cornell_box.p2"


def tri_normals(v):
    v = v.reshape(-1, 3, 3).astype(np.float64)
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    return n / np.linalg.norm(n, axis=1, keepdims=True)


def one_model(geometry, extra="", material="material(matte-opaque) @m = [image albedo \"0.5 0.5 0.5\"];"):
    return f"""{material}
{geometry}
actor(model) @a = [geometry geometry @g] [material material @m];
{extra}"""


def test_rectangle_transform_order():
    text = one_model("geometry(rectangle) @g = [real width 2] [real height 1];",
                     """actor(model).scale(@a) = [vector3 amount "3 3 3"];
actor(model).rotate(@a) = [vector3 axis "0 1 0"] [real degrees 90];
translate(@a) = [vector3 amount "1 2 3"];""")
    sc, st = phg.Scene.parse_p2(text)
    assert sc.num_tris == 2 and st["warnings"] == 0
    v = sc.verts.reshape(-1, 3)
    # local (-1, 0.5, 0) -> scaled (-3, 1.5, 0) -> 90 deg about y -> (0, 1.5, 3) -> + (1, 2, 3)
    assert np.any(np.all(np.abs(v - [1, 3.5, 6]) < 1e-5, axis=1))
    # the rectangle faces +z locally; after the rotation it faces +x
    assert np.allclose(tri_normals(sc.verts), [1, 0, 0], atol=1e-6)


def test_rotations_compose_new_after_old():
    text = one_model("geometry(rectangle) @g = [real width 1] [real height 1];",
                     """rotate(@a) = [vector3 axis "1 0 0"] [real degrees 90];
rotate(@a) = [vector3 axis "0 1 0"] [real degrees 90];""")
    sc, _ = phg.Scene.parse_p2(text)
    # +z --(x, 90)--> -y --(y, 90)--> -y
    assert np.allclose(tri_normals(sc.verts), [0, -1, 0], atol=1e-6)
    # a local +x edge: x --(x, 90)--> x --(y, 90)--> -z
    v = sc.verts.reshape(-1, 3, 3)[0].astype(np.float64)  # (-w/2, h/2), (-w/2, -h/2), (w/2, h/2)
    e = v[2] - v[0]
    assert np.allclose(e / np.linalg.norm(e), [0, 0, -1], atol=1e-6)


def test_cuboid_faces_outward_and_sphere_is_closed():
    sc, _ = phg.Scene.parse_p2(one_model('geometry(cuboid) @g = [vector3 size "2 4 6"] [vector3 offset "1 0 0"];'))
    assert sc.num_tris == 12
    v = sc.verts.reshape(-1, 3, 3).astype(np.float64)
    c = v.mean(axis=1) - [1, 0, 0]
    assert np.all(np.einsum("ij,ij->i", tri_normals(sc.verts), c) > 0)
    assert np.allclose(v.min(axis=(0, 1)), [0, -2, -3]) and np.allclose(v.max(axis=(0, 1)), [2, 2, 3])
    sp, _ = phg.Scene.parse_p2(one_model("geometry(sphere) @g = [real radius 2];"))
    p = sp.verts.reshape(-1, 3).astype(np.float64)
    assert sp.num_tris == 20480 and np.allclose(np.linalg.norm(p, axis=1), 2, atol=1e-5)
    c = sp.verts.reshape(-1, 3, 3).mean(axis=1)
    assert np.all(np.einsum("ij,ij->i", tri_normals(sp.verts), c) > 0)


def test_area_light_power_and_orientation():
    text = """actor(rectangle-light) @l = [spectrum color "1 0.5 0.25"] [real watts 200] [real width 2] [real height 3];
rotate(@l) = [vector3 axis "1 0 0"] [real degrees 90];
translate(@l) = [vector3 amount "0 4 0"];
geometry(rectangle) @g = [real width 1] [real height 1];
actor(model) @floor = [geometry geometry @g];"""
    sc, st = phg.Scene.parse_p2(text)
    assert len(sc.emitters) == 1 and list(sc.emitter_tris) == [0, 1]
    # watts over colour energy x area x pi (AAreaLight::buildEmitter, Spectrum::putEnergy); the
    # colour energy is the reference's estimate (tests/golden/color_energy.json constants)
    import json
    E = json.load(open(os.path.join(HERE, "golden", "color_energy.json")))["constants"]
    energy = E["white"] * 0.25 + E["yellow"] * (0.5 - 0.25) + E["red"] * (1 - 0.5)   # b < g < r
    k = 200 / (energy * 6 * np.pi)
    assert np.allclose(sc.emitters["radiance"][0], [k, 0.5 * k, 0.25 * k], rtol=1e-5)
    assert np.allclose(tri_normals(sc.verts[:2]), [0, -1, 0], atol=1e-6)   # faces down
    assert sc.materials["type"][sc.tri_material[0]] == phg.MAT_ABSORBER


def test_materials_and_settings():
    sc, st = phg.Scene.parse_p2(path=SCENE)
    assert (st["width"], st["height"], st["spp"], st["max_bounces"]) == (64, 48, 16, 12)
    assert st["estimator"] == phg.PHG_ESTIMATOR_BNEEPT and st["filter"] == phg.PHG_FILTER_BOX
    assert st["sampler"] == phg.SAMPLER_STRATIFIED   # sample-source(stratified)
    assert st["warnings"] == 0
    kinds = sorted(set(sc.materials["type"]))
    assert kinds == [phg.MAT_ABSORBER, phg.MAT_LAMBERT, phg.MAT_CONDUCTOR, phg.MAT_DIELECTRIC]
    gold = sc.materials[sc.materials["type"] == phg.MAT_CONDUCTOR][0]
    assert np.allclose(gold["rgb"], [1.0, 0.765557, 0.336057]) and np.isclose(gold["roughness"], 0.35)
    glass = sc.materials[sc.materials["type"] == phg.MAT_DIELECTRIC][0]
    assert np.isclose(glass["ior_inner"], 1.5) and np.isclose(glass["ior_outer"], 1.0)
    assert np.isclose(sc.camera["tan_half_fov"], np.tan(np.radians(27.5)))
    assert np.allclose(sc.camera["pos"], [0, 0, 14]) and np.allclose(sc.camera["forward"], [0, 0, -1])
    assert sc.num_tris == 2 + 5 * 2 + 20480 + 12


def test_roughness_mappings_and_sRGB():
    def mat(m):
        sc, _ = phg.Scene.parse_p2(one_model("geometry(rectangle) @g = [real width 1] [real height 1];", "", m))
        return sc.materials[sc.tri_material[0]]
    eq = mat('material(abraded-opaque) @m = [real roughness 0.25] [enum roughness-to-alpha equaled] [spectrum f0 "0.9 0.9 0.9"];')
    assert np.isclose(eq["roughness"] ** 2, 0.25)   # the device squares: alpha = 0.25
    srgb = mat('image(constant) @c = [real-array values "0.5 0.5 0.5"] [enum color-space SRGB];\n'
               'material(matte-opaque) @m = [image albedo @c];')
    assert np.allclose(srgb["rgb"], ((0.5 + 0.055) / 1.055) ** 2.4)


def test_conductor_fresnel_selection():
    """ConductiveInterfaceInfo::genFresnelEffect: with ior-inner-n / -k, `fresnel exact` is the
    exact conductor Fresnel and `schlick` (the default) Schlick with that index's normal
    reflectance; f0 alone is Schlick, and an exact request without the index falls back to it
    with a warning."""
    def mat(m):
        sc, st = phg.Scene.parse_p2(one_model("geometry(rectangle) @g = [real width 1] [real height 1];", "", m))
        return sc.materials[sc.tri_material[0]], st["warnings"]
    nk = '[spectrum ior-inner-n "0.143 0.374 1.442"] [spectrum ior-inner-k "3.983 2.385 1.603"] [real ior-outer 1.2]'
    ex, w = mat(f'material(abraded-opaque) @m = [real roughness 0.3] [enum fresnel exact] {nk};')
    assert w == 0 and ex["type"] == phg.MAT_CONDUCTOR and ex["fresnel"] == phg.FRESNEL_EXACT
    assert np.allclose(ex["n"], [0.143, 0.374, 1.442]) and np.allclose(ex["k"], [3.983, 2.385, 1.603])
    assert np.isclose(ex["ior_outer"], 1.2)
    for clause in ("[enum fresnel schlick] ", ""):
        sl, w = mat(f'material(abraded-opaque) @m = [real roughness 0.3] {clause}{nk};')
        n, k = np.array([0.143, 0.374, 1.442]), np.array([3.983, 2.385, 1.603])
        assert w == 0 and sl["fresnel"] == phg.FRESNEL_SCHLICK
        assert np.allclose(sl["rgb"], ((n - 1.2) ** 2 + k ** 2) / ((n + 1.2) ** 2 + k ** 2), rtol=1e-5)
    fb, w = mat('material(abraded-opaque) @m = [enum fresnel exact] [spectrum f0 "0.9 0.8 0.7"];')
    assert w == 1 and fb["fresnel"] == phg.FRESNEL_SCHLICK and np.allclose(fb["rgb"], [0.9, 0.8, 0.7])


def test_syntax_errors_and_skipped_commands():
    with pytest.raises(phg.PhgError):
        phg.Scene.parse_p2("geometry(rectangle) @g = [real width 1")          # unterminated clause
    with pytest.raises(phg.PhgError):
        phg.Scene.parse_p2("visualizer(path-tracing) @v = [enum estimator pm];"
                           "geometry(rectangle) @g = [real width 1] [real height 1];"
                           "actor(model) @a = [geometry geometry @g];")        # estimator off this path
    with pytest.raises(phg.PhgError):
        phg.Scene.parse_p2("#version 1.1.0;")                                 # no triangles
    sc, st = phg.Scene.parse_p2(one_model("geometry(rectangle) @g = [real width 1] [real height 1];",
                                          "actor(dome) @d = [real energy-scale 1];\nmystery(@a) = [real x 1];"))
    assert sc.num_tris == 2 and st["warnings"] == 2


@pytest.mark.skipif(not os.path.exists(REF_CORNELL), reason="reference scene files not present")
def test_reference_cornell_scene_loads():
    sc, st = phg.Scene.parse_p2(path=REF_CORNELL)
    assert (st["width"], st["height"], st["spp"]) == (300, 300, 40)
    assert st["estimator"] == phg.PHG_ESTIMATOR_BNEEPT and st["warnings"] == 0
    assert len(sc.emitters) == 1 and sc.num_tris == 2 + 5 * 2 + 20480
    # a white light's colour energy is the white basis energy (tests/golden/color_energy.json)
    import json
    e_white = json.load(open(os.path.join(HERE, "golden", "color_energy.json")))["constants"]["white"]
    assert np.allclose(sc.emitters["radiance"][0], 500 / (e_white * 9 * np.pi), rtol=1e-5)
    v = sc.verts.reshape(-1, 3)
    assert np.allclose(v.min(axis=0), -5) and np.allclose(v.max(axis=0), 5)


def test_pfm_writer_layout(tmp_path):
    """phg_film_write_pfm writes what PfmFile::save writes: 'PF', '<w> <h>', '-1' lines, then
    developed float32 RGB, bottom row first."""
    rng = np.random.default_rng(3)
    H, W = 3, 4
    film = np.zeros((H, W, 8))
    film[..., :3] = rng.random((H, W, 3)) * 5
    film[..., 3] = rng.integers(1, 9, (H, W))
    film[1, 2, 3] = 0.0                                    # no samples: develops to 0
    path = str(tmp_path / "frame.pfm")
    phg.write_pfm(film, path)
    data = open(path, "rb").read()
    header = b"PF\n4 3\n-1\n"
    assert data[:len(header)] == header
    px = np.frombuffer(data[len(header):], "<f4").reshape(H, W, 3)
    expect = np.where(film[..., 3:4] > 0, film[..., :3] / np.where(film[..., 3:4] > 0, film[..., 3:4], 1), 0)
    assert np.allclose(px, expect.astype(np.float32)) and np.all(px[1, 2] == 0)


@pytest.mark.parametrize("clauses,want", [
    ("", phg.SAMPLER_HALTON | phg.HALTON_PERM_FIXED | phg.HALTON_SEQ_ORIGINAL),
    ("[enum permutation owen] [enum sequence random-start]", phg.SAMPLER_HALTON | phg.HALTON_PERM_OWEN
     | phg.HALTON_SEQ_RANDOM_START),
    ("[enum permutation per-digit] [enum sequence leap]", phg.SAMPLER_HALTON | phg.HALTON_PERM_PER_DIGIT
     | phg.HALTON_SEQ_LEAP),
    ("[enum permutation none] [enum sequence random-leap]", phg.SAMPLER_HALTON | phg.HALTON_PERM_NONE
     | phg.HALTON_SEQ_RANDOM_LEAP),
])
def test_halton_sample_source(clauses, want):
    """sample-source(halton) with HaltonSampleSource's permutation (default fixed) and sequence
    (default original)."""
    text = f"sample-source(halton) @s = [integer samples 8] {clauses};\n" + \
        "geometry(rectangle) @g = [real width 1] [real height 1];\nactor(model) @m = [geometry geometry @g];"
    _, st = phg.Scene.parse_p2(text)
    assert st["sampler"] == want and st["spp"] == 8 and st["warnings"] == 0
