/* CPU oracle of the Photon-v2 BNEEPT hot path — test infrastructure only.
 * Loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg; never by the
 * product.  The scene structs mirror the product's flattened scene format field by field
 * (include/phg.h) so that one set of host arrays feeds both. */
#ifndef BNEEPT_ORACLE_H
#define BNEEPT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_MAT_ABSORBER = 0, ORC_MAT_LAMBERT = 1, ORC_MAT_CONDUCTOR = 2, ORC_MAT_DIELECTRIC = 3 };

typedef struct orc_material {
    uint32_t type;
    float rgb[3];      /* albedo (Lambert) or F0 (conductor, Schlick) */
    float roughness;   /* alpha = max(roughness^2, 0.001) */
    float ior_outer;
    float ior_inner;
    uint32_t fresnel;  /* conductor: 0 Schlick from rgb, 1 exact from ior_outer, n, k */
    float n[3];
    float k[3];
    float pad[2];
} orc_material;

typedef struct orc_emitter {
    float radiance[3];
    uint32_t tri_begin; /* into emitter_tris */
    uint32_t tri_count;
    uint32_t pad[3];
} orc_emitter;

typedef struct orc_camera {
    float pos[3];
    float right[3];
    float up[3];
    float forward[3];
    float tan_half_fov; /* tan(horizontal FOV / 2) */
    float pad[3];
} orc_camera;

typedef struct orc_scene {
    uint32_t num_tris;
    uint32_t num_materials;
    uint32_t num_emitters;
    uint32_t num_emitter_tris;
    const float* verts;            /* 9 floats per triangle: A, B, C */
    const uint32_t* tri_material;  /* per triangle */
    const int32_t* tri_emitter;    /* per triangle, -1 if not emissive */
    const orc_material* materials;
    const orc_emitter* emitters;
    const uint32_t* emitter_tris;
    orc_camera camera;
} orc_scene;

typedef struct orc_render_opts {
    uint32_t width, height;
    uint32_t spp_begin, spp;
    uint32_t max_bounces;
    uint32_t nthreads;
    uint64_t seed;
    uint32_t row_step;  /* 0 or 1 = all rows; k = only rows y with y % k == 0 */
    uint32_t estimator; /* 0 BNEEPT, 1 BVPT, 2 BVPTDL */
    uint32_t sampler;   /* 0 independent, 1 stratified first-vertex dimensions, 2 Halton (with the
                           permutation in bits 8-9 and the sequence in bits 12-13) */
    uint32_t pad;
} orc_render_opts;

typedef struct orc_render_stats {
    uint64_t rays_primary, rays_extension, rays_shadow;
    uint64_t paths;
    double render_ms;   /* path tracing only, after the BVH build */
} orc_render_stats;

#ifdef __cplusplus
}
#endif
#endif
